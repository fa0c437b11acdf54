// flowfilter_amd.hpp -- C++ interface of libflowfilter_amd.so: the classes of
// namespace flowfilter::gpu (include/flowfilter/gpu/*.h of the reference)
// as thin, header-only wrappers over the C ABI in flowfilter_amd.h.
//
// Same class names, constructors, method names and argument meanings as the
// reference; failures are thrown as the reference's exception types
// (std::invalid_argument, std::logic_error, std::runtime_error for the
// reference's bare std::exception) instead of terminating the process.
// Objects are movable handles; GPUImage copies share the device buffer
// like the reference's shared_ptr-based GPUImage.
#ifndef FLOWFILTER_AMD_HPP_
#define FLOWFILTER_AMD_HPP_

#include <cstddef>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "flowfilter_amd.h"

namespace flowfilter {

// flowfilter::image_t (include/flowfilter/image.h:30-49), layout-identical.
typedef ff_image_t image_t;

inline void check(ff_status st) {
    if (st == FF_OK) return;
    const std::string msg = ff_last_error();
    switch (st) {
        case FF_ERR_INVALID_ARGUMENT:
        case FF_ERR_NULL_HANDLE: throw std::invalid_argument(msg);
        case FF_ERR_LOGIC:
        case FF_ERR_WRONG_TYPE: throw std::logic_error(msg);
        case FF_ERR_BAD_ALLOC: throw std::bad_alloc();
        default: throw std::runtime_error(msg);
    }
}

// createImage / destroyImage (include/flowfilter/image.h:59-78).
inline image_t createImage(int height, int width, int depth, std::size_t pixelSize) {
    image_t img;
    check(ff_create_image(height, width, depth, pixelSize, &img));
    return img;
}
inline image_t createImage(int height, int width, std::size_t pixelSize) {
    return createImage(height, width, 1, pixelSize);
}
inline void destroyImage(image_t& img) { check(ff_destroy_image(&img)); }
// getColorWheelRGBA (include/flowfilter/colorwheel.h:26): the 256 x 256
// RGBA colour wheel of FlowToColor (read-only library data).
inline image_t getColorWheelRGBA() {
    image_t w{};
    check(ff_get_color_wheel_rgba(&w));
    return w;
}

namespace gpu {

class GPUImage {
public:
    GPUImage() {
        ff_gpuimage* h = nullptr;
        check(ff_gpuimage_create_empty(&h));
        reset(h);
    }
    GPUImage(int height, int width, int depth = 1, int itemSize = sizeof(char)) {
        ff_gpuimage* h = nullptr;
        check(ff_gpuimage_create(height, width, depth, itemSize, &h));
        reset(h);
    }
    static GPUImage adopt(ff_gpuimage* h) {
        GPUImage g(nullptr);
        g.reset(h);
        return g;
    }
    int height() const { return geti(ff_gpuimage_height); }
    int width() const { return geti(ff_gpuimage_width); }
    int depth() const { return geti(ff_gpuimage_depth); }
    int pitch() const { return geti(ff_gpuimage_pitch); }
    int itemSize() const { return geti(ff_gpuimage_item_size); }
    void* data() const {
        void* p = nullptr;
        check(ff_gpuimage_data(h_.get(), &p));
        return p;
    }
    void upload(image_t& img) { check(ff_gpuimage_upload(h_.get(), &img)); }
    void download(image_t& img) const { check(ff_gpuimage_download(h_.get(), &img)); }
    void copyFrom(GPUImage& img) { check(ff_gpuimage_copy_from(h_.get(), img.h_.get())); }
    void clear() { check(ff_gpuimage_clear(h_.get())); }
    void uploadDevice(const void* dev, std::size_t pitch) { check(ff_gpuimage_upload_device(h_.get(), dev, pitch)); }
    ff_gpuimage* handle() const { return h_.get(); }

private:
    explicit GPUImage(std::nullptr_t) {}
    void reset(ff_gpuimage* h) { h_ = std::shared_ptr<ff_gpuimage>(h, [](ff_gpuimage* p) { ff_gpuimage_destroy(p); }); }
    int geti(ff_status (*f)(const ff_gpuimage*, int*)) const {
        int v = 0;
        check(f(h_.get(), &v));
        return v;
    }
    std::shared_ptr<ff_gpuimage> h_;
};

// Stage (include/flowfilter/gpu/pipeline.h:29-78).
class Stage {
public:
    Stage() = default;
    Stage(Stage&&) = default;
    Stage& operator=(Stage&&) = default;
    virtual ~Stage() = default;
    void configure() { check(ff_stage_configure(s_.get())); }
    void compute() { check(ff_stage_compute(s_.get())); }
    float elapsedTime() const {
        float ms = 0.0f;
        check(ff_stage_elapsed_time(s_.get(), &ms));
        return ms;
    }
    ff_stage* handle() const { return s_.get(); }

protected:
    void reset(ff_stage* s) { s_ = std::unique_ptr<ff_stage, Del>(s); }
    static const ff_gpuimage* H(const GPUImage& g) { return g.handle(); }
    GPUImage img(ff_status (*f)(ff_stage*, ff_gpuimage**)) const {
        ff_gpuimage* h = nullptr;
        check(f(s_.get(), &h));
        return GPUImage::adopt(h);
    }
    int geti(ff_status (*f)(ff_stage*, int*)) const {
        int v = 0;
        check(f(s_.get(), &v));
        return v;
    }
    float getf(ff_status (*f)(ff_stage*, float*)) const {
        float v = 0;
        check(f(s_.get(), &v));
        return v;
    }
    struct Del {
        void operator()(ff_stage* s) const { ff_stage_destroy(s); }
    };
    std::unique_ptr<ff_stage, Del> s_;
};

// ImageModel (include/flowfilter/gpu/imagemodel.h:22-63).
class ImageModel : public Stage {
public:
    ImageModel() { make(nullptr); }
    explicit ImageModel(GPUImage inputImage) { make(H(inputImage)); }
    void setInputImage(GPUImage img) { check(ff_imagemodel_set_input_image(s_.get(), H(img))); }
    GPUImage getImageConstant() { return img(ff_imagemodel_get_image_constant); }
    GPUImage getImageGradient() { return img(ff_imagemodel_get_image_gradient); }

private:
    void make(const ff_gpuimage* in) {
        ff_stage* s = nullptr;
        check(ff_imagemodel_create(in, &s));
        reset(s);
    }
};

// FlowPropagator (include/flowfilter/gpu/propagation.h:25-105).
class FlowPropagator : public Stage {
public:
    FlowPropagator() { make(nullptr, 1); }
    FlowPropagator(GPUImage inputFlow, const int iterations = 1) { make(H(inputFlow), iterations); }
    void setIterations(const int N) { check(ff_flowpropagator_set_iterations(s_.get(), N)); }
    int getIterations() const { return geti(ff_flowpropagator_get_iterations); }
    float getDt() const { return getf(ff_flowpropagator_get_dt); }
    void setBorder(const int border) { check(ff_flowpropagator_set_border(s_.get(), border)); }
    int getBorder() const { return geti(ff_flowpropagator_get_border); }
    void setInvertInputFlow(const bool invert) { check(ff_flowpropagator_set_invert_input_flow(s_.get(), invert)); }
    bool getInvertInputFlow() const { return geti(ff_flowpropagator_get_invert_input_flow) != 0; }
    void setInputFlow(GPUImage flow) { check(ff_flowpropagator_set_input_flow(s_.get(), H(flow))); }
    GPUImage getPropagatedFlow() { return img(ff_flowpropagator_get_propagated_flow); }

private:
    void make(const ff_gpuimage* in, int it) {
        ff_stage* s = nullptr;
        check(ff_flowpropagator_create(in, it, &s));
        reset(s);
    }
};

// FlowPropagatorPayload (include/flowfilter/gpu/propagation.h:108-200).
class FlowPropagatorPayload : public Stage {
public:
    FlowPropagatorPayload() { make(nullptr, nullptr, nullptr, 1); }
    FlowPropagatorPayload(GPUImage inputFlow, GPUImage scalarPayload, GPUImage vectorPayload,
                          const int iterations = 1) {
        make(H(inputFlow), H(scalarPayload), H(vectorPayload), iterations);
    }
    void setIterations(const int N) { check(ff_flowpropagatorpayload_set_iterations(s_.get(), N)); }
    int getIterations() const { return geti(ff_flowpropagatorpayload_get_iterations); }
    float getDt() const { return getf(ff_flowpropagatorpayload_get_dt); }
    void setBorder(const int border) { check(ff_flowpropagatorpayload_set_border(s_.get(), border)); }
    int getBorder() const { return geti(ff_flowpropagatorpayload_get_border); }
    void setInputFlow(GPUImage f) { check(ff_flowpropagatorpayload_set_input_flow(s_.get(), H(f))); }
    void setScalarPayload(GPUImage f) { check(ff_flowpropagatorpayload_set_scalar_payload(s_.get(), H(f))); }
    void setVectorPayload(GPUImage f) { check(ff_flowpropagatorpayload_set_vector_payload(s_.get(), H(f))); }
    GPUImage getPropagatedFlow() { return img(ff_flowpropagatorpayload_get_propagated_flow); }
    GPUImage getPropagatedScalar() { return img(ff_flowpropagatorpayload_get_propagated_scalar); }
    GPUImage getPropagatedVector() { return img(ff_flowpropagatorpayload_get_propagated_vector); }

private:
    void make(const ff_gpuimage* f, const ff_gpuimage* sc, const ff_gpuimage* v, int it) {
        ff_stage* s = nullptr;
        check(ff_flowpropagatorpayload_create(f, sc, v, it, &s));
        reset(s);
    }
};

// FlowUpdate (include/flowfilter/gpu/update.h:21-90).
class FlowUpdate : public Stage {
public:
    FlowUpdate() { make(nullptr, nullptr, nullptr, 1.0f, 1.0f); }
    FlowUpdate(GPUImage inputFlow, GPUImage inputImage, GPUImage inputImageGradient, const float gamma = 1.0,
               const float maxflow = 1.0) {
        make(H(inputFlow), H(inputImage), H(inputImageGradient), gamma, maxflow);
    }
    float getGamma() const { return getf(ff_flowupdate_get_gamma); }
    void setGamma(const float gamma) { check(ff_flowupdate_set_gamma(s_.get(), gamma)); }
    float getMaxFlow() const { return getf(ff_flowupdate_get_maxflow); }
    void setMaxFlow(const float maxflow) { check(ff_flowupdate_set_maxflow(s_.get(), maxflow)); }
    void setInputFlow(GPUImage f) { check(ff_flowupdate_set_input_flow(s_.get(), H(f))); }
    void setInputImage(GPUImage f) { check(ff_flowupdate_set_input_image(s_.get(), H(f))); }
    void setInputImageGradient(GPUImage f) { check(ff_flowupdate_set_input_image_gradient(s_.get(), H(f))); }
    GPUImage getUpdatedFlow() { return img(ff_flowupdate_get_updated_flow); }
    GPUImage getUpdatedImage() { return img(ff_flowupdate_get_updated_image); }

private:
    void make(const ff_gpuimage* f, const ff_gpuimage* i, const ff_gpuimage* g, float gamma, float maxflow) {
        ff_stage* s = nullptr;
        check(ff_flowupdate_create(f, i, g, gamma, maxflow, &s));
        reset(s);
    }
};

// DeltaFlowUpdate (include/flowfilter/gpu/update.h:93-174).
class DeltaFlowUpdate : public Stage {
public:
    DeltaFlowUpdate() { make(nullptr, nullptr, nullptr, nullptr, nullptr, 1.0f, 1.0f); }
    DeltaFlowUpdate(GPUImage inputFlow, GPUImage inputDeltaFlow, GPUImage inputImageOld, GPUImage inputImage,
                    GPUImage inputImageGradient, const float gamma = 1.0, const float maxflow = 1.0) {
        make(H(inputFlow), H(inputDeltaFlow), H(inputImageOld), H(inputImage), H(inputImageGradient), gamma,
             maxflow);
    }
    float getGamma() const { return getf(ff_deltaflowupdate_get_gamma); }
    void setGamma(const float gamma) { check(ff_deltaflowupdate_set_gamma(s_.get(), gamma)); }
    float getMaxFlow() const { return getf(ff_deltaflowupdate_get_maxflow); }
    void setMaxFlow(const float maxflow) { check(ff_deltaflowupdate_set_maxflow(s_.get(), maxflow)); }
    void setInputFlow(GPUImage f) { check(ff_deltaflowupdate_set_input_flow(s_.get(), H(f))); }
    void setInputDeltaFlow(GPUImage f) { check(ff_deltaflowupdate_set_input_delta_flow(s_.get(), H(f))); }
    void setInputImageOld(GPUImage f) { check(ff_deltaflowupdate_set_input_image_old(s_.get(), H(f))); }
    void setInputImage(GPUImage f) { check(ff_deltaflowupdate_set_input_image(s_.get(), H(f))); }
    void setInputImageGradient(GPUImage f) { check(ff_deltaflowupdate_set_input_image_gradient(s_.get(), H(f))); }
    GPUImage getUpdatedFlow() { return img(ff_deltaflowupdate_get_updated_flow); }
    GPUImage getUpdatedDeltaFlow() { return img(ff_deltaflowupdate_get_updated_delta_flow); }
    GPUImage getUpdatedImage() { return img(ff_deltaflowupdate_get_updated_image); }

private:
    void make(const ff_gpuimage* f, const ff_gpuimage* d, const ff_gpuimage* o, const ff_gpuimage* i,
              const ff_gpuimage* g, float gamma, float maxflow) {
        ff_stage* s = nullptr;
        check(ff_deltaflowupdate_create(f, d, o, i, g, gamma, maxflow, &s));
        reset(s);
    }
};

// FlowSmoother (include/flowfilter/gpu/flowsmoothing.h:22-57).
class FlowSmoother : public Stage {
public:
    FlowSmoother() { make(nullptr, 1); }
    FlowSmoother(GPUImage inputFlow, const int iterations) { make(H(inputFlow), iterations); }
    int getIterations() const { return geti(ff_flowsmoother_get_iterations); }
    void setIterations(const int N) { check(ff_flowsmoother_set_iterations(s_.get(), N)); }
    void setInputFlow(GPUImage f) { check(ff_flowsmoother_set_input_flow(s_.get(), H(f))); }
    GPUImage getSmoothedFlow() { return img(ff_flowsmoother_get_smoothed_flow); }

private:
    void make(const ff_gpuimage* f, int it) {
        ff_stage* s = nullptr;
        check(ff_flowsmoother_create(f, it, &s));
        reset(s);
    }
};

// ImagePyramid (include/flowfilter/gpu/pyramid.h:22-89).
class ImagePyramid : public Stage {
public:
    ImagePyramid() { make(nullptr, 1); }
    ImagePyramid(GPUImage image, const int levels) { make(H(image), levels); }
    void setInputImage(GPUImage img_) { check(ff_imagepyramid_set_input_image(s_.get(), H(img_))); }
    void setLevels(const int levels) { check(ff_imagepyramid_set_levels(s_.get(), levels)); }
    int getLevels() const { return geti(ff_imagepyramid_get_levels); }
    GPUImage getImage(int level) {
        ff_gpuimage* h = nullptr;
        check(ff_imagepyramid_get_image(s_.get(), level, &h));
        return GPUImage::adopt(h);
    }

private:
    void make(const ff_gpuimage* i, int levels) {
        ff_stage* s = nullptr;
        check(ff_imagepyramid_create(i, levels, &s));
        reset(s);
    }
};

// FlowToColor (include/flowfilter/gpu/display.h:23-93).
class FlowToColor : public Stage {
public:
    FlowToColor() { make(nullptr, 1.0f); }
    FlowToColor(GPUImage inputFlow, const float maxflow) { make(H(inputFlow), maxflow); }
    void downloadColorFlow(image_t& colorFlow) { check(ff_flowtocolor_download_color_flow(s_.get(), &colorFlow)); }
    void setInputFlow(GPUImage f) { check(ff_flowtocolor_set_input_flow(s_.get(), H(f))); }
    GPUImage getColorFlow() { return img(ff_flowtocolor_get_color_flow); }
    float getMaxFlow() const { return getf(ff_flowtocolor_get_max_flow); }
    void setMaxFlow(const float maxflow) { check(ff_flowtocolor_set_max_flow(s_.get(), maxflow)); }

private:
    void make(const ff_gpuimage* f, float m) {
        ff_stage* s = nullptr;
        check(ff_flowtocolor_create(f, m, &s));
        reset(s);
    }
};

// FlowFilter (include/flowfilter/gpu/flowfilter.h:31-145).
class FlowFilter : public Stage {
public:
    explicit FlowFilter(GPUImage inputImage) {
        ff_stage* s = nullptr;
        check(ff_flowfilter_create_from_image(H(inputImage), &s));
        reset(s);
    }
    FlowFilter(const int height, const int width) : FlowFilter(height, width, 1, 1.0f, 1.0f) {}
    FlowFilter(const int height, const int width, const int smoothIterations, const float maxflow,
               const float gamma) {
        ff_stage* s = nullptr;
        check(ff_flowfilter_create(height, width, smoothIterations, maxflow, gamma, &s));
        reset(s);
    }
    void computeImageModel() { check(ff_flowfilter_compute_image_model(s_.get())); }
    void computePropagation() { check(ff_flowfilter_compute_propagation(s_.get())); }
    void computeUpdate() { check(ff_flowfilter_compute_update(s_.get())); }
    void setInputImage(GPUImage img_) { check(ff_flowfilter_set_input_image(s_.get(), H(img_))); }
    GPUImage getFlow() { return img(ff_flowfilter_get_flow); }
    void loadImage(image_t& image) { check(ff_flowfilter_load_image(s_.get(), &image)); }
    void downloadFlow(image_t& flow) { check(ff_flowfilter_download_flow(s_.get(), &flow)); }
    void downloadImage(image_t& image) { check(ff_flowfilter_download_image(s_.get(), &image)); }
    float getGamma() const { return getf(ff_flowfilter_get_gamma); }
    void setGamma(const float gamma) { check(ff_flowfilter_set_gamma(s_.get(), gamma)); }
    float getMaxFlow() const { return getf(ff_flowfilter_get_maxflow); }
    void setMaxFlow(const float maxflow) { check(ff_flowfilter_set_maxflow(s_.get(), maxflow)); }
    int getSmoothIterations() const { return geti(ff_flowfilter_get_smooth_iterations); }
    void setSmoothIterations(const int N) { check(ff_flowfilter_set_smooth_iterations(s_.get(), N)); }
    void setPropagationBorder(const int border) { check(ff_flowfilter_set_propagation_border(s_.get(), border)); }
    int getPropagationBorder() const { return geti(ff_flowfilter_get_propagation_border); }
    int getPropagationIterations() const { return geti(ff_flowfilter_get_propagation_iterations); }
    int height() const { return geti(ff_flowfilter_height); }
    int width() const { return geti(ff_flowfilter_width); }
    void loadImageDevice(const void* dev, std::size_t pitch) {
        check(ff_flowfilter_load_image_device(s_.get(), dev, pitch));
    }
};

// DeltaFlowFilter (include/flowfilter/gpu/flowfilter.h:148-235).
class DeltaFlowFilter : public Stage {
public:
    DeltaFlowFilter(GPUImage inputImage, GPUImage inputFlow) {
        ff_stage* s = nullptr;
        check(ff_deltaflowfilter_create(H(inputImage), H(inputFlow), &s));
        reset(s);
    }
    void computeImageModel() { check(ff_deltaflowfilter_compute_image_model(s_.get())); }
    void computePropagation() { check(ff_deltaflowfilter_compute_propagation(s_.get())); }
    void computeUpdate() { check(ff_deltaflowfilter_compute_update(s_.get())); }
    void setInputImage(GPUImage i) { check(ff_deltaflowfilter_set_input_image(s_.get(), H(i))); }
    void setInputFlow(GPUImage f) { check(ff_deltaflowfilter_set_input_flow(s_.get(), H(f))); }
    GPUImage getFlow() { return img(ff_deltaflowfilter_get_flow); }
    GPUImage getImage() { return img(ff_deltaflowfilter_get_image); }
    float getGamma() const { return getf(ff_deltaflowfilter_get_gamma); }
    void setGamma(const float gamma) { check(ff_deltaflowfilter_set_gamma(s_.get(), gamma)); }
    float getMaxFlow() const { return getf(ff_deltaflowfilter_get_maxflow); }
    void setMaxFlow(const float maxflow) { check(ff_deltaflowfilter_set_maxflow(s_.get(), maxflow)); }
    int getSmoothIterations() const { return geti(ff_deltaflowfilter_get_smooth_iterations); }
    void setSmoothIterations(const int N) { check(ff_deltaflowfilter_set_smooth_iterations(s_.get(), N)); }
    void setPropagationBorder(const int b) { check(ff_deltaflowfilter_set_propagation_border(s_.get(), b)); }
    int getPropagationBorder() const { return geti(ff_deltaflowfilter_get_propagation_border); }
    int getPropagationIterations() const { return geti(ff_deltaflowfilter_get_propagation_iterations); }
    int height() const { return geti(ff_deltaflowfilter_height); }
    int width() const { return geti(ff_deltaflowfilter_width); }
};

// PyramidalFlowFilter (include/flowfilter/gpu/flowfilter.h:238-332).
class PyramidalFlowFilter : public Stage {
public:
    PyramidalFlowFilter(const int height, const int width, const int levels) {
        ff_stage* s = nullptr;
        check(ff_pyramidal_create(height, width, levels, &s));
        reset(s);
    }
    GPUImage getFlow() { return img(ff_pyramidal_get_flow); }
    void loadImage(image_t& image) { check(ff_pyramidal_load_image(s_.get(), &image)); }
    void downloadFlow(image_t& flow) { check(ff_pyramidal_download_flow(s_.get(), &flow)); }
    void downloadImage(image_t& image) { check(ff_pyramidal_download_image(s_.get(), &image)); }
    float getGamma(const int level) const {
        float v = 0;
        check(ff_pyramidal_get_gamma(s_.get(), level, &v));
        return v;
    }
    void setGamma(const int level, const float gamma) { check(ff_pyramidal_set_gamma(s_.get(), level, gamma)); }
    void setGamma(const std::vector<float>& gamma) {
        check(ff_pyramidal_set_gamma_vector(s_.get(), gamma.data(), (int)gamma.size()));
    }
    float getMaxFlow() const { return getf(ff_pyramidal_get_maxflow); }
    void setMaxFlow(const float maxflow) { check(ff_pyramidal_set_maxflow(s_.get(), maxflow)); }
    int getSmoothIterations(const int level) const {
        int v = 0;
        check(ff_pyramidal_get_smooth_iterations(s_.get(), level, &v));
        return v;
    }
    void setSmoothIterations(const int level, const int N) {
        check(ff_pyramidal_set_smooth_iterations(s_.get(), level, N));
    }
    void setSmoothIterations(const std::vector<int>& it) {
        check(ff_pyramidal_set_smooth_iterations_vector(s_.get(), it.data(), (int)it.size()));
    }
    void setPropagationBorder(const int border) { check(ff_pyramidal_set_propagation_border(s_.get(), border)); }
    int getPropagationBorder() const { return geti(ff_pyramidal_get_propagation_border); }
    int height() const { return geti(ff_pyramidal_height); }
    int width() const { return geti(ff_pyramidal_width); }
    int levels() const { return geti(ff_pyramidal_levels); }
    void loadImageDevice(const void* dev, std::size_t pitch) {
        check(ff_pyramidal_load_image_device(s_.get(), dev, pitch));
    }
    // the next compute() reads the frame in place (ff_pyramidal_set_image_device)
    void setImageDevice(const void* dev, std::size_t pitch) {
        check(ff_pyramidal_set_image_device(s_.get(), dev, pitch));
    }
};

// Many PyramidalFlowFilters advanced by one compute() in a single fused
// launch (no reference counterpart; see ff_batch_create). The batch shares
// ownership of the filters added; frames and results go through the
// filters themselves.
class BatchedFlowFilter : public Stage {
public:
    BatchedFlowFilter() {
        ff_stage* s = nullptr;
        check(ff_batch_create(&s));
        reset(s);
    }
    void add(const PyramidalFlowFilter& f) { check(ff_batch_add(s_.get(), f.handle())); }
    int size() const { return geti(ff_batch_size); }
    void setGraphs(bool g) { check(ff_filter_set_graphs(s_.get(), g ? 1 : 0)); }
};

}  // namespace gpu
}  // namespace flowfilter

#endif  // FLOWFILTER_AMD_HPP_
