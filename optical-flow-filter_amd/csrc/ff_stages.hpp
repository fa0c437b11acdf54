// ff_stages.hpp -- the pipeline stages of flowfilter::gpu, re-implemented
// over the HIP kernels of this library.
//
// Each class keeps the reference's configure()/compute() contract, its
// setters' validation and its wiring-by-shared-buffer semantics:
//   Stage                 include/flowfilter/gpu/pipeline.h:29-78
//   ImageModel            include/flowfilter/gpu/imagemodel.h:22-91
//   FlowPropagator        include/flowfilter/gpu/propagation.h:25-105
//   FlowPropagatorPayload include/flowfilter/gpu/propagation.h:108-200
//   FlowUpdate            include/flowfilter/gpu/update.h:21-90
//   DeltaFlowUpdate       include/flowfilter/gpu/update.h:93-174
//   FlowSmoother          include/flowfilter/gpu/flowsmoothing.h:22-91
//   ImagePyramid          include/flowfilter/gpu/pyramid.h:22-89
//
// Inputs the reference reads through texture objects are bound when the
// stage is configured (the texture is created in configure()); inputs it
// reads through wrap() are taken when compute() runs. The same split is
// kept here, so re-wiring an input without configure() behaves as in the
// reference.
#pragma once

#include <vector>

#include "ff_runtime.hpp"

namespace ffamd {

class Stage {
public:
    Stage() = default;
    Stage(Stage&&) = default;
    Stage& operator=(Stage&&) = default;
    virtual ~Stage() = default;
    virtual void configure() = 0;
    virtual void compute() = 0;
    virtual const char* name() const = 0;
    virtual float elapsedTime() { return timer_.elapsed(); }
    // The stream box of a filter or batch; null for the other stages, which
    // run on the stream bound by their caller (the library stream, or the
    // stream of the filter they belong to).
    const std::shared_ptr<StreamBox>& box() const { return box_; }
    // A filter made without arguments: give it a stream of its own.
    void ensureOwnStream() {
        if (!box_) box_ = StreamBox::create();
    }

protected:
    void startTiming() { timer_.start(); }
    void stopTiming() { timer_.stop(); }
    // Filters and batches: share the bound box (a level built inside a
    // pyramid) or create a stream of their own.
    void ownStream() { box_ = stream_bound() ? current_box() : StreamBox::create(); }
    Timer timer_;
    std::shared_ptr<StreamBox> box_;
};

// ---------------------------------------------------------------- ImageModel
class ImageModel : public Stage {
public:
    ImageModel() = default;
    explicit ImageModel(const Image& input);
    void configure() override;
    void compute() override;
    const char* name() const override { return "ImageModel"; }
    void setInputImage(const Image& img);
    Image getImageConstant() const { return a0_; }
    Image getImageGradient() const { return grad_; }
    // The input image as bound at configure().
    const Image& boundImage() const { return bound_; }
    // Enqueue the kernel without touching the timer (used by the filters).
    void run();

private:
    bool configured_ = false, inputSet_ = false;
    Image input_, bound_;
    Image a0_, grad_;
};

// ------------------------------------------------------------ FlowPropagator
class FlowPropagator : public Stage {
public:
    FlowPropagator() = default;
    FlowPropagator(const Image& inputFlow, int iterations = 1);
    void configure() override;
    void compute() override;
    const char* name() const override { return "FlowPropagator"; }
    void setIterations(int n);
    int getIterations() const { return iterations_; }
    float getDt() const { return dt_; }
    void setBorder(int b);
    int getBorder() const { return border_; }
    void setInvertInputFlow(bool v) { invert_ = v; }
    bool getInvertInputFlow() const { return invert_; }
    void setInputFlow(const Image& flow);
    Image getPropagatedFlow() const { return bufY_; }
    void run();

private:
    bool configured_ = false, inputSet_ = false, invert_ = false;
    int iterations_ = 0, border_ = 3;
    float dt_ = 0.0f;
    Image input_, bound_, bufX_, bufY_;
};

// ----------------------------------------------------- FlowPropagatorPayload
class FlowPropagatorPayload : public Stage {
public:
    FlowPropagatorPayload() = default;
    FlowPropagatorPayload(const Image& flow, const Image& scalar, const Image& vector, int iterations = 1);
    void configure() override;
    void compute() override;
    const char* name() const override { return "FlowPropagatorPayload"; }
    void setIterations(int n);
    int getIterations() const { return iterations_; }
    float getDt() const { return dt_; }
    void setBorder(int b);
    int getBorder() const { return border_; }
    void setInputFlow(const Image& flow);
    void setScalarPayload(const Image& img);
    void setVectorPayload(const Image& img);
    Image getPropagatedFlow() const { return flowY_; }
    Image getPropagatedScalar() const { return scalarY_; }
    Image getPropagatedVector() const { return vectorY_; }
    void run();

private:
    bool configured_ = false, flowSet_ = false, scalarSet_ = false, vectorSet_ = false;
    int iterations_ = 0, border_ = 3;
    float dt_ = 0.0f;
    Image flow_, scalar_, vector_;
    Image bFlow_, bScalar_, bVector_;
    Image flowX_, flowY_, scalarX_, scalarY_, vectorX_, vectorY_;
};

// ---------------------------------------------------------------- FlowUpdate
class FlowUpdate : public Stage {
public:
    FlowUpdate() = default;
    FlowUpdate(const Image& flow, const Image& image, const Image& gradient, float gamma = 1.0f,
               float maxflow = 1.0f);
    void configure() override;
    void compute() override;
    const char* name() const override { return "FlowUpdate"; }
    float getGamma() const { return gamma_; }
    void setGamma(float g);
    float getMaxFlow() const { return maxflow_; }
    void setMaxFlow(float m) { maxflow_ = m; }
    void setInputFlow(const Image& img);
    void setInputImage(const Image& img);
    void setInputImageGradient(const Image& img);
    Image getInputFlow() const { return flow_; }
    Image getUpdatedFlow() const { return flowUpd_; }
    Image getUpdatedImage() const { return imageUpd_; }
    void run();

private:
    bool configured_ = false, flowSet_ = false, imageSet_ = false, gradSet_ = false;
    float gamma_ = 1.0f, maxflow_ = 1.0f;
    Image flow_, image_, grad_;
    Image flowUpd_, imageUpd_;
};

// ----------------------------------------------------------- DeltaFlowUpdate
class DeltaFlowUpdate : public Stage {
public:
    DeltaFlowUpdate() = default;
    DeltaFlowUpdate(const Image& flow, const Image& deltaFlow, const Image& imageOld, const Image& image,
                    const Image& gradient, float gamma = 1.0f, float maxflow = 1.0f);
    void configure() override;
    void compute() override;
    const char* name() const override { return "DeltaFlowUpdate"; }
    float getGamma() const { return gamma_; }
    void setGamma(float g);
    float getMaxFlow() const { return maxflow_; }
    void setMaxFlow(float m) { maxflow_ = m; }
    void setInputFlow(const Image& img);
    void setInputDeltaFlow(const Image& img);
    void setInputImageOld(const Image& img);
    void setInputImage(const Image& img);
    void setInputImageGradient(const Image& img);
    Image getUpdatedFlow() const { return flowUpd_; }
    Image getUpdatedDeltaFlow() const { return dflowUpd_; }
    Image getUpdatedImage() const { return imageUpd_; }
    // Upper-level flow as bound at configure(), with its per-column and
    // per-row texel index maps (used by the fused delta-level kernel too).
    const Image& boundUpperFlow() const { return bFlow_; }
    const int* xmap() const { return (const int*)xmap_.data(); }
    const int* ymap() const { return (const int*)ymap_.data(); }
    void run();

private:
    bool configured_ = false, flowSet_ = false, dflowSet_ = false, oldSet_ = false, imageSet_ = false,
         gradSet_ = false;
    float gamma_ = 1.0f, maxflow_ = 1.0f;
    Image flow_, dflow_, imageOld_, image_, grad_;
    Image bFlow_;
    DeviceArray xmap_, ymap_;
    Image flowUpd_, dflowUpd_, imageUpd_;
};

// -------------------------------------------------------------- FlowSmoother
class FlowSmoother : public Stage {
public:
    FlowSmoother() = default;
    FlowSmoother(const Image& flow, int iterations);
    void configure() override;
    void compute() override;
    const char* name() const override { return "FlowSmoother"; }
    int getIterations() const { return iterations_; }
    void setIterations(int n);
    void setInputFlow(const Image& img);
    Image getSmoothedFlow() const { return bufY_; }
    // The smoothed-flow image itself (for buffer exchange by the filters).
    Image& smoothedFlowRef() { return bufY_; }
    void run();

private:
    bool configured_ = false, inputSet_ = false;
    int iterations_ = 0;
    Image input_, bound_, bufX_, bufY_;
};

// -------------------------------------------------------------- ImagePyramid
class ImagePyramid : public Stage {
public:
    ImagePyramid() = default;
    ImagePyramid(const Image& image, int levels);
    void configure() override;
    void compute() override;
    const char* name() const override { return "ImagePyramid"; }
    void setInputImage(const Image& img);
    void setLevels(int levels);
    int getLevels() const { return levels_; }
    Image getImage(int level) const;
    void run();

private:
    bool configured_ = false, inputSet_ = false;
    int levels_ = 0;
    Image input_, bound_;
    std::vector<Image> px_, py_;
};

// --------------------------------------------------------------- FlowToColor
// include/flowfilter/gpu/display.h:23-93: RGBA colour coding of a flow field
// (the demos' display of the filter's output).
class FlowToColor : public Stage {
public:
    FlowToColor() = default;
    FlowToColor(const Image& inputFlow, float maxflow);
    void configure() override;
    void compute() override;
    const char* name() const override { return "FlowToColor"; }
    void setInputFlow(const Image& flow);
    Image getColorFlow() const { return color_; }
    float getMaxFlow() const { return maxflow_; }
    void setMaxFlow(float m);
    void downloadColorFlow(ff_image_t& img) const { color_.download(img); }
    void run();

private:
    bool configured_ = false, inputSet_ = false;
    float maxflow_ = 1.0f;
    Image input_, wheel_, color_;
};

// ----------------------------------------------------- LaxWendroffPropagator
// include/flowfilter/gpu/propagation.h:218-291: N iterations of a Y and an X
// Lax-Wendroff pass (dt = 1/N) of a float image (1 or 4 channels) along a
// flow field; 0 iterations (never set) runs one pair with dt = 0.
class LaxWendroffPropagator : public Stage {
public:
    LaxWendroffPropagator() = default;
    LaxWendroffPropagator(const Image& inputFlow, const Image& inputImage);
    void configure() override;
    void compute() override;
    const char* name() const override { return "LaxWendroffPropagator"; }
    void setIterations(int n);
    int getIterations() const { return iterations_; }
    float getDt() const { return dt_; }
    void setInputFlow(const Image& flow);
    void setInputImage(const Image& img);
    Image getFlow() const { return flow_; }
    Image getPropagatedImage() const { return bufX_; }
    void run();

private:
    bool configured_ = false, flowSet_ = false, imageSet_ = false;
    int iterations_ = 0;
    float dt_ = 0.0f;
    Image flow_, image_, bufX_, bufY_;
};

// ---------------------------------------------- RotationalFlowImagePredictor
// include/flowfilter/gpu/rotation.h:24-88: the optical flow of a rotating
// camera and the image it predicts (the input image propagated along that
// flow with LaxWendroffPropagator).
class RotationalFlowImagePredictor : public Stage {
public:
    RotationalFlowImagePredictor() = default;
    explicit RotationalFlowImagePredictor(const Camera& cam) { setCamera(cam); }
    RotationalFlowImagePredictor(const Camera& cam, const Image& inputImage);
    void configure() override;
    void compute() override;
    const char* name() const override { return "RotationalFlowImagePredictor"; }
    void setInputImage(const Image& img);
    Image getPredictedImage() const { return propagator_.getPropagatedImage(); }
    Image getOpticalFlow() const { return flow_; }
    void setCamera(const Camera& cam) { camera_ = cam; }
    const Camera& getCamera() const { return camera_; }
    void setAngularVelocity(float wx, float wy, float wz);
    void setIterations(int n) { propagator_.setIterations(n); }
    int getIterations() const { return propagator_.getIterations(); }

private:
    bool configured_ = false, imageSet_ = false;
    Camera camera_{0.0f, 0.0f, 0.0f, 0.0f};
    float w_[3] = {0.0f, 0.0f, 0.0f};
    Image image_, flow_;
    LaxWendroffPropagator propagator_;
};

// createPerspectiveCamera (src/gpu/camera.cu:14-29).
Camera createPerspectiveCamera(float focalLength, int height, int width, float sensorHeight, float sensorWidth);

// The colour wheel as a host image descriptor (getColorWheelRGBA,
// src/colorwheel.cpp:21862-21873): 256 x 256 x 4 u8, read-only data of the
// library.
ff_image_t colorWheelRGBA();

// Upper-level texel index of every output column (or row) for the
// normalized-coordinate point sampling of deltaFlowUpdate_k
// (update_k.cu:151-155): i = min(floor(fl(x/(n-1)) * nu), nu-1).
std::vector<int> upper_index_map(int n, int nu);

}  // namespace ffamd
