// ff_capi.cpp -- extern "C" entry points declared in
// include/flowfilter_amd.h. Each one validates its handles, calls the C++
// object and converts exceptions into a status code plus a thread-local
// message; nothing propagates across the C boundary.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/flowfilter_amd.h"
#include "ff_filters.hpp"

using namespace ffamd;

struct ff_gpuimage {
    Image img;
};
struct ff_stage {
    std::shared_ptr<Stage> obj;  // shared with batches that hold the filter
};

namespace {

thread_local std::string t_error;

ff_status fail(int code, const std::string& msg) {
    t_error = msg;
    return code;
}

template <class F>
ff_status guard(F&& f) {
    try {
        f();
        return FF_OK;
    } catch (const Error& e) {
        return fail(e.code(), e.what());
    } catch (const std::bad_alloc& e) {
        return fail(FF_ERR_BAD_ALLOC, e.what());
    } catch (const std::exception& e) {
        return fail(FF_ERR_RUNTIME, e.what());
    } catch (...) {
        return fail(FF_ERR_RUNTIME, "unknown C++ exception");
    }
}

// guard() with the stream of stage `s` (a filter or batch) bound to the
// calling thread for the call.
template <class F>
ff_status guard_on(const ff_stage* s, F&& f) {
    return guard([&] {
        StreamScope sc(s && s->obj ? s->obj->box() : nullptr);
        f();
    });
}

Error null_handle(const char* what) { return Error(FF_ERR_NULL_HANDLE, std::string("null ") + what); }

const Image& img_of(const ff_gpuimage* h) {
    if (!h) throw null_handle("ff_gpuimage");
    return h->img;
}
Image& img_of(ff_gpuimage* h) {
    if (!h) throw null_handle("ff_gpuimage");
    return h->img;
}

template <class T>
T& as(ff_stage* s, const char* type) {
    if (!s || !s->obj) throw null_handle("ff_stage");
    T* p = dynamic_cast<T*>(s->obj.get());
    if (!p)
        throw Error(FF_ERR_WRONG_TYPE,
                    std::string("stage is a ") + s->obj->name() + ", expected a " + type);
    return *p;
}

template <class T>
void out_ptr(T* p, const char* what) {
    if (!p) throw Error(FF_ERR_INVALID_ARGUMENT, std::string("null output pointer: ") + what);
}

void new_image(const Image& img, ff_gpuimage** out) {
    out_ptr(out, "ff_gpuimage**");
    *out = new ff_gpuimage{img};
}

template <class F>
void on_fused_filter(ff_stage* s, F f) {
    if (!s || !s->obj) throw null_handle("ff_stage");
    Stage* o = s->obj.get();
    if (auto* a = dynamic_cast<FlowFilter*>(o)) f(*a);
    else if (auto* b = dynamic_cast<DeltaFlowFilter*>(o)) f(*b);
    else if (auto* c = dynamic_cast<PyramidalFlowFilter*>(o)) f(*c);
    else if (auto* d = dynamic_cast<BatchedFlowFilter*>(o)) f(*d);
    else throw Error(FF_ERR_WRONG_TYPE, std::string("stage is a ") + o->name() + ", expected a filter");
}

template <class T, class... A>
void new_stage(ff_stage** out, A&&... args) {
    out_ptr(out, "ff_stage**");
    std::shared_ptr<Stage> obj = std::make_shared<T>(std::forward<A>(args)...);
    *out = new ff_stage{std::move(obj)};
}

const ff_image_t& host_img(const ff_image_t* p) {
    if (!p) throw Error(FF_ERR_INVALID_ARGUMENT, "null ff_image_t");
    return *p;
}
ff_image_t& host_img(ff_image_t* p) {
    if (!p) throw Error(FF_ERR_INVALID_ARGUMENT, "null ff_image_t");
    return *p;
}

}  // namespace

extern "C" {

// ----------------------------------------------------------------- library
const char* ff_last_error(void) { return t_error.c_str(); }

ff_status ff_version(int* major, int* minor) {
    return guard([&] {
        out_ptr(major, "major");
        out_ptr(minor, "minor");
        *major = 1;
        *minor = 0;
    });
}

ff_status ff_synchronize(void) { return guard([] { synchronize(); }); }
ff_status ff_set_stream(void* s) { return guard([&] { set_stream((hipStream_t)s); }); }
ff_status ff_get_stream(void** s) {
    return guard([&] {
        out_ptr(s, "stream");
        *s = (void*)stream();
    });
}

// ------------------------------------------------------------- host images
// createImage (src/image.cpp:14-40): pitch = width * depth * pixelSize.
ff_status ff_create_image(int height, int width, int depth, size_t pixelSize, ff_image_t* out) {
    return guard([&] {
        out_ptr(out, "ff_image_t*");
        if (height <= 0 || width <= 0 || depth <= 0 || pixelSize == 0)
            throw Error(FF_ERR_RUNTIME, "createImage(): dimensions must be positive");
        const size_t pitch = (size_t)width * depth * pixelSize;
        void* buf = std::calloc((size_t)height, pitch);
        if (!buf) throw std::bad_alloc();
        out->height = height;
        out->width = width;
        out->depth = depth;
        out->pitch = pitch;
        out->itemSize = pixelSize;
        out->data = buf;
    });
}

// destroyImage (src/image.cpp:43-52); frees what ff_create_image allocated.
ff_status ff_destroy_image(ff_image_t* image) {
    return guard([&] {
        out_ptr(image, "ff_image_t*");
        std::free(image->data);
        image->data = nullptr;
    });
}

// ---------------------------------------------------------------- GPUImage
ff_status ff_gpuimage_create(int h, int w, int depth, int itemSize, ff_gpuimage** out) {
    return guard([&] { new_image(Image(h, w, depth, itemSize), out); });
}
ff_status ff_gpuimage_create_empty(ff_gpuimage** out) {
    return guard([&] { new_image(Image(), out); });
}
ff_status ff_gpuimage_share(const ff_gpuimage* img, ff_gpuimage** out) {
    return guard([&] { new_image(img_of(img), out); });
}
ff_status ff_gpuimage_destroy(ff_gpuimage* img) {
    return guard([&] { delete img; });
}
#define GETI(NAME, EXPR)                                         \
    ff_status ff_gpuimage_##NAME(const ff_gpuimage* img, int* out) { \
        return guard([&] {                                       \
            out_ptr(out, #NAME);                                 \
            *out = (int)(img_of(img).EXPR);                      \
        });                                                      \
    }
GETI(height, height())
GETI(width, width())
GETI(depth, depth())
GETI(pitch, pitch())
GETI(item_size, itemSize())
#undef GETI
ff_status ff_gpuimage_data(const ff_gpuimage* img, void** out) {
    return guard([&] {
        out_ptr(out, "data");
        *out = img_of(img).data();
    });
}
ff_status ff_gpuimage_upload(ff_gpuimage* img, const ff_image_t* host) {
    return guard([&] { img_of(img).upload(host_img(host)); });
}
ff_status ff_gpuimage_download(const ff_gpuimage* img, ff_image_t* host) {
    return guard([&] { img_of(img).download(host_img(host)); });
}
ff_status ff_gpuimage_copy_from(ff_gpuimage* img, const ff_gpuimage* src) {
    return guard([&] { img_of(img).copyFrom(img_of(src)); });
}
ff_status ff_gpuimage_clear(ff_gpuimage* img) {
    return guard([&] { img_of(img).clear(); });
}
ff_status ff_gpuimage_upload_device(ff_gpuimage* img, const void* dev, size_t pitch) {
    return guard([&] { img_of(img).upload_device(dev, pitch); });
}
ff_status ff_gpuimage_download_device(const ff_gpuimage* img, void* dev, size_t pitch) {
    return guard([&] { img_of(img).download_device(dev, pitch); });
}

// ------------------------------------------------------------------ DLPack
// The DLPack exchange structs (dmlc/dlpack, the ABI of v0.8 that consumers
// read from a "dltensor" capsule).
namespace {
struct DLDevice {
    int32_t device_type;
    int32_t device_id;
};
struct DLDataType {
    uint8_t code;
    uint8_t bits;
    uint16_t lanes;
};
struct DLTensor {
    void* data;
    DLDevice device;
    int32_t ndim;
    DLDataType dtype;
    int64_t* shape;
    int64_t* strides;
    uint64_t byte_offset;
};
struct DLManagedTensor {
    DLTensor dl_tensor;
    void* manager_ctx;
    void (*deleter)(DLManagedTensor*);
};
constexpr int32_t kDLROCM = 10;
constexpr uint8_t kDLUInt = 1, kDLFloat = 2;

// An exported view: the managed tensor, its shape and strides, and a share
// of the image (keeps the buffer alive until the consumer lets go).
struct DLExport {
    DLManagedTensor m;
    int64_t shape[3];
    int64_t strides[3];
    Image img;
};
void dl_delete(DLManagedTensor* m) {
    if (m) delete static_cast<DLExport*>(m->manager_ctx);
}
}  // namespace

ff_status ff_gpuimage_to_dlpack(const ff_gpuimage* img, void* consumer_stream, int sync, void** out) {
    return guard([&] {
        out_ptr(out, "managed tensor");
        const Image& im = img_of(img);
        if (!im.allocated()) throw logic_error("GPUImage to DLPack: unallocated image");
        if (im.itemSize() != 1 && im.itemSize() != 4)
            throw invalid_argument("GPUImage to DLPack: item size must be 1 (uint8) or 4 (float32)");
        if (im.pitch() % (size_t)im.itemSize())
            throw invalid_argument("GPUImage to DLPack: pitch is not a multiple of the item size");
        if (sync) {
            // the consumer's stream waits for the work enqueued on the image
            hipStream_t home = im.homeStream();
            if ((hipStream_t)consumer_stream != home) {
                hipEvent_t e = im.home()->event();
                FF_HIP(hipEventRecord(e, home));
                FF_HIP(hipStreamWaitEvent((hipStream_t)consumer_stream, e, 0));
            }
        }
        int dev = 0;
        FF_HIP(hipGetDevice(&dev));
        auto* x = new DLExport();
        x->img = im;
        const int nd = im.depth() > 1 ? 3 : 2;
        x->shape[0] = im.height();
        x->shape[1] = im.width();
        x->shape[2] = im.depth();
        x->strides[0] = (int64_t)(im.pitch() / (size_t)im.itemSize());
        x->strides[1] = im.depth();
        x->strides[2] = 1;
        DLTensor& t = x->m.dl_tensor;
        t.data = im.data();
        t.device = DLDevice{kDLROCM, dev};
        t.ndim = nd;
        t.dtype = im.itemSize() == 1 ? DLDataType{kDLUInt, 8, 1} : DLDataType{kDLFloat, 32, 1};
        t.shape = x->shape;
        t.strides = x->strides;
        t.byte_offset = 0;
        x->m.manager_ctx = x;
        x->m.deleter = dl_delete;
        *out = &x->m;
    });
}

ff_status ff_dlpack_delete(void* managed) {
    return guard([&] {
        auto* m = static_cast<DLManagedTensor*>(managed);
        if (m && m->deleter) m->deleter(m);
    });
}

// ------------------------------------------------------------------- Stage
ff_status ff_stage_configure(ff_stage* s) {
    return guard_on(s, [&] { as<Stage>(s, "Stage").configure(); });
}
ff_status ff_stage_compute(ff_stage* s) {
    return guard_on(s, [&] { as<Stage>(s, "Stage").compute(); });
}
ff_status ff_stage_elapsed_time(ff_stage* s, float* ms) {
    return guard_on(s, [&] {
        out_ptr(ms, "ms");
        *ms = as<Stage>(s, "Stage").elapsedTime();
    });
}
ff_status ff_stage_get_stream(ff_stage* s, void** out) {
    return guard_on(s, [&] {
        out_ptr(out, "stream");
        if (!s || !s->obj) throw null_handle("ff_stage");
        *out = (void*)stream();
    });
}
ff_status ff_stage_set_stream(ff_stage* s, void* hip_stream) {
    return guard([&] {
        if (!s || !s->obj) throw null_handle("ff_stage");
        if (!s->obj->box())
            throw Error(FF_ERR_WRONG_TYPE, std::string("stage is a ") + s->obj->name() +
                                               ", expected a filter or a batch (only they have streams of their own)");
        s->obj->box()->use((hipStream_t)hip_stream);
    });
}
ff_status ff_stage_destroy(ff_stage* s) {
    return guard([&] { delete s; });
}

// -------------------------------------------------------- helper macros
#define SETTER(FN, CLS, TYPE, CALL)                                  \
    ff_status FN(ff_stage* s, TYPE v) {                              \
        return guard_on(s, [&] { as<CLS>(s, #CLS).CALL(v); });             \
    }
#define GETTER(FN, CLS, TYPE, CALL)                                  \
    ff_status FN(ff_stage* s, TYPE* out) {                           \
        return guard_on(s, [&] {                                           \
            out_ptr(out, #FN);                                       \
            *out = (TYPE)as<CLS>(s, #CLS).CALL();                    \
        });                                                          \
    }
#define IMG_SETTER(FN, CLS, CALL)                                    \
    ff_status FN(ff_stage* s, const ff_gpuimage* img) {              \
        return guard_on(s, [&] { as<CLS>(s, #CLS).CALL(img_of(img)); });   \
    }
#define IMG_GETTER(FN, CLS, CALL)                                    \
    ff_status FN(ff_stage* s, ff_gpuimage** out) {                   \
        return guard_on(s, [&] { new_image(as<CLS>(s, #CLS).CALL(), out); }); \
    }
#define ACTION(FN, CLS, CALL)                                        \
    ff_status FN(ff_stage* s) {                                      \
        return guard_on(s, [&] { as<CLS>(s, #CLS).CALL(); });              \
    }

// -------------------------------------------------------------- ImageModel
ff_status ff_imagemodel_create(const ff_gpuimage* input, ff_stage** out) {
    return guard([&] {
        if (input) new_stage<ImageModel>(out, input->img);
        else new_stage<ImageModel>(out);
    });
}
IMG_SETTER(ff_imagemodel_set_input_image, ImageModel, setInputImage)
IMG_GETTER(ff_imagemodel_get_image_constant, ImageModel, getImageConstant)
IMG_GETTER(ff_imagemodel_get_image_gradient, ImageModel, getImageGradient)

// ---------------------------------------------------------- FlowPropagator
ff_status ff_flowpropagator_create(const ff_gpuimage* inputFlow, int iterations, ff_stage** out) {
    return guard([&] {
        if (inputFlow) new_stage<FlowPropagator>(out, inputFlow->img, iterations);
        else new_stage<FlowPropagator>(out);
    });
}
SETTER(ff_flowpropagator_set_iterations, FlowPropagator, int, setIterations)
GETTER(ff_flowpropagator_get_iterations, FlowPropagator, int, getIterations)
GETTER(ff_flowpropagator_get_dt, FlowPropagator, float, getDt)
SETTER(ff_flowpropagator_set_border, FlowPropagator, int, setBorder)
GETTER(ff_flowpropagator_get_border, FlowPropagator, int, getBorder)
ff_status ff_flowpropagator_set_invert_input_flow(ff_stage* s, int invert) {
    return guard_on(s, [&] { as<FlowPropagator>(s, "FlowPropagator").setInvertInputFlow(invert != 0); });
}
GETTER(ff_flowpropagator_get_invert_input_flow, FlowPropagator, int, getInvertInputFlow)
IMG_SETTER(ff_flowpropagator_set_input_flow, FlowPropagator, setInputFlow)
IMG_GETTER(ff_flowpropagator_get_propagated_flow, FlowPropagator, getPropagatedFlow)

// --------------------------------------------------- FlowPropagatorPayload
ff_status ff_flowpropagatorpayload_create(const ff_gpuimage* flow, const ff_gpuimage* scalar,
                                          const ff_gpuimage* vector, int iterations, ff_stage** out) {
    return guard([&] {
        if (flow && scalar && vector)
            new_stage<FlowPropagatorPayload>(out, flow->img, scalar->img, vector->img, iterations);
        else new_stage<FlowPropagatorPayload>(out);
    });
}
SETTER(ff_flowpropagatorpayload_set_iterations, FlowPropagatorPayload, int, setIterations)
GETTER(ff_flowpropagatorpayload_get_iterations, FlowPropagatorPayload, int, getIterations)
GETTER(ff_flowpropagatorpayload_get_dt, FlowPropagatorPayload, float, getDt)
SETTER(ff_flowpropagatorpayload_set_border, FlowPropagatorPayload, int, setBorder)
GETTER(ff_flowpropagatorpayload_get_border, FlowPropagatorPayload, int, getBorder)
IMG_SETTER(ff_flowpropagatorpayload_set_input_flow, FlowPropagatorPayload, setInputFlow)
IMG_SETTER(ff_flowpropagatorpayload_set_scalar_payload, FlowPropagatorPayload, setScalarPayload)
IMG_SETTER(ff_flowpropagatorpayload_set_vector_payload, FlowPropagatorPayload, setVectorPayload)
IMG_GETTER(ff_flowpropagatorpayload_get_propagated_flow, FlowPropagatorPayload, getPropagatedFlow)
IMG_GETTER(ff_flowpropagatorpayload_get_propagated_scalar, FlowPropagatorPayload, getPropagatedScalar)
IMG_GETTER(ff_flowpropagatorpayload_get_propagated_vector, FlowPropagatorPayload, getPropagatedVector)

// -------------------------------------------------------------- FlowUpdate
ff_status ff_flowupdate_create(const ff_gpuimage* flow, const ff_gpuimage* image, const ff_gpuimage* grad,
                               float gamma, float maxflow, ff_stage** out) {
    return guard([&] {
        if (flow && image && grad) {
            new_stage<FlowUpdate>(out, flow->img, image->img, grad->img, gamma, maxflow);
        } else {
            // FlowUpdate() then the parameter setters (python/flowfilter/gpu/update.pyx:14-30)
            auto* u = new FlowUpdate();
            std::unique_ptr<Stage> keep(u);
            u->setGamma(gamma);
            u->setMaxFlow(maxflow);
            out_ptr(out, "ff_stage**");
            *out = new ff_stage{std::move(keep)};
        }
    });
}
GETTER(ff_flowupdate_get_gamma, FlowUpdate, float, getGamma)
SETTER(ff_flowupdate_set_gamma, FlowUpdate, float, setGamma)
GETTER(ff_flowupdate_get_maxflow, FlowUpdate, float, getMaxFlow)
SETTER(ff_flowupdate_set_maxflow, FlowUpdate, float, setMaxFlow)
IMG_SETTER(ff_flowupdate_set_input_flow, FlowUpdate, setInputFlow)
IMG_SETTER(ff_flowupdate_set_input_image, FlowUpdate, setInputImage)
IMG_SETTER(ff_flowupdate_set_input_image_gradient, FlowUpdate, setInputImageGradient)
IMG_GETTER(ff_flowupdate_get_updated_flow, FlowUpdate, getUpdatedFlow)
IMG_GETTER(ff_flowupdate_get_updated_image, FlowUpdate, getUpdatedImage)

// --------------------------------------------------------- DeltaFlowUpdate
ff_status ff_deltaflowupdate_create(const ff_gpuimage* flow, const ff_gpuimage* dflow, const ff_gpuimage* old,
                                    const ff_gpuimage* image, const ff_gpuimage* grad, float gamma, float maxflow,
                                    ff_stage** out) {
    return guard([&] {
        if (flow && dflow && old && image && grad) {
            new_stage<DeltaFlowUpdate>(out, flow->img, dflow->img, old->img, image->img, grad->img, gamma, maxflow);
        } else {
            auto* u = new DeltaFlowUpdate();
            std::unique_ptr<Stage> keep(u);
            u->setGamma(gamma);
            u->setMaxFlow(maxflow);
            out_ptr(out, "ff_stage**");
            *out = new ff_stage{std::move(keep)};
        }
    });
}
GETTER(ff_deltaflowupdate_get_gamma, DeltaFlowUpdate, float, getGamma)
SETTER(ff_deltaflowupdate_set_gamma, DeltaFlowUpdate, float, setGamma)
GETTER(ff_deltaflowupdate_get_maxflow, DeltaFlowUpdate, float, getMaxFlow)
SETTER(ff_deltaflowupdate_set_maxflow, DeltaFlowUpdate, float, setMaxFlow)
IMG_SETTER(ff_deltaflowupdate_set_input_flow, DeltaFlowUpdate, setInputFlow)
IMG_SETTER(ff_deltaflowupdate_set_input_delta_flow, DeltaFlowUpdate, setInputDeltaFlow)
IMG_SETTER(ff_deltaflowupdate_set_input_image_old, DeltaFlowUpdate, setInputImageOld)
IMG_SETTER(ff_deltaflowupdate_set_input_image, DeltaFlowUpdate, setInputImage)
IMG_SETTER(ff_deltaflowupdate_set_input_image_gradient, DeltaFlowUpdate, setInputImageGradient)
IMG_GETTER(ff_deltaflowupdate_get_updated_flow, DeltaFlowUpdate, getUpdatedFlow)
IMG_GETTER(ff_deltaflowupdate_get_updated_delta_flow, DeltaFlowUpdate, getUpdatedDeltaFlow)
IMG_GETTER(ff_deltaflowupdate_get_updated_image, DeltaFlowUpdate, getUpdatedImage)

// ------------------------------------------------------------ FlowSmoother
ff_status ff_flowsmoother_create(const ff_gpuimage* flow, int iterations, ff_stage** out) {
    return guard([&] {
        if (flow) new_stage<FlowSmoother>(out, flow->img, iterations);
        else new_stage<FlowSmoother>(out);
    });
}
GETTER(ff_flowsmoother_get_iterations, FlowSmoother, int, getIterations)
SETTER(ff_flowsmoother_set_iterations, FlowSmoother, int, setIterations)
IMG_SETTER(ff_flowsmoother_set_input_flow, FlowSmoother, setInputFlow)
IMG_GETTER(ff_flowsmoother_get_smoothed_flow, FlowSmoother, getSmoothedFlow)

// ------------------------------------------------------------ ImagePyramid
ff_status ff_imagepyramid_create(const ff_gpuimage* image, int levels, ff_stage** out) {
    return guard([&] {
        if (image) new_stage<ImagePyramid>(out, image->img, levels);
        else {
            auto* p = new ImagePyramid();
            std::unique_ptr<Stage> keep(p);
            p->setLevels(levels);
            out_ptr(out, "ff_stage**");
            *out = new ff_stage{std::move(keep)};
        }
    });
}
IMG_SETTER(ff_imagepyramid_set_input_image, ImagePyramid, setInputImage)
SETTER(ff_imagepyramid_set_levels, ImagePyramid, int, setLevels)
GETTER(ff_imagepyramid_get_levels, ImagePyramid, int, getLevels)
ff_status ff_imagepyramid_get_image(ff_stage* s, int level, ff_gpuimage** out) {
    return guard_on(s, [&] { new_image(as<ImagePyramid>(s, "ImagePyramid").getImage(level), out); });
}

// -------------------------------------------------------------- FlowFilter
ff_status ff_flowfilter_create(int h, int w, int smooth, float maxflow, float gamma, ff_stage** out) {
    return guard([&] { new_stage<FlowFilter>(out, h, w, smooth, maxflow, gamma); });
}
ff_status ff_flowfilter_create_from_image(const ff_gpuimage* input, ff_stage** out) {
    return guard([&] { new_stage<FlowFilter>(out, img_of(input)); });
}
ACTION(ff_flowfilter_compute_image_model, FlowFilter, computeImageModel)
ACTION(ff_flowfilter_compute_propagation, FlowFilter, computePropagation)
ACTION(ff_flowfilter_compute_update, FlowFilter, computeUpdate)
IMG_SETTER(ff_flowfilter_set_input_image, FlowFilter, setInputImage)
ff_status ff_flowfilter_load_image(ff_stage* s, const ff_image_t* image) {
    return guard_on(s, [&] { as<FlowFilter>(s, "FlowFilter").loadImage(host_img(image)); });
}
ff_status ff_flowfilter_download_flow(ff_stage* s, ff_image_t* flow) {
    return guard_on(s, [&] { as<FlowFilter>(s, "FlowFilter").downloadFlow(host_img(flow)); });
}
ff_status ff_flowfilter_download_image(ff_stage* s, ff_image_t* image) {
    return guard_on(s, [&] { as<FlowFilter>(s, "FlowFilter").downloadImage(host_img(image)); });
}
IMG_GETTER(ff_flowfilter_get_flow, FlowFilter, getFlow)
GETTER(ff_flowfilter_get_gamma, FlowFilter, float, getGamma)
SETTER(ff_flowfilter_set_gamma, FlowFilter, float, setGamma)
GETTER(ff_flowfilter_get_maxflow, FlowFilter, float, getMaxFlow)
SETTER(ff_flowfilter_set_maxflow, FlowFilter, float, setMaxFlow)
GETTER(ff_flowfilter_get_smooth_iterations, FlowFilter, int, getSmoothIterations)
SETTER(ff_flowfilter_set_smooth_iterations, FlowFilter, int, setSmoothIterations)
SETTER(ff_flowfilter_set_propagation_border, FlowFilter, int, setPropagationBorder)
GETTER(ff_flowfilter_get_propagation_border, FlowFilter, int, getPropagationBorder)
GETTER(ff_flowfilter_get_propagation_iterations, FlowFilter, int, getPropagationIterations)
GETTER(ff_flowfilter_height, FlowFilter, int, height)
GETTER(ff_flowfilter_width, FlowFilter, int, width)
ff_status ff_flowfilter_load_image_device(ff_stage* s, const void* dev, size_t pitch) {
    return guard_on(s, [&] { as<FlowFilter>(s, "FlowFilter").loadImageDevice(dev, pitch); });
}

// --------------------------------------------------------- DeltaFlowFilter
ff_status ff_deltaflowfilter_create(const ff_gpuimage* image, const ff_gpuimage* flow, ff_stage** out) {
    return guard([&] {
        if (image && flow) {
            new_stage<DeltaFlowFilter>(out, image->img, flow->img);
        } else {
            new_stage<DeltaFlowFilter>(out);
            (*out)->obj->ensureOwnStream();
        }
    });
}
ACTION(ff_deltaflowfilter_compute_image_model, DeltaFlowFilter, computeImageModel)
ACTION(ff_deltaflowfilter_compute_propagation, DeltaFlowFilter, computePropagation)
ACTION(ff_deltaflowfilter_compute_update, DeltaFlowFilter, computeUpdate)
IMG_SETTER(ff_deltaflowfilter_set_input_image, DeltaFlowFilter, setInputImage)
IMG_SETTER(ff_deltaflowfilter_set_input_flow, DeltaFlowFilter, setInputFlow)
IMG_GETTER(ff_deltaflowfilter_get_flow, DeltaFlowFilter, getFlow)
IMG_GETTER(ff_deltaflowfilter_get_image, DeltaFlowFilter, getImage)
GETTER(ff_deltaflowfilter_get_gamma, DeltaFlowFilter, float, getGamma)
SETTER(ff_deltaflowfilter_set_gamma, DeltaFlowFilter, float, setGamma)
GETTER(ff_deltaflowfilter_get_maxflow, DeltaFlowFilter, float, getMaxFlow)
SETTER(ff_deltaflowfilter_set_maxflow, DeltaFlowFilter, float, setMaxFlow)
GETTER(ff_deltaflowfilter_get_smooth_iterations, DeltaFlowFilter, int, getSmoothIterations)
SETTER(ff_deltaflowfilter_set_smooth_iterations, DeltaFlowFilter, int, setSmoothIterations)
SETTER(ff_deltaflowfilter_set_propagation_border, DeltaFlowFilter, int, setPropagationBorder)
GETTER(ff_deltaflowfilter_get_propagation_border, DeltaFlowFilter, int, getPropagationBorder)
GETTER(ff_deltaflowfilter_get_propagation_iterations, DeltaFlowFilter, int, getPropagationIterations)
GETTER(ff_deltaflowfilter_height, DeltaFlowFilter, int, height)
GETTER(ff_deltaflowfilter_width, DeltaFlowFilter, int, width)

// ----------------------------------------------------- PyramidalFlowFilter
ff_status ff_pyramidal_create(int h, int w, int levels, ff_stage** out) {
    return guard([&] { new_stage<PyramidalFlowFilter>(out, h, w, levels); });
}
IMG_GETTER(ff_pyramidal_get_flow, PyramidalFlowFilter, getFlow)
ff_status ff_pyramidal_load_image(ff_stage* s, const ff_image_t* image) {
    return guard_on(s, [&] { as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").loadImage(host_img(image)); });
}
ff_status ff_pyramidal_download_flow(ff_stage* s, ff_image_t* flow) {
    return guard_on(s, [&] { as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").downloadFlow(host_img(flow)); });
}
ff_status ff_pyramidal_download_image(ff_stage* s, ff_image_t* image) {
    return guard_on(s, [&] { as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").downloadImage(host_img(image)); });
}
ff_status ff_pyramidal_get_gamma(ff_stage* s, int level, float* out) {
    return guard_on(s, [&] {
        out_ptr(out, "gamma");
        *out = as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").getGamma(level);
    });
}
ff_status ff_pyramidal_set_gamma(ff_stage* s, int level, float gamma) {
    return guard_on(s, [&] { as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").setGamma(level, gamma); });
}
ff_status ff_pyramidal_set_gamma_vector(ff_stage* s, const float* gamma, int n) {
    return guard_on(s, [&] {
        if (n > 0 && !gamma) throw Error(FF_ERR_INVALID_ARGUMENT, "null gamma vector");
        std::vector<float> v(gamma, gamma + (n > 0 ? n : 0));
        as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").setGamma(v);
    });
}
ff_status ff_pyramidal_get_smooth_iterations(ff_stage* s, int level, int* out) {
    return guard_on(s, [&] {
        out_ptr(out, "iterations");
        *out = as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").getSmoothIterations(level);
    });
}
ff_status ff_pyramidal_set_smooth_iterations(ff_stage* s, int level, int n) {
    return guard_on(s, [&] { as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").setSmoothIterations(level, n); });
}
ff_status ff_pyramidal_set_smooth_iterations_vector(ff_stage* s, const int* iters, int n) {
    return guard_on(s, [&] {
        if (n > 0 && !iters) throw Error(FF_ERR_INVALID_ARGUMENT, "null iterations vector");
        std::vector<int> v(iters, iters + (n > 0 ? n : 0));
        as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").setSmoothIterations(v);
    });
}
GETTER(ff_pyramidal_get_maxflow, PyramidalFlowFilter, float, getMaxFlow)
SETTER(ff_pyramidal_set_maxflow, PyramidalFlowFilter, float, setMaxFlow)
SETTER(ff_pyramidal_set_propagation_border, PyramidalFlowFilter, int, setPropagationBorder)
GETTER(ff_pyramidal_get_propagation_border, PyramidalFlowFilter, int, getPropagationBorder)
GETTER(ff_pyramidal_height, PyramidalFlowFilter, int, height)
GETTER(ff_pyramidal_width, PyramidalFlowFilter, int, width)
GETTER(ff_pyramidal_levels, PyramidalFlowFilter, int, levels)
ff_status ff_pyramidal_load_image_device(ff_stage* s, const void* dev, size_t pitch) {
    return guard_on(s, [&] { as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").loadImageDevice(dev, pitch); });
}
ff_status ff_pyramidal_set_image_device(ff_stage* s, const void* dev, size_t pitch) {
    return guard_on(s, [&] { as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").setImageDevice(dev, pitch); });
}

// ------------------------------------------------------- execution controls
ff_status ff_filter_set_fused(ff_stage* s, int fused) {
    return guard_on(s, [&] {
        if (!s || !s->obj) throw null_handle("ff_stage");
        if (auto* f = dynamic_cast<FlowFilter*>(s->obj.get())) f->setFused(fused != 0);
        else if (auto* p = dynamic_cast<PyramidalFlowFilter*>(s->obj.get())) p->setFused(fused != 0);
        else if (auto* d = dynamic_cast<DeltaFlowFilter*>(s->obj.get())) d->setFused(fused != 0);
        else throw Error(FF_ERR_WRONG_TYPE, std::string("stage is a ") + s->obj->name() + ", expected a filter");
    });
}
ff_status ff_filter_get_fused(ff_stage* s, int* out) {
    return guard_on(s, [&] {
        out_ptr(out, "fused");
        if (!s || !s->obj) throw null_handle("ff_stage");
        if (auto* f = dynamic_cast<FlowFilter*>(s->obj.get())) *out = f->fused();
        else if (auto* p = dynamic_cast<PyramidalFlowFilter*>(s->obj.get())) *out = p->fused();
        else if (auto* d = dynamic_cast<DeltaFlowFilter*>(s->obj.get())) *out = d->fused();
        else throw Error(FF_ERR_WRONG_TYPE, std::string("stage is a ") + s->obj->name() + ", expected a filter");
    });
}

ff_status ff_filter_set_graphs(ff_stage* s, int graphs) {
    return guard_on(s, [&] {
        if (!s || !s->obj) throw null_handle("ff_stage");
        if (auto* b = dynamic_cast<BatchedFlowFilter*>(s->obj.get())) b->setGraphs(graphs != 0);
        else as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").setGraphs(graphs != 0);
    });
}
ff_status ff_filter_get_graphs(ff_stage* s, int* out) {
    return guard_on(s, [&] {
        out_ptr(out, "graphs");
        if (!s || !s->obj) throw null_handle("ff_stage");
        if (auto* b = dynamic_cast<BatchedFlowFilter*>(s->obj.get())) *out = b->graphs();
        else *out = as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").graphs();
    });
}

ff_status ff_filter_get_wait_timeouts(ff_stage* s, unsigned long long* out) {
    return guard_on(s, [&] {
        out_ptr(out, "timeouts");
        if (!s || !s->obj) throw null_handle("ff_stage");
        if (auto* b = dynamic_cast<BatchedFlowFilter*>(s->obj.get())) *out = b->waitTimeouts();
        else *out = as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").waitTimeouts();
    });
}

ff_status ff_filter_set_wait_test(ff_stage* s, int spins) {
    return guard_on(s, [&] {
        if (!s || !s->obj) throw null_handle("ff_stage");
        if (auto* b = dynamic_cast<BatchedFlowFilter*>(s->obj.get())) b->setWaitTest(spins);
        else as<PyramidalFlowFilter>(s, "PyramidalFlowFilter").setWaitTest(spins);
    });
}

ff_status ff_filter_set_tile_waves(ff_stage* s, int waves) {
    return guard_on(s, [&] { on_fused_filter(s, [&](auto& f) { f.setTileWaves(waves); }); });
}
ff_status ff_filter_get_tile_waves(ff_stage* s, int* out) {
    return guard_on(s, [&] {
        out_ptr(out, "waves");
        on_fused_filter(s, [&](auto& f) { *out = f.tileWaves(); });
    });
}
ff_status ff_filter_set_tile_top_rows(ff_stage* s, int rows) {
    return guard_on(s, [&] { on_fused_filter(s, [&](auto& f) { f.setTileTopRows(rows); }); });
}
ff_status ff_filter_get_tile_top_rows(ff_stage* s, int* out) {
    return guard_on(s, [&] {
        out_ptr(out, "rows");
        on_fused_filter(s, [&](auto& f) { *out = f.tileTopRows(); });
    });
}
ff_status ff_filter_set_tile_rows(ff_stage* s, int rows) {
    return guard_on(s, [&] { on_fused_filter(s, [&](auto& f) { f.setTileRows(rows); }); });
}
ff_status ff_filter_get_tile_rows(ff_stage* s, int* out) {
    return guard_on(s, [&] {
        out_ptr(out, "rows");
        on_fused_filter(s, [&](auto& f) { *out = f.tileRows(); });
    });
}
ff_status ff_filter_set_tile_interior(ff_stage* s, int mode) {
    return guard_on(s, [&] { on_fused_filter(s, [&](auto& f) { f.setTileInterior(mode); }); });
}
ff_status ff_filter_get_tile_interior(ff_stage* s, int* out) {
    return guard_on(s, [&] {
        out_ptr(out, "mode");
        on_fused_filter(s, [&](auto& f) { *out = f.tileInterior(); });
    });
}

// ------------------------------------------------------------------ batches
ff_status ff_batch_create(ff_stage** out) {
    return guard([&] { new_stage<BatchedFlowFilter>(out); });
}
ff_status ff_batch_add(ff_stage* batch, ff_stage* filter) {
    return guard_on(batch, [&] {
        BatchedFlowFilter& b = as<BatchedFlowFilter>(batch, "BatchedFlowFilter");
        as<PyramidalFlowFilter>(filter, "PyramidalFlowFilter");
        b.add(std::static_pointer_cast<PyramidalFlowFilter>(filter->obj));
    });
}
ff_status ff_batch_size(ff_stage* batch, int* out) {
    return guard_on(batch, [&] {
        out_ptr(out, "size");
        *out = as<BatchedFlowFilter>(batch, "BatchedFlowFilter").size();
    });
}
ff_status ff_batch_load_images_device(ff_stage* batch, const void* const* frames, const size_t* pitches, int n) {
    return guard_on(batch, [&] { as<BatchedFlowFilter>(batch, "BatchedFlowFilter").loadImagesDevice(frames, pitches, n); });
}
ff_status ff_batch_set_images_device(ff_stage* batch, const void* const* frames, const size_t* pitches, int n) {
    return guard_on(batch, [&] { as<BatchedFlowFilter>(batch, "BatchedFlowFilter").setImagesDevice(frames, pitches, n); });
}
ff_status ff_batch_get_blocks(ff_stage* batch, int* out) {
    return guard_on(batch, [&] {
        out_ptr(out, "blocks");
        *out = as<BatchedFlowFilter>(batch, "BatchedFlowFilter").lastBlocks();
    });
}

// ------------------------------------------------------------- FlowToColor
ff_status ff_get_color_wheel_rgba(ff_image_t* out) {
    return guard([&] {
        out_ptr(out, "ff_image_t*");
        *out = colorWheelRGBA();
    });
}
ff_status ff_flowtocolor_create(const ff_gpuimage* inputFlow, float maxflow, ff_stage** out) {
    return guard([&] {
        if (inputFlow) new_stage<FlowToColor>(out, inputFlow->img, maxflow);
        else new_stage<FlowToColor>(out);
    });
}
IMG_SETTER(ff_flowtocolor_set_input_flow, FlowToColor, setInputFlow)
IMG_GETTER(ff_flowtocolor_get_color_flow, FlowToColor, getColorFlow)
GETTER(ff_flowtocolor_get_max_flow, FlowToColor, float, getMaxFlow)
SETTER(ff_flowtocolor_set_max_flow, FlowToColor, float, setMaxFlow)
ff_status ff_flowtocolor_download_color_flow(ff_stage* s, ff_image_t* color) {
    return guard_on(s, [&] { as<FlowToColor>(s, "FlowToColor").downloadColorFlow(host_img(color)); });
}

// ------------------------------------------- camera, rotation, Lax-Wendroff
static_assert(sizeof(ff_camera) == sizeof(Camera), "ff_camera and Camera differ");
namespace {
Camera cam_of(const ff_camera* c) {
    if (!c) throw Error(FF_ERR_INVALID_ARGUMENT, "null ff_camera");
    return Camera{c->alphaX, c->alphaY, c->centerX, c->centerY};
}
}  // namespace
ff_status ff_create_perspective_camera(float focalLength, int height, int width, float sensorHeight,
                                       float sensorWidth, ff_camera* out) {
    return guard([&] {
        out_ptr(out, "ff_camera*");
        const Camera c = createPerspectiveCamera(focalLength, height, width, sensorHeight, sensorWidth);
        *out = ff_camera{c.alphaX, c.alphaY, c.centerX, c.centerY};
    });
}
ff_status ff_laxwendroff_create(const ff_gpuimage* inputFlow, const ff_gpuimage* inputImage, ff_stage** out) {
    return guard([&] {
        if (inputFlow && inputImage) new_stage<LaxWendroffPropagator>(out, inputFlow->img, inputImage->img);
        else new_stage<LaxWendroffPropagator>(out);
    });
}
SETTER(ff_laxwendroff_set_iterations, LaxWendroffPropagator, int, setIterations)
GETTER(ff_laxwendroff_get_iterations, LaxWendroffPropagator, int, getIterations)
GETTER(ff_laxwendroff_get_dt, LaxWendroffPropagator, float, getDt)
IMG_SETTER(ff_laxwendroff_set_input_flow, LaxWendroffPropagator, setInputFlow)
IMG_SETTER(ff_laxwendroff_set_input_image, LaxWendroffPropagator, setInputImage)
IMG_GETTER(ff_laxwendroff_get_flow, LaxWendroffPropagator, getFlow)
IMG_GETTER(ff_laxwendroff_get_propagated_image, LaxWendroffPropagator, getPropagatedImage)

ff_status ff_rotation_create(const ff_camera* cam, const ff_gpuimage* inputImage, ff_stage** out) {
    return guard([&] {
        if (!cam) new_stage<RotationalFlowImagePredictor>(out);
        else if (!inputImage) new_stage<RotationalFlowImagePredictor>(out, cam_of(cam));
        else new_stage<RotationalFlowImagePredictor>(out, cam_of(cam), inputImage->img);
    });
}
IMG_SETTER(ff_rotation_set_input_image, RotationalFlowImagePredictor, setInputImage)
IMG_GETTER(ff_rotation_get_predicted_image, RotationalFlowImagePredictor, getPredictedImage)
IMG_GETTER(ff_rotation_get_optical_flow, RotationalFlowImagePredictor, getOpticalFlow)
ff_status ff_rotation_set_camera(ff_stage* s, const ff_camera* cam) {
    return guard_on(s, [&] { as<RotationalFlowImagePredictor>(s, "RotationalFlowImagePredictor").setCamera(cam_of(cam)); });
}
ff_status ff_rotation_get_camera(ff_stage* s, ff_camera* out) {
    return guard_on(s, [&] {
        out_ptr(out, "ff_camera*");
        const Camera& c = as<RotationalFlowImagePredictor>(s, "RotationalFlowImagePredictor").getCamera();
        *out = ff_camera{c.alphaX, c.alphaY, c.centerX, c.centerY};
    });
}
ff_status ff_rotation_set_angular_velocity(ff_stage* s, float wx, float wy, float wz) {
    return guard_on(s, [&] {
        as<RotationalFlowImagePredictor>(s, "RotationalFlowImagePredictor").setAngularVelocity(wx, wy, wz);
    });
}
SETTER(ff_rotation_set_iterations, RotationalFlowImagePredictor, int, setIterations)
GETTER(ff_rotation_get_iterations, RotationalFlowImagePredictor, int, getIterations)

// ------------------------------------------------------ differentiable form
namespace {
void need(const void* p, const char* what) {
    if (!p) throw invalid_argument(std::string("null pointer: ") + what);
}
void dims(int B, int H, int W) {
    if (B <= 0 || H <= 0 || W <= 0)
        throw invalid_argument("batch, height and width must be positive, got " + std::to_string(B) + ", " +
                               std::to_string(H) + ", " + std::to_string(W));
}
hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
}  // namespace

ff_status ff_ad_partial_count(int height, int width, int* out) {
    return guard([&] {
        out_ptr(out, "count");
        dims(1, height, width);
        *out = ad::blocks(height, width);
    });
}
ff_status ff_ad_image_model(const float* image, float* a0, float* grad, int B, int H, int W, void* stream) {
    return guard([&] {
        dims(B, H, W);
        need(image, "image"), need(a0, "a0"), need(grad, "grad");
        ad::image_model(image, a0, grad, B, H, W, as_stream(stream));
        check_launch("ff_ad_image_model");
    });
}
ff_status ff_ad_image_model_backward(const float* g_a0, const float* g_grad, float* g_image, float* work, int B, int H,
                                     int W, int accumulate, void* stream) {
    return guard([&] {
        dims(B, H, W);
        need(g_a0, "g_a0"), need(g_grad, "g_grad"), need(g_image, "g_image"), need(work, "work");
        ad::image_model_bwd(g_a0, g_grad, g_image, work, B, H, W, accumulate != 0 ? 1 : 0, as_stream(stream));
        check_launch("ff_ad_image_model_backward");
    });
}
ff_status ff_ad_down(const float* image, float* half, float* work, int B, int H, int W, void* stream) {
    return guard([&] {
        dims(B, H / 2 > 0 ? H : 0, W / 2 > 0 ? W : 0);
        need(image, "image"), need(half, "half"), need(work, "work");
        ad::down(image, half, work, B, H, W, as_stream(stream));
        check_launch("ff_ad_down");
    });
}
ff_status ff_ad_down_backward(const float* g_half, float* g_image, float* work, int B, int H, int W, int accumulate,
                              void* stream) {
    return guard([&] {
        dims(B, H / 2 > 0 ? H : 0, W / 2 > 0 ? W : 0);
        need(g_half, "g_half"), need(g_image, "g_image"), need(work, "work");
        ad::down_bwd(g_half, g_image, work, B, H, W, accumulate != 0 ? 1 : 0, as_stream(stream));
        check_launch("ff_ad_down_backward");
    });
}
ff_status ff_ad_propagate_pass(int axis, const float* flow, const float* scalar, const float* vector, float* flow_out,
                               float* scalar_out, float* vector_out, int B, int H, int W, float dt, int border,
                               void* stream) {
    return guard([&] {
        dims(B, H, W);
        if (axis != 0 && axis != 1) throw invalid_argument("axis must be 0 (x) or 1 (y)");
        if (border < 1) throw invalid_argument("border must be at least 1");
        need(flow, "flow"), need(flow_out, "flow_out");
        if ((scalar == nullptr) != (vector == nullptr) || (scalar == nullptr) != (scalar_out == nullptr) ||
            (vector == nullptr) != (vector_out == nullptr))
            throw invalid_argument("payloads: pass all of scalar, vector and their outputs, or none");
        ad::prop(axis, flow, scalar, vector, flow_out, scalar_out, vector_out, B, H, W, dt, border, as_stream(stream));
        check_launch("ff_ad_propagate_pass");
    });
}
ff_status ff_ad_propagate_pair(const float* flow, const float* scalar, const float* vector, float* flow_mid,
                               float* scalar_mid, float* vector_mid, float* flow_out, float* scalar_out,
                               float* vector_out, int B, int H, int W, float dt, int border, void* stream) {
    return guard([&] {
        dims(B, H, W);
        if (border < 1) throw invalid_argument("border must be at least 1");
        need(flow, "flow"), need(flow_out, "flow_out");
        const bool pay = scalar != nullptr;
        if (pay != (vector != nullptr) || pay != (scalar_out != nullptr) || pay != (vector_out != nullptr))
            throw invalid_argument("payloads: pass all of scalar, vector and their outputs, or none");
        const bool save = flow_mid != nullptr;
        if (save && (pay ? (!scalar_mid || !vector_mid) : (scalar_mid || vector_mid)))
            throw invalid_argument("intermediates: pass those of the payloads with the flow's, or none");
        ad::prop_pair(flow, scalar, vector, flow_mid, save ? scalar_mid : nullptr, save ? vector_mid : nullptr, flow_out,
                      scalar_out, vector_out, B, H, W, dt, border, as_stream(stream));
        check_launch("ff_ad_propagate_pair");
    });
}
ff_status ff_ad_propagate_quad(const float* flow, const float* scalar, const float* vector, float* flow_out,
                               float* scalar_out, float* vector_out, int B, int H, int W, float dt, int border,
                               void* stream) {
    return guard([&] {
        dims(B, H, W);
        if (border < 1) throw invalid_argument("border must be at least 1");
        need(flow, "flow"), need(flow_out, "flow_out");
        const bool pay = scalar != nullptr;
        if (pay != (vector != nullptr) || pay != (scalar_out != nullptr) || pay != (vector_out != nullptr))
            throw invalid_argument("payloads: pass all of scalar, vector and their outputs, or none");
        if (flow_out == flow || (pay && (scalar_out == scalar || vector_out == vector)))
            throw invalid_argument("outputs must not be the inputs (the passes read neighbours)");
        ad::prop_quad(flow, scalar, vector, flow_out, scalar_out, vector_out, B, H, W, dt, border, as_stream(stream));
        check_launch("ff_ad_propagate_quad");
    });
}
ff_status ff_ad_propagate_quad_backward(const float* flow, const float* scalar, const float* vector,
                                        const float* g_flow_out, const float* g_scalar_out, const float* g_vector_out,
                                        float* g_flow, float* g_scalar, float* g_vector, int B, int H, int W, float dt,
                                        int border, int accumulate, void* stream) {
    return guard([&] {
        dims(B, H, W);
        if (border < 1) throw invalid_argument("border must be at least 1");
        need(flow, "flow"), need(g_flow, "g_flow");
        const bool pay = scalar != nullptr;
        if (pay && (!vector || !g_scalar || !g_vector))
            throw invalid_argument("payloads: pass scalar, vector and both of their gradients");
        if (g_flow == g_flow_out || (pay && ((g_scalar_out && g_scalar == g_scalar_out) ||
                                             (g_vector_out && g_vector == g_vector_out))))
            throw invalid_argument("input gradients must not be the output gradients (the adjoint reads neighbours)");
        ad::prop_quad_bwd(flow, scalar, vector, g_flow_out, pay ? g_scalar_out : nullptr,
                          pay ? g_vector_out : nullptr, g_flow, g_scalar, g_vector, B, H, W, dt, border, accumulate,
                          as_stream(stream));
        check_launch("ff_ad_propagate_quad_backward");
    });
}
ff_status ff_ad_propagate_pair_backward(const float* flow, const float* scalar, const float* vector,
                                        const float* flow_mid, const float* scalar_mid, const float* vector_mid,
                                        const float* g_flow_out, const float* g_scalar_out, const float* g_vector_out,
                                        float* g_flow, float* g_scalar, float* g_vector, int B, int H, int W, float dt,
                                        int border, int accumulate, void* stream) {
    return guard([&] {
        dims(B, H, W);
        if (border < 1) throw invalid_argument("border must be at least 1");
        need(flow, "flow"), need(g_flow, "g_flow");
        const bool pay = scalar != nullptr, recompute = flow_mid == nullptr;
        if (recompute && (scalar_mid || vector_mid))
            throw invalid_argument("flow_mid is NULL (the X pass is recomputed): scalar_mid and vector_mid must be NULL too");
        if (pay && (!vector || !g_scalar || !g_vector))
            throw invalid_argument("payloads: pass scalar, vector and both of their gradients");
        if (pay && !recompute && (!scalar_mid || !vector_mid))
            throw invalid_argument("payloads: pass the X pass's scalar_mid and vector_mid with flow_mid");
        ad::prop_pair_bwd(flow, scalar, vector, flow_mid, pay ? scalar_mid : nullptr, pay ? vector_mid : nullptr,
                          g_flow_out, pay ? g_scalar_out : nullptr, pay ? g_vector_out : nullptr, g_flow, g_scalar,
                          g_vector, B, H, W, dt, border, accumulate, as_stream(stream));
        check_launch("ff_ad_propagate_pair_backward");
    });
}
ff_status ff_ad_propagate_pass_backward(int axis, const float* flow, const float* scalar, const float* vector,
                                        const float* g_flow_out, const float* g_scalar_out, const float* g_vector_out,
                                        float* g_flow, float* g_scalar, float* g_vector, int B, int H, int W, float dt,
                                        int border, int accumulate, void* stream) {
    return guard([&] {
        dims(B, H, W);
        if (axis != 0 && axis != 1) throw invalid_argument("axis must be 0 (x) or 1 (y)");
        if (border < 1) throw invalid_argument("border must be at least 1");
        need(flow, "flow"), need(g_flow, "g_flow");
        const bool pay = scalar != nullptr;
        if (pay && (!vector || !g_scalar || !g_vector))
            throw invalid_argument("payloads: pass scalar, vector and both of their gradients");
        ad::prop_bwd(axis, flow, scalar, vector, g_flow_out, pay ? g_scalar_out : nullptr,
                     pay ? g_vector_out : nullptr, g_flow, g_scalar, g_vector, B, H, W, dt, border, accumulate,
                     as_stream(stream));
        check_launch("ff_ad_propagate_pass_backward");
    });
}
ff_status ff_ad_update(const float* a0, const float* grad, const float* a0_old, const float* flow, float* flow_out,
                       int B, int H, int W, float gamma, float maxflow, void* stream) {
    return guard([&] {
        dims(B, H, W);
        need(a0, "a0"), need(grad, "grad"), need(a0_old, "a0_old"), need(flow, "flow"), need(flow_out, "flow_out");
        ad::update(a0, grad, a0_old, flow, flow_out, B, H, W, gamma, maxflow, as_stream(stream));
        check_launch("ff_ad_update");
    });
}
ff_status ff_ad_update_backward(const float* a0, const float* grad, const float* a0_old, const float* flow,
                                const float* g_flow_out, float* g_a0, float* g_grad, float* g_a0_old, float* g_flow,
                                float* g_gamma_partial, int B, int H, int W, float gamma, float maxflow,
                                int accumulate, void* stream) {
    return guard([&] {
        dims(B, H, W);
        for (const void* p : {(const void*)a0, (const void*)grad, (const void*)a0_old, (const void*)flow,
                              (const void*)g_flow_out, (const void*)g_a0, (const void*)g_grad, (const void*)g_a0_old,
                              (const void*)g_flow, (const void*)g_gamma_partial})
            need(p, "argument");
        ad::update_bwd(a0, grad, a0_old, flow, g_flow_out, g_a0, g_grad, g_a0_old, g_flow, g_gamma_partial, B, H, W,
                       gamma, maxflow, accumulate, as_stream(stream));
        check_launch("ff_ad_update_backward");
    });
}
ff_status ff_ad_update_backward_noalloc(const float* a0, const float* grad, const float* a0_old, const float* flow,
                                        const float* g_flow_out, float* g_a0, float* g_grad, float* g_a0_old,
                                        float* g_flow, double* g_gamma, int B, int H, int W, float gamma,
                                        float maxflow, int accumulate, void* stream) {
    return guard([&] {
        dims(B, H, W);
        for (const void* p : {(const void*)a0, (const void*)grad, (const void*)a0_old, (const void*)flow,
                              (const void*)g_flow_out, (const void*)g_a0, (const void*)g_grad, (const void*)g_a0_old,
                              (const void*)g_flow, (const void*)g_gamma})
            need(p, "argument");
        ad::update_bwd_noalloc(a0, grad, a0_old, flow, g_flow_out, g_a0, g_grad, g_a0_old, g_flow, g_gamma, B, H, W,
                               gamma, maxflow, accumulate, as_stream(stream));
        check_launch("ff_ad_update_backward_noalloc");
    });
}
ff_status ff_ad_delta_update(const float* a0, const float* grad, const float* a0_prop, const float* dflow_prop,
                             const float* upper, const int* xmap, const int* ymap, float* dflow_out, float* flow_out,
                             int B, int H, int W, int Hu, int Wu, float gamma, float maxflow, void* stream) {
    return guard([&] {
        dims(B, H, W);
        dims(B, Hu, Wu);
        for (const void* p : {(const void*)a0, (const void*)grad, (const void*)a0_prop, (const void*)dflow_prop,
                              (const void*)upper, (const void*)xmap, (const void*)ymap, (const void*)dflow_out,
                              (const void*)flow_out})
            need(p, "argument");
        ad::delta_update(a0, grad, a0_prop, dflow_prop, upper, xmap, ymap, dflow_out, flow_out, B, H, W, Hu, Wu, gamma,
                         maxflow, as_stream(stream));
        check_launch("ff_ad_delta_update");
    });
}
ff_status ff_ad_delta_update_backward(const float* a0, const float* grad, const float* a0_prop,
                                      const float* dflow_prop, const float* upper, const int* xmap, const int* ymap,
                                      const float* g_dflow_out, const float* g_flow_out, float* g_a0, float* g_grad,
                                      float* g_a0_prop, float* g_dflow_prop, float* g_upper_pixel,
                                      float* g_gamma_partial, int B, int H, int W, int Hu, int Wu, float gamma,
                                      float maxflow, int accumulate, void* stream) {
    return guard([&] {
        dims(B, H, W);
        dims(B, Hu, Wu);
        for (const void* p : {(const void*)a0, (const void*)grad, (const void*)a0_prop, (const void*)dflow_prop,
                              (const void*)upper, (const void*)xmap, (const void*)ymap,
                              (const void*)g_flow_out, (const void*)g_a0, (const void*)g_grad, (const void*)g_a0_prop,
                              (const void*)g_dflow_prop, (const void*)g_upper_pixel, (const void*)g_gamma_partial})
            need(p, "argument");
        ad::delta_update_bwd(a0, grad, a0_prop, dflow_prop, upper, xmap, ymap, g_dflow_out, g_flow_out, g_a0, g_grad,
                             g_a0_prop, g_dflow_prop, g_upper_pixel, g_gamma_partial, B, H, W, Hu, Wu, gamma, maxflow,
                             accumulate, as_stream(stream));
        check_launch("ff_ad_delta_update_backward");
    });
}
ff_status ff_ad_upper_backward(const float* g_upper_pixel, const int* xlo, const int* xhi, const int* ylo,
                               const int* yhi, float* g_upper, int B, int H, int W, int Hu, int Wu, int accumulate,
                               void* stream) {
    return guard([&] {
        dims(B, H, W);
        dims(B, Hu, Wu);
        for (const void* p : {(const void*)g_upper_pixel, (const void*)xlo, (const void*)xhi, (const void*)ylo,
                              (const void*)yhi, (const void*)g_upper})
            need(p, "argument");
        ad::upper_gather(g_upper_pixel, xlo, xhi, ylo, yhi, g_upper, B, H, W, Hu, Wu, accumulate != 0 ? 1 : 0,
                         as_stream(stream));
        check_launch("ff_ad_upper_backward");
    });
}
ff_status ff_ad_smooth_pass(int axis, const float* flow, float* flow_out, int B, int H, int W, void* stream) {
    return guard([&] {
        dims(B, H, W);
        if (axis != 0 && axis != 1) throw invalid_argument("axis must be 0 (x) or 1 (y)");
        need(flow, "flow"), need(flow_out, "flow_out");
        ad::smooth(axis, flow, flow_out, B, H, W, as_stream(stream));
        check_launch("ff_ad_smooth_pass");
    });
}
ff_status ff_ad_smooth_pass_backward(int axis, const float* g_flow_out, const float* g_flow_out2, float* g_flow, int B,
                                     int H, int W, int accumulate, void* stream) {
    return guard([&] {
        dims(B, H, W);
        if (axis != 0 && axis != 1) throw invalid_argument("axis must be 0 (x) or 1 (y)");
        need(g_flow_out, "g_flow_out"), need(g_flow, "g_flow");
        ad::smooth_bwd(axis, g_flow_out, g_flow_out2, g_flow, B, H, W, accumulate != 0 ? 1 : 0, as_stream(stream));
        check_launch("ff_ad_smooth_pass_backward");
    });
}

ff_status ff_ad_smooth_pair(const float* flow, float* flow_out, int B, int H, int W, void* stream) {
    return guard([&] {
        dims(B, H, W);
        need(flow, "flow"), need(flow_out, "flow_out");
        if (flow == flow_out) throw invalid_argument("flow_out must not be flow (the passes read neighbours)");
        ad::smooth_pair(flow, flow_out, B, H, W, as_stream(stream));
        check_launch("ff_ad_smooth_pair");
    });
}
ff_status ff_ad_smooth_pair_backward(const float* g_flow_out, const float* g_flow_out2, float* g_flow, int B, int H,
                                     int W, int accumulate, void* stream) {
    return guard([&] {
        dims(B, H, W);
        need(g_flow_out, "g_flow_out"), need(g_flow, "g_flow");
        if (g_flow == g_flow_out || g_flow == g_flow_out2)
            throw invalid_argument("g_flow must not be an output gradient (the adjoint reads neighbours)");
        ad::smooth_pair_bwd(g_flow_out, g_flow_out2, g_flow, B, H, W, accumulate != 0 ? 1 : 0, as_stream(stream));
        check_launch("ff_ad_smooth_pair_backward");
    });
}

ff_status ff_ad_photometric_partial_count(int B, int T, int height, int width, int* out) {
    return guard([&] {
        out_ptr(out, "count");
        if (T < 2) throw invalid_argument("photometric term: needs at least two frames, got T = " + std::to_string(T));
        dims(B, height, width);
        *out = ad::photometric_partials(B, T, height, width);
    });
}
ff_status ff_ad_photometric(const float* frames, const float* flows, float* partial, int B, int T, int H, int W,
                            void* stream) {
    return guard([&] {
        if (T < 2) throw invalid_argument("photometric term: needs at least two frames, got T = " + std::to_string(T));
        dims(B, H, W);
        need(frames, "frames"), need(flows, "flows"), need(partial, "partial");
        ad::photometric(frames, flows, partial, B, T, H, W, as_stream(stream));
        check_launch("ff_ad_photometric");
    });
}
ff_status ff_ad_photometric_backward(const float* frames, const float* flows, const float* scale, float* g_frames,
                                     float* g_flows, int B, int T, int H, int W, void* stream) {
    return guard([&] {
        if (T < 2) throw invalid_argument("photometric term: needs at least two frames, got T = " + std::to_string(T));
        dims(B, H, W);
        need(frames, "frames"), need(flows, "flows"), need(scale, "scale"), need(g_frames, "g_frames"),
            need(g_flows, "g_flows");
        if ((const float*)g_frames == frames || (const float*)g_flows == flows)
            throw invalid_argument("photometric adjoint: the gradients must not alias the inputs");
        ad::photometric_bwd(frames, flows, scale, g_frames, g_flows, B, T, H, W, as_stream(stream));
        check_launch("ff_ad_photometric_backward");
    });
}

// --------------------------------------------------------------- profiling
ff_status ff_profile_begin(const char* filter) {
    return guard([&] { profile_begin(filter ? filter : ""); });
}
ff_status ff_profile_end(double* total_ms, long long* count) {
    return guard([&] { profile_end(total_ms, count); });
}

// ------------------------------------------------------------ stream markers
ff_status ff_marker_record(ff_marker** out) {
    return guard([&] {
        out_ptr(out, "ff_marker**");
        hipEvent_t e = nullptr;
        FF_HIP(hipEventCreate(&e));
        const hipError_t r = hipEventRecord(e, stream());
        if (r != hipSuccess) {
            (void)hipEventDestroy(e);
            throw_hip(r, "hipEventRecord", __FILE__, __LINE__);
        }
        *out = reinterpret_cast<ff_marker*>(e);
    });
}
ff_status ff_marker_record_on(ff_stage* s, ff_marker** out) {
    return guard_on(s, [&] {
        if (!s || !s->obj) throw null_handle("ff_stage");
        out_ptr(out, "ff_marker**");
        hipEvent_t e = nullptr;
        FF_HIP(hipEventCreate(&e));
        const hipError_t r = hipEventRecord(e, stream());
        if (r != hipSuccess) {
            (void)hipEventDestroy(e);
            throw_hip(r, "hipEventRecord", __FILE__, __LINE__);
        }
        *out = reinterpret_cast<ff_marker*>(e);
    });
}
ff_status ff_marker_elapsed(ff_marker* start, ff_marker* stop, float* ms) {
    return guard([&] {
        if (!start || !stop) throw null_handle("ff_marker");
        out_ptr(ms, "ms");
        FF_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
        FF_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)));
    });
}
ff_status ff_marker_destroy(ff_marker* m) {
    return guard([&] {
        if (m) FF_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(m)));
    });
}

}  // extern "C"
