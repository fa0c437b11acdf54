/*
 * flowfilter_amd.h -- C ABI of the MI355X (gfx950) optical-flow filter
 * library, libflowfilter_amd.so.
 *
 * The entry points mirror the C++ classes of namespace flowfilter::gpu
 * one to one: each function replaces the reference method cited next to it
 * (paths relative to the reference repository root). Plain pointers, sizes
 * and opaque handles only; no HIP or C++ types cross this boundary.
 *
 * Conventions
 *  - Every function returns ff_status. FF_OK is 0; on any other value
 *    ff_last_error() returns a message for the calling thread. The library
 *    never exits the process (the reference calls exit() on CUDA errors,
 *    src/gpu/error.cu:15-24).
 *  - Status codes map to the reference's exception types:
 *    FF_ERR_INVALID_ARGUMENT <-> std::invalid_argument (shape mismatch,
 *    wrong depth or item size), FF_ERR_LOGIC <-> std::logic_error
 *    (unconfigured stage), FF_ERR_RUNTIME <-> bare std::exception thrown by
 *    the setters (gamma <= 0, iterations <= 0, level out of range).
 *  - Device images (ff_gpuimage) are reference-counted buffer handles with
 *    the shared-pointer semantics of flowfilter::gpu::GPUImage
 *    (include/flowfilter/gpu/image.h:47-123): every getter returns a new
 *    handle sharing the buffer; release it with ff_gpuimage_destroy. Stages
 *    hold their own references, so a buffer lives until its last holder is
 *    gone.
 *  - Streams. Every filter (FlowFilter, DeltaFlowFilter,
 *    PyramidalFlowFilter) and every batch enqueues its work on a HIP stream
 *    of its own (ff_stage_get_stream; ff_stage_set_stream substitutes a
 *    caller's); a filter added to a batch follows the batch's stream. The
 *    other stages, and images created outside a filter, use the library
 *    stream (ff_get_stream / ff_set_stream). A device image's own
 *    transfers run on the stream of the object that created it, so a
 *    download through a filter's image handle waits for that filter's
 *    work; where a stage reads an image another object owns, the two
 *    streams are ordered with events. compute() returns without waiting;
 *    elapsed_time, download and upload calls wait for the work they depend
 *    on. Objects are not thread-safe: use each object from one thread at a
 *    time; filters on different threads do not wait for each other.
 */
#ifndef FLOWFILTER_AMD_H_
#define FLOWFILTER_AMD_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FF_API __attribute__((visibility("default")))

typedef int ff_status;
enum {
    FF_OK = 0,
    FF_ERR_INVALID_ARGUMENT = 1,
    FF_ERR_LOGIC = 2,
    FF_ERR_RUNTIME = 3,
    FF_ERR_HIP = 4,
    FF_ERR_BAD_ALLOC = 5,
    FF_ERR_NULL_HANDLE = 6,
    FF_ERR_WRONG_TYPE = 7
};

/* Host image descriptor, byte-identical to flowfilter::image_t
 * (include/flowfilter/image.h:30-49; 40 bytes on LP64). */
typedef struct ff_image_t {
    int height;
    int width;
    int depth;
    size_t pitch;    /* row pitch in bytes */
    size_t itemSize; /* bytes per channel element */
    void* data;
} ff_image_t;

typedef struct ff_gpuimage ff_gpuimage; /* flowfilter::gpu::GPUImage */
typedef struct ff_stage ff_stage;       /* flowfilter::gpu::Stage and subclasses */

/* ------------------------------------------------------------- library */
FF_API const char* ff_last_error(void);
FF_API ff_status ff_version(int* major, int* minor);
/* Wait for all work enqueued by the library (device-wide). */
FF_API ff_status ff_synchronize(void);
/* The library stream: standalone stages and images created outside a
 * filter enqueue on it. set substitutes the given hipStream_t (NULL
 * restores the library's own); call it while no other thread is using the
 * library. */
FF_API ff_status ff_set_stream(void* hip_stream);
FF_API ff_status ff_get_stream(void** hip_stream);

/* --------------------------------------------- host images (image.cpp) */
/* createImage (include/flowfilter/image.h:59,69; src/image.cpp:14-40) */
FF_API ff_status ff_create_image(int height, int width, int depth, size_t pixelSize, ff_image_t* out);
/* destroyImage (include/flowfilter/image.h:78; src/image.cpp:43-52) */
FF_API ff_status ff_destroy_image(ff_image_t* image);

/* ------------------------------------ GPUImage (gpu/image.h:47-123) */
/* GPUImage(height, width, depth, itemSize) (src/gpu/image.cu:32-42) */
FF_API ff_status ff_gpuimage_create(int height, int width, int depth, int itemSize, ff_gpuimage** out);
/* GPUImage() (src/gpu/image.cu:24-30): unallocated; upload allocates */
FF_API ff_status ff_gpuimage_create_empty(ff_gpuimage** out);
/* copy constructor: a second handle to the same buffer */
FF_API ff_status ff_gpuimage_share(const ff_gpuimage* img, ff_gpuimage** out);
FF_API ff_status ff_gpuimage_destroy(ff_gpuimage* img);
/* height()/width()/depth()/pitch()/itemSize() (src/gpu/image.cu:52-70) */
FF_API ff_status ff_gpuimage_height(const ff_gpuimage* img, int* out);
FF_API ff_status ff_gpuimage_width(const ff_gpuimage* img, int* out);
FF_API ff_status ff_gpuimage_depth(const ff_gpuimage* img, int* out);
FF_API ff_status ff_gpuimage_pitch(const ff_gpuimage* img, int* out);
FF_API ff_status ff_gpuimage_item_size(const ff_gpuimage* img, int* out);
/* data() (src/gpu/image.cu:72-74): device pointer */
FF_API ff_status ff_gpuimage_data(const ff_gpuimage* img, void** out);
/* upload (src/gpu/image.cu:81-121) */
FF_API ff_status ff_gpuimage_upload(ff_gpuimage* img, const ff_image_t* host);
/* download (src/gpu/image.cu:123-150) */
FF_API ff_status ff_gpuimage_download(const ff_gpuimage* img, ff_image_t* host);
/* copyFrom (src/gpu/image.cu:152-170) */
FF_API ff_status ff_gpuimage_copy_from(ff_gpuimage* img, const ff_gpuimage* src);
/* clear (src/gpu/image.cu:172-176) */
FF_API ff_status ff_gpuimage_clear(ff_gpuimage* img);
/* Device-resident variants of upload/download (no reference counterpart):
 * the source/destination is device memory with the given row pitch, the
 * copy is enqueued on the library stream without waiting. */
FF_API ff_status ff_gpuimage_upload_device(ff_gpuimage* img, const void* dev, size_t pitch);
FF_API ff_status ff_gpuimage_download_device(const ff_gpuimage* img, void* dev, size_t pitch);

/* DLPack export of a device image (no reference counterpart; the zero-copy
 * form of getFlowDevice, python/flowfilter/gpu/flowfilters.pyx:244-251).
 * Writes a DLManagedTensor* (DLPack ABI v0.8: ROCm device, uint8 or
 * float32, shape (height, width) or (height, width, depth), strides in
 * elements from the row pitch) viewing the image's buffer as it is now; the
 * tensor holds a share of the buffer until its deleter runs. With sync, the
 * consumer's hipStream_t first waits for the work enqueued on the image's
 * stream. A filter's state images alternate between two buffers from one
 * compute() to the next: export again after each compute(). */
FF_API ff_status ff_gpuimage_to_dlpack(const ff_gpuimage* img, void* consumer_stream, int sync, void** managed);
/* Runs the deleter of a managed tensor that no consumer took. */
FF_API ff_status ff_dlpack_delete(void* managed);

/* ------------------------------------------ Stage (gpu/pipeline.h:29-78) */
FF_API ff_status ff_stage_configure(ff_stage* s);              /* configure() */
FF_API ff_status ff_stage_compute(ff_stage* s);                /* compute() */
FF_API ff_status ff_stage_elapsed_time(ff_stage* s, float* ms); /* elapsedTime() */
FF_API ff_status ff_stage_destroy(ff_stage* s);
/* The hipStream_t the stage's work is enqueued on (no reference
 * counterpart; the reference uses stream 0 for everything,
 * src/gpu/pipeline.cu:21-89): a filter's or batch's own, else the library
 * stream. */
FF_API ff_status ff_stage_get_stream(ff_stage* s, void** hip_stream);
/* Filters and batches (no reference counterpart): enqueue on the caller's
 * hip_stream from now on (NULL returns to the filter's own stream). Waits
 * for the work already enqueued on the previous stream. Other stages:
 * FF_ERR_WRONG_TYPE. */
FF_API ff_status ff_stage_set_stream(ff_stage* s, void* hip_stream);

/* ------------------------------ ImageModel (gpu/imagemodel.h:22-63) */
/* ImageModel() / ImageModel(GPUImage) (src/gpu/imagemodel.cu:24-43); input may be NULL */
FF_API ff_status ff_imagemodel_create(const ff_gpuimage* input, ff_stage** out);
FF_API ff_status ff_imagemodel_set_input_image(ff_stage* s, const ff_gpuimage* img);  /* :119-134 */
FF_API ff_status ff_imagemodel_get_image_constant(ff_stage* s, ff_gpuimage** out);     /* :139-141 */
FF_API ff_status ff_imagemodel_get_image_gradient(ff_stage* s, ff_gpuimage** out);     /* :144-146 */

/* -------------------------- FlowPropagator (gpu/propagation.h:25-105) */
/* FlowPropagator(GPUImage, iterations) (src/gpu/propagation.cu:20-44); inputFlow may be NULL */
FF_API ff_status ff_flowpropagator_create(const ff_gpuimage* inputFlow, int iterations, ff_stage** out);
FF_API ff_status ff_flowpropagator_set_iterations(ff_stage* s, int n);      /* :137-148 */
FF_API ff_status ff_flowpropagator_get_iterations(ff_stage* s, int* out);   /* :151-153 */
FF_API ff_status ff_flowpropagator_get_dt(ff_stage* s, float* out);        /* :156-158 */
FF_API ff_status ff_flowpropagator_set_border(ff_stage* s, int border);     /* :160-170 */
FF_API ff_status ff_flowpropagator_get_border(ff_stage* s, int* out);       /* :173-175 */
FF_API ff_status ff_flowpropagator_set_invert_input_flow(ff_stage* s, int invert); /* :204-206 */
FF_API ff_status ff_flowpropagator_get_invert_input_flow(ff_stage* s, int* out);   /* :209-211 */
FF_API ff_status ff_flowpropagator_set_input_flow(ff_stage* s, const ff_gpuimage* flow); /* :178-195 */
FF_API ff_status ff_flowpropagator_get_propagated_flow(ff_stage* s, ff_gpuimage** out);  /* :198-201 */

/* ------------- FlowPropagatorPayload (gpu/propagation.h:108-200) */
/* (src/gpu/propagation.cu:217-252); any input may be NULL */
FF_API ff_status ff_flowpropagatorpayload_create(const ff_gpuimage* inputFlow, const ff_gpuimage* scalarPayload,
                                                 const ff_gpuimage* vectorPayload, int iterations, ff_stage** out);
FF_API ff_status ff_flowpropagatorpayload_set_iterations(ff_stage* s, int n);    /* :376-387 */
FF_API ff_status ff_flowpropagatorpayload_get_iterations(ff_stage* s, int* out); /* :389-391 */
FF_API ff_status ff_flowpropagatorpayload_get_dt(ff_stage* s, float* out);      /* :393-395 */
FF_API ff_status ff_flowpropagatorpayload_set_border(ff_stage* s, int border);   /* :398-408 */
FF_API ff_status ff_flowpropagatorpayload_get_border(ff_stage* s, int* out);     /* :411-413 */
FF_API ff_status ff_flowpropagatorpayload_set_input_flow(ff_stage* s, const ff_gpuimage* flow);      /* :418-434 */
FF_API ff_status ff_flowpropagatorpayload_set_scalar_payload(ff_stage* s, const ff_gpuimage* img);   /* :436-460 */
FF_API ff_status ff_flowpropagatorpayload_set_vector_payload(ff_stage* s, const ff_gpuimage* img);   /* :462-486 */
FF_API ff_status ff_flowpropagatorpayload_get_propagated_flow(ff_stage* s, ff_gpuimage** out);       /* :492-494 */
FF_API ff_status ff_flowpropagatorpayload_get_propagated_scalar(ff_stage* s, ff_gpuimage** out);     /* :496-498 */
FF_API ff_status ff_flowpropagatorpayload_get_propagated_vector(ff_stage* s, ff_gpuimage** out);     /* :500-502 */

/* ------------------------------------ FlowUpdate (gpu/update.h:21-90) */
/* FlowUpdate(inputFlow, inputImage, inputImageGradient, gamma, maxflow) (src/gpu/update.cu:20-50) */
FF_API ff_status ff_flowupdate_create(const ff_gpuimage* inputFlow, const ff_gpuimage* inputImage,
                                      const ff_gpuimage* inputImageGradient, float gamma, float maxflow,
                                      ff_stage** out);
FF_API ff_status ff_flowupdate_get_gamma(ff_stage* s, float* out);    /* :119-121 */
FF_API ff_status ff_flowupdate_set_gamma(ff_stage* s, float gamma);   /* :124-132 */
FF_API ff_status ff_flowupdate_get_maxflow(ff_stage* s, float* out);  /* :135-137 */
FF_API ff_status ff_flowupdate_set_maxflow(ff_stage* s, float m);     /* :140-143 */
FF_API ff_status ff_flowupdate_set_input_flow(ff_stage* s, const ff_gpuimage* img);           /* :146-162 */
FF_API ff_status ff_flowupdate_set_input_image(ff_stage* s, const ff_gpuimage* img);          /* :165-181 */
FF_API ff_status ff_flowupdate_set_input_image_gradient(ff_stage* s, const ff_gpuimage* img); /* :184-200 */
FF_API ff_status ff_flowupdate_get_updated_flow(ff_stage* s, ff_gpuimage** out);              /* :203-206 */
FF_API ff_status ff_flowupdate_get_updated_image(ff_stage* s, ff_gpuimage** out);             /* :209-212 */

/* ------------------------------- DeltaFlowUpdate (gpu/update.h:93-174) */
/* (src/gpu/update.cu:221-259) */
FF_API ff_status ff_deltaflowupdate_create(const ff_gpuimage* inputFlow, const ff_gpuimage* inputDeltaFlow,
                                           const ff_gpuimage* inputImageOld, const ff_gpuimage* inputImage,
                                           const ff_gpuimage* inputImageGradient, float gamma, float maxflow,
                                           ff_stage** out);
FF_API ff_status ff_deltaflowupdate_get_gamma(ff_stage* s, float* out);    /* :351-353 */
FF_API ff_status ff_deltaflowupdate_set_gamma(ff_stage* s, float gamma);   /* :355-363 */
FF_API ff_status ff_deltaflowupdate_get_maxflow(ff_stage* s, float* out);  /* :366-368 */
FF_API ff_status ff_deltaflowupdate_set_maxflow(ff_stage* s, float m);     /* :371-374 */
FF_API ff_status ff_deltaflowupdate_set_input_flow(ff_stage* s, const ff_gpuimage* img);           /* :377-393 */
FF_API ff_status ff_deltaflowupdate_set_input_delta_flow(ff_stage* s, const ff_gpuimage* img);     /* :396-412 */
FF_API ff_status ff_deltaflowupdate_set_input_image_old(ff_stage* s, const ff_gpuimage* img);      /* :414-430 */
FF_API ff_status ff_deltaflowupdate_set_input_image(ff_stage* s, const ff_gpuimage* img);          /* :433-449 */
FF_API ff_status ff_deltaflowupdate_set_input_image_gradient(ff_stage* s, const ff_gpuimage* img); /* :452-468 */
FF_API ff_status ff_deltaflowupdate_get_updated_flow(ff_stage* s, ff_gpuimage** out);              /* :471-474 */
FF_API ff_status ff_deltaflowupdate_get_updated_delta_flow(ff_stage* s, ff_gpuimage** out);        /* :476-479 */
FF_API ff_status ff_deltaflowupdate_get_updated_image(ff_stage* s, ff_gpuimage** out);             /* :482-485 */

/* ------------------------- FlowSmoother (gpu/flowsmoothing.h:22-57) */
/* FlowSmoother(inputFlow, iterations) (src/gpu/flowsmoothing.cu:20-38) */
FF_API ff_status ff_flowsmoother_create(const ff_gpuimage* inputFlow, int iterations, ff_stage** out);
FF_API ff_status ff_flowsmoother_get_iterations(ff_stage* s, int* out);   /* :111-114 */
FF_API ff_status ff_flowsmoother_set_iterations(ff_stage* s, int n);      /* :117-127 */
FF_API ff_status ff_flowsmoother_set_input_flow(ff_stage* s, const ff_gpuimage* img); /* :130-146 */
FF_API ff_status ff_flowsmoother_get_smoothed_flow(ff_stage* s, ff_gpuimage** out);   /* :149-152 */

/* -------------------------------- ImagePyramid (gpu/pyramid.h:22-89) */
/* ImagePyramid(GPUImage, levels) (src/gpu/pyramid.cu:22-41) */
FF_API ff_status ff_imagepyramid_create(const ff_gpuimage* image, int levels, ff_stage** out);
FF_API ff_status ff_imagepyramid_set_input_image(ff_stage* s, const ff_gpuimage* img); /* :173-190 */
FF_API ff_status ff_imagepyramid_set_levels(ff_stage* s, int levels);                  /* :192-201 */
FF_API ff_status ff_imagepyramid_get_levels(ff_stage* s, int* out);                    /* :222-224 */
FF_API ff_status ff_imagepyramid_get_image(ff_stage* s, int level, ff_gpuimage** out); /* :207-219 */

/* ------------------------------- FlowToColor (gpu/display.h:23-93) */
/* getColorWheelRGBA (include/flowfilter/colorwheel.h:26, src/colorwheel.cpp:
 * 21862-21873): the 256 x 256 RGBA colour wheel (u8, pitch 1024) as a host
 * image; its data is the library's and read-only. */
FF_API ff_status ff_get_color_wheel_rgba(ff_image_t* out);
/* FlowToColor() / FlowToColor(inputFlow, maxflow) (src/gpu/display.cu:20-42);
 * inputFlow may be NULL. Colours flow vectors (float2) as RGBA (4 x u8): the
 * wheel read at ((u + m) / 2m, (v + m) / 2m) with clamped,
 * normalised-coordinate point sampling (display_k.cu:14-43). */
FF_API ff_status ff_flowtocolor_create(const ff_gpuimage* inputFlow, float maxflow, ff_stage** out);
FF_API ff_status ff_flowtocolor_set_input_flow(ff_stage* s, const ff_gpuimage* flow);   /* :100-117 */
FF_API ff_status ff_flowtocolor_get_color_flow(ff_stage* s, ff_gpuimage** out);         /* :119-121 */
FF_API ff_status ff_flowtocolor_get_max_flow(ff_stage* s, float* out);                  /* :124-126 */
FF_API ff_status ff_flowtocolor_set_max_flow(ff_stage* s, float maxflow);               /* :129-137 */
FF_API ff_status ff_flowtocolor_download_color_flow(ff_stage* s, ff_image_t* color);    /* :140-142 */

/* ----------- camera, rotational prediction, Lax-Wendroff propagation */
/* perspectiveCamera (include/flowfilter/gpu/camera.h:20-25), layout-identical. */
typedef struct ff_camera {
    float alphaX;
    float alphaY;
    float centerX;
    float centerY;
} ff_camera;
/* createPerspectiveCamera (src/gpu/camera.cu:14-29) */
FF_API ff_status ff_create_perspective_camera(float focalLength, int height, int width, float sensorHeight,
                                              float sensorWidth, ff_camera* out);
/* LaxWendroffPropagator() / (inputFlow, inputImage) (src/gpu/propagation.cu:
 * 509-525; gpu/propagation.h:218-291); both inputs or neither. Propagates a
 * float image (depth 1 or 4) along a flow field: iterations of a Y then an
 * X Lax-Wendroff pass (propagation_k.h:49-132), dt = 1/iterations. */
FF_API ff_status ff_laxwendroff_create(const ff_gpuimage* inputFlow, const ff_gpuimage* inputImage, ff_stage** out);
FF_API ff_status ff_laxwendroff_set_iterations(ff_stage* s, int n);                     /* :647-658 */
FF_API ff_status ff_laxwendroff_get_iterations(ff_stage* s, int* out);                  /* :660-662 */
FF_API ff_status ff_laxwendroff_get_dt(ff_stage* s, float* out);                        /* :665-667 */
FF_API ff_status ff_laxwendroff_set_input_flow(ff_stage* s, const ff_gpuimage* flow);   /* :671-687 */
FF_API ff_status ff_laxwendroff_set_input_image(ff_stage* s, const ff_gpuimage* img);   /* :690-714 */
FF_API ff_status ff_laxwendroff_get_flow(ff_stage* s, ff_gpuimage** out);               /* :717-719 */
FF_API ff_status ff_laxwendroff_get_propagated_image(ff_stage* s, ff_gpuimage** out);   /* :722-724 */
/* RotationalFlowImagePredictor() / (cam) / (cam, inputImage) (src/gpu/
 * rotation.cu:22-42; gpu/rotation.h:24-88): cam NULL -> the first, image NULL
 * -> the second. compute() writes the rotational flow of the camera's
 * angular velocity (rotation_k.cu:17-43) and propagates the input image
 * along it. */
FF_API ff_status ff_rotation_create(const ff_camera* cam, const ff_gpuimage* inputImage, ff_stage** out);
FF_API ff_status ff_rotation_set_input_image(ff_stage* s, const ff_gpuimage* img);      /* :91-97 */
FF_API ff_status ff_rotation_get_predicted_image(ff_stage* s, ff_gpuimage** out);       /* :100-103 */
FF_API ff_status ff_rotation_get_optical_flow(ff_stage* s, ff_gpuimage** out);          /* :105-108 */
FF_API ff_status ff_rotation_set_camera(ff_stage* s, const ff_camera* cam);             /* :111-113 */
/* the camera set last (no reference counterpart: the Python wrapper keeps it) */
FF_API ff_status ff_rotation_get_camera(ff_stage* s, ff_camera* out);
FF_API ff_status ff_rotation_set_angular_velocity(ff_stage* s, float wx, float wy, float wz); /* :116-121 */
FF_API ff_status ff_rotation_set_iterations(ff_stage* s, int n);                        /* :124-127 */
FF_API ff_status ff_rotation_get_iterations(ff_stage* s, int* out);                     /* :129-132 */

/* ------------------------------ FlowFilter (gpu/flowfilter.h:31-145) */
/* FlowFilter(h, w, smoothIterations, maxflow, gamma) (src/gpu/flowfilter.cu:45-76) */
FF_API ff_status ff_flowfilter_create(int height, int width, int smoothIterations, float maxflow, float gamma,
                                      ff_stage** out);
/* FlowFilter(GPUImage inputImage) (src/gpu/flowfilter.cu:32-38) */
FF_API ff_status ff_flowfilter_create_from_image(const ff_gpuimage* inputImage, ff_stage** out);
FF_API ff_status ff_flowfilter_compute_image_model(ff_stage* s);  /* :155-162 */
FF_API ff_status ff_flowfilter_compute_propagation(ff_stage* s);  /* :165-172 */
FF_API ff_status ff_flowfilter_compute_update(ff_stage* s);       /* :175-197 */
FF_API ff_status ff_flowfilter_set_input_image(ff_stage* s, const ff_gpuimage* img); /* :200-216 */
FF_API ff_status ff_flowfilter_load_image(ff_stage* s, const ff_image_t* image);     /* :218-236 */
FF_API ff_status ff_flowfilter_download_flow(ff_stage* s, ff_image_t* flow);         /* :238-240 */
FF_API ff_status ff_flowfilter_download_image(ff_stage* s, ff_image_t* image);       /* :242-244 */
FF_API ff_status ff_flowfilter_get_flow(ff_stage* s, ff_gpuimage** out);             /* :266-268 */
FF_API ff_status ff_flowfilter_get_gamma(ff_stage* s, float* out);                   /* :271-273 */
FF_API ff_status ff_flowfilter_set_gamma(ff_stage* s, float gamma);                  /* :276-285 */
FF_API ff_status ff_flowfilter_get_maxflow(ff_stage* s, float* out);                 /* :288-290 */
FF_API ff_status ff_flowfilter_set_maxflow(ff_stage* s, float maxflow);              /* :293-296 */
FF_API ff_status ff_flowfilter_get_smooth_iterations(ff_stage* s, int* out);         /* :299-301 */
FF_API ff_status ff_flowfilter_set_smooth_iterations(ff_stage* s, int n);            /* :304-306 */
FF_API ff_status ff_flowfilter_set_propagation_border(ff_stage* s, int border);      /* :309-311 */
FF_API ff_status ff_flowfilter_get_propagation_border(ff_stage* s, int* out);        /* :314-316 */
FF_API ff_status ff_flowfilter_get_propagation_iterations(ff_stage* s, int* out);    /* :319-321 */
FF_API ff_status ff_flowfilter_height(ff_stage* s, int* out);                        /* :323-326 */
FF_API ff_status ff_flowfilter_width(ff_stage* s, int* out);                         /* :328-330 */
/* loadImage from device memory (no reference counterpart; enqueued, no wait) */
FF_API ff_status ff_flowfilter_load_image_device(ff_stage* s, const void* dev, size_t pitch);

/* ------------------------- DeltaFlowFilter (gpu/flowfilter.h:148-235) */
/* DeltaFlowFilter(inputImage, inputFlow) (src/gpu/flowfilter.cu:349-362) */
FF_API ff_status ff_deltaflowfilter_create(const ff_gpuimage* inputImage, const ff_gpuimage* inputFlow,
                                           ff_stage** out);
FF_API ff_status ff_deltaflowfilter_compute_image_model(ff_stage* s);   /* :460-467 */
FF_API ff_status ff_deltaflowfilter_compute_propagation(ff_stage* s);   /* :470-477 */
FF_API ff_status ff_deltaflowfilter_compute_update(ff_stage* s);        /* :480-503 */
FF_API ff_status ff_deltaflowfilter_set_input_image(ff_stage* s, const ff_gpuimage* img); /* :506-523 */
FF_API ff_status ff_deltaflowfilter_set_input_flow(ff_stage* s, const ff_gpuimage* img);  /* :526-542 */
FF_API ff_status ff_deltaflowfilter_get_flow(ff_stage* s, ff_gpuimage** out);             /* :545-547 */
FF_API ff_status ff_deltaflowfilter_get_image(ff_stage* s, ff_gpuimage** out);            /* :549-551 */
FF_API ff_status ff_deltaflowfilter_get_gamma(ff_stage* s, float* out);                   /* :554-556 */
FF_API ff_status ff_deltaflowfilter_set_gamma(ff_stage* s, float gamma);                  /* :559-568 */
FF_API ff_status ff_deltaflowfilter_get_maxflow(ff_stage* s, float* out);                 /* :571-573 */
FF_API ff_status ff_deltaflowfilter_set_maxflow(ff_stage* s, float maxflow);              /* :576-579 */
FF_API ff_status ff_deltaflowfilter_get_smooth_iterations(ff_stage* s, int* out);         /* :582-584 */
FF_API ff_status ff_deltaflowfilter_set_smooth_iterations(ff_stage* s, int n);            /* :587-589 */
FF_API ff_status ff_deltaflowfilter_set_propagation_border(ff_stage* s, int border);      /* :592-594 */
FF_API ff_status ff_deltaflowfilter_get_propagation_border(ff_stage* s, int* out);        /* :597-599 */
FF_API ff_status ff_deltaflowfilter_get_propagation_iterations(ff_stage* s, int* out);    /* :602-604 */
FF_API ff_status ff_deltaflowfilter_height(ff_stage* s, int* out);                        /* :607-610 */
FF_API ff_status ff_deltaflowfilter_width(ff_stage* s, int* out);                         /* :613-615 */

/* --------------------- PyramidalFlowFilter (gpu/flowfilter.h:238-332) */
/* PyramidalFlowFilter(h, w, levels) (src/gpu/flowfilter.cu:631-641) */
FF_API ff_status ff_pyramidal_create(int height, int width, int levels, ff_stage** out);
FF_API ff_status ff_pyramidal_get_flow(ff_stage* s, ff_gpuimage** out);              /* :718-725 */
FF_API ff_status ff_pyramidal_load_image(ff_stage* s, const ff_image_t* image);      /* :728-731 */
FF_API ff_status ff_pyramidal_download_flow(ff_stage* s, ff_image_t* flow);          /* :734-741 */
FF_API ff_status ff_pyramidal_download_image(ff_stage* s, ff_image_t* image);        /* :744-751 */
FF_API ff_status ff_pyramidal_get_gamma(ff_stage* s, int level, float* out);         /* :754-766 */
FF_API ff_status ff_pyramidal_set_gamma(ff_stage* s, int level, float gamma);        /* :769-781 */
FF_API ff_status ff_pyramidal_set_gamma_vector(ff_stage* s, const float* gamma, int n);        /* :784-794 */
FF_API ff_status ff_pyramidal_get_smooth_iterations(ff_stage* s, int level, int* out);         /* :797-809 */
FF_API ff_status ff_pyramidal_set_smooth_iterations(ff_stage* s, int level, int n);            /* :812-824 */
FF_API ff_status ff_pyramidal_set_smooth_iterations_vector(ff_stage* s, const int* iters, int n); /* :826-836 */
FF_API ff_status ff_pyramidal_get_maxflow(ff_stage* s, float* out);                  /* :839-846 */
FF_API ff_status ff_pyramidal_set_maxflow(ff_stage* s, float maxflow);               /* :849-866 */
FF_API ff_status ff_pyramidal_set_propagation_border(ff_stage* s, int border);       /* :869-877 */
FF_API ff_status ff_pyramidal_get_propagation_border(ff_stage* s, int* out);         /* :880-882 */
FF_API ff_status ff_pyramidal_height(ff_stage* s, int* out);                         /* :885-887 */
FF_API ff_status ff_pyramidal_width(ff_stage* s, int* out);                          /* :890-892 */
FF_API ff_status ff_pyramidal_levels(ff_stage* s, int* out);                         /* :895-897 */
/* loadImage from device memory (no reference counterpart; enqueued, no wait) */
FF_API ff_status ff_pyramidal_load_image_device(ff_stage* s, const void* dev, size_t pitch);
/* Frame in place (no reference counterpart): the next compute() reads its
 * frame from `dev` (u8, the filter's height x width, row pitch `pitch`
 * bytes) where it lies, without the copy into the input image; the memory
 * must hold the frame until that compute() has run on the filter's stream.
 * The filter copies it in instead on its first frame, for pyramids deeper
 * than three levels, on the staged path, when `dev` or `pitch` is not a
 * multiple of 4 or pitch * height reaches 2^31, and in a batch beyond its
 * first 128 frames in place. A later load_image* replaces it. */
FF_API ff_status ff_pyramidal_set_image_device(ff_stage* s, const void* dev, size_t pitch);

/* ----------------------------------------- execution controls (new) */
/* Select how compute() runs a filter: 1 = fused per-level kernels (the
 * default), 0 = one launch per reference stage kernel. Both produce
 * identical results. Applies to FlowFilter, DeltaFlowFilter and
 * PyramidalFlowFilter. */
FF_API ff_status ff_filter_set_fused(ff_stage* s, int fused);
FF_API ff_status ff_filter_get_fused(ff_stage* s, int* out);
/* PyramidalFlowFilter and batches: replay each fused frame as a captured
 * HIP graph (1, the default) or launch its kernels one by one (0). A single
 * filter's frames that are one launch (pyramids of up to three levels) are
 * launched directly either way: replaying the graph costs more than the
 * launch. A batch's step is replayed as a graph whenever this is on (about
 * 1 % of a 64-stream step). */
FF_API ff_status ff_filter_set_graphs(ff_stage* s, int graphs);
FF_API ff_status ff_filter_get_graphs(ff_stage* s, int* out);
/* PyramidalFlowFilter and batches (diagnostic, no reference counterpart):
 * number of times a level of a fused frame stopped waiting for the level
 * above it after the spin limit instead of seeing it complete; any nonzero
 * count means the flow of those frames is not valid. Waits for the filter's
 * stream. Such a wait is also reported as FF_ERR_RUNTIME by the next
 * ff_stage_elapsed_time, ff_pyramidal_download_flow or
 * ff_pyramidal_download_image of the filter (of every filter of the batch,
 * and ff_stage_elapsed_time of the batch). */
FF_API ff_status ff_filter_get_wait_timeouts(ff_stage* s, unsigned long long* out);
/* Test hook (no reference counterpart), PyramidalFlowFilter and batches:
 * spins > 0 makes the waits of later fused frames wait for a completion
 * count that no launch reaches, so each gives up after `spins` polls and
 * counts a timeout (every kernel still ends; the flow of those frames is
 * not valid); 0 restores the normal waits. For tests of the error path. */
FF_API ff_status ff_filter_set_wait_test(ff_stage* s, int spins);

/* Tile height of the fused frame kernel, in waves per workgroup (no
 * reference counterpart): 0 (the default) times the candidates 16, 15, 8, 5
 * and 4 on the first fused frame of a set of level shapes and keeps the
 * fastest; one of them fixes it. Results do not depend on it. get returns
 * the fixed or tuned value (0 before the first fused frame). FlowFilter,
 * DeltaFlowFilter, PyramidalFlowFilter and batches. */
FF_API ff_status ff_filter_set_tile_waves(ff_stage* s, int waves);
FF_API ff_status ff_filter_get_tile_waves(ff_stage* s, int* out);
/* Rows per thread of the fused kernel's tiles on a pyramid's top level
 * (which reads the input frame reduced while loading; no reference
 * counterpart): 0 (the default) tunes it with the tile height, 4 (more,
 * shorter tiles, which the level below waits for less on) or 8 (less halo)
 * fixes it. Results do not depend on it. get returns the fixed or tuned
 * value (0 before the first fused frame). Same filters as above. */
FF_API ff_status ff_filter_set_tile_top_rows(ff_stage* s, int rows);
FF_API ff_status ff_filter_get_tile_top_rows(ff_stage* s, int* out);
/* Rows per thread of the fused kernel's tiles on the other levels (no
 * reference counterpart): 0 (the default) tunes it with the tile height, 8
 * or 4 (16- or 15-wave workgroups only: the same tiles in half the
 * registers per thread, twice as many waves per CU) fixes it; with the tile
 * height and the top rows both fixed and this 0, 8. Results do not depend
 * on it. get returns the fixed or tuned value (0 before the first fused
 * frame). Same filters as above. */
FF_API ff_status ff_filter_set_tile_rows(ff_stage* s, int rows);
FF_API ff_status ff_filter_get_tile_rows(ff_stage* s, int* out);
/* Interior tiles of the fused kernel in a tile body of their own (no
 * reference counterpart): tiles inside the propagation border and away from
 * the image edges run one branch-free body, in a kernel instantiation of
 * its own, with 4 rows per thread on the levels below the top (fewer VALU
 * instructions: a 64-stream batch runs about 2 % faster). 0 (the default)
 * tunes it with the tiling, 1 turns it on (with 4 rows per thread and 16
 * or 15 waves per workgroup), 2 off. Results do not depend on it. get
 * returns the fixed or tuned choice, 1 or 2 (0 before the first fused
 * frame). Same filters as above. */
FF_API ff_status ff_filter_set_tile_interior(ff_stage* s, int mode);
FF_API ff_status ff_filter_get_tile_interior(ff_stage* s, int* out);

/* Batched filters (no reference counterpart; replaces a caller's loop of
 * PyramidalFlowFilter::compute over camera streams, e.g. the per-camera
 * filters of demos/highSpeedDemo/src/main.cpp:140-142). A batch holds
 * shared references to PyramidalFlowFilters of any size and depth; each
 * ff_stage_compute(batch) advances every filter by one frame with a single
 * fused frame launch, replayed as a HIP graph (ff_filter_set_graphs and
 * ff_filter_get_wait_timeouts accept a batch). Results are identical to
 * calling ff_stage_compute on each filter. Frames are loaded, and results
 * read, through the filters' own handles; ff_stage_elapsed_time(batch) is
 * the time of the whole step. Destroy with ff_stage_destroy; a filter stays
 * valid while its handle or a batch holding it exists. */
FF_API ff_status ff_batch_create(ff_stage** out);
FF_API ff_status ff_batch_add(ff_stage* batch, ff_stage* pyramidal_filter);
FF_API ff_status ff_batch_size(ff_stage* batch, int* out);
/* Load the next frame of every filter of the batch (in the order added)
 * from device memory: frames[i] (u8, row pitch pitches[i] bytes, the
 * filter's height x width) into filter i's input image, with one copy
 * launch for all of them; n must equal the batch size. Enqueued on the
 * library stream, like ff_pyramidal_load_image_device. */
FF_API ff_status ff_batch_load_images_device(ff_stage* batch, const void* const* frames, const size_t* pitches, int n);
/* The next ff_stage_compute(batch) reads every filter's frame in place:
 * ff_pyramidal_set_image_device of filter i with frames[i], pitches[i]. The
 * batched launch reads up to 128 such frames where they lie (no copy
 * launch); the memory must hold the frames until that compute has run on
 * the batch's stream. n must equal the batch size. */
FF_API ff_status ff_batch_set_images_device(ff_stage* batch, const void* const* frames, const size_t* pitches, int n);
/* Workgroups of the last fused launch (diagnostic). */
FF_API ff_status ff_batch_get_blocks(ff_stage* batch, int* out);

/* ---------------------------------------------------------------------
 * Differentiable form (no reference counterpart). Batched stages of the
 * filter on float images and their adjoints (vector-Jacobian products),
 * for the autograd function of the Python package
 * (flowfilter.autograd). All arrays are device memory, B images of H x W,
 * contiguous: scalar fields float[B][H][W], vector fields (flow, image
 * gradient) float2[B][H][W]. `stream` is a hipStream_t (NULL: the null
 * stream). Forward stages compute what the corresponding stage kernels
 * compute (same arithmetic); each *_backward takes the forward inputs and
 * the gradient of the outputs and writes the gradient of the inputs.
 * `accumulate`: for a backward with one output, nonzero adds it to the
 * destination; with several outputs it is a bit mask, bit k adding output
 * k (in argument order, gamma partial sums and g_upper_pixel excluded).
 * An output gradient passed as NULL where noted reads as zero. Upwind
 * directions, clamps and the zeroing of non-finite values are held at
 * their forward values.
 * ------------------------------------------------------------------- */
/* Number of per-block partial sums per image written by the update
 * backwards' gamma gradient (g_gamma_partial holds B times this). */
FF_API ff_status ff_ad_partial_count(int height, int width, int* out);
/* Image model (imagemodel_k.cu:31-133): a0 and grad (gx, gy) of a float
 * image. Backward: work holds 2*B*H*W floats. */
FF_API ff_status ff_ad_image_model(const float* image, float* a0, float* grad, int B, int H, int W, void* stream);
FF_API ff_status ff_ad_image_model_backward(const float* g_a0, const float* g_grad, float* g_image, float* work, int B,
                                            int H, int W, int accumulate, void* stream);
/* One float pyramid step (pyramid_k.cu:81-145): half is (H/2) x (W/2);
 * work holds B*H*(W/2) floats. */
FF_API ff_status ff_ad_down(const float* image, float* half, float* work, int B, int H, int W, void* stream);
FF_API ff_status ff_ad_down_backward(const float* g_half, float* g_image, float* work, int B, int H, int W,
                                     int accumulate, void* stream);
/* One upwind pass along axis 0 (x) or 1 (y) (propagation_k.cu:16-334) of
 * the flow and, when scalar/vector are given, of a scalar and a vector
 * payload; border >= 1. Backward: g_flow_out, g_scalar_out and g_vector_out
 * may be NULL (zero). */
FF_API ff_status ff_ad_propagate_pass(int axis, const float* flow, const float* scalar, const float* vector,
                                      float* flow_out, float* scalar_out, float* vector_out, int B, int H, int W,
                                      float dt, int border, void* stream);
/* An X pass and then a Y pass in one launch (the same results as two
 * ff_ad_propagate_pass calls). With flow_mid (and scalar_mid, vector_mid
 * when there are payloads) the X pass's result is written too. */
FF_API ff_status ff_ad_propagate_pair(const float* flow, const float* scalar, const float* vector, float* flow_mid,
                                      float* scalar_mid, float* vector_mid, float* flow_out, float* scalar_out,
                                      float* vector_out, int B, int H, int W, float dt, int border, void* stream);
/* Two iterations -- X, Y, X and Y passes -- in one launch (the same results
 * as two ff_ad_propagate_pair calls), for forwards that keep no
 * intermediates; no output may be its input. */
FF_API ff_status ff_ad_propagate_quad(const float* flow, const float* scalar, const float* vector, float* flow_out,
                                      float* scalar_out, float* vector_out, int B, int H, int W, float dt, int border,
                                      void* stream);
/* Adjoint of ff_ad_propagate_quad given the quad's input only (its three
 * inner passes are recomputed inside the launch); bitwise the adjoints of
 * the two pairs. g_*_out may be NULL (zero); accumulate as for the pass
 * backward; no input gradient may be an output gradient. */
FF_API ff_status ff_ad_propagate_quad_backward(const float* flow, const float* scalar, const float* vector,
                                               const float* g_flow_out, const float* g_scalar_out,
                                               const float* g_vector_out, float* g_flow, float* g_scalar,
                                               float* g_vector, int B, int H, int W, float dt, int border,
                                               int accumulate, void* stream);
/* Adjoint of ff_ad_propagate_pair, given the pair's input and the X pass's
 * result (flow_mid ...); the same results as the two passes' adjoints.
 * With flow_mid, scalar_mid and vector_mid all NULL the X pass's result is
 * recomputed from the pair's input inside the launch (bitwise the same
 * gradients; the forward then need not keep it). g_*_out may be NULL
 * (zero); accumulate as for the pass backward. */
FF_API ff_status ff_ad_propagate_pair_backward(const float* flow, const float* scalar, const float* vector,
                                               const float* flow_mid, const float* scalar_mid,
                                               const float* vector_mid, const float* g_flow_out,
                                               const float* g_scalar_out, const float* g_vector_out, float* g_flow,
                                               float* g_scalar, float* g_vector, int B, int H, int W, float dt,
                                               int border, int accumulate, void* stream);
FF_API ff_status ff_ad_propagate_pass_backward(int axis, const float* flow, const float* scalar, const float* vector,
                                               const float* g_flow_out, const float* g_scalar_out,
                                               const float* g_vector_out, float* g_flow, float* g_scalar,
                                               float* g_vector, int B, int H, int W, float dt, int border,
                                               int accumulate, void* stream);
/* Top-level update (update_k.cu:15-79). Backward writes the gradients of
 * a0, grad, a0_old and flow and B*ff_ad_partial_count partial sums of the
 * gamma gradient. */
FF_API ff_status ff_ad_update(const float* a0, const float* grad, const float* a0_old, const float* flow,
                              float* flow_out, int B, int H, int W, float gamma, float maxflow, void* stream);
FF_API ff_status ff_ad_update_backward(const float* a0, const float* grad, const float* a0_old, const float* flow,
                                       const float* g_flow_out, float* g_a0, float* g_grad, float* g_a0_old,
                                       float* g_flow, float* g_gamma_partial, int B, int H, int W, float gamma,
                                       float maxflow, int accumulate, void* stream);
/* Allocation-free form of ff_ad_update_backward. It has no counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results, except that the gamma gradient's
 * per-workgroup partial sums stay in a library-owned device array and are
 * summed in the library, in double and in a fixed order, into *g_gamma (one
 * double in device memory; bit 4 of accumulate adds to it instead of
 * overwriting it). The array holds 1,048,576 partial sums (4 MiB on each GPU
 * the library runs on, kept until the process ends); B *
 * ff_ad_partial_count(H, W) above that is FF_ERR_INVALID_ARGUMENT. The call
 * allocates nothing. Every call on one GPU uses that GPU's array: calls that
 * overlap on one GPU (different streams, no ordering between them) race. */
FF_API ff_status ff_ad_update_backward_noalloc(const float* a0, const float* grad, const float* a0_old,
                                               const float* flow, const float* g_flow_out, float* g_a0, float* g_grad,
                                               float* g_a0_old, float* g_flow, double* g_gamma, int B, int H, int W,
                                               float gamma, float maxflow, int accumulate, void* stream);
/* Delta-level update (update_k.cu:82-173); upper is the upper level's flow
 * (Hu x Wu), xmap/ymap the upper texel index of every column/row.
 * Backward leaves the upper-flow gradient per pixel of this level in
 * g_upper_pixel (float2[B][H][W]); ff_ad_upper_backward sums it onto the
 * upper level using, for every upper column X (row Y), the range
 * [xlo[X], xhi[X]) ([ylo[Y], yhi[Y])) of this level's columns (rows) that
 * map to it. g_dflow_out may be NULL (zero). */
FF_API ff_status ff_ad_delta_update(const float* a0, const float* grad, const float* a0_prop, const float* dflow_prop,
                                    const float* upper, const int* xmap, const int* ymap, float* dflow_out,
                                    float* flow_out, int B, int H, int W, int Hu, int Wu, float gamma, float maxflow,
                                    void* stream);
FF_API ff_status ff_ad_delta_update_backward(const float* a0, const float* grad, const float* a0_prop,
                                             const float* dflow_prop, const float* upper, const int* xmap,
                                             const int* ymap, const float* g_dflow_out, const float* g_flow_out,
                                             float* g_a0, float* g_grad, float* g_a0_prop, float* g_dflow_prop,
                                             float* g_upper_pixel, float* g_gamma_partial, int B, int H, int W, int Hu,
                                             int Wu, float gamma, float maxflow, int accumulate, void* stream);
FF_API ff_status ff_ad_upper_backward(const float* g_upper_pixel, const int* xlo, const int* xhi, const int* ylo,
                                      const int* yhi, float* g_upper, int B, int H, int W, int Hu, int Wu,
                                      int accumulate, void* stream);
/* One box-smoothing pass along axis 0 (x) or 1 (y) (flowsmoothing_k.cu:18-87).
 * Backward of the pass for the output gradient g_flow_out + g_flow_out2
 * (g_flow_out2 may be NULL). */
FF_API ff_status ff_ad_smooth_pass(int axis, const float* flow, float* flow_out, int B, int H, int W, void* stream);
FF_API ff_status ff_ad_smooth_pass_backward(int axis, const float* g_flow_out, const float* g_flow_out2, float* g_flow,
                                            int B, int H, int W, int accumulate, void* stream);
/* One smoothing iteration, an X pass and then a Y pass, in one launch
 * (flowsmoothing_k.cu:18-87 as FlowSmoother runs it, flowsmoothing.cu:83-104;
 * the same results as two ff_ad_smooth_pass calls). flow_out must not be
 * flow. Backward: the Y pass's adjoint and then the X pass's, for the
 * output gradient g_flow_out + g_flow_out2 (g_flow_out2 may be NULL); g_flow
 * must be neither of them. */
FF_API ff_status ff_ad_smooth_pair(const float* flow, float* flow_out, int B, int H, int W, void* stream);
FF_API ff_status ff_ad_smooth_pair_backward(const float* g_flow_out, const float* g_flow_out2, float* g_flow, int B,
                                            int H, int W, int accumulate, void* stream);
/* Photometric term of the differentiable window's loss, fused (no
 * reference counterpart; it replaces the grid_sample form of
 * flowfilter.autograd.photometric_error). Frames [B][T][H][W] float, flows
 * [B][T][H][W] float2, T >= 2: for every frame t >= 1, frame t-1 is sampled
 * bilinearly at x - flow_t(x) with the coordinates clamped into the image,
 * and partial (ff_ad_photometric_partial_count floats, one per workgroup)
 * receives the sums of the squared differences to frame t. Backward: for the
 * loss scale *scale (device float, d loss / d sum of squares) adds the
 * gradients into g_frames and g_flows (flows of frame 0 untouched); the
 * sampled pixels' gradients are atomic adds, so their summation order
 * varies from run to run. */
FF_API ff_status ff_ad_photometric_partial_count(int B, int T, int height, int width, int* out);
FF_API ff_status ff_ad_photometric(const float* frames, const float* flows, float* partial, int B, int T, int H, int W,
                                   void* stream);
FF_API ff_status ff_ad_photometric_backward(const float* frames, const float* flows, const float* scale,
                                            float* g_frames, float* g_flows, int B, int T, int H, int W,
                                            void* stream);

/* Kernel profiling (no reference counterpart; the reference times whole
 * stages only, src/gpu/pipeline.cu:63-89). Between begin and end, every
 * kernel launch whose name contains `filter` is bracketed by HIP events on
 * its stream; end waits for them and returns the summed milliseconds and
 * the number of launches. Kernel names: image_model, propagate_x/_y,
 * propagate_payload_x/_y, update, delta_update, smooth_x/_y, down_x/_y,
 * scale_f2, copy2d, fused_top, fused_delta. */
FF_API ff_status ff_profile_begin(const char* filter);
FF_API ff_status ff_profile_end(double* total_ms, long long* count);

/* Stream markers (no reference counterpart): a timing event recorded on the
 * library stream; elapsed waits for `stop` and returns the milliseconds
 * between the two markers. For timing loops of compute() calls. */
typedef struct ff_marker ff_marker;
FF_API ff_status ff_marker_record(ff_marker** out);
/* The same on the stream of stage s (ff_stage_get_stream). */
FF_API ff_status ff_marker_record_on(ff_stage* s, ff_marker** out);
FF_API ff_status ff_marker_elapsed(ff_marker* start, ff_marker* stop, float* ms);
FF_API ff_status ff_marker_destroy(ff_marker* m);

#ifdef __cplusplus
}
#endif

#endif /* FLOWFILTER_AMD_H_ */
