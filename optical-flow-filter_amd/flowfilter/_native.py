"""ctypes binding of libflowfilter_amd.so (the C ABI in include/flowfilter_amd.h).

The library is loaded from this package's parent directory, where the build
places it. There is no fallback: if the shared library is missing or fails to
load, importing any ``flowfilter.gpu`` module raises ImportError.
"""
import ctypes
import os

from .errors import raise_for_status

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FLOWFILTER_AMD_LIB", os.path.join(os.path.dirname(_HERE), "libflowfilter_amd.so"))


class image_t(ctypes.Structure):
    """ff_image_t / flowfilter::image_t (include/flowfilter/image.h:30-49)."""
    _fields_ = [("height", ctypes.c_int), ("width", ctypes.c_int), ("depth", ctypes.c_int),
                ("pitch", ctypes.c_size_t), ("itemSize", ctypes.c_size_t), ("data", ctypes.c_void_p)]


class ff_camera(ctypes.Structure):
    """ff_camera / flowfilter::gpu::perspectiveCamera (camera.h:20-25)."""
    _fields_ = [("alphaX", ctypes.c_float), ("alphaY", ctypes.c_float), ("centerX", ctypes.c_float),
                ("centerY", ctypes.c_float)]


_S = ctypes.c_void_p             # ff_stage*
_I = ctypes.c_void_p             # ff_gpuimage*
_pS = ctypes.POINTER(ctypes.c_void_p)
_pI = ctypes.POINTER(ctypes.c_void_p)
_i, _f, _z, _v = ctypes.c_int, ctypes.c_float, ctypes.c_size_t, ctypes.c_void_p
_pi, _pf, _pv = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_void_p)
_h = ctypes.POINTER(image_t)

# name -> argument types; every function returns ff_status (int) except
# ff_last_error.
SIGNATURES = {
    "ff_version": [_pi, _pi],
    "ff_synchronize": [],
    "ff_set_stream": [_v],
    "ff_get_stream": [_pv],
    "ff_create_image": [_i, _i, _i, _z, _h],
    "ff_destroy_image": [_h],
    # GPUImage
    "ff_gpuimage_create": [_i, _i, _i, _i, _pI],
    "ff_gpuimage_create_empty": [_pI],
    "ff_gpuimage_share": [_I, _pI],
    "ff_gpuimage_destroy": [_I],
    "ff_gpuimage_height": [_I, _pi],
    "ff_gpuimage_width": [_I, _pi],
    "ff_gpuimage_depth": [_I, _pi],
    "ff_gpuimage_pitch": [_I, _pi],
    "ff_gpuimage_item_size": [_I, _pi],
    "ff_gpuimage_data": [_I, _pv],
    "ff_gpuimage_upload": [_I, _h],
    "ff_gpuimage_download": [_I, _h],
    "ff_gpuimage_copy_from": [_I, _I],
    "ff_gpuimage_clear": [_I],
    "ff_gpuimage_upload_device": [_I, _v, _z],
    "ff_gpuimage_download_device": [_I, _v, _z],
    # Stage
    "ff_gpuimage_to_dlpack": [_I, _v, _i, _pv],
    "ff_dlpack_delete": [_v],
    # FlowToColor
    "ff_get_color_wheel_rgba": [_h],
    "ff_flowtocolor_create": [_I, _f, _pS],
    "ff_flowtocolor_set_input_flow": [_S, _I],
    "ff_flowtocolor_get_color_flow": [_S, _pI],
    "ff_flowtocolor_get_max_flow": [_S, _pf],
    "ff_flowtocolor_set_max_flow": [_S, _f],
    "ff_flowtocolor_download_color_flow": [_S, _h],
    # camera, rotation, Lax-Wendroff
    "ff_create_perspective_camera": [_f, _i, _i, _f, _f, ctypes.POINTER(ff_camera)],
    "ff_laxwendroff_create": [_I, _I, _pS],
    "ff_laxwendroff_set_iterations": [_S, _i],
    "ff_laxwendroff_get_iterations": [_S, _pi],
    "ff_laxwendroff_get_dt": [_S, _pf],
    "ff_laxwendroff_set_input_flow": [_S, _I],
    "ff_laxwendroff_set_input_image": [_S, _I],
    "ff_laxwendroff_get_flow": [_S, _pI],
    "ff_laxwendroff_get_propagated_image": [_S, _pI],
    "ff_rotation_create": [ctypes.POINTER(ff_camera), _I, _pS],
    "ff_rotation_set_input_image": [_S, _I],
    "ff_rotation_get_predicted_image": [_S, _pI],
    "ff_rotation_get_optical_flow": [_S, _pI],
    "ff_rotation_set_camera": [_S, ctypes.POINTER(ff_camera)],
    "ff_rotation_get_camera": [_S, ctypes.POINTER(ff_camera)],
    "ff_rotation_set_angular_velocity": [_S, _f, _f, _f],
    "ff_rotation_set_iterations": [_S, _i],
    "ff_rotation_get_iterations": [_S, _pi],
    "ff_stage_configure": [_S],
    "ff_stage_compute": [_S],
    "ff_stage_elapsed_time": [_S, _pf],
    "ff_stage_destroy": [_S],
    "ff_stage_get_stream": [_S, _pv],
    "ff_stage_set_stream": [_S, _v],
    # ImageModel
    "ff_imagemodel_create": [_I, _pS],
    "ff_imagemodel_set_input_image": [_S, _I],
    "ff_imagemodel_get_image_constant": [_S, _pI],
    "ff_imagemodel_get_image_gradient": [_S, _pI],
    # FlowPropagator
    "ff_flowpropagator_create": [_I, _i, _pS],
    "ff_flowpropagator_set_iterations": [_S, _i],
    "ff_flowpropagator_get_iterations": [_S, _pi],
    "ff_flowpropagator_get_dt": [_S, _pf],
    "ff_flowpropagator_set_border": [_S, _i],
    "ff_flowpropagator_get_border": [_S, _pi],
    "ff_flowpropagator_set_invert_input_flow": [_S, _i],
    "ff_flowpropagator_get_invert_input_flow": [_S, _pi],
    "ff_flowpropagator_set_input_flow": [_S, _I],
    "ff_flowpropagator_get_propagated_flow": [_S, _pI],
    # FlowPropagatorPayload
    "ff_flowpropagatorpayload_create": [_I, _I, _I, _i, _pS],
    "ff_flowpropagatorpayload_set_iterations": [_S, _i],
    "ff_flowpropagatorpayload_get_iterations": [_S, _pi],
    "ff_flowpropagatorpayload_get_dt": [_S, _pf],
    "ff_flowpropagatorpayload_set_border": [_S, _i],
    "ff_flowpropagatorpayload_get_border": [_S, _pi],
    "ff_flowpropagatorpayload_set_input_flow": [_S, _I],
    "ff_flowpropagatorpayload_set_scalar_payload": [_S, _I],
    "ff_flowpropagatorpayload_set_vector_payload": [_S, _I],
    "ff_flowpropagatorpayload_get_propagated_flow": [_S, _pI],
    "ff_flowpropagatorpayload_get_propagated_scalar": [_S, _pI],
    "ff_flowpropagatorpayload_get_propagated_vector": [_S, _pI],
    # FlowUpdate
    "ff_flowupdate_create": [_I, _I, _I, _f, _f, _pS],
    "ff_flowupdate_get_gamma": [_S, _pf],
    "ff_flowupdate_set_gamma": [_S, _f],
    "ff_flowupdate_get_maxflow": [_S, _pf],
    "ff_flowupdate_set_maxflow": [_S, _f],
    "ff_flowupdate_set_input_flow": [_S, _I],
    "ff_flowupdate_set_input_image": [_S, _I],
    "ff_flowupdate_set_input_image_gradient": [_S, _I],
    "ff_flowupdate_get_updated_flow": [_S, _pI],
    "ff_flowupdate_get_updated_image": [_S, _pI],
    # DeltaFlowUpdate
    "ff_deltaflowupdate_create": [_I, _I, _I, _I, _I, _f, _f, _pS],
    "ff_deltaflowupdate_get_gamma": [_S, _pf],
    "ff_deltaflowupdate_set_gamma": [_S, _f],
    "ff_deltaflowupdate_get_maxflow": [_S, _pf],
    "ff_deltaflowupdate_set_maxflow": [_S, _f],
    "ff_deltaflowupdate_set_input_flow": [_S, _I],
    "ff_deltaflowupdate_set_input_delta_flow": [_S, _I],
    "ff_deltaflowupdate_set_input_image_old": [_S, _I],
    "ff_deltaflowupdate_set_input_image": [_S, _I],
    "ff_deltaflowupdate_set_input_image_gradient": [_S, _I],
    "ff_deltaflowupdate_get_updated_flow": [_S, _pI],
    "ff_deltaflowupdate_get_updated_delta_flow": [_S, _pI],
    "ff_deltaflowupdate_get_updated_image": [_S, _pI],
    # FlowSmoother
    "ff_flowsmoother_create": [_I, _i, _pS],
    "ff_flowsmoother_get_iterations": [_S, _pi],
    "ff_flowsmoother_set_iterations": [_S, _i],
    "ff_flowsmoother_set_input_flow": [_S, _I],
    "ff_flowsmoother_get_smoothed_flow": [_S, _pI],
    # ImagePyramid
    "ff_imagepyramid_create": [_I, _i, _pS],
    "ff_imagepyramid_set_input_image": [_S, _I],
    "ff_imagepyramid_set_levels": [_S, _i],
    "ff_imagepyramid_get_levels": [_S, _pi],
    "ff_imagepyramid_get_image": [_S, _i, _pI],
    # FlowFilter
    "ff_flowfilter_create": [_i, _i, _i, _f, _f, _pS],
    "ff_flowfilter_create_from_image": [_I, _pS],
    "ff_flowfilter_compute_image_model": [_S],
    "ff_flowfilter_compute_propagation": [_S],
    "ff_flowfilter_compute_update": [_S],
    "ff_flowfilter_set_input_image": [_S, _I],
    "ff_flowfilter_load_image": [_S, _h],
    "ff_flowfilter_download_flow": [_S, _h],
    "ff_flowfilter_download_image": [_S, _h],
    "ff_flowfilter_get_flow": [_S, _pI],
    "ff_flowfilter_get_gamma": [_S, _pf],
    "ff_flowfilter_set_gamma": [_S, _f],
    "ff_flowfilter_get_maxflow": [_S, _pf],
    "ff_flowfilter_set_maxflow": [_S, _f],
    "ff_flowfilter_get_smooth_iterations": [_S, _pi],
    "ff_flowfilter_set_smooth_iterations": [_S, _i],
    "ff_flowfilter_set_propagation_border": [_S, _i],
    "ff_flowfilter_get_propagation_border": [_S, _pi],
    "ff_flowfilter_get_propagation_iterations": [_S, _pi],
    "ff_flowfilter_height": [_S, _pi],
    "ff_flowfilter_width": [_S, _pi],
    "ff_flowfilter_load_image_device": [_S, _v, _z],
    # DeltaFlowFilter
    "ff_deltaflowfilter_create": [_I, _I, _pS],
    "ff_deltaflowfilter_compute_image_model": [_S],
    "ff_deltaflowfilter_compute_propagation": [_S],
    "ff_deltaflowfilter_compute_update": [_S],
    "ff_deltaflowfilter_set_input_image": [_S, _I],
    "ff_deltaflowfilter_set_input_flow": [_S, _I],
    "ff_deltaflowfilter_get_flow": [_S, _pI],
    "ff_deltaflowfilter_get_image": [_S, _pI],
    "ff_deltaflowfilter_get_gamma": [_S, _pf],
    "ff_deltaflowfilter_set_gamma": [_S, _f],
    "ff_deltaflowfilter_get_maxflow": [_S, _pf],
    "ff_deltaflowfilter_set_maxflow": [_S, _f],
    "ff_deltaflowfilter_get_smooth_iterations": [_S, _pi],
    "ff_deltaflowfilter_set_smooth_iterations": [_S, _i],
    "ff_deltaflowfilter_set_propagation_border": [_S, _i],
    "ff_deltaflowfilter_get_propagation_border": [_S, _pi],
    "ff_deltaflowfilter_get_propagation_iterations": [_S, _pi],
    "ff_deltaflowfilter_height": [_S, _pi],
    "ff_deltaflowfilter_width": [_S, _pi],
    # PyramidalFlowFilter
    "ff_pyramidal_create": [_i, _i, _i, _pS],
    "ff_pyramidal_get_flow": [_S, _pI],
    "ff_pyramidal_load_image": [_S, _h],
    "ff_pyramidal_download_flow": [_S, _h],
    "ff_pyramidal_download_image": [_S, _h],
    "ff_pyramidal_get_gamma": [_S, _i, _pf],
    "ff_pyramidal_set_gamma": [_S, _i, _f],
    "ff_pyramidal_set_gamma_vector": [_S, _pf, _i],
    "ff_pyramidal_get_smooth_iterations": [_S, _i, _pi],
    "ff_pyramidal_set_smooth_iterations": [_S, _i, _i],
    "ff_pyramidal_set_smooth_iterations_vector": [_S, _pi, _i],
    "ff_pyramidal_get_maxflow": [_S, _pf],
    "ff_pyramidal_set_maxflow": [_S, _f],
    "ff_pyramidal_set_propagation_border": [_S, _i],
    "ff_pyramidal_get_propagation_border": [_S, _pi],
    "ff_pyramidal_height": [_S, _pi],
    "ff_pyramidal_width": [_S, _pi],
    "ff_pyramidal_levels": [_S, _pi],
    "ff_pyramidal_load_image_device": [_S, _v, _z],
    "ff_pyramidal_set_image_device": [_S, _v, _z],
    # execution controls
    "ff_filter_set_fused": [_S, _i],
    "ff_filter_get_fused": [_S, _pi],
    "ff_filter_set_graphs": [_S, _i],
    "ff_filter_get_graphs": [_S, _pi],
    "ff_filter_get_wait_timeouts": [_S, ctypes.POINTER(ctypes.c_ulonglong)],
    "ff_filter_set_wait_test": [_S, _i],
    "ff_filter_set_tile_waves": [_S, _i],
    "ff_filter_get_tile_waves": [_S, _pi],
    "ff_filter_set_tile_top_rows": [_S, _i],
    "ff_filter_get_tile_top_rows": [_S, _pi],
    "ff_filter_set_tile_rows": [_S, _i],
    "ff_filter_get_tile_rows": [_S, _pi],
    "ff_filter_set_tile_interior": [_S, _i],
    "ff_filter_get_tile_interior": [_S, _pi],
    "ff_batch_create": [_pS],
    "ff_ad_partial_count": [_i, _i, _pi],
    "ff_ad_image_model": [_v, _v, _v, _i, _i, _i, _v],
    "ff_ad_image_model_backward": [_v, _v, _v, _v, _i, _i, _i, _i, _v],
    "ff_ad_down": [_v, _v, _v, _i, _i, _i, _v],
    "ff_ad_down_backward": [_v, _v, _v, _i, _i, _i, _i, _v],
    "ff_ad_propagate_pass": [_i, _v, _v, _v, _v, _v, _v, _i, _i, _i, _f, _i, _v],
    "ff_ad_propagate_pair": [_v, _v, _v, _v, _v, _v, _v, _v, _v, _i, _i, _i, _f, _i, _v],
    "ff_ad_propagate_quad": [_v, _v, _v, _v, _v, _v, _i, _i, _i, _f, _i, _v],
    "ff_ad_propagate_quad_backward": [_v, _v, _v, _v, _v, _v, _v, _v, _v, _i, _i, _i, _f, _i, _i, _v],
    "ff_ad_propagate_pair_backward": [_v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _i, _i, _i, _f, _i, _i, _v],
    "ff_ad_propagate_pass_backward": [_i, _v, _v, _v, _v, _v, _v, _v, _v, _v, _i, _i, _i, _f, _i, _i, _v],
    "ff_ad_update": [_v, _v, _v, _v, _v, _i, _i, _i, _f, _f, _v],
    "ff_ad_update_backward": [_v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _i, _i, _i, _f, _f, _i, _v],
    "ff_ad_update_backward_noalloc": [_v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _i, _i, _i, _f, _f, _i, _v],
    "ff_ad_delta_update": [_v, _v, _v, _v, _v, _v, _v, _v, _v, _i, _i, _i, _i, _i, _f, _f, _v],
    "ff_ad_delta_update_backward": [_v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _v, _v,
                                    _i, _i, _i, _i, _i, _f, _f, _i, _v],
    "ff_ad_upper_backward": [_v, _v, _v, _v, _v, _v, _i, _i, _i, _i, _i, _i, _v],
    "ff_ad_smooth_pass": [_i, _v, _v, _i, _i, _i, _v],
    "ff_ad_smooth_pass_backward": [_i, _v, _v, _v, _i, _i, _i, _i, _v],
    "ff_ad_smooth_pair": [_v, _v, _i, _i, _i, _v],
    "ff_ad_smooth_pair_backward": [_v, _v, _v, _i, _i, _i, _i, _v],
    "ff_ad_photometric_partial_count": [_i, _i, _i, _i, _pi],
    "ff_ad_photometric": [_v, _v, _v, _i, _i, _i, _i, _v],
    "ff_ad_photometric_backward": [_v, _v, _v, _v, _v, _i, _i, _i, _i, _v],
    "ff_batch_add": [_S, _S],
    "ff_batch_size": [_S, _pi],
    "ff_batch_load_images_device": [_S, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t), _i],
    "ff_batch_set_images_device": [_S, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t), _i],
    "ff_batch_get_blocks": [_S, _pi],
    # profiling
    "ff_profile_begin": [ctypes.c_char_p],
    "ff_profile_end": [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)],
    # stream markers
    "ff_marker_record": [_pv],
    "ff_marker_record_on": [_S, _pv],
    "ff_marker_elapsed": [_v, _v, _pf],
    "ff_marker_destroy": [_v],
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libflowfilter_amd.so not found at %s: build it first "
                          "(python -c 'import __graft_entry__ as g; g.build()' or make -C csrc)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.ff_last_error.restype = ctypes.c_char_p
    lib.ff_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    return lib


lib = _load()


def call(name, *args):
    """Call ``name`` and raise the mapped Python exception on a non-zero status."""
    st = getattr(lib, name)(*args)
    if st != 0:
        raise_for_status(st, lib.ff_last_error().decode("utf-8", "replace"))


def get_int(name, *args):
    v = ctypes.c_int()
    call(name, *args, ctypes.byref(v))
    return v.value


def get_float(name, *args):
    v = ctypes.c_float()
    call(name, *args, ctypes.byref(v))
    return v.value


def get_handle(name, *args):
    v = ctypes.c_void_p()
    call(name, *args, ctypes.byref(v))
    return v.value


def synchronize():
    call("ff_synchronize")


def set_stream(stream_ptr):
    """Route library work onto an existing hipStream_t (int pointer or None)."""
    call("ff_set_stream", ctypes.c_void_p(stream_ptr) if stream_ptr else None)


def get_stream():
    return get_handle("ff_get_stream")


class Marker(object):
    """A timing event recorded at construction on the stream of ``stage`` (a
    filter or batch object, or its handle), or on the library stream."""

    def __init__(self, stage=None):
        if stage is None:
            self._h = get_handle("ff_marker_record")
        else:
            self._h = get_handle("ff_marker_record_on", getattr(stage, "handle", stage))

    def elapsed_ms(self, stop):
        """Milliseconds from this marker to ``stop`` (waits for ``stop``)."""
        v = ctypes.c_float()
        call("ff_marker_elapsed", ctypes.c_void_p(self._h), ctypes.c_void_p(stop._h), ctypes.byref(v))
        return v.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib.ff_marker_destroy(ctypes.c_void_p(h))
            except Exception:
                pass
            self._h = None


def profile_begin(kernel_filter):
    """Start timing every launch of kernels whose name contains ``kernel_filter``."""
    call("ff_profile_begin", kernel_filter.encode())


def profile_end():
    """Stop profiling; returns (total_ms, launches) of the matched kernels."""
    t, n = ctypes.c_double(), ctypes.c_longlong()
    call("ff_profile_end", ctypes.byref(t), ctypes.byref(n))
    return t.value, n.value
