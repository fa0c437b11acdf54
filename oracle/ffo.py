"""numpy front end of the CPU oracle (oracle/libfforacle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library. Imported by
tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the
product package never imports it.

Every function takes a ``prec`` argument: ``"f32"`` is the parity oracle
(same operations and rounding points as the reference kernels), ``"f64"``
the FP64 oracle (same algorithm in double precision).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfforacle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_c_int, _c_float, _vp = ctypes.c_int, ctypes.c_float, ctypes.c_void_p


def lib():
    """Load (building if needed) the oracle shared library."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE])
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.ffo_pyramidal_create_f32.restype = _vp
        _lib.ffo_pyramidal_create_f64.restype = _vp
        _lib.ffo_flowfilter_create_f32.restype = _vp
        _lib.ffo_flowfilter_create_f64.restype = _vp
        _lib.ffo_deltafilter_create_f32.restype = _vp
        _lib.ffo_deltafilter_create_f64.restype = _vp
        _lib.ffo_down_u8_f32.restype = ctypes.c_uint8
        _lib.ffo_down_u8_f64.restype = ctypes.c_uint8
    return _lib


def set_threads(n):
    lib().ffo_set_threads(int(n))


# Planted mistakes (ffo::Mutation in ff_oracle.hpp), for the tests that show
# a pinning check fails on a wrong oracle.
MUT_NONE = 0
MUT_SMOOTHED_TOP_FEED = 1
MUT_DESCENDING_UPDATE = 2
MUT_TOP_UPDATES_LAST = 3
MUT_PYR_TAP_SHIFT = 4
MUT_PYR_SUM_DOUBLE = 5


class mutated:
    """Context manager: run the oracle with planted mistake ``m``."""

    def __init__(self, m):
        self.m = int(m)

    def __enter__(self):
        self.prev = lib().ffo_get_mutation()
        lib().ffo_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().ffo_set_mutation(self.prev)
        return False


def _dt(prec):
    return np.float32 if prec == "f32" else np.float64


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _fn(name, prec):
    return getattr(lib(), "%s_%s" % (name, prec))


def image_model(img, prec="f32"):
    """a0 (h,w) and gradient (h,w,2) of a u8 or float32 image."""
    h, w = img.shape
    a0 = np.empty((h, w), _dt(prec))
    g = np.empty((h, w, 2), _dt(prec))
    if img.dtype == np.uint8:
        _fn("ffo_image_model_u8", prec)(h, w, _p(np.ascontiguousarray(img)), _p(a0), _p(g))
    else:
        _fn("ffo_image_model_f32", prec)(h, w, _p(_f32(img)), _p(a0), _p(g))
    return a0, g


def propagate(flow, iterations, border=3, invert=False, prec="f32"):
    h, w, _ = flow.shape
    out = np.empty((h, w, 2), _dt(prec))
    _fn("ffo_propagate", prec)(h, w, _p(_f32(flow)), int(iterations), int(border), int(invert), _p(out))
    return out


def propagate_pass(axis, flow, scalar=None, vector=None, dt=1.0, border=3, prec="f32"):
    h, w, _ = flow.shape
    fo = np.empty((h, w, 2), _dt(prec))
    so = np.empty((h, w), _dt(prec)) if scalar is not None else None
    vo = np.empty((h, w, 2), _dt(prec)) if vector is not None else None
    _fn("ffo_propagate_pass", prec)(int(axis), h, w, _p(_f32(flow)), _p(_f32(scalar)), _p(_f32(vector)),
                                    ctypes.c_float(dt), int(border), _p(fo), _p(so), _p(vo))
    return fo, so, vo


def propagate_payload(flow, scalar, vector, iterations, border=3, prec="f32"):
    h, w, _ = flow.shape
    fo = np.empty((h, w, 2), _dt(prec))
    so = np.empty((h, w), _dt(prec))
    vo = np.empty((h, w, 2), _dt(prec))
    _fn("ffo_propagate_payload", prec)(h, w, _p(_f32(flow)), _p(_f32(scalar)), _p(_f32(vector)),
                                       int(iterations), int(border), _p(fo), _p(so), _p(vo))
    return fo, so, vo


def update(a0, grad, a0old, flow_old, gamma, maxflow, prec="f32"):
    h, w = a0.shape
    fo = np.empty((h, w, 2), _dt(prec))
    io = np.empty((h, w), _dt(prec))
    _fn("ffo_update", prec)(h, w, _p(_f32(a0)), _p(_f32(grad)), _p(_f32(a0old)), _p(_f32(flow_old)),
                            ctypes.c_float(gamma), ctypes.c_float(maxflow), _p(fo), _p(io))
    return fo, io


def delta_update(a0, grad, a0old, dflow_old, upper, gamma, maxflow, prec="f32"):
    h, w = a0.shape
    hu, wu, _ = upper.shape
    do = np.empty((h, w, 2), _dt(prec))
    fo = np.empty((h, w, 2), _dt(prec))
    io = np.empty((h, w), _dt(prec))
    _fn("ffo_delta_update", prec)(h, w, _p(_f32(a0)), _p(_f32(grad)), _p(_f32(a0old)), _p(_f32(dflow_old)),
                                  hu, wu, _p(_f32(upper)), ctypes.c_float(gamma), ctypes.c_float(maxflow),
                                  _p(do), _p(fo), _p(io))
    return do, fo, io


def upper_index(x, w, wu, prec="f32"):
    return _fn("ffo_upper_index", prec)(int(x), int(w), int(wu))


def smooth(flow, iterations, prec="f32"):
    h, w, _ = flow.shape
    out = np.empty((h, w, 2), _dt(prec))
    _fn("ffo_smooth", prec)(h, w, _p(_f32(flow)), int(iterations), _p(out))
    return out


def level_sizes(h, w, levels):
    sizes = [(h, w)]
    for _ in range(1, levels):
        h, w = h // 2, w // 2
        sizes.append((h, w))
    return sizes


def pyramid(img, levels, prec="f32"):
    """Levels 0..levels-1 of the image pyramid (level 0 is the input)."""
    h, w = img.shape
    sizes = level_sizes(h, w, levels)
    n = sum(a * b for a, b in sizes[1:])
    if img.dtype == np.uint8:
        out = np.empty(max(n, 1), np.uint8)
        _fn("ffo_pyramid_u8", prec)(h, w, int(levels), _p(np.ascontiguousarray(img)), _p(out))
    else:
        out = np.empty(max(n, 1), np.float32)
        _fn("ffo_pyramid_f32", prec)(h, w, int(levels), _p(_f32(img)), _p(out))
    res, off = [img], 0
    for (a, b) in sizes[1:]:
        res.append(out[off:off + a * b].reshape(a, b).copy())
        off += a * b
    return res


def down_u8(km, k0, kp, prec="f32"):
    return _fn("ffo_down_u8", prec)(int(km), int(k0), int(kp))


def flow_to_color(flow, maxflow, wheel):
    """RGBA colours (h, w, 4) uint8 of a flow field (h, w, 2) read from the
    colour wheel (wh, ww, 4) uint8 (flowToColor_k)."""
    h, w, _ = flow.shape
    wh, ww, _ = wheel.shape
    out = np.empty((h, w, 4), np.uint8)
    lib().ffo_flow_to_color(h, w, _p(_f32(flow)), ctypes.c_float(maxflow), wh, ww,
                            _p(np.ascontiguousarray(wheel, np.uint8)), _p(out))
    return out


def wheel_texel_index(c, n):
    f = lib().ffo_wheel_texel_index
    f.argtypes = [ctypes.c_float, ctypes.c_int]
    return f(float(c), int(n))


def create_perspective_camera(focal, h, w, sensor_h, sensor_w):
    """(alphaX, alphaY, centerX, centerY) as float32."""
    out = np.empty(4, np.float32)
    lib().ffo_create_perspective_camera(ctypes.c_float(focal), int(h), int(w), ctypes.c_float(sensor_h),
                                        ctypes.c_float(sensor_w), _p(out))
    return out


def rotational_flow(cam, wx, wy, wz, h, w):
    """Flow (h, w, 2) of a camera (alphaX, alphaY, centerX, centerY) rotating
    with angular velocity (wx, wy, wz)."""
    out = np.empty((h, w, 2), np.float32)
    lib().ffo_rotational_flow(_p(_f32(np.asarray(cam))), ctypes.c_float(wx), ctypes.c_float(wy), ctypes.c_float(wz),
                              int(h), int(w), _p(out))
    return out


def lax_wendroff(flow, img, iterations):
    """LaxWendroffPropagator on a float image (h, w) or (h, w, 4)."""
    h, w = img.shape[:2]
    c = 1 if img.ndim == 2 else img.shape[2]
    out = np.empty(img.shape, np.float32)
    lib().ffo_lax_wendroff(h, w, c, _p(_f32(flow)), _p(_f32(img)), int(iterations), _p(out))
    return out


def lax_wendroff_pass(axis, flow, img, dt):
    h, w = img.shape[:2]
    c = 1 if img.ndim == 2 else img.shape[2]
    out = np.empty(img.shape, np.float32)
    lib().ffo_lax_wendroff_pass(int(axis), h, w, c, _p(_f32(flow)), _p(_f32(img)), ctypes.c_float(dt), _p(out))
    return out


class FlowFilter:
    """Single-level filter (FlowFilter) on the CPU.

    ``configure_only=True`` mirrors FlowFilter(GPUImage): raw gamma 1.0,
    maxflow 1.0, one smoothing iteration, one propagation iteration.
    """

    def __init__(self, height, width, smoothIterations=1, maxflow=1.0, gamma=1.0, u8=True,
                 prec="f32", configure_only=False):
        self.h, self.w, self.prec, self.u8 = height, width, prec, u8
        self._h = _fn("ffo_flowfilter_create", prec)(height, width, int(u8), int(smoothIterations),
                                                     ctypes.c_float(maxflow), ctypes.c_float(gamma),
                                                     int(configure_only))

    def __del__(self):
        if getattr(self, "_h", None):
            _fn("ffo_flowfilter_destroy", self.prec)(ctypes.c_void_p(self._h))
            self._h = None

    def setGamma(self, g):
        _fn("ffo_flowfilter_set_gamma", self.prec)(ctypes.c_void_p(self._h), ctypes.c_float(g))

    def setMaxFlow(self, m):
        _fn("ffo_flowfilter_set_maxflow", self.prec)(ctypes.c_void_p(self._h), ctypes.c_float(m))

    def setSmoothIterations(self, n):
        _fn("ffo_flowfilter_set_smooth_iterations", self.prec)(ctypes.c_void_p(self._h), int(n))

    def setPropagationBorder(self, b):
        _fn("ffo_flowfilter_set_border", self.prec)(ctypes.c_void_p(self._h), int(b))

    def compute(self, img):
        if self.u8:
            _fn("ffo_flowfilter_compute_u8", self.prec)(ctypes.c_void_p(self._h), _p(np.ascontiguousarray(img, np.uint8)))
        else:
            _fn("ffo_flowfilter_compute_f32", self.prec)(ctypes.c_void_p(self._h), _p(_f32(img)))

    def _get(self, which, ch):
        shape = (self.h, self.w, ch) if ch > 1 else (self.h, self.w)
        out = np.empty(shape, _dt(self.prec))
        _fn("ffo_flowfilter_get", self.prec)(ctypes.c_void_p(self._h), int(which), _p(out))
        return out

    def downloadFlow(self):
        return self._get(0, 2)

    def getFlow(self):
        return self._get(1, 2)

    def downloadImage(self):
        return self._get(2, 1)


class DeltaFlowFilter:
    """Standalone DeltaFlowFilter (its own compute(), flowfilter.cu:429-457) on
    the CPU. The upper-level flow is passed to every compute(): the flow the
    reference filter's inputFlow image holds at that point."""

    def __init__(self, height, width, u8=True, prec="f32"):
        self.h, self.w, self.prec, self.u8 = height, width, prec, u8
        self._h = _fn("ffo_deltafilter_create", prec)(height, width, int(u8))

    def __del__(self):
        if getattr(self, "_h", None):
            _fn("ffo_deltafilter_destroy", self.prec)(ctypes.c_void_p(self._h))
            self._h = None

    def setGamma(self, g):
        _fn("ffo_deltafilter_set_gamma", self.prec)(ctypes.c_void_p(self._h), ctypes.c_float(g))

    def setMaxFlow(self, m):
        _fn("ffo_deltafilter_set_maxflow", self.prec)(ctypes.c_void_p(self._h), ctypes.c_float(m))

    def setSmoothIterations(self, n):
        _fn("ffo_deltafilter_set_smooth_iterations", self.prec)(ctypes.c_void_p(self._h), int(n))

    def setPropagationBorder(self, b):
        _fn("ffo_deltafilter_set_border", self.prec)(ctypes.c_void_p(self._h), int(b))

    def compute(self, img, upper):
        up = _f32(upper)
        hu, wu = up.shape[:2]
        if self.u8:
            _fn("ffo_deltafilter_compute_u8", self.prec)(ctypes.c_void_p(self._h),
                                                         _p(np.ascontiguousarray(img, np.uint8)), hu, wu, _p(up))
        else:
            _fn("ffo_deltafilter_compute_f32", self.prec)(ctypes.c_void_p(self._h), _p(_f32(img)), hu, wu, _p(up))

    def _get(self, which, ch):
        shape = (self.h, self.w, ch) if ch > 1 else (self.h, self.w)
        out = np.empty(shape, _dt(self.prec))
        _fn("ffo_deltafilter_get", self.prec)(ctypes.c_void_p(self._h), int(which), _p(out))
        return out

    def getFlow(self):
        """Smoothed flow (DeltaFlowFilter::getFlow, flowfilter.cu:545-547)."""
        return self._get(0, 2)

    def updatedFlow(self):
        return self._get(1, 2)

    def getImage(self):
        return self._get(2, 1)

    def deltaFlow(self):
        return self._get(3, 2)

    def propagatedScalar(self):
        return self._get(4, 1)


class PyramidalFlowFilter:
    """PyramidalFlowFilter on the CPU (u8 input)."""

    def __init__(self, height, width, levels, prec="f32"):
        self.h, self.w, self.levels, self.prec = height, width, levels, prec
        self.sizes = level_sizes(height, width, levels)
        self._h = _fn("ffo_pyramidal_create", prec)(height, width, int(levels))

    def __del__(self):
        if getattr(self, "_h", None):
            _fn("ffo_pyramidal_destroy", self.prec)(ctypes.c_void_p(self._h))
            self._h = None

    def setGamma(self, level, g):
        _fn("ffo_pyramidal_set_gamma", self.prec)(ctypes.c_void_p(self._h), int(level), ctypes.c_float(g))

    def setMaxFlow(self, m):
        _fn("ffo_pyramidal_set_maxflow", self.prec)(ctypes.c_void_p(self._h), ctypes.c_float(m))

    def setSmoothIterations(self, level, n):
        _fn("ffo_pyramidal_set_smooth_iterations", self.prec)(ctypes.c_void_p(self._h), int(level), int(n))

    def setPropagationBorder(self, b):
        _fn("ffo_pyramidal_set_border", self.prec)(ctypes.c_void_p(self._h), int(b))

    def configure_like(self, gamma, maxflow, smooth):
        for l, g in enumerate(gamma):
            self.setGamma(l, g)
        self.setMaxFlow(maxflow)
        for l, n in enumerate(smooth):
            self.setSmoothIterations(l, n)

    def compute(self, img):
        _fn("ffo_pyramidal_compute", self.prec)(ctypes.c_void_p(self._h), _p(np.ascontiguousarray(img, np.uint8)))

    def _get(self, which, ch, level=0):
        h, w = self.sizes[level]
        shape = (h, w, ch) if ch > 1 else (h, w)
        out = np.empty(shape, _dt(self.prec))
        _fn("ffo_pyramidal_get", self.prec)(ctypes.c_void_p(self._h), int(which), _p(out))
        return out

    def downloadFlow(self):
        return self._get(0, 2)

    def downloadImage(self):
        return self._get(1, 1)

    def getFlow(self):
        lvl = 0 if self.levels > 1 else 0
        return self._get(2, 2, lvl)

    def level_state(self, level, which):
        """0 smoothed flow, 1 updated flow, 2 updated image, 3 delta flow."""
        h, w = self.sizes[level]
        ch = 1 if which == 2 else 2
        shape = (h, w, ch) if ch > 1 else (h, w)
        out = np.empty(shape, _dt(self.prec))
        _fn("ffo_pyramidal_get_level", self.prec)(ctypes.c_void_p(self._h), int(level), int(which), _p(out))
        return out
