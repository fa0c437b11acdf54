"""flowfilter.gpu.flowfilters -- FlowFilter and PyramidalFlowFilter.

Mirrors python/flowfilter/gpu/flowfilters.pyx:15-444 of the reference:
constructor arguments, loadImage/compute/getFlow/getImage/elapsedTime,
getFlowDevice and the gamma/maxflow/smoothIterations/propagationBorder
properties. Also exposes DeltaFlowFilter (C++ only in the reference).

Additions without a reference counterpart: ``loadImageDevice`` (input from
device memory, e.g. another GPUImage's data pointer), the ``fused`` property
selecting the fused frame kernel (default) or one launch per stage kernel,
and ``BatchedFlowFilter`` (many filters advanced by one launch per frame).
"""
import ctypes

import numpy as np

from .. import _native as N
from ..image import Image
from ._stage import Stage
from .image import GPUImage


def _ro(name):
    def _set(self, v):
        raise RuntimeError("Filter %s cannot be set after instantiation" % name)
    return _set


class FlowFilter(Stage):
    """Single-level optical flow filter (flowfilters.pyx:15-210)."""

    def __init__(self, height, width, smoothIterations=1, maxflow=1.0, gamma=1.0):
        self._set_handle(N.get_handle("ff_flowfilter_create", int(height), int(width), int(smoothIterations),
                                      ctypes.c_float(maxflow), ctypes.c_float(gamma)))

    @classmethod
    def fromGPUImage(cls, inputImage):
        """FlowFilter(GPUImage) constructor of the C++ class."""
        self = cls.__new__(cls)
        self._set_handle(N.get_handle("ff_flowfilter_create_from_image", inputImage.handle))
        return self

    def loadImage(self, img):
        w = Image(img)
        N.call("ff_flowfilter_load_image", self.handle, w.ref())

    def loadImageDevice(self, ptr, pitch):
        N.call("ff_flowfilter_load_image_device", self.handle, ctypes.c_void_p(ptr), ctypes.c_size_t(pitch))

    def getFlow(self, flow=None):
        """Smoothed flow as an (H, W, 2) float32 array (downloadFlow)."""
        if flow is None:
            flow = np.zeros((self.height, self.width, 2), dtype=np.float32)
        w = Image(flow)
        N.call("ff_flowfilter_download_flow", self.handle, w.ref())
        return flow

    def getFlowDevice(self):
        """Device buffer of the updated (unsmoothed) flow (FlowFilter::getFlow)."""
        return self._img("ff_flowfilter_get_flow")

    def getImage(self, image=None):
        if image is None:
            image = np.zeros((self.height, self.width), dtype=np.float32)
        w = Image(image)
        N.call("ff_flowfilter_download_image", self.handle, w.ref())
        return image

    def computeImageModel(self):
        N.call("ff_flowfilter_compute_image_model", self.handle)

    def computePropagation(self):
        N.call("ff_flowfilter_compute_propagation", self.handle)

    def computeUpdate(self):
        N.call("ff_flowfilter_compute_update", self.handle)

    def setInputImage(self, img):
        N.call("ff_flowfilter_set_input_image", self.handle, img.handle)

    height = property(lambda self: self._geti("ff_flowfilter_height"), _ro("height"))
    width = property(lambda self: self._geti("ff_flowfilter_width"), _ro("width"))

    @property
    def gamma(self):
        return self._getf("ff_flowfilter_get_gamma")

    @gamma.setter
    def gamma(self, v):
        N.call("ff_flowfilter_set_gamma", self.handle, ctypes.c_float(v))

    @property
    def maxflow(self):
        return self._getf("ff_flowfilter_get_maxflow")

    @maxflow.setter
    def maxflow(self, v):
        N.call("ff_flowfilter_set_maxflow", self.handle, ctypes.c_float(v))

    @property
    def smoothIterations(self):
        return self._geti("ff_flowfilter_get_smooth_iterations")

    @smoothIterations.setter
    def smoothIterations(self, v):
        N.call("ff_flowfilter_set_smooth_iterations", self.handle, int(v))

    @property
    def propagationIterations(self):
        return self._geti("ff_flowfilter_get_propagation_iterations")

    @propagationIterations.setter
    def propagationIterations(self, v):
        raise RuntimeError("propagation iterations cannot be set, use maxflow instead")

    @property
    def propagationBorder(self):
        return self._geti("ff_flowfilter_get_propagation_border")

    @propagationBorder.setter
    def propagationBorder(self, v):
        N.call("ff_flowfilter_set_propagation_border", self.handle, int(v))

    @property
    def fused(self):
        return bool(self._geti("ff_filter_get_fused"))

    @fused.setter
    def fused(self, v):
        N.call("ff_filter_set_fused", self.handle, int(bool(v)))


class DeltaFlowFilter(Stage):
    """Non-top pyramid level filter (include/flowfilter/gpu/flowfilter.h:148-235)."""

    def __init__(self, inputImage=None, inputFlow=None):
        a = inputImage.handle if inputImage is not None else None
        b = inputFlow.handle if inputFlow is not None else None
        self._set_handle(N.get_handle("ff_deltaflowfilter_create", a, b))

    def computeImageModel(self):
        N.call("ff_deltaflowfilter_compute_image_model", self.handle)

    def computePropagation(self):
        N.call("ff_deltaflowfilter_compute_propagation", self.handle)

    def computeUpdate(self):
        N.call("ff_deltaflowfilter_compute_update", self.handle)

    def setInputImage(self, img):
        N.call("ff_deltaflowfilter_set_input_image", self.handle, img.handle)

    def setInputFlow(self, flow):
        N.call("ff_deltaflowfilter_set_input_flow", self.handle, flow.handle)

    def getFlow(self):
        return self._img("ff_deltaflowfilter_get_flow")

    def getImage(self):
        return self._img("ff_deltaflowfilter_get_image")

    height = property(lambda self: self._geti("ff_deltaflowfilter_height"), _ro("height"))
    width = property(lambda self: self._geti("ff_deltaflowfilter_width"), _ro("width"))

    @property
    def gamma(self):
        return self._getf("ff_deltaflowfilter_get_gamma")

    @gamma.setter
    def gamma(self, v):
        N.call("ff_deltaflowfilter_set_gamma", self.handle, ctypes.c_float(v))

    @property
    def maxflow(self):
        return self._getf("ff_deltaflowfilter_get_maxflow")

    @maxflow.setter
    def maxflow(self, v):
        N.call("ff_deltaflowfilter_set_maxflow", self.handle, ctypes.c_float(v))

    @property
    def smoothIterations(self):
        return self._geti("ff_deltaflowfilter_get_smooth_iterations")

    @smoothIterations.setter
    def smoothIterations(self, v):
        N.call("ff_deltaflowfilter_set_smooth_iterations", self.handle, int(v))

    @property
    def propagationIterations(self):
        return self._geti("ff_deltaflowfilter_get_propagation_iterations")

    @property
    def propagationBorder(self):
        return self._geti("ff_deltaflowfilter_get_propagation_border")

    @propagationBorder.setter
    def propagationBorder(self, v):
        N.call("ff_deltaflowfilter_set_propagation_border", self.handle, int(v))

    @property
    def fused(self):
        return bool(self._geti("ff_filter_get_fused"))

    @fused.setter
    def fused(self, v):
        N.call("ff_filter_set_fused", self.handle, int(bool(v)))


class PyramidalFlowFilter(Stage):
    """Pyramidal optical flow filter (flowfilters.pyx:213-444)."""

    def __init__(self, height, width, levels):
        self._set_handle(N.get_handle("ff_pyramidal_create", int(height), int(width), int(levels)))

    def loadImage(self, img):
        w = Image(img)
        N.call("ff_pyramidal_load_image", self.handle, w.ref())

    def loadImageDevice(self, ptr, pitch):
        N.call("ff_pyramidal_load_image_device", self.handle, ctypes.c_void_p(ptr), ctypes.c_size_t(pitch))

    def setImageDevice(self, ptr, pitch):
        """The next compute() reads its u8 frame in place from device memory
        (ptr, row pitch in bytes) instead of copying it into the input image;
        the memory must hold the frame until that compute() has run."""
        N.call("ff_pyramidal_set_image_device", self.handle, ctypes.c_void_p(ptr), ctypes.c_size_t(pitch))

    def getFlow(self, flow=None):
        if flow is None:
            flow = np.zeros((self.height, self.width, 2), dtype=np.float32)
        w = Image(flow)
        N.call("ff_pyramidal_download_flow", self.handle, w.ref())
        return flow

    def getFlowDevice(self):
        """Returns GPU buffer with the optical flow."""
        return self._img("ff_pyramidal_get_flow")

    def getImage(self, image=None):
        if image is None:
            image = np.zeros((self.height, self.width), dtype=np.float32)
        w = Image(image)
        N.call("ff_pyramidal_download_image", self.handle, w.ref())
        return image

    height = property(lambda self: self._geti("ff_pyramidal_height"), _ro("height"))
    width = property(lambda self: self._geti("ff_pyramidal_width"), _ro("width"))
    levels = property(lambda self: self._geti("ff_pyramidal_levels"), _ro("levels"))

    @property
    def gamma(self):
        if self.levels == 1:
            return N.get_float("ff_pyramidal_get_gamma", self.handle, 0)
        return [N.get_float("ff_pyramidal_get_gamma", self.handle, h) for h in range(self.levels)]

    @gamma.setter
    def gamma(self, g):
        if self.levels == 1:
            N.call("ff_pyramidal_set_gamma", self.handle, 0, ctypes.c_float(g))
        else:
            for h in range(self.levels):
                N.call("ff_pyramidal_set_gamma", self.handle, h, ctypes.c_float(g[h]))

    @property
    def maxflow(self):
        return self._getf("ff_pyramidal_get_maxflow")

    @maxflow.setter
    def maxflow(self, v):
        N.call("ff_pyramidal_set_maxflow", self.handle, ctypes.c_float(v))

    @property
    def smoothIterations(self):
        if self.levels == 1:
            return N.get_int("ff_pyramidal_get_smooth_iterations", self.handle, 0)
        return [N.get_int("ff_pyramidal_get_smooth_iterations", self.handle, h) for h in range(self.levels)]

    @smoothIterations.setter
    def smoothIterations(self, n):
        if self.levels == 1:
            N.call("ff_pyramidal_set_smooth_iterations", self.handle, 0, int(n))
        else:
            for h in range(self.levels):
                N.call("ff_pyramidal_set_smooth_iterations", self.handle, h, int(n[h]))

    @property
    def propagationBorder(self):
        return self._geti("ff_pyramidal_get_propagation_border")

    @propagationBorder.setter
    def propagationBorder(self, v):
        N.call("ff_pyramidal_set_propagation_border", self.handle, int(v))

    def setGammaVector(self, gamma):
        arr = (ctypes.c_float * len(gamma))(*gamma)
        N.call("ff_pyramidal_set_gamma_vector", self.handle, arr, len(gamma))

    def setSmoothIterationsVector(self, iters):
        arr = (ctypes.c_int * len(iters))(*iters)
        N.call("ff_pyramidal_set_smooth_iterations_vector", self.handle, arr, len(iters))

    @property
    def fused(self):
        return bool(self._geti("ff_filter_get_fused"))

    @fused.setter
    def fused(self, v):
        N.call("ff_filter_set_fused", self.handle, int(bool(v)))

    @property
    def graphs(self):
        """Replay each fused frame as a captured HIP graph (default True)."""
        return bool(self._geti("ff_filter_get_graphs"))

    @graphs.setter
    def graphs(self, v):
        N.call("ff_filter_set_graphs", self.handle, int(bool(v)))

    def waitTimeouts(self):
        """Waits of a fused frame that ended at the spin limit (0 when every
        level saw the level above it complete). Waits for the filter's stream."""
        v = ctypes.c_ulonglong(0)
        N.call("ff_filter_get_wait_timeouts", self.handle, ctypes.byref(v))
        return v.value


class BatchedFlowFilter(Stage):
    """Many PyramidalFlowFilters advanced by one compute() (no reference
    counterpart; replaces a loop of compute() calls over camera streams).

    Every filter added keeps its own parameters, input and outputs (load
    frames with its loadImage/loadImageDevice, read results with its
    getFlow); compute() moves all of them one frame forward with a single
    fused frame launch, with results identical to each filter's own
    compute(). elapsedTime() is the time of the whole step.
    """

    def __init__(self, filters=()):
        self._set_handle(N.get_handle("ff_batch_create"))
        self._filters = []
        for f in filters:
            self.add(f)

    def add(self, f):
        if not isinstance(f, PyramidalFlowFilter):
            raise TypeError("BatchedFlowFilter.add() takes a PyramidalFlowFilter")
        N.call("ff_batch_add", self.handle, f.handle)
        self._filters.append(f)
        return len(self._filters) - 1

    def __len__(self):
        return N.get_int("ff_batch_size", self.handle)

    def loadImagesDevice(self, ptrs, pitches):
        """Next frame of every filter (in the order added) from device
        memory, in one copy launch: ptrs[i] with row pitch pitches[i]."""
        n = len(ptrs)
        if len(pitches) != n:
            raise ValueError("loadImagesDevice: %d pointers, %d pitches" % (n, len(pitches)))
        p = (ctypes.c_void_p * max(n, 1))(*[int(v) for v in ptrs])
        q = (ctypes.c_size_t * max(n, 1))(*[int(v) for v in pitches])
        N.call("ff_batch_load_images_device", self.handle, p, q, n)

    def setImagesDevice(self, ptrs, pitches):
        """Next frame of every filter read in place by the next compute():
        ptrs[i] (device u8, row pitch pitches[i]) must hold filter i's frame
        until that compute() has run on the batch's stream."""
        n = len(ptrs)
        if len(pitches) != n:
            raise ValueError("setImagesDevice: %d pointers, %d pitches" % (n, len(pitches)))
        p = (ctypes.c_void_p * max(n, 1))(*[int(v) for v in ptrs])
        q = (ctypes.c_size_t * max(n, 1))(*[int(v) for v in pitches])
        N.call("ff_batch_set_images_device", self.handle, p, q, n)

    def __getitem__(self, i):
        return self._filters[i]

    @property
    def graphs(self):
        """Replay each fused frame step as a captured HIP graph (default True)."""
        return bool(N.get_int("ff_filter_get_graphs", self.handle))

    @graphs.setter
    def graphs(self, v):
        N.call("ff_filter_set_graphs", self.handle, int(bool(v)))

    def lastBlocks(self):
        """Workgroups of the last fused launch."""
        return N.get_int("ff_batch_get_blocks", self.handle)

    def waitTimeouts(self):
        """Waits of a fused step that ended at the spin limit (0 when every
        level saw the level above it complete). Waits for the batch's stream."""
        v = ctypes.c_ulonglong(0)
        N.call("ff_filter_get_wait_timeouts", self.handle, ctypes.byref(v))
        return v.value
