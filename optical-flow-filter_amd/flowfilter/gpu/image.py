"""flowfilter.gpu.image -- device image buffers (GPUImage).

Mirrors python/flowfilter/gpu/image.pyx:16-65 of the reference
(GPUImage(shape, itemSize), upload, download(dtype, output), shape) over
ff_gpuimage_* of the C ABI. Handles share the device buffer like the C++
GPUImage (shared pointer semantics). Adds the DLPack protocol
(``__dlpack__`` / ``__dlpack_device__``), so ``torch.from_dlpack`` takes a
device image (e.g. a filter's getFlowDevice()) without a copy.
"""
import ctypes

import numpy as np

from .. import _native as N
from ..image import Image

_KDLROCM = 10

# PyCapsule plumbing for the DLPack exchange (CPython C API through ctypes).
_PyCapsule_New = ctypes.pythonapi.PyCapsule_New
_PyCapsule_New.restype = ctypes.py_object
_PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
_PyCapsule_IsValid = ctypes.pythonapi.PyCapsule_IsValid
_PyCapsule_IsValid.restype = ctypes.c_int
_PyCapsule_IsValid.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
_PyCapsule_GetPointer = ctypes.pythonapi.PyCapsule_GetPointer
_PyCapsule_GetPointer.restype = ctypes.c_void_p
_PyCapsule_GetPointer.argtypes = [ctypes.c_void_p, ctypes.c_char_p]


@ctypes.CFUNCTYPE(None, ctypes.c_void_p)
def _capsule_destructor(capsule):
    # a capsule no consumer took (a consumer renames it "used_dltensor" and
    # calls the tensor's deleter itself)
    if _PyCapsule_IsValid(capsule, b"dltensor"):
        N.lib.ff_dlpack_delete(ctypes.c_void_p(_PyCapsule_GetPointer(capsule, b"dltensor")))


class GPUImage(object):

    def __init__(self, shape=None, itemSize=4, _handle=None):
        self._h = None
        if _handle is not None:
            self._h = _handle
            return
        if shape is not None:
            ndim = len(shape)
            if ndim not in (2, 3):
                raise ValueError("number of dimensions must be 2 or 3, got: {0}".format(ndim))
            depth = 1 if ndim == 2 else shape[2]
            self._h = N.get_handle("ff_gpuimage_create", int(shape[0]), int(shape[1]), int(depth), int(itemSize))
        else:
            self._h = N.get_handle("ff_gpuimage_create_empty")

    @classmethod
    def _wrap(cls, handle):
        return cls(_handle=handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib.ff_gpuimage_destroy(ctypes.c_void_p(h))
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return ctypes.c_void_p(self._h)

    def upload(self, img):
        """Copy a numpy array (host) into this device image."""
        w = Image(np.ascontiguousarray(img) if not img.flags["C_CONTIGUOUS"] else img)
        N.call("ff_gpuimage_upload", self.handle, w.ref())

    def download(self, dtype, output=None):
        """Download to a numpy array of the given dtype (allocated when output is None)."""
        if output is None:
            output = np.zeros(self.shape, dtype=dtype)
        w = Image(output)
        N.call("ff_gpuimage_download", self.handle, w.ref())
        return output

    def copyFrom(self, other):
        N.call("ff_gpuimage_copy_from", self.handle, other.handle)

    def clear(self):
        N.call("ff_gpuimage_clear", self.handle)

    def uploadDevice(self, ptr, pitch):
        """Enqueue a device-to-device copy from ``ptr`` (row pitch in bytes)."""
        N.call("ff_gpuimage_upload_device", self.handle, ctypes.c_void_p(ptr), ctypes.c_size_t(pitch))

    def downloadDevice(self, ptr, pitch):
        N.call("ff_gpuimage_download_device", self.handle, ctypes.c_void_p(ptr), ctypes.c_size_t(pitch))

    def __dlpack_device__(self):
        """(kDLROCM, device id) -- the DLPack device of the buffer."""
        return (_KDLROCM, 0)

    def __dlpack__(self, stream=None, max_version=None, dl_device=None):
        """DLPack capsule viewing this image's buffer (uint8 or float32; shape
        (H, W) or (H, W, depth), row stride from the pitch), zero-copy. The
        consumer's stream (an int hipStream_t; None: the null stream; -1: no
        synchronization) first waits for the work enqueued on the image. The
        view holds the buffer alive. A filter's state images alternate
        between two buffers from one compute() to the next: take a new view
        after each compute()."""
        if dl_device is not None and tuple(dl_device) != self.__dlpack_device__():
            raise BufferError("GPUImage lives on device %s, not %s" % (self.__dlpack_device__(), dl_device))
        sync = 0 if stream == -1 else 1
        s = None if stream in (None, -1) else ctypes.c_void_p(int(stream))
        m = N.get_handle("ff_gpuimage_to_dlpack", self.handle, s, sync)
        return _PyCapsule_New(m, b"dltensor", ctypes.cast(_capsule_destructor, ctypes.c_void_p))

    @property
    def height(self):
        return N.get_int("ff_gpuimage_height", self.handle)

    @property
    def width(self):
        return N.get_int("ff_gpuimage_width", self.handle)

    @property
    def depth(self):
        return N.get_int("ff_gpuimage_depth", self.handle)

    @property
    def pitch(self):
        return N.get_int("ff_gpuimage_pitch", self.handle)

    @property
    def itemSize(self):
        return N.get_int("ff_gpuimage_item_size", self.handle)

    @property
    def data(self):
        """Device pointer of the buffer (int)."""
        return N.get_handle("ff_gpuimage_data", self.handle) or 0

    @property
    def shape(self):
        d = self.depth
        if d == 1:
            return (self.height, self.width)
        return (self.height, self.width, d)
