"""flowfilter.image -- numpy wrapper producing an image_t descriptor.

Mirrors python/flowfilter/image.pyx:15-42 of the reference: the array must be
C-contiguous with 2 or 3 dimensions; pitch is the first stride and itemSize
the last stride.
"""
import ctypes

import numpy as np

from ._native import image_t


class Image(object):
    """Wraps a numpy array (without copying) as an ``image_t``."""

    def __init__(self, arr=None):
        self.numpyArray = None
        self.img = image_t()
        if arr is None:
            return
        if not isinstance(arr, np.ndarray):
            raise TypeError("arr must be a numpy.ndarray")
        if not arr.flags["C_CONTIGUOUS"]:
            raise ValueError("arr must be C_CONTIGUOUS")
        ndim = arr.ndim
        if ndim != 2 and ndim != 3:
            raise ValueError("Incorrect number of image dimensions. Expecting 2 or 3: {0}".format(ndim))
        self.numpyArray = arr
        self.img.height = arr.shape[0]
        self.img.width = arr.shape[1]
        self.img.depth = arr.shape[2] if ndim == 3 else 1
        self.img.pitch = arr.strides[0]
        self.img.itemSize = arr.strides[ndim - 1]
        self.img.data = arr.ctypes.data_as(ctypes.c_void_p)

    def ref(self):
        return ctypes.byref(self.img)

    @property
    def width(self):
        return self.img.width

    @property
    def height(self):
        return self.img.height

    @property
    def depth(self):
        return self.img.depth

    @property
    def itemSize(self):
        return self.img.itemSize

    @property
    def pitch(self):
        return self.img.pitch
