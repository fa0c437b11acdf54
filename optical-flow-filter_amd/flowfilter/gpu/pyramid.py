"""flowfilter.gpu.pyramid (python/flowfilter/gpu/pyramid.pyx:14-58)."""
from .. import _native as N
from ._stage import Stage, _hnd
from .image import GPUImage


class ImagePyramid(Stage):
    """Image pyramid by repeated 2x down-sampling (X then Y)."""

    def __init__(self, inputImage=None, levels=1):
        self._set_handle(N.get_handle("ff_imagepyramid_create", _hnd(inputImage), int(levels)))

    def setInputImage(self, inputImage):
        N.call("ff_imagepyramid_set_input_image", self.handle, inputImage.handle)

    def getImage(self, level):
        return GPUImage._wrap(N.get_handle("ff_imagepyramid_get_image", self.handle, int(level)))

    @property
    def levels(self):
        return self._geti("ff_imagepyramid_get_levels")

    @levels.setter
    def levels(self, n):
        N.call("ff_imagepyramid_set_levels", self.handle, int(n))
