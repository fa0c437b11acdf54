"""flowfilter.gpu.imagemodel (python/flowfilter/gpu/imagemodel.pyx:14-64)."""
from .. import _native as N
from ._stage import Stage, _hnd


class ImageModel(Stage):
    """Brightness model: constant term a0 and gradient of the smoothed image."""

    def __init__(self, inputImage=None):
        self._set_handle(N.get_handle("ff_imagemodel_create", _hnd(inputImage)))

    def setInputImage(self, inputImage):
        N.call("ff_imagemodel_set_input_image", self.handle, inputImage.handle)

    def getImageConstant(self):
        return self._img("ff_imagemodel_get_image_constant")

    def getImageGradient(self):
        return self._img("ff_imagemodel_get_image_gradient")
