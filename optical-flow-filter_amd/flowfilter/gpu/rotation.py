"""flowfilter.gpu.rotation -- image prediction from camera rotation.

Mirrors python/flowfilter/gpu/rotation.pyx:17-89 of the reference:
RotationalFlowImagePredictor(cam, inputImage=None) with configure, compute,
elapsedTime, setInputImage, getPredictedImage, getOpticalFlow, setCamera,
setAngularVelocity and the iterations property, over ff_rotation_* of the C
ABI. (The reference's elapsedTime calls itself; here it returns the stage
time.)
"""
import ctypes

from .. import _native as N
from ._stage import Stage, _hnd
from .camera import PerspectiveCamera


class RotationalFlowImagePredictor(Stage):
    """Optical flow of a rotating camera and the image it predicts (the input
    image propagated along that flow with a Lax-Wendroff scheme)."""

    def __init__(self, cam=None, inputImage=None):
        c = ctypes.byref(cam.cam) if cam is not None else None
        self._set_handle(N.get_handle("ff_rotation_create", c, _hnd(inputImage) if cam is not None else None))

    def setInputImage(self, inputImage):
        N.call("ff_rotation_set_input_image", self.handle, inputImage.handle)

    def getPredictedImage(self):
        return self._img("ff_rotation_get_predicted_image")

    def getOpticalFlow(self):
        return self._img("ff_rotation_get_optical_flow")

    def setCamera(self, cam):
        N.call("ff_rotation_set_camera", self.handle, ctypes.byref(cam.cam))

    def getCamera(self):
        c = N.ff_camera()
        N.call("ff_rotation_get_camera", self.handle, ctypes.byref(c))
        return PerspectiveCamera._of(c)

    def setAngularVelocity(self, wx, wy, wz):
        N.call("ff_rotation_set_angular_velocity", self.handle, ctypes.c_float(wx), ctypes.c_float(wy),
               ctypes.c_float(wz))

    @property
    def iterations(self):
        return self._geti("ff_rotation_get_iterations")

    @iterations.setter
    def iterations(self, n):
        N.call("ff_rotation_set_iterations", self.handle, int(n))
