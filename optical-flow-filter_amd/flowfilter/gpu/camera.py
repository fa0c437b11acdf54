"""flowfilter.gpu.camera -- perspective camera intrinsics.

Mirrors python/flowfilter/gpu/camera.pyx:14-73 of the reference:
PerspectiveCamera (read-only alphaX, alphaY, centerX, centerY and
getIntrinsicsMatrix) and createPerspectiveCamera, over ff_camera and
ff_create_perspective_camera of the C ABI.
"""
import ctypes

import numpy as np

from .. import _native as N


class PerspectiveCamera(object):
    """Intrinsics of a perspective camera (perspectiveCamera,
    include/flowfilter/gpu/camera.h:20-25)."""

    def __init__(self):
        self.cam = N.ff_camera()

    @classmethod
    def _of(cls, c):
        self = cls()
        self.cam = c
        return self

    def getIntrinsicsMatrix(self):
        K = np.zeros((3, 3), dtype=np.float32)
        K[0, 0] = self.alphaX
        K[0, 2] = self.centerX
        K[1, 1] = self.alphaY
        K[1, 2] = self.centerY
        K[2, 2] = 1.0
        return K

    def _ro(name):
        def get(self):
            return getattr(self.cam, name)

        def put(self, value):
            raise RuntimeError("%s cannot be set" % name)
        return property(get, put)

    alphaX = _ro("alphaX")
    alphaY = _ro("alphaY")
    centerX = _ro("centerX")
    centerY = _ro("centerY")
    del _ro


def createPerspectiveCamera(focalLength, height, width, sensorHeight, sensorWidth):
    """createPerspectiveCamera (src/gpu/camera.cu:14-29): focal length and
    sensor size in millimetres, image size in pixels."""
    c = N.ff_camera()
    N.call("ff_create_perspective_camera", ctypes.c_float(focalLength), int(height), int(width),
           ctypes.c_float(sensorHeight), ctypes.c_float(sensorWidth), ctypes.byref(c))
    return PerspectiveCamera._of(c)
