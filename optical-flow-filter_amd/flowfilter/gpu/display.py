"""flowfilter.gpu.display -- colour coding of flow fields (FlowToColor).

Mirrors python/flowfilter/gpu/display.pyx:18-78 of the reference:
FlowToColor(inputFlow=None, maxflow=1.0), configure, compute, elapsedTime,
setInputFlow, getColorFlow, download and the maxflow property, over
ff_flowtocolor_* of the C ABI. ``colorWheel()`` returns the 256 x 256 RGBA
wheel the colours are read from (getColorWheelRGBA of the C++ library).
"""
import ctypes

import numpy as np

from .. import _native as N
from ._stage import Stage, _hnd


def colorWheel():
    """The colour wheel, a (256, 256, 4) uint8 array (a copy)."""
    img = N.image_t()
    N.call("ff_get_color_wheel_rgba", ctypes.byref(img))
    n = img.height * img.pitch
    buf = (ctypes.c_uint8 * n).from_address(img.data)
    return np.frombuffer(buf, dtype=np.uint8).reshape(img.height, img.width, img.depth).copy()


class FlowToColor(Stage):
    """RGBA colour coding of a flow field (display.pyx:18-78)."""

    def __init__(self, inputFlow=None, maxflow=1.0):
        self._set_handle(N.get_handle("ff_flowtocolor_create", _hnd(inputFlow), ctypes.c_float(maxflow)))

    def setInputFlow(self, inputFlow):
        N.call("ff_flowtocolor_set_input_flow", self.handle, inputFlow.handle)

    def getColorFlow(self):
        """Device image of the colours, (H, W, 4) uint8 (shares the buffer)."""
        return self._img("ff_flowtocolor_get_color_flow")

    def download(self):
        """The colours as an (H, W, 4) uint8 array."""
        return self.getColorFlow().download(np.uint8)

    @property
    def maxflow(self):
        return self._getf("ff_flowtocolor_get_max_flow")

    @maxflow.setter
    def maxflow(self, v):
        N.call("ff_flowtocolor_set_max_flow", self.handle, ctypes.c_float(v))
