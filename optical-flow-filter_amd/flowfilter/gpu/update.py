"""flowfilter.gpu.update (python/flowfilter/gpu/update.pyx:17-260)."""
import ctypes

from .. import _native as N
from ._stage import Stage, _hnd


class FlowUpdate(Stage):
    """Per-pixel regularised 2x2 update of the flow."""

    def __init__(self, inputFlow=None, inputImage=None, inputImageGradient=None, gamma=1.0, maxflow=1.0):
        self._set_handle(N.get_handle("ff_flowupdate_create", _hnd(inputFlow), _hnd(inputImage),
                                      _hnd(inputImageGradient), ctypes.c_float(gamma), ctypes.c_float(maxflow)))

    def setInputFlow(self, inputFlow):
        N.call("ff_flowupdate_set_input_flow", self.handle, inputFlow.handle)

    def setInputImage(self, image):
        N.call("ff_flowupdate_set_input_image", self.handle, image.handle)

    def setInputImageGradient(self, imageGradient):
        N.call("ff_flowupdate_set_input_image_gradient", self.handle, imageGradient.handle)

    def getUpdatedFlow(self):
        return self._img("ff_flowupdate_get_updated_flow")

    def getUpdatedImage(self):
        return self._img("ff_flowupdate_get_updated_image")

    @property
    def gamma(self):
        return self._getf("ff_flowupdate_get_gamma")

    @gamma.setter
    def gamma(self, v):
        N.call("ff_flowupdate_set_gamma", self.handle, ctypes.c_float(v))

    @property
    def maxflow(self):
        return self._getf("ff_flowupdate_get_maxflow")

    @maxflow.setter
    def maxflow(self, v):
        N.call("ff_flowupdate_set_maxflow", self.handle, ctypes.c_float(v))


class DeltaFlowUpdate(Stage):
    """Update of the delta flow of a non-top pyramid level."""

    def __init__(self, inputFlow=None, inputDeltaFlow=None, inputImageOld=None, inputImage=None,
                 inputImageGradient=None, gamma=1.0, maxflow=1.0):
        self._set_handle(N.get_handle("ff_deltaflowupdate_create", _hnd(inputFlow), _hnd(inputDeltaFlow),
                                      _hnd(inputImageOld), _hnd(inputImage), _hnd(inputImageGradient),
                                      ctypes.c_float(gamma), ctypes.c_float(maxflow)))

    def setInputFlow(self, inputFlow):
        N.call("ff_deltaflowupdate_set_input_flow", self.handle, inputFlow.handle)

    def setInputDeltaFlow(self, inputDeltaFlow):
        N.call("ff_deltaflowupdate_set_input_delta_flow", self.handle, inputDeltaFlow.handle)

    def setInputImageOld(self, image):
        N.call("ff_deltaflowupdate_set_input_image_old", self.handle, image.handle)

    def setInputImage(self, image):
        N.call("ff_deltaflowupdate_set_input_image", self.handle, image.handle)

    def setInputImageGradient(self, imageGradient):
        N.call("ff_deltaflowupdate_set_input_image_gradient", self.handle, imageGradient.handle)

    def getUpdatedFlow(self):
        return self._img("ff_deltaflowupdate_get_updated_flow")

    def getUpdatedDeltaFlow(self):
        return self._img("ff_deltaflowupdate_get_updated_delta_flow")

    def getUpdatedImage(self):
        return self._img("ff_deltaflowupdate_get_updated_image")

    @property
    def gamma(self):
        return self._getf("ff_deltaflowupdate_get_gamma")

    @gamma.setter
    def gamma(self, v):
        N.call("ff_deltaflowupdate_set_gamma", self.handle, ctypes.c_float(v))

    @property
    def maxflow(self):
        return self._getf("ff_deltaflowupdate_get_maxflow")

    @maxflow.setter
    def maxflow(self, v):
        N.call("ff_deltaflowupdate_set_maxflow", self.handle, ctypes.c_float(v))
