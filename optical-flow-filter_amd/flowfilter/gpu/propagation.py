"""flowfilter.gpu.propagation (python/flowfilter/gpu/propagation.pyx:78-330):
FlowPropagator, FlowPropagatorPayload and LaxWendroffPropagator (the image
propagation of the rotational image-prediction path, :254-330).
"""
from .. import _native as N
from ._stage import Stage, _hnd


class FlowPropagator(Stage):
    """Upwind propagation of the optical flow by itself."""

    def __init__(self, inputFlow=None, iterations=1):
        self._set_handle(N.get_handle("ff_flowpropagator_create", _hnd(inputFlow), int(iterations)))

    def setInputFlow(self, inputFlow):
        N.call("ff_flowpropagator_set_input_flow", self.handle, inputFlow.handle)

    def getPropagatedFlow(self):
        return self._img("ff_flowpropagator_get_propagated_flow")

    @property
    def iterations(self):
        return self._geti("ff_flowpropagator_get_iterations")

    @iterations.setter
    def iterations(self, N_):
        N.call("ff_flowpropagator_set_iterations", self.handle, int(N_))

    @property
    def dt(self):
        return self._getf("ff_flowpropagator_get_dt")

    @dt.setter
    def dt(self, v):
        raise RuntimeError("dt cannot be set, use iterations property instead")

    @property
    def invertInputFlow(self):
        return bool(self._geti("ff_flowpropagator_get_invert_input_flow"))

    @invertInputFlow.setter
    def invertInputFlow(self, invert):
        N.call("ff_flowpropagator_set_invert_input_flow", self.handle, int(bool(invert)))

    @property
    def border(self):
        return self._geti("ff_flowpropagator_get_border")

    @border.setter
    def border(self, b):
        N.call("ff_flowpropagator_set_border", self.handle, int(b))


class FlowPropagatorPayload(Stage):
    """Upwind propagation of the flow with a scalar and a 2-vector payload."""

    def __init__(self, inputFlow=None, inputScalar=None, inputVector=None, iterations=1):
        self._set_handle(N.get_handle("ff_flowpropagatorpayload_create", _hnd(inputFlow), _hnd(inputScalar),
                                      _hnd(inputVector), int(iterations)))

    def setInputFlow(self, inputFlow):
        N.call("ff_flowpropagatorpayload_set_input_flow", self.handle, inputFlow.handle)

    def setScalarPayload(self, scalarField):
        N.call("ff_flowpropagatorpayload_set_scalar_payload", self.handle, scalarField.handle)

    def setVectorPayload(self, vectorField):
        N.call("ff_flowpropagatorpayload_set_vector_payload", self.handle, vectorField.handle)

    def getPropagatedFlow(self):
        return self._img("ff_flowpropagatorpayload_get_propagated_flow")

    def getPropagatedScalar(self):
        return self._img("ff_flowpropagatorpayload_get_propagated_scalar")

    def getPropagatedVector(self):
        return self._img("ff_flowpropagatorpayload_get_propagated_vector")

    @property
    def iterations(self):
        return self._geti("ff_flowpropagatorpayload_get_iterations")

    @iterations.setter
    def iterations(self, N_):
        N.call("ff_flowpropagatorpayload_set_iterations", self.handle, int(N_))

    @property
    def dt(self):
        return self._getf("ff_flowpropagatorpayload_get_dt")

    @dt.setter
    def dt(self, v):
        raise RuntimeError("dt cannot be set, use iterations property instead")

    @property
    def border(self):
        return self._geti("ff_flowpropagatorpayload_get_border")

    @border.setter
    def border(self, b):
        N.call("ff_flowpropagatorpayload_set_border", self.handle, int(b))


class LaxWendroffPropagator(Stage):
    """Lax-Wendroff propagation of a float image (1 or 4 channels) along a
    flow field (propagation.pyx:254-330)."""

    def __init__(self, inputFlow=None, inputImage=None):
        if inputFlow is None or inputImage is None:
            inputFlow = inputImage = None
        self._set_handle(N.get_handle("ff_laxwendroff_create", _hnd(inputFlow), _hnd(inputImage)))

    def setInputFlow(self, inputFlow):
        N.call("ff_laxwendroff_set_input_flow", self.handle, inputFlow.handle)

    def setInputImage(self, img):
        N.call("ff_laxwendroff_set_input_image", self.handle, img.handle)

    def getFlow(self):
        return self._img("ff_laxwendroff_get_flow")

    def getPropagatedImage(self):
        return self._img("ff_laxwendroff_get_propagated_image")

    @property
    def iterations(self):
        return self._geti("ff_laxwendroff_get_iterations")

    @iterations.setter
    def iterations(self, n):
        N.call("ff_laxwendroff_set_iterations", self.handle, int(n))

    @property
    def dt(self):
        return self._getf("ff_laxwendroff_get_dt")
