"""flowfilter.gpu.flowsmoothing -- FlowSmoother
(include/flowfilter/gpu/flowsmoothing.h:22-57; the reference binds it only
in C++)."""
from .. import _native as N
from ._stage import Stage, _hnd


class FlowSmoother(Stage):
    """Iterated separable 5-tap box smoothing of a flow field."""

    def __init__(self, inputFlow=None, iterations=1):
        self._set_handle(N.get_handle("ff_flowsmoother_create", _hnd(inputFlow), int(iterations)))

    def setInputFlow(self, inputFlow):
        N.call("ff_flowsmoother_set_input_flow", self.handle, inputFlow.handle)

    def getSmoothedFlow(self):
        return self._img("ff_flowsmoother_get_smoothed_flow")

    @property
    def iterations(self):
        return self._geti("ff_flowsmoother_get_iterations")

    @iterations.setter
    def iterations(self, n):
        N.call("ff_flowsmoother_set_iterations", self.handle, int(n))
