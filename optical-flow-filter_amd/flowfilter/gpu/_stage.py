"""Common base of the stage wrappers: owns one ff_stage handle."""
import ctypes

from .. import _native as N
from .image import GPUImage


class Stage(object):
    """flowfilter::gpu::Stage (include/flowfilter/gpu/pipeline.h:29-78)."""

    _h = None

    def _set_handle(self, h):
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib.ff_stage_destroy(ctypes.c_void_p(h))
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("stage has no underlying object")
        return ctypes.c_void_p(self._h)

    def configure(self):
        N.call("ff_stage_configure", self.handle)

    def compute(self):
        N.call("ff_stage_compute", self.handle)

    def elapsedTime(self):
        """Milliseconds of the last compute() (waits for it to finish)."""
        return N.get_float("ff_stage_elapsed_time", self.handle)

    @property
    def stream(self):
        """The hipStream_t (as an int) this stage's work is enqueued on: a
        filter's or batch's own, else the library stream."""
        return N.get_handle("ff_stage_get_stream", self.handle)

    @stream.setter
    def stream(self, hip_stream):
        """Filters and batches: enqueue on a caller's hipStream_t (an int
        pointer; None returns to the filter's own stream)."""
        N.call("ff_stage_set_stream", self.handle, ctypes.c_void_p(hip_stream) if hip_stream else None)

    def setWaitTest(self, spins):
        """Test hook (PyramidalFlowFilter and batches): the fused kernel's
        waits give up after ``spins`` polls (0 restores them)."""
        N.call("ff_filter_set_wait_test", self.handle, int(spins))

    # helpers for subclasses
    def _img(self, fn):
        return GPUImage._wrap(N.get_handle(fn, self.handle))

    def _geti(self, fn):
        return N.get_int(fn, self.handle)

    def _getf(self, fn):
        return N.get_float(fn, self.handle)

    @property
    def tileWaves(self):
        """Filters only: tile height of the fused frame kernel in waves per
        workgroup (0 = autotune on the first fused frame, the default; or 16,
        15, 8, 5, 4). Reads the fixed or tuned value."""
        return N.get_int("ff_filter_get_tile_waves", self.handle)

    @tileWaves.setter
    def tileWaves(self, v):
        N.call("ff_filter_set_tile_waves", self.handle, int(v))

    @property
    def tileTopRows(self):
        """Filters only: rows per thread of the fused kernel's tiles on a
        pyramid's top level (0 = tuned with the tile height, the default; or
        4, 8). Reads the fixed or tuned value."""
        return N.get_int("ff_filter_get_tile_top_rows", self.handle)

    @tileTopRows.setter
    def tileTopRows(self, v):
        N.call("ff_filter_set_tile_top_rows", self.handle, int(v))

    @property
    def tileRows(self):
        """Filters only: rows per thread of the fused kernel's tiles below
        the top level (0 = tuned with the tile height, the default; or 4
        with 16 or 15 waves, 8). Reads the fixed or tuned value."""
        return N.get_int("ff_filter_get_tile_rows", self.handle)

    @tileRows.setter
    def tileRows(self, v):
        N.call("ff_filter_set_tile_rows", self.handle, int(v))

    @property
    def tileInterior(self):
        """Filters only: interior tiles of the fused kernel in a tile body of
        their own (0 = tuned with the tiling, the default; 1 = on, with 4
        rows per thread and 16 or 15 waves; 2 = off). Reads the fixed or
        tuned choice."""
        return N.get_int("ff_filter_get_tile_interior", self.handle)

    @tileInterior.setter
    def tileInterior(self, v):
        N.call("ff_filter_set_tile_interior", self.handle, int(v))


def _hnd(img):
    return img.handle if img is not None else None
