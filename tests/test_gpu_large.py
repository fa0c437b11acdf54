"""The largest images: buffers at the 32-bit limit of the fused frame
kernel's buffer offsets (16382 x 16382: a level-0 flow image of
2,147,221,504 bytes, read and written by the fused kernel) and just beyond
it (16384 x 16384: 2 GiB flow images, which the filter runs with the stage
kernels, whose offsets are 64-bit).

The CPU oracle is too slow for 268 Mpixel frames, so it runs on a 1024 x
1024 crop at an even offset near the far corner of the image. A pixel's
flow after t frames depends only on the frame pixels within a bounded
distance of it (per frame: N propagation iterations, the image model's and
the smoothing's taps, and the top level's own such radius at twice the
scale), and the upper-level index map is x >> 1 for even sizes, so away
from the crop's edges (by a margin larger than that distance) the crop's
flow must be bitwise the full image's.

Bar: bitwise equality on the crop's interior after every frame."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FRAMES = 3
CROP = 1024
MARGIN = 160  # > 3 frames x (level 0: 4 + 2 + 4 px; top level: (2 + 3 + 4) x 2 px) with room to spare


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("size", [16382, 16384])
def test_largest_images_match_oracle_on_a_crop(ffo, size):
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import _native as N
    h = w = size
    rng = np.random.default_rng(size)
    base = rng.integers(0, 256, (h, w + FRAMES), dtype=np.uint8)
    # a crop at an even offset near the far corner (large buffer offsets)
    oy, ox = h - CROP - 64, w - CROP - 128
    assert oy % 2 == 0 and ox % 2 == 0
    say("largest images", size, "crop at", oy, ox)
    g = gf.PyramidalFlowFilter(h, w, 2)
    g.gamma = [10.0, 50.0]
    g.maxflow = 4.0
    g.smoothIterations = [2, 2]
    o = ffo.PyramidalFlowFilter(CROP, CROP, 2)
    o.configure_like([10.0, 50.0], 4.0, [2, 2])
    inner = (slice(MARGIN, CROP - MARGIN), slice(MARGIN, CROP - MARGIN))
    for t in range(FRAMES):
        img = np.ascontiguousarray(base[:, t:t + w])  # content moves 1 px/frame to the left
        g.loadImage(img)
        N.profile_begin("fused_frame")
        g.compute()
        _, fused_launches = N.profile_end()
        o.compute(np.ascontiguousarray(img[oy:oy + CROP, ox:ox + CROP]))
        if t == 0:
            continue  # the first frame's flow is zero everywhere
        # 16382: the fused frame kernel; 16384: the stage kernels
        assert (fused_launches > 0) == (size == 16382), fused_launches
        got = g.getFlow()[oy:oy + CROP, ox:ox + CROP][inner]
        want = o.downloadFlow()[inner]
        d = float(np.abs(got - want).max())
        say("  frame", t, "max |diff| on the crop's interior", d, "max |flow|", float(np.abs(want).max()))
        assert np.array_equal(got, want), (size, t, d)
        assert float(np.abs(want).max()) > 0.0
