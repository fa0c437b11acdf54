"""Pins the oracle's float pyramid step, its level wiring and its standalone
DeltaFlowFilter sequencing, and shows each check failing on a planted mistake.

The checks:
  * float pyramid (pyramid_k.cu:81-145): constants survive exactly; a linear
    ramp maps level-1 pixel (i, j) to level-0 pixel (2i, 2j) exactly inside,
    and to the clamped closed form at the first row and column; hand-worked
    float cases, one of which fixes that the two side taps are added in float
    (tests/golden/hand_cases.json, "pyramid_float").
  * level wiring (flowfilter.cu:663-672, 685-716): where no clamp is active, a
    delta level's updated flow is its delta flow plus twice the texel (y>>1,
    x>>1) of the flow it reads from the level above. For two levels that flow
    is the top level's unsmoothed flow of the same frame; for three levels,
    level 0 reads level 1's smoothed flow of the previous frame (the lower
    levels update in ascending order) and level 1 reads the top level's
    unsmoothed flow of the same frame.
  * standalone DeltaFlowFilter.compute() (flowfilter.cu:429-457): the first
    frame's old-image copy precedes propagation, so the propagated scalar
    payload is a0 carried by the zero flow; for finite frames the filter
    agrees bitwise with the same level run inside a pyramid (whose
    computeUpdate copies after propagation, :480-503).

Each planted mistake (ffo.mutated) is a plausible transcription error; the
tests assert that the corresponding check rejects it.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
F32 = np.float32


def golden():
    return json.load(open(os.path.join(HERE, "golden", "hand_cases.json")))


# ---------------------------------------------------------- float pyramid

def _num(v):
    return float.fromhex(v) if isinstance(v, str) else float(v)


def _ramp(h, w, a, b):
    y, x = np.mgrid[0:h, 0:w]
    return (a * x + b * y).astype(F32)


def pyramid_float_violations(ffo):
    """Number of failed float-pyramid checks (0 for a faithful oracle)."""
    bad = 0
    for c in (0.0, 1.0, 0.1, 1.0 / 3.0, 255.0, -7.25, 3.0e-30):
        img = np.full((20, 34), c, F32)
        for lvl in ffo.pyramid(img, 3, "f32")[1:]:
            bad += int(np.count_nonzero(lvl != F32(c)))
    for a, b in ((1.0, 0.0), (0.0, 1.0), (3.0, -2.0), (0.5, 0.25)):
        img = _ramp(24, 40, a, b)
        one = ffo.pyramid(img, 2, "f32")[1]
        # interior: the taps 2i-1, 2i, 2i+1 lie inside and average to 2i
        bad += int(np.count_nonzero(one[1:, 1:] != img[2:24:2, 2:40:2][: one.shape[0] - 1, : one.shape[1] - 1]))
        # first column: 0.5*I(0) + 0.25*(I(0) + I(1)) = I(0) + a/4 (x), the
        # same along y for the first row; the corner carries both
        i = np.arange(1, one.shape[0])
        j = np.arange(1, one.shape[1])
        bad += int(np.count_nonzero(one[i, 0] != img[2 * i, 0] + F32(a / 4)))
        bad += int(np.count_nonzero(one[0, j] != img[0, 2 * j] + F32(b / 4)))
        bad += int(one[0, 0] != F32(a / 4 + b / 4))
    for case in golden()["pyramid_float"]:
        row = np.array([_num(v) for v in case["row"]], F32)
        img = np.stack([row, row])  # two equal rows: the Y step returns them
        got = ffo.pyramid(img, 2, "f32")[1]
        want = np.array([[_num(v) for v in case["expect"]]], F32)
        bad += int(got.shape != want.shape or np.count_nonzero(got != want))
    return bad


def test_float_pyramid_pinned(ffo):
    assert pyramid_float_violations(ffo) == 0


@pytest.mark.parametrize("m", ["MUT_PYR_TAP_SHIFT", "MUT_PYR_SUM_DOUBLE"])
def test_float_pyramid_checks_reject_mistakes(ffo, m):
    with ffo.mutated(getattr(ffo, m)):
        assert pyramid_float_violations(ffo) > 0


def test_u8_pyramid_check_rejects_tap_shift(ffo):
    """The integer-numpy pyramid comparison (test_oracle) fails on shifted taps."""
    img = np.random.default_rng(3).integers(0, 256, (33, 47), dtype=np.uint8)
    good = ffo.pyramid(img, 2, "f32")[1]
    with ffo.mutated(ffo.MUT_PYR_TAP_SHIFT):
        bad = ffo.pyramid(img, 2, "f32")[1]
    assert np.count_nonzero(good != bad) > 0


# ------------------------------------------------------------- level wiring

def _upper_at(up, h, w):
    """Texel (y>>1, x>>1) of the upper flow for every pixel of an h x w level
    (the point-sampled index for even sizes, pinned in test_oracle)."""
    return up[(np.arange(h) >> 1)[:, None], (np.arange(w) >> 1)[None, :]]


def _feed_mismatch(flow_upd, dflow, upper, maxflow):
    """Unclamped pixels whose updated flow is not fl(dflow + 2 * upper texel);
    also returns how many pixels were unclamped."""
    h, w = flow_upd.shape[:2]
    pred = (dflow + F32(2) * _upper_at(upper, h, w)).astype(F32)
    free = np.all(np.abs(pred) < F32(maxflow), axis=-1)
    miss = np.any(flow_upd != pred, axis=-1) & free
    return int(np.count_nonzero(miss)), int(np.count_nonzero(free))


def wiring_violations(ffo, levels, frames=8, h=64, w=96, maxflow=6.0):
    """Failed wiring checks over frames 2.. of a translating sequence (0 for a
    faithful oracle); asserts along the way that the check has teeth (most
    pixels unclamped, fields that differ between the candidates)."""
    from flowfilter import synthetic
    p = ffo.PyramidalFlowFilter(h, w, levels)
    p.configure_like([10.0, 30.0, 50.0][:levels], maxflow, [2] * levels)
    m = [maxflow / 2 ** l for l in range(levels)]
    bad = 0
    for t, img in enumerate(synthetic.translating_noise(h, w, frames, (1.7, -1.1))):
        prev_smoothed = [p.level_state(l, 0) for l in range(levels)]
        p.compute(img)
        if t < 2:
            continue
        top = levels - 1
        top_upd = p.level_state(top, 1)
        # the top level's unsmoothed and smoothed flows differ (else the
        # feed check could not tell them apart)
        assert np.count_nonzero(top_upd != p.level_state(top, 0)) > top_upd.size // 4
        # level L-2 reads the top level's unsmoothed flow of this frame
        lo = levels - 2
        miss, free = _feed_mismatch(p.level_state(lo, 1), p.level_state(lo, 3), top_upd, m[lo])
        assert free > 0.9 * p.sizes[lo][0] * p.sizes[lo][1]
        bad += miss
        # level l < L-2 reads level l+1's smoothed flow of the previous frame
        for l in range(levels - 2):
            upper = prev_smoothed[l + 1]
            assert np.count_nonzero(upper != p.level_state(l + 1, 0)) > upper.size // 4
            miss, free = _feed_mismatch(p.level_state(l, 1), p.level_state(l, 3), upper, m[l])
            assert free > 0.9 * p.sizes[l][0] * p.sizes[l][1]
            bad += miss
    return bad


@pytest.mark.parametrize("levels", [2, 3])
def test_level_wiring_pinned(ffo, levels):
    assert wiring_violations(ffo, levels) == 0


@pytest.mark.parametrize("m,levels", [("MUT_SMOOTHED_TOP_FEED", 2), ("MUT_SMOOTHED_TOP_FEED", 3),
                                      ("MUT_TOP_UPDATES_LAST", 2), ("MUT_DESCENDING_UPDATE", 3)])
def test_level_wiring_checks_reject_mistakes(ffo, m, levels):
    with ffo.mutated(getattr(ffo, m)):
        assert wiring_violations(ffo, levels) > 0


# ------------------------------------------- standalone DeltaFlowFilter

def test_standalone_delta_first_frame_payload(ffo):
    """compute()'s first-frame copy precedes propagation: the zero flow carries
    a0 unchanged for a finite frame, and spreads a NaN pixel's hole by one
    pixel per pass (which a copy after propagation would not)."""
    rng = np.random.default_rng(11)
    h, w = 40, 56
    img = (rng.random((h, w)) * 255).astype(F32)
    up = np.zeros((h // 2, w // 2, 2), F32)
    d = ffo.DeltaFlowFilter(h, w, u8=False)
    d.setMaxFlow(2.0)
    d.compute(img, up)
    a0, _ = ffo.image_model(img, "f32")
    assert np.array_equal(d.propagatedScalar(), a0)
    img[20, 28] = np.nan
    d = ffo.DeltaFlowFilter(h, w, u8=False)
    d.setMaxFlow(2.0)  # two iterations: four passes
    d.compute(img, up)
    a0, _ = ffo.image_model(img, "f32")
    holes_a0 = np.count_nonzero(np.isnan(a0))
    holes_prop = np.count_nonzero(np.isnan(d.propagatedScalar()))
    assert holes_a0 > 0 and holes_prop > holes_a0
    assert np.all(np.isnan(d.propagatedScalar())[np.isnan(a0)])


def test_standalone_delta_matches_pyramid_level(ffo):
    """A standalone DeltaFlowFilter fed, frame by frame, the top level's
    unsmoothed flow of a 2-level pyramid computes exactly the pyramid's
    level 0 (finite frames: the two first-frame orders agree)."""
    from flowfilter import synthetic
    h, w = 48, 64
    p = ffo.PyramidalFlowFilter(h, w, 2)
    p.configure_like([10.0, 50.0], 4.0, [2, 2])
    d = ffo.DeltaFlowFilter(h, w)
    d.setGamma(10.0)
    d.setMaxFlow(4.0)
    d.setSmoothIterations(2)
    for img in synthetic.translating_noise(h, w, 6, (1.2, 0.4)):
        p.compute(img)
        d.compute(img, p.level_state(1, 1))
        assert np.array_equal(d.getFlow(), p.level_state(0, 0))
        assert np.array_equal(d.deltaFlow(), p.level_state(0, 3))
        assert np.array_equal(d.getImage(), p.level_state(0, 2))
    assert np.abs(d.getFlow()).max() > 0.1
