"""Pins the CPU oracle (oracle/ff_oracle.hpp) before it is trusted as the checker.

The reference has no tests, fixtures or golden vectors for this path, so each
oracle function is pinned here by what the mathematics fixes: closed forms,
an independent formulation (scipy's correlation filters, numpy's linear
solver, exact integer arithmetic), invariants that a dropped term, a wrong
sign, a swapped component or a transposed index would break, hand-worked
small cases (tests/golden/hand_cases.json), and FP32-vs-FP64 agreement.
"""
import ctypes
import json
import os

import numpy as np
import pytest
from scipy import ndimage

HERE = os.path.dirname(os.path.abspath(__file__))
S = np.array([0.0625, 0.25, 0.375, 0.25, 0.0625])
D = np.array([-0.125, -0.25, 0.0, 0.25, 0.125])


def rng(seed):
    return np.random.default_rng(seed)


# ------------------------------------------------------- arithmetic identities

def test_unorm8_fma_correction_is_exact(ffo):
    """The GPU's 3-instruction u8 -> float equals (float)v / 255.0f for all 256 v."""
    assert ffo.lib().ffo_count_unorm8_fma_mismatch() == 0


def test_pyramid_u8_integer_form_is_exact(ffo):
    """The reference's float/double pyramid tap equals (2*k0+km+kp)>>2 for all 2^24 inputs."""
    f = ffo.lib().ffo_count_down_u8_int_mismatch
    f.restype = ctypes.c_longlong
    assert f() == 0


def test_unorm8_roundtrip(ffo):
    """(uchar)(255 * (v/255.0f)) == v for every v (constant images survive the pyramid)."""
    v = np.arange(256, dtype=np.float32)
    t = v / np.float32(255.0)
    assert np.array_equal((np.float32(255) * t).astype(np.uint8), v.astype(np.uint8))


# ------------------------------------------------------------------- pyramid

def test_pyramid_hand_table(ffo):
    cases = json.load(open(os.path.join(HERE, "golden", "hand_cases.json")))["pyramid_tap"]
    for km, k0, kp, want in cases:
        assert ffo.down_u8(km, k0, kp, "f32") == want, (km, k0, kp)
        assert ffo.down_u8(km, k0, kp, "f64") == want, (km, k0, kp)


@pytest.mark.parametrize("shape,levels", [((1088, 2048), 2), ((481, 641), 3), ((37, 53), 4), ((8, 8), 3)])
def test_pyramid_vs_integer_numpy(ffo, shape, levels):
    img = rng(3).integers(0, 256, shape, dtype=np.uint8)
    got = ffo.pyramid(img, levels, "f32")
    assert [g.shape for g in got] == ffo.level_sizes(shape[0], shape[1], levels)
    ref = img.astype(np.int64)
    for lvl in range(1, levels):
        w2 = ref.shape[1] // 2
        p = np.pad(ref, ((0, 0), (1, 1)), mode="edge")
        x = np.arange(w2)
        fx = (p[:, 2 * x] + 2 * p[:, 2 * x + 1] + p[:, 2 * x + 2]) >> 2
        h2 = fx.shape[0] // 2
        p = np.pad(fx, ((1, 1), (0, 0)), mode="edge")
        y = np.arange(h2)
        ref = (p[2 * y, :] + 2 * p[2 * y + 1, :] + p[2 * y + 2, :]) >> 2
        assert np.array_equal(got[lvl], ref.astype(np.uint8)), lvl


def test_pyramid_constant_images(ffo):
    for c in (0, 1, 127, 128, 254, 255):
        img = np.full((40, 60), c, np.uint8)
        for lvl in ffo.pyramid(img, 3, "f32"):
            assert np.all(lvl == c)


# --------------------------------------------------------------- image model

def _model_f64(img):
    """Independent formulation: separable correlations with clamp-to-edge."""
    I = img.astype(np.float64)
    sy = ndimage.correlate1d(I, S, axis=0, mode="nearest")
    sx = ndimage.correlate1d(I, S, axis=1, mode="nearest")
    a0 = ndimage.correlate1d(sy, S, axis=1, mode="nearest")
    gx = ndimage.correlate1d(sy, D, axis=1, mode="nearest")
    gy = ndimage.correlate1d(sx, D, axis=0, mode="nearest")
    return a0, np.stack([gx, gy], -1)


def test_image_model_vs_scipy(ffo):
    img = rng(5).standard_normal((23, 31)).astype(np.float32)
    a0, g = ffo.image_model(img, "f64")
    ra0, rg = _model_f64(img)
    assert np.abs(a0 - ra0).max() < 1e-12
    assert np.abs(g - rg).max() < 1e-12
    a32, g32 = ffo.image_model(img, "f32")
    assert np.abs(a32 - ra0).max() < 2e-6 and np.abs(g32 - rg).max() < 2e-6


def test_image_model_u8_is_normalized(ffo):
    img = rng(6).integers(0, 256, (19, 27), dtype=np.uint8)
    a0, g = ffo.image_model(img, "f64")
    ra0, rg = _model_f64(img / 255.0)
    assert np.abs(a0 - ra0).max() < 1e-12 and np.abs(g - rg).max() < 1e-12


def test_image_model_components_not_swapped(ffo):
    yy, xx = np.mgrid[0:20, 0:24].astype(np.float32)
    _, gx = ffo.image_model(xx, "f32")
    _, gy = ffo.image_model(yy, "f32")
    inner = (slice(4, -4), slice(4, -4))
    assert np.allclose(gx[inner][..., 0], 1.0, atol=1e-6) and np.allclose(gx[inner][..., 1], 0.0, atol=1e-6)
    assert np.allclose(gy[inner][..., 0], 0.0, atol=1e-6) and np.allclose(gy[inner][..., 1], 1.0, atol=1e-6)


def test_image_model_closed_forms(ffo):
    """Binomial mask: sum s = 1, sum s c^2 = 1; derivative mask: sum d c = 1, sum d c^2 = 0."""
    yy, xx = np.mgrid[0:24, 0:30].astype(np.float32)
    a0, g = ffo.image_model(xx * xx, "f64")
    inner = (slice(4, -4), slice(4, -4))
    assert np.allclose(a0[inner], (xx * xx + 1)[inner], atol=1e-9)
    assert np.allclose(g[inner][..., 0], 2 * xx[inner], atol=1e-9)
    assert np.allclose(g[inner][..., 1], 0.0, atol=1e-9)
    # u8 ramp: gradient (alpha, beta) / 255 away from the edges
    ramp = (3 * xx + 2 * yy).astype(np.uint8)
    a0, g = ffo.image_model(ramp, "f32")
    assert np.allclose(g[inner][..., 0], 3 / 255.0, rtol=1e-5)
    assert np.allclose(g[inner][..., 1], 2 / 255.0, rtol=1e-5)
    assert np.allclose(a0[inner], ramp[inner] / 255.0, rtol=1e-5)


def test_image_model_constant(ffo):
    for c in (0, 1, 100, 255):
        a0, g = ffo.image_model(np.full((12, 15), c, np.uint8), "f32")
        assert np.allclose(a0, c / 255.0, rtol=1e-6, atol=0)
        assert np.abs(g).max() <= 1e-6


# ---------------------------------------------------------------- propagation

def test_propagation_hand_cases(ffo):
    """1-D upwind steps worked by hand (tests/golden/hand_cases.json)."""
    for case in json.load(open(os.path.join(HERE, "golden", "hand_cases.json")))["upwind_x"]:
        row = np.array(case["flow"], np.float32)  # (W, 2)
        h = 7
        flow = np.repeat(row[None], h, 0)
        fo, _, _ = ffo.propagate_pass(0, flow, dt=case["dt"], border=case["border"], prec="f32")
        for x, want in case["expect"].items():
            got = fo[3, int(x)]
            assert np.allclose(got, want, atol=1e-6), (case["name"], x, got, want)


def test_propagation_constant_flow_invariant(ffo):
    flow = np.empty((20, 30, 2), np.float32)
    flow[..., 0], flow[..., 1] = 0.37, -1.25
    for it in (1, 2, 4):
        assert np.array_equal(ffo.propagate(flow, it, 3, prec="f32"), flow)


def test_propagation_border_pixels_unchanged(ffo):
    flow = rng(7).uniform(-2, 2, (25, 33, 2)).astype(np.float32)
    for b in (1, 3, 5):
        out = ffo.propagate(flow, 3, b, prec="f32")
        m = np.ones(flow.shape[:2], bool)
        m[b:-b, b:-b] = False
        assert np.array_equal(out[m], flow[m])
        assert not np.array_equal(out[~m], flow[~m])


def test_propagation_mirror_symmetry(ffo):
    """Flipping x and negating u mirrors the result exactly (same for y / v)."""
    flow = rng(8).uniform(-2, 2, (21, 29, 2)).astype(np.float32)
    for axis, comp in ((0, 0), (1, 1)):
        out, _, _ = ffo.propagate_pass(axis, flow, dt=0.25, border=2, prec="f32")
        fl = np.flip(flow, axis=1 - axis).copy()
        fl[..., comp] *= -1
        mo, _, _ = ffo.propagate_pass(axis, fl, dt=0.25, border=2, prec="f32")
        mo = np.flip(mo, axis=1 - axis)
        mo[..., comp] *= -1
        assert np.array_equal(mo, out)


def test_propagation_payload_zero_velocity(ffo):
    flow = np.zeros((16, 18, 2), np.float32)
    s = rng(9).standard_normal((16, 18)).astype(np.float32)
    v = rng(10).standard_normal((16, 18, 2)).astype(np.float32)
    fo, so, vo = ffo.propagate_payload(flow, s, v, 4, 3, prec="f32")
    assert np.array_equal(so, s) and np.array_equal(vo, v) and np.array_equal(fo, flow)


def test_propagation_payload_matches_flow_only(ffo):
    """The payload propagator advects the flow exactly like the plain propagator."""
    flow = rng(11).uniform(-3, 3, (24, 28, 2)).astype(np.float32)
    s = rng(12).standard_normal((24, 28)).astype(np.float32)
    v = rng(13).standard_normal((24, 28, 2)).astype(np.float32)
    fo, _, _ = ffo.propagate_payload(flow, s, v, 4, 3, prec="f32")
    assert np.array_equal(fo, ffo.propagate(flow, 4, 3, prec="f32"))


def test_propagation_translates_scalar(ffo):
    """With a uniform velocity u=1 and dt=1, one X pass shifts a payload by one pixel."""
    flow = np.zeros((9, 20, 2), np.float32)
    flow[..., 0] = 1.0
    s = np.tile(np.arange(20, dtype=np.float32) ** 2, (9, 1))
    v = np.zeros((9, 20, 2), np.float32)
    _, so, _ = ffo.propagate_pass(0, flow, s, v, dt=1.0, border=1, prec="f32")
    assert np.array_equal(so[4, 1:-1], s[4, :-2][:18])


def test_propagation_invert(ffo):
    flow = rng(14).uniform(-2, 2, (15, 17, 2)).astype(np.float32)
    assert np.array_equal(ffo.propagate(flow, 2, 3, invert=True, prec="f32"),
                          ffo.propagate(-flow, 2, 3, invert=False, prec="f32"))


def test_propagation_fp32_vs_fp64(ffo):
    f = ndimage.gaussian_filter(rng(15).standard_normal((40, 50, 2)), (4, 4, 0)).astype(np.float32) * 10
    d = np.abs(ffo.propagate(f, 4, 3, prec="f32") - ffo.propagate(f, 4, 3, prec="f64")).max()
    assert d < 1e-5


# --------------------------------------------------------------------- update

def _solve_f64(a0, g, a0old, f, gamma):
    a0, g, a0old, f = (np.asarray(v, np.float64) for v in (a0, g, a0old, f))
    Yt = a0old - a0
    out = np.empty_like(f, dtype=np.float64)
    for idx in np.ndindex(a0.shape):
        gv = g[idx].astype(np.float64)
        M = gamma * np.eye(2) + np.outer(gv, gv)
        out[idx] = np.linalg.solve(M, gamma * f[idx] + gv * Yt[idx])
    return out


def test_update_vs_linear_solve(ffo):
    r = rng(16)
    h, w = 9, 11
    a0 = r.uniform(0, 1, (h, w)).astype(np.float32)
    a0old = r.uniform(0, 1, (h, w)).astype(np.float32)
    g = r.normal(0, 0.05, (h, w, 2)).astype(np.float32)
    f = r.uniform(-1, 1, (h, w, 2)).astype(np.float32)
    gamma = float(np.float32(10 / 65025.0))  # the oracle receives gamma as a float
    ref = _solve_f64(a0, g, a0old, f, gamma)
    fo, io = ffo.update(a0, g, a0old, f, gamma, 1e6, "f64")
    assert np.abs(fo - ref).max() < 1e-9 * max(1, np.abs(ref).max())
    assert np.array_equal(io, a0)
    f32, _ = ffo.update(a0, g, a0old, f, gamma, 1e6, "f32")
    assert np.abs(f32 - ref).max() < 1e-4 * max(1, np.abs(ref).max())
    # Sherman-Morrison form: f + g (Yt - g.f) / (gamma + |g|^2)
    Yt = a0old.astype(np.float64) - a0.astype(np.float64)
    gd, fd = g.astype(np.float64), f.astype(np.float64)
    sm = fd + gd * ((Yt - (gd * fd).sum(-1)) / (gamma + (gd * gd).sum(-1)))[..., None]
    assert np.abs(fo - sm).max() < 1e-9 * max(1, np.abs(sm).max())


def test_update_special_cases(ffo):
    h, w = 4, 5
    a0 = np.full((h, w), 0.5, np.float32)
    f = rng(17).uniform(-3, 3, (h, w, 2)).astype(np.float32)
    zero_g = np.zeros((h, w, 2), np.float32)
    # g = 0, gamma = 1: f' = clamp(f, +-m) exactly
    fo, _ = ffo.update(a0, zero_g, a0, f, 1.0, 2.0, "f32")
    assert np.array_equal(fo, np.clip(f, -2, 2))
    # Yt = 0 and f = 0: zero
    fo, _ = ffo.update(a0, rng(18).standard_normal((h, w, 2)).astype(np.float32), a0,
                       np.zeros((h, w, 2), np.float32), 0.1, 4.0, "f32")
    assert np.array_equal(fo, np.zeros_like(fo))
    # NaN gradient: fminf/fmaxf map NaN to +maxflow (the sanitize step comes after the clamp)
    gn = zero_g.copy()
    gn[1, 2, 0] = np.nan
    fo, _ = ffo.update(a0, gn, a0, f, 1.0, 2.0, "f32")
    assert np.array_equal(fo[1, 2], [2.0, 2.0])
    # large temporal difference saturates at +-maxflow
    g1 = np.zeros((h, w, 2), np.float32)
    g1[..., 0] = 1.0
    fo, _ = ffo.update(a0, g1, a0 + 100, np.zeros((h, w, 2), np.float32), 1e-3, 4.0, "f32")
    assert np.all(fo[..., 0] == 4.0)


def test_delta_update_rules(ffo):
    h, w = 6, 8
    a0 = np.zeros((h, w), np.float32)
    g = np.zeros((h, w, 2), np.float32)
    up = np.zeros((3, 4, 2), np.float32)
    up[..., 0] = 1.0
    up[..., 1] = 1.5
    d = np.zeros((h, w, 2), np.float32)
    d[..., 0] = 10.0
    d[..., 1] = -10.0
    # g = 0, gamma = 1: d' = clamp(d, +-m/2); flow = clamp(d' + 2 up, +-m)
    do, fo, io = ffo.delta_update(a0, g, a0, d, up, 1.0, 4.0, "f32")
    assert np.all(do[..., 0] == 2.0) and np.all(do[..., 1] == -2.0)
    assert np.all(fo[..., 0] == 4.0) and np.all(fo[..., 1] == 1.0)
    # NaN delta flow is zeroed before the half clamp
    gn = g.copy()
    gn[2, 3, 1] = np.nan
    do, fo, _ = ffo.delta_update(a0, gn, a0, d, up, 1.0, 4.0, "f32")
    assert np.array_equal(do[2, 3], [0.0, 0.0]) and np.array_equal(fo[2, 3], [2.0, 3.0])


@pytest.mark.parametrize("n", [2048, 1088, 1024, 544, 640, 480, 320, 240, 512, 256, 3840, 2160, 1920, 1080,
                               1280, 720, 960, 540, 64, 32, 2])
def test_upper_index_is_half_for_even_sizes(ffo, n):
    nu = n // 2
    lib = ffo.lib()
    for x in range(n):
        assert lib.ffo_upper_index_f32(x, n, nu) == x >> 1
        assert lib.ffo_upper_index_f64(x, n, nu) == x >> 1


def test_delta_update_vs_linear_solve(ffo):
    r = rng(19)
    h, w = 8, 10
    a0 = r.uniform(0, 1, (h, w)).astype(np.float32)
    a0old = r.uniform(0, 1, (h, w)).astype(np.float32)
    g = r.normal(0, 0.05, (h, w, 2)).astype(np.float32)
    d = r.uniform(-0.5, 0.5, (h, w, 2)).astype(np.float32)
    up = r.uniform(-1, 1, (4, 5, 2)).astype(np.float32)
    gamma = float(np.float32(50 / 65025.0))
    ref = _solve_f64(a0, g, a0old, d, gamma)
    do, fo, _ = ffo.delta_update(a0, g, a0old, d, up, gamma, 1e6, "f64")
    assert np.abs(do - ref).max() < 1e-9 * max(1, np.abs(ref).max())
    yy, xx = np.mgrid[0:h, 0:w]
    assert np.allclose(fo, do + 2 * up[yy >> 1, xx >> 1], atol=1e-12)


# ------------------------------------------------------------------ smoothing

def test_smoothing_vs_uniform_filter(ffo):
    f = rng(20).standard_normal((17, 23, 2)).astype(np.float32)
    for K in (1, 2, 3):
        ref = f.astype(np.float64)
        for _ in range(K):
            ref = ndimage.uniform_filter1d(ref, 5, axis=1, mode="nearest")
            ref = ndimage.uniform_filter1d(ref, 5, axis=0, mode="nearest")
        out = ffo.smooth(f, K, "f64")
        # the reference weight is the float 0.2f, not 1/5
        assert np.abs(out - ref).max() < 1e-6
        assert np.abs(ffo.smooth(f, K, "f32") - out).max() < 1e-6


def test_smoothing_edge_and_constants(ffo):
    f = np.zeros((9, 9, 2), np.float32)
    f[:, :, 0] = np.arange(9, dtype=np.float32)[None]
    out = ffo.smooth(f, 1, "f64")
    w = np.float64(np.float32(0.2))
    # x = 0: taps at clamp(-2..2) = 0, 0, 0, 1, 2; the Y pass of a field
    # constant along y multiplies by 5 * 0.2f
    assert abs(out[4, 0, 0] - w * (0 + 0 + 0 + 1 + 2) * 5 * w) < 1e-12
    # interior of a linear field is preserved up to the 0.2f weight
    assert np.allclose(out[4, 2:-2, 0], 25 * w * w * np.arange(2, 7), atol=1e-12)
    c = np.full((7, 8, 2), 3.5, np.float32)
    assert np.abs(ffo.smooth(c, 2, "f32") - 3.5).max() < 4e-6


# -------------------------------------------------------------------- filters

def test_first_frame_gives_zero_flow(ffo):
    img = rng(21).integers(0, 256, (48, 64), dtype=np.uint8)
    f = ffo.FlowFilter(48, 64, 2, 4.0, 10.0)
    f.compute(img)
    assert np.array_equal(f.downloadFlow(), np.zeros((48, 64, 2), np.float32))
    a0, _ = ffo.image_model(img, "f32")
    assert np.array_equal(f.downloadImage(), a0)
    p = ffo.PyramidalFlowFilter(48, 64, 2)
    p.configure_like([10, 50], 4, [2, 2])
    p.compute(img)
    assert np.array_equal(p.downloadFlow(), np.zeros((48, 64, 2), np.float32))


def test_filter_recovers_translation(ffo):
    """End to end on FP64: the filter converges to the known translation."""
    from flowfilter import synthetic
    h, w, v = 96, 128, (0.7, -0.4)
    f = ffo.FlowFilter(h, w, 2, 4.0, 10.0, prec="f64")
    for img in synthetic.translating_noise(h, w, 40, v):
        f.compute(img)
    flow = f.downloadFlow()[16:-16, 16:-16]
    med = np.median(flow.reshape(-1, 2), axis=0)
    assert abs(med[0] - v[0]) < 0.05 and abs(med[1] - v[1]) < 0.05, med


def test_pyramidal_recovers_translation(ffo):
    """The pyramid tracks a 1.6 px/frame motion and reduces the single level's
    under-estimate of a 2.6 px/frame motion (linearisation bias)."""
    from flowfilter import synthetic
    h, w = 96, 128

    def run(v, levels):
        p = ffo.PyramidalFlowFilter(h, w, levels, prec="f64")
        p.configure_like([10, 50][:levels] if levels > 1 else [10], 4, [2] * levels)
        for img in synthetic.translating_noise(h, w, 40, v):
            p.compute(img)
        return np.median(p.downloadFlow()[16:-16, 16:-16].reshape(-1, 2), axis=0)

    med = run((1.6, -1.1), 2)
    assert abs(med[0] - 1.6) < 0.12 and abs(med[1] + 1.1) < 0.12, med
    big = np.array((2.6, -1.7))
    e1 = np.abs(run(tuple(big), 1) - big).max()
    e2 = np.abs(run(tuple(big), 2) - big).max()
    assert e2 < e1, (e1, e2)


def test_pyramidal_fp32_vs_fp64(ffo):
    from flowfilter import synthetic
    h, w = 64, 96
    a = ffo.PyramidalFlowFilter(h, w, 2, prec="f32")
    b = ffo.PyramidalFlowFilter(h, w, 2, prec="f64")
    for p in (a, b):
        p.configure_like([10, 50], 4, [2, 2])
    worst = 0.0
    for img in synthetic.translating_noise(h, w, 12, (1.3, 0.6)):
        a.compute(img)
        b.compute(img)
        worst = max(worst, float(np.abs(a.downloadFlow() - b.downloadFlow()).max()))
    assert worst < 1e-3, worst
