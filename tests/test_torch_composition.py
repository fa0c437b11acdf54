"""The pure-PyTorch composition used to check the differentiable form,
pinned against the CPU oracle (float64 builds of the oracle's stages and
of its pyramidal filter), on the CPU.

Bar: agreement to 1e-9 (absolute, float64 values of order 1) for every
stage and for whole pyramidal-filter sequences fed the oracle's own u8
pyramid; the float pyramid step to float32 rounding (2e-7)."""
import numpy as np
import pytest
import torch

import torch_composition as tc


def _spec(h, w, levels, maxflow=4.0, smooth=2):
    from flowfilter.autograd import FilterSpec
    return FilterSpec(h, w, levels, maxflow=maxflow, smooth=smooth, device="cpu")


def _rng(seed):
    return np.random.default_rng(seed)


def t64(a):
    return torch.from_numpy(np.asarray(a, dtype=np.float64))[None]


def close(a, b, tol=1e-9):
    a = a.detach().numpy()[0] if isinstance(a, torch.Tensor) else a
    assert a.shape == b.shape, (a.shape, b.shape)
    err = np.abs(a - b).max()
    assert err <= tol, err


def test_image_model_matches_oracle(ffo):
    img = _rng(1).uniform(0, 1, size=(37, 53)).astype(np.float32)
    a0, g = tc.image_model(t64(img))
    ra0, rg = ffo.image_model(img, prec="f64")
    close(a0, ra0)
    close(g, rg)


def test_down_matches_oracle(ffo):
    img = _rng(2).uniform(0, 1, size=(41, 66)).astype(np.float32)
    p = ffo.pyramid(img, 3, prec="f64")
    x = t64(img)
    for l in (1, 2):
        x = tc.down(x)
        close(x, p[l], 2e-7)
        x = t64(p[l])


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("payload", [False, True])
def test_propagation_pass_matches_oracle(ffo, axis, payload):
    r = _rng(3 + axis + 2 * payload)
    h, w = 29, 41
    f = r.uniform(-3, 3, size=(h, w, 2)).astype(np.float32)
    s = r.uniform(0, 1, size=(h, w)).astype(np.float32) if payload else None
    v = r.uniform(-1, 1, size=(h, w, 2)).astype(np.float32) if payload else None
    dt = float(np.float32(1) / np.float32(3))
    fo, so, vo = tc.prop_pass(axis, t64(f), t64(s) if payload else None, t64(v) if payload else None, dt, 3)
    rf, rs, rv = ffo.propagate_pass(axis, f, s, v, dt=dt, border=3, prec="f64")
    close(fo, rf)
    if payload:
        close(so, rs)
        close(vo, rv)


def test_update_matches_oracle(ffo):
    r = _rng(7)
    h, w = 23, 31
    a0 = r.uniform(0, 1, size=(h, w)).astype(np.float32)
    a0o = r.uniform(0, 1, size=(h, w)).astype(np.float32)
    g = r.uniform(-0.2, 0.2, size=(h, w, 2)).astype(np.float32)
    f = r.uniform(-5, 5, size=(h, w, 2)).astype(np.float32)
    gamma = float(np.float32(10.0) / np.float32(65025.0))
    out = tc.update(t64(a0), t64(g), t64(a0o), t64(f), gamma, 4.0)
    ref, _ = ffo.update(a0, g, a0o, f, gamma, 4.0, prec="f64")
    close(out, ref)


def test_delta_update_matches_oracle(ffo):
    from flowfilter.autograd import _index_map
    r = _rng(8)
    h, w, hu, wu = 24, 34, 12, 17
    # the float32 index rule and the oracle's exact one agree at these sizes
    assert all(_index_map(w, wu)[x] == ffo.upper_index(x, w, wu, prec="f64") for x in range(w))
    assert all(_index_map(h, hu)[y] == ffo.upper_index(y, h, hu, prec="f64") for y in range(h))
    a0 = r.uniform(0, 1, size=(h, w)).astype(np.float32)
    a0p = r.uniform(0, 1, size=(h, w)).astype(np.float32)
    g = r.uniform(-0.2, 0.2, size=(h, w, 2)).astype(np.float32)
    d = r.uniform(-3, 3, size=(h, w, 2)).astype(np.float32)
    up = r.uniform(-2, 2, size=(hu, wu, 2)).astype(np.float32)
    gamma = float(np.float32(50.0) / np.float32(65025.0))
    xm = torch.from_numpy(_index_map(w, wu))
    ym = torch.from_numpy(_index_map(h, hu))
    do, fo = tc.delta_update(t64(a0), t64(g), t64(a0p), t64(d), t64(up), xm, ym, gamma, 4.0)
    rd, rf, _ = ffo.delta_update(a0, g, a0p, d, up, gamma, 4.0, prec="f64")
    close(do, rd)
    close(fo, rf)


def test_smooth_matches_oracle(ffo):
    f = _rng(9).uniform(-4, 4, size=(19, 27, 2)).astype(np.float32)
    close(tc.smooth(t64(f), 3), ffo.smooth(f, 3, prec="f64"))


@pytest.mark.parametrize("levels", [1, 2, 3])
def test_pyramidal_sequence_matches_oracle(ffo, levels):
    """Whole filter over a sequence (level order, first-frame rules, the
    unsmoothed top flow read by level L-2) with the oracle's u8 pyramid."""
    from flowfilter import synthetic
    from flowfilter.autograd import _index_map
    h, w, n = 64, 96, 5
    gamma, maxflow, smooth = [10.0, 30.0, 50.0][:levels], 4.0, [2, 1, 1][:levels]
    spec = _spec(h, w, levels, maxflow, smooth)
    for l in range(levels - 1):
        hh, ww = spec.sizes[l]
        hu, wu = spec.sizes[l + 1]
        assert all(_index_map(ww, wu)[x] == ffo.upper_index(x, ww, wu, prec="f64") for x in range(ww))
        assert all(_index_map(hh, hu)[y] == ffo.upper_index(y, hh, hu, prec="f64") for y in range(hh))
    frames = list(synthetic.translating_noise(h, w, n, (1.3, -0.6), seed=5))
    o = ffo.PyramidalFlowFilter(h, w, levels, prec="f64")
    o.configure_like(gamma, maxflow, smooth)
    gam = torch.tensor(gamma, dtype=torch.float64)
    gammas = [float(np.float32(g) / np.float32(65025.0)) for g in gamma]
    state = None
    for t in range(n):
        pyr = [t64(p.astype(np.float64) / 255.0) for p in ffo.pyramid(frames[t], levels)]
        state = tc.frame(spec, state, pyr[0], gammas, levels=pyr)
        o.compute(frames[t])
        close(state[0]["f"], o.downloadFlow().astype(np.float64), 1e-9)
    assert gam.numel() == levels


def test_warp_matches_grid_sample():
    from flowfilter.autograd import photometric_error
    r = _rng(10)
    frames = torch.from_numpy(r.uniform(0, 1, size=(2, 3, 20, 30)))
    flows = torch.from_numpy(r.uniform(-4, 4, size=(2, 3, 20, 30, 2))).requires_grad_(True)
    a = photometric_error(frames, flows)
    ga, = torch.autograd.grad(a, flows)
    flows2 = flows.detach().clone().requires_grad_(True)
    b = sum(((frames[:, t] - tc.warp_previous(frames[:, t - 1], flows2[:, t])) ** 2).mean() for t in (1, 2)) / 2
    gb, = torch.autograd.grad(b, flows2)
    assert abs(float(a.detach()) - float(b.detach())) <= 1e-13
    assert float((ga - gb).abs().max()) <= 1e-12


def test_frozen_finite_differences_match_autograd():
    """Central differences with the piecewise choices frozen at the base
    point reproduce float64 autograd of the composition."""
    from flowfilter import synthetic
    h, w, n = 32, 40, 4
    spec = _spec(h, w, 2, maxflow=3.0, smooth=1)
    frames = torch.from_numpy(np.stack(list(synthetic.translating_noise(h, w, n, (1.2, -0.7), seed=3)))
                              .astype(np.float64)[None] / 255.0)
    gamma = torch.tensor([10.0, 40.0], dtype=torch.float64)
    truth = torch.tensor([[1.2, -0.7]], dtype=torch.float64)
    x = frames.clone().requires_grad_(True)
    tc.flow_loss(x, tc.window(spec, x, gamma), truth).backward()
    rec = tc.Branches()
    with tc.frozen(rec):
        tc.flow_loss(frames, tc.window(spec, frames, gamma), truth)
    r = _rng(11)
    step = 1e-4
    for _ in range(8):
        t, yy, xx = int(r.integers(n)), int(r.integers(h)), int(r.integers(w))
        vals = []
        for sgn in (1, -1):
            p = frames.clone()
            p[0, t, yy, xx] += sgn * step
            with tc.frozen(rec):
                vals.append(float(tc.flow_loss(p, tc.window(spec, p, gamma), truth)))
        fd = (vals[0] - vals[1]) / (2 * step)
        g = float(x.grad[0, t, yy, xx])
        assert abs(fd - g) <= 1e-5 * max(abs(g), 1e-3 * float(x.grad.abs().max())), (t, yy, xx, fd, g)


def test_upper_index_map_and_gather_ranges(ffo):
    """The differentiable form's upper-level index map equals the oracle's
    FP32 point-sampling rule (the filter library's) for even and odd sizes,
    and the exact rule too for even sizes (for odd ones float rounding can
    move a column to the next texel, e.g. 481 -> 240); the gather ranges of
    its adjoint (k_ad_upper_gather) hold exactly the columns (rows) that map
    to each upper texel."""
    from flowfilter.autograd import _index_map, _ranges
    for n in list(range(1, 40)) + [63, 64, 65, 127, 128, 255, 480, 481, 1088, 2047, 2048, 3840]:
        nu = max(n // 2, 1)
        m = _index_map(n, nu)
        assert all(int(m[x]) == ffo.upper_index(x, n, nu) for x in range(n)), n
        if n % 2 == 0:
            assert all(int(m[x]) == ffo.upper_index(x, n, nu, prec="f64") for x in range(n)), n
        lo, hi = _ranges(m, nu)
        for v in range(nu):
            want = [x for x in range(n) if m[x] == v]
            assert list(range(lo[v], hi[v])) == want, (n, v)
