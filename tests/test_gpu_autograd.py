"""Differentiable form (flowfilter.autograd): forward stages against the
CPU oracle, hand-written adjoints against torch autograd through the
pure-PyTorch composition (tests/torch_composition.py, itself pinned to the
oracle on the CPU), whole-window gradients against that composition and
against float64 central finite differences.

Bars:
  * forward stages: bitwise equal to the oracle's float32 stages; the
    single-level window bitwise equal to the oracle's u8 filter; deeper
    windows within 1e-4 px of the float64 composition;
  * stage adjoints: normwise relative error <= 1e-5 against float64 autograd;
  * window gradients (frames and per-level gamma): normwise relative error
    <= 1e-4 against float64 autograd of the composition;
  * finite differences (float64 composition, central, step 1e-4, its
    piecewise choices frozen at the base point) at 32 random frame
    coordinates and every gamma: |ad - fd| <= 1e-3 x max(|fd|, 1e-2 x
    max |fd|) per coordinate (a relative bound, floored at 1 % of the
    largest derivative checked so that near-zero derivatives are judged on
    an absolute scale);
  * the checkpoint interval changes nothing (bitwise).
"""
import sys

import numpy as np
import pytest
import torch

import torch_composition as tc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ad():
    import flowfilter.autograd as mod
    assert torch.cuda.is_available()
    return mod


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def rng(seed):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------- stages

def _call(ad, name, *args):
    from flowfilter import _native as N
    conv = [ad._p(a) if isinstance(a, torch.Tensor) else a for a in args]
    N.call(name, *conv, ad._stream())


def test_forward_stages_match_oracle(ad, ffo):
    from flowfilter import _native as N
    r = rng(1)
    B, H, W = 2, 45, 70
    img = r.uniform(0, 1, size=(B, H, W)).astype(np.float32)
    I = cuda(img)
    a0, g = torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W, 2, device="cuda")
    _call(ad, "ff_ad_image_model", I, a0, g, B, H, W)
    half, work = torch.empty(B, H // 2, W // 2, device="cuda"), torch.empty(B, H, W // 2, device="cuda")
    _call(ad, "ff_ad_down", I, half, work, B, H, W)
    f = r.uniform(-3, 3, size=(B, H, W, 2)).astype(np.float32)
    s = r.uniform(0, 1, size=(B, H, W)).astype(np.float32)
    v = r.uniform(-1, 1, size=(B, H, W, 2)).astype(np.float32)
    dt = float(np.float32(1) / np.float32(3))
    props = {}
    for axis in (0, 1):
        fo, so, vo = torch.empty_like(cuda(f)), torch.empty_like(cuda(s)), torch.empty_like(cuda(v))
        _call(ad, "ff_ad_propagate_pass", axis, cuda(f), cuda(s), cuda(v), fo, so, vo, B, H, W, ctypes_f(dt), 3)
        props[axis] = (fo, so, vo)
    gam = float(np.float32(10) / np.float32(65025))
    a0o = r.uniform(0, 1, size=(B, H, W)).astype(np.float32)
    gr = r.uniform(-0.2, 0.2, size=(B, H, W, 2)).astype(np.float32)
    fu = torch.empty(B, H, W, 2, device="cuda")
    _call(ad, "ff_ad_update", cuda(img), cuda(gr), cuda(a0o), cuda(f), fu, B, H, W, ctypes_f(gam), ctypes_f(4.0))
    Hu, Wu = H // 2, W // 2
    up = r.uniform(-2, 2, size=(B, Hu, Wu, 2)).astype(np.float32)
    xm, ym = cuda(ad._index_map(W, Wu)), cuda(ad._index_map(H, Hu))
    do, fl = torch.empty(B, H, W, 2, device="cuda"), torch.empty(B, H, W, 2, device="cuda")
    _call(ad, "ff_ad_delta_update", cuda(img), cuda(gr), cuda(a0o), cuda(v), cuda(up), xm, ym, do, fl, B, H, W, Hu, Wu,
          ctypes_f(gam), ctypes_f(4.0))
    sm1, sm2 = torch.empty_like(cuda(f)), torch.empty_like(cuda(f))
    _call(ad, "ff_ad_smooth_pass", 0, cuda(f), sm1, B, H, W)
    _call(ad, "ff_ad_smooth_pass", 1, sm1, sm2, B, H, W)
    torch.cuda.synchronize()
    for b in range(B):
        ra0, rg = ffo.image_model(img[b])
        assert np.array_equal(a0[b].cpu().numpy(), ra0) and np.array_equal(g[b].cpu().numpy(), rg), b
        assert np.array_equal(half[b].cpu().numpy(), ffo.pyramid(img[b], 2)[1]), b
        for axis in (0, 1):
            rf, rs, rv = ffo.propagate_pass(axis, f[b], s[b], v[b], dt=dt, border=3)
            fo, so, vo = props[axis]
            assert np.array_equal(fo[b].cpu().numpy(), rf), (axis, b)
            assert np.array_equal(so[b].cpu().numpy(), rs), (axis, b)
            assert np.array_equal(vo[b].cpu().numpy(), rv), (axis, b)
        rfu, _ = ffo.update(img[b], gr[b], a0o[b], f[b], gam, 4.0)
        assert np.array_equal(fu[b].cpu().numpy(), rfu), b
        rd, rfl, _ = ffo.delta_update(img[b], gr[b], a0o[b], v[b], up[b], gam, 4.0)
        assert np.array_equal(do[b].cpu().numpy(), rd), b
        assert np.array_equal(fl[b].cpu().numpy(), rfl), b
        assert np.array_equal(sm2[b].cpu().numpy(), ffo.smooth(f[b], 1)), b
    N.synchronize()


@pytest.mark.parametrize("shape", [(2, 45, 70), (1, 9, 130), (3, 31, 64), (1, 1, 5)])
@pytest.mark.parametrize("pay", [False, True])
def test_pass_pair_equals_two_passes(ad, shape, pay):
    """ff_ad_propagate_pair (X then Y in one launch, optionally keeping the
    X pass's result) is bitwise the two single passes."""
    B, H, W = shape
    r = rng(21 + H)
    f = cuda(r.uniform(-3, 3, size=(B, H, W, 2)).astype(np.float32))
    sc = cuda(r.uniform(0, 1, size=(B, H, W)).astype(np.float32)) if pay else None
    vc = cuda(r.uniform(-1, 1, size=(B, H, W, 2)).astype(np.float32)) if pay else None
    dt = ctypes_f(float(np.float32(1) / np.float32(3)))
    e2, e1 = (lambda: torch.empty(B, H, W, 2, device="cuda")), (lambda: torch.empty(B, H, W, device="cuda"))
    m = [e2(), e1() if pay else None, e2() if pay else None]
    o = [e2(), e1() if pay else None, e2() if pay else None]
    _call(ad, "ff_ad_propagate_pass", 0, f, sc, vc, *m, B, H, W, dt, 1)
    _call(ad, "ff_ad_propagate_pass", 1, *m, *o, B, H, W, dt, 1)
    pm = [e2(), e1() if pay else None, e2() if pay else None]
    po = [e2(), e1() if pay else None, e2() if pay else None]
    _call(ad, "ff_ad_propagate_pair", f, sc, vc, *pm, *po, B, H, W, dt, 1)
    qo = [e2(), e1() if pay else None, e2() if pay else None]
    _call(ad, "ff_ad_propagate_pair", f, sc, vc, None, None, None, *qo, B, H, W, dt, 1)
    torch.cuda.synchronize()
    for a, b_ in zip(m + o + o, pm + po + qo):
        if a is not None:
            assert torch.equal(a, b_)


@pytest.mark.parametrize("shape", [(2, 45, 70), (1, 9, 130), (3, 31, 64), (1, 1, 5), (1, 13, 62), (2, 25, 124),
                                   (1, 3, 3), (1, 40, 2), (1, 12, 63)])
@pytest.mark.parametrize("pay", [False, True])
def test_pass_quad_equals_two_pairs(ad, shape, pay):
    """ff_ad_propagate_quad (two iterations in one launch) is bitwise two
    ff_ad_propagate_pair launches."""
    B, H, W = shape
    r = rng(51 + H + W)
    f = cuda(r.uniform(-3, 3, size=(B, H, W, 2)).astype(np.float32))
    sc = cuda(r.uniform(0, 1, size=(B, H, W)).astype(np.float32)) if pay else None
    vc = cuda(r.uniform(-1, 1, size=(B, H, W, 2)).astype(np.float32)) if pay else None
    dt = ctypes_f(float(np.float32(1) / np.float32(3)))
    e2, e1 = (lambda: torch.empty(B, H, W, 2, device="cuda")), (lambda: torch.empty(B, H, W, device="cuda"))
    m = [e2(), e1() if pay else None, e2() if pay else None]
    o = [e2(), e1() if pay else None, e2() if pay else None]
    _call(ad, "ff_ad_propagate_pair", f, sc, vc, None, None, None, *m, B, H, W, dt, 2)
    _call(ad, "ff_ad_propagate_pair", *m, None, None, None, *o, B, H, W, dt, 2)
    q = [e2(), e1() if pay else None, e2() if pay else None]
    _call(ad, "ff_ad_propagate_quad", f, sc, vc, *q, B, H, W, dt, 2)
    torch.cuda.synchronize()
    for a, b_ in zip(o, q):
        if a is not None:
            assert torch.equal(a, b_)


@pytest.mark.parametrize("shape", [(2, 45, 70), (1, 9, 130), (3, 33, 62), (1, 1, 5), (2, 17, 117), (1, 40, 2),
                                   (1, 8, 58), (1, 3, 3)])
@pytest.mark.parametrize("pay", [False, True])
def test_pass_quad_adjoint_equals_two_pair_adjoints(ad, shape, pay):
    """ff_ad_propagate_quad_backward (recomputing its inner passes) is
    bitwise the adjoints of the two pairs, zero and accumulated output
    gradients included."""
    B, H, W = shape
    r = rng(61 + H + W)
    rnd = lambda *sh, lo=-1.0, hi=1.0: cuda(r.uniform(lo, hi, size=sh).astype(np.float32))
    f = rnd(B, H, W, 2, lo=-3, hi=3)
    sc, vc = (rnd(B, H, W, lo=0.0), rnd(B, H, W, 2)) if pay else (None, None)
    dt = ctypes_f(float(np.float32(1) / np.float32(4)))
    e2, e1 = (lambda: torch.empty(B, H, W, 2, device="cuda")), (lambda: torch.empty(B, H, W, device="cuda"))
    m = [e2(), e1() if pay else None, e2() if pay else None]  # the first pair's output
    _call(ad, "ff_ad_propagate_pair", f, sc, vc, None, None, None, *m, B, H, W, dt, 2)
    for zero_flow in (False, True):
        g = [None if zero_flow else rnd(B, H, W, 2), rnd(B, H, W) if pay else None, rnd(B, H, W, 2) if pay else None]
        gm = [e2(), e1() if pay else None, e2() if pay else None]
        ref = [rnd(B, H, W, 2), e1() if pay else None, e2() if pay else None]
        pre = ref[0].clone()
        _call(ad, "ff_ad_propagate_pair_backward", *m, None, None, None, *g, *gm, B, H, W, dt, 2, 0)
        _call(ad, "ff_ad_propagate_pair_backward", f, sc, vc, None, None, None, *gm, *ref, B, H, W, dt, 2, 1)
        got = [pre.clone(), e1() if pay else None, e2() if pay else None]
        _call(ad, "ff_ad_propagate_quad_backward", f, sc, vc, *g, *got, B, H, W, dt, 2, 1)
        torch.cuda.synchronize()
        for a, b_ in zip(ref, got):
            if a is not None:
                assert torch.equal(a, b_), zero_flow


@pytest.mark.parametrize("shape", [(2, 45, 70), (1, 9, 130), (3, 33, 62), (1, 1, 5), (2, 17, 121), (1, 40, 2)])
@pytest.mark.parametrize("pay", [False, True])
def test_pass_pair_adjoint_equals_two_adjoints(ad, shape, pay):
    """ff_ad_propagate_pair_backward is bitwise the Y pass's adjoint followed
    by the X pass's (zero and accumulated output gradients included), both
    given the X pass's result and recomputing it (NULL intermediates)."""
    B, H, W = shape
    r = rng(31 + H)
    rnd = lambda *sh, lo=-1.0, hi=1.0: cuda(r.uniform(lo, hi, size=sh).astype(np.float32))
    f = rnd(B, H, W, 2, lo=-3, hi=3)
    sc, vc = (rnd(B, H, W, lo=0.0), rnd(B, H, W, 2)) if pay else (None, None)
    dt = ctypes_f(float(np.float32(1) / np.float32(4)))
    e2, e1 = (lambda: torch.empty(B, H, W, 2, device="cuda")), (lambda: torch.empty(B, H, W, device="cuda"))
    m = [e2(), e1() if pay else None, e2() if pay else None]
    o = [e2(), e1() if pay else None, e2() if pay else None]
    _call(ad, "ff_ad_propagate_pair", f, sc, vc, *m, *o, B, H, W, dt, 2)
    for zero_flow in (False, True):
        g = [None if zero_flow else rnd(B, H, W, 2), rnd(B, H, W) if pay else None, rnd(B, H, W, 2) if pay else None]
        mid = [e2(), e1() if pay else None, e2() if pay else None]
        ref = [rnd(B, H, W, 2), e1() if pay else None, e2() if pay else None]
        pre = ref[0].clone()
        _call(ad, "ff_ad_propagate_pass_backward", 1, *m, *g, *mid, B, H, W, dt, 2, 0)
        _call(ad, "ff_ad_propagate_pass_backward", 0, f, sc, vc, *mid, *ref, B, H, W, dt, 2, 1)
        for mids in (m, [None, None, None]):
            got = [pre.clone(), e1() if pay else None, e2() if pay else None]
            _call(ad, "ff_ad_propagate_pair_backward", f, sc, vc, *mids, *g, *got, B, H, W, dt, 2, 1)
            torch.cuda.synchronize()
            for a, b_ in zip(ref, got):
                if a is not None:
                    assert torch.equal(a, b_), (zero_flow, mids[0] is None)


@pytest.mark.parametrize("shape", [(2, 45, 70), (1, 13, 130), (3, 33, 61), (1, 1, 5), (2, 4, 3), (1, 40, 2)])
def test_smooth_pair_equals_two_passes(ad, shape):
    """ff_ad_smooth_pair is bitwise an X pass then a Y pass, and its adjoint
    bitwise the Y pass's adjoint then the X pass's (with a second output
    gradient, and accumulating)."""
    B, H, W = shape
    r = rng(41 + H + W)
    rnd = lambda *sh: cuda(r.uniform(-2, 2, size=sh).astype(np.float32))
    e2 = lambda: torch.empty(B, H, W, 2, device="cuda")
    f = rnd(B, H, W, 2)
    mid, ref, got = e2(), e2(), e2()
    _call(ad, "ff_ad_smooth_pass", 0, f, mid, B, H, W)
    _call(ad, "ff_ad_smooth_pass", 1, mid, ref, B, H, W)
    _call(ad, "ff_ad_smooth_pair", f, got, B, H, W)
    torch.cuda.synchronize()
    assert torch.equal(ref, got)
    for second in (False, True):
        for acc in (0, 1):
            g, g2 = rnd(B, H, W, 2), (rnd(B, H, W, 2) if second else None)
            base = rnd(B, H, W, 2)
            ref, got = base.clone(), base.clone()
            _call(ad, "ff_ad_smooth_pass_backward", 1, g, g2, mid, B, H, W, 0)
            _call(ad, "ff_ad_smooth_pass_backward", 0, mid, None, ref, B, H, W, acc)
            _call(ad, "ff_ad_smooth_pair_backward", g, g2, got, B, H, W, acc)
            torch.cuda.synchronize()
            assert torch.equal(ref, got), (second, acc)


def ctypes_f(x):
    import ctypes
    return ctypes.c_float(x)


def _vjp64(fn, inputs, cot):
    """Gradients of sum(cot_i * out_i) w.r.t. inputs, float64 autograd."""
    xs = [x.detach().double().cpu().requires_grad_(True) for x in inputs]
    outs = fn(*xs)
    outs = outs if isinstance(outs, (tuple, list)) else (outs,)
    total = sum((o * c.double().cpu()).sum() for o, c in zip(outs, cot))
    return torch.autograd.grad(total, xs, allow_unused=True)


def test_stage_adjoints_match_autograd(ad):
    r = rng(2)
    B, H, W = 2, 38, 52
    rnd = lambda *shape, lo=-1.0, hi=1.0: cuda(r.uniform(lo, hi, size=shape).astype(np.float32))
    # image model
    I = rnd(B, H, W, lo=0.0)
    ga0, gg = rnd(B, H, W), rnd(B, H, W, 2)
    gI, work = torch.empty(B, H, W, device="cuda"), torch.empty(B, 2 * H, W, device="cuda")
    _call(ad, "ff_ad_image_model_backward", ga0, gg, gI, work, B, H, W, 0)
    (ref,) = _vjp64(lambda x: tc.image_model(x), [I], [ga0, gg])
    assert rel(gI, ref) <= 1e-5, ("image_model", rel(gI, ref))
    # pyramid step
    gh = rnd(B, H // 2, W // 2)
    gI2, work2 = torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W // 2, device="cuda")
    _call(ad, "ff_ad_down_backward", gh, gI2, work2, B, H, W, 0)
    (ref,) = _vjp64(tc.down, [I], [gh])
    assert rel(gI2, ref) <= 1e-5, ("down", rel(gI2, ref))
    # propagation passes, plain and with payloads, near ties and away
    dt = float(np.float32(1) / np.float32(4))
    for axis in (0, 1):
        for pay in (False, True):
            f, s, v = rnd(B, H, W, 2, lo=-4, hi=4), rnd(B, H, W, lo=0.0), rnd(B, H, W, 2)
            gfo, gso, gvo = rnd(B, H, W, 2), rnd(B, H, W), rnd(B, H, W, 2)
            gf, gs, gv = torch.empty_like(f), torch.empty_like(s), torch.empty_like(v)
            if pay:
                _call(ad, "ff_ad_propagate_pass_backward", axis, f, s, v, gfo, gso, gvo, gf, gs, gv, B, H, W,
                      ctypes_f(dt), 3, 0)
                refs = _vjp64(lambda a, b_, c: tc.prop_pass(axis, a, b_, c, dt, 3), [f, s, v], [gfo, gso, gvo])
                for mine, ref, nm in zip((gf, gs, gv), refs, "fsv"):
                    assert rel(mine, ref) <= 1e-5, ("prop", axis, nm, rel(mine, ref))
            else:
                _call(ad, "ff_ad_propagate_pass_backward", axis, f, None, None, gfo, None, None, gf, None, None, B, H,
                      W, ctypes_f(dt), 3, 0)
                (ref,) = _vjp64(lambda a: tc.prop_pass(axis, a, None, None, dt, 3)[0], [f], [gfo])
                assert rel(gf, ref) <= 1e-5, ("prop", axis, rel(gf, ref))
    # top update (with saturated pixels) and its gamma gradient
    a0, a0o = rnd(B, H, W, lo=0.0), rnd(B, H, W, lo=0.0)
    g = rnd(B, H, W, 2, lo=-0.2, hi=0.2)
    f = rnd(B, H, W, 2, lo=-5, hi=5)
    gam = float(np.float32(20) / np.float32(65025))
    gout = rnd(B, H, W, 2)
    ga0, gg, gao, gf = (torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W, 2, device="cuda"),
                        torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W, 2, device="cuda"))
    from flowfilter import _native as N
    nb = N.get_int("ff_ad_partial_count", H, W)
    part = torch.empty(B * nb, device="cuda")
    _call(ad, "ff_ad_update_backward", a0, g, a0o, f, gout, ga0, gg, gao, gf, part, B, H, W, ctypes_f(gam),
          ctypes_f(4.0), 0)
    gam64 = torch.tensor(gam, dtype=torch.float64)
    refs = _vjp64(lambda a, b_, c, d, e: tc.update(a, b_, c, d, e, 4.0), [a0, g, a0o, f, gam64], [gout])
    for mine, ref, nm in zip((ga0, gg, gao, gf), refs, ("a0", "grad", "a0_old", "flow")):
        assert rel(mine, ref) <= 1e-5, ("update", nm, rel(mine, ref))
    assert abs(float(part.double().sum()) - float(refs[4])) <= 1e-4 * abs(float(refs[4])), "update gamma"
    # delta update + the upper gather
    Hu, Wu = H // 2, W // 2
    sp, dp = rnd(B, H, W, lo=0.0), rnd(B, H, W, 2, lo=-3, hi=3)
    up = rnd(B, Hu, Wu, 2, lo=-2, hi=2)
    xm, ym = ad._index_map(W, Wu), ad._index_map(H, Hu)
    xlo, xhi = ad._ranges(xm, Wu)
    ylo, yhi = ad._ranges(ym, Hu)
    xm, ym, xlo, xhi, ylo, yhi = (cuda(x) for x in (xm, ym, xlo, xhi, ylo, yhi))
    gdo, gfo = rnd(B, H, W, 2), rnd(B, H, W, 2)
    ga0, gg, gsp, gdp, gup = (torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W, 2, device="cuda"),
                              torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W, 2, device="cuda"),
                              torch.empty(B, H, W, 2, device="cuda"))
    _call(ad, "ff_ad_delta_update_backward", a0, g, sp, dp, up, xm, ym, gdo, gfo, ga0, gg, gsp, gdp, gup, part, B, H,
          W, Hu, Wu, ctypes_f(gam), ctypes_f(4.0), 0)
    gupper = torch.empty(B, Hu, Wu, 2, device="cuda")
    _call(ad, "ff_ad_upper_backward", gup, xlo, xhi, ylo, yhi, gupper, B, H, W, Hu, Wu, 0)
    refs = _vjp64(lambda a, b_, c, d, e, gm: tc.delta_update(a, b_, c, d, e, xm.cpu(), ym.cpu(), gm, 4.0),
                  [a0, g, sp, dp, up, gam64], [gdo, gfo])
    for mine, ref, nm in zip((ga0, gg, gsp, gdp, gupper), refs, ("a0", "grad", "a0_prop", "dflow_prop", "upper")):
        assert rel(mine, ref) <= 1e-5, ("delta_update", nm, rel(mine, ref))
    assert abs(float(part.double().sum()) - float(refs[5])) <= 1e-4 * abs(float(refs[5])), "delta gamma"
    # smoothing passes
    for axis in (0, 1):
        gsm = rnd(B, H, W, 2)
        gin = torch.empty_like(gsm)
        _call(ad, "ff_ad_smooth_pass_backward", axis, gsm, None, gin, B, H, W, 0)
        dim = 2 if axis == 0 else 1
        (ref,) = _vjp64(lambda x: sum(tc.BOX * tc.sh(x, dim, k) for k in range(-2, 3)), [gsm], [gsm])
        assert rel(gin, ref) <= 1e-5, ("smooth", axis, rel(gin, ref))


def test_adjoint_accumulate_masks_and_zero_inputs(ad):
    """Bit k of `accumulate` adds output k in place; NULL output gradients
    read as zero; the smoothing adjoint of g1 + g2 equals that of the sum."""
    from flowfilter import _native as N
    r = rng(3)
    B, H, W = 2, 30, 44
    rnd = lambda *shape, lo=-1.0, hi=1.0: cuda(r.uniform(lo, hi, size=shape).astype(np.float32))
    a0, a0o, g, f, gout = (rnd(B, H, W, lo=0.0), rnd(B, H, W, lo=0.0), rnd(B, H, W, 2, lo=-0.2, hi=0.2),
                           rnd(B, H, W, 2, lo=-3, hi=3), rnd(B, H, W, 2))
    nb = N.get_int("ff_ad_partial_count", H, W)
    part = torch.empty(B * nb, device="cuda")
    gam = ctypes_f(float(np.float32(20) / np.float32(65025)))
    fresh = [torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W, 2, device="cuda"),
             torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W, 2, device="cuda")]
    _call(ad, "ff_ad_update_backward", a0, g, a0o, f, gout, *fresh, part, B, H, W, gam, ctypes_f(4.0), 0)
    for mask in (1, 2, 4, 8, 5, 15):
        pre = [rnd(*t.shape) for t in fresh]
        outs = [t.clone() for t in pre]
        _call(ad, "ff_ad_update_backward", a0, g, a0o, f, gout, *outs, part, B, H, W, gam, ctypes_f(4.0), mask)
        for k in range(4):
            want = pre[k] + fresh[k] if mask & (1 << k) else fresh[k]
            assert torch.equal(outs[k], want), (mask, k)
    # propagation: NULL output gradients are zeros
    s, v = rnd(B, H, W, lo=0.0), rnd(B, H, W, 2)
    zeros2, zeros1 = torch.zeros(B, H, W, 2, device="cuda"), torch.zeros(B, H, W, device="cuda")
    gso, gvo = rnd(B, H, W), rnd(B, H, W, 2)
    dt = ctypes_f(0.25)
    res = []
    for gfo in (None, zeros2):
        o = [torch.empty(B, H, W, 2, device="cuda"), torch.empty(B, H, W, device="cuda"),
             torch.empty(B, H, W, 2, device="cuda")]
        _call(ad, "ff_ad_propagate_pass_backward", 0, f, s, v, gfo, gso, gvo, *o, B, H, W, dt, 3, 0)
        res.append(o)
    assert all(torch.equal(x, y) for x, y in zip(*res))
    pre = rnd(B, H, W, 2)
    acc = pre.clone()
    _call(ad, "ff_ad_propagate_pass_backward", 0, f, s, v, None, gso, gvo, acc, res[0][1].clone(),
          res[0][2].clone(), B, H, W, dt, 3, 1)
    assert torch.equal(acc, pre + res[0][0])
    # smoothing with two inputs
    g1, g2 = rnd(B, H, W, 2), rnd(B, H, W, 2)
    a, b = torch.empty_like(g1), torch.empty_like(g1)
    _call(ad, "ff_ad_smooth_pass_backward", 1, g1, g2, a, B, H, W, 0)
    _call(ad, "ff_ad_smooth_pass_backward", 1, g1 + g2, None, b, B, H, W, 0)
    assert torch.equal(a, b)
    torch.cuda.synchronize()
    assert zeros1.abs().sum() == 0


# ---------------------------------------------------------------- windows

def _frames(B, T, H, W, seed, speed=2.0):
    from flowfilter import synthetic
    r = rng(seed)
    out, vel = [], []
    for b in range(B):
        v = tuple(r.uniform(-speed, speed, size=2))
        vel.append(v)
        out.append(np.stack(list(synthetic.translating_noise(H, W, T, v, seed=seed * 100 + b))))
    return (torch.from_numpy(np.stack(out).astype(np.float32) / np.float32(255.0)).cuda(),
            torch.tensor(vel, dtype=torch.float32).cuda())


def test_single_level_window_matches_oracle(ad, ffo):
    """L = 1 on u8-valued frames is the plain filter: bitwise equal."""
    B, T, H, W = 2, 5, 60, 84
    frames, _ = _frames(B, T, H, W, 3)
    spec = ad.FilterSpec(H, W, 1, maxflow=3.5, smooth=2)
    gamma = torch.tensor([12.0], device="cuda")
    with torch.no_grad():
        flows = ad.flow_window(frames, gamma, spec, checkpoint=2)
    torch.cuda.synchronize()
    u8 = np.rint(frames.cpu().numpy() * 255.0).astype(np.uint8)
    for b in range(B):
        o = ffo.FlowFilter(H, W, smoothIterations=2, maxflow=3.5, gamma=12.0)
        for t in range(T):
            o.compute(u8[b, t])
            assert np.array_equal(flows[b, t].cpu().numpy(), o.downloadFlow()), (b, t)


@pytest.mark.parametrize("levels", [2, 3])
def test_window_forward_matches_composition(ad, levels):
    B, T, H, W = 2, 5, 96, 128
    frames, _ = _frames(B, T, H, W, 4)
    spec = ad.FilterSpec(H, W, levels, maxflow=4.0, smooth=[2, 1, 1][:levels])
    gamma = torch.tensor([10.0, 30.0, 50.0][:levels], device="cuda")
    with torch.no_grad():
        mine = ad.flow_window(frames, gamma, spec)
        ref = tc.window(spec, frames.double(), gamma.double())
    err = float((mine.double() - ref).abs().max())
    say("levels", levels, "max |flow - float64 composition|", err)
    assert err <= 1e-4


def _gradients(ad, spec, frames, gamma, weights, checkpoint):
    x = frames.clone().requires_grad_(True)
    gm = gamma.clone().requires_grad_(True)
    flows = ad.flow_window(x, gm, spec, checkpoint=checkpoint)
    (flows * weights).sum().backward()
    return x.grad, gm.grad


@pytest.mark.parametrize("levels", [1, 2, 3])
def test_window_gradients_match_composition(ad, levels):
    B, T, H, W = 2, 5, 64, 96
    frames, _ = _frames(B, T, H, W, 5 + levels)
    spec = ad.FilterSpec(H, W, levels, maxflow=3.0, smooth=[2, 1, 1][:levels])
    gamma = torch.tensor([10.0, 30.0, 50.0][:levels], device="cuda")
    weights = torch.from_numpy(rng(6).standard_normal((B, T, H, W, 2)).astype(np.float32)).cuda()
    gx, gg = _gradients(ad, spec, frames, gamma, weights, checkpoint=2)
    x64 = frames.double().requires_grad_(True)
    g64 = gamma.double().requires_grad_(True)
    (tc.window(spec, x64, g64) * weights.double()).sum().backward()
    say("levels", levels, "frames rel", rel(gx, x64.grad), "gamma", gg.tolist(), g64.grad.tolist())
    assert rel(gx, x64.grad) <= 1e-4
    assert rel(gg, g64.grad) <= 1e-4


def _random_window(k):
    """Seeded window k of the sweep below: batch, frames, size (at least 8
    rows and columns on the top level), levels, maxflow, per-level
    smoothing, border, gains and checkpoint interval. maxflow stays well
    above the frames' motion (at most 1.5 px per frame along each axis):
    where an update lands within float32 rounding of the ±maxflow clamp,
    the clamp's derivative (0 or 1) is decided by rounding, and the kernels
    and the composition, which round in different orders, can each take
    either side (with motion close to maxflow such pixels occur; a
    1e-4 change of maxflow moves them off the threshold)."""
    r = np.random.default_rng(8800 + k)
    levels = int(r.integers(1, 4))
    lo = 8 << (levels - 1)
    H = int(r.integers(lo, max(lo + 1, 80)))
    W = int(r.integers(lo, max(lo + 1, 120)))
    B, T = int(r.integers(1, 4)), int(r.integers(2, 6))
    maxflow = float(np.round(r.uniform(4.5, 6.0), 2))
    smooth = [int(r.integers(1, 3)) for _ in range(levels)]
    border = int(r.integers(1, 4))
    gamma = [float(np.round(r.uniform(5.0, 60.0), 2)) for _ in range(levels)]
    checkpoint = int(r.integers(1, T + 1))
    return B, T, H, W, levels, maxflow, smooth, border, gamma, checkpoint


@pytest.mark.parametrize("k", range(16))
def test_window_gradients_random_configurations(ad, k):
    """Sixteen seeded windows (batch 1-3, 2-5 frames, odd and even sizes,
    1-3 levels, maxflow 4.5-6, per-level smoothing, border 1-3, any
    checkpoint interval): frame and gain gradients of a random weighting of
    the flows against float64 autograd of the composition, normwise within
    1e-4."""
    B, T, H, W, levels, maxflow, smooth, border, gamma, checkpoint = _random_window(k)
    say("random window", k, (B, T, H, W), levels, maxflow, smooth, border, gamma, checkpoint)
    frames, _ = _frames(B, T, H, W, 40 + k, speed=1.5)
    spec = ad.FilterSpec(H, W, levels, maxflow=maxflow, smooth=smooth, border=border)
    gm = torch.tensor(gamma, device="cuda")
    weights = torch.from_numpy(rng(50 + k).standard_normal((B, T, H, W, 2)).astype(np.float32)).cuda()
    gx, gg = _gradients(ad, spec, frames, gm, weights, checkpoint=checkpoint)
    x64 = frames.double().requires_grad_(True)
    g64 = gm.double().requires_grad_(True)
    (tc.window(spec, x64, g64) * weights.double()).sum().backward()
    say("   frames rel", rel(gx, x64.grad), "gamma rel", rel(gg, g64.grad))
    assert rel(gx, x64.grad) <= 1e-4
    assert rel(gg, g64.grad) <= 1e-4


def test_checkpoint_interval_changes_nothing(ad):
    B, T, H, W = 1, 6, 48, 64
    frames, _ = _frames(B, T, H, W, 9)
    spec = ad.FilterSpec(H, W, 2, maxflow=2.0, smooth=1)
    gamma = torch.tensor([10.0, 40.0], device="cuda")
    weights = torch.from_numpy(rng(7).standard_normal((B, T, H, W, 2)).astype(np.float32)).cuda()
    base = _gradients(ad, spec, frames, gamma, weights, checkpoint=T)
    for c in (1, 4):
        other = _gradients(ad, spec, frames, gamma, weights, checkpoint=c)
        assert torch.equal(base[0], other[0]) and torch.equal(base[1], other[1]), c


def _fd_check(ad, spec, frames, gamma, truth, seed, coords=32, h=1e-4, plain=False):
    """Gradient of flow_loss by the adjoint (float32) against float64
    central differences of the composition. The differences are taken with
    the composition's piecewise choices frozen at the base point (see
    torch_composition.Branches); with `plain`, also without freezing, which
    is reported only: a step of 1e-4 in one pixel moves flows by up to
    ~1e-2 px, enough to cross some of the dense kinks (upwind side, bilinear
    cell, clamps) that the secant then averages over."""
    x = frames.clone().requires_grad_(True)
    gm = gamma.clone().requires_grad_(True)
    loss = ad.flow_loss(x, ad.flow_window(x, gm, spec, checkpoint=4), truth)
    loss.backward()
    B, T, H, W = frames.shape
    r = rng(seed)
    pts = [(int(r.integers(B)), int(r.integers(T)), int(r.integers(H)), int(r.integers(W))) for _ in range(coords)]
    f64 = frames.double()
    g64 = gamma.double()
    t64 = truth.double()
    recs = {}

    def loss_b(b, fr, gam, freeze=True):
        # sequences are independent and every term is a mean over them:
        # the full loss is the mean over b of the per-sequence losses
        if not freeze:
            return float(tc.flow_loss(fr[b:b + 1], tc.window(spec, fr[b:b + 1], gam), t64[b:b + 1]))
        first = b not in recs
        rec = recs.setdefault(b, tc.Branches())
        with tc.frozen(rec):
            fl = tc.window(spec, fr[b:b + 1], gam)
            val = tc.flow_loss(fr[b:b + 1], fl, t64[b:b + 1])
        if first:  # the base point; the composition's loss is the product's
            ref = ad.flow_loss(fr[b:b + 1], fl, t64[b:b + 1])
            assert abs(float(val) - float(ref)) <= 1e-12 * abs(float(ref)), (float(val), float(ref))
        return float(val)

    fd, fdp, mine = [], [], []
    with torch.no_grad():
        for b in sorted(set(range(B))):
            loss_b(b, f64, g64)
        for (b, t, yy, xx) in pts:
            p, m = f64.clone(), f64.clone()
            p[b, t, yy, xx] += h
            m[b, t, yy, xx] -= h
            fd.append((loss_b(b, p, g64) - loss_b(b, m, g64)) / (2 * h) / B)
            if plain:
                fdp.append((loss_b(b, p, g64, False) - loss_b(b, m, g64, False)) / (2 * h) / B)
            mine.append(float(x.grad[b, t, yy, xx]))
        for l in range(spec.L):
            p, m = g64.clone(), g64.clone()
            p[l] += h
            m[l] -= h
            fd.append(sum(loss_b(b, f64, p) - loss_b(b, f64, m) for b in range(B)) / (2 * h) / B)
            mine.append(float(gm.grad[l]))
    fd, mine = np.array(fd), np.array(mine)
    nf = len(pts)
    scale_x = np.maximum(np.abs(fd[:nf]), 1e-2 * np.abs(fd[:nf]).max())
    err_x = np.abs(mine[:nf] - fd[:nf]) / scale_x
    err_g = np.abs(mine[nf:] - fd[nf:]) / np.abs(fd[nf:])
    say("fd (frozen choices): frames max rel", err_x.max(), "median", np.median(err_x), "gamma rel", err_g.tolist())
    if plain:
        fdp = np.array(fdp)
        ep = np.abs(mine[:nf] - fdp) / np.maximum(np.abs(fdp), 1e-2 * np.abs(fdp).max())
        say("fd (plain secant): frames max rel", ep.max(), "median", np.median(ep),
            "within 1e-3:", int((ep <= 1e-3).sum()), "of", nf)
    k = int(np.argmax(err_x))
    say("worst", pts[k], "fd", fd[k], "adjoint", mine[k], "max |fd|", np.abs(fd[:nf]).max())
    return err_x, err_g


@pytest.mark.parametrize("levels", [2, 3])
def test_gradients_match_finite_differences(ad, levels):
    B, T, H, W = 2, 5, 64, 80
    frames, vel = _frames(B, T, H, W, 11, speed=3.0)
    spec = ad.FilterSpec(H, W, levels, maxflow=4.0, smooth=2)
    gamma = torch.tensor([10.0, 50.0, 30.0][:levels], device="cuda")
    err_x, err_g = _fd_check(ad, spec, frames, gamma, vel, seed=12, plain=True)
    assert err_x.max() <= 1e-3
    assert err_g.max() <= 1e-3


def test_sixteen_sequence_workload_finite_differences(ad):
    """16 sequences of 512 x 512 noise translating at U[-3, 3]^2 px/frame,
    8 frames, 2 levels; loss: L2 to the true flow + 0.1 x photometric."""
    B, T, H, W = 16, 8, 512, 512
    frames, vel = _frames(B, T, H, W, 13, speed=3.0)
    spec = ad.FilterSpec(H, W, 2, maxflow=4.0, smooth=2)
    gamma = torch.tensor([10.0, 50.0], device="cuda")
    err_x, err_g = _fd_check(ad, spec, frames, gamma, vel, seed=14)
    assert err_x.max() <= 1e-3
    assert err_g.max() <= 1e-3


def test_errors(ad):
    spec = ad.FilterSpec(32, 32, 2)
    with pytest.raises(ValueError):
        ad.flow_window(torch.zeros(1, 2, 32, 32, device="cuda"), torch.ones(3, device="cuda"), spec)
    with pytest.raises(ValueError):
        ad.flow_window(torch.zeros(1, 2, 16, 32, device="cuda"), torch.ones(2, device="cuda"), spec)
    with pytest.raises(ValueError):
        ad.FilterSpec(32, 32, 2, border=0)
    with pytest.raises(ValueError):
        ad.flow_window(torch.zeros(1, 2, 32, 32, device="cuda"), torch.ones(2, device="cuda"), spec, checkpoint=0)
    # pair adjoint: intermediates all given or all NULL (recomputed)
    f, gf = torch.zeros(1, 8, 8, 2, device="cuda"), torch.zeros(1, 8, 8, 2, device="cuda")
    sc, gs = torch.zeros(1, 8, 8, device="cuda"), torch.zeros(1, 8, 8, device="cuda")
    dt = ctypes_f(0.25)
    with pytest.raises(ValueError):
        _call(ad, "ff_ad_propagate_pair_backward", f, sc, f, None, sc, None, gf, gs, gf, gf, gs, gf, 1, 8, 8, dt, 1, 0)
    with pytest.raises(ValueError):
        _call(ad, "ff_ad_propagate_pair_backward", f, sc, f, f, None, None, gf, gs, gf, gf, gs, gf, 1, 8, 8, dt, 1, 0)
    with pytest.raises(ValueError):
        _call(ad, "ff_ad_propagate_quad", f, None, None, f, None, None, 1, 8, 8, dt, 1)
    # smoothing pairs read neighbours: no output may be an input
    with pytest.raises(ValueError):
        _call(ad, "ff_ad_smooth_pair", f, f, 1, 8, 8)
    with pytest.raises(ValueError):
        _call(ad, "ff_ad_smooth_pair_backward", gf, None, gf, 1, 8, 8, 0)
