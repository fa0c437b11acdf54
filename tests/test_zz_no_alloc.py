"""The allocation-free update adjoint (ff_ad_update_backward_noalloc) against
the allocating entry point (ff_ad_update_backward), through the C ABI.

The four gradient images must be bitwise equal. The gamma gradient is the
allocating entry point's per-workgroup partial sums, summed in double in the
documented fixed order (restated below in numpy), so it must be bitwise
equal to that restatement. Two calls in a row, the second on a smaller
input, show that nothing of the first call's scratch contents leaks into the
second.
"""
import ctypes
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fixed_order_sum(p):
    """Thread t of 256 adds partials t, t + 256, ... in sequence (from 0.0);
    then s[t] += s[t + k] for k = 128, 64, ..., 1."""
    p = np.asarray(p, np.float32)
    m = -(-p.size // 256)
    q = np.zeros(m * 256, np.float64)
    q[:p.size] = p
    s = np.zeros(256, np.float64)
    for row in q.reshape(m, 256):
        s = s + row
    k = 128
    while k:
        s[:k] = s[:k] + s[k:2 * k]
        k //= 2
    return s[0]


def _inputs(seed, B, H, W):
    r = np.random.default_rng(seed)
    rnd = lambda *shape, lo=-1.0, hi=1.0: torch.from_numpy(
        r.uniform(lo, hi, size=shape).astype(np.float32)).cuda()
    return (rnd(B, H, W, lo=0.0), rnd(B, H, W, 2, lo=-0.2, hi=0.2), rnd(B, H, W, lo=0.0),
            rnd(B, H, W, 2, lo=-3, hi=3), rnd(B, H, W, 2))


def _outs(B, H, W):
    return [torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W, 2, device="cuda"),
            torch.empty(B, H, W, device="cuda"), torch.empty(B, H, W, 2, device="cuda")]


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def test_noalloc_update_adjoint_matches_allocating():
    import flowfilter.autograd  # noqa: F401  (checks the torch/HIP runtime order)
    from flowfilter import _native as N
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    gam, mf = ctypes.c_float(float(np.float32(20) / np.float32(65025))), ctypes.c_float(4.0)
    for call, (B, H, W) in enumerate([(3, 200, 300), (2, 37, 70)]):
        say("no-alloc update adjoint, call", call, "shape", (B, H, W))
        a0, g, a0o, f, gout = _inputs(call, B, H, W)
        want = _outs(B, H, W)
        part = torch.empty(B * N.get_int("ff_ad_partial_count", H, W), device="cuda")
        N.call("ff_ad_update_backward", _p(a0), _p(g), _p(a0o), _p(f), _p(gout), *[_p(t) for t in want], _p(part),
               B, H, W, gam, mf, 0, s)
        got = _outs(B, H, W)
        g_gamma = torch.full((1,), 123.0, dtype=torch.float64, device="cuda")
        N.call("ff_ad_update_backward_noalloc", _p(a0), _p(g), _p(a0o), _p(f), _p(gout), *[_p(t) for t in got],
               _p(g_gamma), B, H, W, gam, mf, 0, s)
        torch.cuda.synchronize()
        for k in range(4):
            assert torch.equal(got[k], want[k]), (call, k)
        ref = fixed_order_sum(part.cpu().numpy())
        assert g_gamma.item() == ref, (call, g_gamma.item(), ref)
        assert abs(ref - part.double().sum().item()) <= 1e-12 * max(1.0, abs(ref))
        assert ref != 0.0
        # bit 4 of accumulate adds the gamma gradient; bits 0-3 as before
        pre = [torch.ones_like(t) for t in got]
        acc = [t.clone() for t in pre]
        g_acc = torch.full((1,), 0.5, dtype=torch.float64, device="cuda")
        N.call("ff_ad_update_backward_noalloc", _p(a0), _p(g), _p(a0o), _p(f), _p(gout), *[_p(t) for t in acc],
               _p(g_acc), B, H, W, gam, mf, 16 | 2, s)
        torch.cuda.synchronize()
        assert g_acc.item() == 0.5 + ref
        for k in range(4):
            assert torch.equal(acc[k], pre[k] + want[k] if k == 1 else want[k]), (call, k)


def test_noalloc_update_adjoint_rejects_oversize():
    """More partial sums than the scratch array holds: an error before any
    launch (the tiny tensors are never read)."""
    import flowfilter.autograd  # noqa: F401
    from flowfilter import _native as N
    from flowfilter.errors import InvalidArgument
    B, H, W = 4, 4096, 64 * 257  # 4 * 1024 * 257 > 2**20 partial sums
    assert B * N.get_int("ff_ad_partial_count", H, W) > 1 << 20
    t = torch.zeros(4, device="cuda")
    g_gamma = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(InvalidArgument, match="scratch"):
        N.call("ff_ad_update_backward_noalloc", *[_p(t)] * 9, _p(g_gamma), B, H, W, ctypes.c_float(1.0),
               ctypes.c_float(1.0), 0, None)
