"""The fused photometric term of the differentiable window's loss
(ff_ad_photometric and its adjoint, used by flowfilter.autograd for float32
tensors on the GPU) against float64 references on the CPU: the grid_sample
form of photometric_error and the composition's warp_previous
(tests/torch_composition.py), which agree with each other to 1e-12
(tests/test_torch_composition.py). Flows reach well outside the image, so
clamped coordinates (whose derivative is zero) are covered; odd sizes and
two sequences check the indexing of frame pairs.

Bars: the loss within 1e-6 relative; the gradients normwise within 1e-5
(float32 arithmetic against float64; the sampled pixels' gradients are
atomic adds, so the last bits vary from run to run, and twice the same call
must agree to 1e-6)."""
import numpy as np
import pytest
import torch

import torch_composition as tc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ad():
    import flowfilter.autograd as mod
    assert torch.cuda.is_available()
    return mod


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def _inputs(B, T, H, W, seed, reach=6.0):
    r = np.random.default_rng(seed)
    frames = r.uniform(0, 1, size=(B, T, H, W))
    flows = r.uniform(-reach, reach, size=(B, T, H, W, 2))
    return frames, flows


def _fused(ad, frames, flows):
    x = torch.tensor(frames, dtype=torch.float32, device="cuda", requires_grad=True)
    f = torch.tensor(flows, dtype=torch.float32, device="cuda", requires_grad=True)
    e = ad.photometric_error(x, f)
    e.backward()
    return e.detach(), x.grad, f.grad


def _reference(ad, frames, flows):
    # float64 on the CPU: the grid_sample form
    x = torch.tensor(np.asarray(frames, np.float32).astype(np.float64), requires_grad=True)
    f = torch.tensor(np.asarray(flows, np.float32).astype(np.float64), requires_grad=True)
    e = ad.photometric_error(x, f)
    e.backward()
    return e.detach(), x.grad, f.grad


@pytest.mark.parametrize("shape", [(1, 2, 16, 24), (2, 4, 37, 53), (1, 5, 64, 96)])
def test_fused_matches_grid_sample_float64(ad, shape):
    frames, flows = _inputs(*shape, seed=sum(shape))
    e, gx, gf = _fused(ad, frames, flows)
    e64, gx64, gf64 = _reference(ad, frames, flows)
    assert abs(float(e) - float(e64)) <= 1e-6 * abs(float(e64))
    assert rel(gx, gx64) <= 1e-5, rel(gx, gx64)
    assert rel(gf, gf64) <= 1e-5, rel(gf, gf64)
    # frame 0's flow is not read: its gradient is exactly zero
    assert float(gf[:, 0].abs().max()) == 0.0


def test_fused_matches_composition_warp(ad):
    B, T, H, W = 2, 3, 20, 30
    frames, flows = _inputs(B, T, H, W, seed=10, reach=4.0)
    e, gx, gf = _fused(ad, frames, flows)
    x = torch.tensor(np.asarray(frames, np.float32).astype(np.float64), requires_grad=True)
    f = torch.tensor(np.asarray(flows, np.float32).astype(np.float64), requires_grad=True)
    ref = sum(((x[:, t] - tc.warp_previous(x[:, t - 1], f[:, t])) ** 2).mean() for t in range(1, T)) / (T - 1)
    ref.backward()
    assert abs(float(e) - float(ref)) <= 1e-6 * abs(float(ref))
    assert rel(gx, x.grad) <= 1e-5 and rel(gf, f.grad) <= 1e-5


def test_zero_flow_and_identical_frames(ad):
    """Zero flow samples frame t-1 at the pixel itself: the term is the mean
    squared frame difference exactly, and identical frames give zero loss and
    zero gradients."""
    B, T, H, W = 1, 3, 12, 18
    r = np.random.default_rng(3)
    frames = r.integers(0, 256, size=(B, T, H, W)) / 256.0  # dyadic: exact differences
    flows = np.zeros((B, T, H, W, 2))
    e, gx, gf = _fused(ad, frames, flows)
    want = np.mean((frames[:, 1:] - frames[:, :-1]) ** 2)
    assert abs(float(e) - want) <= 1e-7 * want
    same = np.repeat(frames[:, :1], T, axis=1)
    e, gx, gf = _fused(ad, same, flows)
    assert float(e) == 0.0 and float(gx.abs().max()) == 0.0 and float(gf.abs().max()) == 0.0


def test_repeat_agrees(ad):
    frames, flows = _inputs(1, 4, 48, 64, seed=5)
    a = _fused(ad, frames, flows)
    b = _fused(ad, frames, flows)
    assert float(a[0]) == float(b[0])  # the partial sums are added in a fixed order
    assert rel(a[1], b[1]) <= 1e-6 and torch.equal(a[2], b[2])


def test_errors(ad):
    from flowfilter import _native as N
    from flowfilter.errors import InvalidArgument
    with pytest.raises(InvalidArgument):
        N.get_int("ff_ad_photometric_partial_count", 1, 1, 8, 8)
    x = torch.zeros(1, 2, 8, 8, device="cuda")
    with pytest.raises(InvalidArgument):
        N.call("ff_ad_photometric_backward", ad._p(x), ad._p(x), ad._p(x), ad._p(x), ad._p(x), 1, 2, 8, 8,
               ad._stream())
