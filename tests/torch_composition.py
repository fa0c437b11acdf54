"""The differentiable filter written as a plain PyTorch composition of the
same stencils, in any floating dtype (float64 for the checks). Test
infrastructure: the gradients of flowfilter.autograd are checked against
torch autograd through this composition and against finite differences of
it in float64.

Each stage restates the reference kernel it cites; the composition of a
frame follows PyramidalFlowFilter::compute (flowfilter.cu:685-716).
"""
import numpy as np
import torch

SMOOTH = (0.0625, 0.25, 0.375, 0.25, 0.0625)   # imagemodel_k.cu:21
DIFF = (-0.125, -0.25, 0.0, 0.25, 0.125)       # imagemodel_k.cu:22
BOX = float(np.float32(0.2))                   # 0.2f, flowsmoothing_k.cu:19


class Branches(object):
    """Piecewise choices of one evaluation (upwind side and direction,
    clamps, non-finite masks, bilinear cells), recorded on the first
    evaluation under `frozen` and replayed on later ones: finite differences
    then measure the derivative of the smooth piece the base point lies on,
    which is what an adjoint computes, instead of a secant across the kinks
    that the perturbation crosses."""

    def __init__(self):
        self.log = []
        self.pos = None

    def decide(self, value):
        if self.pos is None:
            self.log.append(value)
            return value
        v = self.log[self.pos]
        self.pos += 1
        return v

    def replay(self):
        self.pos = 0


_BRANCHES = [None]


class frozen(object):
    """Context: route the composition's piecewise choices through `rec`."""

    def __init__(self, rec):
        self.rec = rec

    def __enter__(self):
        _BRANCHES[0] = self.rec
        if self.rec.log:
            self.rec.replay()
        return self.rec

    def __exit__(self, *exc):
        _BRANCHES[0] = None
        return False


def _decide(value):
    rec = _BRANCHES[0]
    return value if rec is None else rec.decide(value)


def _clamp(t, lo, hi):
    below, above = _decide(t < lo), _decide(t > hi)
    return torch.where(above, torch.full_like(t, hi), torch.where(below, torch.full_like(t, lo), t))


def sh(t, dim, k):
    """t shifted by k along dim with clamped (edge-replicating) reads:
    out[i] = t[clamp(i + k, 0, n - 1)]."""
    n = t.shape[dim]
    idx = torch.clamp(torch.arange(n, device=t.device) + k, 0, n - 1)
    return t.index_select(dim, idx)


def image_model(img):
    """imagePrefilter_k + imageModel_k (imagemodel_k.cu:31-133): a0 =
    Sx(Sy I), gx = Dx(Sy I), gy = Dy(Sx I) with clamped reads; img [B,H,W]."""
    sy = sum(w * sh(img, 1, r - 2) for r, w in enumerate(SMOOTH))
    sx = sum(w * sh(img, 2, c - 2) for c, w in enumerate(SMOOTH))
    a0 = sum(w * sh(sy, 2, c - 2) for c, w in enumerate(SMOOTH))
    gx = sum(w * sh(sy, 2, c - 2) for c, w in enumerate(DIFF))
    gy = sum(w * sh(sx, 1, r - 2) for r, w in enumerate(DIFF))
    return a0, torch.stack([gx, gy], dim=-1)


def down(img):
    """imageDownX_float_k then imageDownY_float_k (pyramid_k.cu:81-145):
    out[i] = 0.5 in[2i] + 0.25 (in[2i-1] + in[2i+1]), clamped reads."""
    for dim in (2, 1):
        n = img.shape[dim]
        i = torch.arange(n // 2, device=img.device)
        take = lambda k: img.index_select(dim, torch.clamp(2 * i + k, 0, n - 1))
        img = 0.5 * take(0) + 0.25 * (take(-1) + take(1))
    return img


def _interior(H, W, border, device):
    y = torch.arange(H, device=device)[:, None]
    x = torch.arange(W, device=device)[None, :]
    return ((x >= border) & (x < W - border) & (y >= border) & (y < H - border))[None]


def prop_pass(axis, f, s, v, dt, border):
    """One upwind pass along x (axis 0) or y (axis 1) of the flow f [B,H,W,2]
    and optional payloads s [B,H,W], v [B,H,W,2] (propagation_k.cu:16-334)."""
    dim, c = (2, 0) if axis == 0 else (1, 1)
    fm, fp = sh(f, dim, -1), sh(f, dim, 1)
    um, up = fm[..., c], fp[..., c]
    Ud = torch.where(_decide(up.abs() - um.abs() > 0), up, um)
    back = _decide(Ud >= 0)
    nd = -(dt * Ud)
    mask = _interior(f.shape[1], f.shape[2], border, f.device)

    def one(t, vec):
        tm, tp = sh(t, dim, -1), sh(t, dim, 1)
        b = back[..., None] if vec else back
        n = nd[..., None] if vec else nd
        o = t + n * torch.where(b, t - tm, tp - t)
        return torch.where(mask[..., None] if vec else mask, o, t)

    return one(f, True), (one(s, False) if s is not None else None), (one(v, True) if v is not None else None)


def solve2x2(g, Y, f, gamma):
    """Closed-form regularised solve (update_k.cu:43-63)."""
    gx, gy = g[..., 0], g[..., 1]
    ax2, ay2 = gx * gx, gy * gy
    N00, N01, N11 = gamma + ay2, -gx * gy, gamma + ax2
    rdet = 1.0 / (gamma * ((gamma + ax2) + ay2))
    qx = gamma * f[..., 0] + gx * Y
    qy = gamma * f[..., 1] + gy * Y
    return (N00 * qx + N01 * qy) * rdet, (N01 * qx + N11 * qy) * rdet


def _finite(t):
    return torch.where(_decide(torch.isfinite(t)), t, torch.zeros_like(t))


def update(a0, g, a0_old, f, gamma, maxflow):
    """flowUpdate_k (update_k.cu:15-79): solve, clamp to +-maxflow, zero
    non-finite values."""
    ox, oy = solve2x2(g, a0_old - a0, f, gamma)
    return torch.stack([_finite(_clamp(ox, -maxflow, maxflow)), _finite(_clamp(oy, -maxflow, maxflow))], -1)


def delta_update(a0, g, a0_prop, dflow_prop, upper, xmap, ymap, gamma, maxflow):
    """deltaFlowUpdate_k (update_k.cu:82-173): the delta flow (zeroed if
    non-finite, clamped to +-maxflow/2) and the level flow, delta + 2 x the
    upper flow point-sampled at (xmap, ymap), clamped to +-maxflow."""
    ox, oy = solve2x2(g, a0_prop - a0, dflow_prop, gamma)
    half = 0.5 * maxflow
    d = torch.stack([_clamp(_finite(ox), -half, half), _clamp(_finite(oy), -half, half)], -1)
    u = upper.index_select(1, ymap.long()).index_select(2, xmap.long())
    return d, _clamp(d + 2.0 * u, -maxflow, maxflow)


def smooth(f, K):
    """FlowSmoother (flowsmoothing_k.cu:18-87): K x (box-5 along x, then y)."""
    for _ in range(K):
        for dim in (2, 1):
            f = sum(BOX * sh(f, dim, k) for k in range(-2, 3))
    return f


def frame(spec, state, image, gammas, levels=None):
    """One frame of the pyramidal filter on float image [B,H,W]; gammas are
    the per-level gains in the update's units (tensors allowed). `levels`,
    if given, replaces the float pyramid: the list of level images."""
    L = spec.L
    imgs = [image]
    for _ in range(1, L):
        imgs.append(down(imgs[-1]))
    if levels is not None:
        imgs = levels
    models = [image_model(im) for im in imgs]
    if state is None:
        state = []
        for l in range(L):
            a0 = models[l][0]
            st = dict(f=torch.zeros_like(models[l][1]), a0=a0)
            if l < L - 1:
                st["d"] = torch.zeros_like(models[l][1])
            state.append(st)
    prop = []
    for l in range(L):
        lv = spec.levels[l]
        f = state[l]["f"]
        s = state[l]["a0"] if l < L - 1 else None
        v = state[l]["d"] if l < L - 1 else None
        for _ in range(lv["N"]):
            for axis in (0, 1):
                f, s, v = prop_pass(axis, f, s, v, lv["dt"], spec.border)
        prop.append((f, s, v))
    top = L - 1
    new = [None] * L
    lv = spec.levels[top]
    fu = update(models[top][0], models[top][1], state[top]["a0"], prop[top][0], gammas[top], lv["maxflow"])
    new[top] = dict(f=smooth(fu, lv["K"]), a0=models[top][0])
    for l in range(L - 1):
        lv = spec.levels[l]
        upper = fu if l == L - 2 else state[l + 1]["f"]
        d, fl = delta_update(models[l][0], models[l][1], prop[l][1], prop[l][2], upper, lv["xmap"], lv["ymap"],
                             gammas[l], lv["maxflow"])
        new[l] = dict(f=smooth(fl, lv["K"]), a0=models[l][0], d=d)
    return new


def window(spec, frames, gamma, pyramids=None):
    """Level-0 smoothed flow of every frame, [B,T,H,W,2]; frames [B,T,H,W]
    in [0, 1]; gamma [L] in setGamma units for u8 input. `pyramids[t]`, if
    given, is frame t's list of level images."""
    scale = 255.0 * 255.0
    if frames.dtype == torch.float32:
        gammas = [g / torch.tensor(np.float32(scale), device=gamma.device) for g in gamma]
    else:
        gammas = [g / scale for g in gamma]
    state, outs = None, []
    for t in range(frames.shape[1]):
        state = frame(spec, state, frames[:, t], gammas, None if pyramids is None else pyramids[t])
        outs.append(state[0]["f"])
    return torch.stack(outs, 1)


def warp_previous(prev, flow):
    """prev [B,H,W] sampled bilinearly at x - flow(x), coordinates clamped
    to the image (the warp of grid_sample with align_corners=True and
    border padding), with its cell choices routed through `frozen`."""
    B, H, W = prev.shape
    dev, dt = prev.device, prev.dtype
    xs = torch.arange(W, device=dev, dtype=dt)[None, None, :]
    ys = torch.arange(H, device=dev, dtype=dt)[None, :, None]
    sx = _clamp(xs - flow[..., 0], 0.0, W - 1.0)
    sy = _clamp(ys - flow[..., 1], 0.0, H - 1.0)
    x0, y0 = _decide(torch.floor(sx)), _decide(torch.floor(sy))
    wx, wy = sx - x0, sy - y0
    x0, y0 = x0.long(), y0.long()
    x1, y1 = torch.clamp(x0 + 1, max=W - 1), torch.clamp(y0 + 1, max=H - 1)
    flat = prev.reshape(B, H * W)
    at = lambda yy, xx: flat.gather(1, (yy * W + xx).reshape(B, -1)).reshape(B, H, W)
    top = (1 - wx) * at(y0, x0) + wx * at(y0, x1)
    bot = (1 - wx) * at(y1, x0) + wx * at(y1, x1)
    return (1 - wy) * top + wy * bot


def flow_loss(frames, flows, truth, photometric_weight=0.1):
    """The loss of flowfilter.autograd.flow_loss (truth [B,2] per sequence,
    or a field [B,H,W,2] / [B,T,H,W,2]) with warp_previous in place of
    grid_sample."""
    T = frames.shape[1]
    if truth.dim() == 2:
        truth = truth[:, None, None, None, :]
    elif truth.dim() == 4:
        truth = truth[:, None]
    else:
        truth = truth[:, 1:]
    l2 = ((flows[:, 1:] - truth) ** 2).mean()
    ph = sum(((frames[:, t] - warp_previous(frames[:, t - 1], flows[:, t])) ** 2).mean() for t in range(1, T))
    return l2 + photometric_weight * ph / (T - 1)
