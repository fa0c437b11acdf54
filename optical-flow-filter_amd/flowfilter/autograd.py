"""flowfilter.autograd -- the pyramidal filter as a differentiable function of
its input frames and per-level gains (no counterpart in the reference).

``flow_window(frames, gamma, ...)`` runs the pyramidal filter over a window
of frames for a batch of sequences and returns the level-0 smoothed flow of
every frame; it is a ``torch.autograd.Function`` whose backward pass runs
hand-written adjoint kernels of every stage (image model, float pyramid,
upwind propagation with payloads, top and delta updates, box smoothing)
through the library's C ABI (``ff_ad_*``).

The filter is the one of PyramidalFlowFilter (flowfilter.cu:621-897) on
float images in [0, 1] (a u8 frame divided by 255), with the float pyramid
(imageDown{X,Y}_float_k) in place of the u8 one, and ``gamma`` in the units
of setGamma for u8 input (divided by 255^2 internally, flowfilter.cu:
276-285). The first frame starts from zero flow with the old brightness
equal to the frame's own (the reference's first-load copies), so its flow
is zero.

Memory: the forward keeps only the filter state every ``checkpoint`` frames;
the backward recomputes each segment of frames from its checkpoint,
keeping that segment's intermediates (image models, upper flows, and the
input of every two propagation iterations: their adjoint kernel reruns the
passes in between), and runs the adjoints frame by frame in reverse.

PyTorch must be imported before the library is loaded into the process:
torch's HIP runtime and the system one share a soname, so the library then
binds to torch's runtime and both use one set of devices and streams. This
module imports torch first; import it before any other flowfilter module.
"""
import ctypes
import math
import sys

if "torch" not in sys.modules and "flowfilter._native" in sys.modules:
    raise ImportError("flowfilter.autograd: import torch (or this module) before any other flowfilter module; "
                      "the library must bind to torch's HIP runtime")

import numpy as np
import torch

from . import _native as N

__all__ = ["FilterSpec", "flow_window", "photometric_error", "flow_loss"]

_U8_SCALE = 255.0 * 255.0


def _index_map(n, nu):
    """Upper-level texel index of every column (row): the normalized-
    coordinate point-sampling rule, as the library computes it."""
    x = np.arange(n, dtype=np.float32)
    if n <= 1:
        return np.zeros(n, dtype=np.int32)
    u = x / np.float32(n - 1)
    i = np.floor(u * np.float32(nu)).astype(np.int64)
    return np.clip(i, 0, nu - 1).astype(np.int32)


def _ranges(m, nu):
    lo = np.zeros(nu, dtype=np.int32)
    hi = np.zeros(nu, dtype=np.int32)
    for x, v in enumerate(m):
        if hi[v] == 0:
            lo[v] = x
        hi[v] = x + 1
    return lo, hi


class FilterSpec(object):
    """Shape and fixed parameters of the differentiable filter: image size,
    levels, maxflow (halved per level in float32 as in setMaxFlow,
    flowfilter.cu:849-866; ceil(maxflow) propagation iterations with
    dt = 1/iterations per level), smoothing iterations per level and the
    propagation border."""

    def __init__(self, height, width, levels, maxflow=1.0, smooth=1, border=3, device="cuda"):
        if levels < 1:
            raise ValueError("levels should be greater than zero")
        if border < 1:
            raise ValueError("the differentiable filter needs a propagation border of at least 1")
        self.H, self.W, self.L = int(height), int(width), int(levels)
        smooth = list(smooth) if isinstance(smooth, (list, tuple)) else [int(smooth)] * self.L
        if len(smooth) != self.L or min(smooth) < 0:
            raise ValueError("one non-negative smoothing iteration count per level expected")
        if maxflow <= 0:
            raise ValueError("maxflow must be positive")
        self.border = int(border)
        self.maxflow = float(maxflow)
        self.sizes = [(self.H, self.W)]
        for _ in range(1, self.L):
            h, w = self.sizes[-1]
            self.sizes.append((h // 2, w // 2))
        if min(min(hw) for hw in self.sizes) < 1:
            raise ValueError("the image is too small for %d levels" % self.L)
        ml = np.float32(maxflow)
        self.levels = []
        for l in range(self.L):
            n = max(1, int(math.ceil(float(ml))))
            self.levels.append(dict(H=self.sizes[l][0], W=self.sizes[l][1], maxflow=float(ml), N=n,
                                    dt=float(np.float32(1.0) / np.float32(n)), K=int(smooth[l])))
            if l < self.L - 1:
                ml = np.float32(ml / np.float32(2.0))
        self.device = torch.device(device)
        for l in range(self.L - 1):
            h, w = self.sizes[l]
            hu, wu = self.sizes[l + 1]
            xm, ym = _index_map(w, wu), _index_map(h, hu)
            xlo, xhi = _ranges(xm, wu)
            ylo, yhi = _ranges(ym, hu)
            lv = self.levels[l]
            for k, v in (("xmap", xm), ("ymap", ym), ("xlo", xlo), ("xhi", xhi), ("ylo", ylo), ("yhi", yhi)):
                lv[k] = torch.from_numpy(v).to(self.device)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _stream_device():
    return torch.device("cuda", torch.cuda.current_device())


def _empty(B, h, w, c=None, like=None):
    shape = (B, h, w) if c is None else (B, h, w, c)
    return torch.empty(shape, dtype=torch.float32, device=like.device)


def _effective_gammas(gamma):
    """Per-level gains in the update's units: setGamma's u8 scaling,
    gamma / (255 * 255) in float32 (flowfilter.cu:276-285)."""
    g = gamma.detach().to("cpu", torch.float32).numpy().reshape(-1)
    return [float(np.float32(v) / np.float32(_U8_SCALE)) for v in g]


# ------------------------------------------------------------ one frame

def _frame_forward(spec, state, image, gammas, save):
    """One frame for all levels. state: per level dict(f, a0[, d]) or None
    for the first frame. Returns (new_state, output_flow, saved)."""
    B = image.shape[0]
    L = spec.L
    s = _stream()
    imgs = [image]
    for l in range(1, L):
        h, w = spec.sizes[l - 1]
        out = _empty(B, h // 2, w // 2, like=image)
        work = _empty(B, h, w // 2, like=image)
        N.call("ff_ad_down", _p(imgs[-1]), _p(out), _p(work), B, h, w, s)
        imgs.append(out)
    a0, g = [], []
    for l in range(L):
        h, w = spec.sizes[l]
        a, gg = _empty(B, h, w, like=image), _empty(B, h, w, 2, like=image)
        N.call("ff_ad_image_model", _p(imgs[l]), _p(a), _p(gg), B, h, w, s)
        a0.append(a)
        g.append(gg)
    first = state is None
    if first:
        state = []
        for l in range(L):
            h, w = spec.sizes[l]
            st = dict(f=torch.zeros(B, h, w, 2, device=image.device), a0=a0[l].clone())
            if l < L - 1:
                st["d"] = torch.zeros(B, h, w, 2, device=image.device)
            state.append(st)
    saved = dict(first=first, a0=a0, g=g, passes=[None] * L, prop=[None] * L, upper=[None] * L,
                 state=state) if save else None
    top = L - 1
    prop = [None] * L
    for l in range(L):
        h, w = spec.sizes[l]
        lv = spec.levels[l]
        f = state[l]["f"]
        sc = state[l]["a0"] if l < top else None
        vc = state[l]["d"] if l < top else None
        passes = []
        pay = sc is not None
        it = 0
        while it < lv["N"]:
            fo = _empty(B, h, w, 2, like=image)
            so = _empty(B, h, w, like=image) if pay else None
            vo = _empty(B, h, w, 2, like=image) if pay else None
            # two iterations per launch, then one if N is odd; only their
            # input is kept: the adjoints recompute the passes inside
            n = 2 if it + 2 <= lv["N"] else 1
            if n == 2:
                N.call("ff_ad_propagate_quad", _p(f), _p(sc), _p(vc), _p(fo), _p(so), _p(vo), B, h, w,
                       ctypes.c_float(lv["dt"]), spec.border, s)
            else:
                N.call("ff_ad_propagate_pair", _p(f), _p(sc), _p(vc), None, None, None, _p(fo), _p(so), _p(vo),
                       B, h, w, ctypes.c_float(lv["dt"]), spec.border, s)
            if save:
                passes.append((n, f, sc, vc))
            it += n
            f, sc, vc = fo, so, vo
        prop[l] = (f, sc, vc)
        if save:
            saved["passes"][l] = passes
            saved["prop"][l] = prop[l]
    new_state = [None] * L
    upd = [None] * L
    # top level (flowfilter.cu:685-716: the top level updates first)
    h, w = spec.sizes[top]
    lv = spec.levels[top]
    fu = _empty(B, h, w, 2, like=image)
    N.call("ff_ad_update", _p(a0[top]), _p(g[top]), _p(state[top]["a0"]), _p(prop[top][0]), _p(fu), B, h, w,
           ctypes.c_float(gammas[top]), ctypes.c_float(lv["maxflow"]), s)
    upd[top] = fu
    new_state[top] = dict(f=_smooth(fu, lv["K"], B, h, w, s), a0=a0[top])
    # levels 0 .. L-2 in ascending order: level L-2 reads the top level's
    # unsmoothed flow of this frame, lower levels the level above's
    # smoothed flow from before its update
    for l in range(L - 1):
        h, w = spec.sizes[l]
        hu, wu = spec.sizes[l + 1]
        lv = spec.levels[l]
        upper = upd[top] if l == L - 2 else state[l + 1]["f"]
        dn, fl = _empty(B, h, w, 2, like=image), _empty(B, h, w, 2, like=image)
        N.call("ff_ad_delta_update", _p(a0[l]), _p(g[l]), _p(prop[l][1]), _p(prop[l][2]), _p(upper),
               _p(lv["xmap"]), _p(lv["ymap"]), _p(dn), _p(fl), B, h, w, hu, wu, ctypes.c_float(gammas[l]),
               ctypes.c_float(lv["maxflow"]), s)
        upd[l] = fl
        new_state[l] = dict(f=_smooth(fl, lv["K"], B, h, w, s), a0=a0[l], d=dn)
        if save:
            saved["upper"][l] = upper
    if save:
        saved["upd"] = upd
    return new_state, new_state[0]["f"], saved


def _smooth(f, K, B, h, w, s):
    for _ in range(K):  # an X and a Y pass per launch
        o = torch.empty_like(f)
        N.call("ff_ad_smooth_pair", _p(f), _p(o), B, h, w, s)
        f = o
    return f


def _smooth_bwd(parts, K, B, h, w, s, out=None):
    """Adjoint of K smoothing iterations for the output gradient sum(parts)
    (one or two tensors, summed inside the first pass). The result is
    written to a new tensor, or added to `out`."""
    parts = [p for p in parts if p is not None]
    if not parts:  # no gradient reaches the smoothed flow
        return out if out is not None else torch.zeros(B, h, w, 2, device=_stream_device())
    if K == 0:
        total = parts[0] if len(parts) == 1 else parts[0] + parts[1]
        if out is None:
            return total
        out.add_(total)
        return out
    g, g2 = parts[0], (parts[1] if len(parts) > 1 else None)
    for i in range(K):  # the Y pass's and then the X pass's adjoint per launch
        last = i == K - 1
        o = out if (last and out is not None) else torch.empty_like(g)
        N.call("ff_ad_smooth_pair_backward", _p(g), _p(g2), _p(o), B, h, w, 1 if (last and out is not None) else 0, s)
        g, g2 = o, None
    return g


def _frame_backward(spec, saved, g_new, g_out, gammas, g_image_out):
    """Adjoint of one frame. g_new: per level dict of gradients of the new
    state (missing entries are zero; the tensors are consumed: the adjoint
    adds into them in place); g_out: gradient of the output flow (level 0's
    smoothed flow); g_image_out: where the frame's image gradient goes.
    Returns (g_old_state or None for the first frame, g_gamma_eff [L])."""
    L = spec.L
    top = L - 1
    a0, g = saved["a0"], saved["g"]
    B = a0[0].shape[0]
    s = _stream()
    dev = a0[0].device
    g_old = [dict() for _ in range(L)]
    g_gamma = torch.zeros(L, dtype=torch.float64, device=dev)
    g_a0 = [None] * L
    g_g = [None] * L
    g_upd_top = None
    new = (lambda l, k: g_new[l].get(k)) if g_new is not None else (lambda l, k: None)
    f32 = ctypes.c_float

    for l in range(L - 1):
        h, w = spec.sizes[l]
        hu, wu = spec.sizes[l + 1]
        lv = spec.levels[l]
        gfl = _smooth_bwd([new(l, "f"), g_out if l == 0 else None], lv["K"], B, h, w, s)
        ga = new(l, "a0")
        mask = 1 if ga is not None else 0
        if ga is None:
            ga = _empty(B, h, w, like=gfl)
        gg, gsp, gvp, gup = (_empty(B, h, w, 2, like=gfl), _empty(B, h, w, like=gfl), _empty(B, h, w, 2, like=gfl),
                             _empty(B, h, w, 2, like=gfl))
        part = torch.empty(B * N.get_int("ff_ad_partial_count", h, w), device=dev)
        f, sc, vc = saved["prop"][l]
        N.call("ff_ad_delta_update_backward", _p(a0[l]), _p(g[l]), _p(sc), _p(vc), _p(saved["upper"][l]),
               _p(lv["xmap"]), _p(lv["ymap"]), _p(new(l, "d")), _p(gfl), _p(ga), _p(gg), _p(gsp), _p(gvp), _p(gup),
               _p(part), B, h, w, hu, wu, f32(gammas[l]), f32(lv["maxflow"]), mask, s)
        g_gamma[l] += part.double().sum()
        # the upper flow: this frame's top update (l = L-2) or the old
        # smoothed flow of level l+1 (whose own adjoint comes later and
        # adds into it)
        gupper = _empty(B, hu, wu, 2, like=gfl)
        N.call("ff_ad_upper_backward", _p(gup), _p(lv["xlo"]), _p(lv["xhi"]), _p(lv["ylo"]), _p(lv["yhi"]),
               _p(gupper), B, h, w, hu, wu, 0, s)
        if l == L - 2:
            g_upd_top = gupper
        else:
            g_old[l + 1]["f"] = gupper
        g_a0[l] = ga
        g_g[l] = gg
        # propagation pass pairs in reverse; the propagated flow itself is
        # not read by the delta update, so its gradient enters as zero (None)
        gf, gs, gv = None, gsp, gvp
        passes = saved["passes"][l]
        for i, (n, fi, si, vi) in enumerate(reversed(passes)):
            acc_f = i == len(passes) - 1 and "f" in g_old[l]
            nf = g_old[l]["f"] if acc_f else _empty(B, h, w, 2, like=gfl)
            ns, nv = _empty(B, h, w, like=gfl), _empty(B, h, w, 2, like=gfl)
            if n == 2:
                N.call("ff_ad_propagate_quad_backward", _p(fi), _p(si), _p(vi), _p(gf), _p(gs), _p(gv), _p(nf),
                       _p(ns), _p(nv), B, h, w, f32(lv["dt"]), spec.border, 1 if acc_f else 0, s)
            else:
                N.call("ff_ad_propagate_pair_backward", _p(fi), _p(si), _p(vi), None, None, None, _p(gf), _p(gs),
                       _p(gv), _p(nf), _p(ns), _p(nv), B, h, w, f32(lv["dt"]), spec.border, 1 if acc_f else 0, s)
            gf, gs, gv = nf, ns, nv
        g_old[l]["f"] = gf
        g_old[l]["a0"] = gs
        g_old[l]["d"] = gv
    # top level
    h, w = spec.sizes[top]
    lv = spec.levels[top]
    g_upd_top = _smooth_bwd([new(top, "f"), g_out if L == 1 else None], lv["K"], B, h, w, s, out=g_upd_top)
    ga = new(top, "a0")
    mask = 1 if ga is not None else 0
    if ga is None:
        ga = _empty(B, h, w, like=g_upd_top)
    gg, gao, gfp = (_empty(B, h, w, 2, like=g_upd_top), _empty(B, h, w, like=g_upd_top),
                    _empty(B, h, w, 2, like=g_upd_top))
    part = torch.empty(B * N.get_int("ff_ad_partial_count", h, w), device=dev)
    st = saved["state"][top]
    N.call("ff_ad_update_backward", _p(a0[top]), _p(g[top]), _p(st["a0"]), _p(saved["prop"][top][0]),
           _p(g_upd_top), _p(ga), _p(gg), _p(gao), _p(gfp), _p(part), B, h, w, f32(gammas[top]),
           f32(lv["maxflow"]), mask, s)
    g_gamma[top] += part.double().sum()
    g_a0[top] = ga
    g_g[top] = gg
    g_old[top]["a0"] = gao
    gf = gfp
    for n, fi, _, _ in reversed(saved["passes"][top]):
        nf = _empty(B, h, w, 2, like=gf)
        if n == 2:
            N.call("ff_ad_propagate_quad_backward", _p(fi), None, None, _p(gf), None, None, _p(nf), None, None,
                   B, h, w, f32(lv["dt"]), spec.border, 0, s)
        else:
            N.call("ff_ad_propagate_pair_backward", _p(fi), None, None, None, None, None, _p(gf), None, None,
                   _p(nf), None, None, B, h, w, f32(lv["dt"]), spec.border, 0, s)
        gf = nf
    g_old[top]["f"] = gf
    if saved["first"]:
        # first frame: the old brightness (and the delta levels' scalar
        # payload) was this frame's a0; the flow and delta flow were zero
        for l in range(L):
            g_a0[l].add_(g_old[l]["a0"])
        g_old = None
    # image models, then the pyramid from the top down
    g_img = [None] * L
    for l in range(L):
        h, w = spec.sizes[l]
        gi = g_image_out if l == 0 else _empty(B, h, w, like=g_a0[l])
        work = _empty(B, 2 * h, w, like=g_a0[l])
        N.call("ff_ad_image_model_backward", _p(g_a0[l]), _p(g_g[l]), _p(gi), _p(work), B, h, w, 0, s)
        g_img[l] = gi
    for l in range(L - 1, 0, -1):
        h, w = spec.sizes[l - 1]
        work = _empty(B, h, w // 2, like=g_img[l])
        N.call("ff_ad_down_backward", _p(g_img[l]), _p(g_img[l - 1]), _p(work), B, h, w, 1, s)
    return g_old, g_gamma


# ---------------------------------------------------------- the function

class _FlowWindow(torch.autograd.Function):
    @staticmethod
    def forward(ctx, frames, gamma, spec, checkpoint):
        if frames.dim() != 4 or frames.shape[2:] != (spec.H, spec.W):
            raise ValueError("frames must be [batch, frames, %d, %d]" % (spec.H, spec.W))
        if frames.dtype != torch.float32 or not frames.is_cuda:
            raise ValueError("frames must be a float32 CUDA tensor")
        if gamma.numel() != spec.L:
            raise ValueError("one gamma per level expected")
        B, T = frames.shape[:2]
        gammas = _effective_gammas(gamma)
        state, outs, ckpts = None, [], {}
        for t in range(T):
            # states are never written in place: keeping references suffices
            if t % checkpoint == 0:
                ckpts[t] = state
            state, out, _ = _frame_forward(spec, state, frames[:, t].contiguous(), gammas, False)
            outs.append(out)
        ctx.spec, ctx.checkpoint, ctx.ckpts, ctx.gammas = spec, checkpoint, ckpts, gammas
        ctx.gamma_like = (gamma.dtype, gamma.device, gamma.shape)
        ctx.save_for_backward(frames)
        return torch.stack(outs, dim=1)

    @staticmethod
    def backward(ctx, g_flows):
        (frames,) = ctx.saved_tensors
        spec, C = ctx.spec, ctx.checkpoint
        B, T = frames.shape[:2]
        # image gradients frame-major, so each frame's is one contiguous block
        g_tb = torch.empty((T, B) + tuple(frames.shape[2:]), dtype=frames.dtype, device=frames.device)
        g_gamma = torch.zeros(spec.L, dtype=torch.float64, device=frames.device)
        g_state = None
        for t0 in reversed(range(0, T, C)):
            t1 = min(t0 + C, T)
            state, saved = ctx.ckpts[t0], []
            for t in range(t0, t1):
                state, _, sv = _frame_forward(spec, state, frames[:, t].contiguous(), ctx.gammas, True)
                saved.append(sv)
            for t in reversed(range(t0, t1)):
                g_state, g_gam = _frame_backward(spec, saved[t - t0], g_state, g_flows[:, t].contiguous(),
                                                 ctx.gammas, g_tb[t])
                g_gamma += g_gam
                saved[t - t0] = None
            del saved
        ctx.ckpts = None
        dtype, device, shape = ctx.gamma_like
        g_gamma = (g_gamma / _U8_SCALE).to(dtype=dtype, device=device).reshape(shape)
        return g_tb.transpose(0, 1), g_gamma, None, None


def flow_window(frames, gamma, spec, checkpoint=4):
    """Level-0 smoothed flow of every frame, [batch, frames, H, W, 2].

    frames: float32 CUDA tensor [batch, frames, H, W] with values in [0, 1];
    gamma: tensor of one gain per level (level 0 first), in setGamma units
    for u8 frames; spec: FilterSpec; checkpoint: frames between the states
    kept for the backward pass."""
    if checkpoint < 1:
        raise ValueError("checkpoint must be at least 1")
    return _FlowWindow.apply(frames, gamma, spec, int(checkpoint))


# ------------------------------------------------------------------ losses

class _Photometric(torch.autograd.Function):
    """The photometric term on the GPU in two kernels (ff_ad_photometric and
    its adjoint): the sum of squared differences as per-workgroup partial
    sums, added in double in a fixed order; the backward adds the frame and
    flow gradients in one pass."""

    @staticmethod
    def forward(ctx, frames, flows):
        B, T, H, W = frames.shape
        frames = frames.contiguous()
        flows = flows.contiguous()
        n = N.get_int("ff_ad_photometric_partial_count", B, T, H, W)
        part = torch.empty(n, dtype=torch.float32, device=frames.device)
        N.call("ff_ad_photometric", _p(frames), _p(flows), _p(part), B, T, H, W, _stream())
        count = B * (T - 1) * H * W
        ctx.save_for_backward(frames, flows)
        ctx.count = count
        return (part.double().sum() / count).to(torch.float32)

    @staticmethod
    def backward(ctx, g):
        frames, flows = ctx.saved_tensors
        B, T, H, W = frames.shape
        scale = (g.to(torch.float32) / ctx.count).reshape(1).contiguous()
        gframes = torch.zeros_like(frames)
        gflows = torch.zeros_like(flows)
        N.call("ff_ad_photometric_backward", _p(frames), _p(flows), _p(scale), _p(gframes), _p(gflows), B, T, H, W,
               _stream())
        return gframes, gflows


def photometric_error(frames, flows):
    """Mean over frames t >= 1 of the mean squared difference between frame
    t and frame t-1 sampled bilinearly at x - flow_t(x) (edge-clamped).
    frames [B,T,H,W], flows [B,T,H,W,2] (pixels per frame). Float32 tensors
    on the GPU go through the fused kernels (_Photometric); other tensors
    (the CPU and float64 checks of the composition) through one grid_sample
    call over all frame pairs."""
    B, T, H, W = frames.shape
    if T < 2:
        return frames.sum() * 0.0
    if frames.is_cuda and frames.dtype == torch.float32 and flows.dtype == torch.float32:
        return _Photometric.apply(frames, flows)
    dev, dt = frames.device, frames.dtype
    xs = torch.arange(W, device=dev, dtype=dt)
    ys = torch.arange(H, device=dev, dtype=dt)[:, None]
    fl = flows[:, 1:]
    gx = (xs - fl[..., 0]) * (2.0 / max(W - 1, 1)) - 1.0
    gy = (ys - fl[..., 1]) * (2.0 / max(H - 1, 1)) - 1.0
    grid = torch.stack([gx, gy], -1).reshape(B * (T - 1), H, W, 2)
    prev = frames[:, :-1].reshape(B * (T - 1), 1, H, W)
    warped = torch.nn.functional.grid_sample(prev, grid, mode="bilinear", padding_mode="border",
                                             align_corners=True).reshape(B, T - 1, H, W)
    return ((frames[:, 1:] - warped) ** 2).mean()


def flow_loss(frames, flows, truth, photometric_weight=0.1):
    """Mean squared error of the flow of frames t >= 1 against the known
    flow (truth: [B,2] constant per sequence, or [B,T,H,W,2] / [B,H,W,2]
    fields) plus photometric_weight x photometric_error."""
    if truth.dim() == 2:
        truth = truth[:, None, None, None, :]
    elif truth.dim() == 4:
        truth = truth[:, None]
    else:
        truth = truth[:, 1:]
    l2 = ((flows[:, 1:] - truth) ** 2).mean()
    return l2 + photometric_weight * photometric_error(frames, flows)
