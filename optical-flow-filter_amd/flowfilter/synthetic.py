"""Seeded synthetic image sequences with known motion.

The reference ships no test data (its notebook reads private Blender
renders), so parity tests and benchmarks run on these generated sequences:

* ``translating_noise``: band-limited noise (white noise low-passed with a
  Gaussian, rescaled to [0, 255]) translated by a constant velocity per frame
  (bilinear resampling of the float frame 0). Ground truth is the velocity.
  Default: 640x480, (0.7, -0.4) px/frame, sigma 3, seed 1.
* ``blob_affine``: a sum of random Gaussian blobs on mid-grey, advanced per
  frame by an affine motion about the image centre (rotation plus
  expansion, displacement capped), with fresh Gaussian sensor noise per
  frame. Default: 2048x1088, 400 blobs (sigma 2-20 px, amplitude 20-80),
  0.5 deg rotation + 1% expansion per frame, cap 4 px, noise sigma 2, seed 2.

* ``camera_streams``: many independent streams for the batched filter, each
  with its own resolution (320x240, 640x480, 1280x720 or 1920x1080), pyramid
  depth (1-3), gamma per level (uniform 5-50), smoothing iterations (1-2 per
  level) and band-limited noise texture moved by a random affine motion
  (translation up to 2 px plus rotation up to 0.5 deg and scale within 1%
  about the centre, per frame, displacement capped at 4 px). Seed 3.

* ``smooth_flow_sequence``: band-limited noise texture moved by a smooth
  random flow field (seeded Gaussian noise low-passed with sigma 64 px,
  scaled to a maximum of 8 px per frame): frame t(x) = texture(x - t v(x)),
  quantized to u8 levels. Built with torch on a device (FFT low-pass and
  bilinear warp), so 4K windows take well under a second on a GPU.

All randomness comes from numpy.random.default_rng(seed) (PCG64), so a seed
gives the same frames on every machine with this numpy (for
``smooth_flow_sequence``, up to the device FFT's rounding before the u8
quantization).
"""
import numpy as np
from scipy import ndimage


def _quantize(f):
    return np.clip(np.rint(f), 0, 255).astype(np.uint8)


def band_limited_noise(h, w, sigma=3.0, seed=1, pad=0):
    """Float image of low-passed white noise rescaled to [0, 255]."""
    rng = np.random.default_rng(seed)
    H, W = h + 2 * pad, w + 2 * pad
    n = rng.standard_normal((H, W))
    n = ndimage.gaussian_filter(n, sigma, mode="reflect", truncate=4.0)
    n -= n.min()
    n *= 255.0 / max(n.max(), 1e-12)
    return n


def translating_noise(h=480, w=640, frames=200, velocity=(0.7, -0.4), sigma=3.0, seed=1):
    """Yield uint8 frames of noise translated by ``velocity`` (vx, vy) per frame."""
    vx, vy = velocity
    pad = int(np.ceil(max(abs(vx), abs(vy)) * frames)) + 16
    base = band_limited_noise(h, w, sigma, seed, pad)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    for t in range(frames):
        # frame t(x) = base(x - v t): content moves by +v per frame
        coords = np.stack([yy + pad - vy * t, xx + pad - vx * t])
        yield _quantize(ndimage.map_coordinates(base, coords, order=1, mode="nearest"))


def blob_texture(h, w, blobs=400, seed=2, sigma_range=(2.0, 20.0), amp_range=(20.0, 80.0), background=128.0):
    rng = np.random.default_rng(seed)
    img = np.full((h, w), background, np.float64)
    cy = rng.uniform(0, h, blobs)
    cx = rng.uniform(0, w, blobs)
    sg = rng.uniform(sigma_range[0], sigma_range[1], blobs)
    am = rng.uniform(amp_range[0], amp_range[1], blobs)
    for y0, x0, s, a in zip(cy, cx, sg, am):
        r = int(np.ceil(4 * s))
        ya, yb = max(0, int(y0) - r), min(h, int(y0) + r + 1)
        xa, xb = max(0, int(x0) - r), min(w, int(x0) + r + 1)
        if ya >= yb or xa >= xb:
            continue
        yy, xx = np.mgrid[ya:yb, xa:xb]
        img[ya:yb, xa:xb] += a * np.exp(-((yy - y0) ** 2 + (xx - x0) ** 2) / (2 * s * s))
    return np.clip(img, 0, 255)


def affine_displacement(h, w, rot_deg=0.5, scale=1.01, cap=4.0):
    """Per-pixel displacement (dx, dy) of the affine motion about the centre, capped in norm."""
    th = np.deg2rad(rot_deg)
    A = scale * np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]) - np.eye(2)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    px, py = xx - (w - 1) / 2.0, yy - (h - 1) / 2.0
    dx = A[0, 0] * px + A[0, 1] * py
    dy = A[1, 0] * px + A[1, 1] * py
    n = np.sqrt(dx * dx + dy * dy)
    f = np.where(n > cap, cap / np.maximum(n, 1e-12), 1.0)
    return dx * f, dy * f


def blob_affine(h=1088, w=2048, frames=300, blobs=400, seed=2, rot_deg=0.5, scale=1.01, cap=4.0,
                noise=2.0):
    """Yield uint8 frames of a blob texture advected by a capped affine motion."""
    f = blob_texture(h, w, blobs, seed)
    dx, dy = affine_displacement(h, w, rot_deg, scale, cap)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    coords = np.stack([yy - dy, xx - dx])
    for t in range(frames):
        rng = np.random.default_rng(seed + t)
        yield _quantize(f + rng.normal(0.0, noise, (h, w)))
        f = ndimage.map_coordinates(f, coords, order=1, mode="nearest")


def random_u8(h, w, seed=0):
    return np.random.default_rng(seed).integers(0, 256, (h, w), dtype=np.uint8)


STREAM_SIZES = ((240, 320), (480, 640), (720, 1280), (1080, 1920))


def camera_streams(n=64, frames=3, seed=3, maxflow=4.0):
    """Specs and frames of ``n`` independent camera streams.

    Returns a list of dicts with keys h, w, levels, gamma (list, one per
    level), maxflow, smooth (list) and frames (list of uint8 arrays; frame
    t+1 is frame t's texture moved by the stream's affine motion)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        h, w = STREAM_SIZES[int(rng.integers(len(STREAM_SIZES)))]
        levels = int(rng.integers(1, 4))
        gamma = [float(g) for g in rng.uniform(5.0, 50.0, size=levels)]
        smooth = [int(k) for k in rng.integers(1, 3, size=levels)]
        tx, ty = rng.uniform(-2.0, 2.0, size=2)
        rot = np.deg2rad(rng.uniform(-0.5, 0.5))
        scale = 1.0 + rng.uniform(-0.01, 0.01)
        pad = int(np.ceil(4.0 * frames)) + 8
        base = band_limited_noise(h, w, 3.0, int(rng.integers(1 << 30)), pad)
        yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
        cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
        c, s_ = np.cos(rot) * scale, np.sin(rot) * scale
        dx = (c - 1.0) * (xx - cx) - s_ * (yy - cy) + tx
        dy = s_ * (xx - cx) + (c - 1.0) * (yy - cy) + ty
        mag = np.maximum(np.hypot(dx, dy) / maxflow, 1.0)
        dx, dy = dx / mag, dy / mag
        seq = []
        for t in range(frames):
            coords = np.stack([yy + pad - dy * t, xx + pad - dx * t])
            seq.append(_quantize(ndimage.map_coordinates(base, coords, order=1, mode="nearest")))
        out.append(dict(h=h, w=w, levels=levels, gamma=gamma, maxflow=maxflow, smooth=smooth, frames=seq))
    return out


def smooth_flow_sequence(h, w, frames, maxflow=8.0, flow_sigma=64.0, texture_sigma=3.0, seed=7, device="cuda"):
    """Frames [frames, h, w] (float32 torch tensor, u8 levels / 255) of a
    band-limited noise texture advected by a smooth random flow field, and
    that field v [h, w, 2] (px/frame): frame t(x) = texture(x - t v(x)).
    The field is Gaussian white noise low-passed (sigma flow_sigma px, via
    FFT) and scaled so that max |v| = maxflow."""
    import torch

    rng = np.random.default_rng(seed)
    pad = int(np.ceil(maxflow * frames)) + 16

    def lowpass(x, sigma):
        H, W = x.shape[-2:]
        ky = torch.fft.fftfreq(H, device=x.device, dtype=torch.float64)[:, None]
        kx = torch.fft.rfftfreq(W, device=x.device, dtype=torch.float64)[None, :]
        g = torch.exp(-2.0 * (np.pi * sigma) ** 2 * (kx * kx + ky * ky))
        return torch.fft.irfft2(torch.fft.rfft2(x) * g, s=(H, W))

    v = lowpass(torch.from_numpy(rng.standard_normal((2, h, w))).to(device), flow_sigma)
    v = v * (maxflow / float(torch.sqrt(v[0] ** 2 + v[1] ** 2).max()))
    tex = lowpass(torch.from_numpy(rng.standard_normal((h + 2 * pad, w + 2 * pad))).to(device), texture_sigma)
    tex = (tex - tex.min()) * (255.0 / float(tex.max() - tex.min()))
    ys = torch.arange(h, device=device, dtype=torch.float64)[:, None] + pad
    xs = torch.arange(w, device=device, dtype=torch.float64)[None, :] + pad
    Hp, Wp = tex.shape
    out = torch.empty(frames, h, w, dtype=torch.float32, device=device)
    for t in range(frames):
        gx = 2.0 * (xs - t * v[0]) / (Wp - 1) - 1.0
        gy = 2.0 * (ys - t * v[1]) / (Hp - 1) - 1.0
        grid = torch.stack([gx.expand(h, w), gy.expand(h, w)], -1)[None]
        f = torch.nn.functional.grid_sample(tex[None, None], grid, mode="bilinear", padding_mode="border",
                                            align_corners=True)[0, 0]
        out[t] = (torch.clamp(torch.round(f), 0, 255) / 255.0).float()
    return out, v.permute(1, 2, 0).float().contiguous()
