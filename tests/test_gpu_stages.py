"""GPU parity of every pipeline stage against the CPU oracle (FP32), through
the C ABI. The stage kernels follow the oracle's operation order and
rounding points, so results are compared bit for bit; the filter-level bar
of 1e-4 px is checked in test_gpu_filters.py.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (3, 2), (5, 7), (17, 33), (64, 64), (67, 131), (240, 320)]


def rng(seed):
    return np.random.default_rng(seed)


def gimg(arr, itemSize=None):
    from flowfilter.gpu.image import GPUImage
    arr = np.ascontiguousarray(arr)
    g = GPUImage(arr.shape, arr.dtype.itemsize if itemSize is None else itemSize)
    g.upload(arr)
    return g


def dl(g, dtype=np.float32):
    return g.download(dtype)


def smooth_flow(shape, seed, scale=3.0):
    from scipy import ndimage
    f = rng(seed).standard_normal(shape + (2,))
    f = ndimage.gaussian_filter(f, (2, 2, 0)) * scale * 4
    return f.astype(np.float32)


# ------------------------------------------------------------------ GPUImage

def test_gpuimage_roundtrip_and_shapes():
    from flowfilter import errors
    from flowfilter.gpu.image import GPUImage
    for dtype, shape in ((np.uint8, (13, 29)), (np.float32, (7, 9)), (np.float32, (5, 6, 2))):
        a = (rng(1).uniform(0, 200, shape)).astype(dtype)
        g = gimg(a)
        assert g.shape == shape and g.itemSize == a.dtype.itemsize
        assert g.pitch % 256 == 0 and g.pitch >= shape[1] * a.dtype.itemsize * (shape[2] if len(shape) == 3 else 1)
        assert np.array_equal(g.download(dtype), a)
    g = GPUImage((4, 5), 4)
    with pytest.raises(errors.InvalidArgument):
        g.upload(np.zeros((4, 6), np.float32))
    with pytest.raises(errors.InvalidArgument):
        g.download(np.uint8)


def test_gpuimage_host_pitch_honoured():
    """upload/download respect the host row pitch of the image_t."""
    from flowfilter import _native as N
    from flowfilter.gpu.image import GPUImage
    wide = rng(2).integers(0, 255, (6, 40), dtype=np.uint8)
    g = GPUImage((6, 25), 1)
    d = N.image_t(6, 25, 1, wide.strides[0], 1, wide.ctypes.data_as(ctypes.c_void_p))
    N.call("ff_gpuimage_upload", g.handle, ctypes.byref(d))
    assert np.array_equal(g.download(np.uint8), wide[:, :25])
    out = np.zeros((6, 40), np.uint8)
    d2 = N.image_t(6, 25, 1, out.strides[0], 1, out.ctypes.data_as(ctypes.c_void_p))
    N.call("ff_gpuimage_download", g.handle, ctypes.byref(d2))
    assert np.array_equal(out[:, :25], wide[:, :25]) and not out[:, 25:].any()


def test_gpuimage_copy_clear_share():
    from flowfilter.gpu.image import GPUImage
    a = rng(3).standard_normal((9, 11)).astype(np.float32)
    g = gimg(a)
    h = GPUImage((9, 11), 4)
    h.copyFrom(g)
    assert np.array_equal(h.download(np.float32), a)
    g.clear()
    assert not g.download(np.float32).any()
    # an empty GPUImage takes the shape of its first upload
    e = GPUImage()
    e.upload(a)
    assert e.shape == (9, 11) and np.array_equal(e.download(np.float32), a)


def test_gpuimage_device_copies():
    """Device-to-device upload/download with the source/destination pitch honoured."""
    from flowfilter import _native as N
    from flowfilter.gpu.image import GPUImage
    a = rng(4).integers(0, 255, (10, 70), dtype=np.uint8)
    src = gimg(a)
    g = GPUImage((10, 70), 1)
    g.uploadDevice(src.data, src.pitch)
    out = GPUImage((10, 70), 1)
    g.downloadDevice(out.data, out.pitch)
    N.synchronize()
    assert np.array_equal(out.download(np.uint8), a)
    # a window of a wider device image (pitch larger than the row)
    wide = gimg(rng(5).integers(0, 255, (10, 300), dtype=np.uint8))
    g.uploadDevice(wide.data + 17, wide.pitch)
    assert np.array_equal(g.download(np.uint8), wide.download(np.uint8)[:, 17:87])


@pytest.mark.parametrize("shape,dtype,offset", [
    ((7, 64), np.uint8, 0),        # 64-byte rows: 16-byte accesses, 28 of them (one partial block)
    ((1, 16), np.uint8, 0),        # a single 16-byte access (the second of the thread's pair unused)
    ((300, 2048), np.uint8, 0),    # many blocks of 16-byte accesses
    ((9, 5), np.float32, 0),       # 20-byte rows: 4-byte accesses
    ((33, 3, 2), np.float32, 0),   # 24-byte rows (float2): 4-byte accesses, odd row count
    ((11, 37), np.uint8, 0),       # 37-byte rows: byte accesses
    ((5, 64), np.uint8, 3),        # 16-byte-wide rows read at a 3-byte offset: byte accesses
])
def test_device_copies_every_access_width(shape, dtype, offset):
    """The region copy kernel behind device uploads: rows copied with 16-,
    4- or 1-byte accesses depending on alignment, a thread's two accesses
    possibly past the region's end; every byte lands where it belongs."""
    from flowfilter.gpu.image import GPUImage
    from flowfilter import _native as N
    r = rng(11 + shape[0])
    if np.issubdtype(dtype, np.integer):
        a = r.integers(0, 255, shape, dtype=dtype)
    else:
        a = r.standard_normal(shape).astype(dtype)
    item = np.dtype(dtype).itemsize
    # the source: a wider device image whose rows hold `a` at `offset` bytes
    row = a.shape[1] * (a.shape[2] if a.ndim == 3 else 1) * item
    wide = np.zeros((a.shape[0], row + offset + 32), np.uint8)
    wide[:, offset:offset + row] = a.view(np.uint8).reshape(a.shape[0], row)
    src = gimg(wide)
    g = GPUImage(a.shape, item)
    g.uploadDevice(src.data + offset, src.pitch)
    N.synchronize()
    got = g.download(dtype)
    assert got.shape == a.shape and np.array_equal(got.view(np.uint8), a.view(np.uint8))


def test_device_copy_of_2_gib_with_byte_accesses():
    """A 16384 x 131072 u8 image (2^31 bytes) copied from an odd device
    address, so with one-byte accesses: 2^31 accesses, one more than a
    32-bit count holds. The copy is split into bands of rows; every byte
    lands where it belongs (compared on the device)."""
    import torch
    from flowfilter.gpu.image import GPUImage
    from flowfilter import _native as N
    H, W = 16384, 131072
    pitch = W + 256
    print("2 GiB device copy with byte accesses", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(3)
    buf = torch.randint(0, 256, (H * pitch + 1,), dtype=torch.uint8, device="cuda", generator=gen)
    torch.cuda.synchronize()
    g = GPUImage((H, W), 1)
    assert g.pitch == W
    g.uploadDevice(buf.data_ptr() + 1, pitch)
    N.synchronize()
    got = torch.from_dlpack(g)
    want = torch.as_strided(buf, (H, W), (pitch, 1), 1)
    assert got.shape == (H, W) and torch.equal(got, want)


# ---------------------------------------------------------------- ImageModel

@pytest.mark.parametrize("shape", SIZES)
@pytest.mark.parametrize("u8", [True, False])
def test_image_model(ffo, shape, u8):
    from flowfilter.gpu.imagemodel import ImageModel
    img = rng(5).integers(0, 256, shape, dtype=np.uint8) if u8 else \
        rng(5).uniform(-2, 2, shape).astype(np.float32)
    m = ImageModel(gimg(img))
    m.compute()
    a0, g = ffo.image_model(img, "f32")
    assert np.array_equal(dl(m.getImageConstant()), a0)
    assert np.array_equal(dl(m.getImageGradient()), g)
    assert m.elapsedTime() >= 0.0


# ------------------------------------------------------------ FlowPropagator

@pytest.mark.parametrize("shape", SIZES)
@pytest.mark.parametrize("iters,border,invert", [(1, 3, False), (4, 3, False), (2, 0, True), (3, 1, False)])
def test_flow_propagator(ffo, shape, iters, border, invert):
    from flowfilter.gpu.propagation import FlowPropagator
    f = smooth_flow(shape, 6)
    p = FlowPropagator(gimg(f), iters)
    p.border = border
    p.invertInputFlow = invert
    assert p.iterations == iters and abs(p.dt - np.float32(1.0) / np.float32(iters)) == 0
    p.compute()
    want = ffo.propagate(f, iters, border, invert, "f32")
    assert np.array_equal(dl(p.getPropagatedFlow()), want)


@pytest.mark.parametrize("shape", SIZES)
@pytest.mark.parametrize("iters,border", [(1, 3), (4, 3), (2, 0)])
def test_flow_propagator_payload(ffo, shape, iters, border):
    from flowfilter.gpu.propagation import FlowPropagatorPayload
    f = smooth_flow(shape, 7)
    s = rng(8).uniform(0, 1, shape).astype(np.float32)
    v = smooth_flow(shape, 9, 0.5)
    p = FlowPropagatorPayload(gimg(f), gimg(s), gimg(v), iters)
    p.border = border
    p.compute()
    fo, so, vo = ffo.propagate_payload(f, s, v, iters, border, "f32")
    assert np.array_equal(dl(p.getPropagatedFlow()), fo)
    assert np.array_equal(dl(p.getPropagatedScalar()), so)
    assert np.array_equal(dl(p.getPropagatedVector()), vo)


def test_propagator_errors():
    from flowfilter import errors
    from flowfilter.gpu.propagation import FlowPropagator, FlowPropagatorPayload
    f = gimg(np.zeros((8, 8, 2), np.float32))
    p = FlowPropagator(f, 2)
    with pytest.raises(errors.FilterError):
        p.iterations = 0
    with pytest.raises(errors.FilterError):
        p.border = -1
    with pytest.raises(errors.FilterError):
        p.setInputFlow(gimg(np.zeros((8, 8), np.float32)))
    with pytest.raises(RuntimeError):
        p.dt = 3
    q = FlowPropagatorPayload()
    with pytest.raises(errors.FilterError):
        q.configure()
    with pytest.raises(errors.LogicError):
        q.compute()


# ---------------------------------------------------------------- FlowUpdate

@pytest.mark.parametrize("shape", SIZES)
def test_flow_update(ffo, shape):
    from flowfilter.gpu.update import FlowUpdate
    r = rng(10)
    a0 = r.uniform(0, 1, shape).astype(np.float32)
    old = r.uniform(0, 1, shape).astype(np.float32)
    g = r.normal(0, 0.05, shape + (2,)).astype(np.float32)
    f = r.uniform(-3, 3, shape + (2,)).astype(np.float32)
    if shape[0] > 2:
        g[1, 0, 0] = np.nan      # NaN -> +maxflow
        a0[2, 0] = np.inf
    gamma, maxflow = 10 / 65025.0, 2.5
    u = FlowUpdate(gimg(f), gimg(a0), gimg(g), gamma, maxflow)
    u.getUpdatedImage().upload(old)   # the old image is the stage's own updated image
    u.compute()
    fo, io = ffo.update(a0, g, old, f, np.float32(gamma), maxflow, "f32")
    assert np.array_equal(dl(u.getUpdatedFlow()), fo)
    assert np.array_equal(dl(u.getUpdatedImage()), io)
    assert u.gamma == np.float32(gamma) and u.maxflow == np.float32(maxflow)


def test_flow_update_errors():
    from flowfilter import errors
    from flowfilter.gpu.update import FlowUpdate
    z2, z1 = gimg(np.zeros((4, 4, 2), np.float32)), gimg(np.zeros((4, 4), np.float32))
    u = FlowUpdate(z2, z1, z2, 1.0, 1.0)
    for bad in (0.0, -1.0):
        with pytest.raises(errors.FilterError):
            u.gamma = bad
    with pytest.raises(errors.FilterError):
        u.setInputImage(z2)
    with pytest.raises(errors.FilterError):
        FlowUpdate(z2, gimg(np.zeros((4, 5), np.float32)), z2, 1.0, 1.0)


@pytest.mark.parametrize("shape,ushape", [((64, 64), (32, 32)), ((17, 33), (8, 16)), ((67, 131), (33, 65)),
                                          ((240, 320), (120, 160)), ((5, 7), (3, 2))])
def test_delta_flow_update(ffo, shape, ushape):
    from flowfilter.gpu.update import DeltaFlowUpdate
    r = rng(11)
    a0 = r.uniform(0, 1, shape).astype(np.float32)
    old = r.uniform(0, 1, shape).astype(np.float32)
    g = r.normal(0, 0.05, shape + (2,)).astype(np.float32)
    d = r.uniform(-1, 1, shape + (2,)).astype(np.float32)
    up = r.uniform(-2, 2, ushape + (2,)).astype(np.float32)
    if shape[0] > 2:
        g[1, 1, 1] = np.nan
    gamma, maxflow = 50 / 65025.0, 4.0
    u = DeltaFlowUpdate(gimg(up), gimg(d), gimg(old), gimg(a0), gimg(g), gamma, maxflow)
    u.compute()
    do, fo, io = ffo.delta_update(a0, g, old, d, up, np.float32(gamma), maxflow, "f32")
    assert np.array_equal(dl(u.getUpdatedDeltaFlow()), do)
    assert np.array_equal(dl(u.getUpdatedFlow()), fo)
    assert np.array_equal(dl(u.getUpdatedImage()), io)


# -------------------------------------------------------------- FlowSmoother

@pytest.mark.parametrize("shape", SIZES)
@pytest.mark.parametrize("iters", [1, 2, 3])
def test_flow_smoother(ffo, shape, iters):
    from flowfilter.gpu.flowsmoothing import FlowSmoother
    f = rng(12).standard_normal(shape + (2,)).astype(np.float32)
    s = FlowSmoother(gimg(f), iters)
    assert s.iterations == iters
    s.compute()
    assert np.array_equal(dl(s.getSmoothedFlow()), ffo.smooth(f, iters, "f32"))


# -------------------------------------------------------------- ImagePyramid

@pytest.mark.parametrize("shape,levels", [((1088, 2048), 2), ((481, 641), 3), ((37, 53), 4), ((8, 8), 3),
                                          ((480, 640), 1)])
@pytest.mark.parametrize("u8", [True, False])
def test_image_pyramid(ffo, shape, levels, u8):
    from flowfilter.gpu.pyramid import ImagePyramid
    img = rng(13).integers(0, 256, shape, dtype=np.uint8) if u8 else \
        rng(13).uniform(0, 1, shape).astype(np.float32)
    p = ImagePyramid(gimg(img), levels)
    assert p.levels == levels
    p.compute()
    want = ffo.pyramid(img, levels, "f32")
    for lvl in range(levels):
        got = p.getImage(lvl).download(np.uint8 if u8 else np.float32)
        assert np.array_equal(got, want[lvl]), lvl
    from flowfilter import errors
    with pytest.raises(errors.FilterError):
        p.getImage(levels)
