"""Streams of their own per filter and batch, the ordering of work across
them, host frames loaded double buffered, and the DLPack export of device
images, through the C ABI.

Bar: every flow bitwise equal to the CPU oracle's (or to the same filter run
another way, itself pinned to the oracle elsewhere).
"""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pyr(h, w, levels=2):
    import flowfilter.gpu.flowfilters as gf
    f = gf.PyramidalFlowFilter(h, w, levels)
    f.gamma = [10.0, 50.0, 60.0][:levels] if levels > 1 else 10.0
    f.maxflow = 4.0
    f.smoothIterations = [2] * levels if levels > 1 else 2
    return f


def _blobs(h, w, n, seed=2):
    from flowfilter import synthetic
    return list(synthetic.blob_affine(h, w, n, blobs=40, seed=seed))


def test_each_filter_and_batch_has_its_own_stream():
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import _native as N
    say("streams of filters and batches")
    a, b, c = _pyr(64, 96), _pyr(64, 96), gf.FlowFilter(48, 64)
    lib = N.get_stream()
    streams = {a.stream, b.stream, c.stream}
    assert len(streams) == 3 and lib not in streams
    batch = gf.BatchedFlowFilter([a])
    assert a.stream == batch.stream and batch.stream not in (b.stream, lib)
    # a caller's stream, and back
    s = torch.cuda.Stream()
    b.stream = s.cuda_stream
    assert b.stream == s.cuda_stream
    b.stream = None
    assert b.stream not in (s.cuda_stream, lib)


def test_filter_on_a_caller_stream_matches_oracle(ffo):
    """The filter enqueues on torch's stream; markers on it time compute()."""
    from flowfilter import _native as N
    h, w = 272, 512
    say("filter on a caller stream", (h, w))
    s = torch.cuda.Stream()
    f = _pyr(h, w)
    f.stream = s.cuda_stream
    o = ffo.PyramidalFlowFilter(h, w, 2)
    o.configure_like([10.0, 50.0], 4.0, [2, 2])
    for t, img in enumerate(_blobs(h, w, 6)):
        m0 = N.Marker(f)
        f.loadImage(img)
        f.compute()
        m1 = N.Marker(f)
        o.compute(img)
        assert np.array_equal(f.getFlow(), o.downloadFlow()), t
        assert m0.elapsed_ms(m1) > 0.0
    assert f.stream == s.cuda_stream


def test_host_loads_double_buffered_match_oracle(ffo):
    """Host frames go through the filter's spare input buffer on a copy
    stream while the previous compute() may still run: enqueue many frames
    back to back (no synchronizing call in between) and read the flow only
    at the end, then compare a recomputation frame by frame."""
    h, w = 544, 1024
    say("double-buffered host loads", (h, w))
    frames = _blobs(h, w, 12, seed=9)
    f = _pyr(h, w)
    for img in frames:
        f.loadImage(img)
        f.compute()
    last = f.getFlow()
    o = ffo.PyramidalFlowFilter(h, w, 2)
    o.configure_like([10.0, 50.0], 4.0, [2, 2])
    for img in frames:
        o.compute(img)
    assert np.array_equal(last, o.downloadFlow())
    # the single-level filter that owns its input does the same
    import flowfilter.gpu.flowfilters as gf
    g = gf.FlowFilter(h, w, 2, 4.0, 10.0)
    q = ffo.FlowFilter(h, w, 2, 4.0, 10.0)
    for img in frames[:6]:
        g.loadImage(img)
        g.compute()
        q.compute(img)
    assert np.array_equal(g.getFlow(), q.downloadFlow())


def test_input_image_on_another_stream(ffo):
    """FlowFilter(GPUImage): the input image is the caller's (library
    stream). Frames copied into it from device memory are enqueued on the
    library stream; the filter's compute() on its own stream must see each
    frame, and the next copy must not overwrite a frame compute() still
    reads."""
    import flowfilter.gpu.flowfilters as gf
    from flowfilter.gpu.image import GPUImage
    h, w = 240, 320
    say("input image on the library stream", (h, w))
    frames = [(img / 255.0).astype(np.float32) for img in _blobs(h, w, 8, seed=4)]
    dev = GPUImage((len(frames) * h, w), 4)
    dev.upload(np.concatenate(frames, axis=0))
    inp = GPUImage((h, w), 4)
    g = gf.FlowFilter.fromGPUImage(inp)
    g.gamma = 0.01
    g.maxflow = 2.0
    g.smoothIterations = 2
    o = ffo.FlowFilter(h, w, 2, 2.0, 0.01, u8=False)
    for t, f in enumerate(frames):
        inp.uploadDevice(dev.data + t * h * dev.pitch, dev.pitch)
        g.compute()
        o.compute(f)
        assert np.array_equal(g.getFlow(), o.downloadFlow()), t


def test_stage_reading_a_filters_image(ffo):
    """A standalone smoother (library stream) reading a filter's flow
    (filter's stream) smooths that frame's flow."""
    from flowfilter.gpu.flowsmoothing import FlowSmoother
    h, w = 272, 512
    say("stage reading a filter's image", (h, w))
    f = _pyr(h, w)
    o = ffo.PyramidalFlowFilter(h, w, 2)
    o.configure_like([10.0, 50.0], 4.0, [2, 2])
    for img in _blobs(h, w, 4, seed=3):
        f.loadImage(img)
        f.compute()
        o.compute(img)
        sm = FlowSmoother(f.getFlowDevice(), 1)
        sm.compute()
        got = sm.getSmoothedFlow().download(np.float32)
        assert np.array_equal(got, ffo.smooth(o.downloadFlow(), 1))


def test_dlpack_export_of_flow(ffo):
    """torch.from_dlpack(getFlowDevice()) shares the flow buffer (no copy)
    and equals getFlow() bitwise, also when read on another torch stream."""
    from flowfilter.gpu.image import GPUImage
    h, w = 272, 512
    say("DLPack export", (h, w))
    f = _pyr(h, w)
    o = ffo.PyramidalFlowFilter(h, w, 2)
    o.configure_like([10.0, 50.0], 4.0, [2, 2])
    s = torch.cuda.Stream()
    for t, img in enumerate(_blobs(h, w, 5, seed=8)):
        f.loadImage(img)
        f.compute()
        o.compute(img)
        dimg = f.getFlowDevice()
        with torch.cuda.stream(s):
            tt = torch.from_dlpack(dimg)
            cp = tt.clone()
        assert tt.device.type == "cuda" and tt.dtype == torch.float32 and tuple(tt.shape) == (h, w, 2)
        assert tt.data_ptr() == dimg.data
        assert tt.stride() == (dimg.pitch // 4, 2, 1)
        torch.cuda.synchronize()
        want = o.downloadFlow()
        assert np.array_equal(cp.cpu().numpy(), want), t
        assert np.array_equal(f.getFlow(), want), t
    # u8 images export as uint8; the view keeps the buffer alive
    u = GPUImage((16, 40), 1)
    u.upload(np.arange(16 * 40, dtype=np.uint8).reshape(16, 40))
    tu = torch.from_dlpack(u)
    del u
    assert tu.dtype == torch.uint8 and np.array_equal(tu.cpu().numpy(), np.arange(16 * 40, dtype=np.uint8).reshape(16, 40))
    # a capsule nobody consumes is released by its destructor
    cap = GPUImage((8, 8), 4).__dlpack__()
    del cap


def test_pyramids_on_many_streams_interleaved():
    """Sixteen 2- and 3-level pyramids of 1024x544 on sixteen streams, their
    frames read in place from HBM and launched interleaved for 40 frames:
    fused launches whose tiles wait on the same launch's top level are
    ordered one after another on the device (two side by side can each hold
    an XCD with waiting tiles while the other's producers queue behind them
    there), so no wait gives up, and every flow is bitwise the flow of the
    same filter run alone on one stream."""
    from flowfilter import _native as N
    from flowfilter.gpu.image import GPUImage
    h, w, n, F = 544, 1024, 16, 4
    frames = np.stack(_blobs(h, w, F, seed=11))
    dev = GPUImage((F * h, w), 1)
    dev.upload(frames.reshape(F * h, w))
    ptr = lambda t: dev.data + (t % F) * h * dev.pitch
    say("16 pyramids on 16 streams, 40 interleaved frames")
    fs = [_pyr(h, w, 2 + k % 2) for k in range(n)]
    assert len({f.stream for f in fs}) == n
    for t in range(40):
        for f in fs:
            f.setImageDevice(ptr(t), dev.pitch)
            f.compute()
    got = [f.getFlow() for f in fs]
    assert all(f.waitTimeouts() == 0 for f in fs)
    say("the same filters alone")
    for levels in (2, 3):
        ref = _pyr(h, w, levels)
        for t in range(40):
            ref.setImageDevice(ptr(t), dev.pitch)
            ref.compute()
        want = ref.getFlow()
        for k, f in enumerate(fs):
            if 2 + k % 2 == levels:
                assert np.array_equal(got[k], want), (k, levels)
    N.synchronize()


def _hip_runtime():
    """The HIP runtime this process has loaded (torch's copy, which the
    library binds to as well)."""
    import ctypes
    path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line)
    hip = ctypes.CDLL(path)
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def test_pyramids_on_a_destroyed_caller_stream_and_a_library_stream():
    """A 2-level pyramid on a caller's stream (made and destroyed with the
    HIP runtime) interleaved with a 3-level one on its own stream, then the
    caller's stream destroyed right after the first pyramid's last launch
    and the second pyramid run on: launches stay ordered across the streams
    without the destroyed handle being used again, no wait gives up, and
    both flows are bitwise those of the same filters run alone."""
    import ctypes
    from flowfilter.gpu.image import GPUImage
    hip = _hip_runtime()
    h, w, F = 544, 1024, 4
    frames = np.stack(_blobs(h, w, F, seed=13))
    dev = GPUImage((F * h, w), 1)
    dev.upload(frames.reshape(F * h, w))
    ptr = lambda t: dev.data + (t % F) * h * dev.pitch

    def step(f, t):
        f.setImageDevice(ptr(t), dev.pitch)
        f.compute()

    raw = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(raw)) == 0
    say("pyramid on a caller stream, pyramid on its own stream")
    a, b = _pyr(h, w, 2), _pyr(h, w, 3)
    a.stream = raw.value
    for t in range(12):
        step(a, t)
        step(b, t)
    step(a, 12)
    fa = a.getFlow()
    a.stream = None
    assert hip.hipStreamDestroy(raw) == 0
    say("caller stream destroyed; the other pyramid goes on")
    for t in range(12, 20):
        step(b, t)
    fb = b.getFlow()
    assert a.waitTimeouts() == 0 and b.waitTimeouts() == 0
    say("the same filters alone")
    ra, rb = _pyr(h, w, 2), _pyr(h, w, 3)
    for t in range(13):
        step(ra, t)
    for t in range(20):
        step(rb, t)
    assert np.array_equal(fa, ra.getFlow())
    assert np.array_equal(fb, rb.getFlow())
