"""GPU parity of the filters (FlowFilter, DeltaFlowFilter, PyramidalFlowFilter)
against the CPU oracle over whole synthetic sequences, through the C ABI.

Bar: max |flow_gpu - flow_oracle_fp32| <= 1e-4 px on
every frame. The kernels follow the oracle's rounding points, so the
observed difference is expected to be exactly 0; both are asserted. The
GPU result is also compared with the FP64 oracle and the error printed.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-4  # px, max abs flow error per frame


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def seq_noise(h, w, n, v=(0.7, -0.4)):
    from flowfilter import synthetic
    return list(synthetic.translating_noise(h, w, n, v))


def seq_blobs(h, w, n):
    from flowfilter import synthetic
    return list(synthetic.blob_affine(h, w, n, blobs=max(20, h * w // 5600)))


def check_frame(t, got, want, label):
    d = float(np.abs(got - want).max()) if got.size else 0.0
    assert d <= TOL, "%s frame %d: max |diff| %.3g > %.g" % (label, t, d, TOL)
    assert np.array_equal(got, want), "%s frame %d: not bitwise equal (max |diff| %.3g)" % (label, t, d)
    return d


# ---------------------------------------------------------------- FlowFilter

@pytest.mark.parametrize("shape,frames", [((48, 64), 12), ((480, 640), 30)])
@pytest.mark.parametrize("fused", [True, False])
def test_flowfilter_sequence(ffo, shape, frames, fused):
    import flowfilter.gpu.flowfilters as gf
    h, w = shape
    say("flowfilter", shape, frames, "fused" if fused else "staged")
    g = gf.FlowFilter(h, w, 2, 4.0, 10.0)
    g.fused = fused
    assert (g.height, g.width, g.smoothIterations, g.propagationIterations) == (h, w, 2, 4)
    assert g.maxflow == 4.0 and g.gamma == np.float32(np.float32(10) / (np.float32(255) * np.float32(255)))
    o = ffo.FlowFilter(h, w, 2, 4.0, 10.0)
    o64 = ffo.FlowFilter(h, w, 2, 4.0, 10.0, prec="f64")
    worst64 = 0.0
    for t, img in enumerate(seq_noise(h, w, frames)):
        g.loadImage(img)
        g.compute()
        o.compute(img)
        o64.compute(img)
        got = g.getFlow()
        check_frame(t, got, o.downloadFlow(), "FlowFilter flow")
        check_frame(t, g.getImage(), o.downloadImage(), "FlowFilter image")
        worst64 = max(worst64, float(np.abs(got - o64.downloadFlow()).max()))
    # getFlowDevice is the unsmoothed updated flow (FlowFilter::getFlow)
    check_frame(frames, g.getFlowDevice().download(np.float32), o.getFlow(), "FlowFilter updated flow")
    assert g.elapsedTime() > 0.0
    say("  max |gpu - fp64 oracle| = %.3g px" % worst64)
    assert worst64 < 1e-3


def test_flowfilter_split_methods_match_compute(ffo):
    """computeImageModel + computePropagation + computeUpdate == compute()."""
    import flowfilter.gpu.flowfilters as gf
    h, w = 60, 90
    a, b = gf.FlowFilter(h, w, 2, 3.0, 5.0), gf.FlowFilter(h, w, 2, 3.0, 5.0)
    for img in seq_noise(h, w, 8, (0.5, 0.3)):
        a.loadImage(img)
        b.loadImage(img)
        a.compute()
        b.computeImageModel()
        b.computePropagation()
        b.computeUpdate()
        assert np.array_equal(a.getFlow(), b.getFlow())


def test_flowfilter_float_input(ffo):
    import flowfilter.gpu.flowfilters as gf
    from flowfilter.gpu.image import GPUImage
    h, w = 40, 56
    inp = GPUImage((h, w), 4)
    g = gf.FlowFilter.fromGPUImage(inp)
    # FlowFilter(GPUImage): raw gamma 1.0, maxflow 1.0, one smoothing iteration
    assert g.gamma == 1.0 and g.maxflow == 1.0 and g.smoothIterations == 1
    g.gamma = 0.01
    g.maxflow = 2.0
    o = ffo.FlowFilter(h, w, 1, 2.0, 0.01, u8=False)
    for t, img in enumerate(seq_noise(h, w, 8)):
        f = (img / 255.0).astype(np.float32)
        inp.upload(f)
        g.compute()
        o.compute(f)
        check_frame(t, g.getFlow(), o.downloadFlow(), "float input")


@pytest.mark.parametrize("fused", [True, False])
def test_flowfilter_float_input_with_nan_and_inf(ffo, fused):
    """Float frames holding NaN and +-Inf pixels: the update's clamp and
    sanitize steps (update_k.cu:66-71) turn what reaches the flow into
    +-maxflow or 0 exactly as the oracle does; the flow stays bitwise the
    oracle's, the brightness state too (NaN where the oracle has NaN)."""
    import flowfilter.gpu.flowfilters as gf
    from flowfilter.gpu.image import GPUImage
    h, w = 96, 160
    inp = GPUImage((h, w), 4)
    g = gf.FlowFilter.fromGPUImage(inp)
    g.fused = fused
    g.gamma = 0.01
    g.maxflow = 2.0
    g.smoothIterations = 2
    o = ffo.FlowFilter(h, w, 2, 2.0, 0.01, u8=False)
    rng = np.random.default_rng(17)
    for t, img in enumerate(seq_noise(h, w, 8)):
        f = (img / 255.0).astype(np.float32)
        if t in (2, 5):
            ys, xs = rng.integers(0, h, 6), rng.integers(0, w, 6)
            f[ys[:2], xs[:2]] = np.nan
            f[ys[2:4], xs[2:4]] = np.inf
            f[ys[4:], xs[4:]] = -np.inf
        inp.upload(f)
        g.compute()
        o.compute(f)
        check_frame(t, g.getFlow(), o.downloadFlow(), "float input with NaN/Inf")
        assert np.array_equal(g.getImage(), o.downloadImage(), equal_nan=True), t


def device_frames(frames):
    """All frames in one tall device image; returns (image, base pointer, pitch)."""
    from flowfilter.gpu.image import GPUImage
    f = np.stack(frames)
    dev = GPUImage((f.shape[0] * f.shape[1], f.shape[2]), 1)
    dev.upload(f.reshape(-1, f.shape[2]))
    return dev, dev.data, dev.pitch


def test_flowfilter_device_input(ffo):
    import flowfilter.gpu.flowfilters as gf
    h, w = 200, 100
    g = gf.FlowFilter(h, w, 2, 4.0, 10.0)
    o = ffo.FlowFilter(h, w, 2, 4.0, 10.0)
    frames = seq_noise(h, w, 6)
    dev, base, pitch = device_frames(frames)
    for t, img in enumerate(frames):
        g.loadImageDevice(base + t * h * pitch, pitch)
        g.compute()
        o.compute(img)
        check_frame(t, g.getFlow(), o.downloadFlow(), "device input")


def test_flowfilter_errors():
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import errors
    g = gf.FlowFilter(16, 16)
    with pytest.raises(errors.InvalidArgument):
        g.loadImage(np.zeros((16, 17), np.uint8))
    with pytest.raises(errors.FilterError):
        g.gamma = 0.0
    with pytest.raises(errors.FilterError):
        g.maxflow = 0.0
    with pytest.raises(errors.FilterError):
        g.smoothIterations = 0
    with pytest.raises(RuntimeError):
        g.height = 3
    with pytest.raises(RuntimeError):
        g.propagationIterations = 3
    with pytest.raises(errors.InvalidArgument):
        g.getFlow(np.zeros((16, 16, 2), np.float64))


# ------------------------------------------------------- PyramidalFlowFilter

def _pyr_pair(ffo, h, w, levels, gamma, maxflow, smooth, fused=True):
    import flowfilter.gpu.flowfilters as gf
    g = gf.PyramidalFlowFilter(h, w, levels)
    g.fused = fused
    g.gamma = gamma if levels > 1 else gamma[0]
    g.maxflow = maxflow
    g.smoothIterations = smooth if levels > 1 else smooth[0]
    o = ffo.PyramidalFlowFilter(h, w, levels)
    o.configure_like(gamma, maxflow, smooth)
    return g, o


@pytest.mark.parametrize("shape,levels,frames", [((68, 128), 2, 16), ((272, 512), 2, 12), ((96, 160), 3, 12),
                                                  ((50, 70), 1, 8), ((75, 131), 2, 8)])
@pytest.mark.parametrize("fused", [True, False])
def test_pyramidal_sequence(ffo, shape, levels, frames, fused):
    h, w = shape
    say("pyramidal", shape, levels, frames, "fused" if fused else "staged")
    gamma = [10.0, 50.0, 80.0][:levels]
    smooth = [2, 2, 1][:levels]
    g, o = _pyr_pair(ffo, h, w, levels, gamma, 4.0, smooth, fused)
    for t, img in enumerate(seq_blobs(h, w, frames)):
        g.loadImage(img)
        g.compute()
        o.compute(img)
        check_frame(t, g.getFlow(), o.downloadFlow(), "pyramidal flow")
        check_frame(t, g.getImage(), o.downloadImage(), "pyramidal image")
        for lvl in range(levels):
            o_s = o.level_state(lvl, 0)
            assert o_s.shape[:2] == ffo.level_sizes(h, w, levels)[lvl]
    dev = g.getFlowDevice().download(np.float32)
    check_frame(frames, dev, o.getFlow(), "pyramidal getFlowDevice")


def _prefetched(gen):
    """Yield the items of ``gen``, producing the next one on a worker thread
    while the caller works on the current one (the oracle's C calls release
    the interpreter lock, the generator's numpy work holds it)."""
    from concurrent.futures import ThreadPoolExecutor
    it = iter(gen)
    with ThreadPoolExecutor(1) as pool:
        fut = pool.submit(next, it, None)
        while True:
            item = fut.result()
            if item is None:
                return
            fut = pool.submit(next, it, None)
            yield item


def test_flowfilter_full_sequence_640x480(ffo):
    """The single-level workload over its whole sequence: 640x480, 200 frames
    of band-limited noise (sigma 3) translating by (0.7, -0.4) px/frame;
    gamma 10, maxflow 4, 2 smoothing iterations. Every frame bitwise equal to
    the FP32 oracle; the FP64 oracle's error reported for every frame."""
    import flowfilter.gpu.flowfilters as gf
    h, w, frames = 480, 640, 200
    ffo.set_threads(16)
    say("flowfilter full sequence", h, w, frames)
    g = gf.FlowFilter(h, w, 2, 4.0, 10.0)
    o = ffo.FlowFilter(h, w, 2, 4.0, 10.0)
    o64 = ffo.FlowFilter(h, w, 2, 4.0, 10.0, prec="f64")
    worst64 = 0.0
    for t, img in enumerate(_prefetched(seq_noise(h, w, frames))):
        g.loadImage(img)
        g.compute()
        o.compute(img)
        o64.compute(img)
        got = g.getFlow()
        check_frame(t, got, o.downloadFlow(), "640x480 flow")
        worst64 = max(worst64, float(np.abs(got - o64.downloadFlow()).max()))
        if t % 50 == 49:
            say("  frame", t, "ok, max |gpu - fp64 oracle| so far %.3g px" % worst64)
    say("  max |gpu - fp64 oracle| over %d frames = %.3g px" % (frames, worst64))
    assert worst64 < 1e-3


def test_pyramidal_full_sequence_2048x1088(ffo):
    """The benchmark workload over its whole sequence: 2048x1088, 300 frames
    of 400 Gaussian blobs (sigma 2-20 px, amplitude 20-80) on mid-grey under
    the capped affine motion (0.5 deg/frame, 1 % expansion, at most 4 px),
    N(0, 2) sensor noise; 2 levels, gamma [10, 50], maxflow 4, smooth [2, 2].
    Every frame bitwise equal to the FP32 oracle; the FP64 oracle runs over
    the first 30 frames (its cost is three times the FP32 oracle's) and its
    error is reported."""
    from flowfilter import synthetic
    h, w, frames, f64_frames = 1088, 2048, 300, 30
    ffo.set_threads(16)
    say("pyramidal full sequence", h, w, frames)
    g, o = _pyr_pair(ffo, h, w, 2, [10.0, 50.0], 4.0, [2, 2])
    o64 = ffo.PyramidalFlowFilter(h, w, 2, prec="f64")
    o64.configure_like([10.0, 50.0], 4.0, [2, 2])
    worst64 = 0.0
    seq = synthetic.blob_affine(h, w, frames, blobs=400)
    for t, img in enumerate(_prefetched(seq)):
        g.loadImage(img)
        g.compute()
        o.compute(img)
        got = g.getFlow()
        check_frame(t, got, o.downloadFlow(), "2048x1088 flow")
        if t < f64_frames:
            o64.compute(img)
            worst64 = max(worst64, float(np.abs(got - o64.downloadFlow()).max()))
        if t % 50 == 49:
            say("  frame", t, "ok")
    assert t == frames - 1
    say("  max |gpu - fp64 oracle| over the first %d frames = %.3g px" % (f64_frames, worst64))
    assert worst64 < 1e-3


def test_pyramidal_parameters_and_errors(ffo):
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import errors
    g = gf.PyramidalFlowFilter(64, 96, 3)
    assert (g.height, g.width, g.levels) == (64, 96, 3)
    # defaults: raw gamma 1.0, maxflow 1.0, one smoothing iteration per level
    assert g.gamma == [1.0, 1.0, 1.0] and g.smoothIterations == [1, 1, 1]
    g.gamma = [10, 50, 100]
    assert np.allclose(g.gamma, np.array([10, 50, 100]) / 65025.0)
    g.maxflow = 8.0
    assert g.maxflow == 8.0
    g.setGammaVector([1.0, 2.0, 3.0])
    assert np.allclose(g.gamma, np.array([1, 2, 3]) / 65025.0)
    with pytest.raises(errors.FilterError):
        g.setGammaVector([1.0, 2.0])
    from flowfilter import _native as N
    with pytest.raises(errors.FilterError):
        N.get_float("ff_pyramidal_get_gamma", g.handle, 3)
    g.propagationBorder = 2
    assert g.propagationBorder == 2
    with pytest.raises(errors.FilterError):
        gf.PyramidalFlowFilter(64, 96, 0)


@pytest.mark.parametrize("graphs", [True, False])
def test_pyramidal_device_input_and_timing(ffo, graphs):
    """Device-resident frames (the benchmark's input path), graph replay on and off."""
    h, w = 272, 320
    g, o = _pyr_pair(ffo, h, w, 2, [10.0, 50.0], 4.0, [2, 2])
    g.graphs = graphs
    assert g.graphs == graphs
    frames = seq_blobs(h, w, 9)
    dev, base, pitch = device_frames(frames)
    for t, img in enumerate(frames):
        g.loadImageDevice(base + t * h * pitch, pitch)
        g.compute()
        o.compute(img)
        check_frame(t, g.getFlow(), o.downloadFlow(), "pyramidal device input")
    assert g.elapsedTime() > 0.0


@pytest.mark.parametrize("shape", [(48, 64), (272, 512)])
@pytest.mark.parametrize("fused", [True, False])
def test_delta_flow_filter_standalone(ffo, shape, fused):
    """DeltaFlowFilter.compute() on its own (image model, first-frame copy,
    propagation, update, smoothing; flowfilter.cu:429-457) against the
    oracle's standalone delta filter, bitwise, with a new upper-level flow
    uploaded before every frame."""
    import flowfilter.gpu.flowfilters as gf
    from flowfilter.gpu.image import GPUImage
    h, w = shape
    say("standalone delta filter", shape, "fused" if fused else "staged")
    img_g = GPUImage((h, w), 1)
    up = GPUImage((h // 2, w // 2, 2), 4)
    rng = np.random.default_rng(3)
    d = gf.DeltaFlowFilter(img_g, up)
    d.fused = fused
    d.gamma = 50.0
    d.maxflow = 2.0
    d.smoothIterations = 2
    assert d.propagationIterations == 2 and (d.height, d.width) == (h, w)
    o = ffo.DeltaFlowFilter(h, w)
    o.setGamma(50.0)
    o.setMaxFlow(2.0)
    o.setSmoothIterations(2)
    for t, img in enumerate(seq_noise(h, w, 6, (0.9, -0.6))):
        upv = rng.uniform(-0.5, 0.5, (h // 2, w // 2, 2)).astype(np.float32)
        up.upload(upv)
        img_g.upload(img)
        d.compute()
        o.compute(img, upv)
        check_frame(t, d.getFlow().download(np.float32), o.getFlow(), "standalone delta flow")
        check_frame(t, d.getImage().download(np.float32), o.getImage(), "standalone delta image")
    assert np.abs(o.getFlow()).max() > 0.1


def test_cpp_api_demo(tmp_path):
    """The C++ classes of include/flowfilter_amd.hpp drive the filter end to end."""
    import subprocess
    from test_abi import _build_cpp_demo
    exe = _build_cpp_demo(tmp_path)
    out = subprocess.run([exe, "240", "320", "2", "8"], capture_output=True, text=True, timeout=120)
    say(out.stdout.strip(), out.stderr.strip()[-500:])
    assert out.returncode == 0 and out.stdout.startswith("ok")


# ------------------------------------------------ fused kernel coverage
# One fused frame launch per frame after the first. These cases exercise it
# against the oracle: tile clamping at the right/bottom edges, odd sizes,
# levels narrower than 64 columns or shorter than 128 rows (a single tile
# whose buffer extends past the image), several propagation/smoothing counts
# and borders, 1-3 levels.

@pytest.mark.parametrize("shape,levels,maxflow,smooth,border", [
    ((272, 512), 2, 4.0, [2, 2], 3),
    ((300, 200), 2, 4.0, [2, 2], 3),
    ((257, 333), 2, 2.5, [1, 3], 3),
    ((512, 700), 3, 4.0, [2, 2, 1], 3),
    ((480, 640), 1, 4.0, [2], 3),
    ((256, 130), 2, 1.0, [2, 2], 1),
    ((260, 400), 2, 6.0, [3, 2], 5),
    ((60, 50), 1, 4.0, [2], 3),
    ((130, 63), 1, 2.0, [1], 1),
    ((129, 65), 1, 4.0, [2], 3),
    ((240, 320), 3, 4.0, [2, 2, 2], 3),
    ((37, 200), 2, 3.0, [2, 1], 2),
    ((5, 7), 1, 1.0, [1], 1),
    ((300, 420), 2, 12.0, [2, 2], 3),
    ((256, 300), 1, 26.0, [1], 2),
])
def test_fused_matches_oracle(ffo, shape, levels, maxflow, smooth, border):
    from flowfilter import _native as N
    h, w = shape
    say("fused", shape, levels, maxflow, smooth, border)
    gamma = [10.0, 50.0, 80.0][:levels]
    g, o = _pyr_pair(ffo, h, w, levels, gamma, maxflow, smooth, fused=True)
    s, _ = _pyr_pair(ffo, h, w, levels, gamma, maxflow, smooth, fused=False)
    g.propagationBorder = border
    s.propagationBorder = border
    o.setPropagationBorder(border)
    frames = 10
    for t, img in enumerate(seq_blobs(h, w, frames)):
        for f in (g, s):
            if f is g:
                N.profile_begin("fused_frame")
            f.loadImage(img)
            f.compute()
            if f is g:
                _, n = N.profile_end()
                # the first frame runs the stage kernels, the second also
                # times the candidate tile shapes, every later frame is
                # one fused frame launch
                assert (n == 0) if t == 0 else ((n >= 1) if t == 1 else (n == 1)), (t, n)
        o.compute(img)
        want = o.downloadFlow()
        check_frame(t, g.getFlow(), want, "fused")
        check_frame(t, s.getFlow(), want, "staged")
        check_frame(t, g.getImage(), o.downloadImage(), "fused image")
    assert g.waitTimeouts() == 0


@pytest.mark.parametrize("levels", [1, 2])
def test_halo_beyond_a_tile_runs_stage_kernels(ffo, levels):
    """maxflow 40 (40 propagation iterations at level 0): the halo of a
    fused tile would exceed its 64 columns, so every frame runs the stage
    kernels, with the oracle's results."""
    from flowfilter import _native as N
    h, w = 200, 260
    say("halo beyond a tile", levels)
    g, o = _pyr_pair(ffo, h, w, levels, [10.0, 40.0][:levels], 40.0, [2, 1][:levels], fused=True)
    for t, img in enumerate(seq_blobs(h, w, 4)):
        g.loadImage(img)
        N.profile_begin("fused_frame")
        g.compute()
        _, n = N.profile_end()
        assert n == 0, n
        o.compute(img)
        check_frame(t, g.getFlow(), o.downloadFlow(), "maxflow 40")


_TILES = [(waves, top_rows, 8, 2) for waves in (16, 15, 8, 5, 4) for top_rows in (4, 8)] + \
    [(waves, 4, 4, interior) for waves in (16, 15) for interior in (2, 1)]


@pytest.mark.parametrize("levels", [2, 3])
@pytest.mark.parametrize("waves,top_rows,rows,interior", _TILES)
def test_fused_tile_heights_match_oracle(ffo, waves, top_rows, rows, interior, levels):
    """Every tile height of the fused frame kernel, with either row count on
    the reduced top level and, in 16- and 15-wave workgroups, four rows per
    thread on every level (the kernel held to 64 VGPRs), with and without
    the interior tiles' own body, gives the oracle's flow."""
    from flowfilter import _native as N
    h, w = 300, 410
    say("tile waves", waves, "top rows", top_rows, "rows", rows, "interior", interior, "levels", levels)
    g, o = _pyr_pair(ffo, h, w, levels, [10.0, 30.0, 60.0][:levels], 4.0, [2, 2, 1][:levels], fused=True)
    g.tileWaves = waves
    g.tileTopRows = top_rows
    g.tileRows = rows
    g.tileInterior = interior
    assert (g.tileWaves, g.tileTopRows, g.tileRows, g.tileInterior) == (waves, top_rows, rows, interior)
    for t, img in enumerate(seq_blobs(h, w, 5)):
        g.loadImage(img)
        N.profile_begin("fused_frame")
        g.compute()
        _, n = N.profile_end()
        assert n == (0 if t == 0 else 1)
        o.compute(img)
        check_frame(t, g.getFlow(), o.downloadFlow(), "waves %d top rows %d rows %d" % (waves, top_rows, rows))


def test_fused_autotune_picks_a_tiling(ffo):
    g, o = _pyr_pair(ffo, 272, 512, 2, [10.0, 50.0], 4.0, [2, 2], fused=True)
    assert g.tileWaves == 0
    for t, img in enumerate(seq_blobs(272, 512, 3)):
        g.loadImage(img)
        g.compute()
        o.compute(img)
        check_frame(t, g.getFlow(), o.downloadFlow(), "autotuned")
    assert g.tileWaves in (16, 15, 8, 5, 4)
    assert g.tileTopRows in (4, 8)
    assert g.tileRows == 8 or (g.tileRows == 4 and g.tileWaves in (16, 15))
    assert g.tileInterior == 2 or (g.tileInterior == 1 and g.tileRows == 4)
    from flowfilter import errors
    with pytest.raises(errors.InvalidArgument):
        g.tileWaves = 7
    with pytest.raises(errors.InvalidArgument):
        g.tileTopRows = 6
    with pytest.raises(errors.InvalidArgument):
        g.tileRows = 2
    with pytest.raises(errors.InvalidArgument):
        g.tileInterior = 3
    # the interior body with a fixed 8-row tiling: refused at the next frame
    f, _ = _pyr_pair(ffo, 272, 512, 2, [10.0, 50.0], 4.0, [2, 2], fused=True)
    f.tileWaves, f.tileTopRows, f.tileRows, f.tileInterior = 8, 4, 8, 1
    imgs = list(seq_blobs(272, 512, 2))
    f.loadImage(imgs[0])
    f.compute()  # the first frame runs the stage kernels
    f.loadImage(imgs[1])
    with pytest.raises(errors.InvalidArgument):
        f.compute()


def _random_config(k):
    """Seeded configuration k of the sweep below: level count, odd and
    even sizes (at least 8 rows and columns on the top level), per-level
    gamma and smoothing, maxflow (hence 1-9 propagation iterations and the
    halo) and border."""
    rng = np.random.default_rng(7000 + k)
    levels = int(rng.integers(1, 5))
    lo = 8 << (levels - 1)
    h = int(rng.integers(lo, max(lo + 1, 560)))
    w = int(rng.integers(lo, max(lo + 1, 720)))
    gamma = [float(np.round(rng.uniform(1.0, 100.0), 3)) for _ in range(levels)]
    maxflow = float(np.round(rng.uniform(0.3, 9.0), 2))
    smooth = [int(rng.integers(1, 4)) for _ in range(levels)]
    border = int(rng.integers(1, 5))
    return (h, w), levels, gamma, maxflow, smooth, border


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("k", range(48))
def test_fused_random_configurations_match_oracle(ffo, k, fused):
    """Forty-eight seeded configurations of the pyramidal filter (1-4 levels,
    odd and even sizes, per-level gamma and smoothing, maxflow 0.3-9,
    border 1-4), six frames each, on the fused frame kernel (the interior
    tiles' own body tuned, on or off) and on the stage kernels: every flow
    and brightness bitwise the oracle's, and no in-launch wait gives up."""
    (h, w), levels, gamma, maxflow, smooth, border = _random_config(k)
    say("random configuration", k, (h, w), levels, gamma, maxflow, smooth, border, "fused" if fused else "staged")
    g, o = _pyr_pair(ffo, h, w, levels, gamma, maxflow, smooth, fused=fused)
    g.propagationBorder = border
    o.setPropagationBorder(border)
    if fused:  # interior tile body tuned, forced on, forced off
        g.tileInterior = (0, 1, 2)[k % 3]
    for t, img in enumerate(seq_blobs(h, w, 6)):
        g.loadImage(img)
        g.compute()
        o.compute(img)
        check_frame(t, g.getFlow(), o.downloadFlow(), "random %d" % k)
        check_frame(t, g.getImage(), o.downloadImage(), "random %d image" % k)
    assert g.waitTimeouts() == 0
