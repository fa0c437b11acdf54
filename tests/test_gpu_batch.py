"""BatchedFlowFilter: many PyramidalFlowFilters of uneven size and depth
advanced by one fused launch per frame step.

Bar: every stream's flow and brightness after every step are bitwise equal
to those of an identical filter computed on its own (itself checked against
the CPU oracle by test_gpu_filters.py), and one stream is also checked
against the oracle directly.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (height, width, levels, gamma per level, maxflow, smooth iterations per level)
STREAMS = [
    (240, 320, 1, [12.0], 4.0, [2]),
    (240, 320, 2, [10.0, 40.0], 4.0, [2, 2]),
    (480, 640, 3, [8.0, 25.0, 50.0], 4.0, [2, 2, 1]),
    (120, 160, 2, [20.0, 30.0], 2.5, [1, 2]),
    (60, 80, 1, [5.0], 3.0, [2]),
    (272, 512, 2, [10.0, 50.0], 4.0, [2, 2]),
    (96, 160, 3, [15.0, 30.0, 45.0], 4.0, [1, 1, 1]),
    (300, 200, 2, [10.0, 50.0], 1.5, [3, 2]),
    (480, 640, 4, [8.0, 20.0, 35.0, 50.0], 4.0, [2, 2, 1, 1]),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(h, w, levels, gamma, maxflow, smooth):
    import flowfilter.gpu.flowfilters as gf
    f = gf.PyramidalFlowFilter(h, w, levels)
    f.gamma = gamma if levels > 1 else gamma[0]
    f.maxflow = maxflow
    f.smoothIterations = smooth if levels > 1 else smooth[0]
    return f


def frames_for(i, h, w, n):
    from flowfilter import synthetic
    rng = np.random.default_rng(100 + i)
    v = tuple(rng.uniform(-2.0, 2.0, size=2))
    return list(synthetic.translating_noise(h, w, n, v, seed=200 + i))


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_matches_separate_filters(ffo, graphs):
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import _native as N
    n = 6
    batch = gf.BatchedFlowFilter()
    batch.graphs = graphs
    alone = []
    seqs = []
    for i, (h, w, levels, gamma, maxflow, smooth) in enumerate(STREAMS):
        batch.add(make(h, w, levels, gamma, maxflow, smooth))
        alone.append(make(h, w, levels, gamma, maxflow, smooth))
        seqs.append(frames_for(i, h, w, n))
    assert len(batch) == len(STREAMS)
    h0, w0, l0, g0, m0, s0 = STREAMS[2]
    o = ffo.PyramidalFlowFilter(h0, w0, l0)
    o.configure_like(g0, m0, s0)
    for t in range(n):
        say("batch step", t, "graphs" if graphs else "no graphs")
        for i in range(len(STREAMS)):
            batch[i].loadImage(seqs[i][t])
            alone[i].loadImage(seqs[i][t])
        N.profile_begin("fused_frame")
        batch.compute()
        _, launches = N.profile_end()
        # step 0: stage kernels; step 1: one launch after timing the tile
        # shapes; later steps: one launch
        assert (launches == 0) if t == 0 else ((launches >= 1) if t == 1 else (launches == 1)), (t, launches)
        for f in alone:
            f.compute()
        for i in range(len(STREAMS)):
            a, b = batch[i].getFlow(), alone[i].getFlow()
            assert np.array_equal(a, b), "stream %d step %d: max |diff| %g" % (i, t, np.abs(a - b).max())
            assert np.array_equal(batch[i].getImage(), alone[i].getImage()), (i, t)
        o.compute(seqs[2][t])
        assert np.array_equal(batch[2].getFlow(), o.downloadFlow()), t
    assert batch.waitTimeouts() == 0
    assert batch.lastBlocks() > 0


def test_batch_streams_joining_later(ffo):
    """A stream added after others have started runs its first frame with
    the stage kernels inside the same step; all results stay identical."""
    import flowfilter.gpu.flowfilters as gf
    n = 5
    specs = [STREAMS[1], STREAMS[2], STREAMS[5]]
    batch = gf.BatchedFlowFilter()
    alone, seqs = [], []
    for i, sp in enumerate(specs):
        seqs.append(frames_for(10 + i, sp[0], sp[1], n + 2))
    for t in range(n + 2):
        if t in (0, 2, 3):  # stream k joins at step 0, 2, 3
            k = (0, 2, 3).index(t)
            batch.add(make(*specs[k]))
            alone.append(make(*specs[k]))
        for k in range(len(alone)):
            start = (0, 2, 3)[k]
            batch[k].loadImage(seqs[k][t - start])
            alone[k].loadImage(seqs[k][t - start])
        batch.compute()
        for f in alone:
            f.compute()
        for k in range(len(alone)):
            assert np.array_equal(batch[k].getFlow(), alone[k].getFlow()), (k, t)


def test_batch_errors():
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import errors
    batch = gf.BatchedFlowFilter()
    with pytest.raises(errors.LogicError):
        batch.compute()
    f = make(*STREAMS[0])
    batch.add(f)
    with pytest.raises(errors.InvalidArgument):
        batch.add(f)
    with pytest.raises(TypeError):
        batch.add(gf.FlowFilter(64, 64))


def test_batch_device_frames_match_host_loads(ffo):
    """loadImagesDevice (one copy launch for all streams) gives the same
    results as loading each filter's frame from the host."""
    import flowfilter.gpu.flowfilters as gf
    from flowfilter.gpu.image import GPUImage
    n = 4
    specs = STREAMS[:5]
    a = gf.BatchedFlowFilter([make(*sp) for sp in specs])
    b = gf.BatchedFlowFilter([make(*sp) for sp in specs])
    seqs = [frames_for(20 + i, sp[0], sp[1], n) for i, sp in enumerate(specs)]
    dev = []
    for i, sp in enumerate(specs):
        g = GPUImage((n * sp[0], sp[1]), 1)
        g.upload(np.concatenate(seqs[i], axis=0))
        dev.append(g)
    for t in range(n):
        a.loadImagesDevice([g.data + t * sp[0] * g.pitch for g, sp in zip(dev, specs)], [g.pitch for g in dev])
        for i in range(len(specs)):
            b[i].loadImage(seqs[i][t])
        a.compute()
        b.compute()
        for i in range(len(specs)):
            assert np.array_equal(a[i].getFlow(), b[i].getFlow()), (i, t)
    from flowfilter import errors
    with pytest.raises(errors.InvalidArgument):
        a.loadImagesDevice([dev[0].data], [dev[0].pitch])


def test_batch_frames_in_place_match_copies(ffo):
    """setImagesDevice (every stream's frame read where it lies by the
    batched launch; the 4-level stream and the first frames copy instead)
    gives the same results as loadImagesDevice (one copy launch)."""
    import flowfilter.gpu.flowfilters as gf
    from flowfilter.gpu.image import GPUImage
    n = 5
    specs = STREAMS
    a = gf.BatchedFlowFilter([make(*sp) for sp in specs])
    b = gf.BatchedFlowFilter([make(*sp) for sp in specs])
    seqs = [frames_for(30 + i, sp[0], sp[1], n) for i, sp in enumerate(specs)]
    dev = []
    for i, sp in enumerate(specs):
        g = GPUImage((n * sp[0], sp[1]), 1)
        g.upload(np.concatenate(seqs[i], axis=0))
        dev.append(g)
    for t in range(n):
        say("batch frames in place, step", t)
        ptrs = [g.data + t * sp[0] * g.pitch for g, sp in zip(dev, specs)]
        a.setImagesDevice(ptrs, [g.pitch for g in dev])
        b.loadImagesDevice(ptrs, [g.pitch for g in dev])
        a.compute()
        b.compute()
        for i in range(len(specs)):
            assert np.array_equal(a[i].getFlow(), b[i].getFlow()), (i, t)
            assert np.array_equal(a[i].getImage(), b[i].getImage()), (i, t)
    h0, w0, l0, g0, m0, s0 = specs[2]
    o = ffo.PyramidalFlowFilter(h0, w0, l0)
    o.configure_like(g0, m0, s0)
    for img in seqs[2]:
        o.compute(img)
    assert np.array_equal(a[2].getFlow(), o.downloadFlow())
    assert a.waitTimeouts() == 0
    from flowfilter import errors
    with pytest.raises(errors.InvalidArgument):
        a.setImagesDevice([dev[0].data], [dev[0].pitch])


@pytest.mark.parametrize("in_place", [False, True])
def test_batch_sixty_four_streams_hundred_twenty_steps(in_place):
    """The 64-stream workload (seed 3: 320x240 to 1920x1080, 1-3 levels,
    gamma U[5,50], smooth 1-2, affine motion), 120 steps over each stream's
    frames (cycled), frames loaded from device memory in one launch (or read
    in place by the batched launch): every stream's flow and brightness
    bitwise equal to its filter computed alone (loop of loadImageDevice +
    compute, or setImageDevice + compute) at steps 1, 60 and 119."""
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import synthetic
    from flowfilter.gpu.image import GPUImage
    streams = synthetic.camera_streams(64, 3)
    dev = []
    for sp in streams:
        g = GPUImage((3 * sp["h"], sp["w"]), 1)
        g.upload(np.concatenate(sp["frames"], axis=0))
        dev.append(g)
    pitches = [g.pitch for g in dev]

    def build():
        return [make(sp["h"], sp["w"], sp["levels"], sp["gamma"], sp["maxflow"], sp["smooth"]) for sp in streams]

    batched, alone = build(), build()
    batch = gf.BatchedFlowFilter(batched)
    for t in range(120):
        ptrs = [g.data + (t % 3) * sp["h"] * g.pitch for sp, g in zip(streams, dev)]
        if in_place:
            batch.setImagesDevice(ptrs, pitches)
        else:
            batch.loadImagesDevice(ptrs, pitches)
        batch.compute()
        for f, p, q in zip(alone, ptrs, pitches):
            if in_place:
                f.setImageDevice(p, q)
            else:
                f.loadImageDevice(p, q)
            f.compute()
        if t in (1, 60, 119):
            say("64 streams, step", t, "in place" if in_place else "copied")
            for i in range(len(streams)):
                a, b = batched[i].getFlow(), alone[i].getFlow()
                assert np.array_equal(a, b), "stream %d step %d: max |diff| %g" % (i, t, np.abs(a - b).max())
                assert np.array_equal(batched[i].getImage(), alone[i].getImage()), (i, t)
    assert batch.waitTimeouts() == 0


def _random_mix(k):
    """Seeded batch k of the sweep below: 3-8 streams, each of random size
    (at least 8 rows and columns on its top level), 1-4 levels, per-level
    gamma and smoothing, maxflow and border."""
    rng = np.random.default_rng(9100 + k)
    streams = []
    for _ in range(int(rng.integers(3, 9))):
        levels = int(rng.integers(1, 5))
        lo = 8 << (levels - 1)
        h = int(rng.integers(lo, max(lo + 1, 500)))
        w = int(rng.integers(lo, max(lo + 1, 640)))
        gamma = [float(np.round(rng.uniform(2.0, 80.0), 3)) for _ in range(levels)]
        maxflow = float(np.round(rng.uniform(0.5, 6.0), 2))
        smooth = [int(rng.integers(1, 4)) for _ in range(levels)]
        border = int(rng.integers(1, 4))
        streams.append((h, w, levels, gamma, maxflow, smooth, border))
    return streams


@pytest.mark.parametrize("k", range(12))
def test_batch_random_mixes_match_oracle(ffo, k):
    """Twelve seeded batches of 3-8 streams of random size and depth
    (1-4 levels, per-level gamma and smoothing, maxflow 0.5-6, border 1-3),
    replayed as graphs or launched directly, the interior tiles' own body
    tuned, on or off, five steps each: every stream's
    flow and brightness bitwise the oracle's after every step, and no
    in-launch wait gives up."""
    import flowfilter.gpu.flowfilters as gf
    mix = _random_mix(k)
    say("random batch", k, [(h, w, levels) for h, w, levels, *_ in mix])
    batch = gf.BatchedFlowFilter()
    batch.graphs = k % 2 == 0
    batch.tileInterior = (0, 1, 2)[k % 3]  # interior tile body tuned, on, off
    oracles, seqs = [], []
    for i, (h, w, levels, gamma, maxflow, smooth, border) in enumerate(mix):
        f = make(h, w, levels, gamma, maxflow, smooth)
        f.propagationBorder = border
        batch.add(f)
        o = ffo.PyramidalFlowFilter(h, w, levels)
        o.configure_like(gamma, maxflow, smooth)
        o.setPropagationBorder(border)
        oracles.append(o)
        seqs.append(frames_for(300 + 10 * k + i, h, w, 5))
    for t in range(5):
        for i in range(len(mix)):
            batch[i].loadImage(seqs[i][t])
        batch.compute()
        for i, o in enumerate(oracles):
            o.compute(seqs[i][t])
            a, b = batch[i].getFlow(), o.downloadFlow()
            assert np.array_equal(a, b), "batch %d stream %d step %d: max |diff| %g" % (k, i, t, np.abs(a - b).max())
            assert np.array_equal(batch[i].getImage(), o.downloadImage()), (k, i, t)
    assert batch.waitTimeouts() == 0
