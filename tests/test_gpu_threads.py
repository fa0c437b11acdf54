"""Filters driven from several host threads at once: one filter per thread,
as the library's contract allows (objects are not shared between threads;
the library itself is), plus a batch of filters on a thread of its own.
Each filter and batch enqueues on a stream of its own. Through the C ABI via
ctypes, which releases the interpreter lock during every call, so the calls
overlap.

Bar: every thread's flow is bitwise equal, frame by frame, to the FP32
oracle's over the same sequence computed alone."""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (height, width, levels); the 4-level pyramid replays its frames as HIP
# graphs captured on its own stream while the other threads enqueue work
CASES = [(96, 160, 2), (128, 200, 3), (75, 131, 1), (272, 512, 2), (64, 96, 1), (150, 250, 3), (544, 1024, 2),
         (300, 410, 3), (480, 640, 4)]
GAMMA, SMOOTH, FRAMES = [10.0, 30.0, 60.0, 80.0], [2, 2, 1, 1], 10
# streams of the batch run on a thread of its own beside the filters
BATCH = [(120, 160, 1), (240, 320, 2), (200, 300, 3)]


def test_one_filter_per_thread(ffo):
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import synthetic
    seqs = [list(synthetic.blob_affine(h, w, FRAMES, blobs=30, seed=10 + i)) for i, (h, w, _) in enumerate(CASES)]
    want = []
    for (h, w, L), frames in zip(CASES, seqs):
        o = ffo.PyramidalFlowFilter(h, w, L)
        o.configure_like(GAMMA[:L], 4.0, SMOOTH[:L])
        per = []
        for img in frames:
            o.compute(img)
            per.append(o.downloadFlow())
        want.append(per)

    bseqs = [list(synthetic.blob_affine(h, w, FRAMES, blobs=30, seed=40 + i)) for i, (h, w, _) in enumerate(BATCH)]
    bwant = []
    for (h, w, L), frames in zip(BATCH, bseqs):
        o = ffo.PyramidalFlowFilter(h, w, L)
        o.configure_like(GAMMA[:L], 4.0, SMOOTH[:L])
        per = []
        for img in frames:
            o.compute(img)
            per.append(o.downloadFlow())
        bwant.append(per)

    failures = []
    start = threading.Barrier(len(CASES) + 1)

    def run_batch():
        try:
            fs = []
            for h, w, L in BATCH:
                g = gf.PyramidalFlowFilter(h, w, L)
                g.gamma = GAMMA[:L] if L > 1 else GAMMA[0]
                g.maxflow = 4.0
                g.smoothIterations = SMOOTH[:L] if L > 1 else SMOOTH[0]
                fs.append(g)
            batch = gf.BatchedFlowFilter(fs)
            start.wait(timeout=60)
            for t in range(FRAMES):
                for g, seq in zip(fs, bseqs):
                    g.loadImage(seq[t])
                batch.compute()
                for k, g in enumerate(fs):
                    got = g.getFlow()
                    if not np.array_equal(got, bwant[k][t]):
                        failures.append(("batch", BATCH[k], t, float(np.abs(got - bwant[k][t]).max())))
        except Exception as e:
            failures.append(("batch", repr(e)))

    def run(i):
        try:
            h, w, L = CASES[i]
            g = gf.PyramidalFlowFilter(h, w, L)
            g.gamma = GAMMA[:L] if L > 1 else GAMMA[0]
            g.maxflow = 4.0
            g.smoothIterations = SMOOTH[:L] if L > 1 else SMOOTH[0]
            start.wait(timeout=60)
            for t, img in enumerate(seqs[i]):
                g.loadImage(img)
                g.compute()
                got = g.getFlow()
                if not np.array_equal(got, want[i][t]):
                    failures.append((CASES[i], t, float(np.abs(got - want[i][t]).max())))
        except Exception as e:  # reported below with the case that raised it
            failures.append((CASES[i], repr(e)))

    print("threads:", len(CASES) + 1, flush=True)
    sys.stdout.flush()
    threads = [threading.Thread(target=run, args=(i,)) for i in range(len(CASES))]
    threads.append(threading.Thread(target=run_batch))
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "a thread did not finish"
    assert not failures, failures
