"""FP32-vs-FP64 oracle error over a whole synthetic sequence (CPU only).

TEST INFRASTRUCTURE: runs the FP32 oracle (the parity oracle, which the GPU
filter matches bitwise frame by frame) and the FP64 oracle side by side over
every frame and reports the per-frame max |flow difference| in pixels. Since
the GPU output equals the FP32 oracle's bit for bit, this is also the GPU
filter's FP64 error.

    python oracle/fp64_error.py [--frames 300] [--threads N] [--json out.json]

Default workload: PyramidalFlowFilter 2048x1088, 2 levels, gamma [10, 50],
maxflow 4, smooth [2, 2], the seeded blob texture under the capped affine
motion (flowfilter.synthetic.blob_affine).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "optical-flow-filter_amd"))

import numpy as np  # noqa: E402

import ffo  # noqa: E402
from flowfilter import synthetic  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--height", type=int, default=1088)
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--threads", type=int, default=len(os.sched_getaffinity(0)))
    ap.add_argument("--json", default=None)
    a = ap.parse_args()
    ffo.set_threads(a.threads)
    h, w = a.height, a.width
    f32 = ffo.PyramidalFlowFilter(h, w, 2, prec="f32")
    f64 = ffo.PyramidalFlowFilter(h, w, 2, prec="f64")
    for f in (f32, f64):
        f.configure_like([10.0, 50.0], 4.0, [2, 2])
    per_frame = []
    t0 = time.time()
    for t, img in enumerate(synthetic.blob_affine(h, w, a.frames)):
        f32.compute(img)
        f64.compute(img)
        d = np.abs(f32.downloadFlow().astype(np.float64) - f64.downloadFlow())
        per_frame.append(float(d.max()))
        if t % 25 == 0 or t == a.frames - 1:
            print("frame %d: max |fp32 - fp64| = %.3g px (worst so far %.3g, %.0f s)"
                  % (t, per_frame[-1], max(per_frame), time.time() - t0), flush=True)
    worst = int(np.argmax(per_frame))
    res = {"workload": "PyramidalFlowFilter %dx%d, 2 levels, gamma [10,50], maxflow 4, smooth [2,2], "
                       "blob texture under capped affine motion" % (w, h),
           "frames": a.frames, "threads": a.threads,
           "max_px": max(per_frame), "max_frame": worst, "median_px": float(np.median(per_frame)),
           "per_frame_px": per_frame}
    print(json.dumps({k: v for k, v in res.items() if k != "per_frame_px"}))
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(res, fh)


if __name__ == "__main__":
    main()
