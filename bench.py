#!/usr/bin/env python3
"""Throughput of the optical-flow filter's per-frame hot path on one MI355X.

Workload: PyramidalFlowFilter at 2048x1088 with 2 levels, gamma [10, 50],
maxflow 4 (4 propagation sub-iterations at level 0, 2 at level 1),
2 smoothing iterations per level, uint8 frames, FP32 flow. Frames are a
seeded synthetic sequence (400 Gaussian blobs on mid-grey, affine motion of
0.5 deg rotation + 1% expansion per frame capped at 4 px, N(0, 2) sensor
noise), uploaded to HBM before timing and cycled.

One step = compute() of one frame read in place from HBM (the frame's
device address handed over with setImageDevice; the fused frame kernel
loads it directly). W warm-up steps, then K timed steps between two HIP
events on the filter's stream, with the device synchronized before and
after. The same K steps with a device loadImage (a copy into the filter's
input image) before each compute() are reported as `value_with_copy`, and
with a host loadImage (the reference's boundary, PCIe included; wall clock,
best of three runs, each listed in `host_upload_ms_per_step_runs`) as
`value_host_upload`.

Prints one JSON line: Mpix/s of level-0 input pixels, the roofline of the
dominant kernel (algorithmic bytes per launch / its average duration from
HIP events on its stream, over a second pass of the same K steps; `traffic`
is the HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
passes, calibrated on a copy kernel, read from
profiles/fused_frame_2048x1088_2lvl_traffic.json; `valu_issue` is the
kernel's VALU-issue floor, its SQ_INSTS_VALU count from
profiles/fused_frame_2048x1088_2lvl_valu.json over the CUs' issue rate,
against the measured launch time: how busy the VALU is), the 64-stream
batched measurement, FlowToColor and the rotational predictor, the reference
notebook's own runtime measurement (640x480, 2 levels, mean elapsedTime()
beside the numbers the notebook published), the differentiable-window
measurements (forward and forward + backward of a 3-level 3840x2160
32-frame window and of 16 sequences of 512x512 x 8 frames, with peak HBM),
and a CPU baseline (the FP32 CPU oracle on the host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "optical-flow-filter_amd"))

H, W, LEVELS = 1088, 2048, 2
GAMMA, MAXFLOW, SMOOTH = [10.0, 50.0], 4.0, [2, 2]
HBM_PEAK_GBPS = 8000.0


def algorithmic_bytes():
    """Compulsory HBM bytes of one frame of the filter at this workload
    (the reference algorithm's state in + state out + input, halos excluded).

    Per pixel of its level: delta level 43 B (u8 image 1, flow 8, delta flow 8,
    old brightness 4, upper-level flow 8 B at 1/4 density = 2, new flow 8,
    delta flow 8, brightness 4); top level 33 B (u8 image 1, flow 8,
    brightness 4, new smoothed flow 8, brightness 4, unsmoothed flow 8, which
    the level below reads); pyramid 1 B per level-1 pixel (the u8 write of
    the reduced image). 43 * 2,228,224 + 33 * 557,056 + 557,056 =
    114,753,536 B per frame, which the fused frame kernel moves in one launch
    (its top level reduces the level-0 image while loading it, so its
    actual compulsory traffic is 1.1 MB lower). One payload propagation
    pass of the staged path: 40 B (read and write flow 8, scalar 4, vector 8)."""
    n0 = H * W
    n1 = (H // 2) * (W // 2)
    return {
        "fused_frame": 43 * n0 + 33 * n1 + n1,
        "propagate_payload": 40 * n0,
    }


_LSCPU_FIELDS = ("Model name", "BIOS Model name", "CPU family", "Model", "Stepping", "Socket(s)",
                 "Core(s) per socket", "Thread(s) per core", "CPU max MHz", "NUMA node(s)")


def _lscpu():
    """Identifying fields of lscpu's report ({} without lscpu)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return {}
    info = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            if k.strip() in _LSCPU_FIELDS:
                info[k.strip()] = v.strip()
    return info


def _cpu_model():
    """The host CPU's model name (lscpu), else /proc/cpuinfo's."""
    name = _lscpu().get("Model name")
    if name:
        return name
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Cores this process may run on (its CPU affinity)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(frames, budget_s=12.0, threads=None):
    """The FP32 CPU oracle on the host: OpenMP over rows on every core this
    process may use (the value), then on one thread (reported beside it),
    each for about budget_s / 2 seconds of frames."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ffo

    def run(nthreads, budget):
        ffo.set_threads(nthreads)
        o = ffo.PyramidalFlowFilter(H, W, LEVELS)
        o.configure_like(GAMMA, MAXFLOW, SMOOTH)
        n, t0 = 0, time.perf_counter()
        while True:
            o.compute(frames[n % len(frames)])
            n += 1
            el = time.perf_counter() - t0
            if (el >= budget and n >= 2) or n >= 2000:
                return n, el

    threads = threads or host_cores()
    n, el = run(threads, budget_s / 2)
    n1, el1 = run(1, budget_s / 2)
    return {"value": round(H * W * n / el / 1e6, 3), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "single_thread_value": round(H * W * n1 / el1 / 1e6, 3), "cpu_model": _cpu_model(),
            "cpu_lscpu": _lscpu(),
            "sample": "%d frames (%d threads, %.1f s) and %d frames (1 thread, %.1f s) of the same 2048x1088 "
                      "2-level workload, FP32 CPU oracle (oracle/ff_oracle.hpp, OpenMP over rows)"
                      % (n, threads, el, n1, el1)}


def batched_bench(args, N, gf, GPUImage):
    """64 camera streams (synthetic.camera_streams, seed 3: 320x240 to
    1920x1080, 1-3 levels each) resident in HBM. One step hands every
    stream its next frame in place (device address, no copy) and advances
    all of them: batched = BatchedFlowFilter.setImagesDevice and compute()
    (one fused launch), plain = each filter's setImageDevice and compute()
    in a loop, the filters sharing one stream (the reference runs every
    filter on stream 0, one after the other). `plain_own_streams` repeats
    the loop with each filter on its own stream (the library's default,
    which lets the GPU overlap the filters' launches). Each mode is timed
    three times (best run kept). Returns aggregate level-0 Mpix/s and the
    ratios."""
    import numpy as np
    from flowfilter import synthetic
    t0 = time.perf_counter()
    streams = synthetic.camera_streams(64, 3)
    print("generated 64 streams in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    dev = []
    for sp in streams:
        F = len(sp["frames"])
        g = GPUImage((F * sp["h"], sp["w"]), 1)
        g.upload(np.concatenate(sp["frames"], axis=0))
        dev.append(g)

    def build():
        fs = []
        for sp in streams:
            f = gf.PyramidalFlowFilter(sp["h"], sp["w"], sp["levels"])
            f.gamma = sp["gamma"] if sp["levels"] > 1 else sp["gamma"][0]
            f.maxflow = sp["maxflow"]
            f.smoothIterations = sp["smooth"] if sp["levels"] > 1 else sp["smooth"][0]
            fs.append(f)
        return fs

    def frame_ptrs(i):
        return [g.data + (i % len(sp["frames"])) * sp["h"] * g.pitch for sp, g in zip(streams, dev)]

    pitches = [g.pitch for g in dev]

    px = sum(sp["h"] * sp["w"] for sp in streams)
    res = {}
    for mode in ("batched", "plain", "plain_own_streams"):
        fs = build()
        batch = gf.BatchedFlowFilter(fs) if mode == "batched" else None
        if mode == "plain":
            # every filter on the first one's stream
            for f in fs[1:]:
                f.stream = fs[0].stream

        def step(i):
            ptrs = frame_ptrs(i)
            if batch is not None:
                batch.setImagesDevice(ptrs, pitches)
                batch.compute()
            else:
                for f, p, q in zip(fs, ptrs, pitches):
                    f.setImageDevice(p, q)
                    f.compute()

        for i in range(5):
            step(i)
        N.synchronize()
        # best of three timed runs per mode (the host-driven plain loop is
        # noisy run to run); wall clock between device-wide waits, which
        # covers every stream the mode uses
        ms = None
        for rep in range(3):
            N.synchronize()
            t0 = time.perf_counter()
            for i in range(args.batched_steps):
                step(5 + rep * args.batched_steps + i)
            N.synchronize()
            t = (time.perf_counter() - t0) * 1e3
            ms = t if ms is None else min(ms, t)
        res[mode] = px * args.batched_steps / (ms / 1e3) / 1e6
        if batch is not None:
            res["blocks"] = batch.lastBlocks()
            res["tile"] = {"waves": batch.tileWaves, "top_rows": batch.tileTopRows, "rows": batch.tileRows,
                           "interior": batch.tileInterior}
            res["timeouts"] = batch.waitTimeouts()
        print("%s: %.3f ms/step, %.1f Mpix/s" % (mode, ms / args.batched_steps, res[mode]), file=sys.stderr,
              flush=True)
        if mode == "plain":
            for f in fs[1:]:
                f.stream = None  # before the first filter (and its stream) goes
    return {"streams": len(streams), "mpix_per_step": round(px / 1e6, 3), "steps": args.batched_steps,
            "value": round(res["batched"], 3), "plain_value": round(res["plain"], 3), "unit": "Mpix/s",
            "speedup": round(res["batched"] / res["plain"], 3),
            "plain_own_streams_value": round(res["plain_own_streams"], 3),
            "speedup_vs_own_streams": round(res["batched"] / res["plain_own_streams"], 3),
            "launch_blocks": res["blocks"],
            "tile": res["tile"],
            "wait_timeouts": res["timeouts"],
            "workload": "64 camera streams (seed 3): 320x240..1920x1080, 1-3 levels, gamma U[5,50], "
                        "maxflow 4, smooth 1-2, affine motion; step = hand each stream its next frame in HBM "
                        "(in place, no copy) + advance 64 filters; plain = the same loop of filters on one stream"}


def reference_notebook_bench(args, gf):
    """The runtime measurement of the reference's own notebook
    (notebooks/GPU_example.ipynb, cells 7 and 10): a 640x480 two-level
    PyramidalFlowFilter, gamma [10, 50], maxflow 4, smooth [2, 4]; every
    frame loadImage, compute() and elapsedTime() (the GPU time of compute()),
    averaged. Cell 10 feeds 1000 all-zero frames (published 0.995867583513
    ms); cell 7 averages 20 frames of a rendered sequence (published
    1.17239360809 ms), here 20 frames of the synthetic blob sequence after
    100 warm-up frames. The reference's hardware is not stated."""
    import numpy as np
    from flowfilter import synthetic

    def make():
        f = gf.PyramidalFlowFilter(480, 640, 2)
        f.gamma = [10.0, 50.0]
        f.maxflow = 4.0
        f.smoothIterations = [2, 4]
        return f

    f = make()
    zero = np.zeros((480, 640), np.uint8)
    et = []
    for k in range(1000):
        f.loadImage(zero)
        f.compute()
        et.append(f.elapsedTime())
    zero_ms = float(np.mean(et))
    f = make()
    frames = list(synthetic.blob_affine(480, 640, 120, seed=5))
    et = []
    for k, img in enumerate(frames):
        f.loadImage(img)
        f.compute()
        if k >= 100:
            et.append(f.elapsedTime())
    tex_ms = float(np.mean(et))
    ref_zero, ref_tex = 0.995867583513, 1.17239360809
    return {"workload": "640x480 PyramidalFlowFilter, 2 levels, gamma [10,50], maxflow 4, smooth [2,4]; "
                        "per frame loadImage + compute() + elapsedTime(), mean of elapsedTime()",
            "zero_frames": {"frames": 1000, "ms_per_compute": round(zero_ms, 5),
                            "reference_published_ms": ref_zero, "speedup": round(ref_zero / zero_ms, 1)},
            "textured_frames": {"frames": 20, "ms_per_compute": round(tex_ms, 5),
                                "reference_published_ms": ref_tex, "speedup": round(ref_tex / tex_ms, 1),
                                "note": "synthetic blob sequence; the reference's rendered frames are not available"},
            "reference_hardware": "not stated"}


def display_rotation_bench(args, N, GPUImage):
    """The display and image-prediction stages at 2048x1088 (library stream):
    FlowToColor of a flow field (12 B/px: flow 8 in, RGBA 4 out; the wheel
    is L2-resident), and RotationalFlowImagePredictor with 4 iterations on a
    float image (rotational flow 8 B/px written, then 8 Lax-Wendroff passes
    of 16 B/px: flow 8 + image 4 in, image 4 out). Kernel times from HIP
    events around each launch; GB/s = algorithmic bytes / kernel time."""
    import numpy as np
    from flowfilter.gpu.camera import createPerspectiveCamera
    from flowfilter.gpu.display import FlowToColor
    from flowfilter.gpu.rotation import RotationalFlowImagePredictor
    rng = np.random.default_rng(5)
    n = H * W
    flow = GPUImage((H, W, 2), 4)
    flow.upload(rng.uniform(-5, 5, (H, W, 2)).astype(np.float32))
    color = FlowToColor(flow, 4.0)
    img = GPUImage((H, W), 4)
    img.upload(rng.uniform(0, 1, (H, W)).astype(np.float32))
    pred = RotationalFlowImagePredictor(createPerspectiveCamera(4.0, H, W, 3.6, 4.8), img)
    pred.iterations = 4
    pred.setAngularVelocity(0.01, -0.02, 0.005)
    out = {}
    for name, stage, kernels in (("flow_to_color", color, {"flow_to_color": 12}),
                                 ("rotational_predictor", pred, {"rotational_flow": 8, "lax_wendroff": 16})):
        for _ in range(args.warmup):
            stage.compute()
        N.synchronize()
        m0 = N.Marker()
        for _ in range(args.steps):
            stage.compute()
        m1 = N.Marker()
        ms = m0.elapsed_ms(m1) / args.steps
        res = {"ms_per_compute": round(ms, 5), "mpix_s": round(n / (ms / 1e3) / 1e6, 1)}
        for k, bpp in kernels.items():
            N.profile_begin(k)
            for _ in range(args.steps):
                stage.compute()
            tot, launches = N.profile_end()
            avg = tot / max(launches, 1)
            gbs = bpp * n / (avg / 1e3) / 1e9
            res[k] = {"launches_per_compute": launches // args.steps, "avg_launch_ms": round(avg, 5),
                      "bytes_per_px": bpp, "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBPS, 4)}
        out[name] = res
        print("%s: %s" % (name, res), file=sys.stderr, flush=True)
    out["workload"] = ("2048x1088: FlowToColor (maxflow 4) of a random flow; RotationalFlowImagePredictor "
                       "(f 4 mm, sensor 3.6x4.8 mm, w (0.01, -0.02, 0.005), 4 iterations) of a float image")
    return out


def copy_bench(args, N, GPUImage):
    """GPUImage.copyFrom of one float2 flow image, K times: a kernel whose HBM
    traffic is known exactly (H*W*8 bytes read, the same written)."""
    import numpy as np
    a, b = GPUImage((H, W, 2), 4), GPUImage((H, W, 2), 4)
    a.upload(np.random.default_rng(0).random((H, W, 2), dtype=np.float32))
    for _ in range(args.warmup):
        b.copyFrom(a)
    N.synchronize()
    m0 = N.Marker()
    for _ in range(args.steps):
        b.copyFrom(a)
    m1 = N.Marker()
    ms = m0.elapsed_ms(m1) / args.steps
    nbytes = 2 * H * W * 8
    print(json.dumps({"kernel": "k_copy2d_u64", "bytes_read": H * W * 8, "bytes_written": H * W * 8,
                      "ms_per_launch": round(ms, 5), "GB/s": round(nbytes / (ms / 1e3) / 1e9, 2),
                      "launches": args.steps}), flush=True)


def differentiable_bench(args):
    """Forward + backward of the differentiable filter over an unrolled
    window: 3 levels at 3840x2160, 32 frames, maxflow 8, checkpoint every 4
    frames, loss = L2 to the true flow field + 0.1 x photometric error;
    and the 16-sequence 512x512 8-frame 2-level window."""
    import torch
    import numpy as np
    from flowfilter import autograd as ad, synthetic

    def measure(frames, gamma, spec, truth, reps, checkpoint):
        def once(timing):
            x = frames.detach().requires_grad_(True)
            gm = gamma.detach().clone().requires_grad_(True)
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record()
            flows = ad.flow_window(x, gm, spec, checkpoint=checkpoint)
            e1.record()
            ad.flow_loss(x, flows, truth).backward()
            e2.record()
            torch.cuda.synchronize()
            if timing is not None:
                timing.append((e0.elapsed_time(e1), e0.elapsed_time(e2)))
            return x.grad, gm.grad
        once(None)
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        times = []
        g = gg = None
        for _ in range(reps):
            g = gg = None  # the peak is one call's, not two calls' gradients
            g, gg = once(times)
        assert bool(torch.isfinite(g).all()) and bool(torch.isfinite(gg).all())
        fwd = min(t[0] for t in times)
        tot = min(t[1] for t in times)
        pix = frames.shape[0] * frames.shape[1] * frames.shape[2] * frames.shape[3]
        return {"ms_forward": round(fwd, 2), "ms_forward_backward": round(tot, 2),
                "forward_mpix_s": round(pix / (fwd / 1e3) / 1e6, 1),
                "forward_backward_mpix_s": round(pix / (tot / 1e3) / 1e6, 1),
                "peak_hbm_gib": round(torch.cuda.max_memory_allocated() / 2 ** 30, 3), "reps": reps}

    t0 = time.perf_counter()
    frames, v = synthetic.smooth_flow_sequence(2160, 3840, 32, maxflow=8.0, flow_sigma=64.0, seed=7)
    torch.cuda.synchronize()
    print("generated the 3840x2160 window in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    spec = ad.FilterSpec(2160, 3840, 3, maxflow=8.0, smooth=2)
    big = measure(frames[None], torch.tensor([10.0, 30.0, 50.0], device="cuda"), spec, v[None], args.ad_reps, 4)
    big["workload"] = ("3-level filter, 3840x2160 u8-valued frames, 32-frame window, maxflow 8, smooth 2, gamma "
                       "[10,30,50], checkpoint every 4 frames; loss: L2 to the flow field + 0.1 x photometric; "
                       "gradients w.r.t. frames and gamma; Mpix/s = level-0 pixels x frames / time")
    del frames, v
    print("differentiable 4K window: %s" % big, file=sys.stderr, flush=True)
    rng = np.random.default_rng(13)
    seqs, vel = [], []
    for b in range(16):
        vv = tuple(rng.uniform(-3.0, 3.0, size=2))
        vel.append(vv)
        seqs.append(np.stack(list(synthetic.translating_noise(512, 512, 8, vv, seed=1300 + b))))
    frames = torch.from_numpy(np.stack(seqs).astype(np.float32) / np.float32(255.0)).cuda()
    spec = ad.FilterSpec(512, 512, 2, maxflow=4.0, smooth=2)
    small = measure(frames, torch.tensor([10.0, 50.0], device="cuda"), spec,
                    torch.tensor(vel, dtype=torch.float32, device="cuda"), args.ad_reps, 4)
    small["workload"] = ("16 sequences of 512x512 translating noise (U[-3,3]^2 px/frame), 8 frames, 2 levels, "
                         "maxflow 4; loss: L2 to the true flow + 0.1 x photometric")
    print("differentiable 16x512^2 window: %s" % small, file=sys.stderr, flush=True)
    return {"window_3840x2160": big, "batch16_512x512": small}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--frames", type=int, default=24, help="distinct synthetic frames kept in HBM")
    ap.add_argument("--staged", action="store_true", help="one launch per stage kernel instead of fused")
    ap.add_argument("--tile", default="", help="fix the fused kernel's tile shape: waves,top_rows,rows "
                                               "(e.g. 16,4,4); default: tuned")
    ap.add_argument("--no-graphs", action="store_true", help="launch the fused kernels one by one")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-stream batched measurement")
    ap.add_argument("--no-copy-variant", action="store_true",
                    help="skip the pass that loads each frame with a copy (value_with_copy)")
    ap.add_argument("--batched-steps", type=int, default=60)
    ap.add_argument("--no-display", action="store_true",
                    help="skip the FlowToColor / rotational predictor and the reference notebook measurements")
    ap.add_argument("--no-differentiable", action="store_true",
                    help="skip the differentiable-window measurements (and the torch import they need)")
    ap.add_argument("--ad-reps", type=int, default=2)
    ap.add_argument("--ad-only", action="store_true", help="only the differentiable-window measurements")
    ap.add_argument("--copy", action="store_true",
                    help="instead of the filter, time a plain device copy of one 2048x1088 flow image "
                         "(17.8 MB read + 17.8 MB written per launch; the reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this path has no multi-GPU form "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if not args.no_differentiable:
        # torch before the library: one HIP runtime in the process (torch
        # ships its own under the system runtime's soname)
        import torch  # noqa: F401
        if args.ad_only:
            print(json.dumps({"differentiable": differentiable_bench(args)}), flush=True)
            return

    import numpy as np
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import _native as N, synthetic
    from flowfilter.gpu.image import GPUImage

    if args.copy:
        copy_bench(args, N, GPUImage)
        return

    t0 = time.perf_counter()
    frames = np.stack(list(synthetic.blob_affine(H, W, args.frames)))
    print("generated %d frames in %.1f s" % (len(frames), time.perf_counter() - t0), file=sys.stderr, flush=True)
    F = len(frames)
    # all frames resident in HBM as one tall u8 image, frame t at rows t*H..
    dev = GPUImage((F * H, W), 1)
    dev.upload(frames.reshape(F * H, W))
    base, pitch = dev.data, dev.pitch

    f = gf.PyramidalFlowFilter(H, W, LEVELS)
    f.gamma = GAMMA
    f.maxflow = MAXFLOW
    f.smoothIterations = SMOOTH
    f.fused = not args.staged
    f.graphs = not args.no_graphs
    if args.tile:
        # a fixed tile shape (profiling runs); by default the tuner picks it
        f.tileWaves, f.tileTopRows, f.tileRows = (int(v) for v in args.tile.split(","))

    def step(i):
        f.setImageDevice(base + (i % F) * H * pitch, pitch)
        f.compute()

    def step_copy(i):
        f.loadImageDevice(base + (i % F) * H * pitch, pitch)
        f.compute()

    for i in range(args.warmup):
        step(i)
    N.synchronize()
    print("warm-up done", file=sys.stderr, flush=True)

    def timed(first, fn=step):
        N.synchronize()
        w0 = time.perf_counter()
        m0 = N.Marker(f)
        for i in range(args.steps):
            fn(first + i)
        m1 = N.Marker(f)
        enq = time.perf_counter() - w0
        ms = m0.elapsed_ms(m1)
        return ms, time.perf_counter() - w0, enq

    ms, wall, enq = timed(args.warmup)
    ms_per_step = ms / args.steps
    value = H * W * args.steps / (ms / 1e3) / 1e6
    print("timed %d steps: %.3f ms/step (wall %.3f s; host enqueue %.3f ms/step)"
          % (args.steps, ms_per_step, wall, enq * 1e3 / args.steps), file=sys.stderr, flush=True)
    value_copy = None
    if not args.no_copy_variant:
        ms_copy, _, _ = timed(args.warmup, step_copy)
        value_copy = round(H * W * args.steps / (ms_copy / 1e3) / 1e6, 3)
        print("with a device loadImage copy per step: %.3f ms/step" % (ms_copy / args.steps), file=sys.stderr,
              flush=True)
    # the reference's own boundary: loadImage of a host frame (PCIe
    # included), then compute(); wall clock over the same number of steps
    value_host = None
    if not args.no_copy_variant:
        # untimed pass over every frame first: the first copy from a host
        # page pays a one-time cost in the driver (0.34 against 0.13 ms/step
        # measured on 2048x1088 u8 frames)
        for i in range(max(F, args.warmup)):
            f.loadImage(frames[i % F])
            f.compute()
        # best of three timed runs (wall clock, so host threads and the
        # driver's queues add run-to-run noise); every run is reported
        host_runs = []
        for rep in range(3):
            N.synchronize()
            w0 = time.perf_counter()
            for i in range(args.steps):
                f.loadImage(frames[i % F])
                f.compute()
            N.synchronize()
            host_runs.append((time.perf_counter() - w0) * 1e3)
        host_ms = min(host_runs)
        value_host = round(H * W * args.steps / (host_ms / 1e3) / 1e6, 3)
        print("with a host loadImage per step (PCIe included): %.3f ms/step (runs: %s)"
              % (host_ms / args.steps, ", ".join("%.3f" % (r / args.steps) for r in host_runs)),
              file=sys.stderr, flush=True)

    # roofline of the dominant kernel: second pass, HIP events per launch
    dominant = "propagate_payload" if args.staged else "fused_frame"
    N.profile_begin(dominant)
    timed(args.warmup + args.steps)
    tot_ms, launches = N.profile_end()
    avg_ms = tot_ms / max(launches, 1)
    alg_bytes = algorithmic_bytes()[dominant]
    achieved = alg_bytes / (avg_ms / 1e3) / 1e9
    print("dominant kernel %s: %d launches, %.4f ms avg, %.1f GB/s" % (dominant, launches, avg_ms, achieved),
          file=sys.stderr, flush=True)

    traffic = None
    tf = os.path.join(ROOT, "profiles", "fused_frame_2048x1088_2lvl_traffic.json")
    if dominant == "fused_frame" and os.path.exists(tf):
        with open(tf) as fh:
            t = json.load(fh)
        if t.get("algorithmic_bytes_per_launch") == algorithmic_bytes()[dominant]:
            traffic = t["traffic_bytes_per_launch"]
    # what actually bounds the fused kernel: VALU issue. Its instruction
    # count per launch (rocprofv3 SQ_INSTS_VALU, profiles/) over the CUs'
    # issue rate (one wave-wide VALU instruction per CU per cycle, 256 CUs
    # at 2.4 GHz) is the floor; frac = floor / this run's launch time.
    valu = None
    vf = os.path.join(ROOT, "profiles", "fused_frame_2048x1088_2lvl_valu.json")
    if dominant == "fused_frame" and os.path.exists(vf):
        with open(vf) as fh:
            v = json.load(fh)
        floor_ms = v["sq_insts_valu_per_launch"] / v["valu_issue_rate_per_s"] * 1e3
        valu = {"instructions_per_launch": v["sq_insts_valu_per_launch"], "issue_floor_ms": round(floor_ms, 5),
                "frac": round(floor_ms / avg_ms, 4), "source": "profiles/fused_frame_2048x1088_2lvl_valu_counters.md"}

    out = {
        "metric": "Mpix/s at 2048x1088 2-lvl; HBM GB/s % of peak; batched/plain speedup",
        "value": round(value, 3),
        "value_with_copy": value_copy,
        "value_host_upload": value_host,
        "host_upload_ms_per_step_runs": None if value_host is None else [round(r / args.steps, 4) for r in host_runs],
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded blob texture, affine motion, sensor noise; %d frames resident in HBM, cycled)" % F,
        "config": {"workload": "PyramidalFlowFilter 2048x1088, 2 levels, gamma [10,50], maxflow 4, "
                               "smooth [2,2], u8 frames, step = compute() of a frame read in place from HBM "
                               "(setImageDevice)",
                   "path": "staged" if args.staged else "fused",
                   "tile": None if args.staged else {"waves": f.tileWaves, "top_rows": f.tileTopRows,
                                                     "rows": f.tileRows, "interior": f.tileInterior},
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg_bytes, "avg_launch_ms": round(avg_ms, 5),
                     "launches": launches, "valu_issue": valu},
    }
    if not args.no_batched:
        out["batched"] = batched_bench(args, N, gf, GPUImage)
    if not args.no_display:
        out["display_rotation"] = display_rotation_bench(args, N, GPUImage)
        out["reference_notebook"] = reference_notebook_bench(args, gf)
    if not args.no_differentiable:
        out["differentiable"] = differentiable_bench(args)
    if not args.no_cpu:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(frames[: min(F, 8)], args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
