"""bench.py on the CPU: the one-GPU contract, the algorithmic byte model of
the fused frame kernel, and the profile files it reads beside its own
measurement (HBM traffic and VALU instruction counts), which must describe
the same workload and launch shape."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_per_frame():
    b = _bench()
    n0, n1 = 2048 * 1088, 1024 * 544
    # delta level 43 B/px, top level 33 B/px, pyramid write 1 B per level-1 px
    assert b.algorithmic_bytes()["fused_frame"] == 43 * n0 + 33 * n1 + n1 == 114753536


def test_profiles_describe_the_benchmark_launch():
    b = _bench()
    with open(os.path.join(ROOT, "profiles", "fused_frame_2048x1088_2lvl_traffic.json")) as fh:
        traffic = json.load(fh)
    with open(os.path.join(ROOT, "profiles", "fused_frame_2048x1088_2lvl_valu.json")) as fh:
        valu = json.load(fh)
    for p in (traffic, valu):
        assert p["kernel"] == "fused_frame"
        assert p["algorithmic_bytes_per_launch"] == b.algorithmic_bytes()["fused_frame"]
        assert "2048x1088, 2 levels" in p["workload"]
    assert traffic["launch_shape"] == valu["launch_shape"]
    # HBM bytes from the two counter passes, each scaled by the copy kernel's factor
    cal = traffic["calibration"]
    read = traffic["fetch_size_kib"] * 1024 * cal["fetch_factor"]
    write = traffic["write_size_kib"] * 1024 * cal["write_factor"]
    assert abs(read - traffic["hbm_read_bytes"]) < 1e-3 * read
    assert abs(write - traffic["hbm_write_bytes"]) < 1e-3 * write
    assert traffic["traffic_bytes_per_launch"] == traffic["hbm_read_bytes"] + traffic["hbm_write_bytes"]
    # waves per launch: workgroups x waves per workgroup
    shape = valu["launch_shape"]
    assert valu["sq_waves_per_launch"] == shape["workgroups"] * shape["waves_per_workgroup"]
