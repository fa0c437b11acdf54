"""The C ABI boundary on the CPU: the library loads, exports every symbol that
include/flowfilter_amd.h declares, the Python binding declares each one with
the header's arity, the image descriptor has the reference's layout, and
argument errors detected before any device call come back as status codes.
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flowfilter_amd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"FF_API\s+[\w\s\*]+?\b(ff_\w+)\s*\(([^)]*)\)\s*;", text):
        args = m.group(2).strip()
        n = 0 if args in ("", "void") else len(args.split(","))
        out[m.group(1)] = n
    return out


def test_header_parses():
    fns = header_functions()
    assert len(fns) >= 140
    assert "ff_pyramidal_create" in fns and fns["ff_pyramidal_create"] == 4


def test_library_exports_every_declared_symbol():
    from flowfilter import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_python_binding_matches_header():
    from flowfilter import _native
    fns = header_functions()
    declared = dict(_native.SIGNATURES)
    declared["ff_last_error"] = []
    assert set(declared) == set(fns), set(declared) ^ set(fns)
    for name, n in fns.items():
        assert len(declared[name]) == n, name


def test_image_t_layout():
    """ff_image_t is byte-identical to flowfilter::image_t (LP64: 40 bytes)."""
    from flowfilter._native import image_t
    assert ctypes.sizeof(image_t) == 40
    assert image_t.height.offset == 0 and image_t.width.offset == 4 and image_t.depth.offset == 8
    assert image_t.pitch.offset == 16 and image_t.itemSize.offset == 24 and image_t.data.offset == 32


def test_numpy_image_wrapper():
    from flowfilter.image import Image
    a = np.zeros((5, 7, 2), np.float32)
    w = Image(a)
    assert (w.height, w.width, w.depth, w.pitch, w.itemSize) == (5, 7, 2, 56, 4)
    with pytest.raises(ValueError):
        Image(np.zeros((2, 3, 4, 5), np.float32))
    with pytest.raises(ValueError):
        Image(np.zeros((6, 8), np.float32)[:, ::2])


def test_errors_before_device_work():
    """Argument checks that precede any allocation report the reference's exception kinds."""
    from flowfilter import errors
    import flowfilter.gpu.flowfilters as ff
    with pytest.raises(errors.InvalidArgument):
        ff.FlowFilter(0, 10)
    with pytest.raises(errors.InvalidArgument):
        ff.FlowFilter(10, -3)
    from flowfilter import _native as N
    st = N.lib.ff_stage_compute(None)
    assert st == errors.FF_ERR_NULL_HANDLE
    assert b"null" in N.lib.ff_last_error()
    st = N.lib.ff_gpuimage_upload(None, None)
    assert st == errors.FF_ERR_NULL_HANDLE


def test_host_image_create_destroy():
    from flowfilter import _native as N
    img = N.image_t()
    N.call("ff_create_image", 4, 6, 2, 4, ctypes.byref(img))
    assert (img.height, img.width, img.depth, img.pitch, img.itemSize) == (4, 6, 2, 48, 4)
    assert img.data
    N.call("ff_destroy_image", ctypes.byref(img))
    assert not img.data
    from flowfilter import errors
    with pytest.raises(errors.FilterError):
        N.call("ff_create_image", 0, 6, 1, 1, ctypes.byref(img))


def test_version():
    from flowfilter import _native as N
    a, b = ctypes.c_int(), ctypes.c_int()
    N.call("ff_version", ctypes.byref(a), ctypes.byref(b))
    assert a.value >= 1


def _build_cpp_demo(out_dir):
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    pkg = os.path.join(ROOT, "optical-flow-filter_amd")
    exe = os.path.join(str(out_dir), "api_demo")
    subprocess.check_call([gxx, "-std=c++17", "-O2", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "api_demo.cpp"), "-L" + pkg, "-lflowfilter_amd",
                           "-Wl,-rpath," + pkg, "-o", exe])
    return exe


def test_cpp_header_compiles_and_links(tmp_path):
    """include/flowfilter_amd.hpp (the flowfilter::gpu classes) builds against the library."""
    exe = _build_cpp_demo(tmp_path)
    assert os.path.exists(exe)
