"""The C-ABI library: it loads and exports what include/licom_sw.h declares (CPU).

No compute entry point is called here (there is no GPU); host-only entry points
(routing table, argument validation that fails before any device call) are.
"""

from __future__ import annotations

import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "licom_sw.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sw_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from licom4py_gpu_amd import _hip
    return _hip.lib()


def test_header_and_binding_agree():
    from licom4py_gpu_amd import _hip
    assert declared_functions() == sorted(_hip.EXPORTED)


def test_library_exports_every_declared_symbol(lib):
    from licom4py_gpu_amd import build
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", build.HIP_SO], capture_output=True, text=True)
    if out.returncode == 0:
        exported = set(re.findall(r"\sT\s+(\w+)$", out.stdout, flags=re.M))
        assert set(declared_functions()) <= exported


def test_routing_table_known_answer(lib):
    from licom4py_gpu_amd import _hip
    with open(os.path.join(ROOT, "tests", "golden", "routing.json")) as f:
        want = np.array(json.load(f), np.int32)
    assert np.array_equal(_hip.routing_table(), want)


def test_null_arguments_are_rejected(lib):
    ctx = ctypes.c_void_p()
    assert lib.sw_create(ctypes.byref(ctx), 48, 0, None, None) == -1
    assert b"null" in lib.sw_last_error()
    assert lib.sw_create(None, 48, 0, None, None) == -1
    assert lib.sw_sync(None) == -1
    assert lib.sw_rk2(None, None, None, 1, 1.0, 0) == -1
    assert lib.sw_routing_table(None) == -1


def test_bad_sizes_are_rejected(lib):
    from licom4py_gpu_amd import _hip
    ctx = ctypes.c_void_p()
    arr, _keep = _hip.ptr_array([1] * 17)
    assert lib.sw_create(ctypes.byref(ctx), 3, 0, arr, None) == -1
    assert b"nx" in lib.sw_last_error()
    assert lib.sw_create(ctypes.byref(ctx), 48, 7, arr, None) == -1
    assert b"dtype" in lib.sw_last_error()


def test_bench_refuses_more_than_one_gpu():
    import sys
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 2
    assert "one GPU" in out.stderr


def test_build_flags_kernels_with_scratch():
    from licom4py_gpu_amd import build
    rep = ("x.hip:1:1: remark: Function Name: k_one [-Rpass-analysis=kernel-resource-usage]\n"
           "x.hip:1:1: remark:     ScratchSize [bytes/lane]: 0 [-Rpass-analysis=kernel-resource-usage]\n"
           "x.hip:2:1: remark: Function Name: k_two [-Rpass-analysis=kernel-resource-usage]\n"
           "x.hip:2:1: remark:     ScratchSize [bytes/lane]: 408 [-Rpass-analysis=kernel-resource-usage]\n")
    assert build.scratch_kernels(rep) == [("k_two", 408)]
