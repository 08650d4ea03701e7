"""The C ABI library: loads, exports every declared symbol, reports errors
without a device, and the bench refuses more than one GPU."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess
import sys

import pytest

from helpers import ROOT_DIR
from joltphysics_amd import xpbd


def _declared_functions():
    text = open(os.path.join(ROOT_DIR, "include", "jolt_xpbd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xpbd_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = xpbd.load_library()
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert set(xpbd.EXPORTED_SYMBOLS) <= set(names)
    # every declared entry point with arguments has its ctypes signature (an undeclared one
    # would pass 64-bit pointers as C ints)
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT_DIR, "include", "jolt_xpbd.h")).read(), flags=re.S)
    no_sig = [n for n in names if not re.search(n + r"\s*\(\s*void\s*\)", text) and getattr(lib, n).argtypes is None]
    assert no_sig == []


def test_context_create_without_device_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = xpbd.load_library()
    h = ctypes.c_void_p()
    rc = lib.xpbd_context_create(0, ctypes.byref(h))
    assert rc != 0 and not h.value
    assert len(lib.xpbd_last_error()) > 0


def test_null_arguments_are_rejected():
    lib = xpbd.load_library()
    assert lib.xpbd_update(None, ctypes.c_float(0.016), 1, 1) == -1
    assert lib.xpbd_synchronize(None) == -1
    assert lib.xpbd_settings_release(None) == -1
    assert lib.xpbd_context_destroy(None) == -1
    h = ctypes.c_void_p()
    assert lib.xpbd_rollout_create(None, None, ctypes.byref(h)) == -1 and not h.value
    assert lib.xpbd_rollout_forward(None, None, None, None, None) == -1
    assert lib.xpbd_rollout_forward_f64(None, None, None, None, None) == -1
    assert lib.xpbd_rollout_backward(None, None, None, None, None) == -1
    assert lib.xpbd_rollout_destroy(None) == -1


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT_DIR, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
