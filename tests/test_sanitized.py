"""The host code of the library (the runtime's schedule and layout code, the settings
builder) under AddressSanitizer and UndefinedBehaviorSanitizer: the CPU tests that exercise it
(settings builder byte equality, the level schedules of every body shape, the C ABI) run
again, in a child process, against libjolt_xpbd_asan.so with the sanitizer runtime preloaded;
any report aborts the child."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

from joltphysics_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_asan_and_ubsan():
    lib, runtime = build.asan_library_path(), build.sanitizer_runtime()
    if not os.path.exists(lib) or runtime is None:
        pytest.skip("sanitized library not built (joltphysics_amd/build.py --sanitized)")
    env = dict(os.environ, JOLT_XPBD_LIBRARY=lib, LD_PRELOAD=runtime,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    tests = [os.path.join(ROOT, "tests", t) for t in ("test_builder.py", "test_schedule.py", "test_capi.py")]
    probe = ("import sys; sys.path.insert(0, %r); from joltphysics_amd import xpbd; xpbd.load_library(); "
             "print(any(%r in l for l in open('/proc/self/maps')))" % (ROOT, os.path.basename(lib)))
    out = subprocess.run([sys.executable, "-c", probe], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("True"), out.stderr[-2000:]
    run = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *tests], env=env,
                         capture_output=True, text=True, timeout=1200, cwd=ROOT)
    assert run.returncode == 0, (run.stdout[-3000:], run.stderr[-3000:])
    assert "passed" in run.stdout
