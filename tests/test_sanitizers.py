"""The library's host code under sanitizers (CPU only): the OBJ parser,
the threaded SAH builder with its four-wide collapse, and the importance
tables, built from their sources with AddressSanitizer + UndefinedBehavior-
Sanitizer and with ThreadSanitizer (tests/sanitize/Makefile) and driven by
tests/sanitize/host_fuzz.cpp.  A sanitizer finding aborts the driver."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


def _build(target):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    r = subprocess.run(["make", "-C", HERE, target], capture_output=True, text=True)
    if r.returncode != 0 and "cannot find -l" in (r.stderr + r.stdout):
        pytest.skip("sanitizer runtime not installed")
    assert r.returncode == 0, r.stderr[-2000:]
    return os.path.join(HERE, "_build", "host_fuzz_" + target)


def _run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-4000:])
    return r.stdout


def test_host_code_under_address_and_undefined_sanitizers():
    out = _run(_build("asan"), "all")
    assert "obj:" in out and "bvh:" in out and "env:" in out


def test_threaded_sah_build_under_thread_sanitizer():
    out = _run(_build("tsan"), "bvh", "quick")
    assert "bvh:" in out
