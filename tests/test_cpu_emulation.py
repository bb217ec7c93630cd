"""CPU build of the library's kernel and host sources (tests/cpu_emu).

The sources compile with g++ against an emulated HIP runtime that runs each
workgroup's threads as fibers (barriers, wave shuffles and ballots with
64-lane waves) and workgroups on concurrent OS threads, under
AddressSanitizer and UndefinedBehaviorSanitizer.  Device memory is heap
memory and workgroup arrays carry their static bounds, so an out-of-bounds
access by any kernel, a shuffle that reads a lane outside the operation and
a barrier that not every live thread reaches abort the run.  The scenarios
drive the C ABI (2D tile stage with every tile shape and both tile orders,
constant tiles, slice sweeps, 1D stage variants, FP32, time loops,
sc_step_noalloc, batches, the request server) and compare with the oracle.
"""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_emu")
BUILD = os.path.join(HERE, "_build")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


@pytest.fixture(scope="module")
def built():
    jobs = str(min(8, os.cpu_count() or 1))
    r = subprocess.run(["make", "-s", "-C", HERE, "-j", jobs], capture_output=True, text=True, timeout=1200)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return BUILD


def _run(args, timeout=900):
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=ENV)


def test_scenarios_sanitized(built):
    r = _run([os.path.join(built, "emu_main")])
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "0 failures" in r.stdout
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
    # every scenario ran and passed
    assert r.stdout.count("PASS ") >= 31, r.stdout


@pytest.mark.parametrize("case,expect", [
    ("heap", "heap-buffer-overflow"),
    ("lds", "out of bounds"),
    ("wave", "does not take part"),
])
def test_negative_controls(built, case, expect):
    """Each deliberate fault is reported and aborts the run."""
    r = _run([os.path.join(built, "emu_selftest"), case], timeout=120)
    assert r.returncode != 0
    assert expect in r.stdout + r.stderr


def test_wellformed_kernel(built):
    r = _run([os.path.join(built, "emu_selftest"), "ok"], timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
