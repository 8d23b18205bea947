"""The CPU oracle and the library's host code run under AddressSanitizer and
UndefinedBehaviorSanitizer (built by make -C oracle asan, make -C
distributed-lsb_amd asan, make -C tests/cxx asan; __graft_entry__.build()
runs all three).  The reference checks the same kinds of invariants with
asserts (mpi/mpi_lsbsort.cpp:364, 392, 487-488, 583)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SELFTEST = os.path.join(ROOT, "oracle", "_asan", "oracle_selftest")
HOST_SELFTEST = os.path.join(ROOT, "tests", "cxx", "host_selftest")


# the ROCm runtime's own process-lifetime allocations, not ours
LSAN_SUPP = os.path.join(ROOT, "tests", "cxx", "lsan_rocm.supp")


def _run(cmd, timeout):
    assert os.path.exists(cmd[0]), f"{cmd[0]} is not built (__graft_entry__.build())"
    env = dict(os.environ, LSAN_OPTIONS=f"suppressions={LSAN_SUPP}")
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    return p.stdout


def test_oracle_under_sanitizers():
    assert "passed" in _run([ORACLE_SELFTEST], 300)


def test_host_code_under_sanitizers_cpu():
    """Planning and layout arithmetic for every size class up to LSB_MAX_N,
    argument checks of every entry point, tune-file parsing with malformed
    lines, and the server's start-up failure path without a device."""
    assert "passed" in _run([HOST_SELFTEST], 300)


@pytest.mark.gpu
def test_host_code_under_sanitizers_gpu():
    """The same library with sanitized host code on the GPU: plain sorts that
    grow and shrink the cached workspace (eager, graph, no-allocation),
    batches in both forms, and a serving run fed by two threads under a small
    budget, each result proven with lsb_verify."""
    print("host_selftest --gpu", flush=True)
    assert "passed" in _run([HOST_SELFTEST, "--gpu"], 240)
