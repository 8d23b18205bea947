"""A C++ program calls the library the way the reference's main calls
mySort (tests/cxx/test_mysort.cpp through include/lsb_mysort.hpp): the
reference's keygen, the reference's timer lines, and its verification
against std::stable_sort by key (mpi/mpi_lsbsort.cpp:731-784)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cxx", "test_mysort")
C1_SORTED_SHA = "edbb62f7b024eaade41bc12278e7cb240d038889a9780f8913a336352530d0c9"
EMPTY_SHA = "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"


def run(*cases, timeout=300):
    assert os.path.exists(BIN), f"{BIN} is not built (make -C tests/cxx)"
    p = subprocess.run([BIN, *cases], capture_output=True, text=True, timeout=timeout)
    lines = [l for l in p.stdout.splitlines() if l.startswith("case ")]
    out = {}
    for l in lines:
        m = re.match(r"case n=(\d+) ranks=(\d+) (.*)", l)
        out[(int(m.group(1)), int(m.group(2)))] = dict(
            kv.split("=", 1) for kv in m.group(3).split() if "=" in kv)
    return p.returncode, out, p.stdout


def test_trivial_sizes_need_no_device():
    """n = 0 and 1 return before any device call, as mySort handles them."""
    rc, out, text = run("0:4", "1:4")
    assert rc == 0, text
    assert out[(0, 4)]["verify"] == "ok" and out[(0, 4)]["sha256"] == EMPTY_SHA
    assert out[(1, 4)]["verify"] == "ok"
    assert "Sorted 0 values in" in text and "M elements sorted / s" in text


@pytest.mark.gpu
def test_cxx_caller_reference_cases():
    """The reference's own edge cases (n in {0, 1, 3, 1001, 65537} at 4 ranks,
    1000 at 3) and its 2^24 / 4-rank input, whose sorted image must hash to
    the reference output's digest; every case equals std::stable_sort."""
    cases = ["0:4", "1:4", "3:4", "1001:4", "65537:4", "1000:3", "1048576:256", "16777216:4"]
    rc, out, text = run(*cases)
    assert rc == 0, text
    for c in cases:
        n, r = (int(x) for x in c.split(":"))
        assert out[(n, r)]["verify"] == "ok", (c, text)
    assert out[(1 << 24, 4)]["sha256"] == C1_SORTED_SHA
