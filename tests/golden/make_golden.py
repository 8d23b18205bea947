"""Writes tests/golden/ref_runs.json from the reference's own program.

Runs oracle/_ref/mpi_lsbsort (the reference's mpi/mpi_lsbsort.cpp compiled
unmodified by oracle/Makefile) under mpirun with --print --verify and records
the rows it prints before and after the sort (`A[i] = (key,val)`, first 10
per rank; mpi_lsbsort.cpp:183-212, 748-752) together with its own verify
verdict.  The JSON holds only data: (n, ranks, printed rows, digests).

Run in a container that has the reference and MPI:
    make -C oracle ref && python tests/golden/make_golden.py
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(REPO, "oracle", "_ref", "mpi_lsbsort")
MPIRUN = os.environ.get("MPIRUN", "/opt/conda/bin/mpirun")

CASES = [(0, 4), (1, 4), (3, 4), (40, 4), (1000, 3), (1001, 4), (65537, 4),
         (1 << 20, 4), (1 << 24, 4)]
ROW = re.compile(r"^A\[(\d+)\] = \(([0-9a-f]{16}),(\d+)\)$")


def run(n, ranks):
    cmd = [MPIRUN, "-n", str(ranks), BIN, "--n", str(n), "--print", "--verify"]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        raise RuntimeError(f"{cmd} failed: {p.stdout}\n{p.stderr}")
    blocks, cur = [], None
    for line in p.stdout.splitlines():
        if line.startswith("A: displaying"):
            cur = []
            blocks.append(cur)
            continue
        m = ROW.match(line.strip())
        if m and cur is not None:
            cur.append([int(m.group(1)), int(m.group(2), 16), int(m.group(3))])
    if n > 0 and len(blocks) != 2:
        raise RuntimeError(f"expected 2 printed blocks, got {len(blocks)}")
    # -DNDEBUG compiles out assert(!failures) (mpi_lsbsort.cpp:782), so a
    # mismatch shows only as "did not match" lines (773-780).
    verified = "Verifying" in p.stdout and "did not match" not in p.stdout
    return {"n": n, "ranks": ranks,
            "input_rows": [[i, f"{k:016x}", v] for i, k, v in blocks[0]] if blocks else [],
            "sorted_rows": [[i, f"{k:016x}", v] for i, k, v in blocks[1]] if blocks else [],
            "reference_verify_ran_and_passed": verified}


def main():
    if not os.path.exists(BIN):
        sys.exit(f"missing {BIN}: run make -C oracle ref")
    out = {"program": "mpi/mpi_lsbsort.cpp (reference), unmodified, -DNDEBUG -O3",
           "runs": [run(n, r) for n, r in CASES]}
    with open(os.path.join(HERE, "ref_runs.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "ref_runs.json"))


if __name__ == "__main__":
    main()
