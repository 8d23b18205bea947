"""Benchmark: one GPU's share at the reference README's scale.

Workload: 2^30 16-byte {uint64 key uniform over [0, 2^64), uint64 val =
index} pairs (16 GiB in A + 16 GiB ping-pong B), keys from pcg64 over 256
virtual ranks exactly like the reference keygen (mpi/mpi_lsbsort.cpp:
676-704), sorted by the full 64-bit key, stable.  A step is one lsb_sort of
that array in the library's automatic form (most significant digit first at
this size and key width; --form picks one): the input is restored from a
pristine device copy before every step (untimed), then the sort is timed
between two device synchronisations.  Metric: Melem/s sorted = n / t_sort /
1e6 (mpi/mpi_lsbsort.cpp:741).

roofline: the dominant kernel of the timed form, the launch with the largest
average duration, read from HIP events recorded on the sort's stream between
its launches (lsb_msd_profile_read / lsb_profile_read), averaged over the
timed steps, against its algorithmic bytes: 32 B per element it moves (16 B
read + 16 B written) for a scatter, a partition or the in-LDS finish, 8 B per
element for a histogram (the key).  The whole sort is also reported against
the sum of its launches' algorithmic bytes.  The LSD form (the reference's
pass order) is timed on the same input beside it: its scatter pass's
roofline and its throughput are under lsd_form.

cpu_baseline: the reference's own mpi_lsbsort (oracle/_ref, built from its
sources) under mpirun on this host's cores at a bounded sample size, before
the GPU is touched; if it cannot run, the oracle's single-thread C LSD sort
(a port) is timed instead.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "distributed-lsb_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "Melem/s sorted, 2^30 16B pairs, % HBM roofline; served p50/p99 vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    """Host CPU and memory facts for the baseline legs (lscpu, /proc/meminfo)."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        f = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        get = lambda k: f.get(k, "").strip()
        info["model"] = get("Model name")
        sockets = int(get("Socket(s)") or 0)
        cores = int(get("Core(s) per socket") or 0)
        info.update({"sockets": sockets, "cores_per_socket": cores,
                     "threads_per_core": int(get("Thread(s) per core") or 0),
                     "physical_cores": sockets * cores})
    except Exception as e:
        log("lscpu failed:", e)
    try:
        mem = dict(line.split(":", 1) for line in open("/proc/meminfo"))
        info["mem_available_gib"] = round(int(mem["MemAvailable"].split()[0]) / 2 ** 20, 1)
    except Exception:
        pass
    return info


# At most this many ranks: a GPU job's share of the host's CPUs on the
# pooled boxes this runs on (the rest belong to other jobs), whatever nproc
# says.
CPU_SHARE = 16


def _ref_leg(mpirun, ref, n, ranks, reps):
    rates, secs = [], []
    for r in range(reps):
        log(f"  reference mpi_lsbsort: n=2^{n.bit_length() - 1}, {ranks} ranks, rep {r + 1}/{reps}")
        p = subprocess.run([mpirun, "-n", str(ranks), "-bind-to", "core", ref, "--n", str(n),
                            "--no-verify"], capture_output=True, text=True, timeout=300)
        m = re.search(r"That's ([0-9.eE+]+) M elements sorted / s", p.stdout)
        t = re.search(r"Sorted \d+ values in ([0-9.eE+]+)", p.stdout)
        if p.returncode != 0 or not m:
            raise RuntimeError(p.stdout[-500:] + p.stderr[-500:])
        rates.append(float(m.group(1)))
        if t:
            secs.append(float(t.group(1)))
    rates.sort()
    secs.sort()
    return {"n": n, "ranks": ranks, "reps": reps, "melem_s": rates[len(rates) // 2],
            "sort_s": secs[len(secs) // 2] if secs else None, "all_melem_s": rates}


def cpu_baseline(reps=5, full=False):
    """The reference's own mpi_lsbsort (oracle/_ref, built unmodified from its
    sources with -DNDEBUG -O3) under mpirun -bind-to core on this host, before
    the GPU is touched: 4 ranks at 2^24 (the oracle config), all usable
    physical cores at 2^24 and at 2^28 (when 24 GiB of RAM are free; 2^30 with
    --cpu-full and 96 GiB free), median of `reps` each.  `value` is the
    largest leg.  If the reference cannot run, the oracle's single-thread C
    LSD sort (a port) is timed instead."""
    ref = os.path.join(ROOT, "oracle", "_ref", "mpi_lsbsort")
    mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
    info = host_info()
    phys = info.get("physical_cores") or info["affinity_cpus"]
    ranks = max(1, min(phys, info["affinity_cpus"], CPU_SHARE))
    if os.path.exists(ref) and os.path.exists(mpirun):
        try:
            probe = subprocess.run([mpirun, "-n", "2", ref, "--n", "1000", "--no-verify"],
                                   capture_output=True, text=True, timeout=60)
            if probe.returncode != 0:
                raise RuntimeError("mpirun probe failed: " + probe.stderr[-300:])
            legs = [_ref_leg(mpirun, ref, 1 << 24, 4, reps),
                    _ref_leg(mpirun, ref, 1 << 24, ranks, reps)]
            free = info.get("mem_available_gib", 0)
            if free >= 24:
                legs.append(_ref_leg(mpirun, ref, 1 << 28, ranks, reps))
            if full and free >= 96:
                legs.append(_ref_leg(mpirun, ref, 1 << 30, ranks, 3))
            top = legs[-1]
            return {"value": top["melem_s"], "unit": "Melem/s", "cores": top["ranks"],
                    "kind": "reference",
                    "sample": f"mpi_lsbsort --n 2^{top['n'].bit_length() - 1} --no-verify, "
                              f"mpirun -n {top['ranks']} -bind-to core, median of "
                              f"{top['reps']} (reference keygen pcg64(rank), 64-bit keys); "
                              f"ranks = min(physical cores, usable CPUs, {CPU_SHARE}-CPU share)",
                    "legs": legs, "host": info}
        except Exception as e:  # fall through to the port
            log("reference CPU baseline failed:", e)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = 1 << 24
    a = O.gen_uniform(n, ranks=4)
    t = time.perf_counter()
    O.lsd_sort(a, 64, 16)
    dt = time.perf_counter() - t
    return {"value": n / dt / 1e6, "unit": "Melem/s", "cores": 1, "kind": "port",
            "sample": f"oracle C LSD sort (16-bit digits, one thread), n={n}", "host": info}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_onesweep_rg_2p30_uniform_traffic.md")
MSD_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_msd_2p30_uniform_traffic.md")


# The sources each form's kernels are compiled from, and the build file.
PLAIN_SORT_SOURCES = ["csrc/lsb_kernels.hip", "csrc/lsb_onesweep.hpp", "csrc/lsb_device.hpp",
                      "csrc/lsb_internal.hpp", "Makefile"]
MSD_SOURCES = ["csrc/lsb_msd.hip", "csrc/lsb_onesweep.hpp", "csrc/lsb_device.hpp",
               "csrc/lsb_internal.hpp", "Makefile"]


def source_digest(sources):
    """SHA-256 over kernel sources and the build file: a committed counter
    summary is used only for the code it was taken on."""
    import hashlib
    h = hashlib.sha256()
    pkg = os.path.join(ROOT, "distributed-lsb_amd")
    for rel in sources:
        h.update(rel.encode())
        h.update(open(os.path.join(pkg, rel), "rb").read())
    return h.hexdigest()[:16]


def traffic_from_profile(path, log2n, sources, kernel, config=None):
    """HBM bytes per launch of `kernel` from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary, when it was taken on this workload and
    these exact sources; else None."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch\[" + re.escape(kernel) + r"\]: (\d+)$", text, re.M)
    c = re.search(r"^traffic_config: (.*)$", text, re.M)
    h = re.search(r"^traffic_sources: ([0-9a-f]+)$", text, re.M)
    want = config or f"log2n={log2n}"
    if not m or not c or c.group(1).strip() != want:
        return None
    if not h or h.group(1) != source_digest(sources):
        log(f"{os.path.basename(path)} was taken on other sources: traffic reported as null")
        return None
    return int(m.group(1))


def lsd_summary(n, wall, profs, d, tile, P, log2n):
    """The LSD form's figures from its profiled steps: its dominant kernel is
    the onesweep scatter pass, 32 B per element per executed pass (16 B read
    + 16 B written); the whole sort moves n * (8 + 32 * passes executed)."""
    ms_wall = sum(wall) / len(wall) * 1e3
    p_exec = sum(profs[-1]["executed"])
    pass_ms = [t for p in profs for t, e in zip(p["pass_ms"], p["executed"]) if e]
    avg_pass_ms = sum(pass_ms) / len(pass_ms) if pass_ms else float("nan")
    hist_ms = sum(p["histogram_ms"] for p in profs) / len(profs)
    achieved = 32 * n / (avg_pass_ms * 1e-3) / 1e9
    model_bytes = n * (8 + 32 * p_exec)
    sort_gbs = model_bytes / (ms_wall * 1e-3) / 1e9
    traffic = traffic_from_profile(TRAFFIC_PROFILE, log2n, PLAIN_SORT_SOURCES, "k_onesweep_rg",
                                   f"log2n={log2n} digit_bits={d} tile={tile}")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": f"k_onesweep_rg scatter pass (d={d}, tile={tile}), 32 B/elem x 2^{log2n}",
                "avg_launch_ms": round(avg_pass_ms, 4)}
    block = {"melem_s": round(n / (ms_wall * 1e-3) / 1e6, 2), "ms_per_step": round(ms_wall, 4),
             "steps": len(wall), "digit_bits": d, "tile": tile, "passes": P,
             "passes_executed": p_exec, "histogram_ms": round(hist_ms, 4), "roofline": roofline,
             "sort_roofline": {"model_bytes": model_bytes, "achieved": round(sort_gbs, 1),
                               "frac": round(sort_gbs / HBM_PEAK_GBS, 4)},
             "what": "lsb_sort forced to the LSD form (the reference's pass order), same input"}
    return {"avg_pass_ms": avg_pass_ms, "model_bytes": model_bytes, "roofline": roofline,
            "passes_executed": p_exec, "block": block}


MSD_KERNELS = {"w": ("k_msd_whist", "k_msd_wplan", "k_msd_wscatter"),
               "l": ("k_msd_hist", "k_msd_plan", "k_msd_scatter (or _rg)")}


def msd_launches(profs):
    """Every launch of the MSD form that moved elements, averaged over the
    profiled steps, with its algorithmic bytes: a scatter or the finish 32 B
    per element it moved (16 B read + 16 B written), a histogram 8 B per
    element of its step (the key read; on uniform keys every range it counts
    is partitioned), a plan none."""
    K, S = profs[-1]["levels"], profs[-1]["wide_steps"]
    steps = [("wide step 0", "w", 0)]
    for k in range(K):
        steps.append((f"level {k}", "l", k))
        if k + 1 < S:
            steps.append((f"wide step {k + 1}", "w", k + 1))
    mean = lambda xs: sum(xs) / len(xs)
    out = []
    for label, kind, i in steps:
        elems = profs[-1]["wide_elems" if kind == "w" else "level_elems"][i]
        if not elems:
            continue
        for j, bpe in enumerate((8, 0, 32)):
            ms = mean([p["wide_ms" if kind == "w" else "level_ms"][i][j] for p in profs])
            b = bpe * elems
            out.append({"kernel": MSD_KERNELS[kind][j], "step": label, "ms": round(ms, 4),
                        "elems": elems, "bytes_per_elem": bpe, "bytes": b,
                        "gbs": round(b / (ms * 1e-3) / 1e9, 1) if b and ms > 0 else None})
    ms = mean([p["finish_ms"] for p in profs])
    fe = profs[-1]["finish_elems"]
    out.append({"kernel": "k_msd_finish", "step": "finish", "ms": round(ms, 4), "elems": fe,
                "bytes_per_elem": 32, "bytes": 32 * fe,
                "gbs": round(32 * fe / (ms * 1e-3) / 1e9, 1) if ms > 0 else None})
    return out


def _counter_rows(csv_dir):
    """(kernel name, dispatch id) -> counter value of every rocprofv3
    counter_collection.csv under csv_dir, summed per dispatch."""
    import csv
    per = {}
    for root, _, files in os.walk(csv_dir):
        for f in files:
            if not f.endswith("counter_collection.csv"):
                continue
            for row in csv.DictReader(open(os.path.join(root, f))):
                key = (row["Kernel_Name"], int(row["Dispatch_Id"]))
                per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    return per


def traffic_summary(fetch_dir, write_dir, log2n, d, tile, cmd_fetch, cmd_write):
    """Writes both traffic profiles from two rocprofv3 --pmc runs of
    `bench.py --traffic-probe` (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch), which sorts the array once in the MSD form and twice in the
    LSD form.  Scale factors come from the probe's plain device copy of the
    same 2^log2n-element array (16 B read + 16 B written per element)."""
    fetch, write = _counter_rows(fetch_dir), _counter_rows(write_dir)
    n = 1 << log2n
    kib = 16 * n / 1024.0

    def base(name):  # "void lsb::k_msd_scatter<1024, 8>(...)" -> "k_msd_scatter"
        return re.split(r"[<(]", name)[0].split("::")[-1].split(" ")[-1]

    def pick(per, pat):
        return [(i, v) for (k, i), v in sorted(per.items(), key=lambda kv: kv[0][1])
                if (pat in k if pat == "copyBuffer" else base(k) == pat)]

    def copy_of(per):
        vals = [v for _, v in pick(per, "copyBuffer") if v > 0.5 * kib]
        return vals[-1] if vals else None

    cf, cw = copy_of(fetch), copy_of(write)
    sf, sw = kib / cf, kib / cw
    head = ["Commands (two passes, one counter each; repo root = /root/repo):", "",
            "    " + cmd_fetch, "    " + cmd_write, "",
            "Summary written by `python3 bench.py --traffic-summary <fetch dir> <write dir>`.", "",
            f"Calibration on the plain copy (`copyBuffer`, 2^{log2n} x 16 B = {kib:.0f} KiB read "
            f"and written; raw FETCH_SIZE {cf:.0f} KiB, WRITE_SIZE {cw:.0f} KiB): FETCH_SIZE scale "
            f"{sf:.3f}, WRITE_SIZE scale {sw:.3f}.", ""]

    # LSD form: the onesweep scatter pass, mean over its dispatches
    pf, pw = [v for _, v in pick(fetch, "k_onesweep_rg")], [v for _, v in pick(write, "k_onesweep_rg")]
    hf = [v for _, v in pick(fetch, "k_histogram")]
    rd = sf * sum(pf) / len(pf) * 1024
    wr = sw * sum(pw) / len(pw) * 1024
    alg = 16 * n
    lines = [f"# HBM traffic per launch, LSD form, 2^{log2n} uniform 16-byte pairs, d = {d}, "
             f"tile {tile}", ""] + head + [
        "Raw counters, KiB per dispatch:", "",
        "| kernel | dispatches | FETCH_SIZE | WRITE_SIZE |",
        "|---|---|---|---|",
        f"| `k_histogram` | {len(hf)} | {' '.join(f'{v:.0f}' for v in hf)} | - |",
        f"| `k_onesweep_rg` | {len(pf)} | {' '.join(f'{v:.0f}' for v in pf)} | "
        f"{' '.join(f'{v:.0f}' for v in pw)} |",
        "",
        "Scatter pass, mean over its dispatches, scaled:",
        "",
        f"- read  = {rd / 1e9:.3f} GB (algorithmic {alg / 1e9:.3f} GB, {100 * (rd / alg - 1):+.1f} %)",
        f"- write = {wr / 1e9:.3f} GB (algorithmic {alg / 1e9:.3f} GB, {100 * (wr / alg - 1):+.1f} %)",
        f"- traffic per launch = {(rd + wr) / 1e9:.3f} GB = {(rd + wr) / n:.2f} B per element "
        "(algorithmic 32 B)",
        "",
        f"traffic_bytes_per_launch[k_onesweep_rg]: {int(round(rd + wr))}",
        f"traffic_config: log2n={log2n} digit_bits={d} tile={tile}",
        f"traffic_sources: {source_digest(PLAIN_SORT_SOURCES)}",
        "",
    ]
    open(TRAFFIC_PROFILE, "w").write("\n".join(lines))
    print("\n".join(lines))

    # MSD form: per kernel, its largest dispatch (the one step that moves the
    # array: level 0's partition, in the region form at this size; wide step
    # 1's; the finish)
    rows, keys = [], []
    for k, alg_r, alg_w in (("k_msd_hist", 8, 0), ("k_msd_scatter_rg", 16, 16),
                            ("k_msd_scatter", 16, 16),
                            ("k_msd_whist", 8, 0), ("k_msd_wscatter", 16, 16),
                            ("k_msd_finish", 16, 16)):
        fd = pick(fetch, k)
        wd = dict(pick(write, k))
        if not fd:
            continue
        i, v = max(fd, key=lambda x: x[1])
        if v < 0.01 * kib:  # no dispatch of this kernel moved the array
            continue
        rb = sf * v * 1024
        wb = sw * wd.get(i, 0.0) * 1024
        ar, aw = alg_r * n, alg_w * n
        rows.append(f"| `{k}` | {len(fd)} | {' '.join(f'{x:.0f}' for _, x in fd)} | "
                    f"{rb / 1e9:.3f} ({100 * (rb / ar - 1):+.1f} %) | "
                    + (f"{wb / 1e9:.3f} ({100 * (wb / aw - 1):+.1f} %)" if aw else f"{wb / 1e9:.3f}")
                    + f" | {(rb + wb) / n:.2f} ({alg_r + alg_w}) |")
        keys.append(f"traffic_bytes_per_launch[{k}]: {int(round(rb + wb))}")
    lines = [f"# HBM traffic per launch, MSD form, 2^{log2n} uniform 16-byte pairs", ""] + head + [
        "Per kernel, its largest dispatch in one sort (FETCH_SIZE KiB of every dispatch listed), "
        "scaled; GB and the excess over the algorithmic bytes:", "",
        "| kernel | dispatches | FETCH_SIZE per dispatch | read GB | write GB | B/elem (alg.) |",
        "|---|---|---|---|---|---|"] + rows + [""] + keys + [
        f"traffic_config: log2n={log2n}",
        f"traffic_sources: {source_digest(MSD_SOURCES)}",
        "",
    ]
    open(MSD_TRAFFIC_PROFILE, "w").write("\n".join(lines))
    print("\n".join(lines))


def serving_classes(jobs=500, seed=77, span_log2=12, reps=3):
    """The serving mix split by class (key width x uniform/Zipf): each class's
    jobs sorted as one batch with lsb_sort_batch (graph-replayed, as the
    server launches them), once in the LSD segmented form and once most
    significant digit first; median of `reps` after one warm-up; the two
    forms' outputs compared bit for bit.  Prints a markdown table."""
    import torch
    import lsb_amd
    from lsb_amd import workloads as W

    trace = W.serving_trace(jobs, seed, lo_log2=12, span_log2=span_log2)
    dev = torch.device("cuda:0")
    classes = {}
    for j, job in enumerate(trace):
        classes.setdefault((job["key_bits"], job["zipf"]), []).append((j, job))
    print("| key bits | keys | jobs | elements | LSD form ms | Gelem/s | MSD form ms | Gelem/s "
          "| same output |")
    print("|---|---|---|---|---|---|---|---|---|")
    for (kb, z), jl in sorted(classes.items()):
        As, Bs, pris, ns = [], [], [], []
        for j, job in jl:
            a = torch.empty((job["n"], 2), dtype=torch.int64, device=dev)
            lsb_amd.generate(a, kind=lsb_amd.GEN_ZIPF if z else lsb_amd.GEN_UNIFORM, ranks=4,
                             key_bits=kb, seed_base=7000 + 16 * j, val_base=0)
            As.append(a)
            pris.append(a.clone())
            ns.append(job["n"])
            Bs.append(torch.empty_like(a) if job["n"] > 8192 else None)
        total = sum(ns)
        row, outs = [], {}
        for lsd in (True, False):
            ts = []
            for _ in range(reps + 1):
                for a, p in zip(As, pris):
                    a.copy_(p)
                ts.append(_timed(lambda: lsb_amd.sort_batch(As, Bs, ns, key_bits=kb, graph=True,
                                                            lsd=lsd)))
            t = sorted(ts[1:])[len(ts[1:]) // 2]
            row += [f"{t * 1e3:.3f}", f"{total / t / 1e9:.2f}"]
            outs[lsd] = [a.clone() for a in As]
        same = all(torch.equal(x, y) for x, y in zip(outs[True], outs[False]))
        print(f"| {kb} | {'Zipf(1.1)' if z else 'uniform'} | {len(jl)} | {total} | "
              + " | ".join(row) + f" | {'yes' if same else 'NO'} |", flush=True)
        del As, Bs, pris, outs
        torch.cuda.empty_cache()


def variant_table(log2ns=(24, 26, 28, 30), reps=3):
    """Every compiled plain-sort variant timed (median of `reps` after one
    warm-up, HIP events around the whole lsb_sort on its stream) on uniform
    keys over 64 virtual ranks, as a markdown table with the time per
    executed pass."""
    import torch
    import lsb_amd

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    vs = lsb_amd.variants()
    forms = {0: "one tile per block", 1: "persistent, one chain", 3: "persistent, 8 regions"}
    print("| log2 n | key bits | digit | tile | form | ms | ms per pass | Melem/s |")
    print("|---|---|---|---|---|---|---|---|")
    for log2n in log2ns:
        n = 1 << log2n
        pristine = torch.empty((n, 2), dtype=torch.int64, device=dev)
        lsb_amd.generate(pristine, kind=lsb_amd.GEN_UNIFORM, ranks=64)
        A = torch.empty_like(pristine)
        B = torch.empty_like(pristine)
        stream = torch.cuda.current_stream()
        for kb in (32, 64):
            for v, d, tile, form in vs:
                ts = []
                for r in range(reps + 1):
                    A.copy_(pristine)
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    lsb_amd.sort(A, B, key_bits=kb, variant=v, stream=stream)
                    e1.record(stream)
                    e1.synchronize()
                    if r:
                        ts.append(e0.elapsed_time(e1))
                ms = sorted(ts)[len(ts) // 2]
                passes = (kb + d - 1) // d
                print(f"| {log2n} | {kb} | {d} | {tile} | {forms.get(form, form)} | {ms:.3f} | "
                      f"{ms / passes:.3f} | {n / ms / 1e3:.0f} |", flush=True)
        del pristine, A, B
        torch.cuda.empty_cache()
    # The write pattern a pass with 2^d digits lays down, without the
    # sorting: 8192-element tiles read in order, each written as R runs of
    # 8192 / R consecutive elements (R = 2^d for uniform keys).
    n = 1 << log2ns[-1]
    src = torch.empty((n, 2), dtype=torch.int64, device=dev)
    lsb_amd.generate(src, kind=lsb_amd.GEN_UNIFORM, ranks=64)
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    print()
    print(f"| write runs per 8192-element tile (2^{log2ns[-1]} elements) | elements per run | "
          "ms | GB/s (32 B per element) |")
    print("|---|---|---|---|")
    for runs in (1, 256, 512, 1024, 2048, 4096):
        ts = []
        for r in range(reps + 1):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            lsb_amd.debug_scatter_floor(src, dst, runs=runs)
            e1.record(stream)
            e1.synchronize()
            if r:
                ts.append(e0.elapsed_time(e1))
        ms = sorted(ts)[len(ts) // 2]
        print(f"| {runs} | {8192 // runs} | {ms:.3f} | {32 * n / (ms * 1e-3) / 1e9:.0f} |",
              flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=30)
    ap.add_argument("--digit-bits", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--form", choices=("auto", "lsd", "msd"), default="auto",
                    help="sort form of the timed steps (auto: the library's choice)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the reference at 2^30 (needs 96 GiB of free RAM)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only the timed sorts and one plain copy of the array, for "
                         "rocprofv3 --pmc runs")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write the traffic profile from two rocprofv3 --pmc output dirs")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-segments", type=int, default=4096)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-jobs", type=int, default=500)
    ap.add_argument("--serving-wide-jobs", type=int, default=96)
    ap.add_argument("--serving-streams", type=int, default=0,
                    help="device streams of the server (0: its default, 8); 1 runs the "
                         "coalesced launches one after another, for per-kernel profiles")
    ap.add_argument("--serving-classes", action="store_true",
                    help="only time each class of the 2^12..2^24 serving mix (key width x "
                         "uniform/Zipf) as one batch in both batched forms and print a "
                         "markdown table")
    ap.add_argument("--variant-table", action="store_true",
                    help="only time every compiled plain-sort variant at 2^24..2^30 for 32- "
                         "and 64-bit keys and print a markdown table")
    args = ap.parse_args()

    if args.traffic_summary:
        probe = ("cd /tmp && export TMPDIR=/tmp && timeout -s KILL 120 rocprofv3 --pmc {c} "
                 "--output-format csv -d /root/repo/gpurun_out/pmc_{l} -o {l} -- python3 "
                 "/root/repo/bench.py --traffic-probe --steps 1 --warmup 0")
        d, tile = args.digit_bits or 8, args.tile or 8192
        traffic_summary(args.traffic_summary[0], args.traffic_summary[1], args.log2n, d, tile,
                        probe.format(c="FETCH_SIZE", l="fetch"),
                        probe.format(c="WRITE_SIZE", l="write"))
        return
    if args.traffic_probe:
        args.no_cpu_baseline = args.no_verify = args.no_batched = args.no_serving = True

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the sort is one rank's share on one MI355X "
              "(the multi-GPU exchange is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.variant_table:
        variant_table()
        return
    if args.serving_classes:
        serving_classes(args.serving_jobs)
        return

    base = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        base = cpu_baseline(args.cpu_reps, args.cpu_full)
        log("cpu baseline:", base)

    import torch
    import lsb_amd

    n = 1 << args.log2n
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    log(f"generating n={n} on {torch.cuda.get_device_name(0)} ...")
    pristine = torch.empty((n, 2), dtype=torch.int64, device=dev)
    lsb_amd.generate(pristine, kind=lsb_amd.GEN_UNIFORM, ranks=256)
    A = torch.empty_like(pristine)
    B = torch.empty_like(pristine)
    lsd = {"auto": None, "lsd": True, "msd": False}[args.form]
    form = lsb_amd.sort_form(n, digit_bits=args.digit_bits, tile=args.tile, lsd=lsd)
    d, tile, P = lsb_amd.plan(n, digit_bits=args.digit_bits, tile=args.tile)
    torch.cuda.synchronize()

    def run_form(lsd_flag, steps, warmup):
        """Times `steps` sorts (after `warmup` untimed ones) in one form, each
        profiled (HIP events between its launches), input restored untimed."""
        read = lsb_amd.profile_read if lsd_flag else lsb_amd.msd_profile_read

        def step(profile):
            A.copy_(pristine)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            lsb_amd.sort(A, B, digit_bits=args.digit_bits, tile=args.tile, profile=profile,
                         lsd=lsd_flag)
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        for i in range(warmup):
            step(False)
        wall, profs = [], []
        for i in range(steps):
            wall.append(step(True))
            profs.append(read())
        return wall, profs

    log(f"form {form}: warmup {args.warmup}, timing {args.steps} steps ...")
    wall, profs = run_form(form == "lsd", args.steps, args.warmup)
    ms_wall = sum(wall) / len(wall) * 1e3
    ms_event = sum(p["total_ms"] for p in profs) / len(profs)

    verified = None
    if not args.no_verify:
        keys = torch.empty(n, dtype=torch.int64, device=dev)
        lsb_amd.extract_keys(pristine, keys)
        bad, first = lsb_amd.verify(A, keys)
        verified = bad == 0
        del keys
        if not verified:
            log(f"VERIFY FAILED: {bad} bad positions, first {first}")

    # The LSD form (the reference's pass order) on the same input: its
    # scatter pass is the kernel the LSD roofline is quoted on.
    if form == "lsd":
        lsd_wall, lsd_profs = wall, profs
    else:
        log("LSD form for comparison ...")
        lsd_wall, lsd_profs = run_form(True, args.steps, 1)
    lsd = lsd_summary(n, lsd_wall, lsd_profs, d, tile, P, args.log2n)

    # Context for the roofline: the HIP runtime's own device-to-device copy of
    # the same array (16 B read + 16 B written per element, as one scatter
    # pass), timed with events on the stream it runs on; mean of 3.
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    copy_ms = []
    for i in range(4):
        e0.record(stream)
        A.copy_(pristine)
        e1.record(stream)
        torch.cuda.synchronize()
        if i:
            copy_ms.append(e0.elapsed_time(e1))
    copy_avg = sum(copy_ms) / len(copy_ms)
    copy_gbs = 32 * n / (copy_avg * 1e-3) / 1e9

    # The pass's access pattern without its sorting (lsb_debug_scatter_floor):
    # 8192-element tiles read in order, written as R runs into R regions
    # (R = 256: a uniform pass's layout; R = 1: a plain copy).  Mean of 3
    # after one warm-up, HIP events on the stream the probe runs on.
    floor = None
    if n % 8192 == 0 and not args.traffic_probe:
        floor = {}
        for runs in (1, 16, 64, 256):
            ms = []
            for i in range(4):
                e0.record(stream)
                lsb_amd.debug_scatter_floor(pristine, A, runs=runs)
                e1.record(stream)
                torch.cuda.synchronize()
                if i:
                    ms.append(e0.elapsed_time(e1))
            m = sum(ms) / len(ms)
            floor[f"runs_{runs}"] = {"ms": round(m, 4),
                                     "gbs": round(32 * n / (m * 1e-3) / 1e9, 1)}
        floor["lsd_pass_over_floor_256"] = round(lsd["avg_pass_ms"] / floor["runs_256"]["ms"], 4)
        floor["what"] = ("scatter pass memory traffic alone: 8192-element tiles read in order, "
                         "written as R runs of consecutive elements into R regions")

    if form == "msd":
        launches = msd_launches(profs)
        dom = max(launches, key=lambda k: k["ms"])
        model_bytes = sum(k["bytes"] for k in launches)
        traffic = traffic_from_profile(MSD_TRAFFIC_PROFILE, args.log2n, MSD_SOURCES,
                                       dom["kernel"])
        roofline = {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": f"{dom['kernel']} ({dom['step']}), {dom['bytes_per_elem']} B/elem x "
                              f"{dom['elems']} elements",
                    "avg_launch_ms": dom["ms"]}
        workload = (f"plain sort (MSD-first form), 2^{args.log2n} 16-byte pairs, uniform 64-bit "
                    "keys, val = index, stable, HBM-resident")
        cfg_extra = {"form": "msd", "levels": profs[-1]["levels"],
                     "msd_launches": [{k: v for k, v in x.items() if k != "bytes"}
                                      for x in launches]}
    else:
        model_bytes = lsd["model_bytes"]
        roofline = dict(lsd["roofline"])
        workload = (f"plain LSD sort, 2^{args.log2n} 16-byte pairs, uniform 64-bit keys, "
                    "val = index, stable, HBM-resident")
        cfg_extra = {"form": "lsd", "digit_bits": d, "tile": tile, "passes": P,
                     "passes_executed": lsd["passes_executed"]}
    sort_gbs = model_bytes / (ms_wall * 1e-3) / 1e9

    value = n / (ms_wall * 1e-3) / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (pcg64 keys over 256 virtual ranks, reference keygen)",
        "config": dict({"workload": workload, "n": n}, **cfg_extra),
        "roofline": roofline,
        "sort_roofline": {"model_bytes": model_bytes, "achieved": round(sort_gbs, 1),
                          "frac": round(sort_gbs / HBM_PEAK_GBS, 4),
                          "ms_event": round(ms_event, 4),
                          "what": "algorithmic bytes of every launch of the sort over its "
                                  "wall time"},
        "lsd_form": lsd["block"],
        "scatter_floor": floor,
        "copy_ceiling": {"achieved": round(copy_gbs, 1), "unit": "GB/s", "ms": round(copy_avg, 4),
                         "frac": round(copy_gbs / HBM_PEAK_GBS, 4),
                         "what": f"device-to-device copy of the 2^{args.log2n}-element input "
                                 "(torch copy_, 32 B moved per element), for context"},
        "verified_bit_exact": verified,
        "cpu_baseline": base,
    }
    del pristine, A, B
    torch.cuda.empty_cache()
    if not args.no_batched:
        log("batched segmented sort ...")
        out["batched"] = bench_batched(args.batched_segments)
        torch.cuda.empty_cache()
    if not args.no_serving:
        log("serving stream, sizes 2^12..2^24 ...")
        out["serving"] = bench_serving(args.serving_jobs, span_log2=12,
                                       streams=args.serving_streams)
        torch.cuda.empty_cache()
        if args.serving_wide_jobs > 0:
            log("serving stream, sizes 2^12..2^28 ...")
            out["serving_wide"] = bench_serving(args.serving_wide_jobs, span_log2=16,
                                                streams=args.serving_streams)
            torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def bench_batched(nseg, seed=2024, reps=3):
    """Batched segmented sort: nseg independent arrays, sizes log-uniform over
    [2^10, 2^22], uniform 64-bit keys, val = index within the segment; one
    segmented launch sequence (replayed from a hipGraph; most-significant-
    digit-first) against the LSD segmented form and one plain lsb_sort
    launch per segment, all on HBM-resident data."""
    import torch
    import lsb_amd
    from lsb_amd import workloads as W
    sizes = W.batched_sizes(nseg, seed)
    off = [0]
    for n in sizes:
        off.append(off[-1] + n)
    total = off[-1]
    dev = torch.device("cuda:0")
    A = torch.empty((total, 2), dtype=torch.int64, device=dev)
    for s, n in enumerate(sizes):
        lsb_amd.generate(A[off[s]:off[s] + n], kind=lsb_amd.GEN_UNIFORM, ranks=1,
                         seed_base=1000 + s, val_base=0)
    pristine = A.clone()
    B = torch.empty_like(A)
    ts = []
    for r in range(reps + 1):
        A.copy_(pristine)
        t = _timed(lambda: lsb_amd.sort_segmented(A, B, off, graph=True))
        if r:
            ts.append(t)
    t_batched = sorted(ts)[len(ts) // 2]
    # the LSD segmented form (LSB_SEG_LSD: one onesweep pass per digit over
    # every large segment) on the same input, for comparison
    tl = []
    for r in range(2):
        A.copy_(pristine)
        tl.append(_timed(lambda: lsb_amd.sort_segmented(A, B, off, graph=True, lsd=True)))
    t_lsd = min(tl)
    A.copy_(pristine)
    lsb_amd.sort_segmented(A, B, off, graph=True)
    # spot-check segments (the test suite checks all of them)
    ok = True
    keys = torch.empty(max(sizes), dtype=torch.int64, device=dev)
    for s in list(range(0, nseg, max(1, nseg // 64))):
        lo, n = off[s], sizes[s]
        lsb_amd.extract_keys(pristine[lo:lo + n], keys, n=n)
        bad, _ = lsb_amd.verify(A[lo:lo + n], keys, n=n)
        ok &= bad == 0
    del keys

    def plain():
        for s, n in enumerate(sizes):
            lsb_amd.sort(A[off[s]:off[s] + n], B[off[s]:off[s] + n])
    tp = []
    for r in range(2):
        A.copy_(pristine)
        tp.append(_timed(plain))
    t_plain = min(tp)
    del A, B, pristine
    return {"workload": f"{nseg} segments, sizes log-uniform [2^10, 2^22], uniform 64-bit keys",
            "elements": total, "ms": round(t_batched * 1e3, 3),
            "melem_s": round(total / t_batched / 1e6, 1),
            "plain_per_segment_ms": round(t_plain * 1e3, 3),
            "plain_per_segment_melem_s": round(total / t_plain / 1e6, 1),
            "speedup_vs_plain": round(t_plain / t_batched, 3),
            "lsd_form_ms": round(t_lsd * 1e3, 3), "msd_vs_lsd_form": round(t_lsd / t_batched, 3),
            "verified_sample": ok}


def bench_serving(jobs, seed=77, lo_log2=12, span_log2=16, streams=0):
    """Serving stream: jobs of log-uniform size over [2^lo, 2^(lo+span)],
    32- or 64-bit keys, uniform or Zipf(1.1), resident in HBM, admitted under
    the server's 200 GB budget (32 B per element: A and B).  Three runs on
    the same jobs: one plain lsb_sort per job in arrival order (each
    synchronised); the server with every job submitted at once (capacity);
    and the server fed by Poisson arrivals at 0.8 of that capacity
    (latency)."""
    import torch
    import lsb_amd
    from lsb_amd import workloads as W
    trace = W.serving_trace(jobs, seed, lo_log2=lo_log2, span_log2=span_log2)
    dev = torch.device("cuda:0")
    bufs, pris = [], []
    for j, job in enumerate(trace):
        a = torch.empty((job["n"], 2), dtype=torch.int64, device=dev)
        lsb_amd.generate(a, kind=lsb_amd.GEN_ZIPF if job["zipf"] else lsb_amd.GEN_UNIFORM,
                         ranks=4, key_bits=job["key_bits"], seed_base=7000 + 16 * j, val_base=0)
        bufs.append(a)
        pris.append(a.clone())
    total = sum(job["n"] for job in trace)

    def restore():
        for a, p in zip(bufs, pris):
            a.copy_(p)
        torch.cuda.synchronize()

    # one plain launch per job, in arrival order
    scratch = torch.empty((max(j["n"] for j in trace), 2), dtype=torch.int64, device=dev)
    restore()
    t_plain = 0.0
    for a, job in zip(bufs, trace):
        t_plain += _timed(lambda: lsb_amd.sort(a, scratch, key_bits=job["key_bits"]))
    del scratch
    torch.cuda.empty_cache()
    out = {"workload": f"{jobs} jobs, sizes log-uniform [2^{lo_log2}, 2^{lo_log2 + span_log2}], "
                       "32/64-bit keys, uniform/Zipf(1.1), HBM-resident",
           "jobs": jobs,
           "elements": total, "plain_melem_s": round(total / t_plain / 1e6, 1)}
    with lsb_amd.Server(streams=streams) as srv:
        # warm the server's streams, workspaces and graphs
        restore()
        for a, job in zip(bufs, trace):
            srv.submit(a, key_bits=job["key_bits"])
        srv.drain()
        # capacity: everything arrives at once (one batch submission, so the
        # scheduler sees the whole queue); the median of 3 runs
        runs = []
        for r in range(3):
            restore()
            srv.reset_stats()
            srv.submit_many(bufs, key_bits=[job["key_bits"] for job in trace])
            srv.drain()
            runs.append(srv.stats())
        runs.sort(key=lambda x: x["span_s"])
        st = runs[1]
        cap_s = st["span_s"]
        out.update({"saturated_melem_s": round(total / cap_s / 1e6, 1),
                    "saturated_p50_ms": round(st["p50_us"] / 1e3, 3),
                    "saturated_p99_ms": round(st["p99_us"] / 1e3, 3),
                    "served_vs_plain": round((total / cap_s) / (total / t_plain), 3),
                    "saturated_runs_melem_s": [round(total / x["span_s"] / 1e6, 1) for x in runs],
                    "batches": st["batches"], "batched_jobs": st["batched_jobs"]})
        # latency: Poisson arrivals at 0.8 of capacity
        lam = 0.8 * jobs / cap_s
        restore()
        srv.reset_stats()
        t0 = time.perf_counter()
        at = 0.0
        for a, job in zip(bufs, trace):
            at += job["gap"] / lam
            while time.perf_counter() - t0 < at:
                pass
            srv.submit(a, key_bits=job["key_bits"])
        srv.drain()
        st = srv.stats()
        out.update({"poisson_rate_jobs_s": round(lam, 2),
                    "poisson_melem_s": round(total / st["span_s"] / 1e6, 1),
                    "p50_ms": round(st["p50_us"] / 1e3, 3), "p99_ms": round(st["p99_us"] / 1e3, 3),
                    "max_ms": round(st["max_us"] / 1e3, 3),
                    "latency_samples": int(st["jobs_done"]),
                    "latency": "p50/p99/max of completion - submission over every job of the "
                               "Poisson run (arrivals at 0.8 of the saturated capacity)"})
    # every job's result is the stable sort of its input
    bad_jobs = []
    keys = torch.empty(max(j["n"] for j in trace), dtype=torch.int64, device=dev)
    for j, (a, p, job) in enumerate(zip(bufs, pris, trace)):
        lsb_amd.extract_keys(p, keys, n=job["n"])
        bad, first = lsb_amd.verify(a, keys, n=job["n"], key_bits=job["key_bits"])
        if bad:
            bad_jobs.append({"job": j, "n": job["n"], "key_bits": job["key_bits"],
                             "zipf": job["zipf"], "bad": bad, "first": first})
    out["verified_bit_exact"] = not bad_jobs
    if bad_jobs:
        out["bad_jobs"] = bad_jobs[:8]
        log("SERVING VERIFY FAILED:", len(bad_jobs), "jobs, first:", bad_jobs[:8])
    del bufs, pris, keys
    return out


if __name__ == "__main__":
    main()
