"""Benchmark: one GPU's share at the reference README's scale.

Workload: 2^30 16-byte {uint64 key uniform over [0, 2^64), uint64 val =
index} pairs (16 GiB in A + 16 GiB ping-pong B), keys from pcg64 over 256
virtual ranks exactly like the reference keygen (mpi/mpi_lsbsort.cpp:
676-704), sorted by the full 64-bit key, stable.  A step is one lsb_sort of
that array: the input is restored from a pristine device copy before every
step (untimed), then the sort is timed between two device synchronisations.
Metric: Melem/s sorted = n / t_sort / 1e6 (mpi/mpi_lsbsort.cpp:741).

roofline: the dominant kernel is the onesweep scatter pass.  Its
algorithmic bytes are 32 B per element per executed pass (16 B read + 16 B
written); its duration is read from HIP events recorded on the sort's stream
between launches (lsb_profile_read), averaged over the timed steps.  The
whole sort is also reported against the model n * (8 + 32 * P_exec) bytes
(one 8-byte key read for the all-digit histogram, 32 B per executed pass).

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


def cpu_baseline(sample_n, max_cores, reps):
    """Reference mpi_lsbsort on the host (kind=reference), else the oracle port."""
    ref = os.path.join(ROOT, "oracle", "_ref", "mpi_lsbsort")
    mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
    cores = max(1, min(max_cores, len(os.sched_getaffinity(0))))
    if os.path.exists(ref) and os.path.exists(mpirun):
        rates = []
        try:
            probe = subprocess.run([mpirun, "-n", "2", ref, "--n", "1000", "--no-verify"],
                                   capture_output=True, text=True, timeout=60)
            if probe.returncode != 0:
                raise RuntimeError("mpirun probe failed: " + probe.stderr[-300:])
            for r in range(reps):
                log(f"  reference mpi_lsbsort rep {r + 1}/{reps}: n={sample_n}, {cores} ranks")
                p = subprocess.run([mpirun, "-n", str(cores), ref, "--n", str(sample_n),
                                    "--no-verify"], capture_output=True, text=True, timeout=150)
                m = re.search(r"That's ([0-9.eE+]+) M elements sorted / s", p.stdout)
                if p.returncode != 0 or not m:
                    raise RuntimeError(p.stdout[-500:] + p.stderr[-500:])
                rates.append(float(m.group(1)))
            rates.sort()
            return {"value": rates[len(rates) // 2], "unit": "Melem/s", "cores": cores,
                    "kind": "reference",
                    "sample": f"mpi_lsbsort --n {sample_n} --no-verify, mpirun -n {cores}, "
                              f"median of {reps} (reference keygen pcg64(rank), 64-bit keys)"}
        except Exception as e:  # fall through to the port
            log("reference CPU baseline failed:", e)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    n = min(sample_n, 1 << 24)
    a = O.gen_uniform(n, ranks=4)
    t = time.perf_counter()
    O.lsd_sort(a, 64, 16)
    dt = time.perf_counter() - t
    return {"value": n / dt / 1e6, "unit": "Melem/s", "cores": 1, "kind": "port",
            "sample": f"oracle C LSD sort (16-bit digits, one thread), n={n}"}


def traffic_from_profile(log2n, d, tile):
    """HBM bytes per scatter launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary, when it was taken on this workload and layout."""
    path = os.path.join(ROOT, "profiles", "k_onesweep_rg_2p30_uniform_traffic.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch: (\d+)$", text, re.M)
    c = re.search(r"^traffic_config: log2n=(\d+) digit_bits=(\d+) tile=(\d+)$", text, re.M)
    if not m or not c or tuple(map(int, c.groups())) != (log2n, d, tile):
        return None
    return int(m.group(1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=30)
    ap.add_argument("--digit-bits", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log2n", type=int, default=26)
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-segments", type=int, default=4096)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-jobs", type=int, default=96)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the sort is one rank's share on one MI355X "
              "(the multi-GPU exchange is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    base = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        base = cpu_baseline(1 << args.cpu_sample_log2n, 16, 3)
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
    d, tile, P = lsb_amd.plan(n, digit_bits=args.digit_bits, tile=args.tile)
    torch.cuda.synchronize()

    def step(profile):
        A.copy_(pristine)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lsb_amd.sort(A, B, digit_bits=args.digit_bits, tile=args.tile, profile=profile)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    for i in range(args.warmup):
        step(False)
    log(f"warmup done ({args.warmup}); timing {args.steps} steps ...")
    wall, profs = [], []
    for i in range(args.steps):
        wall.append(step(True))
        profs.append(lsb_amd.profile_read())
    ms_wall = sum(wall) / len(wall) * 1e3
    ms_event = sum(p["total_ms"] for p in profs) / len(profs)

    executed = profs[-1]["executed"]
    p_exec = sum(executed)
    pass_ms = [t for p in profs for t, e in zip(p["pass_ms"], p["executed"]) if e]
    avg_pass_ms = sum(pass_ms) / len(pass_ms) if pass_ms else float("nan")
    hist_ms = sum(p["histogram_ms"] for p in profs) / len(profs)
    pass_bytes = 32 * n
    achieved = pass_bytes / (avg_pass_ms * 1e-3) / 1e9
    model_bytes = n * (8 + 32 * p_exec)
    sort_gbs = model_bytes / (ms_wall * 1e-3) / 1e9

    verified = None
    if not args.no_verify:
        keys = torch.empty(n, dtype=torch.int64, device=dev)
        lsb_amd.extract_keys(pristine, keys)
        bad, first = lsb_amd.verify(A, keys)
        verified = bad == 0
        del keys
        if not verified:
            log(f"VERIFY FAILED: {bad} bad positions, first {first}")

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

    traffic = traffic_from_profile(args.log2n, d, tile)

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
        "config": {"workload": f"plain LSD sort, 2^{args.log2n} 16-byte pairs, uniform 64-bit "
                               "keys, val = index, stable, HBM-resident",
                   "n": n, "digit_bits": d, "tile": tile, "passes": P,
                   "passes_executed": p_exec},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": f"onesweep scatter pass (d={d}, tile={tile}), 32 B/elem x 2^{args.log2n}",
                     "avg_launch_ms": round(avg_pass_ms, 4)},
        "sort_roofline": {"model_bytes": model_bytes, "achieved": round(sort_gbs, 1),
                          "frac": round(sort_gbs / HBM_PEAK_GBS, 4),
                          "ms_event": round(ms_event, 4), "histogram_ms": round(hist_ms, 4)},
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
        out["serving"] = bench_serving(args.serving_jobs, span_log2=12)
        torch.cuda.empty_cache()
        log("serving stream, sizes 2^12..2^28 ...")
        out["serving_wide"] = bench_serving(args.serving_jobs, span_log2=16)
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
    segmented launch sequence (replayed from a hipGraph) against one plain
    lsb_sort launch per segment, both on HBM-resident data."""
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
            "speedup_vs_plain": round(t_plain / t_batched, 3), "verified_sample": ok}


def bench_serving(jobs, seed=77, lo_log2=12, span_log2=16):
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
           "elements": total, "plain_melem_s": round(total / t_plain / 1e6, 1)}
    with lsb_amd.Server() as srv:
        # warm the server's streams, workspaces and graphs
        restore()
        for a, job in zip(bufs, trace):
            srv.submit(a, key_bits=job["key_bits"])
        srv.drain()
        # capacity: everything arrives at once (one batch submission, so the
        # scheduler sees the whole queue)
        restore()
        srv.reset_stats()
        srv.submit_many(bufs, key_bits=[job["key_bits"] for job in trace])
        srv.drain()
        st = srv.stats()
        cap_s = st["span_s"]
        out.update({"saturated_melem_s": round(total / cap_s / 1e6, 1),
                    "saturated_p50_ms": round(st["p50_us"] / 1e3, 3),
                    "saturated_p99_ms": round(st["p99_us"] / 1e3, 3),
                    "served_vs_plain": round((total / cap_s) / (total / t_plain), 3),
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
                    "p50_ms": round(st["p50_us"] / 1e3, 3), "p99_ms": round(st["p99_us"] / 1e3, 3)})
    # every job's result is the stable sort of its input
    ok = True
    keys = torch.empty(max(j["n"] for j in trace), dtype=torch.int64, device=dev)
    for a, p, job in zip(bufs, pris, trace):
        lsb_amd.extract_keys(p, keys, n=job["n"])
        bad, _ = lsb_amd.verify(a, keys, n=job["n"], key_bits=job["key_bits"])
        ok &= bad == 0
    out["verified_bit_exact"] = ok
    del bufs, pris, keys
    return out


if __name__ == "__main__":
    main()
