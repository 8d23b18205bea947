"""The serving loop returns, for every job, exactly what a lone lsb_sort of
that job returns -- whether the job ran by itself, inside a coalesced batch,
or overlapped with other jobs on other streams -- and its admission respects
the memory budget."""
import threading
import time

import numpy as np
import pytest
import torch

import lsb_amd
import oracle as O
from lsb_amd import workloads as W

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64).reshape(len(a), 2).copy()).to(DEV)


def from_dev(t):
    return t.cpu().numpy().reshape(-1).view(O.ELEM)


def make_jobs(count, seed, max_log2=20):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        n = int(2 ** rng.uniform(0, max_log2))
        kb = 32 if rng.random() < 0.5 else 64
        if rng.random() < 0.5:
            a = O.gen_uniform(n, ranks=2, key_bits=kb, seed_base=1000 * seed + i)
        else:
            a = O.gen_zipf(n, ranks=2, m=1 << 12, seed_base=1000 * seed + i)
            if kb == 32:
                a["key"] &= np.uint64(0xFFFFFFFF)
        jobs.append((a, kb))
    return jobs


def alone(a, kb):
    A = to_dev(a)
    if len(a) > 1:
        B = torch.empty_like(A)
        lsb_amd.sort(A, B, key_bits=kb, sync=True)
    return from_dev(A)


def expected(a, kb):
    k = a["key"] if kb >= 64 else a["key"] & np.uint64((1 << kb) - 1)
    return a[np.argsort(k, kind="stable")]


def test_each_job_same_alone_and_in_the_mix():
    jobs = make_jobs(120, seed=1)
    with lsb_amd.Server(streams=4) as srv:
        bufs = [to_dev(a) for a, _ in jobs]
        torch.cuda.synchronize()
        ids = [srv.submit(buf, key_bits=kb) for buf, (_, kb) in zip(bufs, jobs)]
        srv.drain()
        st = srv.stats()
        for jid in ids:
            assert srv.poll(jid)
            assert srv.latency_us(jid) > 0
    assert st["jobs_done"] == len(jobs)
    assert st["batched_jobs"] + st["plain_jobs"] == len(jobs)
    assert st["p99_us"] >= st["p50_us"] > 0
    for (a, kb), buf in zip(jobs, bufs):
        got = from_dev(buf)
        assert np.array_equal(got, expected(a, kb))
        assert np.array_equal(got, alone(a, kb))


def test_batch_submission():
    """submit_many queues a whole mix under one lock: every job still comes
    out as its lone sort; a bad entry (a misaligned array) refuses the whole
    call and queues nothing."""
    jobs = make_jobs(60, seed=3)
    with lsb_amd.Server(streams=4) as srv:
        bufs = [to_dev(a) for a, _ in jobs]
        torch.cuda.synchronize()
        ids = srv.submit_many(bufs, key_bits=[kb for _, kb in jobs])
        assert len(ids) == len(jobs) and ids == sorted(ids)
        srv.drain()
        assert srv.stats()["jobs_done"] == len(jobs)
        for jid in ids:
            assert srv.poll(jid)
        spare = torch.empty((1001, 2), dtype=torch.int64, device=DEV)
        misaligned = spare.view(-1)[1:2001]  # 8-byte offset from a 16-byte element
        with pytest.raises(lsb_amd.LsbError):
            srv.submit_many([bufs[0], misaligned])
        srv.drain()
        assert srv.stats()["jobs_done"] == len(jobs)
    for (a, kb), buf in zip(jobs, bufs):
        assert np.array_equal(from_dev(buf), expected(a, kb))


def test_key_orders_in_the_mix():
    """Jobs of different key orders are never batched together: signed,
    float and unsigned jobs of equal width submitted at once each come back
    in their own order."""
    rng = np.random.default_rng(21)
    jobs = []
    for i in range(36):
        n = int(2 ** rng.uniform(4, 17))
        order = ("unsigned", "signed", "float")[i % 3]
        a = np.zeros(n, dtype=O.ELEM)
        if order == "float":
            a["key"] = (rng.standard_normal(n) * 100).view(np.uint64)
        else:
            a["key"] = rng.integers(-(1 << 62), 1 << 62, size=n, dtype=np.int64).view(np.uint64)
        a["val"] = np.arange(n, dtype=np.uint64)
        jobs.append((a, order))
    with lsb_amd.Server(streams=2) as srv:
        bufs = [to_dev(a) for a, _ in jobs]
        torch.cuda.synchronize()
        for buf, (_, order) in zip(bufs, jobs):
            srv.submit(buf, key_order=order)
        srv.drain()
        assert srv.stats()["batched_jobs"] > 0
    for (a, order), buf in zip(jobs, bufs):
        assert np.array_equal(from_dev(buf), O.stable_sort_order(a, 64, order)), order


def test_coalescing_happens_under_load():
    """Many small jobs submitted at once: most must run inside batches."""
    rng = np.random.default_rng(5)
    jobs = [O.gen_uniform(int(rng.integers(100, 60000)), ranks=1, seed_base=i) for i in range(200)]
    with lsb_amd.Server(streams=2, batch_window_us=2000) as srv:
        bufs = [to_dev(a) for a in jobs]
        torch.cuda.synchronize()
        for b in bufs:
            srv.submit(b)
        srv.drain()
        st = srv.stats()
    assert st["batches"] >= 1 and st["batched_jobs"] >= 100, st
    for a, b in zip(jobs, bufs):
        assert np.array_equal(from_dev(b), expected(a, 64))


def test_budget_admission_and_rejection():
    """A budget that holds one large job at a time still completes every job
    (FIFO admission waits for memory); a job larger than the budget is
    refused at submission."""
    n = 1 << 20
    budget = 32 * n + 32 * 1000
    jobs = [O.gen_uniform(n, ranks=1, seed_base=s) for s in range(4)]
    with lsb_amd.Server(memory_budget=budget, streams=3) as srv:
        bufs = [to_dev(a) for a in jobs]
        torch.cuda.synchronize()
        ids = [srv.submit(b) for b in bufs]
        with pytest.raises(lsb_amd.LsbError):
            srv.submit(bufs[0], n=2 * n)
        for jid in ids:
            srv.wait(jid, 60.0)
        assert srv.stats()["memory_in_use"] == 0
    for a, b in zip(jobs, bufs):
        assert np.array_equal(from_dev(b), expected(a, 64))


def test_concurrent_submitters():
    jobs = make_jobs(40, seed=7, max_log2=17)
    bufs = [to_dev(a) for a, _ in jobs]
    torch.cuda.synchronize()
    with lsb_amd.Server() as srv:
        def worker(lo):
            for i in range(lo, len(jobs), 2):
                srv.wait(srv.submit(bufs[i], key_bits=jobs[i][1]), 60.0)
        ts = [threading.Thread(target=worker, args=(k,)) for k in (0, 1)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
    for (a, kb), b in zip(jobs, bufs):
        assert np.array_equal(from_dev(b), expected(a, kb))


def test_finished_jobs_retained_up_to_limit():
    """A long-running server keeps the most recent 2^18 finished jobs (and
    latencies) only: a job that finished 2^18 + 1000 jobs ago is forgotten
    (LSB_ENOTFOUND) while the latest ones still answer poll / latency."""
    import ctypes
    retain = 1 << 18
    a0 = O.gen_uniform(1000, ranks=1, seed_base=77)
    buf0 = to_dev(a0)
    one = torch.zeros((1, 2), dtype=torch.int64, device=DEV)
    torch.cuda.synchronize()
    L = lsb_amd.lib()
    with lsb_amd.Server(streams=2, use_graphs=False) as srv:
        first = srv.submit(buf0)
        srv.drain()
        assert srv.poll(first) and srv.latency_us(first) > 0
        total, chunk, last = retain + 1000, 32768, []
        done = 0
        while done < total:
            m = min(chunk, total - done)
            As = (ctypes.c_void_p * m)(*([one.data_ptr()] * m))
            ns = (ctypes.c_size_t * m)(*([1] * m))
            bits = (ctypes.c_uint32 * m)(*([64] * m))
            flags = (ctypes.c_uint32 * m)()
            ids = (ctypes.c_uint64 * m)()
            assert L.lsb_server_submit_batch(srv._h, As, ns, bits, flags, m, ids) == 0
            srv.drain()
            last = list(ids[max(0, m - 16):m])
            done += m
        with pytest.raises(lsb_amd.LsbError) as err:
            srv.poll(first)
        assert err.value.code == lsb_amd.LSB_ENOTFOUND
        for jid in last:
            assert srv.poll(jid) and srv.latency_us(jid) >= 0
        st = srv.stats()
        assert st["jobs_done"] == total + 1 and st["p99_us"] >= st["p50_us"] > 0
        srv.reset_stats()
        with pytest.raises(lsb_amd.LsbError):
            srv.poll(last[-1])
    assert np.array_equal(from_dev(buf0), expected(a0, 64))


def _torch_stable_ok(A_sorted, keys):
    """torch's own stable GPU sort as an independent check of one job:
    sorted keys and vals (= input indices) in stable order."""
    flip = torch.tensor(-(1 << 63), dtype=torch.int64, device=DEV)
    sk, idx = torch.sort(torch.bitwise_xor(keys, flip), stable=True)
    return (torch.equal(torch.bitwise_xor(sk, flip), A_sorted[:, 0])
            and torch.equal(idx, A_sorted[:, 1]))


def _poisson_feed(srv, bufs, trace, rate):
    """Submits job j at time sum(gap[0..j]) / rate from a thread of its own,
    as arrivals that come while earlier jobs are on the GPU."""
    ids = [None] * len(bufs)
    err = []

    def feed():
        try:
            t0, at = time.perf_counter(), 0.0
            for j, (a, job) in enumerate(zip(bufs, trace)):
                at += job["gap"] / rate
                d = at - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                ids[j] = srv.submit(a, key_bits=job["key_bits"])
        except Exception as e:
            err.append(repr(e))

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    assert err == []
    srv.drain()
    return ids


def test_serving_stream_full_workload():
    """The serving config at its own workload: 40 jobs of log-uniform size
    over [2^12, 2^28] (1.15 G elements, largest 2.48e8), 32- or 64-bit keys,
    uniform or Zipf(1.1) keys, arriving as a Poisson stream from another
    thread.  Run twice: with the default server (200 GB budget; jobs up to
    2^24 coalesced, larger ones plain with a scratch slab) and with a 10 GB
    budget and 2^22-element slabs, where admission waits for memory and jobs
    above 2^22 take their scratch from the memory pool.  Every job is proven
    the stable sort of its input on the device; one batched job and one
    slab-plain job are compared with numpy's stable sort, the largest with
    torch's stable sort."""
    trace = W.serving_trace(40, 79, lo_log2=12, span_log2=16)
    bufs, pris, keys = [], [], []
    for j, job in enumerate(trace):
        a = torch.empty((job["n"], 2), dtype=torch.int64, device=DEV)
        lsb_amd.generate(a, kind=lsb_amd.GEN_ZIPF if job["zipf"] else lsb_amd.GEN_UNIFORM,
                         ranks=4, key_bits=job["key_bits"], seed_base=9100 + 16 * j)
        k = torch.empty(job["n"], dtype=torch.int64, device=DEV)
        lsb_amd.extract_keys(a, k)
        bufs.append(a)
        pris.append(a.clone())
        keys.append(k)
    torch.cuda.synchronize()
    ns = [job["n"] for job in trace]
    assert max(ns) > 2 ** 27 and sum(n > 2 ** 24 for n in ns) >= 5
    small = next(j for j, n in enumerate(ns) if 2 ** 16 <= n <= 2 ** 20)
    slab = min((j for j, n in enumerate(ns) if n > 2 ** 24), key=lambda j: ns[j])
    largest = max(range(len(ns)), key=lambda j: ns[j])
    host_in = {j: pris[j].cpu().numpy().reshape(-1).view(O.ELEM).copy() for j in (small, slab)}

    def check_all(tag):
        for j, (a, job) in enumerate(zip(bufs, trace)):
            assert lsb_amd.verify(a, keys[j], key_bits=job["key_bits"]) == (0, -1), (tag, j)
        for j in (small, slab):
            got = from_dev(bufs[j])
            assert np.array_equal(got, expected(host_in[j], trace[j]["key_bits"])), (tag, j)
        assert _torch_stable_ok(bufs[largest], keys[largest]), (tag, largest)

    print(f"serving full workload: {len(ns)} jobs, {sum(ns)} elements, default server",
          flush=True)
    with lsb_amd.Server() as srv:
        _poisson_feed(srv, bufs, trace, rate=2000.0)
        st = srv.stats()
    assert st["jobs_done"] == len(ns)
    # how many small jobs coalesce depends on how many arrive while the
    # device is busy; at least one batch and the slab-plain jobs must run
    assert st["batches"] >= 1 and st["batched_jobs"] >= 2, st
    assert st["plain_jobs"] >= sum(n > 2 ** 24 for n in ns), st
    assert st["pool_plain_jobs"] == 0 and st["memory_in_use"] == 0, st
    check_all("default")

    for a, p in zip(bufs, pris):
        a.copy_(p)
    torch.cuda.synchronize()
    print("serving full workload: 10 GB budget, 2^22-element slabs", flush=True)
    with lsb_amd.Server(memory_budget=10 * 10 ** 9, batch_max_elems=1 << 20,
                        batch_max_total=1 << 22) as srv:
        _poisson_feed(srv, bufs, trace, rate=2000.0)
        st = srv.stats()
    assert st["jobs_done"] == len(ns)
    assert st["pool_plain_jobs"] >= sum(n > 2 ** 22 for n in ns), st
    assert st["admission_waits"] >= 1 and st["batches"] >= 1, st
    assert st["memory_in_use"] == 0
    check_all("small budget")
