"""GPU parity of the HIP sort against the CPU oracle, through the C ABI.

Every comparison is bit-exact on (key, val): integers only.  Sizes the oracle
finishes in seconds are compared element by element; the full-size configs
are proven with the O(n) invariants (permutation, key matches source,
sorted, equal keys in index order) on the device, which together imply
equality with the stable sort."""
import numpy as np
import pytest
import torch

import lsb_amd
import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64).reshape(len(a), 2).copy()).to(DEV)


def from_dev(t):
    return t.cpu().numpy().reshape(-1).view(O.ELEM)


def gpu_sort(a, **kw):
    A = to_dev(a)
    B = torch.empty_like(A)
    lsb_amd.sort(A, B, sync=True, **kw)
    return from_dev(A)


def stable(a, key_bits=64):
    if key_bits >= 64:
        return a[np.argsort(a["key"], kind="stable")]
    return a[np.argsort(a["key"] & np.uint64((1 << key_bits) - 1), kind="stable")]


VARIANTS = None


def variants():
    global VARIANTS
    if VARIANTS is None:
        VARIANTS = lsb_amd.variants()
    return VARIANTS


SIZES = [2, 3, 63, 64, 65, 1000, 2047, 2048, 2049, 4095, 4096, 4097, 6143, 6144, 6145,
         8191, 8192, 8193, 65537, 100003, (1 << 20) + 7]


@pytest.mark.parametrize("n", SIZES)
def test_sort_uniform_all_variants(n):
    a = O.gen_uniform(n, ranks=3)
    want = stable(a)
    for v, d, tile, _ in variants():
        got = gpu_sort(a, variant=v)
        assert np.array_equal(got, want), (n, d, tile, v)


def test_sort_default_and_empty():
    for n in (0, 1):
        A = torch.zeros((max(n, 1), 2), dtype=torch.int64, device=DEV)
        A[0, 0] = 5
        B = torch.empty_like(A)
        lsb_amd.sort(A, B, n=n, sync=True)
        assert A[0, 0].item() == 5


def test_c1_reference_config_digest():
    """n = 2^24 keys of the reference's 4-rank input: the GPU generator
    reproduces the input digest and the sort reproduces the sorted digest."""
    n = 1 << 24
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=4)
    torch.cuda.synchronize()
    host = from_dev(A)
    assert O.sha256(host) == "4a8c88b35e1bc097a11998afab7608c9c3f66c2d4dd56fb501233124a51d450d"
    for v, d, tile, _ in variants():
        A2 = A.clone()
        B = torch.empty_like(A2)
        lsb_amd.sort(A2, B, variant=v, sync=True)
        assert O.sha256(from_dev(A2)) == \
            "edbb62f7b024eaade41bc12278e7cb240d038889a9780f8913a336352530d0c9", (d, tile, v)


@pytest.mark.parametrize("kind,kw", [
    (lsb_amd.GEN_UNIFORM, dict(ranks=3)),
    (lsb_amd.GEN_UNIFORM, dict(ranks=7, key_bits=32, seed_base=11)),
    (lsb_amd.GEN_SORTED, {}),
    (lsb_amd.GEN_REVERSE, {}),
    (lsb_amd.GEN_ZIPF, dict(ranks=5, zipf_m=1 << 20)),
    (lsb_amd.GEN_ZIPF, dict(ranks=3, zipf_m=1 << 16, key_bits=32)),
])
def test_generator_matches_oracle(kind, kw):
    n = (1 << 18) + 12345
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    lsb_amd.generate(A, kind=kind, **kw)
    torch.cuda.synchronize()
    got = from_dev(A)
    if kind == lsb_amd.GEN_UNIFORM:
        want = O.gen_uniform(n, ranks=kw["ranks"], key_bits=kw.get("key_bits", 64),
                             seed_base=kw.get("seed_base", 0))
    elif kind == lsb_amd.GEN_ZIPF:
        want = O.gen_zipf(n, ranks=kw["ranks"], m=kw["zipf_m"], key_bits=kw.get("key_bits", 64))
    else:
        want = O.gen_ordered(n, reverse=kind == lsb_amd.GEN_REVERSE)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("d", [8, 11])
@pytest.mark.parametrize("case", ["uniform", "zipf", "sorted", "low32"])
def test_histogram_kernel_equals_oracle(d, case):
    n = 300_001
    a = {"uniform": lambda: O.gen_uniform(n, ranks=2),
         "zipf": lambda: O.gen_zipf(n, ranks=2),
         "sorted": lambda: O.gen_ordered(n),
         "low32": lambda: O.gen_uniform(n, ranks=2, key_bits=32)}[case]()
    P = O.n_digits(64, d)
    counts = torch.zeros(P * (1 << d), dtype=torch.int32, device=DEV)
    lsb_amd.debug_histogram(to_dev(a), counts, digit_bits=d)
    got = counts.cpu().numpy().view(np.uint32).reshape(P, 1 << d)
    assert np.array_equal(got, O.histograms(a, 64, d))


def _skew_cases(n):
    rng = np.random.default_rng(7)
    out = {
        "low32_keys": O.gen_uniform(n, ranks=2, key_bits=32),
        "zipf": O.gen_zipf(n, ranks=4),
        "sorted": O.gen_ordered(n),
        "reverse": O.gen_ordered(n, reverse=True),
    }
    eq = np.zeros(n, dtype=O.ELEM)
    eq["key"] = 0xDEADBEEF
    eq["val"] = np.arange(n)
    out["all_equal"] = eq
    few = np.zeros(n, dtype=O.ELEM)
    few["key"] = rng.integers(0, 3, n).astype(np.uint64) << np.uint64(61)
    few["val"] = np.arange(n)
    out["three_keys_top_bits"] = few
    one = O.gen_uniform(n, ranks=1)
    one["key"] &= np.uint64(0xFF00)  # a single varying digit
    out["one_digit"] = one
    return out


@pytest.mark.parametrize("case", sorted(_skew_cases(10)))
def test_skew_and_stability(case):
    n = 200_003
    a = _skew_cases(n)[case]
    want = stable(a)
    for v, d, tile, _ in variants():
        assert np.array_equal(gpu_sort(a, variant=v), want), (case, d, tile, v)
        assert np.array_equal(gpu_sort(a, variant=v, no_skip=True), want)


@pytest.mark.parametrize("low_bits", [8, 16, 24])
def test_constant_low_digits(low_bits):
    """Keys whose low bytes are one constant: those passes are skipped and
    the first executed pass is a later digit -- for the region form one the
    up-front histogram counts (digit 1) or one a second read counts (digits
    2, 3)."""
    n = 300_007
    a = O.gen_uniform(n, ranks=2)
    low = np.uint64(0x5AA55A & ((1 << low_bits) - 1))
    a["key"] = (a["key"] << np.uint64(low_bits)) | low
    want = stable(a)
    skipped = low_bits // 8
    for v, d, tile, _ in variants():
        A = to_dev(a)
        B = torch.empty_like(A)
        lsb_amd.sort(A, B, variant=v, sync=True, profile=True)
        executed = lsb_amd.profile_read()["executed"]
        assert np.array_equal(from_dev(A), want), (low_bits, v, d, tile)
        if d == 8:
            assert list(executed) == [0] * skipped + [1] * (8 - skipped), (low_bits, v, executed)


@pytest.mark.parametrize("key_bits", [1, 8, 16, 21, 32, 33, 40, 57, 63])
def test_key_bits_and_per_pass_prefix(key_bits):
    """Sorting by the low key_bits is what the first passes leave behind:
    after pass p the array is stably sorted by the low (p+1)d bits."""
    n = 150_001
    a = O.gen_uniform(n, ranks=2)
    want = stable(a, key_bits)
    for v, d, tile, _ in variants():
        assert np.array_equal(gpu_sort(a, key_bits=key_bits, variant=v), want), (v, key_bits)


def test_non_default_stream_and_caller_workspace():
    n = 500_000
    a = O.gen_uniform(n, ranks=2)
    want = stable(a)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        A = to_dev(a)
        B = torch.empty_like(A)
        ws = torch.empty(lsb_amd.workspace_size(n), dtype=torch.uint8, device=DEV)
        lsb_amd.sort(A, B, stream=s, workspace=ws)
    s.synchronize()
    assert np.array_equal(from_dev(A), want)
    # too-small workspace is rejected, not overrun
    small = torch.empty(256, dtype=torch.uint8, device=DEV)
    with pytest.raises(lsb_amd.LsbError):
        lsb_amd.sort(A, B, workspace=small)


def test_rejects_host_pointer():
    n = 1000
    host = np.zeros(2 * n + 2, dtype=np.int64)
    B = torch.zeros((n, 2), dtype=torch.int64, device=DEV)
    with pytest.raises(lsb_amd.LsbError):
        lsb_amd.sort(host.ctypes.data + (16 - host.ctypes.data % 16) % 16, B.data_ptr(), n=n)


def test_release_frees_stream_state_and_sorts_again():
    """lsb_release gives back a stream's cached workspaces (plain and
    segmented) and the graphs captured against them; sorts on that stream
    afterwards, graph-replayed ones included, allocate and capture again and
    stay bit-exact."""
    n = 1 << 26
    a = O.gen_uniform(1 << 20, ranks=2, seed_base=31)
    want = stable(a)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        big = torch.empty((n, 2), dtype=torch.int64, device=DEV)
        lsb_amd.generate(big, kind=lsb_amd.GEN_UNIFORM, ranks=4, stream=s)
        big_b = torch.empty_like(big)
        lsb_amd.sort(big, big_b, stream=s)  # a workspace of about 20 MiB
        for graph in (False, True, True):
            A = to_dev(a)
            B = torch.empty_like(A)
            lsb_amd.sort(A, B, stream=s, graph=graph)
            s.synchronize()
            assert np.array_equal(from_dev(A), want), graph
        off = [0, 3000, 300000, len(a)]
        A = to_dev(a)
        B = torch.empty_like(A)
        lsb_amd.sort_segmented(A, B, off, stream=s, sync=True)
        got = from_dev(A)
        for lo, hi in zip(off[:-1], off[1:]):
            assert np.array_equal(got[lo:hi], stable(a[lo:hi]))
        s.synchronize()
        del big, big_b
        torch.cuda.empty_cache()
        free0 = torch.cuda.mem_get_info()[0]
        lsb_amd.release(s)
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 - free0 >= 8 << 20, (free0, free1)
        for graph in (True, False):
            A = to_dev(a)
            B = torch.empty_like(A)
            lsb_amd.sort(A, B, stream=s, graph=graph)
            s.synchronize()
            assert np.array_equal(from_dev(A), want), graph
        lsb_amd.release(s)
    lsb_amd.release(torch.cuda.Stream())  # a stream the library never saw


def test_msd_sorts_on_many_streams_at_once():
    """Eight MSD-first sorts of one heavy-duplicate input (Zipf(1.1) keys,
    2^24 + 12,345 elements: the plain sort's MSD form, with runs that take
    every finish path) in flight together on eight streams, six times; every
    result passes the device checker.  Splitting the finish into a kernel for
    most runs and a second kernel for the rest gave wrong elements at run
    boundaries here (about one sort in ten) while every single-stream test
    passed."""
    n = (1 << 24) + 12_345
    src = torch.empty((n, 2), dtype=torch.int64, device="cuda")
    lsb_amd.generate(src, kind=lsb_amd.GEN_ZIPF, ranks=4, seed_base=4242)
    assert lsb_amd.sort_form(n) == "msd"
    keys = torch.empty(n, dtype=torch.int64, device="cuda")
    lsb_amd.extract_keys(src, keys)
    streams = [torch.cuda.Stream() for _ in range(8)]
    As = [torch.empty_like(src) for _ in streams]
    Bs = [torch.empty_like(src) for _ in streams]
    for rep in range(6):
        for A in As:
            A.copy_(src)
        torch.cuda.synchronize()
        for A, B, s in zip(As, Bs, streams):
            lsb_amd.sort(A, B, stream=s)
        torch.cuda.synchronize()
        for k, A in enumerate(As):
            bad, first = lsb_amd.verify(A, keys)
            assert bad == 0, (rep, k, bad, first)
    for s in streams:
        lsb_amd.release(s)


def test_concurrent_host_threads():
    """lsb_sort is reentrant across streams: four host threads, each on its
    own stream, run plain (eager and graph-replayed), segmented and batched
    sorts at the same time (the binding releases the GIL in every call), and
    every result is the stable sort."""
    import threading
    cases = []
    for tid in range(4):
        plain = [O.gen_uniform(n, ranks=2, seed_base=500 + 10 * tid + k)
                 for k, n in enumerate((5000, 200_000, (1 << 20) + 3))]
        seg = O.gen_zipf(300_000, ranks=2, m=1 << 10, seed_base=600 + tid)
        off = [0, 7, 40_000, 41_000, 300_000]
        batch = [O.gen_uniform(n, ranks=1, seed_base=700 + 10 * tid + k)
                 for k, n in enumerate((3000, 90_000, 8192))]
        cases.append((plain, seg, off, batch))
    streams = [torch.cuda.Stream() for _ in range(4)]
    errors = []

    def work(tid):
        plain, seg, off, batch = cases[tid]
        s = streams[tid]
        try:
            with torch.cuda.stream(s):
                for rep in range(3):
                    for a in plain:
                        A = to_dev(a)
                        B = torch.empty_like(A)
                        lsb_amd.sort(A, B, stream=s, graph=rep == 2)
                        s.synchronize()
                        if not np.array_equal(from_dev(A), stable(a)):
                            errors.append(("plain", tid, rep, len(a)))
                    A = to_dev(seg)
                    B = torch.empty_like(A)
                    lsb_amd.sort_segmented(A, B, off, stream=s, sync=True, graph=rep == 2)
                    got = from_dev(A)
                    for lo, hi in zip(off[:-1], off[1:]):
                        if not np.array_equal(got[lo:hi], stable(seg[lo:hi])):
                            errors.append(("segmented", tid, rep, lo))
                    As = [to_dev(x) for x in batch]
                    Bs = [torch.empty_like(x) if x.shape[0] > 8192 else None for x in As]
                    lsb_amd.sort_batch(As, Bs, stream=s, sync=True, graph=rep == 2)
                    for x, X in zip(batch, As):
                        if not np.array_equal(from_dev(X), stable(x)):
                            errors.append(("batch", tid, rep, len(x)))
        except Exception as e:  # reported below, not lost in the thread
            errors.append(("raised", tid, repr(e)))

    threads = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert errors == []
    for s in streams:
        lsb_amd.release(s)


def test_two_threads_one_stream_growing_sizes():
    """Two host threads sorting on the SAME stream with sizes that keep
    growing, so one thread regrows the stream's cached workspace while the
    other's sorts are queued: the old workspace must outlive them (the
    entry's lock is held from lookup until the launches are enqueued)."""
    import threading
    s = torch.cuda.Stream()
    sizes = [[(1 << k) + 3 * t + 1 for k in range(14, 22)] for t in range(2)]
    inputs = [[O.gen_uniform(n, ranks=1, seed_base=900 + 50 * t + i)
               for i, n in enumerate(sizes[t])] for t in range(2)]
    errors = []
    bufs = [[(to_dev(a), torch.empty((len(a), 2), dtype=torch.int64, device=DEV))
             for a in inputs[t]] for t in range(2)]
    torch.cuda.synchronize()
    start = threading.Barrier(2)

    def work(t):
        try:
            start.wait()
            for A, B in bufs[t]:
                lsb_amd.sort(A, B, stream=s)
        except Exception as e:
            errors.append(repr(e))

    ths = [threading.Thread(target=work, args=(t,)) for t in range(2)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    s.synchronize()
    assert errors == []
    for t in range(2):
        for a, (A, _) in zip(inputs[t], bufs[t]):
            assert np.array_equal(from_dev(A), stable(a)), (t, len(a))
    lsb_amd.release(s)


def test_repeated_sorts_reuse_workspace():
    """Back-to-back sorts of different sizes on one stream share the cached
    workspace; stale look-back state must never leak between them."""
    for n in (1 << 20, 5000, 1 << 16, (1 << 20) + 3, 77):
        a = O.gen_uniform(n, ranks=2, seed_base=n)
        assert np.array_equal(gpu_sort(a), stable(a)), n


def _device_check(A_sorted, keys, n, key_bits=64):
    bad, first = lsb_amd.verify(A_sorted, keys, n=n, key_bits=key_bits)
    return bad, first


def _torch_stable_check(A_sorted, keys):
    """Independent full-size check with torch's own stable GPU sort: the
    output's keys must be the sorted keys and its vals (= input indices) the
    stable permutation.  Unsigned order is signed order with the top bit
    flipped."""
    flip = torch.tensor(-(1 << 63), dtype=torch.int64, device=DEV)
    sk, idx = torch.sort(torch.bitwise_xor(keys, flip), stable=True)
    ok_keys = torch.equal(torch.bitwise_xor(sk, flip), A_sorted[:, 0])
    del sk
    ok_vals = torch.equal(idx, A_sorted[:, 1])
    del idx
    return ok_keys and ok_vals


def test_device_verifier_detects_wrong_output():
    n = 100_000
    a = O.gen_zipf(n, ranks=2, m=256)
    s = stable(a)
    keys = torch.from_numpy(a["key"].view(np.int64).copy()).to(DEV)
    assert _device_check(to_dev(s), keys, n) == (0, -1)
    for mutate in ("swap", "unstable", "dup"):
        b = s.copy()
        if mutate == "swap":
            b[[10, 90000]] = b[[90000, 10]]
        elif mutate == "unstable":
            i = int(np.nonzero(b["key"][1:] == b["key"][:-1])[0][0])
            b[[i, i + 1]] = b[[i + 1, i]]
        else:
            b[500]["val"] = b[501]["val"]
        bad, first = _device_check(to_dev(b), keys, n)
        assert bad > 0 and first >= 0, mutate


@pytest.mark.parametrize("cfg", [
    dict(kind=lsb_amd.GEN_UNIFORM, ranks=64, key_bits=32),   # high 32 bits zero: skipped passes
    dict(kind=lsb_amd.GEN_ZIPF, ranks=64),                   # heavy duplicates
    dict(kind=lsb_amd.GEN_SORTED),
    dict(kind=lsb_amd.GEN_REVERSE),
], ids=["low32", "zipf", "sorted", "reverse"])
def test_stress_2p28(cfg):
    """The skew/stability stress configs at their full 2^28 size."""
    n = 1 << 28
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    lsb_amd.generate(A, **cfg)
    keys = torch.empty(n, dtype=torch.int64, device=DEV)
    lsb_amd.extract_keys(A, keys)
    B = torch.empty_like(A)
    lsb_amd.sort(A, B, sync=True)
    assert _device_check(A, keys, n) == (0, -1)
    del B
    kind = cfg["kind"]
    if kind in (lsb_amd.GEN_SORTED, lsb_amd.GEN_REVERSE):
        # closed form: sorted input is its own result; reversed input (key
        # n-1-i at index i) comes out as key j carrying val n-1-j
        j = torch.arange(n, dtype=torch.int64, device=DEV)
        assert torch.equal(A[:, 0], j)
        assert torch.equal(A[:, 1], j if kind == lsb_amd.GEN_SORTED else (n - 1) - j)
        del j
    else:
        assert _torch_stable_check(A, keys)
    del A, keys
    torch.cuda.empty_cache()


def test_full_2p30_uniform():
    """One GPU's share at the README's scale: 2^30 pairs (16 GiB + 16 GiB)."""
    n = 1 << 30
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=256)
    keys = torch.empty(n, dtype=torch.int64, device=DEV)
    lsb_amd.extract_keys(A, keys)
    # the first 2^24 elements are the reference's 4-rank input (same per-rank size)
    head = from_dev(A[: 1 << 16])
    assert np.array_equal(head, O.gen_uniform(1 << 24, ranks=4, count=1 << 16))
    B = torch.empty_like(A)
    lsb_amd.sort(A, B, sync=True)
    del B
    assert _device_check(A, keys, n) == (0, -1)
    assert _torch_stable_check(A, keys)
    del A, keys
    torch.cuda.empty_cache()


def test_max_n_uniform():
    """The largest n the ABI accepts (LSB_MAX_N = 2^31 - 2, 32 GiB per
    buffer): the last tile is partial and ends one short of 2^31."""
    n = lsb_amd.LSB_MAX_N
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=256)
    keys = torch.empty(n, dtype=torch.int64, device=DEV)
    lsb_amd.extract_keys(A, keys)
    B = torch.empty_like(A)
    lsb_amd.sort(A, B, sync=True)
    del B
    assert _device_check(A, keys, n) == (0, -1)
    assert _torch_stable_check(A, keys)
    del A, keys
    torch.cuda.empty_cache()


def test_max_n_one_bucket_no_skip():
    """LSB_MAX_N equal keys, every pass run (LSB_NO_SKIP) by the one-tile
    form, whose look-back prefixes are global: the last tile publishes the
    largest status word there is, (n + 1) | inclusive flag = 0xFFFFFFFF."""
    n = lsb_amd.LSB_MAX_N
    v = [v for v, d, t, pf in variants() if d == 8 and pf == 0 and t == 4096][0]
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=256)  # val = index
    A[:, 0].fill_(0x5A5A5A5A5A5A5A5A)
    keys = torch.empty(n, dtype=torch.int64, device=DEV)
    lsb_amd.extract_keys(A, keys)
    B = torch.empty_like(A)
    lsb_amd.sort(A, B, sync=True, no_skip=True, variant=v)
    del B
    assert _device_check(A, keys, n) == (0, -1)
    del A, keys
    torch.cuda.empty_cache()


def test_autotune_table_and_sorts_after_tuning():
    """The autotuner times every variant per class; sorts that follow use the
    tuned choice and stay bit-exact."""
    try:
        table = lsb_amd.autotune(12, 18, key_bits=64)
        assert [e["log2n"] for e in table] == list(range(12, 19))
        ids = {v for v, *_ in variants()}
        for e in table:
            assert e["variant"] in ids and 0 < e["ms"] <= e["runner_up_ms"]
        for n in (5000, 70000, 250001):
            a = O.gen_uniform(n, ranks=2, seed_base=n)
            assert np.array_equal(gpu_sort(a), stable(a))
    finally:
        lsb_amd.tune_clear()


# ----------------------------------------------------------- key orders

def _signed_input(n, key_bits, seed):
    rng = np.random.default_rng(seed)
    a = np.zeros(n, dtype=O.ELEM)
    lo, hi = -(1 << (key_bits - 1)), (1 << (key_bits - 1))
    v = rng.integers(lo, hi, size=n, dtype=np.int64)
    q = n // 4
    v[:q] = v[q:2 * q]  # duplicates: stability matters
    mask = (1 << key_bits) - 1 if key_bits < 64 else None
    a["key"] = v.view(np.uint64) if mask is None else (v & mask).astype(np.uint64)
    a["val"] = np.arange(n, dtype=np.uint64)
    return a


def _float_input(n, key_bits, seed):
    rng = np.random.default_rng(seed)
    a = np.zeros(n, dtype=O.ELEM)
    f = rng.standard_normal(n) * 10.0 ** rng.integers(-30, 30, n)
    f[: n // 8] = -0.0
    f[n // 8: n // 4] = 0.0
    f[n // 4: n // 4 + 3] = [np.inf, -np.inf, np.nan]
    if key_bits == 32:
        a["key"] = f.astype(np.float32).view(np.uint32).astype(np.uint64)
    else:
        a["key"] = f.view(np.uint64)
    a["val"] = np.arange(n, dtype=np.uint64)
    return a


@pytest.mark.parametrize("order,key_bits", [("signed", 64), ("signed", 13), ("float", 64),
                                            ("float", 32)])
@pytest.mark.parametrize("n", [1000, 8192, 100003])
def test_key_orders_all_variants(order, key_bits, n):
    """Signed and float keys sort bit-exactly as the oracle's order, through
    the one-launch LDS sort and every compiled scatter variant."""
    make = _signed_input if order == "signed" else _float_input
    a = make(n, key_bits, seed=n + key_bits)
    want = O.stable_sort_order(a, key_bits, order)
    assert np.array_equal(gpu_sort(a, key_bits=key_bits, key_order=order), want)
    for v, d, tile, _ in variants():
        got = gpu_sort(a, key_bits=key_bits, key_order=order, variant=v)
        assert np.array_equal(got, want), (order, key_bits, n, v, d, tile)


def test_key_order_verifier_and_skip():
    """The device verifier checks the declared order (an unsigned-sorted
    signed array fails it), and pass skipping works on the mapped keys
    (non-negative signed values below 2^16: six of eight mapped digits are
    constant)."""
    n = 300_000
    rng = np.random.default_rng(11)
    a = np.zeros(n, dtype=O.ELEM)
    a["val"] = np.arange(n, dtype=np.uint64)
    B = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    a["key"] = rng.integers(0, 60000, size=n, dtype=np.int64).view(np.uint64)
    A = to_dev(a)
    lsb_amd.sort(A, B, sync=True, key_order="signed", profile=True)
    assert sum(lsb_amd.profile_read()["executed"]) == 2
    assert np.array_equal(from_dev(A), O.stable_sort_order(a, 64, "signed"))

    a["key"] = rng.integers(-1000, 1000, size=n, dtype=np.int64).view(np.uint64)
    keys = torch.from_numpy(a["key"].view(np.int64).copy()).to(DEV)
    A = to_dev(a)
    lsb_amd.sort(A, B, sync=True, key_order="signed")
    assert np.array_equal(from_dev(A), O.stable_sort_order(a, 64, "signed"))
    assert lsb_amd.verify(A, keys, n=n, key_order="signed")[0] == 0
    assert lsb_amd.verify(A, keys, n=n)[0] > 0  # not the unsigned order
    U = to_dev(a)
    lsb_amd.sort(U, B, sync=True)
    assert lsb_amd.verify(U, keys, n=n, key_order="signed")[0] > 0


def test_randomized_configurations():
    """Forty seeded draws of size (small, mid, or within two of a multiple of
    the tile), key width, key order, distribution, variant (or the automatic
    choice) and pass skipping, each bit-exact against the oracle's stable
    sort in that key order."""
    rng = np.random.default_rng(20261203)
    vs = variants()
    for case in range(40):
        sizes = [int(rng.integers(2, 70_000)), int(rng.integers(70_000, 400_000)),
                 8192 * int(rng.integers(1, 40)) + int(rng.integers(-2, 3))]
        n = sizes[int(rng.integers(0, 3))]
        order = ["unsigned", "unsigned", "signed", "float"][int(rng.integers(0, 4))]
        key_bits = [32, 64][int(rng.integers(0, 2))] if order == "float" else int(rng.integers(1, 65))
        dist = ["uniform", "zipf", "few", "ordered"][int(rng.integers(0, 4))]
        seed = int(rng.integers(0, 1 << 30))
        if order != "unsigned":
            n = max(n, 16)  # the float input plants three special values
        if order == "signed":
            a = _signed_input(n, key_bits, seed)
        elif order == "float":
            a = _float_input(n, key_bits, seed)
        elif dist == "uniform":
            a = O.gen_uniform(n, ranks=3, seed_base=seed)
        elif dist == "zipf":
            a = O.gen_zipf(n, ranks=2, m=1 << 12, seed_base=seed)
        elif dist == "few":
            r2 = np.random.default_rng(seed)
            a = np.zeros(n, dtype=O.ELEM)
            a["key"] = r2.integers(0, 5, n).astype(np.uint64) << np.uint64(int(r2.integers(0, 60)))
            a["val"] = np.arange(n, dtype=np.uint64)
        else:
            a = O.gen_ordered(n, reverse=bool(seed & 1))
        v = vs[int(rng.integers(0, len(vs)))][0] if rng.random() < 0.7 else 0
        no_skip = bool(rng.random() < 0.3)
        want = O.stable_sort_order(a, key_bits, order)
        got = gpu_sort(a, key_bits=key_bits, key_order=order, variant=v, no_skip=no_skip)
        assert np.array_equal(got, want), (case, n, order, key_bits, dist, v, no_skip)


@pytest.mark.parametrize("runs", [1, 16, 256])
def test_scatter_floor_probe_layout(runs):
    """lsb_debug_scatter_floor writes run r of tile t (8192 / runs
    consecutive elements) at dst[r * n / runs + t * 8192 / runs]; bad
    sizes and run counts are refused."""
    T, tiles = 8192, 24
    n = T * tiles
    src = torch.arange(2 * n, dtype=torch.int64, device=DEV).reshape(n, 2)
    dst = torch.full_like(src, -1)
    lsb_amd.debug_scatter_floor(src, dst, runs=runs)
    torch.cuda.synchronize()
    L = T // runs
    j = torch.arange(n, device=DEV)
    t, w = j // T, j % T
    want_pos = (w // L) * (n // runs) + t * L + (w % L)
    expect = torch.empty_like(src)
    expect[want_pos] = src
    assert torch.equal(dst, expect)
    for bad in (dict(runs=3), dict(runs=0), dict(runs=16384), dict(n=n - 1)):
        with pytest.raises(lsb_amd.LsbError):
            lsb_amd.debug_scatter_floor(src, dst, **bad)


def _forms_equal(A_in, **kw):
    """Sorts copies of A_in in the automatic form and in the LSD form and
    returns (automatic form, the two results equal bit for bit)."""
    n = A_in.shape[0]
    form = lsb_amd.sort_form(n, key_bits=kw.get("key_bits", 64))
    A1, A2 = A_in.clone(), A_in.clone()
    B = torch.empty_like(A1)
    lsb_amd.sort(A1, B, sync=True, **kw)
    lsb_amd.sort(A2, B, sync=True, lsd=True, **kw)
    same = torch.equal(A1, A2)
    del A1, A2, B
    return form, same


@pytest.mark.parametrize("case", ["uniform", "zipf", "few", "high_constant", "kb40",
                                  "signed", "float", "sorted"])
def test_plain_msd_form_equals_lsd_form(case):
    """From LSB_MSD_MIN_N elements on, lsb_sort sorts wider-than-32-bit keys
    most significant digit first; its result is the LSD form's bit for bit
    (which the oracle pins at smaller sizes), and both are the stable sort
    (device verifier + torch's stable sort)."""
    n = (1 << 24) + 12345
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    kw = {}
    if case == "uniform":
        lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=4)
    elif case == "zipf":
        lsb_amd.generate(A, kind=lsb_amd.GEN_ZIPF, ranks=4, zipf_m=1 << 16)
    elif case == "few":
        lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=4)
        A[:, 0] = torch.bitwise_and(A[:, 0], 3) << 40
    elif case == "high_constant":   # top 3 bytes one value: a pass-through level
        lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=4)
        A[:, 0] = torch.bitwise_or(torch.bitwise_and(A[:, 0], (1 << 40) - 1), 0x5A << 48)
    elif case == "kb40":
        lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=4)
        kw["key_bits"] = 40
    elif case == "signed":
        lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=4)
        A[:, 0] = A[:, 0] >> 20    # arithmetic: both signs, 44 significant bits
        kw["key_order"] = "signed"
    elif case == "float":
        g = torch.Generator(device=DEV).manual_seed(7)
        f = torch.randn(n, dtype=torch.float64, device=DEV, generator=g) * 1e6
        f[: n // 16] = 0.0
        A[:, 0] = f.view(torch.int64)
        A[:, 1] = torch.arange(n, dtype=torch.int64, device=DEV)
        kw["key_order"] = "float"
    else:
        lsb_amd.generate(A, kind=lsb_amd.GEN_SORTED)
    torch.cuda.synchronize()
    form, same = _forms_equal(A, **kw)
    assert form == "msd" and same, (case, form)
    if kw.get("key_order", "unsigned") == "unsigned" and "key_bits" not in kw:
        keys = torch.empty(n, dtype=torch.int64, device=DEV)
        lsb_amd.extract_keys(A, keys)
        B = torch.empty_like(A)
        lsb_amd.sort(A, B, sync=True)
        del B
        assert _device_check(A, keys, n) == (0, -1)
        assert _torch_stable_check(A, keys)
        del keys
    del A
    torch.cuda.empty_cache()


def test_plain_msd_reference_digest():
    """The reference's 4-rank 2^24 input sorted in the automatic (MSD) form
    reproduces the sorted digest of the reference's own output."""
    n = 1 << 24
    assert lsb_amd.sort_form(n) == "msd"
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    lsb_amd.generate(A, kind=lsb_amd.GEN_UNIFORM, ranks=4)
    B = torch.empty_like(A)
    lsb_amd.sort(A, B, sync=True)
    assert O.sha256(from_dev(A)) == \
        "edbb62f7b024eaade41bc12278e7cb240d038889a9780f8913a336352530d0c9"


@pytest.mark.parametrize("key_bits", [1, 16, 32, 33, 64])
def test_plain_msd_forced_small_sizes(key_bits):
    """lsd=False forces the MSD form at any size and key width: bit-exact
    against the oracle's stable sort, sizes around the run and tile
    capacity (8192) and the one-workgroup paths."""
    for n in (2, 1000, 8192, 8193, 65537, 100003, (1 << 20) + 7):
        a = O.gen_uniform(n, ranks=3, seed_base=n + key_bits)
        if n > 8193:
            a["key"][: n // 3] = a["key"][n // 3: 2 * (n // 3)]  # duplicates
        assert lsb_amd.sort_form(n, key_bits=key_bits, lsd=False) == "msd"
        got = gpu_sort(a, key_bits=key_bits, lsd=False)
        assert np.array_equal(got, stable(a, key_bits)), (n, key_bits)


def test_plain_msd_workspace_graph_and_profile():
    """The MSD form with a caller workspace of workspace_size(n) bytes (and
    too small a one refused), as a captured graph replayed on new data, and
    profiled: lsb_msd_profile_read's element counts follow the input
    (uniform 2^24 + 3: level 0 and level 1 each move every element, the
    finish's runs hold them all) and lsb_profile_read refuses, as
    lsb_msd_profile_read does after an LSD sort."""
    n = (1 << 24) + 3
    a = O.gen_uniform(n, ranks=4, seed_base=99)
    want = stable(a)
    A = to_dev(a)
    B = torch.empty_like(A)
    ws = torch.empty(lsb_amd.workspace_size(n), dtype=torch.uint8, device=DEV)
    lsb_amd.sort(A, B, workspace=ws, sync=True)
    assert np.array_equal(from_dev(A), want)
    small = torch.empty(lsb_amd.workspace_size(n, lsd=True) // 2, dtype=torch.uint8, device=DEV)
    with pytest.raises(lsb_amd.LsbError):
        lsb_amd.sort(A, B, workspace=small)
    del ws, small
    # graph: two replays on fresh inputs, eager work and a sync in between
    for seed in (5, 6, 7):
        b = O.gen_uniform(n, ranks=2, seed_base=seed)
        A.copy_(to_dev(b))
        lsb_amd.sort(A, B, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(A), stable(b)), seed
    # profile
    A.copy_(to_dev(a))
    lsb_amd.sort(A, B, profile=True, sync=True)
    p = lsb_amd.msd_profile_read()
    assert np.array_equal(from_dev(A), want)
    assert p["levels"] == 9 and p["level_elems"][0] == n and p["level_elems"][1] == n
    assert p["wide_steps"] == 3 and p["wide_elems"] == [0, 0, 0]
    assert p["finish_elems"] == n and p["runs"] > 0
    assert all(x == 0 for x in p["level_elems"][2:])
    parts = p["setup_ms"] + p["finish_ms"] + sum(map(sum, p["wide_ms"])) + \
        sum(map(sum, p["level_ms"]))
    assert p["total_ms"] > 0 and abs(parts - p["total_ms"]) < 0.05 * p["total_ms"] + 0.01
    with pytest.raises(lsb_amd.LsbError):
        lsb_amd.profile_read()
    A.copy_(to_dev(a))
    lsb_amd.sort(A, B, profile=True, sync=True, lsd=True)
    assert lsb_amd.profile_read()["total_ms"] > 0
    with pytest.raises(lsb_amd.LsbError):
        lsb_amd.msd_profile_read()
    assert np.array_equal(from_dev(A), want)
