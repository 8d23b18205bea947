"""GPU parity of the batched segmented sort: every segment of a batch must
equal numpy's stable sort of that segment alone, bit for bit, whatever the
other segments hold -- in eager mode and when replayed from a captured
hipGraph."""
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


def expected(a, offsets, key_bits=64):
    out = a.copy()
    m = np.uint64((1 << key_bits) - 1) if key_bits < 64 else None
    for s in range(len(offsets) - 1):
        lo, hi = offsets[s], offsets[s + 1]
        seg = a[lo:hi]
        k = seg["key"] if m is None else seg["key"] & m
        out[lo:hi] = seg[np.argsort(k, kind="stable")]
    return out


def batch(sizes, seed=0, kind="uniform"):
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(offsets[-1])
    if kind == "uniform":
        a = O.gen_uniform(n, ranks=7, seed_base=seed)
    elif kind == "zipf":
        a = O.gen_zipf(n, ranks=3, m=1 << 10, seed_base=seed)
    elif kind == "low32":
        a = O.gen_uniform(n, ranks=5, key_bits=32, seed_base=seed)
    else:
        raise ValueError(kind)
    return a, offsets


MIXED = [0, 1, 2, 3, 64, 1000, 2048, 2049, 4096, 8191, 8192, 8193, 9000, 20000, 70001,
         0, 5, 300000, 2047, 12345]


@pytest.mark.parametrize("kind", ["uniform", "zipf", "low32"])
def test_segmented_mixed_sizes(kind):
    a, off = batch(MIXED, seed=3, kind=kind)
    want = expected(a, off)
    # tile widths compiled for the segmented passes
    for d, tile in [(8, 4096), (8, 8192), (8, 2048), (11, 6144)]:
        A = to_dev(a)
        B = torch.empty_like(A)
        lsb_amd.sort_segmented(A, B, off, digit_bits=d, tile=tile, sync=True)
        assert np.array_equal(from_dev(A), want), (kind, d, tile)


@pytest.mark.parametrize("nseg", [64, 16])
def test_caller_workspace_sized_by_segmented_workspace_size(nseg):
    """A workspace of segmented_workspace_size(total, nseg) bytes holds the
    layout of whatever tile the sort picks for the batch: nseg segments of
    2^14 (small totals pick small tiles, whose layout is the larger one)."""
    sizes = [1 << 14] * nseg
    a, off = batch(sizes, seed=40 + nseg)
    want = expected(a, off)
    total = int(off[-1])
    wbytes = lsb_amd.segmented_workspace_size(total, nseg)
    assert wbytes > 0
    ws = torch.empty(wbytes, dtype=torch.uint8, device=DEV)
    for graph in (False, True):
        A = to_dev(a)
        B = torch.empty_like(A)
        lsb_amd.sort_segmented(A, B, off, sync=True, graph=graph, workspace=ws,
                               workspace_bytes=wbytes)
        assert np.array_equal(from_dev(A), want), graph
        As = [to_dev(a[int(off[i]):int(off[i + 1])]) for i in range(nseg)]
        Bs = [torch.empty_like(x) for x in As]
        lsb_amd.sort_batch(As, Bs, sync=True, graph=graph)
        for i, X in enumerate(As):
            assert np.array_equal(from_dev(X), want[int(off[i]):int(off[i + 1])])


def test_segmented_skips_and_parity():
    """Segments whose keys leave some or all digits in one bucket run an odd
    or even number of passes; each must still end in A."""
    sizes = [50000, 30000, 9000, 100000, 7000]
    a, off = batch(sizes, seed=9)
    a["key"][off[0]:off[1]] = 77                                         # no pass runs
    a["key"][off[1]:off[2]] &= np.uint64(0xFF)                           # one pass
    a["key"][off[2]:off[3]] &= np.uint64(0xFFFFFF)                       # three passes
    a["key"][off[3]:off[4]] = (a["key"][off[3]:off[4]] % np.uint64(5)) << np.uint64(56)
    want = expected(a, off)
    for no_skip in (False, True):
        A = to_dev(a)
        B = torch.empty_like(A)
        lsb_amd.sort_segmented(A, B, off, sync=True, no_skip=no_skip)
        assert np.array_equal(from_dev(A), want)


@pytest.mark.parametrize("key_bits", [8, 21, 32])
def test_segmented_key_bits(key_bits):
    a, off = batch([3000, 40000, 10000, 1], seed=4)
    want = expected(a, off, key_bits)
    A = to_dev(a)
    B = torch.empty_like(A)
    lsb_amd.sort_segmented(A, B, off, key_bits=key_bits, sync=True)
    assert np.array_equal(from_dev(A), want)


@pytest.mark.parametrize("order", ["signed", "float"])
def test_segmented_key_orders(order):
    """Signed / float keys per segment (LDS-sorted small segments and the
    tile pipeline alike) equal the oracle's order of each segment alone."""
    sizes = [3000, 40000, 10000, 1, 8192, 70001, 5]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(off[-1])
    rng = np.random.default_rng(5)
    a = np.zeros(n, dtype=O.ELEM)
    if order == "signed":
        a["key"] = rng.integers(-(1 << 40), 1 << 40, size=n, dtype=np.int64).view(np.uint64)
    else:
        f = rng.standard_normal(n) * 1e6
        f[::7] = -0.0
        a["key"] = f.view(np.uint64)
    a["val"] = np.arange(n, dtype=np.uint64)
    want = a.copy()
    for s_ in range(len(sizes)):
        lo, hi = int(off[s_]), int(off[s_ + 1])
        want[lo:hi] = O.stable_sort_order(a[lo:hi], 64, order)
    A = to_dev(a)
    B = torch.empty_like(A)
    lsb_amd.sort_segmented(A, B, off, key_order=order, sync=True)
    assert np.array_equal(from_dev(A), want)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_of_separate_buffers(graph):
    """lsb_sort_batch: arrays in separate buffers (tiny, small and tile-
    pipeline sizes, scratch only where needed) each come back exactly as a
    lone sort of that array, over several batches of different shapes (with
    graph=True the later ones replay one captured sequence)."""
    rng = np.random.default_rng(31)
    for rep in range(3):
        sizes = [int(x) for x in rng.choice([0, 1, 2, 700, 2048, 2049, 8192, 9000, 70001,
                                             300000], size=12)]
        arrs = [O.gen_uniform(n, ranks=2, seed_base=100 * rep + i) if n else
                np.zeros(0, dtype=O.ELEM) for i, n in enumerate(sizes)]
        As = [to_dev(a) if len(a) else torch.empty((0, 2), dtype=torch.int64, device=DEV)
              for a in arrs]
        Bs = [torch.empty_like(A) if len(a) > 8192 else None for a, A in zip(arrs, As)]
        lsb_amd.sort_batch(As, Bs, ns=sizes, graph=graph, sync=True)
        for a, A in zip(arrs, As):
            if len(a):
                assert np.array_equal(from_dev(A), a[np.argsort(a["key"], kind="stable")])


def test_segment_alone_equals_segment_in_batch():
    """A segment's result does not depend on its neighbours."""
    a, off = batch([20000, 5000, 90000, 3], seed=5)
    A = to_dev(a)
    B = torch.empty_like(A)
    lsb_amd.sort_segmented(A, B, off, sync=True)
    batched = from_dev(A)
    for s in range(len(off) - 1):
        lo, hi = int(off[s]), int(off[s + 1])
        A1 = to_dev(a[lo:hi])
        B1 = torch.empty_like(A1)
        lsb_amd.sort(A1, B1, sync=True)
        assert np.array_equal(from_dev(A1), batched[lo:hi])


@pytest.mark.parametrize("shape", ["dominant", "even", "two"])
def test_group_assignment_shapes(shape):
    """The scatter passes split large segments over 8 groups of blocks only
    when that balances; one dominant segment, two segments, or many even
    ones all sort every segment exactly."""
    rng = np.random.default_rng(5)
    if shape == "dominant":
        sizes = [600_000] + list(rng.integers(9000, 30000, size=20))
    elif shape == "even":
        sizes = [40_000] * 64
    else:
        sizes = [300_000, 250_000]
    a, off = batch(sizes, seed=len(sizes))
    n = int(off[-1])
    A = to_dev(a)
    B = torch.empty_like(A)
    for graph in (False, True):
        A.copy_(to_dev(a))
        lsb_amd.sort_segmented(A, B, off, graph=graph, sync=True)
        assert np.array_equal(from_dev(A[:n]), expected(a, off)), (shape, graph)


def test_graph_replay_many_batches():
    """One captured graph per buffer set, replayed for batches of different
    shapes, with eager work and synchronisations between replays."""
    cap = 400_000
    A = torch.empty((cap, 2), dtype=torch.int64, device=DEV)
    B = torch.empty_like(A)
    rng = np.random.default_rng(11)
    for rep in range(5):
        sizes = list(rng.integers(0, 40000, size=int(rng.integers(1, 12))))
        a, off = batch(sizes, seed=rep)
        n = int(off[-1])
        A[:n].copy_(to_dev(a))
        lsb_amd.sort_segmented(A, B, off, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(A[:n]), expected(a, off)), rep
        # unrelated eager work between replays
        x = torch.arange(1000, device=DEV).sum()
        assert int(x) == 499500


def test_plain_sort_graph_replay():
    n = 300_000
    a = O.gen_uniform(n, ranks=2)
    want = a[np.argsort(a["key"], kind="stable")]
    A = to_dev(a)
    B = torch.empty_like(A)
    src = A.clone()
    for rep in range(3):
        A.copy_(src)
        lsb_amd.sort(A, B, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(A), want), rep


@pytest.mark.parametrize("variant", [0, 1])
def test_plain_sort_graph_replay_changing_skips(variant):
    """One captured sequence replayed over inputs whose single-bucket digits
    differ (none, the low two, all but the top one, all): the plan, and for
    the region form the digit counted by the extra histogram read, are
    decided on the device at every replay."""
    n = 250_000
    base = O.gen_uniform(n, ranks=2)
    inputs = []
    inputs.append(base.copy())
    low = base.copy()
    low["key"] = (low["key"] << np.uint64(16)) | np.uint64(0x1234)
    inputs.append(low)
    top = base.copy()
    top["key"] &= np.uint64(0xFF00000000000000)
    inputs.append(top)
    same = base.copy()
    same["key"] = np.uint64(0xDEADBEEF)
    inputs.append(same)
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    B = torch.empty_like(A)
    for rep in range(2):
        for i, a in enumerate(inputs):
            A.copy_(to_dev(a))
            lsb_amd.sort(A, B, graph=True, variant=variant)
            torch.cuda.synchronize()
            want = a[np.argsort(a["key"], kind="stable")]
            assert np.array_equal(from_dev(A), want), (rep, i, variant)
            x = torch.arange(1000, device=DEV).sum()  # eager work between replays
            assert int(x) == 499500


def test_segmented_rejects_bad_offsets():
    A = torch.empty((100, 2), dtype=torch.int64, device=DEV)
    B = torch.empty_like(A)
    with pytest.raises(lsb_amd.LsbError):
        lsb_amd.sort_segmented(A, B, [0, 50, 40, 100])
    lsb_amd.sort_segmented(A, B, [0])            # no segments
    lsb_amd.sort_segmented(A, B, [0, 0, 1, 1])   # only trivial segments


def test_c4_batched_4096_segments_full_scale():
    """4096 segments, sizes floor(2^(10 + 12u)), u uniform, ~2^31 elements;
    each segment checked on the device with the O(n) invariants."""
    sizes = O.c4_sizes(4096, seed=2024)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    total = int(off[-1])
    A = torch.empty((total, 2), dtype=torch.int64, device=DEV)
    for s in range(len(sizes)):
        lo, n = int(off[s]), int(sizes[s])
        lsb_amd.generate(A[lo:lo + n], kind=lsb_amd.GEN_UNIFORM, ranks=1, seed_base=1000 + s,
                         val_base=0)
    keys = torch.empty(total, dtype=torch.int64, device=DEV)
    lsb_amd.extract_keys(A, keys)
    B = torch.empty_like(A)
    lsb_amd.sort_segmented(A, B, off, sync=True)
    del B
    for s in range(len(sizes)):
        lo, n = int(off[s]), int(sizes[s])
        bad, first = lsb_amd.verify(A[lo:lo + n], keys[lo:lo + n], n=n, val_base=0)
        assert bad == 0, (s, n, first)
    # and a few segments bit-exactly against the oracle
    for s in (0, 1, int(np.argmin(sizes))):
        lo, n = int(off[s]), int(sizes[s])
        want = O.gen_uniform(n, ranks=1, seed_base=1000 + s)
        want = want[np.argsort(want["key"], kind="stable")]
        assert np.array_equal(from_dev(A[lo:lo + n]), want)
    del A, keys
    torch.cuda.empty_cache()


def test_max_size_segment_between_small_ones():
    """A segment of LSB_MAX_N elements between a one-block segment and a
    multi-tile one: 2^31 + 74998 elements in all, so the last segment starts
    past element 2^31."""
    sizes = [5000, lsb_amd.LSB_MAX_N, 70000]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    total = int(off[-1])
    A = torch.empty((total, 2), dtype=torch.int64, device=DEV)
    for s, n in enumerate(sizes):
        lo = int(off[s])
        lsb_amd.generate(A[lo:lo + n], kind=lsb_amd.GEN_UNIFORM, ranks=64 if s == 1 else 1,
                         seed_base=7000 + s, val_base=0)
    keys = torch.empty(total, dtype=torch.int64, device=DEV)
    lsb_amd.extract_keys(A, keys)
    B = torch.empty_like(A)
    lsb_amd.sort_segmented(A, B, off, sync=True)
    del B
    for s, n in enumerate(sizes):
        lo = int(off[s])
        bad, first = lsb_amd.verify(A[lo:lo + n], keys[lo:lo + n], n=n, val_base=0)
        assert bad == 0, (s, n, first)
    for s in (0, 2):
        lo, n = int(off[s]), sizes[s]
        want = O.gen_uniform(n, ranks=1, seed_base=7000 + s)
        want = want[np.argsort(want["key"], kind="stable")]
        assert np.array_equal(from_dev(A[lo:lo + n]), want)
    del A, keys
    torch.cuda.empty_cache()


def test_randomized_batches():
    """Sixteen seeded batches: 1-150 segments of log-uniform sizes from 0 to
    2^18 (empty, single, LDS-sized and multi-tile ones mixed), random key
    width and key order, through lsb_sort_segmented (one buffer) or
    lsb_sort_batch (separate buffers), eager or graph-replayed; every segment
    equals the oracle's stable sort of that segment in that order."""
    rng = np.random.default_rng(777)
    for case in range(16):
        nseg = int(rng.integers(1, 151))
        sizes = [0 if rng.random() < 0.05 else int(2 ** rng.uniform(0, 18)) for _ in range(nseg)]
        order = ["unsigned", "unsigned", "signed", "float"][int(rng.integers(0, 4))]
        key_bits = [32, 64][int(rng.integers(0, 2))] if order == "float" else int(rng.integers(1, 65))
        graph = bool(rng.random() < 0.5)
        segs = []
        for s, n in enumerate(sizes):
            a = O.gen_uniform(n, ranks=1, seed_base=10_000 * case + s, key_bits=64)
            if n and rng.random() < 0.3:
                a["key"] = a["key"] % np.uint64(7)  # heavy duplicates: stability
            if order == "float" and key_bits == 32:
                a["key"] &= np.uint64(0xFFFFFFFF)
            segs.append(a)
        wants = [O.stable_sort_order(a, key_bits, order) for a in segs]
        if rng.random() < 0.5:
            off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
            total = int(off[-1])
            A = to_dev(np.concatenate(segs)) if total else torch.zeros((1, 2), dtype=torch.int64,
                                                                        device=DEV)
            B = torch.empty_like(A)
            lsb_amd.sort_segmented(A, B, off, key_bits=key_bits, key_order=order, sync=True,
                                   graph=graph)
            got = from_dev(A)[:total]
            for s in range(nseg):
                lo, hi = int(off[s]), int(off[s + 1])
                assert np.array_equal(got[lo:hi], wants[s]), (case, "segmented", s, hi - lo)
        else:
            As = [to_dev(a) if len(a) else torch.zeros((1, 2), dtype=torch.int64, device=DEV)
                  for a in segs]
            Bs = [torch.empty_like(x) if n > 8192 else None for x, n in zip(As, sizes)]
            lsb_amd.sort_batch(As, Bs, ns=sizes, key_bits=key_bits, key_order=order, sync=True,
                               graph=graph)
            for s in range(nseg):
                got = from_dev(As[s])[:sizes[s]]
                assert np.array_equal(got, wants[s]), (case, "batch", s, sizes[s])


def _msd_shapes(seed=0):
    """Segments that take every branch of the MSD-first form: ranges whose
    keys are all equal (in A, and as large buckets left in B), constant
    digits above or between varying ones (pass-through levels), heavy
    duplicates, buckets just above and below one workgroup's capacity,
    sorted and reversed input, and 2^20 / 2^22-element segments (two
    partition levels)."""
    rng = np.random.default_rng(seed)
    segs = []
    n = 9000
    a = O.gen_uniform(n, ranks=1, seed_base=seed + 1)
    a["key"] = np.uint64(0xABCDEF0123456789)                       # all equal
    segs.append(a)
    a = O.gen_uniform(200_000, ranks=1, seed_base=seed + 2)
    a["key"] = (a["key"] % np.uint64(3)) << np.uint64(56)          # top digit only: 3 big equal buckets
    segs.append(a)
    a = O.gen_uniform(150_000, ranks=1, seed_base=seed + 3)
    a["key"] = (a["key"] & np.uint64(0xFFFF)) | np.uint64(0x7700000000000000)  # low digits vary
    segs.append(a)
    a = O.gen_uniform(120_000, ranks=1, seed_base=seed + 4)
    a["key"] = ((a["key"] & np.uint64(0xFF)) << np.uint64(40)) | (a["key"] >> np.uint64(60))
    segs.append(a)                                                 # digit 5 and bits 0-3 vary
    a = O.gen_zipf(300_000, ranks=1, m=1 << 8, seed_base=seed + 5)  # heavy duplicates
    segs.append(a)
    a = O.gen_uniform(8193, ranks=1, seed_base=seed + 6)
    segs.append(a)
    a = O.gen_uniform(16384 * 256, ranks=4, seed_base=seed + 7)    # level-1 buckets of ~16384
    segs.append(a)
    a = O.gen_uniform(1 << 20, ranks=2, seed_base=seed + 8)
    a["key"] = np.arange(1 << 20, dtype=np.uint64) * np.uint64(0x9E3779B1)  # sorted
    segs.append(a)
    a = O.gen_uniform(70_000, ranks=1, seed_base=seed + 9)
    a["key"] = np.arange(70_000, 0, -1, dtype=np.uint64) << np.uint64(20)  # reversed
    segs.append(a)
    a = O.gen_uniform(1 << 22, ranks=8, seed_base=seed + 10)
    segs.append(a)
    rng.shuffle(segs)
    return segs


@pytest.mark.parametrize("key_order", ["unsigned", "signed"])
def test_msd_first_branches_equal_lsd_and_oracle(key_order):
    segs = _msd_shapes(3)
    sizes = [len(s) for s in segs]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    a = np.concatenate(segs)
    wants = [O.stable_sort_order(s, 64, key_order) for s in segs]
    outs = {}
    for lsd in (False, True):
        for graph in (False, True):
            A = to_dev(a)
            B = torch.empty_like(A)
            lsb_amd.sort_segmented(A, B, off, key_order=key_order, sync=True, lsd=lsd,
                                   graph=graph)
            outs[(lsd, graph)] = from_dev(A)
    for s in range(len(segs)):
        lo, hi = int(off[s]), int(off[s + 1])
        for k, got in outs.items():
            assert np.array_equal(got[lo:hi], wants[s]), (k, s, sizes[s])


@pytest.mark.parametrize("key_bits", [3, 12, 33, 57])
def test_msd_first_key_widths(key_bits):
    """Key widths whose top digit is narrower than 8 bits; bits above the
    key width are carried along but do not order."""
    sizes = [20_000, 9_000, 300_000, 8192, 100_000]
    a, off = batch(sizes, seed=key_bits)
    want = expected(a, off, key_bits)
    As = [to_dev(a[int(off[i]):int(off[i + 1])]) for i in range(len(sizes))]
    Bs = [torch.empty_like(x) for x in As]
    lsb_amd.sort_batch(As, Bs, key_bits=key_bits, sync=True, lsd=False)
    for i, X in enumerate(As):
        assert np.array_equal(from_dev(X), want[int(off[i]):int(off[i + 1])]), (key_bits, i)


def _wide_shapes(seed=0):
    """Segments of more than 2^21 and at most 2^23 elements, which are first
    partitioned on their top 10 key bits: uniform keys (buckets of about 2900
    elements), all keys equal, a constant top 10 bits (partitioned from the
    next 8-bit digit instead), keys that differ only in their top 11 bits
    (every bucket holds two keys), heavy duplicates, sorted and reversed
    input, the smallest size of the range, and a 2^24-element segment beside
    them that takes two 8-bit partitions."""
    segs = []
    a = O.gen_uniform(3_000_000, ranks=4, seed_base=seed + 1)
    segs.append(a)
    a = O.gen_uniform(1 << 24, ranks=16, seed_base=seed + 2)
    segs.append(a)
    a = O.gen_uniform(2_500_000, ranks=1, seed_base=seed + 3)
    a["key"] = np.uint64(0x0123456789ABCDEF)
    segs.append(a)
    a = O.gen_uniform(3_100_000, ranks=2, seed_base=seed + 4)
    a["key"] = (a["key"] & np.uint64((1 << 53) - 1)) | np.uint64(0x5AB << 53)
    segs.append(a)
    a = O.gen_uniform(2_200_000, ranks=2, seed_base=seed + 5)
    a["key"] = (a["key"] >> np.uint64(53)) << np.uint64(53)
    segs.append(a)
    a = O.gen_zipf(2_400_000, ranks=2, m=1 << 10, seed_base=seed + 6)
    a["key"] = (a["key"] << np.uint64(53)) | (a["key"] * np.uint64(40503) & np.uint64(0xFF))
    segs.append(a)
    a = O.gen_uniform((1 << 21) + 1, ranks=1, seed_base=seed + 7)
    a["key"] = np.arange((1 << 21) + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15 >> 22)
    segs.append(a)
    a = O.gen_uniform(2_300_000, ranks=1, seed_base=seed + 8)
    a["key"] = np.arange(2_300_000, 0, -1, dtype=np.uint64) << np.uint64(40)
    segs.append(a)
    return segs


def test_msd_wide_partition_shapes():
    segs = _wide_shapes(11)
    sizes = [len(s) for s in segs]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    a = np.concatenate(segs)
    wants = [O.stable_sort_order(s, 64, "unsigned") for s in segs]
    for graph in (False, True):
        A = to_dev(a)
        B = torch.empty_like(A)
        lsb_amd.sort_segmented(A, B, off, sync=True, graph=graph)
        got = from_dev(A)
        for s in range(len(segs)):
            lo, hi = int(off[s]), int(off[s + 1])
            assert np.array_equal(got[lo:hi], wants[s]), (graph, s, sizes[s])
    del A, B
    torch.cuda.empty_cache()


@pytest.mark.parametrize("key_bits", [16, 17, 33, 58, 59])
def test_msd_wide_partition_key_widths(key_bits):
    """The 10-bit digit's lowest bit falls at several offsets inside an
    8-bit digit (runs stay within aligned groups of 2 to 256 buckets); bits
    above the key width are carried along but do not order."""
    sizes = [2_200_000, 50_000, 3_000_000]
    a, off = batch(sizes, seed=100 + key_bits)
    a[: sizes[0]]["key"] = a[: sizes[0]]["key"] % np.uint64(1 << 20) * np.uint64(0x10001)
    want = expected(a, off, key_bits)
    A = to_dev(a)
    B = torch.empty_like(A)
    lsb_amd.sort_segmented(A, B, off, key_bits=key_bits, sync=True, lsd=False)
    got = from_dev(A)
    for s in range(len(sizes)):
        lo, hi = int(off[s]), int(off[s + 1])
        assert np.array_equal(got[lo:hi], want[lo:hi]), (key_bits, s)


@pytest.mark.parametrize("key_bits", [20, 32])
def test_automatic_form_for_narrow_keys(key_bits):
    """Keys of at most 32 bits take the LSD form unless MSD is asked for;
    all three calls give every segment its stable sort."""
    sizes = [9_000, 300_000, 2_500_000, 70_000]
    a, off = batch(sizes, seed=300 + key_bits, kind="zipf")
    want = expected(a, off, key_bits)
    for lsd in (None, True, False):
        A = to_dev(a)
        B = torch.empty_like(A)
        lsb_amd.sort_segmented(A, B, off, key_bits=key_bits, sync=True, lsd=lsd)
        got = from_dev(A)
        for s in range(len(sizes)):
            lo, hi = int(off[s]), int(off[s + 1])
            assert np.array_equal(got[lo:hi], want[lo:hi]), (key_bits, lsd, s)


def test_msd_wide_step_after_level_one():
    """Segments whose top byte is one value: level 0 finds it constant and
    level 1 partitions on the second byte.  In the first (9 M) that byte
    takes three values: three buckets of about 3 M elements go through wide
    step 2 (10-bit digit below the second byte) before they finish; in the
    second (9.5 M) one value holds 60 % (a wide-step bucket of 5.7 M) and the
    rest spread over the other 255 (level 2), in both key orders; the plain
    sort of the first in the MSD form moves every element through level 1
    and wide step 2."""
    rng = np.random.default_rng(6)
    a = O.gen_uniform(9_000_000, ranks=4, seed_base=43)
    second = np.array([0x10, 0x80, 0xF0], dtype=np.uint64)[rng.integers(0, 3, len(a))]
    a["key"] = (a["key"] & np.uint64((1 << 48) - 1)) | (second << np.uint64(48)) | \
        np.uint64(0x5A << 56)
    b = O.gen_uniform(9_500_000, ranks=2, seed_base=44)
    heavy = rng.random(len(b)) < 0.6
    b["key"] = (b["key"] & np.uint64((1 << 56) - 1)) | np.uint64(0xC3 << 56)
    b["key"][heavy] = (b["key"][heavy] & np.uint64((1 << 48) - 1)) | np.uint64(0xC377 << 48)
    segs = [a, b]
    sizes = [len(x) for x in segs]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    for order in ("unsigned", "signed"):
        wants = [O.stable_sort_order(x, 64, order) for x in segs]
        A = to_dev(np.concatenate(segs))
        B = torch.empty_like(A)
        lsb_amd.sort_segmented(A, B, off, key_order=order, sync=True, lsd=False, graph=True)
        got = from_dev(A)
        for s in range(len(segs)):
            lo, hi = int(off[s]), int(off[s + 1])
            assert np.array_equal(got[lo:hi], wants[s]), (order, s)
        del A, B
    A = to_dev(a)
    B = torch.empty_like(A)
    lsb_amd.sort(A, B, sync=True, lsd=False, profile=True)
    p = lsb_amd.msd_profile_read()
    assert np.array_equal(from_dev(A), O.stable_sort_order(a, 64, "unsigned"))
    n = len(a)
    assert p["level_elems"][0] == 0 and p["level_elems"][1] == n, p
    assert p["wide_elems"] == [0, 0, n] and p["finish_elems"] == n, p
    del A, B
    torch.cuda.empty_cache()


def test_msd_wide_step_after_level_zero():
    """A segment whose top byte takes three values: level 0 leaves three
    buckets of about 3 M elements, which go through the second wide step
    (10-bit digit just below the top byte) before they finish; beside it a
    segment (9.5 M) with one such bucket of 7.6 M among small ones, in both
    key orders."""
    rng = np.random.default_rng(5)
    a = O.gen_uniform(9_000_000, ranks=4, seed_base=41)
    top = np.array([0x10, 0x20, 0x30], dtype=np.uint64)[rng.integers(0, 3, len(a))]
    a["key"] = (a["key"] & np.uint64((1 << 56) - 1)) | (top << np.uint64(56))
    b = O.gen_uniform(9_500_000, ranks=2, seed_base=42)
    big = rng.random(len(b)) < 0.8
    b["key"][big] = (b["key"][big] & np.uint64((1 << 56) - 1)) | np.uint64(0x77 << 56)
    segs = [a, b]
    sizes = [len(x) for x in segs]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    for order in ("unsigned", "signed"):
        wants = [O.stable_sort_order(x, 64, order) for x in segs]
        A = to_dev(np.concatenate(segs))
        B = torch.empty_like(A)
        lsb_amd.sort_segmented(A, B, off, key_order=order, sync=True, lsd=False, graph=True)
        got = from_dev(A)
        for s in range(len(segs)):
            lo, hi = int(off[s]), int(off[s + 1])
            assert np.array_equal(got[lo:hi], wants[s]), (order, s)
        del A, B
    torch.cuda.empty_cache()
