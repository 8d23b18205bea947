"""lsb_sort_noalloc, the entry point that allocates no device memory, gives
exactly what lsb_sort gives on the same input, call after call, although
every call shares the library's one scratch array per device and the
previous call left its counters, tickets and status words there."""
import ctypes

import numpy as np
import pytest
import torch

import lsb_amd
import oracle as O

DEV = "cuda:0"


def _noalloc_rc(A, B, n, **kw):
    c = lsb_amd.make_config(stream=0, **kw)
    return lsb_amd.lib().lsb_sort_noalloc(A, B, n, ctypes.byref(c))


def test_noalloc_argument_errors_need_no_gpu():
    a, b = 0x10000, 0x20000000  # never dereferenced: every case fails before HIP
    assert _noalloc_rc(None, None, 0) == 0
    assert _noalloc_rc(a, b, 1) == 0
    assert _noalloc_rc(a, b, lsb_amd.LSB_NOALLOC_MAX_N + 1) == lsb_amd.LSB_ERANGE
    assert _noalloc_rc(a, b, 5, flags=lsb_amd.LSB_GRAPH) == lsb_amd.LSB_EINVAL
    assert _noalloc_rc(a, b, 5, flags=lsb_amd.LSB_PROFILE) == lsb_amd.LSB_EINVAL
    assert _noalloc_rc(a, b, 5, workspace=0x30000000, workspace_bytes=1 << 20) == \
        lsb_amd.LSB_EINVAL
    assert _noalloc_rc(a + 8, b, 5) == lsb_amd.LSB_EINVAL
    # the MSD form stages its range table in pinned host memory: refused
    assert _noalloc_rc(a, b, 5, flags=lsb_amd.LSB_FORM_MSD) == lsb_amd.LSB_EINVAL
    # an explicit variant whose workspace exceeds the scratch array is refused
    small_tile = [v for v, d, t, p in lsb_amd.variants() if d == 8 and t == 2048 and p == 0][0]
    assert _noalloc_rc(a, 1 << 40, lsb_amd.LSB_NOALLOC_MAX_N, variant=small_tile) == \
        lsb_amd.LSB_ERANGE
    assert lsb_amd.noalloc_scratch_bytes() >= lsb_amd.workspace_size(
        lsb_amd.LSB_NOALLOC_MAX_N, digit_bits=8, tile=8192)


def _gen(n, kind, key_bits, seed):
    A = torch.empty((n, 2), dtype=torch.int64, device=DEV)
    lsb_amd.generate(A, kind=kind, ranks=4, key_bits=key_bits, seed_base=seed)
    return A


def _both(pristine, key_bits, **kw):
    """(lsb_sort result, lsb_sort_noalloc result) of the same input."""
    want = pristine.clone()
    B = torch.empty_like(want)
    lsb_amd.sort(want, B, key_bits=key_bits, sync=True, **kw)
    got = pristine.clone()
    B.fill_(-1)  # the scratch buffer's old contents must not matter either
    lsb_amd.sort_noalloc(got, B, key_bits=key_bits, sync=True, **kw)
    return want, got


@pytest.mark.gpu
def test_noalloc_equals_lsb_sort_twice_then_smaller():
    # first call: a large array (8192-element tiles, every pass executed)
    print("noalloc call 1: n=2^24+12345 uniform 64-bit", flush=True)
    p1 = _gen((1 << 24) + 12345, lsb_amd.GEN_UNIFORM, 64, 11)
    w1, g1 = _both(p1, 64)
    assert torch.equal(w1, g1)
    keys = torch.empty(p1.shape[0], dtype=torch.int64, device=DEV)
    lsb_amd.extract_keys(p1, keys)
    assert lsb_amd.verify(g1, keys) == (0, -1)
    del w1, g1, keys
    # second call right after, smaller, other tile size and skipped digits:
    # whatever the first call left in the scratch array is not read
    print("noalloc call 2: n=300001 Zipf 32-bit", flush=True)
    n2 = 300001
    a2 = O.gen_zipf(n2, ranks=2, m=1 << 12, seed_base=5)
    a2["key"] &= np.uint64(0xFFFFFFFF)
    p2 = torch.from_numpy(a2.view(np.int64).reshape(n2, 2).copy()).to(DEV)
    w2, g2 = _both(p2, 32)
    assert torch.equal(w2, g2)
    got = g2.cpu().numpy().reshape(-1).view(O.ELEM)
    want = a2[np.argsort(a2["key"], kind="stable")]
    assert np.array_equal(got, want)
    # third: one workgroup's worth, the in-LDS path
    print("noalloc call 3: n=5000 uniform 64-bit, signed order", flush=True)
    a3 = O.gen_uniform(5000, ranks=1, seed_base=9)
    p3 = torch.from_numpy(a3.view(np.int64).reshape(5000, 2).copy()).to(DEV)
    w3, g3 = _both(p3, 64, key_order="signed")
    assert torch.equal(w3, g3)


@pytest.mark.gpu
def test_noalloc_largest_size_and_repeat():
    """LSB_NOALLOC_MAX_N elements (the scratch array's limit), twice in a row
    on different inputs, each equal to lsb_sort's result."""
    n = lsb_amd.LSB_NOALLOC_MAX_N
    for seed, kb in ((21, 64), (22, 40)):
        print(f"noalloc max size: n=2^27 key_bits={kb} seed={seed}", flush=True)
        p = _gen(n, lsb_amd.GEN_UNIFORM, kb, seed)
        w, g = _both(p, kb)
        assert torch.equal(w, g)
        del p, w, g
        torch.cuda.empty_cache()
