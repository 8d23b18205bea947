"""ctypes wrapper over the CPU oracle (liblsb_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product library never
imports it.  The C file lsb_oracle.c documents which reference lines each
function restates (mpi/mpi_lsbsort.cpp keygen 676-704, localShuffle 225-259,
mySort 591-597, verify 755-784).

Arrays of elements are numpy structured arrays with dtype ELEM
({'key': <u8, 'val': <u8}, 16 bytes, key at offset 0), the layout of
SortElement (mpi/mpi_lsbsort.cpp:30-33).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsb_oracle.so")
REF_BIN = os.path.join(HERE, "_ref", "mpi_lsbsort")

ELEM = np.dtype([("key", "<u8"), ("val", "<u8")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        p, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
        L.orc_pcg64_fill.argtypes = [u64, u64, u64, p]
        L.orc_gen_uniform.argtypes = [p, u64, u32, u32, u64, u64, u64]
        L.orc_gen_ordered.argtypes = [p, u64, ctypes.c_int]
        L.orc_zipf_cdf.argtypes = [p, u32, ctypes.c_double]
        L.orc_gen_zipf.argtypes = [p, u64, u32, u64, p, p, u32, u64, u64, u32]
        L.orc_histograms.argtypes = [p, u64, u32, u32, p]
        L.orc_counting_pass.argtypes = [p, p, u64, u32, u32, u32]
        L.orc_lsd_sort.argtypes = [p, p, u64, u32, u32]
        L.orc_stable_sort.argtypes = [p, u64, u32]
        L.orc_check.argtypes = [p, p, u64, u64, u32, ctypes.POINTER(i32)]
        L.orc_check.restype = ctypes.c_int64
        L.orc_sha256.argtypes = [p, u64, p]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def pcg64(seed, count, skip=0):
    out = np.empty(count, dtype=np.uint64)
    lib().orc_pcg64_fill(seed, skip, count, _ptr(out))
    return out


def gen_uniform(n, ranks=1, key_bits=64, seed_base=0, first=0, count=None):
    """Reference keygen: rank r draws ceil(n/R) keys from pcg64(seed_base+r)."""
    count = n - first if count is None else count
    out = np.empty(count, dtype=ELEM)
    lib().orc_gen_uniform(_ptr(out), n, ranks, key_bits, seed_base, first, count)
    return out


def gen_ordered(n, reverse=False):
    out = np.empty(n, dtype=ELEM)
    lib().orc_gen_ordered(_ptr(out), n, 1 if reverse else 0)
    return out


def zipf_cdf(m, s):
    cdf = np.empty(m, dtype=np.float64)
    lib().orc_zipf_cdf(_ptr(cdf), m, s)
    return cdf


def gen_zipf(n, ranks=1, seed_base=0, m=1 << 20, s=1.1, table_seed=12345,
             first=0, count=None, key_bits=64):
    count = n - first if count is None else count
    cdf = zipf_cdf(m, s)
    table = pcg64(table_seed, m)
    out = np.empty(count, dtype=ELEM)
    lib().orc_gen_zipf(_ptr(out), n, ranks, seed_base, _ptr(cdf), _ptr(table),
                       m, first, count, key_bits)
    return out


def c4_sizes(nseg, seed=2024, lo_log2=10, span_log2=12):
    """Batched-sort segment sizes: floor(2^(lo + span*u)), u = 53-bit uniform
    from output s of pcg64(seed)."""
    u = (pcg64(seed, nseg) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.floor(2.0 ** (lo_log2 + span_log2 * u)).astype(np.int64)


def n_digits(key_bits, d):
    return (key_bits + d - 1) // d


def histograms(a, key_bits=64, d=8):
    a = np.ascontiguousarray(a, dtype=ELEM)
    P = n_digits(key_bits, d)
    counts = np.empty(P * (1 << d), dtype=np.uint32)
    lib().orc_histograms(_ptr(a), len(a), key_bits, d, _ptr(counts))
    return counts.reshape(P, 1 << d)


def counting_pass(a, p, d=8, key_bits=64):
    a = np.ascontiguousarray(a, dtype=ELEM)
    out = np.empty_like(a)
    lib().orc_counting_pass(_ptr(a), _ptr(out), len(a), p, d, key_bits)
    return out


def lsd_sort(a, key_bits=64, d=8):
    A = np.array(a, dtype=ELEM, copy=True)
    B = np.empty_like(A)
    lib().orc_lsd_sort(_ptr(A), _ptr(B), len(A), key_bits, d)
    return A


def stable_sort(a, key_bits=64):
    A = np.array(a, dtype=ELEM, copy=True)
    lib().orc_stable_sort(_ptr(A), len(A), key_bits)
    return A


def key_order_fwd(keys, key_bits=64, order="unsigned"):
    """Maps keys (uint64 array) to unsigned integers of the same order:
    "signed" = two's complement of key_bits bits (flip the sign bit),
    "float" = IEEE-754 binary64 / binary32-in-the-low-32-bits (negative:
    flip all key_bits bits; positive: flip the sign bit).  The reference's
    getDigit carries an unused `negs` argument for this
    (chpl/arkouda-radix-sort.chpl:37-39).  Bits above key_bits pass through."""
    k = np.asarray(keys, dtype=np.uint64)
    if order == "unsigned":
        return k.copy()
    sign = np.uint64(1 << (key_bits - 1))
    full = np.uint64((1 << key_bits) - 1) if key_bits < 64 else np.uint64(~0 & (2**64 - 1))
    if order == "signed":
        return k ^ sign
    if order == "float":
        neg = (k & sign) != 0
        return np.where(neg, k ^ full, k ^ sign)
    raise ValueError(order)


def stable_sort_order(a, key_bits=64, order="unsigned"):
    """Stable sort of elements a by the low key_bits of key_order_fwd(key)."""
    a = np.asarray(a, dtype=ELEM)
    mask = np.uint64((1 << key_bits) - 1) if key_bits < 64 else np.uint64(2**64 - 1)
    t = key_order_fwd(a["key"], key_bits, order) & mask
    return a[np.argsort(t, kind="stable")]


CHECK_REASONS = {1: "val out of range", 2: "duplicate val (not a permutation)",
                 3: "key does not match its source", 4: "keys out of order",
                 5: "equal keys out of index order (unstable)"}


def check(out, in_keys, val_base=0, key_bits=64):
    """O(n) proof that out is the stable sort of the input whose key at
    original index i is in_keys[i] and whose vals were val_base + i.
    Returns None on success, else (position, reason)."""
    out = np.ascontiguousarray(out, dtype=ELEM)
    in_keys = np.ascontiguousarray(in_keys, dtype=np.uint64)
    why = ctypes.c_int32(0)
    bad = lib().orc_check(_ptr(out), _ptr(in_keys), len(out), val_base, key_bits,
                          ctypes.byref(why))
    if bad < 0:
        return None
    return int(bad), CHECK_REASONS.get(why.value, str(why.value))


def sha256(a):
    a = np.ascontiguousarray(a)
    out = np.empty(32, dtype=np.uint8)
    lib().orc_sha256(_ptr(a), a.nbytes, _ptr(out))
    return out.tobytes().hex()
