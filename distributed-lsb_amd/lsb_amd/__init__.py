"""Python binding of the C ABI in include/lsb.h (liblsb_amd.so).

This is the host-side mirror of the reference's per-rank sort entry
``mySort(A, B)`` (mpi/mpi_lsbsort.cpp:591-597): ``sort(A, B)`` sorts A stably
by key using B as scratch.  A and B are torch CUDA (HIP) tensors of dtype
int64 and shape [n, 2] (or any contiguous 16-byte-per-element buffer): column
0 is the key, column 1 the value, bit-identical to the reference's
``struct SortElement {uint64_t key; uint64_t val;}`` (30-33).

PyTorch only provides device memory and streams here.  Every computation
runs in the HIP kernels of liblsb_amd.so; if the library is missing, loading
raises instead of falling back to anything else.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "liblsb_amd.so")

LSB_OK, LSB_EINVAL, LSB_ENOMEM, LSB_EHIP, LSB_ERANGE, LSB_EBUSY, LSB_ENOTFOUND = (
    0, -1, -2, -3, -4, -5, -6)
LSB_SYNC, LSB_NO_SKIP, LSB_PROFILE, LSB_GRAPH = 1, 2, 4, 8
LSB_KEYS_SIGNED, LSB_KEYS_FLOAT = 16, 32
LSB_FORM_LSD = 64
LSB_FORM_MSD = 128
LSB_SEG_LSD = LSB_FORM_LSD
LSB_SEG_MSD = LSB_FORM_MSD
LSB_MSD_MIN_N = 1 << 24
LSB_MSD_MAX_LEVELS = 9
LSB_MSD_WIDE_STEPS = 3
# key_order names accepted by the sort, verify and server calls
KEY_ORDERS = {"unsigned": 0, "signed": LSB_KEYS_SIGNED, "float": LSB_KEYS_FLOAT}
LSB_MAX_N = 0x7FFFFFFE
LSB_NOALLOC_MAX_N = 1 << 27
GEN_UNIFORM, GEN_SORTED, GEN_REVERSE, GEN_ZIPF = 0, 1, 2, 3


class LsbError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = strerror(code) if _lib is not None else str(code)
        if code == LSB_EHIP and _lib is not None:
            msg += f" (hipError {_lib.lsb_last_hip_error()})"
        super().__init__(f"{what}: {msg} [{code}]" if what else f"{msg} [{code}]")


class Config(ctypes.Structure):
    """struct lsb_config."""
    _fields_ = [("key_bits", ctypes.c_uint32), ("digit_bits", ctypes.c_uint32),
                ("tile", ctypes.c_uint32), ("variant", ctypes.c_uint32),
                ("stream", ctypes.c_void_p), ("workspace", ctypes.c_void_p),
                ("workspace_bytes", ctypes.c_size_t), ("flags", ctypes.c_uint32),
                ("reserved1", ctypes.c_uint32)]


class GenSpec(ctypes.Structure):
    """struct lsb_gen_spec."""
    _fields_ = [("kind", ctypes.c_uint32), ("ranks", ctypes.c_uint32),
                ("key_bits", ctypes.c_uint32), ("zipf_m", ctypes.c_uint32),
                ("zipf_s", ctypes.c_double), ("zipf_table_seed", ctypes.c_uint64),
                ("seed_base", ctypes.c_uint64), ("val_base", ctypes.c_uint64)]


class Profile(ctypes.Structure):
    """struct lsb_profile."""
    _fields_ = [("passes", ctypes.c_uint32), ("executed_mask", ctypes.c_uint32),
                ("digit_bits", ctypes.c_uint32), ("tile", ctypes.c_uint32),
                ("histogram_ms", ctypes.c_float), ("scan_ms", ctypes.c_float),
                ("copy_ms", ctypes.c_float), ("total_ms", ctypes.c_float),
                ("pass_ms", ctypes.c_float * 16)]


class MsdProfile(ctypes.Structure):
    """struct lsb_msd_profile."""
    _fields_ = [("levels", ctypes.c_uint32), ("wide_steps", ctypes.c_uint32),
                ("runs", ctypes.c_uint32),
                ("wide_elems", ctypes.c_uint32 * LSB_MSD_WIDE_STEPS),
                ("level_elems", ctypes.c_uint32 * LSB_MSD_MAX_LEVELS),
                ("finish_elems", ctypes.c_uint32), ("setup_ms", ctypes.c_float),
                ("wide_ms", (ctypes.c_float * 3) * LSB_MSD_WIDE_STEPS),
                ("level_ms", (ctypes.c_float * 3) * LSB_MSD_MAX_LEVELS),
                ("finish_ms", ctypes.c_float), ("total_ms", ctypes.c_float)]


class ServerConfig(ctypes.Structure):
    """struct lsb_server_config."""
    _fields_ = [("memory_budget", ctypes.c_size_t), ("max_pending", ctypes.c_uint32),
                ("batch_max_elems", ctypes.c_uint32), ("batch_max_total", ctypes.c_uint64),
                ("streams", ctypes.c_uint32), ("use_graphs", ctypes.c_uint32),
                ("batch_window_us", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    """struct lsb_server_stats."""
    _fields_ = [("jobs_done", ctypes.c_uint64), ("elems_done", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("batched_jobs", ctypes.c_uint64),
                ("plain_jobs", ctypes.c_uint64), ("memory_in_use", ctypes.c_uint64),
                ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double),
                ("max_us", ctypes.c_double), ("mean_us", ctypes.c_double),
                ("span_s", ctypes.c_double), ("pool_plain_jobs", ctypes.c_uint64),
                ("admission_waits", ctypes.c_uint64)]


class TuneEntry(ctypes.Structure):
    """struct lsb_tune_entry."""
    _fields_ = [("log2n", ctypes.c_uint32), ("key_bits", ctypes.c_uint32),
                ("digit_bits", ctypes.c_uint32), ("tile", ctypes.c_uint32),
                ("persistent", ctypes.c_uint32), ("variant", ctypes.c_uint32),
                ("ms", ctypes.c_float), ("runner_up_ms", ctypes.c_float)]


_lib = None

# Every function include/lsb.h declares, with its ctypes signature.
_P, _SZ, _U32, _U64, _I = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32,
                           ctypes.c_uint64, ctypes.c_int)
SIGNATURES = {
    "lsb_config_init": (None, [ctypes.POINTER(Config)]),
    "lsb_sort": (_I, [_P, _P, _SZ, ctypes.POINTER(Config)]),
    "lsb_sort_noalloc": (_I, [_P, _P, _SZ, ctypes.POINTER(Config)]),
    "lsb_noalloc_scratch_bytes": (_SZ, []),
    "lsb_workspace_size": (_SZ, [_SZ, ctypes.POINTER(Config)]),
    "lsb_sort_form": (_I, [_SZ, ctypes.POINTER(Config)]),
    "lsb_plan": (_I, [_SZ, ctypes.POINTER(Config), ctypes.POINTER(_U32),
                      ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lsb_variants": (_I, [_P, _P, _P, _I]),
    "lsb_strerror": (ctypes.c_char_p, [_I]),
    "lsb_version": (ctypes.c_char_p, []),
    "lsb_last_hip_error": (_I, []),
    "lsb_sort_segmented": (_I, [_P, _P, ctypes.POINTER(_U64), _U32, ctypes.POINTER(Config)]),
    "lsb_segmented_workspace_size": (_SZ, [_SZ, _U32, ctypes.POINTER(Config)]),
    "lsb_sort_batch": (_I, [_P, _P, _P, _U32, ctypes.POINTER(Config)]),
    "lsb_release": (_I, [_P]),
    "lsb_server_config_init": (None, [ctypes.POINTER(ServerConfig)]),
    "lsb_server_create": (_I, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]),
    "lsb_server_destroy": (_I, [_P]),
    "lsb_server_submit": (_I, [_P, _P, _SZ, _U32, _U32, ctypes.POINTER(_U64)]),
    "lsb_server_submit_batch": (_I, [_P, _P, _P, _P, _P, _SZ, _P]),
    "lsb_server_poll": (_I, [_P, _U64]),
    "lsb_server_wait": (_I, [_P, _U64, ctypes.c_double]),
    "lsb_server_drain": (_I, [_P]),
    "lsb_server_latency": (_I, [_P, _U64, ctypes.POINTER(ctypes.c_double)]),
    "lsb_server_stats_get": (_I, [_P, ctypes.POINTER(ServerStats)]),
    "lsb_server_reset_stats": (_I, [_P]),
    "lsb_autotune": (_I, [_U32, _U32, _U32, _P]),
    "lsb_tune_table": (_I, [ctypes.POINTER(TuneEntry), _I]),
    "lsb_tune_save": (_I, [ctypes.c_char_p]),
    "lsb_tune_load": (_I, [ctypes.c_char_p]),
    "lsb_tune_clear": (_I, []),
    "lsb_profile_read": (_I, [ctypes.POINTER(Profile)]),
    "lsb_msd_profile_read": (_I, [ctypes.POINTER(MsdProfile)]),
    "lsb_debug_histogram": (_I, [_P, _SZ, ctypes.POINTER(Config), _P]),
    "lsb_generate": (_I, [_P, _SZ, ctypes.POINTER(GenSpec), _P]),
    "lsb_verify": (_I, [_P, _P, _SZ, _U64, _U32, _U32, ctypes.POINTER(_U64),
                        ctypes.POINTER(ctypes.c_int64), _P]),
    "lsb_extract_keys": (_I, [_P, _SZ, _P, _P]),
    "lsb_debug_scatter_floor": (_I, [_P, _P, _SZ, _U32, _P]),
}


def lib():
    """Loads liblsb_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"liblsb_amd.so not built at {LIB_PATH}: run `make -C distributed-lsb_amd` "
                "or __graft_entry__.build(); there is no fallback path")
        # PyTorch-ROCm ships its own libamdhip64 and links it by file name,
        # so a process that maps /opt/rocm's copy first (our NEEDED entry)
        # and imports torch afterwards ends up with two HIP runtimes, and
        # allocations fail.  Loading torch first makes this library bind to
        # torch's runtime through the shared soname.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != LSB_OK:
        raise LsbError(rc, what)


def strerror(code):
    return lib().lsb_strerror(code).decode()


def version():
    return lib().lsb_version().decode()


def variants():
    """[(variant id, digit_bits, tile, persistent)] of the compiled scatter
    kernels; pass the id as sort(..., variant=id)."""
    n = lib().lsb_variants(None, None, None, 0)
    d = (ctypes.c_uint32 * n)()
    t = (ctypes.c_uint32 * n)()
    p = (ctypes.c_uint32 * n)()
    lib().lsb_variants(d, t, p, n)
    return [(i + 1, d[i], t[i], p[i]) for i in range(n)]


def make_config(key_bits=64, digit_bits=0, tile=0, stream=None, flags=0,
                workspace=None, workspace_bytes=0, variant=0):
    c = Config()
    lib().lsb_config_init(ctypes.byref(c))
    c.key_bits = key_bits
    c.digit_bits = digit_bits
    c.tile = tile
    c.variant = variant
    c.stream = _stream_ptr(stream)
    c.flags = flags
    if workspace is not None:
        c.workspace = workspace if isinstance(workspace, int) else workspace.data_ptr()
        c.workspace_bytes = workspace_bytes or (
            workspace.numel() * workspace.element_size() if hasattr(workspace, "numel") else 0)
    return c


def _stream_ptr(stream):
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available():
                return torch.cuda.current_stream().cuda_stream
        except ImportError:
            pass
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _nelem(x):
    if hasattr(x, "numel"):
        return x.numel() * x.element_size() // 16
    raise TypeError("pass n explicitly for raw pointers")


def plan(n, key_bits=64, digit_bits=0, tile=0, variant=0):
    c = make_config(key_bits, digit_bits, tile, stream=0, variant=variant)
    d, t, p = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    _check(lib().lsb_plan(n, ctypes.byref(c), ctypes.byref(d), ctypes.byref(t),
                          ctypes.byref(p)), "lsb_plan")
    return d.value, t.value, p.value


def workspace_size(n, key_bits=64, digit_bits=0, tile=0, variant=0, lsd=None):
    c = make_config(key_bits, digit_bits, tile, stream=0, flags=_form_flag(lsd),
                    variant=variant)
    return lib().lsb_workspace_size(n, ctypes.byref(c))


def sort_form(n, key_bits=64, digit_bits=0, tile=0, variant=0, lsd=None, no_skip=False):
    """The form sort() takes for n elements: "lsd" or "msd"."""
    c = make_config(key_bits, digit_bits, tile, stream=0,
                    flags=_form_flag(lsd) | (LSB_NO_SKIP if no_skip else 0), variant=variant)
    rc = lib().lsb_sort_form(n, ctypes.byref(c))
    _check(rc if rc < 0 else LSB_OK, "lsb_sort_form")
    return "msd" if rc == LSB_FORM_MSD else "lsd"


def _key_flags(key_order):
    if key_order not in KEY_ORDERS:
        raise ValueError(f"key_order must be one of {sorted(KEY_ORDERS)}")
    return KEY_ORDERS[key_order]


def sort(A, B, n=None, key_bits=64, digit_bits=0, tile=0, stream=None, sync=False,
         no_skip=False, workspace=None, workspace_bytes=0, profile=False, graph=False,
         variant=0, key_order="unsigned", lsd=None):
    """mySort(A, B): stable sort of A[0:n] by key, B is scratch, result in A.
    key_order: "unsigned" (default), "signed" or "float" keys.  lsd=True
    runs the LSD passes, lsd=False most significant digit first, None the
    library's choice (sort_form tells which)."""
    n = _nelem(A) if n is None else n
    flags = ((LSB_SYNC if sync else 0) | (LSB_NO_SKIP if no_skip else 0)
             | (LSB_PROFILE if profile else 0) | (LSB_GRAPH if graph else 0)
             | _key_flags(key_order) | _form_flag(lsd))
    c = make_config(key_bits, digit_bits, tile, stream, flags, workspace, workspace_bytes,
                    variant)
    _check(lib().lsb_sort(_ptr(A), _ptr(B), n, ctypes.byref(c)), "lsb_sort")


def sort_noalloc(A, B, n=None, key_bits=64, digit_bits=0, tile=0, stream=None, sync=False,
                 no_skip=False, variant=0, key_order="unsigned"):
    """sort() without any device allocation at call time (lsb_sort_noalloc):
    the workspace is the library's fixed per-device scratch array, so
    n <= LSB_NOALLOC_MAX_N and calls that overlap on one device race."""
    n = _nelem(A) if n is None else n
    flags = ((LSB_SYNC if sync else 0) | (LSB_NO_SKIP if no_skip else 0)
             | _key_flags(key_order))
    c = make_config(key_bits, digit_bits, tile, stream, flags, variant=variant)
    _check(lib().lsb_sort_noalloc(_ptr(A), _ptr(B), n, ctypes.byref(c)), "lsb_sort_noalloc")


def noalloc_scratch_bytes():
    return lib().lsb_noalloc_scratch_bytes()


def _form_flag(lsd):
    """Sort form: None (the library's choice), True (LSD), False (MSD first)."""
    if lsd is None:
        return 0
    return LSB_FORM_LSD if lsd else LSB_FORM_MSD


def sort_segmented(A, B, offsets, key_bits=64, digit_bits=0, tile=0, stream=None, sync=False,
                   no_skip=False, graph=False, workspace=None, workspace_bytes=0,
                   key_order="unsigned", lsd=None):
    """Sorts every segment [offsets[s], offsets[s+1]) of A independently and
    stably by key; B is scratch; results in A.  offsets: host sequence of
    nseg + 1 non-decreasing element offsets.  lsd=True sorts the large
    segments with LSD passes, lsd=False most significant digit first, None
    the library's choice (MSD first for keys wider than 32 bits)."""
    import numpy as _np
    off = _np.ascontiguousarray(_np.asarray(offsets, dtype=_np.uint64))
    nseg = len(off) - 1
    flags = ((LSB_SYNC if sync else 0) | (LSB_NO_SKIP if no_skip else 0)
             | (LSB_GRAPH if graph else 0) | _form_flag(lsd) | _key_flags(key_order))
    c = make_config(key_bits, digit_bits, tile, stream, flags, workspace, workspace_bytes)
    _check(lib().lsb_sort_segmented(_ptr(A), _ptr(B),
                                    off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                    max(nseg, 0), ctypes.byref(c)), "lsb_sort_segmented")


def release(stream=None):
    """Frees the library's cached workspaces and graphs for sorts on
    `stream` (default: torch's current stream) after its queued work, and
    returns the library pool's unused memory to the device."""
    _check(lib().lsb_release(_stream_ptr(stream)), "lsb_release")


def sort_batch(As, Bs, ns=None, key_bits=64, digit_bits=0, tile=0, stream=None, sync=False,
               no_skip=False, graph=False, key_order="unsigned", lsd=None):
    """Sorts each buffer As[i] (its first ns[i] elements) on its own, stably
    by key, with Bs[i] as its scratch (None allowed for at most 8192
    elements), in one batched launch sequence without copying anything.
    lsd: as for sort_segmented."""
    nseg = len(As)
    if ns is None:
        ns = [_nelem(a) for a in As]
    Ap = (ctypes.c_void_p * max(nseg, 1))(*[_ptr(a) for a in As])
    Bp = (ctypes.c_void_p * max(nseg, 1))(*[(_ptr(b) if b is not None else None) for b in Bs])
    Np = (ctypes.c_uint64 * max(nseg, 1))(*ns)
    flags = ((LSB_SYNC if sync else 0) | (LSB_NO_SKIP if no_skip else 0)
             | (LSB_GRAPH if graph else 0) | _form_flag(lsd) | _key_flags(key_order))
    c = make_config(key_bits, digit_bits, tile, stream, flags)
    _check(lib().lsb_sort_batch(Ap, Bp, Np, nseg, ctypes.byref(c)), "lsb_sort_batch")


def segmented_workspace_size(total_n, nseg, key_bits=64, digit_bits=0, tile=0):
    c = make_config(key_bits, digit_bits, tile, stream=0)
    return lib().lsb_segmented_workspace_size(total_n, nseg, ctypes.byref(c))


def autotune(lo_log2=12, hi_log2=26, key_bits=64, stream=None):
    """Times every compiled variant per size class and makes the fastest the
    automatic choice; returns the resulting table."""
    _check(lib().lsb_autotune(lo_log2, hi_log2, key_bits, _stream_ptr(stream)), "lsb_autotune")
    return tune_table()


def tune_table():
    n = lib().lsb_tune_table(None, 0)
    arr = (TuneEntry * max(n, 1))()
    lib().lsb_tune_table(arr, n)
    return [{f: getattr(arr[i], f) for f, _ in TuneEntry._fields_} for i in range(n)]


def tune_save(path):
    _check(lib().lsb_tune_save(path.encode()), "lsb_tune_save")


def tune_load(path):
    _check(lib().lsb_tune_load(path.encode()), "lsb_tune_load")


def tune_clear():
    _check(lib().lsb_tune_clear(), "lsb_tune_clear")


def profile_read():
    """Kernel timings (ms) of the last sort(profile=True) on this thread."""
    p = Profile()
    _check(lib().lsb_profile_read(ctypes.byref(p)), "lsb_profile_read")
    P = p.passes
    return {"passes": P, "executed": [bool(p.executed_mask >> q & 1) for q in range(P)],
            "digit_bits": p.digit_bits, "tile": p.tile, "histogram_ms": p.histogram_ms,
            "scan_ms": p.scan_ms, "copy_ms": p.copy_ms, "total_ms": p.total_ms,
            "pass_ms": [p.pass_ms[q] for q in range(P)]}


def msd_profile_read():
    """Kernel timings (ms) and element counts of the last sort(profile=True)
    on this thread that ran the MSD form.  Launch order: wide step 0, then
    level k followed by wide step k+1 while k + 1 < wide_steps; each is
    [histogram, plan, scatter]."""
    p = MsdProfile()
    _check(lib().lsb_msd_profile_read(ctypes.byref(p)), "lsb_msd_profile_read")
    K, S = p.levels, p.wide_steps
    return {"levels": K, "wide_steps": S, "runs": p.runs,
            "wide_elems": [p.wide_elems[w] for w in range(S)],
            "level_elems": [p.level_elems[k] for k in range(K)],
            "finish_elems": p.finish_elems, "setup_ms": p.setup_ms,
            "wide_ms": [list(p.wide_ms[w]) for w in range(S)],
            "level_ms": [list(p.level_ms[k]) for k in range(K)],
            "finish_ms": p.finish_ms, "total_ms": p.total_ms}


def debug_histogram(A, counts, n=None, key_bits=64, digit_bits=0, stream=None, sync=True,
                    key_order="unsigned"):
    n = _nelem(A) if n is None else n
    c = make_config(key_bits, digit_bits, 0, stream,
                    (LSB_SYNC if sync else 0) | _key_flags(key_order))
    _check(lib().lsb_debug_histogram(_ptr(A), n, ctypes.byref(c), _ptr(counts)),
           "lsb_debug_histogram")


def generate(A, n=None, kind=GEN_UNIFORM, ranks=1, key_bits=64, seed_base=0, val_base=0,
             zipf_m=1 << 20, zipf_s=1.1, zipf_table_seed=12345, stream=None):
    n = _nelem(A) if n is None else n
    s = GenSpec(kind=kind, ranks=ranks, key_bits=key_bits, zipf_m=zipf_m, zipf_s=zipf_s,
                zipf_table_seed=zipf_table_seed, seed_base=seed_base, val_base=val_base)
    _check(lib().lsb_generate(_ptr(A), n, ctypes.byref(s), _stream_ptr(stream)),
           "lsb_generate")


def verify(out, in_keys, n=None, val_base=0, key_bits=64, stream=None, key_order="unsigned"):
    """Returns (bad_count, first_bad) -- (0, -1) proves out is the stable sort."""
    n = _nelem(out) if n is None else n
    bad, first = ctypes.c_uint64(), ctypes.c_int64()
    _check(lib().lsb_verify(_ptr(out), _ptr(in_keys), n, val_base, key_bits,
                            _key_flags(key_order), ctypes.byref(bad), ctypes.byref(first),
                            _stream_ptr(stream)),
           "lsb_verify")
    return bad.value, first.value


def debug_scatter_floor(src, dst, runs=256, n=None, stream=None):
    """One scatter pass's memory traffic without the sort: 8192-element tiles
    read in order, each written as `runs` runs into `runs` regions of dst."""
    n = _nelem(src) if n is None else n
    _check(lib().lsb_debug_scatter_floor(_ptr(src), _ptr(dst), n, runs, _stream_ptr(stream)),
           "lsb_debug_scatter_floor")


def extract_keys(A, keys, n=None, stream=None):
    n = _nelem(A) if n is None else n
    _check(lib().lsb_extract_keys(_ptr(A), n, _ptr(keys), _stream_ptr(stream)),
           "lsb_extract_keys")


class Server:
    """Continuous-batching sort service (lsb_server_*): jobs are device
    arrays sorted in place; small jobs arriving while the device is busy are
    coalesced into segmented launches."""

    def __init__(self, memory_budget=0, max_pending=0, batch_max_elems=0, batch_max_total=0,
                 streams=0, use_graphs=True, batch_window_us=0):
        c = ServerConfig()
        lib().lsb_server_config_init(ctypes.byref(c))
        if memory_budget:
            c.memory_budget = memory_budget
        if max_pending:
            c.max_pending = max_pending
        if batch_max_elems:
            c.batch_max_elems = batch_max_elems
        if batch_max_total:
            c.batch_max_total = batch_max_total
        if streams:
            c.streams = streams
        if batch_window_us:
            c.batch_window_us = batch_window_us
        c.use_graphs = 1 if use_graphs else 0
        h = ctypes.c_void_p()
        _check(lib().lsb_server_create(ctypes.byref(c), ctypes.byref(h)), "lsb_server_create")
        self._h = h

    def submit(self, A, n=None, key_bits=64, key_order="unsigned"):
        n = _nelem(A) if n is None else n
        jid = ctypes.c_uint64()
        _check(lib().lsb_server_submit(self._h, _ptr(A), n, key_bits, _key_flags(key_order),
                                       ctypes.byref(jid)),
               "lsb_server_submit")
        return jid.value

    def submit_many(self, arrays, key_bits=64, key_order="unsigned"):
        """Queues every array at once (lsb_server_submit_batch, all or
        nothing); key_bits / key_order are one value or one per array.
        Returns the job ids in order."""
        m = len(arrays)
        kb = list(key_bits) if isinstance(key_bits, (list, tuple)) else [key_bits] * m
        ko = list(key_order) if isinstance(key_order, (list, tuple)) else [key_order] * m
        if len(kb) != m or len(ko) != m:
            raise ValueError("one key_bits and key_order per array")
        As = (ctypes.c_void_p * max(m, 1))(*[_ptr(a) for a in arrays])
        ns = (ctypes.c_size_t * max(m, 1))(*[_nelem(a) for a in arrays])
        bits = (ctypes.c_uint32 * max(m, 1))(*kb)
        flags = (ctypes.c_uint32 * max(m, 1))(*[_key_flags(o) for o in ko])
        ids = (ctypes.c_uint64 * max(m, 1))()
        _check(lib().lsb_server_submit_batch(self._h, As, ns, bits, flags, m, ids),
               "lsb_server_submit_batch")
        return list(ids[:m])

    def poll(self, job):
        rc = lib().lsb_server_poll(self._h, job)
        if rc < 0:
            raise LsbError(rc, "lsb_server_poll")
        return rc == 1

    def wait(self, job, timeout_s=-1.0):
        _check(lib().lsb_server_wait(self._h, job, timeout_s), "lsb_server_wait")

    def drain(self):
        _check(lib().lsb_server_drain(self._h), "lsb_server_drain")

    def latency_us(self, job):
        v = ctypes.c_double()
        _check(lib().lsb_server_latency(self._h, job, ctypes.byref(v)), "lsb_server_latency")
        return v.value

    def stats(self):
        st = ServerStats()
        _check(lib().lsb_server_stats_get(self._h, ctypes.byref(st)), "lsb_server_stats_get")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(lib().lsb_server_reset_stats(self._h), "lsb_server_reset_stats")

    def close(self):
        if self._h:
            _check(lib().lsb_server_destroy(self._h), "lsb_server_destroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
