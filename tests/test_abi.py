"""The C ABI library loads, exports every function include/lsb.h declares,
and rejects bad arguments before touching the GPU.  CPU only: nothing here
launches a kernel."""
import ctypes
import os
import re

import pytest

import lsb_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = open(os.path.join(ROOT, "include", "lsb.h")).read()


def declared_functions():
    body = re.sub(r"/\*.*?\*/", "", HEADER, flags=re.S)
    return sorted(set(re.findall(r"\b(lsb_[a-z_0-9]+)\s*\(", body)))


def test_library_exports_every_declared_symbol():
    L = lsb_amd.lib()
    names = declared_functions()
    assert "lsb_sort" in names and len(names) >= 10
    for name in names:
        assert hasattr(L, name), name
        assert name in lsb_amd.SIGNATURES, f"binding for {name} missing"


def test_struct_layouts_match_header():
    assert ctypes.sizeof(lsb_amd.Config) == 48
    assert lsb_amd.Config.stream.offset == 16
    assert ctypes.sizeof(lsb_amd.GenSpec) == 48
    assert ctypes.sizeof(lsb_amd.ServerStats) == 104


def test_version_and_strerror():
    assert "gfx950" in lsb_amd.version()
    assert lsb_amd.strerror(0) == "success"
    for code in (-1, -2, -3, -4, -5, -6):
        assert lsb_amd.strerror(code) != "unknown error"
    assert lsb_amd.strerror(-99) == "unknown error"


def _sort_rc(A, B, n, **kw):
    c = lsb_amd.make_config(stream=0, **kw)
    return lsb_amd.lib().lsb_sort(A, B, n, ctypes.byref(c))


def test_argument_errors_need_no_gpu():
    a, b = 0x10000, 0x20000000  # never dereferenced: every case fails before HIP
    assert _sort_rc(None, None, 0) == 0                 # n = 0 always succeeds
    assert _sort_rc(a, b, 1) == 0                       # n = 1: nothing to do
    assert _sort_rc(None, b, 5) == lsb_amd.LSB_EINVAL
    assert _sort_rc(a, None, 5) == lsb_amd.LSB_EINVAL
    assert _sort_rc(a + 8, b, 5) == lsb_amd.LSB_EINVAL  # misaligned
    assert _sort_rc(a, a + 16 * 4, 5) == lsb_amd.LSB_EINVAL  # overlapping
    assert _sort_rc(a, b, lsb_amd.LSB_MAX_N + 1) == lsb_amd.LSB_ERANGE
    assert _sort_rc(a, b, 5, key_bits=65) == lsb_amd.LSB_EINVAL
    assert _sort_rc(a, b, 5, digit_bits=16) == lsb_amd.LSB_EINVAL
    assert _sort_rc(a, b, 5, digit_bits=8, tile=12345) == lsb_amd.LSB_EINVAL


def test_plan_and_workspace():
    assert lsb_amd.plan(1 << 24, digit_bits=8)[::2] == (8, 8)
    assert lsb_amd.plan(1 << 24, digit_bits=11)[::2] == (11, 6)
    assert lsb_amd.plan(1 << 24, key_bits=32, digit_bits=8)[2] == 4
    assert lsb_amd.plan(1 << 24, key_bits=1, digit_bits=8)[2] == 1
    assert lsb_amd.workspace_size(0) == 0 and lsb_amd.workspace_size(1) == 0
    w1, w2 = lsb_amd.workspace_size(1 << 20), lsb_amd.workspace_size(1 << 24, lsd=True)
    assert 0 < w1 < w2
    # two look-back slices of tiles x 2^d words dominate
    d, tile, _ = lsb_amd.plan(1 << 24)
    assert w2 >= 2 * ((1 << 24) // tile) * (1 << d) * 4
    vs = lsb_amd.variants()
    assert any(d == 8 and t == 4096 for _, d, t, _ in vs) and any(d == 11 for _, d, _, _ in vs)
    assert any(p for *_, p in vs)
    for v, d, t, _ in vs:
        assert lsb_amd.plan(1 << 20, digit_bits=d, tile=t)[:2] == (d, t)
        assert lsb_amd.plan(1 << 20, variant=v)[:2] == (d, t)
    c = lsb_amd.make_config(stream=0, variant=len(vs) + 1)
    assert lsb_amd.lib().lsb_plan(10, ctypes.byref(c), None, None, None) == lsb_amd.LSB_EINVAL
    # key orders: signed any width, float 32 or 64 bits, never both
    F, S = lsb_amd.LSB_KEYS_FLOAT, lsb_amd.LSB_KEYS_SIGNED
    for kb, flags, ok in [(64, S, True), (13, S, True), (64, F, True), (32, F, True),
                          (48, F, False), (64, F | S, False)]:
        c = lsb_amd.make_config(key_bits=kb, stream=0, flags=flags)
        rc = lsb_amd.lib().lsb_plan(1 << 20, ctypes.byref(c), None, None, None)
        assert (rc == 0) == ok, (kb, flags, rc)
        assert (lsb_amd.lib().lsb_workspace_size(1 << 20, ctypes.byref(c)) > 0) == ok


def test_sort_form_choice():
    """lsb_sort takes the MSD form for keys wider than 32 bits from
    LSB_MSD_MIN_N elements on, with the digit width, tile and variant left
    automatic; the form flags, an explicit LSD configuration or LSB_NO_SKIP
    override it; the workspace size follows the form."""
    m = lsb_amd.LSB_MSD_MIN_N
    f = lsb_amd.sort_form
    assert f(m) == "msd" and f(1 << 30) == "msd" and f(m, key_bits=33) == "msd"
    assert f(m - 1) == "lsd" and f(m, key_bits=32) == "lsd" and f(m, key_bits=20) == "lsd"
    assert f(m, digit_bits=8) == "lsd" and f(m, digit_bits=11) == "lsd"
    assert f(m, no_skip=True) == "lsd" and f(m, lsd=True) == "lsd"
    vs = lsb_amd.variants()
    assert f(m, variant=vs[0][0]) == "lsd"
    # forced MSD: any size and width, unless the configuration is LSD-only
    assert f(100, lsd=False) == "msd" and f(m, key_bits=16, lsd=False) == "msd"
    assert f(m, digit_bits=8, lsd=False) == "msd"
    assert f(m, digit_bits=11, lsd=False) == "lsd" and f(m, tile=4096, lsd=False) == "lsd"
    assert f(m, no_skip=True, lsd=False) == "lsd"
    c = lsb_amd.make_config(key_bits=65, stream=0)
    assert lsb_amd.lib().lsb_sort_form(m, ctypes.byref(c)) == lsb_amd.LSB_EINVAL
    c = lsb_amd.make_config(stream=0)
    assert lsb_amd.lib().lsb_sort_form(lsb_amd.LSB_MAX_N + 1, ctypes.byref(c)) == \
        lsb_amd.LSB_ERANGE
    # the workspace follows the form: the MSD layout holds the level lists
    w_msd, w_lsd = lsb_amd.workspace_size(m), lsb_amd.workspace_size(m, lsd=True)
    assert w_msd > 0 and w_lsd > 0 and w_msd != w_lsd
    assert w_msd == lsb_amd.workspace_size(m, lsd=False)
    assert w_lsd == lsb_amd.workspace_size(m, digit_bits=8)


def test_generate_and_verify_reject_bad_args():
    L = lsb_amd.lib()
    spec = lsb_amd.GenSpec(kind=7)
    assert L.lsb_generate(0x10000, 10, ctypes.byref(spec), None) == lsb_amd.LSB_EINVAL
    spec = lsb_amd.GenSpec(kind=lsb_amd.GEN_ZIPF, zipf_m=0, zipf_s=1.1)
    assert L.lsb_generate(0x10000, 10, ctypes.byref(spec), None) == lsb_amd.LSB_EINVAL
    bad, first = ctypes.c_uint64(), ctypes.c_int64()
    assert L.lsb_verify(None, None, 0, 0, 64, 0, ctypes.byref(bad), ctypes.byref(first),
                        None) == 0
    assert first.value == -1
    assert L.lsb_verify(None, None, 5, 0, 64, 0, ctypes.byref(bad), ctypes.byref(first),
                        None) == lsb_amd.LSB_EINVAL
    # key orders: at most one; floats are 32 or 64 bits wide
    both = lsb_amd.LSB_KEYS_SIGNED | lsb_amd.LSB_KEYS_FLOAT
    assert L.lsb_verify(None, None, 0, 0, 64, both, ctypes.byref(bad), ctypes.byref(first),
                        None) == lsb_amd.LSB_EINVAL
    assert L.lsb_verify(None, None, 0, 0, 48, lsb_amd.LSB_KEYS_FLOAT, ctypes.byref(bad),
                        ctypes.byref(first), None) == lsb_amd.LSB_EINVAL
    assert L.lsb_verify(None, None, 0, 0, 64, 64, ctypes.byref(bad), ctypes.byref(first),
                        None) == lsb_amd.LSB_EINVAL


def _seg_rc(A, B, offsets, **kw):
    off = (ctypes.c_uint64 * len(offsets))(*offsets)
    c = lsb_amd.make_config(stream=0, **kw)
    return lsb_amd.lib().lsb_sort_segmented(A, B, off, len(offsets) - 1, ctypes.byref(c))


def test_segmented_argument_errors_need_no_gpu():
    a, b = 0x10000, 0x80000000
    assert _seg_rc(a, b, [0]) == 0                          # no segments
    assert _seg_rc(None, None, [0, 0, 1, 1, 2]) == 0        # only 0/1-element segments
    assert _seg_rc(a, b, [0, 10, 5]) == lsb_amd.LSB_EINVAL  # decreasing
    assert _seg_rc(None, b, [0, 10]) == lsb_amd.LSB_EINVAL
    assert _seg_rc(a + 8, b, [0, 10]) == lsb_amd.LSB_EINVAL
    assert _seg_rc(a, a + 32, [0, 10]) == lsb_amd.LSB_EINVAL  # A and B overlap
    assert _seg_rc(a, b, [0, lsb_amd.LSB_MAX_N + 1]) == lsb_amd.LSB_ERANGE
    assert _seg_rc(a, b, [0, 10], key_bits=70) == lsb_amd.LSB_EINVAL
    assert _seg_rc(a, b, [0, 10], flags=lsb_amd.LSB_PROFILE) == lsb_amd.LSB_EINVAL


def _batch_rc(As, Bs, ns, **kw):
    k = len(ns)
    Ap = (ctypes.c_void_p * max(k, 1))(*As)
    Bp = (ctypes.c_void_p * max(k, 1))(*Bs) if Bs is not None else None
    Np = (ctypes.c_uint64 * max(k, 1))(*ns)
    c = lsb_amd.make_config(stream=0, **kw)
    return lsb_amd.lib().lsb_sort_batch(Ap, Bp, Np, k, ctypes.byref(c))


def test_batch_argument_errors_need_no_gpu():
    a, b, big = 0x10000, 0x80000000, 20000
    assert _batch_rc([], [], []) == 0                                  # no segments
    assert _batch_rc([None, None], [None, None], [0, 1]) == 0          # nothing to sort
    assert _batch_rc([None], [b], [big]) == lsb_amd.LSB_EINVAL         # null A
    assert _batch_rc([a + 8], [b], [10]) == lsb_amd.LSB_EINVAL         # misaligned A
    assert _batch_rc([a], [None], [big]) == lsb_amd.LSB_EINVAL         # large needs scratch
    assert _batch_rc([a], None, [big]) == lsb_amd.LSB_EINVAL
    assert _batch_rc([a], [b + 8], [big]) == lsb_amd.LSB_EINVAL        # misaligned B
    assert _batch_rc([a], [a + 16 * 100], [big]) == lsb_amd.LSB_EINVAL  # A and B overlap
    assert _batch_rc([a], [b], [lsb_amd.LSB_MAX_N + 1]) == lsb_amd.LSB_ERANGE
    assert _batch_rc([a], [b], [big], key_bits=70) == lsb_amd.LSB_EINVAL
    assert _batch_rc([a], [b], [big], flags=lsb_amd.LSB_PROFILE) == lsb_amd.LSB_EINVAL


def test_segmented_workspace_grows_with_capacity():
    w1 = lsb_amd.segmented_workspace_size(1 << 20, 16)
    w2 = lsb_amd.segmented_workspace_size(1 << 24, 16)
    w3 = lsb_amd.segmented_workspace_size(1 << 24, 4096)
    assert 0 < w1 < w2 < w3


def test_server_rejects_bad_config_without_gpu():
    L = lsb_amd.lib()
    c = lsb_amd.ServerConfig()
    L.lsb_server_config_init(ctypes.byref(c))
    assert c.memory_budget == 200_000_000_000 and c.streams == 8
    h = ctypes.c_void_p()
    bad = lsb_amd.ServerConfig()
    L.lsb_server_config_init(ctypes.byref(bad))
    bad.streams = 99
    assert L.lsb_server_create(ctypes.byref(bad), ctypes.byref(h)) == lsb_amd.LSB_EINVAL
    L.lsb_server_config_init(ctypes.byref(bad))
    bad.batch_max_elems, bad.batch_max_total = 1 << 20, 1 << 10
    assert L.lsb_server_create(ctypes.byref(bad), ctypes.byref(h)) == lsb_amd.LSB_EINVAL
    assert L.lsb_server_create(ctypes.byref(c), None) == lsb_amd.LSB_EINVAL
    jid = ctypes.c_uint64()
    assert L.lsb_server_submit(None, None, 5, 64, 0, ctypes.byref(jid)) == lsb_amd.LSB_EINVAL
    assert L.lsb_server_submit_batch(None, None, None, None, None, 1, None) == lsb_amd.LSB_EINVAL


def test_tune_table_drives_auto_choice(tmp_path):
    """A loaded tune table picks the automatic variant per size class and
    key width (nearest class wins); clearing it restores the built-in table
    (the autotuner's MI355X output: 2048-element tiles for small arrays,
    8192 for large ones)."""
    lsb_amd.tune_clear()
    assert lsb_amd.tune_table() == []
    default = lsb_amd.plan(1 << 20)
    assert lsb_amd.plan(1 << 16)[:2] == (8, 2048)
    assert lsb_amd.plan(1 << 19)[:2] == (8, 4096)
    assert lsb_amd.plan(1 << 22)[:2] == (8, 8192)
    assert lsb_amd.plan(1 << 30)[:2] == (8, 8192)
    assert lsb_amd.plan(1 << 30, key_bits=32)[:2] == (8, 8192)
    p = tmp_path / "tune.txt"
    p.write_text("# comment\n"
                 "20 64 8 2048 0 1.0 2.0\n"
                 "26 64 8 4096 3 9.0 9.5\n"
                 "20 32 8 8192 3 0.5 0.6\n"
                 "22 64 13 999 0 1.0 1.0\n")        # no such variant: dropped
    lsb_amd.tune_load(str(p))
    t = lsb_amd.tune_table()
    assert [e["log2n"] for e in t] == [20, 26, 20]
    assert lsb_amd.plan(1 << 20)[:2] == (8, 2048)
    assert lsb_amd.plan(1 << 19)[:2] == (8, 2048)
    assert lsb_amd.plan(1 << 27)[:2] == (8, 4096)
    assert lsb_amd.plan(1 << 20, key_bits=32)[:2] == (8, 8192)
    # explicit settings still win over the table
    assert lsb_amd.plan(1 << 20, digit_bits=8, tile=4096)[:2] == (8, 4096)
    out = tmp_path / "saved.txt"
    lsb_amd.tune_save(str(out))
    lsb_amd.tune_clear()
    assert lsb_amd.plan(1 << 20) == default
    lsb_amd.tune_load(str(out))
    assert lsb_amd.plan(1 << 20)[:2] == (8, 2048)
    lsb_amd.tune_clear()
    L = lsb_amd.lib()
    assert L.lsb_autotune(5, 20, 64, None) == lsb_amd.LSB_EINVAL
    assert L.lsb_autotune(20, 12, 64, None) == lsb_amd.LSB_EINVAL
    assert L.lsb_tune_load(b"/nonexistent/tune.txt") == lsb_amd.LSB_EINVAL
