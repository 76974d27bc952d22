"""libldpc loads, exports every symbol include/ldpc_mi355x.h declares, and
rejects bad arguments with a status (no GPU compute here)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT
from ldpc_mi355x import _lib as L
from ldpc_mi355x import build as B

HEADER = os.path.join(ROOT, "include", "ldpc_mi355x.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ldpc_[a-z_]+)\s*\(", src)))


def test_library_is_built_in_tree():
    assert os.path.exists(B.LIB), "run __graft_entry__.build() first"
    assert os.path.dirname(B.LIB).startswith(ROOT)


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(B.LIB)
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), "ctypes bindings out of sync with the header"


def test_struct_layouts_match_header():
    # fixed-width POD structs: sizes a C compiler gives them on x86-64
    assert C.sizeof(L.CodeDesc) == 48
    assert C.sizeof(L.DecoderCfg) == 32
    assert C.sizeof(L.CodeInfo) == 32
    assert C.sizeof(L.PointCfg) == 48
    assert C.sizeof(L.PointResult) == 64
    assert C.sizeof(L.TuneResult) == 48


def test_create_rejects_bad_arguments_without_gpu_work():
    lib = L.load()
    h = C.c_void_p()
    bad_code = L.CodeDesc(L.LDPC_CODE_AR4JA, 1000, 1, 2)
    cfg = L.DecoderCfg(L.LDPC_PROFILE_T, 200, 0.9375)
    assert lib.ldpc_create(C.byref(bad_code), C.byref(cfg), 0, C.byref(h)) == L.LDPC_ERR_ARG
    assert b"AR4JA" in lib.ldpc_last_error(None)
    code = L.CodeDesc(L.LDPC_CODE_AR4JA, 1024, 1, 2)
    bad_cfg = L.DecoderCfg(L.LDPC_PROFILE_T, 0, 0.9375)
    assert lib.ldpc_create(C.byref(code), C.byref(bad_cfg), 0, C.byref(h)) == L.LDPC_ERR_ARG
    bad_profile = L.DecoderCfg(99, 200, 0.9375)
    assert lib.ldpc_create(C.byref(code), C.byref(bad_profile), 0, C.byref(h)) == L.LDPC_ERR_UNSUPPORTED
    assert lib.ldpc_create(None, C.byref(cfg), 0, C.byref(h)) == L.LDPC_ERR_ARG
    assert not h.value


def test_null_context_calls_return_status():
    lib = L.load()
    assert lib.ldpc_sync(None) == L.LDPC_ERR_ARG
    assert lib.ldpc_get_code_info(None, None) == L.LDPC_ERR_ARG
    assert lib.ldpc_run_point(None, None, None, None) == L.LDPC_ERR_ARG
    lib.ldpc_destroy(None)  # no-op


def test_compiled_schedule_equals_python_builder(code12):
    from ldpc_mi355x import decoder
    got = decoder.compiled_schedule(1024, "1/2")
    exp = [(f, [(e.base, e.shift) for e in es]) for f, es in code12.layers]
    assert got == exp


@pytest.mark.parametrize("order", [0, 1])
def test_compiled_schedules_of_all_nine_codes_equal_python_builder(order):
    """The kernels' compile-time schedules (ar4ja_schedule.h) equal codes.py's
    construction from the makeH tables, for both layer orders."""
    from ldpc_mi355x import codes, decoder
    br = codes.BLOCK_ROW_ORDER_T if order == 0 else (0, 1, 2)
    for k, r in decoder.AR4JA_CODES:
        code = codes.ar4ja(k, r, block_row_order=br)
        got = decoder.compiled_schedule(k, r, order)
        assert got == [(f, [(e.base, e.shift) for e in es]) for f, es in code.layers], (k, r, order)


def test_schedule_rejects_other_codes_and_orders():
    lib = L.load()
    buf = (C.c_int32 * 64)()
    desc = L.CodeDesc(L.LDPC_CODE_AR4JA, 1000, 1, 2)
    assert lib.ldpc_get_schedule(C.byref(desc), 0, buf, 64) == L.LDPC_ERR_UNSUPPORTED
    ok = L.CodeDesc(L.LDPC_CODE_AR4JA, 4096, 2, 3)
    assert lib.ldpc_get_schedule(C.byref(ok), 2, buf, 64) == L.LDPC_ERR_ARG
    assert lib.ldpc_get_schedule(C.byref(ok), 0, buf, 8) == L.LDPC_ERR_ARG  # too small


def test_variant_tables():
    """Every code has C = 1 and 2 in both layer orders, except K=16384 r1/2
    (z = 2048 > 1024 threads per workgroup: C = 2 only)."""
    from ldpc_mi355x import decoder
    for k, r in decoder.AR4JA_CODES:
        vs = decoder.list_variants(k, r)
        fields = sorted(L.variant_fields(v) for v in vs)
        if (k, r) == (16384, "1/2"):
            assert fields == [(2, 0), (2, 1)]
        else:
            assert fields == [(1, 0), (1, 1), (2, 0), (2, 1)]
        assert all(v == L.variant(*L.variant_fields(v)) for v in vs)


def test_variant_and_batch_arguments_checked_without_gpu():
    lib = L.load()
    h = C.c_void_p()
    code = L.CodeDesc(L.LDPC_CODE_AR4JA, 4096, 4, 5)
    cfg = L.DecoderCfg(L.LDPC_PROFILE_T, 200, 0.9375, 0, 0, 0, 0, L.variant(3, 0))
    assert lib.ldpc_create(C.byref(code), C.byref(cfg), 0, C.byref(h)) == L.LDPC_ERR_UNSUPPORTED
    cfg = L.DecoderCfg(L.LDPC_PROFILE_R, 20, 0.0, 0, 0, 0, 0, L.variant(1, 0))
    assert lib.ldpc_create(C.byref(code), C.byref(cfg), 0, C.byref(h)) == L.LDPC_ERR_ARG
    descs = (L.CodeDesc * 2)(L.CodeDesc(L.LDPC_CODE_AR4JA, 1024, 1, 2), L.CodeDesc(L.LDPC_CODE_AR4JA, 1000, 1, 2))
    cfg_t = L.DecoderCfg(L.LDPC_PROFILE_T, 200, 0.9375)
    assert lib.ldpc_batch_create(descs, None, 2, C.byref(cfg_t), 0, C.byref(h)) == L.LDPC_ERR_ARG
    assert b"batch member 1" in lib.ldpc_batch_last_error(None)
    cfg_r = L.DecoderCfg(L.LDPC_PROFILE_R, 20, 0.0)
    assert lib.ldpc_batch_create(descs, None, 1, C.byref(cfg_r), 0, C.byref(h)) == L.LDPC_ERR_UNSUPPORTED
    best = C.c_uint32()
    out = (L.TuneResult * 4)()
    assert lib.ldpc_autotune(C.byref(descs[0]), C.byref(cfg_r), 0, 2.0, 1, 64, 1, 0, out, 4, C.byref(best)) \
        == L.LDPC_ERR_ARG
    assert lib.ldpc_batch_run(None, None, None, None) == L.LDPC_ERR_ARG
    lib.ldpc_batch_destroy(None)


def test_row_levels_rejects_bad_matrices():
    lib = L.load()
    rp = (C.c_int32 * 3)(0, 2, 15)
    col = (C.c_int32 * 15)(*range(15))
    out = (C.c_int32 * 2)()
    assert lib.ldpc_row_levels(20, 2, rp, col, out) == L.LDPC_ERR_ARG   # degree 13 > 12
    rp2 = (C.c_int32 * 3)(0, 2, 4)
    assert lib.ldpc_row_levels(3, 2, rp2, col, out) == L.LDPC_ERR_ARG   # variable 3 out of range
    assert lib.ldpc_row_levels(4, 2, rp2, col, out) == 1 and list(out) == [1, 1]
    assert lib.ldpc_row_passes(3, 2, rp2, col, 64, out) == L.LDPC_ERR_ARG  # variable 3 out of range
    assert lib.ldpc_row_passes(4, 2, rp2, col, 0, out) == L.LDPC_ERR_ARG  # width 0
    assert lib.ldpc_row_passes(4, 2, rp2, col, 1, out) == 2 and list(out) == [0, 1]


def test_repeated_column_is_refused_and_never_deadlocks_the_schedule():
    """A row listing one variable twice ([[0, 1, 1], [1, 2], [2, 3], [0, 3]])
    once made the list scheduler add an edge from the row to itself, so that
    no row was ever ready and the pass table was built empty.  libldpc refuses
    the matrix; the Python mirror, which takes any rows, schedules it."""
    lib = L.load()
    rows = [[0, 1, 1], [1, 2], [2, 3], [0, 3]]
    rp = (C.c_int32 * 5)(0, 3, 5, 7, 9)
    col = (C.c_int32 * 9)(*[v for r in rows for v in r])
    out = (C.c_int32 * 4)()
    assert lib.ldpc_row_passes(4, 4, rp, col, 64, out) == L.LDPC_ERR_ARG
    assert lib.ldpc_row_levels(4, 4, rp, col, out) == L.LDPC_ERR_ARG
    from ldpc_mi355x import schedule_r
    assert schedule_r.passes(rows, 4, 64) == [0, 1, 2, 3]
    # the same rows without the repeat are accepted and chain one row per pass
    rows_ok = [[0, 1], [1, 2], [2, 3], [0, 3]]
    rp_ok = (C.c_int32 * 5)(0, 2, 4, 6, 8)
    col_ok = (C.c_int32 * 8)(*[v for r in rows_ok for v in r])
    assert lib.ldpc_row_passes(4, 4, rp_ok, col_ok, 64, out) == 4 and list(out) == [0, 1, 2, 3]
    assert schedule_r.passes(rows_ok, 4, 64) == [0, 1, 2, 3]


def test_profile_t_rejects_clamps_and_matrices():
    lib = L.load()
    h = C.c_void_p()
    code = L.CodeDesc(L.LDPC_CODE_AR4JA, 1024, 1, 2)
    cfg = L.DecoderCfg(L.LDPC_PROFILE_T, 200, 0.9375, 8.0)
    assert lib.ldpc_create(C.byref(code), C.byref(cfg), 0, C.byref(h)) == L.LDPC_ERR_ARG
    cfg = L.DecoderCfg(L.LDPC_PROFILE_R, 16, 0.0, -1.0)
    assert lib.ldpc_create(C.byref(code), C.byref(cfg), 0, C.byref(h)) == L.LDPC_ERR_ARG
    bad = L.CodeDesc(L.LDPC_CODE_MATRIX, 0, 0, 0, 10, 0, 0, 0, None, None)
    cfg = L.DecoderCfg(L.LDPC_PROFILE_R, 16, 0.0)
    assert lib.ldpc_create(C.byref(bad), C.byref(cfg), 0, C.byref(h)) == L.LDPC_ERR_ARG


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        L.load()


def test_profile_d_arguments_checked_without_gpu():
    lib = L.load()
    h = C.c_void_p()
    # every AR4JA code has a Profile D kernel: K=16384 r1/2 passes the
    # argument checks and stops at the device (none here)
    big = L.CodeDesc(L.LDPC_CODE_AR4JA, 16384, 1, 2)
    cfg = L.DecoderCfg(L.LDPC_PROFILE_D, 20, 0.0)
    rc = lib.ldpc_create(C.byref(big), C.byref(cfg), 0, C.byref(h))
    assert rc in (L.LDPC_OK, L.LDPC_ERR_NODEV)
    if rc == L.LDPC_OK:
        lib.ldpc_destroy(h)
    code = L.CodeDesc(L.LDPC_CODE_AR4JA, 1024, 1, 2)
    cfg_v = L.DecoderCfg(L.LDPC_PROFILE_D, 20, 0.0, 0, 0, 0, 0, L.variant(1, 0))
    assert lib.ldpc_create(C.byref(code), C.byref(cfg_v), 0, C.byref(h)) == L.LDPC_ERR_ARG
    mat = L.CodeDesc(L.LDPC_CODE_MATRIX, 0, 0, 0, 10, 5, 0, 0, None, None)
    assert lib.ldpc_create(C.byref(mat), C.byref(cfg), 0, C.byref(h)) == L.LDPC_ERR_UNSUPPORTED
    assert lib.ldpc_set_dmin_params(None, None, None, 1, 60, 0, 0.0) == L.LDPC_ERR_ARG
