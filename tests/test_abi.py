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


def test_schedule_of_unbuilt_code_is_unsupported():
    lib = L.load()
    buf = (C.c_int32 * 64)()
    desc = L.CodeDesc(L.LDPC_CODE_AR4JA, 4096, 2, 3)
    assert lib.ldpc_get_schedule(C.byref(desc), buf, 64) == L.LDPC_ERR_UNSUPPORTED


def test_row_levels_rejects_bad_matrices():
    lib = L.load()
    rp = (C.c_int32 * 3)(0, 2, 15)
    col = (C.c_int32 * 15)(*range(15))
    out = (C.c_int32 * 2)()
    assert lib.ldpc_row_levels(20, 2, rp, col, out) == L.LDPC_ERR_ARG   # degree 13 > 12
    rp2 = (C.c_int32 * 3)(0, 2, 4)
    assert lib.ldpc_row_levels(3, 2, rp2, col, out) == L.LDPC_ERR_ARG   # variable 3 out of range
    assert lib.ldpc_row_levels(4, 2, rp2, col, out) == 1 and list(out) == [1, 1]


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
