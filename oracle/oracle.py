"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke() entry point and by
bench.py's cpu_baseline leg, never by the product package.  The C source
(ldpc_oracle.c) cites the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_STEPPER = os.path.join(HERE, "_ref", "stepper")

_lib = None

u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.orc_amin_t.restype = C.c_float
        L.orc_amin_t.argtypes = [C.c_float, C.c_float]
        L.orc_decode_t.restype = C.c_int
        L.orc_decode_t.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, C.c_int, u16p, C.c_int, C.c_float,
                                   u16p, u32p, u8p, C.c_int]
        L.orc_trace_t.restype = C.c_int
        L.orc_trace_t.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, u16p, C.c_int, C.c_float, u32p, u16p]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_log_unit.restype = C.c_float
        L.orc_log_unit.argtypes = [C.c_float]
        L.orc_sqrt_newton.restype = C.c_float
        L.orc_sqrt_newton.argtypes = [C.c_float]
        L.orc_sincos_turn.restype = None
        L.orc_sincos_turn.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_box_muller.restype = None
        L.orc_box_muller.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_gen_channel.restype = None
        L.orc_gen_channel.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int,
                                      u16p]
        L.orc_normals.restype = None
        L.orc_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                  np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")]
        L.orc_run_point_t.restype = C.c_int
        L.orc_run_point_t.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, C.c_int, C.c_int, C.c_uint64,
                                      C.c_uint64, C.c_uint64, C.c_float, C.c_float, C.c_int, C.c_float, C.c_int,
                                      u64p, u64p]
        L.orc_run_llrs_t.restype = C.c_int
        L.orc_run_llrs_t.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, C.c_int, C.c_int, u16p, C.c_int, C.c_float,
                                     C.c_int, u64p, u64p]
        L.orc_gen_channel_cw.restype = C.c_int
        L.orc_gen_channel_cw.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int,
                                         C.c_int, u64p, C.c_int, u16p, C.c_void_p]
        L.orc_run_point_t_cw.restype = C.c_int
        L.orc_run_point_t_cw.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, C.c_int, C.c_int, u64p, C.c_int,
                                         C.c_uint64, C.c_uint64, C.c_uint64, C.c_float, C.c_float, C.c_int, C.c_float,
                                         C.c_int, u64p, u64p]
        L.orc_decode_r.restype = C.c_int
        L.orc_decode_r.argtypes = [C.c_int, C.c_int, i32p, i32p, C.c_int, u16p, C.c_int, C.c_int, C.c_float, C.c_float,
                                   u16p, u32p, u8p, C.c_int]
        L.orc_run_point_r.restype = C.c_int
        L.orc_run_point_r.argtypes = [C.c_int, C.c_int, i32p, i32p, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64,
                                      C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                      u64p, u64p]
        L.orc_gen_channel_clamped.restype = None
        L.orc_gen_channel_clamped.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int,
                                              C.c_int, u16p]
        L.orc_num_threads.restype = C.c_int
        L.orc_num_threads.argtypes = []
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.orc_tree_even.restype = None
        L.orc_tree_even.argtypes = [f32p, C.c_int, C.c_float, f32p]
        L.orc_tree6.restype = None
        L.orc_tree6.argtypes = [f32p, C.c_float, f32p]
        L.orc_check_update_t.restype = C.c_int
        L.orc_check_update_t.argtypes = [C.c_int, u16p, u16p, C.c_int, C.c_float, u16p, u16p]
        _lib = L
    return _lib


def code_arrays(code) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(row_ptr, col, sched) int32 arrays of a codes.Code."""
    rp = np.zeros(code.m + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in code.rows])
    col = np.concatenate([np.asarray(r, dtype=np.int32) for r in code.rows]).astype(np.int32)
    sched = code.schedule_rows().astype(np.int32)
    return rp, col, sched


def amin_t(a: float, b: float) -> float:
    return lib().orc_amin_t(a, b)


def tree_even(m, beta: float = 0.9375) -> np.ndarray:
    """Generalised even-degree tree (pairs j, j+h; prefix/suffix over pairs)."""
    m = np.ascontiguousarray(m, dtype=np.float32)
    r = np.empty_like(m)
    lib().orc_tree_even(m, len(m) // 2, beta, r)
    return r


def tree6(m, beta: float = 0.9375) -> np.ndarray:
    """The golden degree-6 tree as the reference stepper writes it."""
    m = np.ascontiguousarray(m, dtype=np.float32)
    r = np.empty_like(m)
    lib().orc_tree6(m, beta, r)
    return r


def check_update_t(L_in, R_in, iter0: bool = False, beta: float = 0.9375):
    """One Profile T check update on explicit bf16 values: (L_out, R_out)."""
    L_in = np.ascontiguousarray(L_in, dtype=np.uint16)
    R_in = np.ascontiguousarray(R_in, dtype=np.uint16)
    Lo, Ro = np.empty_like(L_in), np.empty_like(R_in)
    if lib().orc_check_update_t(len(L_in), L_in, R_in, int(iter0), beta, Lo, Ro):
        raise ValueError("unsupported degree")
    return Lo, Ro


def decode_t(code, llr: np.ndarray, max_iter: int = 200, beta: float = 0.9375, threads: int = 0):
    llr = np.ascontiguousarray(llr, dtype=np.uint16).reshape(-1, code.n)
    n_cw = llr.shape[0]
    rp, col, sched = code_arrays(code)
    out = np.empty_like(llr)
    iters = np.empty(n_cw, dtype=np.uint32)
    ok = np.empty(n_cw, dtype=np.uint8)
    rc = lib().orc_decode_t(code.n, code.m, rp, col, sched, n_cw, llr, max_iter, beta, out, iters, ok, threads)
    if rc:
        raise ValueError("oracle Profile T supports check degree 3 and even degrees 4..18")
    return out, iters, ok


def trace_t(code, llr_one: np.ndarray, max_iter: int = 200, beta: float = 0.9375):
    llr_one = np.ascontiguousarray(llr_one, dtype=np.uint16).reshape(code.n)
    rp, col, sched = code_arrays(code)
    unsat = np.zeros(max_iter, dtype=np.uint32)
    out = np.empty_like(llr_one)
    it = lib().orc_trace_t(code.n, code.m, rp, col, sched, llr_one, max_iter, beta, unsat, out)
    return unsat[:it], out


def philox4x32_10(ctr, key) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def log_unit(u: float) -> float:
    return lib().orc_log_unit(u)


def sqrt_newton(t: float) -> float:
    return lib().orc_sqrt_newton(t)


def sincos_turn(x: int):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos_turn(x, C.byref(s), C.byref(c))
    return s.value, c.value


def box_muller(x0: int, x1: int):
    a, b = C.c_float(), C.c_float()
    lib().orc_box_muller(x0, x1, C.byref(a), C.byref(b))
    return a.value, b.value


def gen_channel(code, seed: int, first_cw: int, n_cw: int, mu: float, sigma: float) -> np.ndarray:
    out = np.empty((n_cw, code.n), dtype=np.uint16)
    lib().orc_gen_channel(seed, first_cw, n_cw, mu, sigma, code.n_tx, code.n, out)
    return out


def normals(seed: int, first_cw: int, n_cw: int, n_tx: int) -> np.ndarray:
    out = np.empty((n_cw, n_tx), dtype=np.float32)
    lib().orc_normals(seed, first_cw, n_cw, n_tx, out)
    return out


def run_point_t(code, seed: int, first_cw: int, n_cw: int, mu: float, sigma: float, max_iter: int = 200,
                beta: float = 0.9375, threads: int = 0):
    """Returns dict(blocks, frame_errors, bit_errors, iter_sum, hist)."""
    rp, col, sched = code_arrays(code)
    cnt = np.zeros(4, dtype=np.uint64)
    hist = np.zeros(max_iter + 1, dtype=np.uint64)
    rc = lib().orc_run_point_t(code.n, code.m, rp, col, sched, code.n_tx, code.k, seed, first_cw, n_cw, mu, sigma,
                               max_iter, beta, threads, cnt, hist)
    if rc:
        raise ValueError("oracle Profile T supports check degree 3 and even degrees 4..18")
    return dict(blocks=int(cnt[0]), frame_errors=int(cnt[1]), bit_errors=int(cnt[2]), iter_sum=int(cnt[3]),
                hist=hist)


def run_llrs_t(code, llr: np.ndarray, max_iter: int = 200, beta: float = 0.9375, threads: int = 0):
    """Decode given channel LLRs [n_cw, n] and tally them with run_point_t's
    rule.  Returns dict(blocks, frame_errors, bit_errors, iter_sum, hist)."""
    llr = np.ascontiguousarray(llr, dtype=np.uint16).reshape(-1, code.n)
    rp, col, sched = code_arrays(code)
    cnt = np.zeros(4, dtype=np.uint64)
    hist = np.zeros(max_iter + 1, dtype=np.uint64)
    rc = lib().orc_run_llrs_t(code.n, code.m, rp, col, sched, code.k, llr.shape[0], llr, max_iter, beta, threads,
                              cnt, hist)
    if rc:
        raise ValueError("oracle Profile T supports check degree 3 and even degrees 4..18")
    return dict(blocks=int(cnt[0]), frame_errors=int(cnt[1]), bit_errors=int(cnt[2]), iter_sum=int(cnt[3]),
                hist=hist)


# --- random-info-bit codewords (no reference counterpart) ------------------

_GEN_CACHE = {}


def dense_parity_generator(code) -> np.ndarray:
    """Dense systematic generator of a code whose last m variables carry the
    parity: P (k x m over GF(2)) with parity = u P, from Gauss-Jordan
    elimination of [H_p | H_u] on bit-packed rows (H_p the last m columns,
    which must be invertible).  Plain dense linear algebra, independent of
    the product's quasi-cyclic encoder.  Returns P packed: k rows of
    ceil(m / 64) little-endian uint64 words."""
    key = (code.name, code.n, code.m)
    if key in _GEN_CACHE:
        return _GEN_CACHE[key]
    k, m = code.k, code.m
    assert code.n - k == m, "the parity part must be square"
    H = code.dense()
    A = np.concatenate([H[:, k:], H[:, :k]], axis=1).astype(np.uint8)   # [H_p | H_u]
    W = (A.shape[1] + 63) // 64
    bits = np.zeros((m, W * 64), dtype=np.uint8)
    bits[:, :A.shape[1]] = A
    rows = np.packbits(bits.reshape(m, W, 64)[:, :, ::-1], axis=2).view(">u8").reshape(m, W).astype(np.uint64)

    def col_bits(c):
        return (rows[:, c // 64] >> np.uint64(c % 64)) & np.uint64(1)

    for c in range(m):
        piv = c + int(np.argmax(col_bits(c)[c:]))
        if not col_bits(c)[piv]:
            raise ValueError("parity part of H is singular")
        if piv != c:
            rows[[c, piv]] = rows[[piv, c]]
        hit = col_bits(c).astype(bool)
        hit[c] = False
        rows[hit] ^= rows[c]
    # rows now hold [I | H_p^-1 H_u]; P = (H_p^-1 H_u)^T
    right = np.zeros((m, k), dtype=np.uint8)
    for c in range(k):
        right[:, c] = col_bits(m + c).astype(np.uint8)
    Pt = right.T                                                         # k x m
    pw = (m + 63) // 64
    pad = np.zeros((k, pw * 64), dtype=np.uint8)
    pad[:, :m] = Pt
    P = np.packbits(pad.reshape(k, pw, 64)[:, :, ::-1], axis=2).view(">u8").reshape(k, pw).astype(np.uint64)
    P = np.ascontiguousarray(P)
    _GEN_CACHE[key] = P
    return P


def gen_channel_cw(code, seed: int, first_cw: int, n_cw: int, mu: float, sigma: float, with_bits: bool = False):
    """Channel LLRs [n_cw, n] of random-info-bit codewords (and their bits)."""
    P = dense_parity_generator(code)
    out = np.empty((n_cw, code.n), dtype=np.uint16)
    bits = np.empty((n_cw, code.n), dtype=np.uint8) if with_bits else None
    rc = lib().orc_gen_channel_cw(seed, first_cw, n_cw, mu, sigma, code.n_tx, code.n, code.k, P, P.shape[1], out,
                                  bits.ctypes.data if with_bits else None)
    if rc:
        raise ValueError("unsupported code for the random-codeword oracle")
    return (out, bits) if with_bits else out


def run_point_t_cw(code, seed: int, first_cw: int, n_cw: int, mu: float, sigma: float, max_iter: int = 200,
                   beta: float = 0.9375, threads: int = 0):
    """Monte Carlo of random-info-bit codewords: dict(blocks, frame_errors, bit_errors, iter_sum, hist)."""
    P = dense_parity_generator(code)
    rp, col, sched = code_arrays(code)
    cnt = np.zeros(4, dtype=np.uint64)
    hist = np.zeros(max_iter + 1, dtype=np.uint64)
    rc = lib().orc_run_point_t_cw(code.n, code.m, rp, col, sched, code.n_tx, code.k, P, P.shape[1], seed, first_cw,
                                  n_cw, mu, sigma, max_iter, beta, threads, cnt, hist)
    if rc:
        raise ValueError("unsupported code for the random-codeword oracle")
    return dict(blocks=int(cnt[0]), frame_errors=int(cnt[1]), bit_errors=int(cnt[2]), iter_sum=int(cnt[3]),
                hist=hist)


def rows_arrays(rows):
    """(row_ptr, col) int32 arrays of rows given in decoding (file) order."""
    rp = np.zeros(len(rows) + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate([np.asarray(r, dtype=np.int32) for r in rows]).astype(np.int32)
    return rp, col


def decode_r(n, rows, llr, max_iter=16, nms=False, r_max=0.0, l_post_max=0.0, threads=0):
    """Profile R decode of bf16 LLRs [n_cw, n]; rows in file order."""
    llr = np.ascontiguousarray(llr, dtype=np.uint16).reshape(-1, n)
    rp, col = rows_arrays(rows)
    out = np.empty_like(llr)
    iters = np.empty(llr.shape[0], dtype=np.uint32)
    ok = np.empty(llr.shape[0], dtype=np.uint8)
    rc = lib().orc_decode_r(n, len(rows), rp, col, llr.shape[0], llr, max_iter, int(nms), r_max, l_post_max, out,
                            iters, ok, threads)
    if rc:
        raise ValueError("Profile R supports check degrees 1..12")
    return out, iters, ok


def run_point_r(n, rows, n_tx, seed, first_cw, n_cw, mu, sigma, l_max=0.0, max_iter=16, nms=False, r_max=0.0,
                l_post_max=0.0, threads=0):
    rp, col = rows_arrays(rows)
    cnt = np.zeros(4, dtype=np.uint64)
    hist = np.zeros(max_iter + 1, dtype=np.uint64)
    rc = lib().orc_run_point_r(n, len(rows), rp, col, n_tx, seed, first_cw, n_cw, mu, sigma, l_max, max_iter,
                               int(nms), r_max, l_post_max, threads, cnt, hist)
    if rc:
        raise ValueError("Profile R supports check degrees 1..12")
    return dict(blocks=int(cnt[0]), frame_errors=int(cnt[1]), bit_errors=int(cnt[2]), iter_sum=int(cnt[3]),
                hist=hist)


def gen_channel_clamped(n, n_tx, seed, first_cw, n_cw, mu, sigma, l_max=0.0):
    out = np.empty((n_cw, n), dtype=np.uint16)
    lib().orc_gen_channel_clamped(seed, first_cw, n_cw, mu, sigma, l_max, n_tx, n, out)
    return out


def num_threads() -> int:
    return lib().orc_num_threads()


# --- Packed4CW dump format (stepper:879-880, 962-965) ------------------------

def to_packed4(llr4: np.ndarray) -> bytes:
    """[4, n] bf16 codewords -> variable-major Packed4CW bytes (n * 8 B)."""
    llr4 = np.asarray(llr4, dtype=np.uint16)
    assert llr4.shape[0] == 4
    return np.ascontiguousarray(llr4.T).astype("<u2").tobytes()


def from_packed4(buf: bytes, n: int) -> np.ndarray:
    a = np.frombuffer(buf, dtype="<u2", count=4 * n).reshape(n, 4)
    return np.ascontiguousarray(a.T)
