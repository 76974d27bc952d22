"""Host-side API over libldpc: a Monte Carlo LDPC engine on one GPU.

`Engine` is the in-process counterpart of the reference's engine binaries
(engines/scalar_5w/ldpc_sim_scalar.cpp, engines/circulant_dma_4cw/
ldpc_sim_circulant_dma.cpp) and of the golden stepper's dump replay
(tests/test_circulant_dma_stepper.cpp:861-990).  All decoding runs in the HIP
kernels of libldpc; this module only marshals arguments.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional

import numpy as np

from . import _lib as L
from . import codes

RATE_FRACTIONS = {"1/2": (1, 2), "2/3": (2, 3), "4/5": (4, 5)}


@dataclasses.dataclass
class PointResult:
    ebn0_db: float
    blocks: int
    frame_errors: int
    bit_errors: int
    iter_sum: int
    next_cw: int
    elapsed_s: float
    kernel_s: float
    launches: int
    hist: np.ndarray  # hist[i]: converged at iteration i (i >= 1); hist[0]: not converged
    n_tx: int
    k_info: int

    @property
    def fer(self) -> float:
        return self.frame_errors / self.blocks if self.blocks else 0.0

    @property
    def ber(self) -> float:
        """Bit errors per information bit (stepper:824)."""
        return self.bit_errors / (self.blocks * self.k_info) if self.blocks else 0.0

    @property
    def mean_iterations(self) -> float:
        return self.iter_sum / self.blocks if self.blocks else 0.0

    @property
    def gsps(self) -> float:
        """Coded channel symbols per second over the call's wall time, in G."""
        return self.blocks * self.n_tx / self.elapsed_s / 1e9 if self.elapsed_s else 0.0

    @property
    def kernel_gsps(self) -> float:
        return self.blocks * self.n_tx / self.kernel_s / 1e9 if self.kernel_s else 0.0


def row_levels(rows, n):
    """Profile R levels of rows (1-based), as libldpc computes them."""
    lib = L.load()
    rp = np.zeros(len(rows) + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate([np.asarray(r, dtype=np.int32) for r in rows]).astype(np.int32)
    out = np.zeros(len(rows), dtype=np.int32)
    nl = lib.ldpc_row_levels(n, len(rows), rp.ctypes.data, col.ctypes.data, out.ctypes.data)
    if nl < 0:
        raise L.LdpcError(nl, "invalid matrix")
    return out


def row_passes(rows, n, width=64):
    """Pass of every row (0-based) in libldpc's Profile R schedule."""
    lib = L.load()
    rp = np.zeros(len(rows) + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate([np.asarray(r, dtype=np.int32) for r in rows]).astype(np.int32)
    out = np.zeros(len(rows), dtype=np.int32)
    npass = lib.ldpc_row_passes(n, len(rows), rp.ctypes.data, col.ctypes.data, width, out.ctypes.data)
    if npass < 0:
        raise L.LdpcError(npass, "invalid matrix")
    return out


def row_passes_by_degree(rows, n, width=32):
    """(pass of every row, kind of every pass) in libldpc's two-codeword
    Profile R schedule: kind 3 / 6 = every row of the pass has that degree,
    7 = rows of any degree (masked datapath)."""
    lib = L.load()
    rp = np.zeros(len(rows) + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate([np.asarray(r, dtype=np.int32) for r in rows]).astype(np.int32)
    out = np.zeros(len(rows), dtype=np.int32)
    kind = np.zeros(len(rows), dtype=np.uint8)
    npass = lib.ldpc_row_passes_by_degree(n, len(rows), rp.ctypes.data, col.ctypes.data, width, out.ctypes.data,
                                          kind.ctypes.data, len(kind))
    if npass < 0:
        raise L.LdpcError(npass, "invalid matrix")
    return out, kind[:npass].copy()


def generator(k_info: int, rate: str) -> np.ndarray:
    """Quasi-cyclic generator of an AR4JA code as uint32 words
    [12 parity blocks][k_info / z info blocks][z / 32] (ldpc_get_generator):
    parity block r = sum_i g_{r,i} u_i mod x^z + 1."""
    lib = L.load()
    d = ar4ja_desc(k_info, rate)
    n = lib.ldpc_get_generator(C.byref(d), None, 0)
    if n < 0:
        raise L.LdpcError(n, "no generator for this code")
    out = np.zeros(n, dtype=np.uint32)
    lib.ldpc_get_generator(C.byref(d), out.ctypes.data, n)
    z = codes.ar4ja(k_info, rate).z
    return out.reshape(12, k_info // z, z // 32)


def encode(k_info: int, rate: str, info: np.ndarray) -> np.ndarray:
    """Systematic codewords [n_cw, n] (bits as uint8) of info bits [n_cw, k_info] (ldpc_encode)."""
    lib = L.load()
    d = ar4ja_desc(k_info, rate)
    info = np.ascontiguousarray(info, dtype=np.uint8).reshape(-1, k_info)
    n = codes.ar4ja(k_info, rate).n
    out = np.zeros((info.shape[0], n), dtype=np.uint8)
    rc = lib.ldpc_encode(C.byref(d), info.ctypes.data, info.shape[0], out.ctypes.data)
    if rc:
        raise L.LdpcError(rc, "ldpc_encode failed (info bits must be 0 or 1)")
    return out


def ar4ja_desc(k_info: int, rate: str) -> L.CodeDesc:
    num, den = RATE_FRACTIONS[rate]
    return L.CodeDesc(L.LDPC_CODE_AR4JA, k_info, num, den)


AR4JA_CODES = [(k, r) for k in (1024, 4096, 16384) for r in ("1/2", "2/3", "4/5")]


def compiled_schedule(k_info: int = 1024, rate: str = "1/2", layer_order: int = 0):
    """Layer schedule compiled into the kernels: [(first_row, [(base, shift), ...]), ...]."""
    lib = L.load()
    desc = ar4ja_desc(k_info, rate)
    buf = np.zeros(8192, dtype=np.int32)
    n = lib.ldpc_get_schedule(C.byref(desc), layer_order, buf.ctypes.data, buf.size)
    if n < 0:
        raise L.LdpcError(n, "no compiled schedule for this code")
    out, i = [], 0
    while i < n:
        first, deg = int(buf[i]), int(buf[i + 1])
        out.append((first, [(int(buf[i + 2 + 2 * d]), int(buf[i + 3 + 2 * d])) for d in range(deg)]))
        i += 2 + 2 * deg
    return out


def list_variants(k_info: int, rate: str):
    """Profile T kernel variants built for an AR4JA code (LDPC_VARIANT codes)."""
    lib = L.load()
    buf = np.zeros(16, dtype=np.uint32)
    n = lib.ldpc_list_variants(C.byref(ar4ja_desc(k_info, rate)), buf.ctypes.data, buf.size)
    if n < 0:
        raise L.LdpcError(n, "not an AR4JA code")
    return [int(v) for v in buf[:n]]


PROFILES = {"T": L.LDPC_PROFILE_T, "R": L.LDPC_PROFILE_R, "R_NMS": L.LDPC_PROFILE_R_NMS, "D": L.LDPC_PROFILE_D}


def autotune(k_info: int, rate: str, ebn0_db: float, n_cw: int = 1 << 16, reps: int = 2, max_iter: int = 200,
             beta: float = 0.9375, seed: int = 0x5EED2026, any_order: bool = False, device: int = 0):
    """Time every Profile T kernel variant of an AR4JA code on the same codewords.

    Returns (best variant, [dict per variant]); libldpc checks that variants
    with the same layer order produce identical counters.
    """
    lib = L.load()
    cfg = L.DecoderCfg(L.LDPC_PROFILE_T, max_iter, beta)
    out = (L.TuneResult * 16)()
    best = C.c_uint32()
    n = lib.ldpc_autotune(C.byref(ar4ja_desc(k_info, rate)), C.byref(cfg), device, ebn0_db, seed, n_cw, reps,
                          int(any_order), out, 16, C.byref(best))
    if n < 0:
        raise L.LdpcError(n, lib.ldpc_last_error(None).decode())
    rows = []
    for t in out[:n]:
        c, order = L.variant_fields(t.variant)
        rows.append(dict(variant=t.variant, checks_per_thread=c, layer_order=order, threads_per_wg=t.threads_per_wg,
                         cw_per_wg=t.cw_per_wg, wg_per_cu=t.wg_per_cu, gsps=t.gsps,
                         mean_iterations=t.mean_iterations, frame_errors=t.frame_errors,
                         kernel_ms=t.kernel_ns * 1e-6))
    return best.value, rows


class Engine:
    """One libldpc context: a code, a decoder configuration and a device.

    Profiles:
      "T"     the golden stepper's tree Approx-Min* (the nine AR4JA codes;
              ``variant`` = L.variant(checks_per_thread, layer_order) picks a
              kernel variant, 0 the library default);
      "R"     the scalar_5w recursive Approx-Min* that run_simulation.py drives;
      "R_NMS" the same engine's normalized min-sum (its -n flag);
      "D"     the learned layered min-sum of ldpc_mi355x.dmin (set its
              parameters with set_dmin_params).
    A Profile R engine takes either an AR4JA (k_info, rate) code in makeH row
    order or an explicit matrix: ``matrix=(rows, n)`` with rows in decoding
    (file) order, and ``n_punct`` punctured trailing variables.
    """

    def __init__(self, k_info: int = 1024, rate: str = "1/2", max_iter: int = 200, beta: float = 0.9375,
                 profile="T", device: int = 0, matrix=None, n_punct: int = 0, l_max: float = 0.0,
                 r_max: float = 0.0, l_post_max: float = 0.0, gamma: float = 0.0, variant: int = 0):
        self._lib = L.load()
        self._h = None
        prof = PROFILES[profile] if isinstance(profile, str) else int(profile)
        if matrix is not None:
            rows, n = matrix
            rp = np.zeros(len(rows) + 1, dtype=np.int32)
            rp[1:] = np.cumsum([len(r) for r in rows])
            col = np.concatenate([np.asarray(r, dtype=np.int32) for r in rows]).astype(np.int32)
            desc = L.CodeDesc(L.LDPC_CODE_MATRIX, 0, 0, 0, n, len(rows), n_punct, 0, rp.ctypes.data,
                              col.ctypes.data)
        else:
            num, den = RATE_FRACTIONS[rate]
            desc = L.CodeDesc(L.LDPC_CODE_AR4JA, k_info, num, den)
        cfg = L.DecoderCfg(prof, max_iter, beta, l_max, r_max, l_post_max, gamma, variant)
        h = C.c_void_p()
        rc = self._lib.ldpc_create(C.byref(desc), C.byref(cfg), device, C.byref(h))
        if rc != L.LDPC_OK:
            raise L.LdpcError(rc, self._lib.ldpc_last_error(None).decode())
        self._h = h
        info = L.CodeInfo()
        self._check(self._lib.ldpc_get_code_info(self._h, C.byref(info)))
        self.n, self.m, self.k_info, self.n_tx = info.n, info.m, info.k_info, info.n_tx
        self.z, self.n_layers, self.n_edges = info.z, info.n_layers, info.n_edges
        self.max_iter, self.beta, self.rate, self.profile, self.variant = max_iter, beta, rate, prof, variant

    def _check(self, rc: int) -> None:
        if rc != L.LDPC_OK:
            raise L.LdpcError(rc, self._lib.ldpc_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ldpc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- queries -----------------------------------------------------------
    def channel_params(self, ebn0_db: float):
        mu, sg = C.c_float(), C.c_float()
        self._check(self._lib.ldpc_channel_params(self._h, ebn0_db, C.byref(mu), C.byref(sg)))
        return mu.value, sg.value

    def set_capture(self, max_frames: int) -> None:
        """Record the ids of up to max_frames failing codewords in later
        run_point calls (0: off)."""
        self._check(self._lib.ldpc_set_capture(self._h, max_frames))

    def captures(self, cap: int = 1 << 16):
        """(ascending ids of captured failures, failures seen)."""
        ids = np.zeros(cap, dtype=np.uint64)
        seen = C.c_uint64()
        n = self._lib.ldpc_get_captures(self._h, ids.ctypes.data, cap, C.byref(seen))
        self._check(min(n, 0))
        return ids[:n], int(seen.value)

    def set_dmin_params(self, alpha, beta, correction: int = 0, r_max: float = 0.0) -> None:
        """Profile D scale / offset rows [groups, n_params] (float32);
        correction 0 / False min-sum, 1 / True the two-minimum correction."""
        a = np.ascontiguousarray(alpha, dtype=np.float32)
        b = np.ascontiguousarray(beta, dtype=np.float32)
        if a.ndim == 1:
            a, b = a[None, :], b[None, :]
        if a.shape != b.shape:
            raise ValueError("alpha and beta must have the same shape")
        self._check(self._lib.ldpc_set_dmin_params(self._h, a.ctypes.data, b.ctypes.data, a.shape[0], a.shape[1],
                                                   int(correction), float(r_max)))

    def schedule(self):
        order = L.variant_fields(self.variant)[1] if self.variant else 0
        return compiled_schedule(self.k_info, self.rate, order)


    @property
    def stream(self) -> int:
        return self._lib.ldpc_stream(self._h) or 0

    # -- work --------------------------------------------------------------
    def run_point(self, ebn0_db: float, max_blocks: int, target_fe: int = 0, seed: int = 0x5EED2026,
                  first_cw: int = 0, chunk_cw: int = 1 << 20, result: Optional[PointResult] = None) -> PointResult:
        """Monte Carlo at one Eb/N0; accumulates into `result` if given."""
        pt = L.PointCfg(ebn0_db, seed, first_cw, max_blocks, target_fe, chunk_cw)
        res = L.PointResult()
        hist = np.zeros(self.max_iter + 1, dtype=np.uint64)
        self._check(self._lib.ldpc_run_point(self._h, C.byref(pt), C.byref(res), hist.ctypes.data))
        out = PointResult(ebn0_db, res.blocks, res.frame_errors, res.bit_errors, res.iter_sum, res.next_cw,
                          res.elapsed_ns * 1e-9, res.kernel_ns * 1e-9, res.launches, hist, self.n_tx, self.k_info)
        if result is not None:
            result.blocks += out.blocks
            result.frame_errors += out.frame_errors
            result.bit_errors += out.bit_errors
            result.iter_sum += out.iter_sum
            result.next_cw = out.next_cw
            result.elapsed_s += out.elapsed_s
            result.kernel_s += out.kernel_s
            result.launches += out.launches
            result.hist = result.hist + out.hist
            return result
        return out

    def decode_llrs(self, llr: np.ndarray):
        """Decode bf16 channel LLRs [n_cw, N] -> (final LLRs, iterations, converged)."""
        llr = np.ascontiguousarray(llr, dtype=np.uint16).reshape(-1, self.n)
        n_cw = llr.shape[0]
        out = np.empty_like(llr)
        iters = np.empty(n_cw, dtype=np.uint32)
        ok = np.empty(n_cw, dtype=np.uint8)
        self._check(self._lib.ldpc_decode_llrs(self._h, llr.ctypes.data, n_cw, out.ctypes.data, iters.ctypes.data,
                                               ok.ctypes.data))
        return out, iters, ok

    def decode_llrs_device_async(self, d_llr: int, n_cw: int, d_out: int, d_iters: int, d_ok: int) -> None:
        self._check(self._lib.ldpc_decode_llrs_device_async(self._h, d_llr, n_cw, d_out, d_iters, d_ok))

    def noise_tails(self, seed: int, first_cw: int, n_cw: int, thresholds) -> np.ndarray:
        """Counts of channel normals with |z| > each threshold (ldpc_noise_tails)."""
        thr = np.ascontiguousarray(thresholds, dtype=np.float32)
        out = np.zeros(len(thr), dtype=np.uint64)
        self._check(self._lib.ldpc_noise_tails(self._h, seed, first_cw, n_cw, thr.ctypes.data, len(thr),
                                               out.ctypes.data))
        return out

    def set_codeword_mode(self, random_info: bool) -> None:
        """All-zero codeword (the reference's only mode) or random info bits,
        systematically encoded, for run_point and gen_channel."""
        self._check(self._lib.ldpc_set_codeword_mode(
            self._h, L.LDPC_CODEWORD_RANDOM if random_info else L.LDPC_CODEWORD_ZERO))

    def sync(self) -> None:
        self._check(self._lib.ldpc_sync(self._h))

    def gen_channel(self, seed: int, first_cw: int, n_cw: int, ebn0_db: float) -> np.ndarray:
        out = np.empty((n_cw, self.n), dtype=np.uint16)
        self._check(self._lib.ldpc_gen_channel(self._h, seed, first_cw, n_cw, ebn0_db, out.ctypes.data))
        return out


class Batch:
    """Several AR4JA codes (Profile T) simulated together: every round is one
    captured HIP graph holding each code's persistent kernel.

    ``codes``: list of (k_info, rate); ``variants``: optional list of
    LDPC_VARIANT codes (0 = library default).
    """

    def __init__(self, codes, max_iter: int = 200, beta: float = 0.9375, variants=None, device: int = 0):
        self._lib = L.load()
        self._h = None
        self.codes = list(codes)
        n = len(self.codes)
        descs = (L.CodeDesc * n)(*[ar4ja_desc(k, r) for k, r in self.codes])
        var = np.asarray(variants if variants is not None else [0] * n, dtype=np.uint32)
        cfg = L.DecoderCfg(L.LDPC_PROFILE_T, max_iter, beta)
        h = C.c_void_p()
        rc = self._lib.ldpc_batch_create(descs, var.ctypes.data, n, C.byref(cfg), device, C.byref(h))
        if rc != L.LDPC_OK:
            raise L.LdpcError(rc, self._lib.ldpc_batch_last_error(None).decode())
        self._h = h
        self.max_iter = max_iter
        self.info = []
        for i in range(n):
            ci = L.CodeInfo()
            self._check(self._lib.ldpc_batch_code_info(self._h, i, C.byref(ci)))
            self.info.append(ci)

    def _check(self, rc: int) -> None:
        if rc != L.LDPC_OK:
            raise L.LdpcError(rc, self._lib.ldpc_batch_last_error(self._h).decode())

    def run(self, ebn0_db, max_blocks, target_fe=0, seed=0x5EED2026, first_cw=0, chunk_cw=1 << 20):
        """Per-code lists (or scalars for all codes) -> list of PointResult."""
        n = len(self.codes)

        def per(x):
            return list(x) if isinstance(x, (list, tuple, np.ndarray)) else [x] * n

        eb, mb, tf, sd, fc, ch = map(per, (ebn0_db, max_blocks, target_fe, seed, first_cw, chunk_cw))
        pts = (L.PointCfg * n)(*[L.PointCfg(eb[i], sd[i], fc[i], mb[i], tf[i], ch[i]) for i in range(n)])
        res = (L.PointResult * n)()
        hist = np.zeros((n, self.max_iter + 1), dtype=np.uint64)
        self._check(self._lib.ldpc_batch_run(self._h, pts, res, hist.ctypes.data))
        return [PointResult(eb[i], r.blocks, r.frame_errors, r.bit_errors, r.iter_sum, r.next_cw, r.elapsed_ns * 1e-9,
                            r.kernel_ns * 1e-9, r.launches, hist[i], self.info[i].n_tx, self.info[i].k_info)
                for i, r in enumerate(res)]

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ldpc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
