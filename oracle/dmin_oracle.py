"""CPU oracle of the differentiable layered min-sum decoder.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's cpu_baseline
leg, never by the product package.

The reference has no differentiable decoder; this file states the
arithmetic the HIP kernels implement, in numpy, so that the GPU forward can
be compared bit for bit (float32, every op separately rounded, as numpy
does) and the GPU gradient against central
finite differences of the same function evaluated in float64.

Decoder (per layer of the AR4JA layered schedule, checks in parallel):

    q_d   = L_v(d) - R_d
    a_d   = |q_d|,  idx1 = first argmin a,  min1 = a_idx1,
          idx2 = first argmin of a without idx1,  min2 = a_idx2
    g_d   = min2                          if d == idx1
          = min1                          otherwise (min-sum), or
          = max(min1 - corr(min2 - min1), 0)   with the Approx-Min* correction
            (corr = 1), corr(x) = max(C1 + K1 x, C2 + K2 x, 0), bf16-truncated
            constants of the reference (sfpi_host_emu.h:566-584)
    with three minima (corr = 2; idx3 = first argmin without idx1, idx2):
      g_d = amin(min2, min3) if d == idx1, amin(min1, min3) if d == idx2,
            amin(min1, min2) otherwise,  amin(lo, hi) = max(lo - corr(hi - lo), 0)
    z_d   = alpha_p(d) * g_d - beta_p(d)          (learnable, per layer slot)
    r_d   = copysign(max(z_d, 0), s ^ sign(q_d)),  s = XOR of all sign bits
    (with r_max > 0 the magnitude is min(max(z_d, 0), r_max))
    L_v(d) <- q_d + r_d,  R_d <- r_d

p(d) is the edge slot of the layer (one circulant of the quarter-lifted
protograph): P = sum of layer degrees parameters per kind (60 for rate 1/2).
alpha and beta have shape [G, P]; iteration t uses row t if G == T, row 0 if
G == 1.  Loss: mean over codewords and transmitted positions of
softplus(-L) (binary cross-entropy of the soft output against the all-zero
codeword's bits).
"""
from __future__ import annotations

import numpy as np

C1, K1, C2, K2 = 0.69140625, -0.3984375, 0.65625, -0.326171875


def schedule(code):
    """[(deg, Z, var_idx[deg, Z], edge_off, param_off)] in decoding order."""
    out, e_off, p_off = [], 0, 0
    for first, slots in code.layers:
        z = code.z
        j = np.arange(z)
        var = np.stack([s.base + (j + s.shift) % z for s in slots]).astype(np.int64)
        out.append((len(slots), z, var, e_off, p_off))
        e_off += len(slots) * z
        p_off += len(slots)
    return out


def n_params(code) -> int:
    return sum(len(s) for _, s in code.layers)


def _amin(lo, hi, dt):
    """Approx-Min* of two magnitudes lo <= hi: max(lo - corr(hi - lo), 0)."""
    x = (hi - lo).astype(dt)
    l1 = (dt(K1) * x).astype(dt) + dt(C1)
    l2 = (dt(K2) * x).astype(dt) + dt(C2)
    c = np.maximum(np.maximum(l1, l2), dt(0))
    return np.maximum((lo - c).astype(dt), dt(0))


def _layer(L, R, lay, alpha_row, beta_row, corr, dt, tape=None, r_max=0.0):
    deg, z, var, e_off, p_off = lay
    B = L.shape[0]
    Rl = R[:, e_off:e_off + deg * z].reshape(B, deg, z)
    q = (L[:, var] - Rl).astype(dt)
    if tape is not None:
        tape.append(q.copy())
    a = np.abs(q)
    idx1 = np.argmin(a, axis=1)                                     # [B, z], first occurrence
    min1 = np.take_along_axis(a, idx1[:, None, :], 1)[:, 0, :]
    a2 = a.copy()
    np.put_along_axis(a2, idx1[:, None, :], np.inf, 1)
    idx2 = np.argmin(a2, axis=1)
    min2 = np.take_along_axis(a, idx2[:, None, :], 1)[:, 0, :]
    d_ix = np.arange(deg)[None, :, None]
    if corr == 2:
        a3 = a2.copy()
        np.put_along_axis(a3, idx2[:, None, :], np.inf, 1)
        idx3 = np.argmin(a3, axis=1)
        min3 = np.take_along_axis(a, idx3[:, None, :], 1)[:, 0, :]
        o12, o13, o23 = _amin(min1, min2, dt), _amin(min1, min3, dt), _amin(min2, min3, dt)
        g = np.where(d_ix == idx1[:, None, :], o23[:, None, :],
                     np.where(d_ix == idx2[:, None, :], o13[:, None, :], o12[:, None, :])).astype(dt)
    else:
        other = _amin(min1, min2, dt) if corr else min1
        g = np.where(d_ix == idx1[:, None, :], min2[:, None, :], other[:, None, :]).astype(dt)
    al = alpha_row[p_off:p_off + deg].astype(dt)[None, :, None]
    be = beta_row[p_off:p_off + deg].astype(dt)[None, :, None]
    zz = ((al * g).astype(dt) - be).astype(dt)
    mag = np.where(zz > 0, zz, dt(0)).astype(dt)
    if r_max > 0:
        mag = np.where(mag < dt(r_max), mag, dt(r_max)).astype(dt)
    sb = np.signbit(q)
    s = np.bitwise_xor.reduce(sb, axis=1)[:, None, :]
    r = np.copysign(mag, np.where(s ^ sb, dt(-1), dt(1))).astype(dt)
    L[:, var] = (q + r).astype(dt)
    R[:, e_off:e_off + deg * z] = r.reshape(B, deg * z)


def forward(code, llr, alpha, beta, T, corr=False, dtype=np.float32, r_max=0.0):
    """Soft output after T iterations (no early exit), same dtype as computed."""
    dt = np.dtype(dtype).type
    sch = schedule(code)
    E = sum(l[0] * l[1] for l in sch)
    L = np.asarray(llr, dtype=dt).copy()
    R = np.zeros((L.shape[0], E), dtype=dt)
    alpha = np.asarray(alpha, dtype=dt)
    beta = np.asarray(beta, dtype=dt)
    for t in range(T):
        g = t if alpha.shape[0] == T else 0
        for lay in sch:
            _layer(L, R, lay, alpha[g], beta[g], corr, dt, r_max=r_max)
    return L


def loss(L, n_tx):
    """mean softplus(-L) over the transmitted positions (all-zero codeword)."""
    x = -np.asarray(L[:, :n_tx], dtype=np.float64)
    return float(np.mean(np.logaddexp(0.0, x)))


def fd_grad(code, llr, alpha, beta, T, which, indices, eps=1e-3, corr=False, r_max=0.0):
    """Central finite differences (float64) of the loss w.r.t. alpha or beta
    entries given as (g, p) pairs."""
    a64 = np.asarray(alpha, np.float64)
    b64 = np.asarray(beta, np.float64)
    out = []
    for g, p in indices:
        vals = []
        for sgn in (1.0, -1.0):
            a, b = a64.copy(), b64.copy()
            (a if which == "alpha" else b)[g, p] += sgn * eps
            vals.append(loss(forward(code, llr, a, b, T, corr, np.float64, r_max), code.n_tx))
        out.append((vals[0] - vals[1]) / (2 * eps))
    return np.array(out)


def _syndrome_fail(L, sch):
    """Per codeword: some check has odd parity of hard decisions (L < 0)."""
    hard = L < 0
    fail = np.zeros(L.shape[0], dtype=bool)
    for deg, z, var, _, _ in sch:
        par = np.bitwise_xor.reduce(hard[:, var], axis=1)
        fail |= par.any(axis=1)
    return fail


def decode_d(code, llr_bf16, alpha, beta, max_iter, corr=False, r_max=0.0, return_float=False):
    """Profile D (learned min-sum, float32) with the syndrome early exit:
    bf16 channel LLRs in; returns (final L as RNE bf16 bits, iterations,
    converged[, final float32 L]).  Row t of alpha / beta is used at
    iteration t (the last row beyond); converged codewords are frozen."""
    dt = np.float32
    sch = schedule(code)
    E = sum(l[0] * l[1] for l in sch)
    llr = (np.asarray(llr_bf16, np.uint16).astype(np.uint32) << 16).view(np.float32).reshape(-1, code.n)
    B = llr.shape[0]
    L = llr.copy()
    R = np.zeros((B, E), dt)
    alpha = np.asarray(alpha, dt).reshape(-1, sum(l[0] for l in sch))
    beta = np.asarray(beta, dt).reshape(alpha.shape)
    iters = np.full(B, max_iter, np.uint32)
    ok = np.zeros(B, np.uint8)
    active = np.arange(B)
    for t in range(max_iter):
        if active.size == 0:
            break
        g = min(t, alpha.shape[0] - 1)
        La, Ra = L[active], R[active]
        for lay in sch:
            _layer(La, Ra, lay, alpha[g], beta[g], corr, dt, r_max=r_max)
        L[active], R[active] = La, Ra
        conv = ~_syndrome_fail(La, sch)
        iters[active[conv]] = t + 1
        ok[active[conv]] = 1
        active = active[~conv]
    u = L.view(np.uint32)
    out = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return (out, iters, ok, L) if return_float else (out, iters, ok)


def run_point_d(code, llr_bf16, alpha, beta, max_iter, corr=False, r_max=0.0):
    """Monte Carlo tallies of Profile D on given channel LLRs: dict(blocks,
    frame_errors, bit_errors, iter_sum, hist) as the Profile T tallies (bit
    errors = information bits <= 0 of failed codewords, from the float32 L)."""
    _, iters, ok, L = decode_d(code, llr_bf16, alpha, beta, max_iter, corr, r_max, return_float=True)
    hist = np.zeros(max_iter + 1, np.uint64)
    for it, o in zip(iters, ok):
        hist[it if o else 0] += 1
    fail = ok == 0
    be = int((L[fail, :code.k] <= 0).sum())
    return dict(blocks=len(ok), frame_errors=int(fail.sum()), bit_errors=be, iter_sum=int(iters.sum()), hist=hist)
