"""PyTorch-autograd restatement of the differentiable layered min-sum decoder
(the reference gradient for the HIP backward kernel).  Test helper only.

Same arithmetic as oracle/dmin_oracle.py, written with differentiable torch
ops whose subgradients follow the kernel's rules: first-index argmin / second
(/ third) minimum, where() selecting one branch (relu at z > 0, correction line 1 when
l1 >= l2, clamp at a positive corrected minimum), d|q|/dq = sgn(q) with
sgn(0) = 0, sign bits carrying no gradient.
"""
import torch

C1, K1, C2, K2 = 0.69140625, -0.3984375, 0.65625, -0.326171875


def layers_of(code, device):
    out, e_off, p_off = [], 0, 0
    for _, slots in code.layers:
        z = code.z
        j = torch.arange(z, device=device)
        var = torch.stack([s.base + (j + s.shift) % z for s in slots])
        out.append((len(slots), z, var, e_off, p_off))
        e_off += len(slots) * z
        p_off += len(slots)
    return out, e_off


def _amin(lo, hi):
    """Approx-Min* of two magnitudes lo <= hi: max(lo - corr(hi - lo), 0),
    the gradient through the selected line and the positive branch only."""
    x = hi - lo
    l1 = K1 * x + C1
    l2 = K2 * x + C2
    lm = torch.where(l1 >= l2, l1, l2)
    c = torch.where(lm > 0, lm, torch.zeros_like(lm))
    o = lo - c
    return torch.where(o > 0, o, torch.zeros_like(o))


def forward(code, llr, alpha, beta, T, corr=False, r_max=0.0, all_iters=False, trunc=0):
    """corr: 0 / False min-sum, 1 / True two-minimum correction, 2 three minima."""
    lays, E = layers_of(code, llr.device)
    B = llr.shape[0]
    L = llr
    R = torch.zeros(B, E, dtype=llr.dtype, device=llr.device)
    outs = []
    for t in range(T):
        g = t if alpha.shape[0] == T else 0
        for deg, z, var, e_off, p_off in lays:
            Rl = R[:, e_off:e_off + deg * z].reshape(B, deg, z)
            q = L[:, var] - Rl
            a = q.abs()
            idx1 = a.argmin(dim=1, keepdim=True)
            m1 = a.gather(1, idx1)
            a2 = a.detach().scatter(1, idx1, float("inf"))
            idx2 = a2.argmin(dim=1, keepdim=True)
            m2 = a.gather(1, idx2)
            dix = torch.arange(deg, device=llr.device)[None, :, None]
            if corr == 2:
                a3 = a2.scatter(1, idx2, float("inf"))
                idx3 = a3.argmin(dim=1, keepdim=True)
                m3 = a.gather(1, idx3)
                gm = torch.where(dix == idx1, _amin(m2, m3), torch.where(dix == idx2, _amin(m1, m3), _amin(m1, m2)))
            else:
                gm = torch.where(dix == idx1, m2, _amin(m1, m2) if corr else m1)
            zz = alpha[g, p_off:p_off + deg][None, :, None] * gm - beta[g, p_off:p_off + deg][None, :, None]
            mag = torch.where(zz > 0, zz, torch.zeros_like(zz))
            if r_max > 0:
                mag = torch.where(mag < r_max, mag, torch.full_like(mag, r_max))
            sb = torch.signbit(q.detach())
            s = (sb.sum(dim=1, keepdim=True) % 2).bool()
            r = torch.where(s ^ sb, -mag, mag)
            Lnew = q + r
            L = L.scatter(1, var.reshape(1, -1).expand(B, -1), Lnew.reshape(B, -1))
            R = torch.cat([R[:, :e_off], r.reshape(B, -1), R[:, e_off + deg * z:]], dim=1)
        outs.append(L)
        if trunc and (t + 1) % trunc == 0 and t < T - 1:
            L, R = L.detach(), R.detach()
    return torch.stack(outs) if all_iters else L


def loss(L, n_tx):
    return torch.nn.functional.softplus(-L[..., :n_tx]).mean()
