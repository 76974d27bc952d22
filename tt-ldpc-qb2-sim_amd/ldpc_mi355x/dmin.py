"""Differentiable layered min-sum decoder on the GPU (PyTorch).

An unrolled, layered offset / normalized min-sum decoder for the AR4JA codes
with learnable FP32 scale (alpha) and offset (beta) per (layer, circulant):
one pair per edge slot of the layered schedule, i.e. per circulant of the
quarter-lifted protograph (60 of each for rate 1/2), shared by all
iterations or one row per iteration.  The check update is

    r_d = copysign(min(max(alpha_p * g_d - beta_p, 0), r_max), s ^ sign q_d)

with g_d the leave-one-out minimum (second minimum on the argmin edge),
optionally with the Approx-Min* correction of the reference's golden tree
(bf16-truncated constants, sfpi_host_emu.h:566-584): correction 1 corrects
the minimum the other edges get with the second; correction 2 keeps three
minima and gives every edge min* of the two smallest magnitudes among the
others (min*(min2, min3) on the argmin edge, min*(min1, min3) on the second).  Forward and backward
are hand-written HIP kernels (csrc/dmin_torch.hip) behind a
``torch.autograd.Function``; activations are recomputed from (L, R)
checkpoints taken every ``ckpt_every`` layer steps.  The reference has no
counterpart of this decoder; it is evaluated by the plain Monte Carlo FER
of the trained parameters (libldpc Profile D).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np
import torch

from . import codes

try:
    from . import ldpc_dmin as _ext
except ImportError as e:  # no fallback: the decoder only exists as HIP kernels
    raise ImportError("ldpc_mi355x.ldpc_dmin is not built: run `python -m ldpc_mi355x.build` "
                      "(the HIP extension is required; there is no CPU fallback)") from e

MAXD = _ext.MAXD


def schedule_table(code) -> torch.Tensor:
    """int32 CPU tensor [n_layers, 2 + 2*MAXD]: (deg, z, base_0, shift_0, ...)."""
    tab = np.zeros((len(code.layers), 2 + 2 * MAXD), dtype=np.int32)
    for i, (_, slots) in enumerate(code.layers):
        tab[i, 0] = len(slots)
        tab[i, 1] = code.z
        for d, s in enumerate(slots):
            tab[i, 2 + 2 * d] = s.base
            tab[i, 3 + 2 * d] = s.shift
    return torch.from_numpy(tab)


def n_params(code) -> int:
    return sum(len(s) for _, s in code.layers)


class DMinSumFunction(torch.autograd.Function):
    """(llr [B, N], alpha [G, P], beta [G, P]) -> soft output L [B, N] after T
    iterations, or after every iteration [T, B, N] if all_iters.  Gradients
    flow to all three inputs."""

    @staticmethod
    def forward(ctx, llr, alpha, beta, tab, T, ckpt_every, corr, r_max=0.0, all_iters=False, trunc=0):
        need = ctx.needs_input_grad[0] or ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        corr = int(corr)
        outs = _ext.forward(llr.contiguous(), alpha.contiguous(), beta.contiguous(), tab, T, ckpt_every, corr, need,
                            float(r_max), bool(all_iters))
        ctx.ckpts = outs[1:]
        ctx.cfg = (tab, T, ckpt_every, corr, float(r_max), int(trunc))
        ctx.save_for_backward(alpha, beta)
        return outs[0]

    @staticmethod
    def backward(ctx, grad_out):
        alpha, beta = ctx.saved_tensors
        tab, T, ckpt_every, corr, r_max, trunc = ctx.cfg
        g_llr, g_alpha, g_beta = _ext.backward(grad_out.contiguous(), list(ctx.ckpts), alpha.contiguous(),
                                               beta.contiguous(), tab, T, ckpt_every, corr, r_max, trunc)
        ctx.ckpts = None
        return g_llr, g_alpha, g_beta, None, None, None, None, None, None, None


class DMinSum(torch.nn.Module):
    """Learnable layered min-sum decoder for one AR4JA code."""

    def __init__(self, k_info: int = 1024, rate: str = "1/2", T: int = 20, ckpt_every: int = 4,
                 correction: int = 0, per_iteration: bool = False, r_max: float = 0.0, trunc: int = 0,
                 device="cuda"):
        """correction: 0 (or False) plain min-sum, 1 (or True) the two-minimum
        Approx-Min* correction, 2 the three-minimum one."""
        super().__init__()
        correction = int(correction)
        if correction not in (0, 1, 2):
            raise ValueError("correction must be 0, 1 or 2")
        self.code = codes.ar4ja(k_info, rate)
        self.k_info, self.rate, self.T = k_info, rate, T
        self.ckpt_every, self.correction, self.r_max, self.trunc = ckpt_every, correction, r_max, trunc
        self.tab = schedule_table(self.code)
        p = n_params(self.code)
        g = T if per_iteration else 1
        self.alpha = torch.nn.Parameter(torch.ones(g, p, device=device))
        self.beta = torch.nn.Parameter(torch.zeros(g, p, device=device))

    @property
    def n(self) -> int:
        return self.code.n

    @property
    def n_tx(self) -> int:
        return self.code.n_tx

    def forward(self, llr: torch.Tensor, all_iters: bool = False) -> torch.Tensor:
        """Soft output after T iterations [B, N], or after each [T, B, N].
        With gradients disabled (torch.no_grad) no checkpoints are kept: the
        autograd.Function would keep them whenever the parameters require
        grad, T * 12 / ckpt_every copies of (L, R)."""
        if not torch.is_grad_enabled():
            return _ext.forward(llr.contiguous(), self.alpha.detach().contiguous(), self.beta.detach().contiguous(),
                                self.tab, self.T, self.ckpt_every, int(self.correction), False, float(self.r_max),
                                bool(all_iters))[0]
        return DMinSumFunction.apply(llr, self.alpha, self.beta, self.tab, self.T, self.ckpt_every,
                                     self.correction, self.r_max, all_iters, self.trunc)

    def export(self) -> dict:
        """Parameters for the Monte Carlo evaluation (libldpc Profile D)."""
        return dict(k_info=self.k_info, rate=self.rate, T=self.T, correction=self.correction, r_max=self.r_max,
                    alpha=self.alpha.detach().float().cpu().numpy(), beta=self.beta.detach().float().cpu().numpy())


def soft_bce(L: torch.Tensor, n_tx: int) -> torch.Tensor:
    """Binary cross-entropy of the soft output against the all-zero
    codeword's transmitted bits: mean softplus(-L) over positions < n_tx
    (and over iterations for a [T, B, N] output: the multi-loss)."""
    return torch.nn.functional.softplus(-L[..., :n_tx]).mean()


def soft_fer(L: torch.Tensor, n_tx: int) -> torch.Tensor:
    """Soft frame error rate: 1 - prod over positions < n_tx of sigmoid(L),
    i.e. 1 - exp(-sum softplus(-L)), averaged over codewords (and over
    iterations for a [T, B, N] output).  A codeword with many unreliable bits
    saturates at 1 and contributes little gradient, so borderline codewords
    weigh most."""
    s = torch.nn.functional.softplus(-L[..., :n_tx]).sum(dim=-1)
    return (-torch.expm1(-s)).mean()


def channel_llr(code, ebn0_db: torch.Tensor, generator: Optional[torch.Generator] = None,
                bf16: bool = False) -> torch.Tensor:
    """BPSK LLRs of the all-zero codeword, one Eb/N0 per codeword (AWGN,
    mu = 4 R g, sigma = sqrt(8 R g), R = k / n_tx), punctured = 0."""
    B = ebn0_db.numel()
    g = torch.pow(10.0, ebn0_db.float() / 10.0)[:, None]
    r = code.k / code.n_tx
    z = torch.randn(B, code.n_tx, device=ebn0_db.device, generator=generator)
    llr = torch.zeros(B, code.n, device=ebn0_db.device)
    llr[:, :code.n_tx] = 4 * r * g + torch.sqrt(8 * r * g) * z
    if bf16:
        llr = llr.to(torch.bfloat16).float()
    return llr


@dataclasses.dataclass
class TrainLog:
    step: int
    loss: float
    seconds: float


def train(model: DMinSum, steps: int, batch: int = 65536, micro_batch: int = 16384, lr: float = 1e-3,
          ebn0_range=(1.0, 2.5), seed: int = 0, bf16_llr: bool = False, log_every: int = 10,
          callback=None, opt: Optional[torch.optim.Optimizer] = None, multi_loss: bool = False,
          hard: Optional[torch.Tensor] = None, hard_frac: float = 0.0, loss: str = "bce") -> List[TrainLog]:
    """Adam on the soft BCE (of the last iteration, or averaged over all
    iterations with `multi_loss`); every step draws `batch` codewords at
    Eb/N0 uniform in `ebn0_range`, processed in micro-batches (gradients
    summed).  Pass `opt` to continue with an existing optimizer's state.

    `hard` [H, N] (channel LLRs of codewords a decoder failed on) with
    `hard_frac` > 0: that share of every micro-batch is drawn from `hard`
    (uniformly, with replacement) instead of the channel.  `loss`: "bce"
    (soft_bce) or "softfer" (soft_fer)."""
    dev = model.alpha.device
    if opt is None:
        opt = torch.optim.Adam(model.parameters(), lr=lr)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    log: List[TrainLog] = []
    import time
    t0 = time.time()
    n_micro = max(1, math.ceil(batch / micro_batch))
    for step in range(steps):
        opt.zero_grad(set_to_none=True)
        total = 0.0
        for m in range(n_micro):
            b = min(micro_batch, batch - m * micro_batch)
            eb = ebn0_range[0] + (ebn0_range[1] - ebn0_range[0]) * torch.rand(b, device=dev, generator=gen)
            llr = channel_llr(model.code, eb, gen, bf16_llr)
            nh = int(round(hard_frac * b)) if hard is not None and hard.shape[0] else 0
            if nh:
                pick = torch.randint(0, hard.shape[0], (nh,), device=dev, generator=gen)
                llr[:nh] = hard[pick]
            out = model(llr, all_iters=multi_loss)
            val = (soft_fer(out, model.n_tx) if loss == "softfer" else soft_bce(out, model.n_tx)) * (b / batch)
            val.backward()
            total += float(val.detach())
        opt.step()
        if step % log_every == 0 or step == steps - 1:
            log.append(TrainLog(step, total, time.time() - t0))
            if callback:
                callback(log[-1])
    return log
