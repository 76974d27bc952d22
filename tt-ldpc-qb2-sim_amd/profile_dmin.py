#!/usr/bin/env python3
"""HBM roofline of the differentiable decoder's kernels (training path).

The training forward (k_fwd) and backward (k_bwd) of ldpc_mi355x.dmin run one
launch per layer step with one thread per (codeword, check); L [B, N], R
[B, E], the tape and the adjoints live in HBM as fp32, so every kernel is a
streaming pass whose bound is HBM bandwidth.  Algorithmic bytes per edge of
a layer step:

    forward, no tape  (T x 12 launches per forward)      16 B: L read + write, R read + write
    forward with tape (recompute inside the backward)    20 B: the same + the tape write
    backward                                             20 B: tape read, adjoint of L read + write,
                                                              adjoint of R read + write

This tool times each phase with HIP events on torch's stream (warm, several
repetitions), divides the algorithmic bytes by the time and prints one JSON
line with the achieved GB/s and the fraction of the MI355X's 8 TB/s.  Run it
under ``rocprofv3 --kernel-trace --stats`` for per-kernel durations.

    python tt-ldpc-qb2-sim_amd/profile_dmin.py --k 1024 --T 20 --batch 16384 --ckpt-every 4
    python tt-ldpc-qb2-sim_amd/profile_dmin.py --k 16384 --T 50 --batch 1024 --ckpt-every 60
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import torch  # noqa: E402

from ldpc_mi355x import dmin  # noqa: E402

HBM_PEAK_GBS = 8000.0


def timed(fn, reps):
    fn()  # warm
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        out = fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / 1e3 / reps, out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--k", type=int, default=1024)
    ap.add_argument("--rate", default="1/2")
    ap.add_argument("--T", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16384, help="codewords per forward/backward (one micro-batch)")
    ap.add_argument("--ckpt-every", type=int, default=4)
    ap.add_argument("--correction", action="store_true", default=True)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)

    from ldpc_mi355x.dmin import _ext
    model = dmin.DMinSum(args.k, args.rate, T=args.T, ckpt_every=args.ckpt_every, correction=args.correction)
    code = model.code
    B, T, nl = args.batch, args.T, len(code.layers)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    llr = dmin.channel_llr(code, torch.full((B,), 1.5, device="cuda"), gen)
    tab = dmin.schedule_table(code)
    a, b = model.alpha.detach().contiguous(), model.beta.detach().contiguous()
    E = code.n_edges
    steps = T * nl

    t_fwd, _ = timed(lambda: _ext.forward(llr, a, b, tab, T, args.ckpt_every, args.correction, False, 0.0, False),
                     args.reps)
    t_fwd_ck, outs = timed(lambda: _ext.forward(llr, a, b, tab, T, args.ckpt_every, args.correction, True, 0.0,
                                                False), args.reps)
    grad = torch.randn_like(outs[0])
    ckpts = outs[1:]
    t_bwd, _ = timed(lambda: _ext.backward(grad, ckpts, a, b, tab, T, args.ckpt_every, args.correction, 0.0, 0),
                     args.reps)
    n_seg = len(ckpts) // 2
    ck_bytes = n_seg * B * (code.n + E) * 4 * 2            # clone: read + write of (L, R)
    fwd_bytes = T * B * E * 16
    # the backward recomputes every segment with the tape, then runs k_bwd per step,
    # and clones each checkpoint once more before recomputing from it
    bwd_bytes = T * B * E * 20 + T * B * E * 20 + ck_bytes
    res = {
        "code": code.name, "T": T, "batch": B, "ckpt_every": args.ckpt_every, "layer_steps": steps,
        "forward": {"seconds": t_fwd, "ms_per_launch": t_fwd / steps * 1e3, "bytes": fwd_bytes,
                    "gbs": fwd_bytes / t_fwd / 1e9, "frac_hbm": fwd_bytes / t_fwd / 1e9 / HBM_PEAK_GBS},
        "forward_with_checkpoints": {"seconds": t_fwd_ck, "bytes": fwd_bytes + ck_bytes,
                                     "gbs": (fwd_bytes + ck_bytes) / t_fwd_ck / 1e9,
                                     "frac_hbm": (fwd_bytes + ck_bytes) / t_fwd_ck / 1e9 / HBM_PEAK_GBS},
        "backward": {"seconds": t_bwd, "bytes": bwd_bytes, "gbs": bwd_bytes / t_bwd / 1e9,
                     "frac_hbm": bwd_bytes / t_bwd / 1e9 / HBM_PEAK_GBS},
        "gsps_forward": B * code.n_tx / t_fwd / 1e9,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
