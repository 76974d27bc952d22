#!/usr/bin/env python3
"""Train the differentiable layered min-sum decoder on one GPU.

Defaults follow the small training configuration: AR4JA r1/2 K=1024, T = 20
unrolled iterations, one FP32 scale and offset per (layer, circulant),
initialised at 1 / 0, batches of 65536 codewords at Eb/N0 uniform in
[1.0, 2.5] dB, binary cross-entropy of the soft output against the
transmitted (all-zero) bits, Adam at 1e-3, checkpoints every 4 layer steps.
The large configuration is ``--k 16384 --T 50 --ckpt-every 60 --batch 4096
--micro-batch 1024 --bf16-llr`` (recompute from a checkpoint every 5
iterations).

Writes the parameters (numpy .npz: alpha, beta and the decoder settings)
and prints one JSON line per validation.  The validation set is fixed
(seeded) at each ``--val-ebn0`` point: soft BCE and the frame error rate of
the hard decisions after T iterations.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import torch  # noqa: E402

from ldpc_mi355x import dmin  # noqa: E402


def validate(model, sets):
    out = {}
    with torch.no_grad():
        for eb, llr in sets.items():
            loss, fe = 0.0, 0
            for x in llr.split(8192):
                L = model(x)
                loss += float(dmin.soft_bce(L, model.n_tx)) * x.shape[0]
                fe += int(((L[:, :model.code.k] <= 0).any(dim=1)).sum())
            out[f"{eb:.2f}"] = {"bce": loss / llr.shape[0], "fer": fe / llr.shape[0]}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--k", type=int, default=1024)
    ap.add_argument("--rate", default="1/2")
    ap.add_argument("--T", type=int, default=20)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--micro-batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--ebn0-lo", type=float, default=1.0)
    ap.add_argument("--ebn0-hi", type=float, default=2.5)
    ap.add_argument("--ckpt-every", type=int, default=4)
    ap.add_argument("--r-max", type=float, default=0.0, help="cap on check-message magnitude (0: none)")
    ap.add_argument("--correction", action="store_true", help="Approx-Min* correction on the min-sum magnitude")
    ap.add_argument("--per-iteration", action="store_true", help="one parameter row per iteration")
    ap.add_argument("--bf16-llr", action="store_true", help="round channel LLRs to bf16")
    ap.add_argument("--multi-loss", action="store_true", help="average the BCE over every iteration's output")
    ap.add_argument("--trunc", type=int, default=0,
                    help="truncate backpropagation every TRUNC iterations (0: through all T)")
    ap.add_argument("--val-ebn0", type=float, nargs="*", default=[1.5, 2.0])
    ap.add_argument("--val-size", type=int, default=32768)
    ap.add_argument("--val-every", type=int, default=25)
    ap.add_argument("--init-alpha", type=float, default=1.0)
    ap.add_argument("--init-beta", type=float, default=0.0)
    ap.add_argument("--init-from", default=None,
                    help=".npz written by this script: start from its alpha / beta (same code and parameter shape)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default="dmin_params.npz")
    args = ap.parse_args(argv)

    model = dmin.DMinSum(args.k, args.rate, T=args.T, ckpt_every=args.ckpt_every, correction=args.correction,
                         per_iteration=args.per_iteration, r_max=args.r_max, trunc=args.trunc)
    with torch.no_grad():
        model.alpha.fill_(args.init_alpha)
        model.beta.fill_(args.init_beta)
        if args.init_from:
            z = np.load(args.init_from, allow_pickle=False)
            if tuple(z["alpha"].shape) != tuple(model.alpha.shape):
                raise SystemExit(f"--init-from: parameter shape {z['alpha'].shape} does not match {tuple(model.alpha.shape)}")
            model.alpha.copy_(torch.from_numpy(z["alpha"]))
            model.beta.copy_(torch.from_numpy(z["beta"]))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(12345)
    sets = {eb: dmin.channel_llr(model.code, torch.full((args.val_size,), eb, device="cuda"), gen, args.bf16_llr)
            for eb in args.val_ebn0}
    opt = torch.optim.Adam(model.parameters(), lr=args.lr)
    t0 = time.time()
    print(json.dumps({"step": 0, "val": validate(model, sets), "seconds": 0.0}), flush=True)
    done = 0
    while done < args.steps:
        n = min(args.val_every, args.steps - done)
        log = dmin.train(model, n, batch=args.batch, micro_batch=args.micro_batch, lr=args.lr,
                         ebn0_range=(args.ebn0_lo, args.ebn0_hi), seed=args.seed + done, bf16_llr=args.bf16_llr,
                         log_every=max(1, n), opt=opt, multi_loss=args.multi_loss)
        done += n
        torch.cuda.synchronize()
        print(json.dumps({"step": done, "train_loss": log[-1].loss, "val": validate(model, sets),
                          "seconds": time.time() - t0}), flush=True)
    p = model.export()
    np.savez(args.out, alpha=p["alpha"], beta=p["beta"], T=p["T"], k_info=p["k_info"], rate=p["rate"],
             correction=p["correction"], r_max=p["r_max"])
    print(json.dumps({"saved": args.out, "seconds": time.time() - t0}), flush=True)


if __name__ == "__main__":
    main()
