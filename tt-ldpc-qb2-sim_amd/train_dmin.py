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


# Mining draws from its own seed, so that no mined codeword is one that
# eval_dmin.py (seed 0x5EED2026) judges the decoder on.
MINE_SEED = 0x4A2D5EED


def mine_hard(model, args, step, hard, mined):
    """Decode args.hard_pool fresh codewords at args.hard_ebn0 with the current
    parameters (libldpc Profile D: plain Monte Carlo with early exit), capture
    the ids of the failures and regenerate their channels (bf16 LLRs, the
    channel the evaluation uses); they join the mined set, which keeps the
    newest args.hard_keep."""
    from ldpc_mi355x import decoder
    p = model.export()
    first = (step + 1) * args.hard_pool  # a new pool of codeword ids each round
    with decoder.Engine(args.k, args.rate, max_iter=args.hard_iters, profile="D") as eng:
        eng.set_dmin_params(p["alpha"], p["beta"], int(p["correction"]), float(p["r_max"]))
        eng.set_capture(args.hard_keep)
        r = eng.run_point(args.hard_ebn0, args.hard_pool, seed=MINE_SEED, first_cw=first, chunk_cw=1 << 20)
        ids, seen = eng.captures(args.hard_keep)
    with decoder.Engine(args.k, args.rate, max_iter=1) as gen:
        rows = [gen.gen_channel(MINE_SEED, int(i), 1, args.hard_ebn0) for i in ids]
    new = None
    if rows:
        llr16 = torch.from_numpy(np.concatenate(rows).astype(np.int32))
        new = (llr16 << 16).view(torch.float32).cuda()
    if new is not None:
        hard = new if hard is None else torch.cat([hard, new])[-args.hard_keep:]
    return hard, {"pool": int(r.blocks), "failures": int(seen), "fer": r.fer,
                  "set": 0 if hard is None else int(hard.shape[0])}


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
    ap.add_argument("--correction", type=int, nargs="?", const=1, default=0, choices=[0, 1, 2],
                    help="Approx-Min* correction on the min-sum magnitude: 1 (the flag alone) from the two "
                         "smallest magnitudes, 2 from three (every edge's pair excludes the edge itself)")
    ap.add_argument("--per-iteration", action="store_true", help="one parameter row per iteration")
    ap.add_argument("--bf16-llr", action="store_true", help="round channel LLRs to bf16")
    ap.add_argument("--multi-loss", action="store_true", help="average the loss over every iteration's output")
    ap.add_argument("--loss", choices=["bce", "softfer"], default="bce",
                    help="bce: mean binary cross-entropy of the soft output; softfer: soft frame error rate, "
                         "1 - prod sigmoid(L) per codeword")
    ap.add_argument("--trunc", type=int, default=0,
                    help="truncate backpropagation every TRUNC iterations (0: through all T)")
    ap.add_argument("--val-ebn0", type=float, nargs="*", default=[1.5, 2.0])
    ap.add_argument("--val-size", type=int, default=32768)
    ap.add_argument("--val-every", type=int, default=25)
    ap.add_argument("--init-alpha", type=float, default=1.0)
    ap.add_argument("--init-beta", type=float, default=0.0)
    ap.add_argument("--init-from", default=None,
                    help=".npz written by this script: start from its alpha / beta (same code and parameter shape)")
    ap.add_argument("--hard-ebn0", type=float, default=0.0,
                    help="mine hard codewords at this Eb/N0 (0: off): every --val-every steps the current "
                         "parameters decode --hard-pool codewords by plain Monte Carlo (libldpc Profile D, "
                         "--hard-iters iterations, early exit) and the failures' channels join the training set")
    ap.add_argument("--hard-pool", type=int, default=1 << 23, help="codewords decoded per mining round")
    ap.add_argument("--hard-iters", type=int, default=100, help="iteration cap of the mining decode")
    ap.add_argument("--hard-frac", type=float, default=0.25, help="share of each micro-batch drawn from the mined set")
    ap.add_argument("--hard-keep", type=int, default=16384, help="most mined codewords kept (the newest)")
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
    hard, mined = None, {}
    while done < args.steps:
        n = min(args.val_every, args.steps - done)
        if args.hard_ebn0 > 0:
            hard, mined = mine_hard(model, args, done, hard, mined)
        log = dmin.train(model, n, batch=args.batch, micro_batch=args.micro_batch, lr=args.lr,
                         ebn0_range=(args.ebn0_lo, args.ebn0_hi), seed=args.seed + done, bf16_llr=args.bf16_llr,
                         log_every=max(1, n), opt=opt, multi_loss=args.multi_loss, hard=hard,
                         hard_frac=args.hard_frac, loss=args.loss)
        done += n
        torch.cuda.synchronize()
        print(json.dumps({"step": done, "train_loss": log[-1].loss, "val": validate(model, sets),
                          "mined": mined, "seconds": time.time() - t0}), flush=True)
    p = model.export()
    np.savez(args.out, alpha=p["alpha"], beta=p["beta"], T=p["T"], k_info=p["k_info"], rate=p["rate"],
             correction=p["correction"], r_max=p["r_max"])
    print(json.dumps({"saved": args.out, "seconds": time.time() - t0}), flush=True)


if __name__ == "__main__":
    main()
