#!/usr/bin/env python3
"""Evaluate a trained layered min-sum decoder by plain Monte Carlo on one GPU.

Runs libldpc Profile D (the trained parameters, syndrome early exit) at each
Eb/N0 for the given iteration caps, and Profile T (the golden Approx-Min*)
at 200 iterations (or ``--ref-iters``) on the same codeword ids and seed --
the same channel realisations.  Codes without a Profile D kernel (K=16384
r1/2, whose fp32 codeword does not fit in LDS) are decoded by the
differentiable decoder's own forward kernels instead (``--forward``: T
iterations, no early exit, frame error = an information bit <= 0), on the
channel libldpc generates for the same ids.

Prints one JSON line per (decoder, Eb/N0) with FER, its 95% interval, mean
iterations and Gsps, then one summary line with the Eb/N0 gap at equal FER
(log-linear interpolation of each FER curve).
"""
import argparse
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

from ldpc_mi355x import decoder  # noqa: E402


def run_forward(z, ebn0, blocks, seed, target_fe, chunk=1024, T=0):
    """The trained decoder through the extension's forward (no early exit),
    T iterations (0: the T it was trained with)."""
    import time

    import torch  # loaded before libldpc's first use in this process (same HIP runtime)
    from ldpc_mi355x import dmin
    k, rate, T = int(z["k_info"]), str(z["rate"]), (T or int(z["T"]))
    if z["alpha"].shape[0] > 1 and T != int(z["T"]):
        raise SystemExit("--forward-T needs shared (not per-iteration) parameters")
    model = dmin.DMinSum(k, rate, T=T, correction=int(z["correction"]), r_max=float(z["r_max"]),
                         per_iteration=z["alpha"].shape[0] > 1)
    with torch.no_grad():
        model.alpha.copy_(torch.from_numpy(z["alpha"]))
        model.beta.copy_(torch.from_numpy(z["beta"]))
    gen = decoder.Engine(k, rate, max_iter=1)
    fe = be = done = 0
    t0 = time.time()
    while done < blocks and (not target_fe or fe < target_fe):
        n = min(chunk, blocks - done)
        llr16 = torch.from_numpy(gen.gen_channel(seed, done, n, ebn0).astype(np.int32))
        llr = (llr16 << 16).view(torch.float32).cuda()
        with torch.no_grad():
            L = model(llr)
        bad = (L[:, :k] <= 0)
        fe += int(bad.any(dim=1).sum())
        be += int(bad.sum())
        done += n
    torch.cuda.synchronize()
    gen.close()
    return done, fe, be, time.time() - t0, model.n_tx


def poisson_ci95(k):
    from scipy.stats import chi2
    return (0.0 if k == 0 else chi2.ppf(0.025, 2 * k) / 2), chi2.ppf(0.975, 2 * k + 2) / 2


def ebn0_at(points, fer):
    """Eb/N0 where a FER curve [(ebn0, fer)] crosses `fer` (log-linear)."""
    pts = [(e, f) for e, f in points if f > 0]
    for (e0, f0), (e1, f1) in zip(pts, pts[1:]):
        if f0 >= fer >= f1 and f0 > f1:
            return e0 + (math.log(f0) - math.log(fer)) / (math.log(f0) - math.log(f1)) * (e1 - e0)
    return None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("params", help=".npz written by train_dmin.py")
    ap.add_argument("--ebn0", type=float, nargs="+", default=[1.0, 1.25, 1.5, 1.75, 2.0])
    ap.add_argument("--blocks", type=int, default=1 << 22, help="codewords per point")
    ap.add_argument("--target-fe", type=int, default=0, help="stop a point early after this many frame errors")
    ap.add_argument("--iters", type=int, nargs="+", default=[20, 100], help="Profile D iteration caps")
    ap.add_argument("--seed", type=int, default=0x5EED2026)
    ap.add_argument("--first-cw", type=int, default=0,
                    help="first codeword id of every Monte Carlo point (continue an earlier run of the same seed)")
    ap.add_argument("--ref-iters", type=int, nargs="+", default=[200], help="Approx-Min* iteration caps")
    ap.add_argument("--forward", action="store_true",
                    help="decode the trained decoder with its training forward (T iterations, no early exit)")
    ap.add_argument("--forward-T", type=int, default=0, help="iterations of the --forward decode (0: the trained T)")
    ap.add_argument("--forward-chunk", type=int, default=1024, help="codewords per --forward decode call")
    ap.add_argument("--segment", type=int, default=1 << 24,
                    help="codewords per progress line (a point runs as consecutive segments of codeword ids)")
    ap.add_argument("--decoders", nargs="+", choices=["trained", "approx_min_star"],
                    default=["trained", "approx_min_star"], help="which decoders to run")
    args = ap.parse_args(argv)

    z = np.load(args.params, allow_pickle=False)
    k, rate = int(z["k_info"]), str(z["rate"])
    curves = {}
    if args.forward:
        import torch  # noqa: F401  (before libldpc: one HIP runtime for both)
        pts = []
        t_fwd = args.forward_T or int(z["T"])
        for eb in args.ebn0:
            done, fe, be, sec, n_tx = run_forward(z, eb, args.blocks, args.seed, args.target_fe,
                                                  chunk=args.forward_chunk, T=t_fwd)
            lo, hi = poisson_ci95(fe)
            pts.append((eb, fe / done))
            print(json.dumps({"decoder": f"trained_forward_T{t_fwd}", "max_iter": t_fwd, "ebn0_db": eb,
                              "blocks": done, "frame_errors": fe, "fer": fe / done,
                              "fer_ci95": [lo / done, hi / done], "ber": be / (done * k),
                              "gsps": done * n_tx / sec / 1e9}), flush=True)
        curves[f"trained_forward_T{t_fwd}"] = pts
    runs = []
    if "approx_min_star" in args.decoders:
        runs += [(f"approx_min_star_{r}" if r != 200 else "approx_min_star", r, None) for r in args.ref_iters]
    if not args.forward and "trained" in args.decoders:
        runs += [(f"trained_T{t}", t, t) for t in args.iters]
    for name, max_iter, _ in runs:
        if name.startswith("approx_min_star"):
            eng = decoder.Engine(k, rate, max_iter=max_iter)
        else:
            eng = decoder.Engine(k, rate, max_iter=max_iter, profile="D")
            eng.set_dmin_params(z["alpha"], z["beta"], int(z["correction"]), float(z["r_max"]))
        pts = []
        for eb in args.ebn0:
            r = None
            while r is None or (r.blocks < args.blocks and not (args.target_fe and r.frame_errors >= args.target_fe)):
                n = min(args.segment, args.blocks - (r.blocks if r else 0))
                fe_left = max(args.target_fe - (r.frame_errors if r else 0), 1) if args.target_fe else 0
                r = eng.run_point(eb, n, target_fe=fe_left, seed=args.seed, first_cw=r.next_cw if r else args.first_cw,
                                  chunk_cw=1 << 20, result=r)
                print(f"# {name} {eb} dB: {r.blocks} codewords, {r.frame_errors} FE, {r.kernel_gsps:.1f} Gsps",
                      file=sys.stderr, flush=True)
            lo, hi = poisson_ci95(r.frame_errors)
            pts.append((eb, r.fer))
            print(json.dumps({"decoder": name, "max_iter": max_iter, "ebn0_db": eb, "blocks": r.blocks,
                              "frame_errors": r.frame_errors, "fer": r.fer,
                              "fer_ci95": [lo / r.blocks, hi / r.blocks], "ber": r.ber,
                              "mean_iterations": r.mean_iterations, "kernel_gsps": r.kernel_gsps}), flush=True)
        curves[name] = pts
        eng.close()
    refs = [n for n in curves if n.startswith("approx_min_star")]
    if not refs:
        return
    ref_name = refs[-1]
    ref = curves[ref_name]
    summary = {"reference": ref_name}
    for name, pts in curves.items():
        if name.startswith("approx_min_star"):
            continue
        gaps = []
        for _, f in ref:
            if f <= 0:
                continue
            a, b = ebn0_at(ref, f), ebn0_at(pts, f)
            if a is not None and b is not None:
                gaps.append({"fer": f, "gap_db": b - a})
        summary[name] = gaps
    print(json.dumps({"summary": summary}), flush=True)


if __name__ == "__main__":
    main()
