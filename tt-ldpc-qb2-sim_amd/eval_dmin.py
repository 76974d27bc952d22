#!/usr/bin/env python3
"""Judge a trained layered min-sum decoder by plain Monte Carlo on one GPU.

Runs libldpc Profile D (the trained parameters, syndrome early exit) at each
Eb/N0 for the given iteration caps, and Profile T (the golden Approx-Min*)
at 200 iterations on the same codeword ids and seed -- the same channel
realisations.  Prints one JSON line per (decoder, Eb/N0) with FER, its 95%
interval, mean iterations and Gsps, then one summary line with the Eb/N0
gap at equal FER (log-linear interpolation of each FER curve) and the
throughput ratio.
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
    args = ap.parse_args(argv)

    z = np.load(args.params, allow_pickle=False)
    k, rate = int(z["k_info"]), str(z["rate"])
    curves = {}
    runs = [("approx_min_star", 200, None)] + [(f"trained_T{t}", t, t) for t in args.iters]
    for name, max_iter, _ in runs:
        if name == "approx_min_star":
            eng = decoder.Engine(k, rate, max_iter=200)
        else:
            eng = decoder.Engine(k, rate, max_iter=max_iter, profile="D")
            eng.set_dmin_params(z["alpha"], z["beta"], bool(z["correction"]), float(z["r_max"]))
        pts = []
        for eb in args.ebn0:
            r = eng.run_point(eb, args.blocks, target_fe=args.target_fe, seed=args.seed, chunk_cw=1 << 20)
            lo, hi = poisson_ci95(r.frame_errors)
            pts.append((eb, r.fer))
            print(json.dumps({"decoder": name, "max_iter": max_iter, "ebn0_db": eb, "blocks": r.blocks,
                              "frame_errors": r.frame_errors, "fer": r.fer,
                              "fer_ci95": [lo / r.blocks, hi / r.blocks], "ber": r.ber,
                              "mean_iterations": r.mean_iterations, "kernel_gsps": r.kernel_gsps}), flush=True)
        curves[name] = pts
        eng.close()
    ref = curves["approx_min_star"]
    summary = {}
    for name, pts in curves.items():
        if name == "approx_min_star":
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
