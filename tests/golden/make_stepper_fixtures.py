#!/usr/bin/env python3
"""Write golden fixtures for the Profile T decoder from the reference stepper.

The reference's golden layered decoder (tests/test_circulant_dma_stepper.cpp,
built from its own source by ``make -f oracle/Makefile.ref`` into
oracle/_ref/stepper) is run in ``--sfpi --load-init <dump> <max_iter>
<damping>`` mode (stepper:861-990, 1009-1018) on Packed4CW dumps of
channel LLRs.  For every 4-codeword dump we record:

  * the input LLRs (bf16, codeword-major),
  * the final LLRs it writes to stepper_final_vnodes.bin,
  * the per-iteration unsatisfied-check counts it prints ("Unsat="), and
  * the iteration of convergence ("CONV(k)", 0 = never converged).

Inputs are channel LLRs from the oracle's Philox/Box-Muller stream at several
Eb/N0 plus hand-made edge cases (zeros, -0, tiny and large magnitudes,
random bit patterns).  Output: tests/golden/stepper_profile_t.npz.

Run from the repo root:  python tests/golden/make_stepper_fixtures.py
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "tt-ldpc-qb2-sim_amd"), os.path.join(ROOT, "oracle")]
import oracle as O  # noqa: E402
from ldpc_mi355x import codes  # noqa: E402

ITER_RE = re.compile(r"^\[Iter\s+(\d+)\]")
CW_RE = re.compile(r"CW(\d): (?:Unsat=(\d+), BE=(\d+)|CONV\((\d+)\))")


def run_stepper(llr4, max_iter, damping, n):
    with tempfile.TemporaryDirectory() as d:
        dump = os.path.join(d, "init.bin")
        with open(dump, "wb") as f:
            f.write(O.to_packed4(llr4))
        p = subprocess.run([O.REF_STEPPER, "--sfpi", "--load-init", dump, str(max_iter), repr(float(damping))],
                           cwd=d, capture_output=True, text=True)
        with open(os.path.join(d, "stepper_final_vnodes.bin"), "rb") as f:
            final = O.from_packed4(f.read(), n)
    unsat = np.zeros((4, max_iter), dtype=np.int32) - 1
    conv = np.zeros(4, dtype=np.int32)
    for line in p.stdout.splitlines():
        m = ITER_RE.match(line)
        if not m:
            continue
        it = int(m.group(1))
        for cw, u, _be, c in CW_RE.findall(line):
            cw = int(cw)
            if u != "":
                unsat[cw, it - 1] = int(u)
                if int(u) == 0:
                    conv[cw] = it
    return final, unsat, conv


def edge_case_llrs(rng, n, n_tx):
    """Four codewords of bit patterns that stress rounding and signs."""
    out = np.zeros((4, n), dtype=np.uint16)
    # 0: random finite bf16 patterns (exponent 0x60..0x9F), including -0/+0
    e = rng.integers(0x60, 0xA0, size=n)
    m = rng.integers(0, 128, size=n)
    s = rng.integers(0, 2, size=n)
    out[0] = (s << 15) | (e << 7) | m
    out[0, rng.integers(0, n, 40)] = 0x8000
    out[0, rng.integers(0, n, 40)] = 0x0000
    # 1: weak channel with many exact zeros and -0 (erasure-like)
    g = rng.normal(0.6, 1.6, size=n).astype(np.float32)
    out[1] = (g.view(np.uint32) + 0x7FFF + ((g.view(np.uint32) >> 16) & 1) >> 16).astype(np.uint16)
    z = rng.random(n) < 0.2
    out[1, z] = np.where(rng.random(z.sum()) < 0.5, 0x0000, 0x8000)
    # 2: large magnitudes with a few strong wrong signs
    g = rng.normal(40.0, 30.0, size=n).astype(np.float32)
    out[2] = (g.view(np.uint32) + 0x7FFF + ((g.view(np.uint32) >> 16) & 1) >> 16).astype(np.uint16)
    # 3: tiny magnitudes (1e-4 scale), correct-sign bias
    g = (rng.normal(1.0, 1.4, size=n) * 1e-4).astype(np.float32)
    out[3] = (g.view(np.uint32) + 0x7FFF + ((g.view(np.uint32) >> 16) & 1) >> 16).astype(np.uint16)
    out[:, n_tx:] = 0
    out[0, n_tx:] = np.where(rng.random(n - n_tx) < 0.5, 0, 0x8000)  # punctured as +-0 in case 0
    return out


def main():
    code = codes.ar4ja(1024, "1/2")
    n = code.n
    rng = np.random.default_rng(20260)
    cases = []  # (label, llr[4, n], max_iter, damping)
    seed = 0x5EED2026
    first = 0
    for ebn0, n_dumps, max_iter in ((2.0, 3, 200), (1.5, 3, 200), (1.0, 3, 200), (2.5, 1, 200), (0.5, 1, 40)):
        g = 10.0 ** (ebn0 / 10.0)
        mu, sg = np.float32(4 * 0.5 * g), np.float32(np.sqrt(8 * 0.5 * g))
        for _ in range(n_dumps):
            llr = O.gen_channel(code, seed, first, 4, mu, sg)
            first += 4
            cases.append((f"awgn_{ebn0:.1f}dB", llr, max_iter, 0.9375))
    for ebn0 in (2.0, 1.0):
        g = 10.0 ** (ebn0 / 10.0)
        mu, sg = np.float32(4 * 0.5 * g), np.float32(np.sqrt(8 * 0.5 * g))
        llr = O.gen_channel(code, seed, first, 4, mu, sg)
        first += 4
        cases.append((f"awgn_{ebn0:.1f}dB_beta1", llr, 200, 1.0))
    cases.append(("edge_cases", edge_case_llrs(rng, n, code.n_tx), 30, 0.9375))
    cases.append(("edge_cases_beta1", edge_case_llrs(rng, n, code.n_tx), 30, 1.0))

    labels, inp, outp, convs, unsats, iters, damps = [], [], [], [], [], [], []
    max_len = max(c[2] for c in cases)
    for label, llr, max_iter, damping in cases:
        final, unsat, conv = run_stepper(llr, max_iter, damping, n)
        pad = np.full((4, max_len), -1, dtype=np.int32)
        pad[:, :max_iter] = unsat
        labels += [label] * 4
        inp.append(llr)
        outp.append(final)
        convs.append(conv)
        unsats.append(pad)
        iters += [max_iter] * 4
        damps += [damping] * 4
        print(f"{label:22s} max_iter={max_iter:3d} beta={damping} conv={conv.tolist()}")
    np.savez_compressed(
        os.path.join(os.path.dirname(os.path.abspath(__file__)), "stepper_profile_t.npz"),
        label=np.array(labels), llr_in=np.concatenate(inp), llr_out=np.concatenate(outp),
        conv_iter=np.concatenate(convs), unsat=np.concatenate(unsats), max_iter=np.array(iters, dtype=np.int32),
        beta=np.array(damps, dtype=np.float32), code_digest=np.array(code.digest()))


if __name__ == "__main__":
    main()
