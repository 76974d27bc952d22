#!/usr/bin/env python3
"""Record the reference golden's own Monte Carlo totals as a fixture.

The reference's golden layered decoder (tests/test_circulant_dma_stepper.cpp,
built from its own source by ``make -f oracle/Makefile.ref`` into
oracle/_ref/stepper) is run in its Monte Carlo mode with one host thread:

    stepper --sfpi --monte-carlo <ebn0> 1000 1000000000 200 1

With one thread the run is deterministic (mt19937_64 seeded 1337,
std::normal_distribution<float>, stepper:734-757) and decodes 250 batches of 4
codewords (stepper:739-740: the loop stops once 1000 blocks are done).  It
prints Total Blocks, Frame Errors, Bit Errors and Mean Iterations (3
decimals; with 1000 blocks this is the exact iteration sum / 1000).  This
script stores those four numbers per Eb/N0 in tests/golden/golden_mc.json.

tests/test_golden_mc.py rebuilds the same channel draws with
golden_mc_channel.cpp (the C++ standard library's generators, as the stepper
calls them), decodes them with the oracle and tallies them with the oracle's
Monte Carlo rule, and must reproduce all four numbers.

Run from the repo root:  python tests/golden/make_golden_mc.py
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
STEPPER = os.path.join(ROOT, "oracle", "_ref", "stepper")
POINTS = (1.0, 1.5)
BLOCKS = 1000
MAX_ITER = 200

FIELDS = {
    "blocks": re.compile(r"Total Blocks:\s+(\d+)"),
    "frame_errors": re.compile(r"Frame Errors:\s+(\d+)"),
    "bit_errors": re.compile(r"Bit Errors:\s+(\d+)"),
    "mean_iterations": re.compile(r"Mean Iterations:\s+([0-9.]+)"),
}


def run(ebn0):
    cmd = [STEPPER, "--sfpi", "--monte-carlo", f"{ebn0:.2f}", str(BLOCKS), "1000000000", str(MAX_ITER), "1"]
    p = subprocess.run(cmd, capture_output=True, text=True, check=True)
    rec = {"command": "oracle/_ref/stepper " + " ".join(cmd[1:]), "ebn0_db": ebn0, "max_iter": MAX_ITER,
           "beta": 1.0}
    for k, rx in FIELDS.items():
        m = rx.search(p.stdout)
        if not m:
            raise RuntimeError(f"no '{k}' in the stepper's output:\n{p.stdout[-2000:]}")
        rec[k] = m.group(1) if k == "mean_iterations" else int(m.group(1))
    # Mean Iterations is printed with 3 decimals: the exact sum over 1000 blocks
    rec["iter_sum"] = round(float(rec["mean_iterations"]) * rec["blocks"])
    return rec


def main():
    if not os.path.exists(STEPPER):
        sys.exit(f"{STEPPER} missing: build it with  make -f oracle/Makefile.ref")
    out = {"source": "reference golden stepper, --sfpi --monte-carlo, 1 host thread (stepper:708-856)",
           "points": [run(e) for e in POINTS]}
    path = os.path.join(HERE, "golden_mc.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
