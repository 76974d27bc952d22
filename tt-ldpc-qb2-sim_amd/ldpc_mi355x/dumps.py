"""Packed4CW dumps and trapping-set capture.

The reference's circulant engine snapshots the variable nodes of four
codewords (one core's batch) into ``tensix_initial_vnodes.bin`` and
``tensix_final_vnodes.bin`` (engines/circulant_dma_4cw/ldpc_sim_circulant_dma.
cpp:435-472), and its golden stepper replays such a dump with ``--load-init``
(tests/test_circulant_dma_stepper.cpp:861-990).  Packed4CW layout
(stepper:879-880): N entries of 8 bytes, variable-major, each four
little-endian bf16 values for codewords 0..3.

Here a trap is a codeword id (Engine.set_capture / Engine.captures): the
channel of any id is regenerated bit for bit from (seed, id) and its decode
is deterministic, so both dumps of the id's group of four are rebuilt on
demand.  ``trapping_set`` gives the (a, b) signature of a failure -- the
variables in error and the unsatisfied checks of its final hard decision --
and ``python -m ldpc_mi355x.dumps`` turns captured ids into dumps plus that
report.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
from typing import Dict, List, Optional

import numpy as np

INITIAL = "tensix_initial_vnodes.bin"
FINAL = "tensix_final_vnodes.bin"


def to_packed4(llr4: np.ndarray) -> bytes:
    """[4, N] bf16 bit patterns -> Packed4CW bytes."""
    llr4 = np.asarray(llr4, dtype=np.uint16)
    if llr4.ndim != 2 or llr4.shape[0] != 4:
        raise ValueError("Packed4CW holds exactly four codewords: shape [4, N]")
    return np.ascontiguousarray(llr4.T).astype("<u2").tobytes()


def from_packed4(data: bytes, n: int = 2560) -> np.ndarray:
    """Packed4CW bytes -> [4, N] bf16 bit patterns."""
    if len(data) != 8 * n:
        raise ValueError(f"Packed4CW of {n} variables is {8 * n} bytes, got {len(data)}")
    return np.frombuffer(data, dtype="<u2").reshape(n, 4).T.copy()


def write_packed4(path: str, llr4: np.ndarray) -> None:
    with open(path, "wb") as f:
        f.write(to_packed4(llr4))


def read_packed4(path: str, n: int = 2560) -> np.ndarray:
    with open(path, "rb") as f:
        return from_packed4(f.read(), n)


def replay(engine, llr4: np.ndarray) -> Dict:
    """Decode a four-codeword dump (the stepper's --load-init): final state,
    iterations and convergence per codeword."""
    out, iters, ok = engine.decode_llrs(np.asarray(llr4, np.uint16))
    be = ((out[:, :engine.k_info] & 0x8000) != 0) | (out[:, :engine.k_info] == 0)
    return {"final": out, "iterations": iters, "converged": ok, "bit_errors": be.sum(axis=1)}


def trap_dumps(engine, seed: int, ebn0_db: float, cw_id: int, out_dir: str = ".") -> Dict:
    """Rebuild the group of four codewords holding `cw_id` (ids 4g .. 4g+3):
    write its initial and final Packed4CW dumps into out_dir and return the
    trap report."""
    g = int(cw_id) // 4
    init = engine.gen_channel(seed, 4 * g, 4, ebn0_db)
    res = replay(engine, init)
    os.makedirs(out_dir, exist_ok=True)
    write_packed4(os.path.join(out_dir, INITIAL), init)
    write_packed4(os.path.join(out_dir, FINAL), res["final"])
    mask = sum(1 << i for i in range(4) if not res["converged"][i])
    return {"group": g, "ids": [4 * g + i for i in range(4)], "failing_cw": int(cw_id) - 4 * g,
            "active_mask": mask, "iterations": res["iterations"].tolist(),
            "bit_errors": res["bit_errors"].tolist(), "initial": init, "final": res["final"]}


def trapping_set(code, final: np.ndarray) -> Dict:
    """(a, b) signature of a decoder failure on the all-zero codeword: the
    variables whose final hard decision is 1 (bf16 < 0, the golden's hard
    decision, stepper:781) and the checks of H left unsatisfied by that hard
    decision.  `final` holds one codeword's N bf16 bit patterns."""
    final = np.asarray(final, dtype=np.uint16)
    if final.shape != (code.n,):
        raise ValueError(f"expected {code.n} bf16 values, got shape {final.shape}")
    hard = ((final & 0x8000) != 0) & ((final & 0x7FFF) != 0)
    err = np.flatnonzero(hard)
    unsat = [i for i, r in enumerate(code.rows) if int(hard[r].sum()) & 1]
    return {"a": int(err.size), "b": len(unsat), "variables": err.tolist(), "checks": unsat,
            "info_bits": int(hard[:code.k].sum())}


def main(argv: Optional[List[str]] = None) -> int:
    ap = argparse.ArgumentParser(prog="python -m ldpc_mi355x.dumps",
                                 description="Rebuild captured failures: Packed4CW dumps and trapping-set signatures")
    ap.add_argument("--seed", type=lambda x: int(x, 0), required=True, help="the Monte Carlo run's seed")
    ap.add_argument("--ebn0", type=float, required=True)
    ap.add_argument("--id", type=int, action="append", required=True, help="failing codeword id (repeatable)")
    ap.add_argument("--k", type=int, default=1024)
    ap.add_argument("--rate", default="1/2")
    ap.add_argument("--max-iter", type=int, default=200)
    ap.add_argument("--beta", type=float, default=0.9375)
    ap.add_argument("--out", default="traps", help="one sub-folder per id")
    args = ap.parse_args(argv)
    from . import codes, decoder
    code = codes.ar4ja(args.k, args.rate)
    with decoder.Engine(args.k, args.rate, max_iter=args.max_iter, beta=args.beta) as eng:
        for cw in args.id:
            d = os.path.join(args.out, f"id_{cw}")
            rep = trap_dumps(eng, args.seed, args.ebn0, cw, d)
            i = rep["failing_cw"]
            ts = trapping_set(code, rep["final"][i])
            print(json.dumps({"id": cw, "group": rep["group"], "dir": d, "iterations": rep["iterations"][i],
                              "converged": ((rep["active_mask"] >> i) & 1) == 0,
                              "bit_errors": rep["bit_errors"][i], "a": ts["a"], "b": ts["b"],
                              "info_bits_in_error": ts["info_bits"], "variables": ts["variables"][:64],
                              "checks": ts["checks"][:64]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
