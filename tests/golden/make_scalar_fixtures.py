#!/usr/bin/env python3
"""Write golden fixtures for the Profile R decoder from the reference's own
scalar_5w worker (engines/scalar_5w/ldpc_decoder_core.h:214-586).

oracle/_ref/scalar_driver(_nms) (built by ``make -f oracle/Makefile.ref``)
runs run_ldpc_worker on one pair of codewords and records the channel LLRs
the worker drew (xoshiro128+ and polar Marsaglia from the header itself), the
final LLRs and the worker's stats.  Cases cover Approx-Min* and normalized
min-sum, the three clamps, several Eb/N0 and iteration limits, including
limits at which the syndrome (checked from the third iteration on) is never
evaluated.

Output: tests/golden/scalar_profile_r.npz and, as a data file the decoder
needs on machines without the reference tree, a copy of the reference's
matrices/AR4JA_r45_4c_128c_r12.chinn.out.

Run from the repo root:  python tests/golden/make_scalar_fixtures.py
"""
import hashlib
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_MATRIX = "/root/reference/matrices/AR4JA_r45_4c_128c_r12.chinn.out"
MATRIX = os.path.join(HERE, "AR4JA_r45_4c_128c_r12.chinn.out")
DRV = os.path.join(ROOT, "oracle", "_ref", "scalar_driver")

# (nms, ebn0, max_iter, seed, l_max, r_max, l_post_max)
CASES = [
    (0, 2.0, 50, 1, 0, 0, 0), (0, 2.0, 50, 2, 0, 0, 0),
    (0, 1.5, 50, 3, 0, 0, 0), (0, 1.5, 50, 4, 0, 0, 0),
    (0, 1.0, 50, 5, 0, 0, 0), (0, 1.0, 200, 6, 0, 0, 0),
    (0, 2.5, 50, 7, 0, 0, 0), (0, 0.5, 30, 8, 0, 0, 0),
    (0, 1.5, 2, 9, 0, 0, 0), (0, 1.5, 3, 10, 0, 0, 0),
    (0, 1.5, 60, 11, 8.0, 6.0, 20.0), (0, 1.0, 60, 12, 4.0, 0, 0),
    (0, 1.0, 60, 13, 0, 1.5, 0), (0, 1.0, 60, 14, 0, 0, 6.0),
    (1, 2.0, 50, 15, 0, 0, 0), (1, 1.0, 100, 16, 0, 0, 0),
    (1, 1.5, 60, 17, 8.0, 6.0, 20.0), (1, 0.5, 30, 18, 0, 0, 0),
]


def main():
    shutil.copyfile(REF_MATRIX, MATRIX)
    rec = {k: [] for k in ("nms", "ebn0", "max_iter", "seed", "l_max", "r_max", "l_post_max", "llr_in", "llr_out",
                           "stats")}
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "o.bin")
        for nms, ebn0, mi, seed, lmax, rmax, lpost in CASES:
            exe = DRV + ("_nms" if nms else "")
            subprocess.run([exe, MATRIX, repr(ebn0), "512", str(mi), str(seed), repr(float(lmax)),
                            repr(float(rmax)), repr(float(lpost)), out], check=True)
            a = np.fromfile(out, dtype="<u4")
            N = int(a[0])
            inp, fin, st = a[4:4 + N], a[4 + N:4 + 2 * N], a[4 + 2 * N:4 + 2 * N + 4]
            unpack = lambda w: np.stack([(w & 0xFFFF), (w >> 16)]).astype(np.uint16)  # noqa: E731
            for k, v in zip(("nms", "ebn0", "max_iter", "seed", "l_max", "r_max", "l_post_max"),
                            (nms, ebn0, mi, seed, lmax, rmax, lpost)):
                rec[k].append(v)
            rec["llr_in"].append(unpack(inp))
            rec["llr_out"].append(unpack(fin))
            rec["stats"].append(st[:3].astype(np.int64))
            print(f"nms={nms} ebn0={ebn0} max_iter={mi:3d} clamps=({lmax},{rmax},{lpost}) "
                  f"BE={st[0]} FE={st[1]} cw0_iters={st[2]}")
    np.savez_compressed(os.path.join(HERE, "scalar_profile_r.npz"),
                        matrix_sha256=np.array(hashlib.sha256(open(MATRIX, "rb").read()).hexdigest()),
                        **{k: np.array(v) for k, v in rec.items()})


if __name__ == "__main__":
    main()
