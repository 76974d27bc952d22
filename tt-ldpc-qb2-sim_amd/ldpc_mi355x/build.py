"""Build libldpc.so (HIP kernels + C ABI runtime) in-tree for gfx950.

    python -m ldpc_mi355x.build      (from tt-ldpc-qb2-sim_amd/)

hipcc cross-compiles without a GPU.  All HIP code is built with
-ffp-contract=off (the decoder is bit-exact only with separately rounded
fp32 ops) and with packed-fp32 instruction selection disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
LIB = os.path.join(PKG, "libldpc.so")
ARCH = os.environ.get("LDPC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["decode_t.hip", "decode_r.hip", "runtime.cpp"]
# command-line shims speaking the reference engines' CLI, linked to libldpc
CLIS = {"ldpc_sim_scalar": "cli_scalar.cpp", "ldpc_sim_circulant": "cli_circulant.cpp"}
HEADERS = ["ar4ja_schedule.h", "device_math.h", "kernels.h"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libldpc")


def build(force: bool = True, verbose: bool = False) -> str:
    """Compile every HIP/C++ source into libldpc.so; returns its path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    deps = srcs + [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(os.path.dirname(ROOT), "include",
                                                                           "ldpc_mi355x.h")]
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(d) for d in deps):
        return LIB
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", *HIPCC_FLAGS, "-shared", "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)
    os.replace(tmp, LIB)
    build_clis(verbose)
    return LIB


def build_clis(verbose: bool = False) -> None:
    inc = os.path.join(os.path.dirname(ROOT), "include")
    for exe, src in CLIS.items():
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{inc}", os.path.join(CSRC, src), f"-L{PKG}", "-lldpc",
               "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,-rpath,$ORIGIN", "-o",
               os.path.join(PKG, exe)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    print(build(force=True, verbose="-v" in sys.argv))
