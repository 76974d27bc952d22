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
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
LIB = os.path.join(PKG, "libldpc.so")
ARCH = os.environ.get("LDPC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["decode_t.hip", "decode_r.hip", "decode_r2.hip", "runtime.cpp", "batch.cpp", "noalloc.hip", "encoder.cpp"]
# decode_multi.hip is compiled once per AR4JA code (K id, rate id)
MULTI = "decode_multi.hip"
MULTI_CODES = [(k, r) for k in range(3) for r in range(3)]
# command-line shims speaking the reference engines' CLI, linked to libldpc
CLIS = {"ldpc_sim_scalar": "cli_scalar.cpp", "ldpc_sim_circulant": "cli_circulant.cpp"}
HEADERS = ["ar4ja_schedule.h", "device_math.h", "kernels.h", "profile_r.h", "runtime_internal.h"]
OBJ_DIR = os.path.join(PKG, "_build")
# PyTorch extension (differentiable decoder): its own module, loaded into the
# torch process so that its kernels run on torch's HIP runtime and streams
TORCH_EXT = "ldpc_dmin"
TORCH_EXT_SRC = "dmin_torch.hip"
TORCH_EXT_LIB = os.path.join(PKG, TORCH_EXT + ".so")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
]


# Decoder kernels: their device code assumes no NaN (inputs are finite LLRs;
# no operation makes a NaN), so fminf / fmaxf need no canonicalising of their
# operands, which otherwise costs an instruction per leaf and registers.
# Only device code: host code validates user floats with isfinite.
KERNEL_SOURCES = ("decode_t.hip", "decode_r.hip", "decode_r2.hip", MULTI)
KERNEL_FLAGS = ["-Xarch_device", "-fno-honor-nans"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libldpc")


def _compile_jobs():
    jobs = [(os.path.join(CSRC, s), os.path.join(OBJ_DIR, s + ".o"), KERNEL_FLAGS if s in KERNEL_SOURCES else [])
            for s in SOURCES]
    for k, r in MULTI_CODES:
        jobs.append((os.path.join(CSRC, MULTI), os.path.join(OBJ_DIR, f"multi_k{k}_r{r}.o"),
                     [*KERNEL_FLAGS, f"-DLDPC_KID={k}", f"-DLDPC_RID={r}"]))
    return jobs


def build(force: bool = True, verbose: bool = False, jobs: int = 0) -> str:
    """Compile every HIP/C++ source (in parallel) and link libldpc.so; returns its path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES + [MULTI]]
    deps = srcs + [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(os.path.dirname(ROOT), "include",
                                                                           "ldpc_mi355x.h")]
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(d) for d in deps):
        return LIB
    os.makedirs(OBJ_DIR, exist_ok=True)
    cc = hipcc()

    def compile_one(job):
        src, obj, defs = job
        cmd = [cc, f"--offload-arch={ARCH}", *HIPCC_FLAGS, *defs, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, cwd=CSRC)
        return obj

    workers = jobs or min(16, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(compile_one, _compile_jobs()))
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-fPIC", "-shared", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)
    os.replace(tmp, LIB)
    build_clis(verbose)
    return LIB


def build_torch_ext(verbose: bool = False) -> str:
    """Compile csrc/dmin_torch.hip against the installed PyTorch (plain hipcc,
    torch's include and library paths) into ldpc_mi355x/ldpc_dmin.so."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce
    inc = ce.include_paths(device_type="cuda") + [sysconfig.get_paths()["include"]]
    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    tmp = TORCH_EXT_LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", *HIPCC_FLAGS, "-shared", "-Wno-unused-result",
           "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", "-DHIPBLAS_V2", "-D__HIP_NO_HALF_OPERATORS__=1",
           "-D__HIP_NO_HALF_CONVERSIONS__=1", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           f"-DTORCH_EXTENSION_NAME={TORCH_EXT}", "-DTORCH_API_INCLUDE_EXTENSION_H",
           *[f"-I{d}" for d in inc], os.path.join(CSRC, TORCH_EXT_SRC),
           f"-L{torch_lib}", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python",
           f"-Wl,-rpath,{torch_lib}", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)
    os.replace(tmp, TORCH_EXT_LIB)
    return TORCH_EXT_LIB


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
    print(build_torch_ext(verbose="-v" in sys.argv))
