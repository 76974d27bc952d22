"""ctypes binding of libldpc.so (include/ldpc_mi355x.h).

The library is loaded from this package directory.  If it is missing, import
of the binding raises: there is no CPU fallback for the decoder.
"""
from __future__ import annotations

import ctypes as C
import os

from .build import LIB

LDPC_OK = 0
LDPC_ERR_ARG = -1
LDPC_ERR_UNSUPPORTED = -2
LDPC_ERR_HIP = -3
LDPC_ERR_NOMEM = -4
LDPC_ERR_NODEV = -5

LDPC_PROFILE_T = 0
LDPC_PROFILE_R = 1
LDPC_PROFILE_R_NMS = 2
LDPC_PROFILE_D = 3
LDPC_CODEWORD_ZERO = 0
LDPC_CODEWORD_RANDOM = 1
LDPC_CODE_AR4JA = 1
LDPC_CODE_MATRIX = 2


class CodeDesc(C.Structure):
    _fields_ = [("family", C.c_uint32), ("k_info", C.c_uint32), ("rate_num", C.c_uint32), ("rate_den", C.c_uint32),
                ("n", C.c_uint32), ("m", C.c_uint32), ("n_punct", C.c_uint32), ("reserved", C.c_uint32),
                ("row_ptr", C.c_void_p), ("col", C.c_void_p)]


class DecoderCfg(C.Structure):
    _fields_ = [("profile", C.c_uint32), ("max_iter", C.c_uint32), ("beta", C.c_float), ("l_max", C.c_float),
                ("r_max", C.c_float), ("l_post_max", C.c_float), ("gamma", C.c_float), ("variant", C.c_uint32)]


class CodeInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("n", "m", "k_info", "n_tx", "z", "n_layers", "n_edges", "max_iter")]


class PointCfg(C.Structure):
    _fields_ = [("ebn0_db", C.c_double), ("seed", C.c_uint64), ("first_cw", C.c_uint64), ("max_blocks", C.c_uint64),
                ("target_fe", C.c_uint64), ("chunk_cw", C.c_uint64)]


class PointResult(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("blocks", "frame_errors", "bit_errors", "iter_sum", "next_cw", "elapsed_ns",
                                          "kernel_ns", "launches")]


class TuneResult(C.Structure):
    _fields_ = [("variant", C.c_uint32), ("threads_per_wg", C.c_uint32), ("cw_per_wg", C.c_uint32),
                ("wg_per_cu", C.c_uint32), ("gsps", C.c_double), ("mean_iterations", C.c_double),
                ("frame_errors", C.c_uint64), ("kernel_ns", C.c_uint64)]


def variant(checks_per_thread: int, layer_order: int) -> int:
    """LDPC_VARIANT(checks_per_thread, layer_order) of the header."""
    return 0x100 | (checks_per_thread & 0xF) | ((layer_order & 0xF) << 4)


def variant_fields(v: int):
    """(checks_per_thread, layer_order) of a variant code."""
    return v & 0xF, (v >> 4) & 0xF


# name -> (restype, argtypes); every symbol include/ldpc_mi355x.h declares.
SIGNATURES = {
    "ldpc_create": (C.c_int, [C.POINTER(CodeDesc), C.POINTER(DecoderCfg), C.c_int, C.POINTER(C.c_void_p)]),
    "ldpc_destroy": (None, [C.c_void_p]),
    "ldpc_last_error": (C.c_char_p, [C.c_void_p]),
    "ldpc_get_code_info": (C.c_int, [C.c_void_p, C.POINTER(CodeInfo)]),
    "ldpc_get_schedule": (C.c_int, [C.POINTER(CodeDesc), C.c_uint32, C.c_void_p, C.c_int32]),
    "ldpc_list_variants": (C.c_int, [C.POINTER(CodeDesc), C.c_void_p, C.c_int32]),
    "ldpc_autotune": (C.c_int, [C.POINTER(CodeDesc), C.POINTER(DecoderCfg), C.c_int, C.c_double, C.c_uint64,
                                C.c_uint32, C.c_uint32, C.c_int, C.POINTER(TuneResult), C.c_int32,
                                C.POINTER(C.c_uint32)]),
    "ldpc_batch_create": (C.c_int, [C.POINTER(CodeDesc), C.c_void_p, C.c_uint32, C.POINTER(DecoderCfg), C.c_int,
                                    C.POINTER(C.c_void_p)]),
    "ldpc_batch_run": (C.c_int, [C.c_void_p, C.POINTER(PointCfg), C.POINTER(PointResult), C.c_void_p]),
    "ldpc_batch_code_info": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(CodeInfo)]),
    "ldpc_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "ldpc_batch_destroy": (None, [C.c_void_p]),
    "ldpc_row_levels": (C.c_int, [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ldpc_row_passes": (C.c_int, [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "ldpc_row_passes_by_degree": (C.c_int, [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                            C.c_void_p, C.c_int32]),
    "ldpc_run_point": (C.c_int, [C.c_void_p, C.POINTER(PointCfg), C.POINTER(PointResult), C.c_void_p]),
    "ldpc_decode_llrs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ldpc_decode_llrs_device_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                                C.c_void_p]),
    "ldpc_decode_llrs_device_noalloc": (C.c_int, [C.POINTER(CodeDesc), C.POINTER(DecoderCfg), C.c_int, C.c_void_p,
                                                  C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ldpc_get_generator": (C.c_int, [C.POINTER(CodeDesc), C.c_void_p, C.c_int32]),
    "ldpc_encode": (C.c_int, [C.POINTER(CodeDesc), C.c_void_p, C.c_uint32, C.c_void_p]),
    "ldpc_gen_channel": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, C.c_void_p]),
    "ldpc_channel_params": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "ldpc_set_capture": (C.c_int, [C.c_void_p, C.c_uint32]),
    "ldpc_get_captures": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64)]),
    "ldpc_set_dmin_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int,
                                       C.c_float]),
    "ldpc_set_codeword_mode": (C.c_int, [C.c_void_p, C.c_uint32]),
    "ldpc_noise_tails": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint32,
                                   C.c_void_p]),
    "ldpc_sync": (C.c_int, [C.c_void_p]),
    "ldpc_stream": (C.c_void_p, [C.c_void_p]),
}

_lib = None


def load():
    """Load libldpc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} is missing: build it with `python -m ldpc_mi355x.build` "
                              "(the HIP extension is required; there is no CPU fallback)")
        lib = C.CDLL(LIB)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class LdpcError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"libldpc error {status}: {msg}")
        self.status = status
