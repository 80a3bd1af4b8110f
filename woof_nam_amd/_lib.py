"""ctypes binding of libwoofrad.so (the gfx950 HIP library declared in include/woofrad.h).

The library is built in-tree by ``python -m woof_nam_amd.build``.  There is no CPU fallback:
loading fails with an ImportError naming the missing file.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from . import abi

LIB_PATH = Path(os.environ.get("WR_LIB", Path(__file__).resolve().parent / "libwoofrad.so"))
TABLES = Path(__file__).resolve().parent / "data" / "rrtmg_tables.bin"

# symbols include/woofrad.h declares (the CPU test suite checks every one is exported)
EXPORTS = ("wr_default_params", "wr_create", "wr_context_bytes", "wr_run_fp64", "wr_run_mixed", "wr_run_mixed_noalloc", "wr_run_band", "wr_cal_cldfra",
           "wr_mcica_masks", "wr_timing", "wr_timing_read", "wr_kernel_name", "wr_strerror", "wr_destroy",
           "wr_layer_optics", "wr_mlp_gemm", "wr_mlp_splitk_reduce", "wr_mlp_colsum", "wr_mlp_mse", "wr_mlp_adam",
           "wr_mlp_to_bf16", "wr_surrogate_attach", "wr_run_surrogate", "wr_nmmb_prep", "wr_nmmb_post", "wr_nmmb_albedo",
           "wr_sol_update", "wr_coszmn", "wr_rdtemp", "wr_aer_optics", "wr_mlp_forward",
           "wr_slot_create", "wr_slot_launch", "wr_slot_query", "wr_slot_sync", "wr_slot_destroy")

_ERRORS = {
    abi.WR_ERR_ARG: ValueError,
    abi.WR_ERR_TABLES: FileNotFoundError,
    abi.WR_ERR_HIP: RuntimeError,
    abi.WR_ERR_NOMEM: MemoryError,
    abi.WR_ERR_UNSUPPORTED: NotImplementedError,
    abi.WR_ERR_TOO_LARGE: ValueError,
}

_lib = None


def load():
    """Load libwoofrad.so once and declare the argument types of every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built (python -m woof_nam_amd.build); "
                          "the radiation path has no CPU fallback")
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    L.wr_default_params.argtypes = [C.POINTER(abi.WrParams)]
    L.wr_default_params.restype = None
    L.wr_create.argtypes = [C.POINTER(vp), C.c_char_p, C.c_int, C.c_int32]
    L.wr_create.restype = C.c_int
    L.wr_context_bytes.argtypes = [C.c_int32]
    L.wr_context_bytes.restype = C.c_uint64
    for f in ("wr_run_fp64", "wr_run_mixed"):
        getattr(L, f).argtypes = [vp, C.POINTER(abi.WrColumns), C.POINTER(abi.WrParams),
                                  C.POINTER(abi.WrOutputs), C.POINTER(abi.WrDiag), vp]
        getattr(L, f).restype = C.c_int
    L.wr_run_mixed_noalloc.argtypes = [vp, C.POINTER(abi.WrColumns), C.POINTER(abi.WrParams),
                                       C.POINTER(abi.WrOutputs), vp]
    L.wr_run_mixed_noalloc.restype = C.c_int
    L.wr_run_band.argtypes = [vp, C.POINTER(abi.WrColumns), C.POINTER(abi.WrParams), C.POINTER(abi.WrOutputs),
                              C.c_int32, C.c_int32, vp]
    L.wr_run_band.restype = C.c_int
    L.wr_cal_cldfra.argtypes = [C.c_int32] * 9 + [vp, vp, vp, vp]
    L.wr_cal_cldfra.restype = C.c_int
    L.wr_mcica_masks.argtypes = [C.c_int32, C.c_int32, vp, vp, C.c_uint64, C.c_int32, vp, vp, vp]
    L.wr_mcica_masks.restype = C.c_int
    L.wr_timing.argtypes = [vp, C.c_int32]
    L.wr_timing.restype = C.c_int
    L.wr_timing_read.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_int32, C.POINTER(C.c_int32)]
    L.wr_timing_read.restype = C.c_int
    L.wr_kernel_name.argtypes = [C.c_int32]
    L.wr_kernel_name.restype = C.c_char_p
    i32, i64, f32 = C.c_int32, C.c_int64, C.c_float
    L.wr_layer_optics.argtypes = [vp, C.POINTER(abi.WrColumns), C.POINTER(abi.WrParams), vp, vp, vp]
    L.wr_mlp_gemm.argtypes = [i32, i32, i32, vp, i32, i64, i64, vp, i32, i64, i64, vp, vp, i64, i64, vp, vp, vp,
                              i32, i32, i64, vp]
    L.wr_mlp_splitk_reduce.argtypes = [vp, i32, i64, vp, i64, vp]
    L.wr_mlp_colsum.argtypes = [vp, i32, i32, i64, vp, vp]
    L.wr_mlp_mse.argtypes = [vp, vp, i64, f32, vp, vp, i32, vp]
    L.wr_mlp_adam.argtypes = [vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, vp]
    L.wr_mlp_to_bf16.argtypes = [vp, vp, i64, vp]
    for f in ("wr_layer_optics", "wr_mlp_gemm", "wr_mlp_splitk_reduce", "wr_mlp_colsum", "wr_mlp_mse", "wr_mlp_adam",
              "wr_mlp_to_bf16", "wr_surrogate_attach", "wr_run_surrogate"):
        getattr(L, f).restype = C.c_int
    L.wr_surrogate_attach.argtypes = [vp, C.POINTER(abi.WrSurrogate)]
    L.wr_surrogate_attach.restype = C.c_int
    L.wr_run_surrogate.argtypes = [vp, C.POINTER(abi.WrColumns), C.POINTER(abi.WrParams), C.POINTER(abi.WrOutputs),
                                   vp, vp]
    L.wr_run_surrogate.restype = C.c_int
    L.wr_nmmb_prep.argtypes = [C.POINTER(abi.WrNmmbState), C.POINTER(abi.WrNmmbOpts), C.POINTER(abi.WrNmmbCols), vp]
    L.wr_nmmb_prep.restype = C.c_int
    L.wr_nmmb_post.argtypes = [C.POINTER(abi.WrNmmbState), C.POINTER(abi.WrNmmbCols), C.POINTER(abi.WrOutputs), vp,
                               C.c_int32, C.c_int32, C.POINTER(abi.WrNmmbOut), vp]
    L.wr_nmmb_post.restype = C.c_int
    L.wr_nmmb_albedo.argtypes = [C.POINTER(abi.WrNmmbState), vp, vp, vp, vp, vp]
    L.wr_nmmb_albedo.restype = C.c_int
    L.wr_aer_optics.argtypes = [C.POINTER(abi.WrAerTables), C.POINTER(abi.WrColumns), vp, vp, vp, vp, vp, C.c_int64,
                                vp]
    L.wr_aer_optics.restype = C.c_int
    L.wr_mlp_forward.argtypes = [C.POINTER(abi.WrSurrogate), C.c_int64, vp, vp, C.c_int64, vp]
    L.wr_mlp_forward.restype = C.c_int
    L.wr_slot_create.argtypes = [C.POINTER(vp), vp, C.c_int32, C.POINTER(abi.WrParams), C.c_int32,
                                 C.POINTER(abi.WrColumns)]
    L.wr_slot_create.restype = C.c_int
    L.wr_slot_launch.argtypes = [vp, C.POINTER(abi.WrColumns), C.POINTER(abi.WrOutputs), C.c_int32, C.c_int32,
                                 C.c_uint64]
    L.wr_slot_launch.restype = C.c_int
    for f in ("wr_slot_query", "wr_slot_sync"):
        getattr(L, f).argtypes = [vp]
        getattr(L, f).restype = C.c_int
    L.wr_slot_destroy.argtypes = [vp]
    L.wr_slot_destroy.restype = None
    i32, f64 = C.c_int32, C.c_double
    L.wr_sol_update.argtypes = [i32] * 6 + [f64, f64, f64, C.POINTER(abi.WrSol)]
    L.wr_sol_update.restype = C.c_int
    L.wr_coszmn.argtypes = [i32, vp, vp, f64, f64, i32, C.POINTER(abi.WrSol), vp, vp, vp]
    L.wr_coszmn.restype = C.c_int
    L.wr_rdtemp.argtypes = [i32] * 8 + [f64] + [i32] * 4 + [vp] * 7 + [vp]
    L.wr_rdtemp.restype = C.c_int
    L.wr_strerror.argtypes = [C.c_int]
    L.wr_strerror.restype = C.c_char_p
    L.wr_destroy.argtypes = [vp]
    L.wr_destroy.restype = None
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    """Raise the named exception a status code stands for."""
    if rc == abi.WR_OK:
        return
    msg = load().wr_strerror(rc).decode()
    raise _ERRORS.get(rc, RuntimeError)(f"{what}: {msg} (code {rc})")
