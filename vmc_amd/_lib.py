"""ctypes binding of ``libvmc_amd.so`` (declarations in include/vmc_amd.h).

The product path has no CPU fallback: if the library is missing this module
raises at import time.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvmc_amd.so")

VMC_F32 = 0
VMC_F64 = 1
VMC_C64 = 2
VMC_C128 = 3
VMC_NO_TRUNCATION = 0
VMC_ZIPUP = 1
VMC_DENSITY_MATRIX = 2
VMC_VARIATIONAL = 3
VMC_NOALLOC_SCRATCH_BYTES = 64 << 20  # include/vmc_amd.h
VMC_SOLVE_CHOLESKY = 0
VMC_SOLVE_CG = 1
VMC_SOLVE_SVD = 2
VMC_TUNE_TRUNC = 1
VMC_TUNE_THETA = 2
VMC_TRUNC_LPE8 = 0
VMC_TRUNC_LPE16 = 1
VMC_TRUNC_GENERIC = 2
VMC_TILE_AUTO = 0
VMC_TILE_WIDE = 1
VMC_TILE_TALL = 2
VMC_TILE_SQUARE = 3
VMC_TUNE_ENTRY_INTS = 13

EXPORTED = (
    "vmc_plan_create",
    "vmc_plan_destroy",
    "vmc_plan_num_params",
    "vmc_plan_num_sliced",
    "vmc_amp_grad_workspace_size",
    "vmc_amp_grad",
    "vmc_amp_grad_noalloc",
    "vmc_plan_set_terms",
    "vmc_plan_set_terms_c",
    "vmc_cache_bytes",
    "vmc_sweep_workspace_size",
    "vmc_init_cache",
    "vmc_mc_sweep",
    "vmc_transition",
    "vmc_local_estimate",
    "vmc_estimate",
    "vmc_sr_workspace_size",
    "vmc_sr_solve",
    "vmc_sr_solve_ex",
    "vmc_qgt_dense",
    "vmc_sr_qr_workspace_size",
    "vmc_sr_qr_solve",
    "vmc_gi_plan_create",
    "vmc_gi_plan_set_degeneracy",
    "vmc_gi_set_terms",
    "vmc_gi_slices",
    "vmc_gi_workspace_size",
    "vmc_gi_amp_grad",
    "vmc_gi_init_cache",
    "vmc_gi_mc_sweep",
    "vmc_bl_plan_create",
    "vmc_bl_slices",
    "vmc_bl_workspace_size",
    "vmc_bl_amp_grad",
    "vmc_bl_init_cache",
    "vmc_bl_mc_sweep",
    "vmc_plan_autotune",
    "vmc_plan_tuning",
    "vmc_plan_set_tuning",
    "vmc_profile_begin",
    "vmc_profile_end",
    "vmc_last_error",
)


class VMCError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise VMCError(
            f"{LIB_PATH} is missing: build it with `python -m vmc_amd.build` "
            "(there is no CPU fallback)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    c_int, c_i64, c_sz, c_vp = ctypes.c_int, ctypes.c_int64, ctypes.c_size_t, ctypes.c_void_p
    lib.vmc_plan_create.argtypes = [ctypes.POINTER(c_vp), c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int]
    lib.vmc_plan_create.restype = c_int
    lib.vmc_plan_destroy.argtypes = [c_vp]
    lib.vmc_plan_destroy.restype = c_int
    lib.vmc_plan_num_params.argtypes = [c_vp]
    lib.vmc_plan_num_params.restype = c_i64
    lib.vmc_plan_num_sliced.argtypes = [c_vp]
    lib.vmc_plan_num_sliced.restype = c_i64
    lib.vmc_amp_grad_workspace_size.argtypes = [c_vp, c_int]
    lib.vmc_amp_grad_workspace_size.restype = c_sz
    lib.vmc_amp_grad.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_amp_grad.restype = c_int
    lib.vmc_amp_grad_noalloc.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp]
    lib.vmc_amp_grad_noalloc.restype = c_int
    lib.vmc_plan_set_terms.argtypes = [c_vp, c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_int, c_vp, c_vp,
                                       c_int, c_vp, c_vp]
    lib.vmc_plan_set_terms.restype = c_int
    lib.vmc_plan_set_terms_c.argtypes = lib.vmc_plan_set_terms.argtypes
    lib.vmc_plan_set_terms_c.restype = c_int
    lib.vmc_cache_bytes.argtypes = [c_vp, c_int]
    lib.vmc_cache_bytes.restype = c_sz
    lib.vmc_sweep_workspace_size.argtypes = [c_vp, c_int]
    lib.vmc_sweep_workspace_size.restype = c_sz
    lib.vmc_init_cache.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_init_cache.restype = c_int
    lib.vmc_mc_sweep.argtypes = [c_vp, c_vp, c_vp, c_int, ctypes.c_uint64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp,
                                 c_vp, c_sz, c_vp]
    lib.vmc_mc_sweep.restype = c_int
    lib.vmc_transition.argtypes = [c_vp, c_vp, c_vp, c_int, ctypes.c_uint64, c_int, c_vp, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_transition.restype = c_int
    lib.vmc_estimate.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_estimate.restype = c_int
    lib.vmc_local_estimate.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_local_estimate.restype = c_int
    lib.vmc_sr_workspace_size.argtypes = [c_int, c_i64, c_int, c_int, c_int]
    lib.vmc_sr_workspace_size.restype = c_sz
    lib.vmc_sr_solve.argtypes = [c_int, c_i64, c_int, c_vp, c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_int, c_int,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, c_int, c_vp, c_vp, c_vp, c_sz,
                                 c_vp]
    lib.vmc_sr_solve.restype = c_int
    lib.vmc_sr_solve_ex.argtypes = [c_int, c_i64, c_int, c_vp, c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_int, c_int,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, c_int, c_int,
                                    ctypes.c_double, c_int, c_vp, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_sr_solve_ex.restype = c_int
    lib.vmc_qgt_dense.argtypes = [c_int, c_i64, c_int, c_vp, c_int, c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_vp,
                                  c_sz, c_vp]
    lib.vmc_qgt_dense.restype = c_int
    lib.vmc_sr_qr_workspace_size.argtypes = [c_int, c_i64, c_int]
    lib.vmc_sr_qr_workspace_size.restype = c_sz
    lib.vmc_sr_qr_solve.argtypes = [c_int, c_i64, c_vp, c_int, c_vp, c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, c_int, c_int, c_vp, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_sr_qr_solve.restype = c_int
    lib.vmc_gi_plan_create.argtypes = [ctypes.POINTER(c_vp), c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                       c_int, c_int, c_vp]
    lib.vmc_gi_plan_create.restype = c_int
    lib.vmc_gi_plan_set_degeneracy.argtypes = [c_vp, c_vp]
    lib.vmc_gi_plan_set_degeneracy.restype = c_int
    lib.vmc_gi_set_terms.argtypes = [c_vp, c_vp, c_vp]
    lib.vmc_gi_set_terms.restype = c_int
    lib.vmc_gi_slices.argtypes = [c_vp, c_vp, c_int, c_vp, c_vp, c_vp]
    lib.vmc_gi_slices.restype = c_int
    lib.vmc_gi_workspace_size.argtypes = [c_vp, c_int]
    lib.vmc_gi_workspace_size.restype = c_sz
    lib.vmc_gi_amp_grad.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_gi_amp_grad.restype = c_int
    lib.vmc_gi_init_cache.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_gi_init_cache.restype = c_int
    lib.vmc_gi_mc_sweep.argtypes = [c_vp, c_vp, c_vp, c_int, ctypes.c_uint64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp,
                                    c_vp, c_sz, c_vp]
    lib.vmc_gi_mc_sweep.restype = c_int
    lib.vmc_bl_plan_create.argtypes = [ctypes.POINTER(c_vp), c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int]
    lib.vmc_bl_plan_create.restype = c_int
    lib.vmc_bl_slices.argtypes = [c_vp, c_vp, c_int, c_vp, c_vp, c_vp]
    lib.vmc_bl_slices.restype = c_int
    lib.vmc_bl_workspace_size.argtypes = [c_vp, c_int]
    lib.vmc_bl_workspace_size.restype = c_sz
    lib.vmc_bl_amp_grad.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_bl_amp_grad.restype = c_int
    lib.vmc_bl_init_cache.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_sz, c_vp]
    lib.vmc_bl_init_cache.restype = c_int
    lib.vmc_bl_mc_sweep.argtypes = [c_vp, c_vp, c_vp, c_int, ctypes.c_uint64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp,
                                    c_vp, c_sz, c_vp]
    lib.vmc_bl_mc_sweep.restype = c_int
    lib.vmc_plan_autotune.argtypes = [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_int, c_vp]
    lib.vmc_plan_autotune.restype = c_int
    lib.vmc_plan_tuning.argtypes = [c_vp, c_vp, c_int]
    lib.vmc_plan_tuning.restype = c_int
    lib.vmc_plan_set_tuning.argtypes = [c_vp, c_vp, c_int]
    lib.vmc_plan_set_tuning.restype = c_int
    lib.vmc_profile_begin.argtypes = [c_vp]
    lib.vmc_profile_begin.restype = c_int
    lib.vmc_profile_end.argtypes = [c_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_int), c_int]
    lib.vmc_profile_end.restype = c_int
    lib.vmc_last_error.argtypes = []
    lib.vmc_last_error.restype = ctypes.c_char_p
    return lib


LIB = _load()


def check(status: int) -> None:
    if status != 0:
        raise VMCError(LIB.vmc_last_error().decode())
