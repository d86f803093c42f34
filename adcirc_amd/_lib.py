"""ctypes binding to libadcgpu.so (the C ABI in include/adcgpu.h).

The library is the product: there is no CPU fallback.  Loading fails loudly if
the library has not been built (``python -m adcirc_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libadcgpu.so")

dp_ = C.POINTER(C.c_double)
ip_ = C.POINTER(C.c_int32)

FP64, MIXED = 0, 1
FIELDS = dict(eta1=0, eta2=1, etas=2, uu1=3, vv1=4, uu2=5, vv2=6, qx1=7, qy1=8, qx2=9, qy2=10,
              h1=11, h2=12, tk=13, etamax=14, totalarea=15, wsx1=16, wsy1=17, wsx2=18,
              wsy2=19, pr1=20, pr2=21, gwce_lv=22, umax=23, etas_prev=24, tip1=25, tip2=26,
              nodecode=100, noff=101,
              mju=102, etamax_it=103, umax_it=104,
              coef=200, alphal=300)
INT_FIELDS = ("nodecode", "noff", "mju", "etamax_it", "umax_it")

# every symbol include/adcgpu.h declares
EXPORTS = ("adcgpu_create", "adcgpu_destroy", "adcgpu_last_error", "adcgpu_set_field",
           "adcgpu_get_field", "adcgpu_cold_start", "adcgpu_begin_step", "adcgpu_friction",
           "adcgpu_set_forcing", "adcgpu_solve_gwce", "adcgpu_wetdry", "adcgpu_momentum",
           "adcgpu_timestep", "adcgpu_last_numitr", "adcgpu_synchronize", "adcgpu_run",
           "adcgpu_run_iterations", "adcgpu_run_rebuilds", "adcgpu_elevation_bc", "adcgpu_jcg", "adcgpu_holland",
           "adcgpu_profile", "adcgpu_profile_read", "adcgpu_jcg_shape", "adcgpu_run_stats",
           "adcgpu_set_wait_limit", "adcgpu_jcg_noalloc", "adcgpu_ml_params",
           "adcgpu_ml_collect", "adcgpu_ml_samples", "adcgpu_ml_train", "adcgpu_ml_grad",
           "adcgpu_ml_features", "adcgpu_ml_correction", "adcgpu_state_bytes",
           "adcgpu_state_save", "adcgpu_state_load", "adcgpu_last_solve",
           "adcgpu_set_jcg_blocks", "adcgpu_set_tide_potential", "adcgpu_set_hybrid_friction", "adcgpu_set_flux_boundary",
           "adcgpu_set_flux", "adcgpu_set_flux_table", "adcgpu_flux_bc", "adcgpu_batch_create", "adcgpu_batch_run", "adcgpu_batch_destroy")
ML_NPARAM = 22913
# error codes (ADCGPU_E_*)
E_ARG, E_UNSUPPORTED, E_HIP, E_DIAG, E_CAPACITY, E_NONFINITE, E_ERRORELEV = range(-1, -8, -1)
JCG_NOALLOC_MAX_N, JCG_NOALLOC_MAX_NZ = 32768, 16


class Mesh(C.Structure):
    _fields_ = [("np", C.c_int32), ("ne", C.c_int32), ("mnei", C.c_int32), ("neta", C.c_int32),
                ("nvel", C.c_int32), ("nvelme", C.c_int32),
                ("nm", ip_), ("x", dp_), ("y", dp_), ("dp", dp_), ("fdxe", dp_), ("fdye", dp_),
                ("areas", dp_), ("nneigh", ip_), ("neitab", ip_), ("csr_position", ip_),
                ("nbd", ip_), ("nbv", ip_), ("lbcodei", ip_), ("csii", dp_), ("siii", dp_),
                ("me2gw", ip_)]


class Params(C.Structure):
    _fields_ = [("dtdp", C.c_double), ("g", C.c_double), ("h0", C.c_double),
                ("velmin", C.c_double), ("a00", C.c_double), ("b00", C.c_double),
                ("c00", C.c_double), ("convcr", C.c_double), ("prbckgrnd_mh2o", C.c_double),
                ("nolifa", C.c_int32), ("nolica", C.c_int32), ("nolicat", C.c_int32),
                ("nolibf", C.c_int32), ("itmax", C.c_int32), ("nws", C.c_int32),
                ("precision", C.c_int32), ("tau0var", dp_), ("evm", dp_), ("fric", dp_),
                ("corif", dp_), ("ilump", C.c_int32), ("warm_start", C.c_int32),
                ("errorelev", C.c_double)]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m adcirc_amd.build`")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.adcgpu_create.argtypes = [C.POINTER(Mesh), C.POINTER(Params), C.c_int, C.POINTER(vp)]
    lib.adcgpu_destroy.argtypes = [vp]
    lib.adcgpu_destroy.restype = None
    lib.adcgpu_last_error.restype = C.c_char_p
    lib.adcgpu_set_field.argtypes = [vp, C.c_int, vp, C.c_int64]
    lib.adcgpu_get_field.argtypes = [vp, C.c_int, vp, C.c_int64]
    lib.adcgpu_cold_start.argtypes = [vp, dp_, ip_, ip_]
    lib.adcgpu_begin_step.argtypes = [vp, C.c_int]
    lib.adcgpu_friction.argtypes = [vp]
    lib.adcgpu_set_forcing.argtypes = [vp, dp_, dp_, dp_]
    lib.adcgpu_solve_gwce.argtypes = [vp, C.c_int, dp_]
    lib.adcgpu_wetdry.argtypes = [vp, C.c_int]
    lib.adcgpu_momentum.argtypes = [vp]
    lib.adcgpu_timestep.argtypes = [vp, C.c_int, dp_, dp_, dp_, dp_]
    lib.adcgpu_last_numitr.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.adcgpu_synchronize.argtypes = [vp]
    lib.adcgpu_run.argtypes = [vp, C.c_int, C.c_int, dp_, dp_, C.c_int]
    lib.adcgpu_run_iterations.argtypes = [vp]
    lib.adcgpu_run_iterations.restype = C.c_int64
    lib.adcgpu_run_rebuilds.argtypes = [vp]
    lib.adcgpu_run_rebuilds.restype = C.c_int64
    lib.adcgpu_elevation_bc.argtypes = [C.c_int, C.c_int, dp_, dp_, dp_, dp_, dp_, C.c_double,
                                        C.c_double, dp_]
    lib.adcgpu_jcg.argtypes = [C.c_int, C.c_int, C.c_int, ip_, dp_, dp_, dp_, ip_, dp_]
    lib.adcgpu_holland.argtypes = [vp, C.c_double, dp_]
    lib.adcgpu_profile.argtypes = [vp, C.c_int]
    lib.adcgpu_profile_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    lib.adcgpu_jcg_shape.argtypes = [vp, ip_]
    lib.adcgpu_run_stats.argtypes = [vp, C.POINTER(C.c_int64)]
    lib.adcgpu_last_solve.argtypes = [vp, dp_]
    lib.adcgpu_set_jcg_blocks.argtypes = [vp, C.c_int]
    lib.adcgpu_set_tide_potential.argtypes = [vp, C.c_int, dp_, dp_, dp_, dp_, dp_, dp_, dp_,
                                              C.c_int, C.c_double, C.c_double, C.c_double]
    lib.adcgpu_set_hybrid_friction.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    lib.adcgpu_set_flux_boundary.argtypes = [vp, dp_]
    lib.adcgpu_set_flux.argtypes = [vp, dp_]
    lib.adcgpu_set_flux_table.argtypes = [vp, C.c_int, dp_]
    lib.adcgpu_flux_bc.argtypes = [C.c_int, ip_, C.c_int, dp_, dp_, dp_, dp_, dp_, C.c_double,
                                   C.c_double, dp_]
    lib.adcgpu_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
    lib.adcgpu_batch_run.argtypes = [vp, C.c_int, C.c_int, C.POINTER(dp_), C.POINTER(dp_)]
    lib.adcgpu_batch_destroy.argtypes = [vp]
    lib.adcgpu_batch_destroy.restype = None
    lib.adcgpu_set_wait_limit.argtypes = [C.c_double]
    lib.adcgpu_jcg_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, ip_, dp_, dp_, dp_, ip_, dp_]
    lib.adcgpu_state_bytes.argtypes = [vp]
    lib.adcgpu_state_bytes.restype = C.c_int64
    lib.adcgpu_state_save.argtypes = [vp, vp, C.c_int64]
    lib.adcgpu_state_load.argtypes = [vp, vp, C.c_int64]
    fp_ = C.POINTER(C.c_float)
    lib.adcgpu_ml_params.argtypes = [vp, fp_, fp_]
    lib.adcgpu_ml_collect.argtypes = [vp, C.c_int]
    lib.adcgpu_ml_samples.argtypes = [vp]
    lib.adcgpu_ml_samples.restype = C.c_int64
    lib.adcgpu_ml_train.argtypes = [vp, C.c_int, C.c_double, C.c_uint64, dp_]
    lib.adcgpu_ml_grad.argtypes = [vp, C.c_int, fp_, fp_, fp_, fp_, dp_]
    lib.adcgpu_ml_features.argtypes = [vp, fp_, dp_]
    lib.adcgpu_ml_correction.argtypes = [vp, dp_]
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().adcgpu_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")
