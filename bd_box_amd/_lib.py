"""ctypes binding of libbdbox_hip.so (the C ABI declared in include/bdbox_hip.h).

The library is built in-tree (`make -C bd_box_amd/csrc`).  Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbdbox_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u64p = C.POINTER(C.c_uint64)

# every symbol declared in include/bdbox_hip.h
EXPORTS = [
    "bdh_last_error", "bdh_init", "bdh_create", "bdh_destroy", "bdh_step", "bdh_compute_D",
    "bdh_compute_forces", "bdh_factor", "bdh_noise_product", "bdh_compute_Rf",
    "bdh_get_coords", "bdh_set_coords", "bdh_get_forces", "bdh_get_last_displacement",
    "bdh_get_rng_state", "bdh_set_rng_state", "bdh_dpotrf", "bdh_dtrmv", "bdh_gauss",
    "bdh_seed_state", "bdh_probe_mfma", "bdh_time_factor", "bdh_time_steps", "bdh_sync",
    "bdh_phase_timing", "bdh_phase_times", "bdh_neighbour_pairs", "bdh_last_dt",
    "bdh_batch_create", "bdh_batch_destroy", "bdh_batch_step", "bdh_batch_get_coords",
    "bdh_batch_get_forces", "bdh_batch_time_steps", "bdh_batch_phase_timing",
    "bdh_batch_phase_times", "bdh_step_vjp", "bdh_factor_cache", "bdh_save_factor",
    "bdh_step_vjp_cached", "bdh_batch_set_schedule", "bdh_batch_get_schedule",
    "bdh_batch_autotune", "bdh_step_vjp_params", "bdh_set_params", "bdh_compute_Dij",
    "bdh_compute_r", "bdh_ermak_move", "bdh_igt_phase", "bdh_set_forces",
    "bdh_tea_coefficients", "bdh_batch_set_tiles", "bdh_batch_get_tiles",
    "bdh_batch_autotune_tiles", "bdh_set_ewald_split", "bdh_gauss_noalloc", "bdh_set_flow",
    "bdh_set_ewald_cutoff",
]


class Config(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("T", C.c_double), ("visc", C.c_double), ("vfactor", C.c_double),
        ("alpha_lj", C.c_double), ("cutoff_lj", C.c_double), ("lj_6_term", C.c_int),
        ("elec", C.c_int), ("epsilon_c", C.c_double), ("kappa_c", C.c_double),
        ("gamma_c", C.c_double), ("cutoff_c", C.c_double), ("bond_lj_scale", C.c_double),
        ("bond_c_scale", C.c_double), ("check_overlap", C.c_int), ("bc", C.c_int),
        ("box", C.c_double), ("hydro", C.c_int), ("ewald_method", C.c_int),
        ("ewald_real", C.c_int), ("precision", C.c_int), ("refactor_every", C.c_int),
        ("rand_seed", C.c_uint32), ("nb_list", C.c_int), ("algorithm", C.c_int),
        ("move_attempts", C.c_int), ("ewald_alpha", C.c_double),
    ]


class BackendError(RuntimeError):
    pass


_lib = None


def load():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BackendError(f"{LIB_PATH} is not built; run `make -C bd_box_amd/csrc`")
    L = C.CDLL(LIB_PATH)
    L.bdh_last_error.restype = C.c_char_p
    L.bdh_init.argtypes = [C.c_int]
    L.bdh_create.argtypes = [C.POINTER(Config), C.c_int, _dp, _dp, _dp, C.c_int, _ip, _dp,
                             C.POINTER(C.c_void_p)]
    L.bdh_destroy.argtypes = [C.c_void_p]
    L.bdh_destroy.restype = None
    L.bdh_step.argtypes = [C.c_void_p, C.c_int, _dp]
    L.bdh_compute_D.argtypes = [C.c_void_p, _dp, C.c_int]
    L.bdh_compute_forces.argtypes = [C.c_void_p, _dp, _dp, _ip]
    L.bdh_neighbour_pairs.argtypes = [C.c_void_p, C.c_double, _ip, C.c_longlong,
                                      C.POINTER(C.c_longlong)]
    L.bdh_factor.argtypes = [C.c_void_p, _dp, C.c_int]
    L.bdh_noise_product.argtypes = [C.c_void_p, _dp, _dp]
    L.bdh_compute_Rf.argtypes = [C.c_void_p, _dp]
    L.bdh_get_coords.argtypes = [C.c_void_p, _dp]
    L.bdh_set_coords.argtypes = [C.c_void_p, _dp]
    L.bdh_get_forces.argtypes = [C.c_void_p, _dp]
    L.bdh_get_last_displacement.argtypes = [C.c_void_p, _dp, _dp]
    L.bdh_last_dt.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.bdh_get_rng_state.argtypes = [C.c_void_p, _u64p]
    L.bdh_set_rng_state.argtypes = [C.c_void_p, C.c_uint64]
    L.bdh_dpotrf.argtypes = [C.c_char, _dp, C.c_int, C.c_int, C.c_int]
    L.bdh_dtrmv.argtypes = [C.c_char, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
    L.bdh_gauss.argtypes = [_u64p, _dp, C.c_int]
    L.bdh_gauss_noalloc.argtypes = [_u64p, _dp, C.c_int]
    L.bdh_seed_state.argtypes = [C.c_uint32]
    L.bdh_seed_state.restype = C.c_uint64
    L.bdh_probe_mfma.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.bdh_time_factor.argtypes = [C.c_void_p, C.c_int, _dp]
    L.bdh_time_steps.argtypes = [C.c_void_p, C.c_int, _dp]
    L.bdh_sync.argtypes = [C.c_void_p]
    L.bdh_phase_timing.argtypes = [C.c_void_p, C.c_int]
    L.bdh_phase_times.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_longlong)]
    L.bdh_step_vjp.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.bdh_factor_cache.argtypes = [C.c_void_p, C.c_int]
    L.bdh_save_factor.argtypes = [C.c_void_p, C.c_int]
    L.bdh_step_vjp_cached.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]
    L.bdh_step_vjp_params.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.bdh_set_params.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.bdh_compute_Dij.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp]
    L.bdh_compute_r.argtypes = [C.c_void_p, _dp]
    L.bdh_ermak_move.argtypes = [C.c_void_p, C.c_double, _dp]
    L.bdh_igt_phase.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _ip]
    L.bdh_set_forces.argtypes = [C.c_void_p, _dp]
    L.bdh_tea_coefficients.argtypes = [C.c_void_p, _dp, _dp]
    L.bdh_set_ewald_split.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int]
    L.bdh_set_flow.argtypes = [C.c_void_p, _dp]
    L.bdh_set_ewald_cutoff.argtypes = [C.c_void_p, C.c_double]
    L.bdh_batch_create.argtypes = [C.POINTER(Config), C.c_int, _ip, _dp, _dp, _dp, _ip, _ip, _dp,
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_void_p)]
    L.bdh_batch_destroy.argtypes = [C.c_void_p]
    L.bdh_batch_destroy.restype = None
    L.bdh_batch_step.argtypes = [C.c_void_p, C.c_int, _dp]
    L.bdh_batch_get_coords.argtypes = [C.c_void_p, _dp]
    L.bdh_batch_get_forces.argtypes = [C.c_void_p, _dp]
    L.bdh_batch_time_steps.argtypes = [C.c_void_p, C.c_int, _dp]
    L.bdh_batch_phase_timing.argtypes = [C.c_void_p, C.c_int]
    L.bdh_batch_phase_times.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_longlong)]
    L.bdh_batch_set_schedule.argtypes = [C.c_void_p, C.c_int]
    L.bdh_batch_get_schedule.argtypes = [C.c_void_p, _ip]
    L.bdh_batch_autotune.argtypes = [C.c_void_p, C.c_int, _ip, _dp, _ip, _ip]
    L.bdh_batch_set_tiles.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.bdh_batch_get_tiles.argtypes = [C.c_void_p, _ip, _ip]
    L.bdh_batch_autotune_tiles.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, _ip, _ip]
    _lib = L
    return L


def check(rc: int, what: str) -> int:
    if rc < 0:
        msg = load().bdh_last_error().decode(errors="replace")
        raise BackendError(f"{what} failed ({rc}): {msg}")
    return rc


def dptr(a):
    return a.ctypes.data_as(_dp)
