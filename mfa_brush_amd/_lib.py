"""ctypes binding of libmfa_hip.so (include/mfa_hip.h).

The HIP library is the product; there is no CPU fallback.  Loading fails
loudly when the shared object is missing or was not built for this tree.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfa_hip.so")

MFA_PREC_FP64 = 0
MFA_PREC_MIXED = 1

_D3 = C.c_double * 3
_D16 = C.c_double * 16
_D4 = C.c_double * 4
_D2 = C.c_double * 2


class SystemParams(C.Structure):
    """mfa_system_params (include/mfa_hip.h)."""
    _fields_ = [
        ("n_mon", C.c_int32), ("n_chain", C.c_int32), ("n_mon_d", C.c_int32), ("n_chain_d", C.c_int32),
        ("thermostat", C.c_int32), ("dpd_cut_off", C.c_int32), ("force_switch_on", C.c_int32),
        ("bending", C.c_int32), ("bending_melt", C.c_int32), ("poiseuille", C.c_int32),
        ("l_box", C.c_int32), ("bend_form", C.c_int32),
        ("boundary", _D3),
        ("epsil", _D16), ("sigma_2", _D16), ("range_2", _D16), ("e_shift", _D16), ("inv_range_2", _D16),
        ("r_cut_max", C.c_double), ("r_cut_dpd_2", C.c_double), ("skin", C.c_double),
        ("sig", C.c_double), ("friction", C.c_double),
        ("r_2_min_init", C.c_double), ("r_2_min_time", C.c_double),
        ("sigma_wall", _D4), ("a_wall", _D4), ("a_w", C.c_double),
        ("k_chain", C.c_double), ("inv_r_chain_2", C.c_double),
        ("k_bend", C.c_double), ("alpha_eq", C.c_double),
        ("dt", C.c_double),
        ("const_force", _D3), ("head_vel", _D2), ("mass_type", _D4),
        ("seed", C.c_uint64), ("t_idx0", C.c_int64), ("dpd_vv", C.c_int32), ("reserved1", C.c_int32),
    ]


class Energies(C.Structure):
    _fields_ = [("v_fluid_fluid", C.c_double), ("v_fluid_wall", C.c_double), ("v_intra_molec", C.c_double),
                ("v_bend", C.c_double), ("v_bend_melt", C.c_double), ("t_fluid", C.c_double),
                ("virial", C.c_double * 6), ("virial_bond", C.c_double * 6)]


class Flags(C.Structure):
    _fields_ = [("nonfinite", C.c_int32), ("out_of_box_z", C.c_int32), ("fene_stretched", C.c_int32),
                ("rebuilds", C.c_int32), ("list_overflow", C.c_int32)]


# every symbol include/mfa_hip.h declares, with its ctypes signature
_vp, _i32, _i64, _d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
SIGNATURES = {
    "mfa_hip_create": [_vp, _i32, _i32, _vp],
    "mfa_hip_destroy": [_vp],
    "mfa_hip_set_state": [_vp, _i32, _vp, _vp, _vp, _vp],
    "mfa_hip_get_state": [_vp, _i32, _vp, _vp, _vp, _vp],
    "mfa_hip_set_time": [_vp, _i64],
    "mfa_hip_get_time": [_vp, _vp],
    "mfa_hip_verlet_positions": [_vp],
    "mfa_hip_check_skin": [_vp, _vp],
    "mfa_hip_fluid_fluid": [_vp],
    "mfa_hip_fluid_wall": [_vp, _i32],
    "mfa_hip_wall_wall": [_vp, _i32, _d, _vp],
    "mfa_hip_intra_molec": [_vp],
    "mfa_hip_bending": [_vp, _i32],
    "mfa_hip_bending_melt": [_vp, _i32],
    "mfa_hip_constant_force": [_vp, _i32],
    "mfa_hip_verlet_velocities": [_vp],
    "mfa_hip_new_dpd_fd": [_vp],
    "mfa_hip_md_steps": [_vp, _i32, _i32],
    "mfa_hip_prepare_graphs": [_vp],
    "mfa_hip_synchronize": [_vp],
    "mfa_hip_get_energies": [_vp, _i32, _vp],
    "mfa_hip_get_f_on_heads": [_vp, _i32, _vp],
    "mfa_hip_get_flags": [_vp, _vp],
    "mfa_hip_check_fluid_fluid": [_vp, _i32, _vp],
    "mfa_hip_pair_stats": [_vp, _i32, _vp, _vp],
    "mfa_hip_check_fluid_fluid_noalloc": [_vp, _vp, _vp, _i32, _d, _vp, _vp],
    "mfa_hip_profiles_reset": [_vp, _i32],
    "mfa_hip_profiles_accumulate": [_vp],
    "mfa_hip_profiles_get": [_vp, _i32, _vp, _vp, _vp, _vp],
    "mfa_hip_make_binning": [_i32, _vp, _d, _vp, _vp, _vp, _vp, _vp],
    "mfa_hip_get_cells": [_vp, _i32, _vp, _vp, _vp, _vp],
    "mfa_hip_set_timing": [_vp, _i32],
    "mfa_hip_get_kernel_time": [_vp, C.c_char_p, _vp, _vp],
    "mfa_hip_set_virial": [_vp, _i32],
    "mfa_hip_set_neighbor_list": [_vp, _i32],
    "mfa_hip_get_neighbor_list": [_vp, _vp],
    "mfa_hip_set_dpd_queue": [_vp, _i32],
    "mfa_hip_set_pair_grid": [_vp, _i32],
    "mfa_hip_chain_observables": [_vp, _i32, _vp],
    "mfa_hip_list_stats": [_vp, _i32, _vp],
    "mfa_hip_get_pair_count": [_vp, _i32, _vp],
    "mfa_hip_last_error": [],
    "mfa_hip_version": [],
}

_lib = None


class MfaError(RuntimeError):
    pass


def load(path: str = LIB_PATH) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MfaError("libmfa_hip.so is not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.mfa_hip_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().mfa_hip_last_error().decode(errors="replace")
        raise MfaError("%s failed (%d): %s" % (what, rc, msg))
