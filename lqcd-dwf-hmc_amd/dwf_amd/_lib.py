"""ctypes binding of libdwfhip.so (the C ABI declared in include/dwf_hip.h).

There is no fallback: if the library or a GPU is missing, loading fails with
an exception.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdwfhip.so")

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_ip = C.POINTER(C.c_int)

F64, F32, F16, BF16 = 0, 1, 2, 3
GUESS_RHS, GUESS_GIVEN, GUESS_ZERO = 0, 1, 2

OK = 0
ERR_NOT_CONVERGED = 4


class DwfError(RuntimeError):
    def __init__(self, func, code, msg):
        super().__init__(f"{func} failed (status {code}): {msg}")
        self.code = code


class FoprParams(C.Structure):
    _fields_ = [
        ("lattice_size", C.c_int * 4),
        ("extent_of_5th_dimension", C.c_int),
        ("quark_mass", C.c_double),
        ("domain_wall_height", C.c_double),
        ("coefficient_b", _dp),
        ("coefficient_c", _dp),
        ("b_uniform", C.c_double),
        ("c_uniform", C.c_double),
        ("parameter_alpha", C.c_double),
        ("boundary_condition", C.c_int * 4),
        ("precision", C.c_int),
    ]


class RationalParams(C.Structure):
    _fields_ = [
        ("np", C.c_int),
        ("a0", C.c_double),
        ("bl", _dp),
        ("cl", _dp),
        ("max_iter", C.c_int),
        ("stop_cond", C.c_double),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "dwf_last_error": (C.c_char_p, []),
    "dwf_version": (C.c_char_p, []),
    "dwf_stream": (_vp, []),
    "dwf_synchronize": (C.c_int, []),
    "dwf_fopr_create": (C.c_int, [C.POINTER(FoprParams), C.POINTER(_vp)]),
    "dwf_fopr_destroy": (C.c_int, [_vp]),
    "dwf_fopr_set_config": (C.c_int, [_vp, _dp]),
    "dwf_fopr_set_config_device": (C.c_int, [_vp, _vp]),
    "dwf_fopr_set_mode": (C.c_int, [_vp, C.c_char_p]),
    "dwf_fopr_get_mode": (C.c_char_p, [_vp]),
    "dwf_fopr_mult": (C.c_int, [_vp, _vp, _vp]),
    "dwf_fopr_mult_dag": (C.c_int, [_vp, _vp, _vp]),
    "dwf_fopr_mult_noalloc": (C.c_int, [_vp, _vp, _vp]),
    "dwf_fopr_mult_mode": (C.c_int, [_vp, _vp, _vp, C.c_char_p]),
    "dwf_fopr_mult_dag_mode": (C.c_int, [_vp, _vp, _vp, C.c_char_p]),
    "dwf_fopr_mult_gm5": (C.c_int, [_vp, _vp, _vp]),
    "dwf_fopr_mult_gm5R": (C.c_int, [_vp, _vp, _vp]),
    "dwf_fopr_field_nin": (C.c_int, [_vp]),
    "dwf_fopr_field_nvol": (C.c_int, [_vp]),
    "dwf_fopr_field_nex": (C.c_int, [_vp]),
    "dwf_fopr_field_bytes": (C.c_size_t, [_vp]),
    "dwf_fopr_precision": (C.c_int, [_vp]),
    "dwf_fopr_flop_count": (C.c_double, [_vp, C.c_char_p]),
    "dwf_fopr_convert": (C.c_int, [_vp, _vp, _dp]),
    "dwf_fopr_reverse": (C.c_int, [_vp, _dp, _vp]),
    "dwf_fopr_convert_device": (C.c_int, [_vp, _vp, _vp]),
    "dwf_fopr_reverse_device": (C.c_int, [_vp, _vp, _vp]),
    "dwf_field_alloc": (C.c_int, [_vp, C.POINTER(_vp)]),
    "dwf_field_free": (C.c_int, [_vp]),
    "dwf_field_set_zero": (C.c_int, [_vp, _vp]),
    "dwf_field_set": (C.c_int, [_vp, _vp, C.c_double]),
    "dwf_field_copy": (C.c_int, [_vp, _vp, _vp]),
    "dwf_field_axpy": (C.c_int, [_vp, _vp, C.c_double, _vp]),
    "dwf_field_aypx": (C.c_int, [_vp, C.c_double, _vp, _vp]),
    "dwf_field_scal": (C.c_int, [_vp, _vp, C.c_double]),
    "dwf_field_norm2": (C.c_int, [_vp, _vp, _dp]),
    "dwf_field_dot": (C.c_int, [_vp, _vp, _vp, _dp]),
    "dwf_field_dotc": (C.c_int, [_vp, _vp, _vp, _dp, _dp]),
    "dwf_field_convert_precision": (C.c_int, [_vp, _vp, _vp, _vp]),
    "dwf_solver_create_cg": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.POINTER(_vp)]),
    "dwf_solver_create_richardson": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.c_int,
                                               C.c_double, C.POINTER(_vp)]),
    "dwf_solver_tune_richardson": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.c_int, _dp, C.c_int,
                                             C.c_double, C.c_int, _vp, _vp, _dp,
                                             C.POINTER(C.c_int), C.POINTER(_vp)]),
    "dwf_solver_solve": (C.c_int, [_vp, _vp, _vp, _ip, _dp]),
    "dwf_solver_inner_iterations": (C.c_int, [_vp]),
    "dwf_solver_destroy": (C.c_int, [_vp]),
    "dwf_lpc_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "dwf_lpc_destroy": (C.c_int, [_vp]),
    "dwf_lpc_degree": (C.c_int, [_vp]),
    "dwf_lpc_set_weights": (C.c_int, [_vp, _dp]),
    "dwf_lpc_get_weights": (C.c_int, [_vp, _dp]),
    "dwf_lpc_train": (C.c_int, [_vp, _vp, C.c_int, C.c_ulonglong, C.c_double, _dp]),
    "dwf_lpc_apply": (C.c_int, [_vp, C.c_char_p, _vp, _vp]),
    "dwf_lpc_chiral_gram": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), _dp]),
    "dwf_field_fill_uniform": (C.c_int, [_vp, _vp, C.c_ulonglong]),
    "dwf_solver_create_fcg": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.POINTER(_vp)]),
    "dwf_solver_create_richardson_lpc": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.c_int,
                                                   C.c_double, C.POINTER(_vp)]),
    "dwf_shiftsolver_cg_solve": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, _dp,
                                           C.POINTER(_vp), _vp, _ip, _dp]),
    "dwf_msbatch_create": (C.c_int, [C.c_int, C.POINTER(_vp), C.c_int, C.c_int,
                                     C.POINTER(_vp)]),
    "dwf_msbatch_solve": (C.c_int, [_vp, _dp, C.POINTER(_vp), C.POINTER(_vp), C.c_int,
                                    C.c_double, _ip, _dp]),
    "dwf_msbatch_iterations": (C.c_int, [_vp]),
    "dwf_msbatch_solve_mixed": (C.c_int, [_vp, C.POINTER(_vp), _dp, C.POINTER(_vp),
                                          C.POINTER(_vp), C.c_int, C.c_double, C.c_double,
                                          C.c_double, _ip, _dp]),
    "dwf_msbatch_refine_iterations": (C.c_int, [_vp]),
    "dwf_msbatch_destroy": (C.c_int, [_vp]),
    "dwf_ptc_create": (C.c_int, [_vp, C.c_int, _ip, C.POINTER(_vp)]),
    "dwf_ptc_destroy": (C.c_int, [_vp]),
    "dwf_ptc_num_weights": (C.c_size_t, [_vp]),
    "dwf_ptc_set_weights": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "dwf_ptc_get_weights": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "dwf_ptc_apply": (C.c_int, [_vp, _vp, _vp]),
    "dwf_ptc_loss_grad": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), _dp, _dp]),
    "dwf_solver_create_fcg_ptc": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.POINTER(_vp)]),
    "dwf_solver_create_richardson_ptc": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.c_int,
                                                   C.c_double, C.POINTER(_vp)]),
    "dwf_force_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "dwf_force_destroy": (C.c_int, [_vp]),
    "dwf_force_set_mode": (C.c_int, [_vp, C.c_char_p]),
    "dwf_force_udiv1": (C.c_int, [_vp, _vp, _vp, _vp]),
    "dwf_force_udiv": (C.c_int, [_vp, _vp, _vp]),
    "dwf_force_mult_generator": (C.c_int, [_vp, _vp]),
    "dwf_force_core": (C.c_int, [_vp, _vp, _vp]),
    "dwf_force_core1": (C.c_int, [_vp, _vp, _vp, _vp]),
    "dwf_rational_mult": (C.c_int, [_vp, C.POINTER(RationalParams), _vp, _vp, _ip, _dp]),
    "dwf_force_rational_ratio_create": (C.c_int, [_vp, _vp, C.POINTER(RationalParams),
                                                  C.POINTER(RationalParams), C.POINTER(_vp)]),
    "dwf_force_rational_ratio_destroy": (C.c_int, [_vp]),
    "dwf_force_rational_ratio_udiv": (C.c_int, [_vp, _vp, _vp]),
    "dwf_force_rational_ratio_core": (C.c_int, [_vp, _vp, _vp]),
    "dwf_device_alloc": (C.c_int, [C.c_size_t, C.POINTER(_vp)]),
    "dwf_device_free": (C.c_int, [_vp]),
    "dwf_memcpy_to_device": (C.c_int, [_vp, _vp, C.c_size_t]),
    "dwf_memcpy_to_host": (C.c_int, [_vp, _vp, C.c_size_t]),
    "dwf_smear_create": (C.c_int, [_ip, C.c_double, C.c_int, C.POINTER(_vp)]),
    "dwf_smear_destroy": (C.c_int, [_vp]),
    "dwf_smear_set_config": (C.c_int, [_vp, _dp]),
    "dwf_smear_set_config_device": (C.c_int, [_vp, _vp]),
    "dwf_smear_get_config": (C.c_int, [_vp, C.c_int, _dp]),
    "dwf_smear_config_device": (_vp, [_vp, C.c_int]),
    "dwf_smear_force_udiv": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "dwf_smear_mult_jacobian": (C.c_int, [_vp, _vp]),
    "dwf_smear_mult_generator": (C.c_int, [_vp, _vp]),
    "dwf_fopr_set_config_smeared": (C.c_int, [_vp, _vp, _dp]),
    "dwf_plaquette": (C.c_int, [_ip, _dp, _dp]),
    "dwf_plaquette_device": (C.c_int, [_ip, _vp, _dp]),
    "dwf_gauge_read_binary": (C.c_int, [C.c_char_p, _ip, _dp]),
    "dwf_gauge_write_binary": (C.c_int, [C.c_char_p, _ip, _dp]),
    "dwf_rng_create": (C.c_int, [_ip, C.c_ulong, C.POINTER(_vp)]),
    "dwf_rng_reset": (C.c_int, [_vp, C.c_ulong]),
    "dwf_rng_destroy": (C.c_int, [_vp]),
    "dwf_rng_get": (C.c_int, [_vp, _dp]),
    "dwf_rng_gauss_lex_global": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "dwf_rng_gauss_eo_global": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "dwf_hmc_set_iP": (C.c_int, [_ip, _vp, _dp, _dp]),
    "dwf_hmc_exp_update": (C.c_int, [_ip, _vp, C.c_double, _vp, C.c_int]),
    "dwf_hmc_rect_action": (C.c_int, [_ip, _vp, C.c_double, C.c_double, C.c_double, _dp, _dp]),
    "dwf_hmc_rect_force": (C.c_int, [_ip, _vp, _vp, C.c_double, C.c_double, C.c_double]),
    "dwf_device_axpy": (C.c_int, [_vp, C.c_double, _vp, C.c_size_t]),
    "dwf_device_copy": (C.c_int, [_vp, _vp, C.c_size_t]),
    "dwf_device_norm2": (C.c_int, [_vp, C.c_size_t, _dp]),
}

_lib = None


def load():
    """Load libdwfhip.so; raises if it is missing or lacks a declared symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libdwfhip.so is not built ({LIB_PATH}); run "
                           "python lqcd-dwf-hmc_amd/build.py")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return sorted(_SIGS)


def check(func, rc, allow=()):
    if rc != OK and rc not in allow:
        msg = load().dwf_last_error().decode(errors="replace")
        raise DwfError(func, rc, msg)
    return rc
