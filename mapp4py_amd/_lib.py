"""ctypes binding of libmapp_amd.so (the C ABI declared in include/mapp_amd.h).

The library is built in-tree (``python -c "import __graft_entry__ as g; g.build()"``
or ``make -C mapp4py_amd/csrc``). Importing the package fails loudly when the
shared object is missing: there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmapp_amd.so")

_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
i8p = C.POINTER(C.c_int8)
u8p = C.POINTER(C.c_uint8)
vpp = C.POINTER(C.c_void_p)
fp = C.POINTER(C.c_float)

# name -> (restype, argtypes); every exported symbol of include/mapp_amd.h
SIGNATURES = {
    "mapp_last_error": (C.c_char_p, []),
    "mapp_version": (C.c_char_p, []),
    "mapp_eam_create_setfl": (C.c_int, [C.c_char_p, C.POINTER(C.c_char_p), C.c_int, C.c_int, vpp]),
    "mapp_eam_create_fs": (C.c_int, [C.c_char_p, C.POINTER(C.c_char_p), C.c_int, C.c_int, vpp]),
    "mapp_eam_create_funcfl": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.c_int, vpp]),
    "mapp_eam_create_knots": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dp, dp, dp,
                                        C.c_int, C.c_int, vpp]),
    "mapp_eam_destroy": (None, [C.c_void_p]),
    "mapp_eam_cutoff": (C.c_int, [C.c_void_p, ip, dp]),
    "mapp_eam_coefs": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, dp]),
    "mapp_md_create": (C.c_int, [C.c_void_p, C.c_int, dp, dp, u8p, C.c_double, vpp]),
    "mapp_md_destroy": (None, [C.c_void_p]),
    "mapp_md_set_x": (C.c_int, [C.c_void_p, dp]),
    "mapp_md_get_x": (C.c_int, [C.c_void_p, dp]),
    "mapp_md_set_v": (C.c_int, [C.c_void_p, dp]),
    "mapp_md_get_v": (C.c_int, [C.c_void_p, dp]),
    "mapp_md_set_box": (C.c_int, [C.c_void_p, dp, dp]),
    "mapp_md_update": (C.c_int, [C.c_void_p, ip]),
    "mapp_md_force_calc": (C.c_int, [C.c_void_p, dp, dp]),
    "mapp_md_value": (C.c_int, [C.c_void_p, dp]),
    "mapp_md_value_noalloc": (C.c_int, [C.c_void_p, dp]),
    "mapp_md_get_rho": (C.c_int, [C.c_void_p, dp]),
    "mapp_md_pairs_count": (C.c_int, [C.c_void_p, i64p]),
    "mapp_md_pairs": (C.c_int, [C.c_void_p, i32p, i32p, i8p]),
    "mapp_md_list_stats": (C.c_int, [C.c_void_p, ip, i64p, ip]),
    "mapp_md_nvt_run": (C.c_int, [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, ip]),
    "mapp_md_nst_run": (C.c_int, [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, dp, C.c_void_p, C.c_int,
                                  C.c_int, dp, dp, dp, C.c_void_p, ip]),
    "mapp_md_muvt_run": (C.c_int, [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, ip, dp, ip]),
    "mapp_md_natms": (C.c_int, [C.c_void_p, ip]),
    "mapp_md_get_elem": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mapp_md_minimize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, dp, C.c_int, ip]),
    "mapp_dmd_min_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_void_p, ip, ip, dp]),
    "mapp_line_search": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int, dp, C.c_void_p, C.c_void_p,
                                   C.c_void_p, dp, dp, ip]),
    "mapp_md_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "mapp_md_profile_read": (C.c_int, [C.c_void_p, C.c_char_p, dp, i64p]),
    "mapp_md_set_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mapp_md_set_lds": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "mapp_md_lds_window": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mapp_dmd_set_box": (C.c_int, [C.c_void_p, dp, dp]),
    "mapp_dmd_force_calc_static": (C.c_int, [C.c_void_p, dp]),
    "mapp_dmd_set_fp32": (C.c_int, [C.c_void_p, C.c_int]),
    "mapp_dmd_knot_window": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mapp_dmd_set_batch": (C.c_int, [C.c_void_p, C.c_int]),
    "mapp_dmd_c_d": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_double)]),
    "mapp_dmd_J": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mapp_dmd_newton": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                  C.POINTER(C.c_double), C.c_double, C.c_double, C.c_int, C.c_int,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mapp_dmd_bdf_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                   C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int)]),
    "mapp_dmd_Jxa": (C.c_int, [C.c_void_p, dp, dp, dp, dp, C.c_void_p, C.c_void_p]),
    "mapp_dmd_min_newton": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int, ip, C.POINTER(C.c_double)]),
    "mapp_cfg_md_read": (C.c_int, [C.c_char_p, vpp]),
    "mapp_cfg_md_info": (C.c_int, [C.c_void_p, ip, ip, ip]),
    "mapp_cfg_md_get": (C.c_int, [C.c_void_p, dp, dp, dp, u8p, dp]),
    "mapp_cfg_md_element": (C.c_char_p, [C.c_void_p, C.c_int]),
    "mapp_cfg_md_free": (None, [C.c_void_p]),
    "mapp_cfg_md_write": (C.c_int, [C.c_char_p, C.c_int, dp, dp, u8p, dp, C.c_int, dp, C.POINTER(C.c_char_p),
                                    C.c_int, dp, C.POINTER(C.c_char_p)]),
    "mapp_md_export_cfg": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, dp, C.POINTER(C.c_char_p), C.c_int]),
    "mapp_cfg_dmd_read": (C.c_int, [C.c_char_p, vpp]),
    "mapp_cfg_dmd_info": (C.c_int, [C.c_void_p, ip, ip, ip]),
    "mapp_cfg_dmd_get": (C.c_int, [C.c_void_p, dp, dp, u8p, dp, dp, dp]),
    "mapp_cfg_dmd_element": (C.c_char_p, [C.c_void_p, C.c_int]),
    "mapp_cfg_dmd_free": (None, [C.c_void_p]),
    "mapp_cfg_dmd_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, dp, u8p, dp, dp, C.c_int, dp,
                                     C.POINTER(C.c_char_p)]),
    "mapp_dmd_export_cfg": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, dp, C.POINTER(C.c_char_p)]),
    "mapp_dmd_unit_gradient": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mapp_dmd_J_entries": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mapp_dmd_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, dp, dp, u8p, dp, dp, C.c_double,
                                  C.c_double, C.c_double, C.c_double, dp, vpp]),
    "mapp_dmd_destroy": (None, [C.c_void_p]),
    "mapp_dmd_set_state": (C.c_int, [C.c_void_p, dp, dp, dp]),
    "mapp_dmd_get_state": (C.c_int, [C.c_void_p, dp, dp, dp]),
    "mapp_dmd_update": (C.c_int, [C.c_void_p, ip]),
    "mapp_dmd_derivative": (C.c_int, [C.c_void_p, C.c_int, dp, dp, dp, dp, dp]),
    "mapp_dmd_value": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
    "mapp_dmd_pairs_count": (C.c_int, [C.c_void_p, i64p]),
    "mapp_dmd_pairs": (C.c_int, [C.c_void_p, i32p, i32p, i8p]),
    "mapp_dmd_smear": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, dp, dp, dp]),
    "mapp_dmd_min_cg": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, ip, ip, dp]),
    "mapp_dmd_min_lbfgs": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, ip, ip, dp]),
    "mapp_dmd_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "mapp_dmd_profile_read": (C.c_int, [C.c_void_p, C.c_char_p, dp, i64p]),
    "mapp_gauss_hermite": (C.c_int, [C.c_int, dp, dp]),
    "mapp_surr_create": (C.c_int, [C.c_int, fp, C.c_int, vpp]),
    "mapp_surr_destroy": (None, [C.c_void_p]),
    "mapp_surr_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, dp]),
    "mapp_surr_loss_grad": (C.c_int, [C.c_void_p, C.c_int, C.c_int, fp, fp, fp, dp]),
    "mapp_surr_get_params": (C.c_int, [C.c_void_p, C.c_int, fp]),
    "mapp_surr_set_params": (C.c_int, [C.c_void_p, C.c_int, fp]),
    "mapp_surr_get_norm": (C.c_int, [C.c_void_p, C.c_int, fp, fp]),
    "mapp_surr_set_norm": (C.c_int, [C.c_void_p, C.c_int, fp, fp]),
    "mapp_surr_infer": (C.c_int, [C.c_void_p, C.c_int, C.c_int, fp, fp]),
    "mapp_dmd_use_surrogate": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mapp_dmd_surrogate_coverage": (C.c_int, [C.c_void_p, i64p, i64p]),
    "mapp_gcmc_create": (C.c_int, [C.c_void_p, C.c_int, i32p, dp, dp, u8p, i32p, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_double, dp, i32p, vpp]),
    "mapp_gcmc_destroy": (None, [C.c_void_p]),
    "mapp_replica_create": (C.c_int, [C.c_void_p, C.c_int, i32p, dp, dp, u8p, C.c_double, vpp]),
    "mapp_replica_destroy": (None, [C.c_void_p]),
    "mapp_replica_min_cg": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_double, i32p, i32p, dp, dp, i32p]),
    "mapp_replica_tune": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, i32p, dp]),
    "mapp_replica_set_waves": (C.c_int, [C.c_void_p, i32p]),
    "mapp_replica_use_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "mapp_replica_get_x": (C.c_int, [C.c_void_p, C.c_int, dp]),
    "mapp_replica_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "mapp_replica_profile_read": (C.c_int, [C.c_void_p, dp, C.POINTER(C.c_int64)]),
    "mapp_gcmc_attempts": (C.c_int, [C.c_void_p, C.c_int, i8p, dp]),
    "mapp_gcmc_tune": (C.c_int, [C.c_void_p, C.c_int, i32p, dp]),
    "mapp_gcmc_set_waves": (C.c_int, [C.c_void_p, i32p]),
    "mapp_gcmc_use_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "mapp_gcmc_xchng_energy": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, C.c_int, dp]),
    "mapp_gcmc_xchng_commit": (C.c_int, [C.c_void_p, dp]),
    "mapp_gcmc_counts": (C.c_int, [C.c_void_p, i32p, i32p]),
    "mapp_gcmc_get_cell": (C.c_int, [C.c_void_p, C.c_int, dp, u8p, dp, dp, dp, i32p]),
    "mapp_gcmc_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "mapp_gcmc_profile_read": (C.c_int, [C.c_void_p, C.c_char_p, dp, i64p]),
    "mapp_diag_copy": (C.c_int, [C.c_size_t, C.c_int, dp]),
    "mapp_create_T": (C.c_int, [C.c_int, u8p, dp, C.c_double, C.c_double, C.c_int, dp]),
}


class MappError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, status, msg):
        super().__init__(f"mapp status {status}: {msg}")
        self.status = status


def lib():
    """Load libmapp_amd.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build the HIP library first (make -C mapp4py_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise MappError(status, lib().mapp_last_error().decode())


def ptr(a, t=dp):
    return a.ctypes.data_as(t)


class MinParams(C.Structure):
    _fields_ = [("lbfgs", C.c_int), ("m", C.c_int), ("e_tol", C.c_double), ("max_dx", C.c_double),
                ("max_nsteps", C.c_int), ("ls_c", C.c_double), ("ls_rho", C.c_double),
                ("ls_min_alpha", C.c_double), ("ls", C.c_int), ("ls_max_iter", C.c_int), ("ls_brack", C.c_int),
                ("ls_tol", C.c_double), ("ls_zeps", C.c_double)]


class MinReport(C.Structure):
    _fields_ = [("steps", C.c_int), ("status", C.c_int), ("evals", C.c_int), ("force_calcs", C.c_int),
                ("rebuilds", C.c_int), ("e0", C.c_double), ("e", C.c_double), ("f_norm", C.c_double),
                ("f_max", C.c_double)]


class BdfParams(C.Structure):
    _fields_ = [("a_tol", C.c_double), ("min_dt", C.c_double), ("max_nsteps", C.c_int),
                ("max_nnewton_iters", C.c_int), ("max_ngmres_iters", C.c_int), ("nreset", C.c_int),
                ("ntally", C.c_int), ("relax_max_iter", C.c_int)]


class BdfReport(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("max_q", C.c_int), ("t", C.c_double), ("max_dt", C.c_double),
                ("min_dt", C.c_double), ("nonlin_acc", C.c_int), ("nonlin_rej", C.c_int), ("intpol_acc", C.c_int),
                ("intpol_rej", C.c_int), ("integ_acc", C.c_int), ("integ_rej", C.c_int), ("err_mins", C.c_int)]


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a
