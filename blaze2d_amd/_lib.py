"""ctypes binding of libblaze2d_hip.so (include/blaze2d_hip.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared library has not been built.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libblaze2d_hip.so")
MAX_ATOMS = 64

c_dp = ctypes.POINTER(ctypes.c_double)
c_ip = ctypes.POINTER(ctypes.c_int)


class Geometry(ctypes.Structure):
    _fields_ = [
        ("a1", ctypes.c_double * 2),
        ("a2", ctypes.c_double * 2),
        ("eps_bg", ctypes.c_double),
        ("n_atoms", ctypes.c_int),
        ("atoms", (ctypes.c_double * 4) * MAX_ATOMS),
    ]


class ThetaDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("pol", ctypes.c_int),
        ("b1", ctypes.c_double * 2),
        ("b2", ctypes.c_double * 2),
        ("bloch", ctypes.c_double * 2),
        ("eps", c_dp),
        ("tensor", c_dp),
        ("precond", ctypes.c_int),
        ("precision", ctypes.c_int),
    ]


class Job(ctypes.Structure):
    _fields_ = [
        ("geom", Geometry),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("pol", ctypes.c_int),
        ("n_bands", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("block_size", ctypes.c_int),
        ("smoothing", ctypes.c_int),
        ("precond", ctypes.c_int),
        ("subspace_prediction", ctypes.c_int),
        ("extrapolation", ctypes.c_int),
        ("band_tracking", ctypes.c_int),
        ("reuse_gamma", ctypes.c_int),
        ("exit_rule", ctypes.c_int),
        ("nk", ctypes.c_int),
        ("kpath", c_dp),
        ("warm_pw", c_dp),
        ("warm_count", ctypes.c_int),
        ("pw_half", ctypes.c_int),
    ]


class RunOptions(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("precision", ctypes.c_int), ("profile", ctypes.c_int),
                ("groups", ctypes.c_int), ("timed_kernel", ctypes.c_int), ("engines", ctypes.c_int),
                ("gram_mfma_min", ctypes.c_int)]


class RunStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_long),
        ("iterations", ctypes.c_long),
        ("solves", ctypes.c_long),
        ("seconds", ctypes.c_double),
        ("eigh_calls", ctypes.c_long * 2),
        ("eigh_sweeps", ctypes.c_long * 2),
        ("eigh_max_sweeps", ctypes.c_long * 2),
        ("eigh_launch_max_sum", ctypes.c_long * 2),
        ("timed_ms", ctypes.c_double),
        ("timed_launches", ctypes.c_long),
        ("timed_bytes", ctypes.c_double),
        ("timed_flops", ctypes.c_double),
        ("host_wait_seconds", ctypes.c_double),
        ("host_plan_seconds", ctypes.c_double),
        ("rr_problems", ctypes.c_long),
        ("svqb_fallbacks", ctypes.c_long),
        ("rank_drops", ctypes.c_long),
        ("timed_flops_executed", ctypes.c_double),
        ("alg_bytes", ctypes.c_double),
    ]


class RunOutputs(ctypes.Structure):
    _fields_ = [
        ("rel_residuals", c_dp),
        ("rel_residual_iters", c_ip),
        ("pw_half", ctypes.c_int),
        ("pw_coefs", c_dp),
        ("prec_residuals", c_dp),
    ]


class Diagnostics(ctypes.Structure):
    _fields_ = [
        ("k_index", ctypes.c_int),
        ("precond", ctypes.c_int),
        ("lambda_max", ctypes.c_double),
        ("lambda_min", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("pm_lambda_max", ctypes.c_double),
        ("pm_kappa", ctypes.c_double),
        ("self_adjoint_err", ctypes.c_double),
    ]


EXPORTS = [
    "bh_version", "bh_last_error", "bh_device_count", "bh_context_create", "bh_context_destroy", "bh_synchronize",
    "bh_buffer_alloc", "bh_buffer_free", "bh_buffer_upload", "bh_buffer_download", "bh_fft2d", "bh_scale",
    "bh_axpy", "bh_dot", "bh_gram", "bh_lincomb", "bh_theta_create", "bh_theta_destroy", "bh_theta_apply",
    "bh_mass_apply", "bh_precond_apply", "bh_dielectric", "bh_dielectric_batch", "bh_bandstructure_run",
    "bh_hermitian_eigh", "bh_operator_diagnostics", "bh_rayleigh_ritz",
    "bh_bandstructure_run_ex", "bh_gram_noalloc", "bh_cnn_create", "bh_cnn_destroy", "bh_cnn_param_count",
    "bh_cnn_get_params", "bh_cnn_set_params", "bh_cnn_get_grads", "bh_cnn_predict", "bh_cnn_train_step",
    "bh_cnn_predict_geometries", "bh_cnn_set_graphs",
]

_lib = None


def lib() -> ctypes.CDLL:
    """Load the HIP backend.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m blaze2d_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.bh_version.restype = ctypes.c_char_p
    L.bh_last_error.restype = ctypes.c_char_p
    L.bh_device_count.restype = ctypes.c_int
    L.bh_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.bh_context_destroy.argtypes = [vp]
    L.bh_context_destroy.restype = None
    L.bh_synchronize.argtypes = [vp]
    L.bh_buffer_alloc.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.bh_buffer_free.argtypes = [vp]
    L.bh_buffer_free.restype = None
    L.bh_buffer_upload.argtypes = [vp, ctypes.c_int, ctypes.c_int, c_dp]
    L.bh_buffer_download.argtypes = [vp, ctypes.c_int, ctypes.c_int, c_dp]
    L.bh_fft2d.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.bh_scale.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.bh_axpy.argtypes = [vp, ctypes.c_double, ctypes.c_double, vp, ctypes.c_int, vp, ctypes.c_int]
    L.bh_dot.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, c_dp]
    L.bh_gram.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, c_dp]
    L.bh_gram_noalloc.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, c_dp]
    L.bh_hermitian_eigh.argtypes = [vp, ctypes.c_int, ctypes.c_int, c_dp, c_dp, c_dp]
    L.bh_rayleigh_ritz.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_dp, c_dp, c_dp, c_dp, c_ip]
    L.bh_lincomb.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, c_dp, ctypes.c_int, vp, ctypes.c_int]
    L.bh_theta_create.argtypes = [vp, ctypes.POINTER(ThetaDesc), ctypes.POINTER(vp)]
    L.bh_theta_destroy.argtypes = [vp]
    L.bh_theta_destroy.restype = None
    L.bh_theta_apply.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int]
    L.bh_mass_apply.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int]
    L.bh_precond_apply.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    L.bh_dielectric.argtypes = [ctypes.POINTER(Geometry), ctypes.c_int, ctypes.c_int, ctypes.c_int, c_dp, c_dp, c_dp]
    L.bh_dielectric_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(Geometry), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, c_dp, c_dp, c_dp, c_dp]
    L.bh_bandstructure_run.argtypes = [vp, ctypes.POINTER(Job), ctypes.c_int, ctypes.POINTER(RunOptions), c_dp, c_ip,
                                       c_ip, ctypes.POINTER(RunStats)]
    L.bh_bandstructure_run_ex.argtypes = [vp, ctypes.POINTER(Job), ctypes.c_int, ctypes.POINTER(RunOptions), c_dp,
                                          c_ip, c_ip, ctypes.POINTER(RunOutputs), ctypes.POINTER(RunStats)]
    L.bh_operator_diagnostics.argtypes = [vp, ctypes.POINTER(Job), ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(Diagnostics), c_dp]
    c_fp = ctypes.POINTER(ctypes.c_float)
    L.bh_cnn_create.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_ulonglong, ctypes.POINTER(vp)]
    L.bh_cnn_destroy.argtypes = [vp]
    L.bh_cnn_destroy.restype = None
    L.bh_cnn_param_count.argtypes = [vp]
    L.bh_cnn_param_count.restype = ctypes.c_longlong
    L.bh_cnn_get_params.argtypes = [vp, c_fp]
    L.bh_cnn_set_params.argtypes = [vp, c_fp]
    L.bh_cnn_get_grads.argtypes = [vp, c_fp]
    L.bh_cnn_predict.argtypes = [vp, ctypes.c_int, c_fp, c_fp, c_dp]
    L.bh_cnn_predict_geometries.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Geometry), ctypes.c_int, c_fp, c_dp]
    L.bh_cnn_train_step.argtypes = [vp, ctypes.c_int, c_fp, c_fp, c_dp, ctypes.c_float, c_dp]
    L.bh_cnn_set_graphs.argtypes = [vp, ctypes.c_int]
    _lib = L
    return L


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().bh_last_error().decode(errors="replace")
        raise RuntimeError(f"{what}: {msg}" if what else msg)
