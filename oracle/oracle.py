"""ctypes wrapper for the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  The HIP band-solve path's results are compared
against the oracle; only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s
cpu_baseline leg import this module.  The product package ``blaze2d_amd``
never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAX_ATOMS = 64


class _OracleJob(ctypes.Structure):
    _fields_ = [
        ("a1", ctypes.c_double * 2),
        ("a2", ctypes.c_double * 2),
        ("eps_bg", ctypes.c_double),
        ("n_atoms", ctypes.c_int),
        ("atoms", (ctypes.c_double * 4) * MAX_ATOMS),
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
        ("kpath", ctypes.POINTER(ctypes.c_double)),
    ]


def build() -> str:
    """Compile liboracle.so from oracle.cpp (g++)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_bandstructure.argtypes = [ctypes.POINTER(_OracleJob), dp, ip, ip]
        L.oracle_bandstructure.restype = ctypes.c_int
        L.oracle_bandstructure_residuals.argtypes = [ctypes.POINTER(_OracleJob), dp, ip, ip, dp, ip]
        L.oracle_bandstructure_warm.argtypes = [ctypes.POINTER(_OracleJob), dp, ctypes.c_int, ctypes.c_int, dp, ip,
                                                ip, dp, ip]
        L.oracle_bandstructure_residuals.restype = ctypes.c_int
        L.oracle_dielectric.argtypes = [ctypes.POINTER(_OracleJob), dp, dp, dp]
        L.oracle_dielectric.restype = ctypes.c_int
        L.oracle_fft2d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int]
        L.oracle_fft2d.restype = None
        L.oracle_operator_apply.argtypes = [ctypes.POINTER(_OracleJob), dp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_operator_apply.restype = ctypes.c_int
        L.oracle_eigh.argtypes = [ctypes.c_int, dp, dp, dp]
        L.oracle_eigh.restype = None
        L.oracle_svqb.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_svqb.restype = ctypes.c_int
        L.oracle_polar.argtypes = [ctypes.c_int, dp, dp]
        L.oracle_polar.restype = ctypes.c_double
        L.oracle_hungarian.argtypes = [ctypes.c_int, dp, ip]
        L.oracle_hungarian.restype = None
        L.oracle_lobpcg_diagonal.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, dp]
        L.oracle_lobpcg_diagonal.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


@dataclass
class Job:
    """A band-structure job (mirrors bandstructure::BandStructureJob)."""

    lattice: str = "square"  # "square" | "hex"
    eps_bg: float = 13.0
    atoms: List[Tuple[float, float, float, float]] = field(default_factory=lambda: [(0.5, 0.5, 0.3, 1.0)])
    nx: int = 24
    ny: int = 24
    pol: str = "TM"
    n_bands: int = 8
    max_iter: int = 200
    tol: float = 1e-6
    block_size: int = 0
    smoothing: bool = True
    precond: int = -1
    subspace_prediction: bool = True
    extrapolation: bool = True
    band_tracking: bool = True
    reuse_gamma: bool = True
    exit_rule: int = 0
    kpath: Optional[np.ndarray] = None

    def lattice_vectors(self):
        if self.lattice == "square":
            return (1.0, 0.0), (0.0, 1.0)
        if self.lattice in ("hex", "hexagonal", "triangular"):
            return (1.0, 0.0), (0.5, np.sqrt(3.0) * 0.5)
        raise ValueError(self.lattice)

    def to_c(self) -> Tuple[_OracleJob, np.ndarray]:
        j = _OracleJob()
        a1, a2 = self.lattice_vectors()
        j.a1[0], j.a1[1] = a1
        j.a2[0], j.a2[1] = a2
        j.eps_bg = self.eps_bg
        if len(self.atoms) > MAX_ATOMS:
            raise ValueError("too many atoms")
        j.n_atoms = len(self.atoms)
        for i, at in enumerate(self.atoms):
            for t in range(4):
                j.atoms[i][t] = float(at[t])
        j.nx, j.ny = self.nx, self.ny
        j.pol = 0 if self.pol.upper() == "TM" else 1
        j.n_bands = self.n_bands
        j.max_iter = self.max_iter
        j.tol = self.tol
        j.block_size = self.block_size
        j.smoothing = int(self.smoothing)
        j.precond = self.precond
        j.subspace_prediction = int(self.subspace_prediction)
        j.extrapolation = int(self.extrapolation)
        j.band_tracking = int(self.band_tracking)
        j.reuse_gamma = int(self.reuse_gamma)
        j.exit_rule = self.exit_rule
        kp = np.ascontiguousarray(self.kpath if self.kpath is not None else np.zeros((1, 2)), dtype=np.float64)
        j.nk = kp.shape[0]
        j.kpath = _dp(kp)
        return j, kp


def bandstructure(job: Job):
    """Returns (omegas[nk, nb], iterations[nk], converged[nk], total_iterations)."""
    cj, kp = job.to_c()
    nk = cj.nk
    om = np.zeros((nk, job.n_bands))
    it = np.zeros(nk, dtype=np.int32)
    cv = np.zeros(nk, dtype=np.int32)
    tot = lib().oracle_bandstructure(ctypes.byref(cj), _dp(om), _ip(it), _ip(cv))
    if tot < 0:
        raise RuntimeError("oracle_bandstructure failed")
    return om, it, cv, tot


def bandstructure_residuals(job: Job):
    """bandstructure() plus, per k-point, the relative residual B-norms of the
    first n_bands block entries at the solve's last log iteration and that
    iteration (1-based): (omegas, iterations, converged, total, rel[nk, nb], rel_iter[nk])."""
    cj, kp = job.to_c()
    nk = cj.nk
    om = np.zeros((nk, job.n_bands))
    it = np.zeros(nk, dtype=np.int32)
    cv = np.zeros(nk, dtype=np.int32)
    rel = np.zeros((nk, job.n_bands))
    rit = np.zeros(nk, dtype=np.int32)
    tot = lib().oracle_bandstructure_residuals(ctypes.byref(cj), _dp(om), _ip(it), _ip(cv), _dp(rel), _ip(rit))
    if tot < 0:
        raise RuntimeError("oracle_bandstructure_residuals failed")
    return om, it, cv, tot, rel, rit


def bandstructure_warm(job: Job, warm_pw: np.ndarray, pw_half: int, residuals: bool = False):
    """bandstructure() with warm vectors for the first k-point given as
    plane-wave coefficients [count, (2 pw_half + 1)^2] complex (the learned
    start's layout): (omegas, iterations, converged, total), plus
    (rel[nk, nb], rel_iter[nk]) as in bandstructure_residuals when residuals."""
    cj, kp = job.to_c()
    nk = cj.nk
    w = np.ascontiguousarray(warm_pw, dtype=np.complex128).reshape(-1, (2 * pw_half + 1) ** 2)
    om = np.zeros((nk, job.n_bands))
    it = np.zeros(nk, dtype=np.int32)
    cv = np.zeros(nk, dtype=np.int32)
    rel = np.zeros((nk, job.n_bands))
    rit = np.zeros(nk, dtype=np.int32)
    tot = lib().oracle_bandstructure_warm(ctypes.byref(cj), _dp(w.view(np.float64)), w.shape[0], pw_half, _dp(om),
                                          _ip(it), _ip(cv), _dp(rel), _ip(rit))
    if tot < 0:
        raise RuntimeError("oracle_bandstructure_warm failed")
    if residuals:
        return om, it, cv, tot, rel, rit
    return om, it, cv, tot


def dielectric(job: Job):
    cj, kp = job.to_c()
    n = job.nx * job.ny
    eps = np.zeros(n)
    inv = np.zeros(n)
    ten = np.zeros((n, 4))
    lib().oracle_dielectric(ctypes.byref(cj), _dp(eps), _dp(inv), _dp(ten))
    return eps, inv, ten


def fft2d(data: np.ndarray, inverse: bool = False) -> np.ndarray:
    """data: complex [count, ny, nx] or [ny, nx]."""
    a = np.ascontiguousarray(data, dtype=np.complex128).copy()
    shp = a.shape
    if a.ndim == 2:
        a = a[None]
    cnt, ny, nx = a.shape
    lib().oracle_fft2d(nx, ny, cnt, _dp(a.view(np.float64)), 1 if inverse else -1)
    return a.reshape(shp)


def operator_apply(job: Job, kfrac, which: str, x: np.ndarray) -> np.ndarray:
    """which: 'A' (Theta), 'B' (mass) or 'P' (preconditioner). x: [count, N] complex."""
    cj, kp = job.to_c()
    a = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex128)
    out = np.zeros_like(a)
    k = np.ascontiguousarray(kfrac, dtype=np.float64)
    w = {"A": 0, "B": 1, "P": 2}[which]
    lib().oracle_operator_apply(ctypes.byref(cj), _dp(k), w, a.shape[0], _dp(a.view(np.float64)),
                                _dp(out.view(np.float64)))
    return out


def kpg_sq(job: Job, kfrac) -> np.ndarray:
    """Clamped |k+G|^2 on the grid, row-major iy * nx + ix
    (maxwell.rs:1573-1690: FFT index order, |k+G| clamped to >= sqrt(1e-9))."""
    a1, a2 = job.lattice_vectors()
    det = a1[0] * a2[1] - a1[1] * a2[0]
    b1 = (2 * np.pi * a2[1] / det, -2 * np.pi * a2[0] / det)
    b2 = (-2 * np.pi * a1[1] / det, 2 * np.pi * a1[0] / det)
    n1 = np.array([i if i <= job.nx // 2 else i - job.nx for i in range(job.nx)], dtype=float)
    n2 = np.array([i if i <= job.ny // 2 else i - job.ny for i in range(job.ny)], dtype=float)
    N2, N1 = np.meshgrid(n2, n1, indexing="ij")
    bx = kfrac[0] * b1[0] + kfrac[1] * b2[0]
    by = kfrac[0] * b1[1] + kfrac[1] * b2[1]
    kx = N1 * b1[0] + N2 * b2[0] + bx
    ky = N1 * b1[1] + N2 * b2[1] + by
    sq = kx * kx + ky * ky
    small = sq < 1e-9
    sc = np.sqrt(1e-9 / np.where(sq > 0, sq, 1.0))
    kx = np.where(small, np.where(sq > 0, kx * sc, np.sqrt(1e-9)), kx)
    ky = np.where(small, np.where(sq > 0, ky * sc, 0.0), ky)
    return (kx * kx + ky * ky).reshape(-1)


def diagnostic_start_vectors(n: int):
    """The x and y start vectors of maxwell.rs:386-391 and :529-538."""
    i = np.arange(n, dtype=np.float64)
    x = np.sin(i * 0.618033988749895) + 1j * np.cos(i * 0.414213562373095)
    y = np.cos(i * np.sqrt(2.0)) + 1j * np.sin(i * np.sqrt(3.0))
    return x, y


def operator_diagnostics(job: Job, power_iterations: int = 15):
    """run_core's one-time diagnostics (bandstructure.rs:437-498) over the
    oracle's operators: ThetaOperator::check_self_adjointness,
    estimate_condition_number and estimate_preconditioned_condition_number
    (maxwell.rs:382-582) at the first non-Gamma k-point (|k|^2 >= 1e-8,
    bandstructure.rs:72-81).  Returns a dict, or None when the path is all Gamma.
    Pinned by the closed form of the TM power iteration (Theta_TM is diagonal
    in Fourier space) in tests/test_oracle.py."""
    kp = np.atleast_2d(job.kpath if job.kpath is not None else np.zeros((1, 2)))
    idx = next((i for i, k in enumerate(kp) if k[0] * k[0] + k[1] * k[1] >= 1e-8), None)
    if idx is None:
        return None
    k = kp[idx]
    n = job.nx * job.ny
    x, y = diagnostic_start_vectors(n)

    def A(v):
        return operator_apply(job, k, "A", v[None])[0]

    def P(v):
        return operator_apply(job, k, "P", v[None])[0]

    ax, ay = A(x), A(y)
    scale = np.linalg.norm(ax) * np.linalg.norm(y)
    sa = abs(np.vdot(ax, y) - np.vdot(x, ay)) / scale if scale > 1e-15 else 0.0

    def power(pre):
        v = x / np.linalg.norm(x)
        lam = 0.0
        for _ in range(power_iterations):
            av = A(v)
            if pre:
                av = P(av)
            lam = float(np.real(np.vdot(v, av)))
            nr = np.linalg.norm(av)
            if nr > 1e-15:
                av = av / nr
            v = av
        return lam

    prec = job.precond if job.precond >= 0 else (1 if job.pol.upper() == "TM" else 2)
    lam_max = power(False)
    lam_min = float(kpg_sq(job, k).min())
    pm = power(True) if prec else 0.0
    return {
        "k_index": idx, "precond": prec, "lambda_max": lam_max, "lambda_min": lam_min,
        "kappa": lam_max / lam_min if lam_min > 1e-15 else float("inf"),
        "pm_lambda_max": pm, "pm_kappa": pm, "self_adjoint_err": float(sa),
    }


def eigh(a: np.ndarray):
    n = a.shape[0]
    ac = np.asfortranarray(a, dtype=np.complex128)
    w = np.zeros(n)
    v = np.zeros((n, n), dtype=np.complex128, order="F")
    lib().oracle_eigh(n, _dp(ac.reshape(-1, order="F").view(np.float64)), _dp(w),
                      _dp(v.reshape(-1, order="F").view(np.float64)))
    return w, v


def eigh_colmajor(a: np.ndarray):
    n = a.shape[0]
    flat = np.ascontiguousarray(a.T.reshape(-1), dtype=np.complex128)
    w = np.zeros(n)
    v = np.zeros(n * n, dtype=np.complex128)
    lib().oracle_eigh(n, _dp(flat.view(np.float64)), _dp(w), _dp(v.view(np.float64)))
    return w, v.reshape(n, n).T


def svqb(x: np.ndarray, bx: np.ndarray):
    """x, bx: [p, len] complex.  Returns (x_out, bx_out, rank)."""
    xc = np.ascontiguousarray(x, dtype=np.complex128).copy()
    bc = np.ascontiguousarray(bx, dtype=np.complex128).copy()
    p, n = xc.shape
    r = lib().oracle_svqb(n, p, _dp(xc.view(np.float64)), _dp(bc.view(np.float64)))
    return xc, bc, r


def polar(o: np.ndarray):
    m = o.shape[0]
    flat = np.ascontiguousarray(o.T.reshape(-1), dtype=np.complex128)
    u = np.zeros(m * m, dtype=np.complex128)
    s = lib().oracle_polar(m, _dp(flat.view(np.float64)), _dp(u.view(np.float64)))
    return u.reshape(m, m).T, s


def hungarian(w: np.ndarray) -> np.ndarray:
    n = w.shape[0]
    wc = np.ascontiguousarray(w, dtype=np.float64)
    p = np.zeros(n, dtype=np.int32)
    lib().oracle_hungarian(n, _dp(wc), _ip(p))
    return p


def lobpcg_diagonal(diag: Sequence[float], n_bands: int, max_iter: int = 200, tol: float = 1e-6,
                    precond: bool = False):
    d = np.ascontiguousarray(diag, dtype=np.float64)
    ev = np.zeros(n_bands)
    it = lib().oracle_lobpcg_diagonal(len(d), _dp(d), n_bands, max_iter, tol, int(precond), _dp(ev))
    return ev, it


def kpath(lattice: str, segments_per_leg: int) -> np.ndarray:
    """brillouin.rs:220-264 generate_path with the square / hexagonal presets."""
    if lattice == "square":
        nodes = [(0.0, 0.0), (0.5, 0.0), (0.5, 0.5), (0.0, 0.0)]
    else:
        nodes = [(0.0, 0.0), (0.5, 0.0), (1.0 / 3.0, 1.0 / 3.0), (0.0, 0.0)]
    seg = max(segments_per_leg, 1)
    path = [nodes[0]]
    for a, b in zip(nodes[:-1], nodes[1:]):
        for step in range(1, seg + 1):
            t = step / seg
            pt = ((1.0 - t) * a[0] + t * b[0], (1.0 - t) * a[1] + t * b[1])
            if path[-1] != pt:
                path.append(pt)
    return np.array(path, dtype=np.float64)
