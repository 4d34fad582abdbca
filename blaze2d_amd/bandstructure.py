"""Band-structure jobs on the HIP backend.

Mirrors ``BandStructureJob`` / ``RunOptions`` / ``run_with_options``
(crates/core/src/drivers/bandstructure.rs:98-351), ``Geometry2D`` /
``BasisAtom`` (crates/core/src/geometry.rs), ``Lattice2D``
(crates/core/src/lattice.rs:143-180) and the k-path presets
(crates/core/src/brillouin.rs:220-264, symmetry.rs:57-101).  Frequencies are
returned as omega = sqrt(lambda) like ``BandStructureResult.bands``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from .backend import HipBackend


@dataclass
class Lattice2D:
    a1: tuple
    a2: tuple

    @staticmethod
    def square(a: float = 1.0) -> "Lattice2D":
        return Lattice2D((a, 0.0), (0.0, a))

    @staticmethod
    def hexagonal(a: float = 1.0) -> "Lattice2D":
        return Lattice2D((a, 0.0), (0.5 * a, np.sqrt(3.0) * 0.5 * a))

    def reciprocal(self):
        det = self.a1[0] * self.a2[1] - self.a1[1] * self.a2[0]
        inv = 2.0 * np.pi / det
        return (self.a2[1] * inv, -self.a2[0] * inv), (-self.a1[1] * inv, self.a1[0] * inv)


@dataclass
class BasisAtom:
    pos: tuple
    radius: float
    eps_inside: float = 1.0


@dataclass
class Geometry2D:
    lattice: Lattice2D
    eps_bg: float = 12.0
    atoms: List[BasisAtom] = field(default_factory=list)

    def to_c(self) -> L.Geometry:
        g = L.Geometry()
        g.a1[0], g.a1[1] = self.lattice.a1
        g.a2[0], g.a2[1] = self.lattice.a2
        g.eps_bg = self.eps_bg
        if len(self.atoms) > L.MAX_ATOMS:
            raise ValueError(f"at most {L.MAX_ATOMS} atoms")
        g.n_atoms = len(self.atoms)
        for i, a in enumerate(self.atoms):
            g.atoms[i][0], g.atoms[i][1] = a.pos
            g.atoms[i][2] = a.radius
            g.atoms[i][3] = a.eps_inside
        return g


def k_path(lattice: str, segments_per_leg: int) -> np.ndarray:
    """Square Gamma-X-M-Gamma or hexagonal Gamma-M-K-Gamma, linearly densified."""
    if lattice == "square":
        nodes = [(0.0, 0.0), (0.5, 0.0), (0.5, 0.5), (0.0, 0.0)]
    elif lattice in ("hex", "hexagonal", "triangular"):
        nodes = [(0.0, 0.0), (0.5, 0.0), (1.0 / 3.0, 1.0 / 3.0), (0.0, 0.0)]
    else:
        raise ValueError(lattice)
    seg = max(int(segments_per_leg), 1)
    path = [nodes[0]]
    for a, b in zip(nodes[:-1], nodes[1:]):
        for step in range(1, seg + 1):
            t = step / seg
            p = ((1.0 - t) * a[0] + t * b[0], (1.0 - t) * a[1] + t * b[1])
            if path[-1] != p:
                path.append(p)
    return np.array(path, dtype=np.float64)


@dataclass
class RunOptions:
    """bandstructure.rs:98-280 RunOptions (plus the solver exit rule)."""

    precond: int = -1                  # -1 auto: TM fourier-diagonal, TE transverse projection
    use_subspace_prediction: bool = True
    use_extrapolation: bool = True
    disable_band_tracking: bool = False
    reuse_gamma: bool = True
    exit_rule: int = 0                 # 0 Eigensolver::solve() accounting, 1 soft-locked bands only


@dataclass
class BandStructureJob:
    geom: Geometry2D
    nx: int
    ny: int
    pol: str
    k_path: np.ndarray
    n_bands: int = 8
    tol: float = 1e-6
    max_iter: int = 200
    block_size: int = 0
    smoothing: bool = True
    # learned start: warm vectors for the first k-point as plane-wave
    # coefficients [count, (2 pw_half + 1)^2] complex (index (n2 + h)(2h + 1) + n1 + h)
    warm_pw: Optional[np.ndarray] = None
    pw_half: int = 0

    def to_c(self, opts: RunOptions) -> L.Job:
        j = L.Job()
        j.geom = self.geom.to_c()
        j.nx, j.ny = self.nx, self.ny
        j.pol = 0 if self.pol.upper() == "TM" else 1
        j.n_bands = self.n_bands
        j.max_iter = self.max_iter
        j.tol = self.tol
        j.block_size = self.block_size
        j.smoothing = int(self.smoothing)
        j.precond = opts.precond
        j.subspace_prediction = int(opts.use_subspace_prediction)
        j.extrapolation = int(opts.use_extrapolation)
        j.band_tracking = int(not opts.disable_band_tracking)
        j.reuse_gamma = int(opts.reuse_gamma)
        j.exit_rule = opts.exit_rule
        self._kp = np.ascontiguousarray(self.k_path, dtype=np.float64)
        if self._kp.ndim != 2 or self._kp.shape[1] != 2 or self._kp.shape[0] < 1:
            raise ValueError("k_path must be an (nk, 2) array")
        j.nk = self._kp.shape[0]
        j.kpath = self._kp.ctypes.data_as(L.c_dp)
        if self.warm_pw is not None:
            ng = (2 * self.pw_half + 1) ** 2
            self._warm = np.ascontiguousarray(self.warm_pw, dtype=np.complex128).reshape(-1, ng)
            j.warm_pw = self._warm.view(np.float64).ctypes.data_as(L.c_dp)
            j.warm_count = self._warm.shape[0]
            j.pw_half = self.pw_half
        return j


@dataclass
class BandStructureResult:
    k_path: np.ndarray
    bands: np.ndarray           # [nk, n_bands] omega
    iterations: np.ndarray      # [nk]
    converged: np.ndarray       # [nk]
    # relative residual B-norms ||r||_B / |lambda| of the first n_bands block
    # entries at each solve's last log iteration (eigensolver.rs:744-779), and
    # that iteration (1-based)
    relative_residuals: Optional[np.ndarray] = None   # [nk, n_bands]
    residual_iteration: Optional[np.ndarray] = None   # [nk]
    # with run_batch(pw_half=h): the converged eigenvectors (ascending eigenvalue)
    # on the plane waves |n1|, |n2| <= h: [nk, n_bands, (2h + 1)^2] complex
    pw_coefs: Optional[np.ndarray] = None
    # B-norms ||M^-1 r_i||_B of the preconditioned residuals of each solve's
    # first iteration (the start block after its Rayleigh-Ritz): [nk, n_bands]
    first_prec_residuals: Optional[np.ndarray] = None


@dataclass
class BatchStats:
    steps: int
    iterations: int
    solves: int
    seconds: float
    timed_ms: float = 0.0           # launches of the timed kernel class (event-timed on their stream)
    timed_launches: int = 0
    timed_bytes: float = 0.0        # their algorithmic HBM bytes
    timed_flops: float = 0.0        # their algorithmic FLOPs
    eigh_calls: tuple = (0, 0)      # (SVQB, Rayleigh-Ritz), profile mode only
    eigh_sweeps: tuple = (0, 0)
    eigh_max_sweeps: tuple = (0, 0)
    eigh_launch_max_sum: tuple = (0, 0)
    host_wait_seconds: float = 0.0
    host_plan_seconds: float = 0.0
    rr_problems: int = 0            # Rayleigh-Ritz problems solved
    svqb_fallbacks: int = 0         # of them, B-orthonormalised on the SVQB path
    rank_drops: int = 0             # of them, with a direction dropped
    timed_flops_executed: float = 0.0  # FLOPs the timed kernels ran (Rayleigh-Ritz: Jacobi sweeps as run)
    alg_bytes: float = 0.0          # algorithmic HBM bytes of every launch of the run


def run_batch(backend: HipBackend, jobs: Sequence[BandStructureJob], options: Optional[RunOptions] = None,
              slots: int = 64, precision: int = 0, profile: bool = False, groups: int = 4, timed_kernel: int = 0,
              pw_half: int = 0, engines: int = 1, gram_mfma_min: int = 0):
    """Runs many band structures concurrently on one GPU.

    `slots` band structures are resident at once; they are split into `groups`
    pipelined on separate HIP streams.  Jobs may use different grids: each
    grid size runs as an engine of its own, and `engines` > 1 splits the
    largest ones further; the engines run side by side (own streams, own host
    thread), sharing the slots in proportion to their work.  `gram_mfma_min`
    (0: 33) is the width from which FP64 Gram products run on the matrix
    cores.  `timed_kernel` (1 Theta operator,
    2 dual Rayleigh-Ritz Gram, 3 Rayleigh-Ritz solve, 4 Ritz update, 5
    deflation) times every launch of that kernel class with HIP events on its
    stream during the run.
    """
    opts = options or RunOptions()
    jobs = list(jobs)
    if not jobs:
        return [], BatchStats(0, 0, 0, 0.0)
    arr = (L.Job * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = j.to_c(opts)
    tot_om = sum(j.k_path.shape[0] * j.n_bands for j in jobs)
    tot_k = sum(j.k_path.shape[0] for j in jobs)
    om = np.zeros(tot_om)
    it = np.zeros(tot_k, dtype=np.int32)
    cv = np.zeros(tot_k, dtype=np.int32)
    rel = np.zeros(tot_om)
    rit = np.zeros(tot_k, dtype=np.int32)
    prec = np.zeros(tot_om)
    ng = (2 * pw_half + 1) ** 2
    pw = np.zeros(tot_om * ng if pw_half > 0 else 1, dtype=np.complex128)
    ro = L.RunOptions(slots, precision, int(profile), groups, timed_kernel, engines, gram_mfma_min)
    st = L.RunStats()
    extra = L.RunOutputs(rel.ctypes.data_as(L.c_dp), rit.ctypes.data_as(L.c_ip), pw_half,
                         pw.view(np.float64).ctypes.data_as(L.c_dp) if pw_half > 0 else None,
                         prec.ctypes.data_as(L.c_dp))
    L.check(L.lib().bh_bandstructure_run_ex(backend.handle, arr, len(jobs), ctypes.byref(ro),
                                            om.ctypes.data_as(L.c_dp), it.ctypes.data_as(L.c_ip),
                                            cv.ctypes.data_as(L.c_ip), ctypes.byref(extra), ctypes.byref(st)),
            "bh_bandstructure_run_ex")
    out = []
    o = k = 0
    for j in jobs:
        nk = j.k_path.shape[0]
        out.append(BandStructureResult(j.k_path, om[o:o + nk * j.n_bands].reshape(nk, j.n_bands).copy(),
                                       it[k:k + nk].copy(), cv[k:k + nk].copy(),
                                       rel[o:o + nk * j.n_bands].reshape(nk, j.n_bands).copy(),
                                       rit[k:k + nk].copy(),
                                       pw[o * ng:(o + nk * j.n_bands) * ng].reshape(nk, j.n_bands, ng).copy()
                                       if pw_half > 0 else None,
                                       prec[o:o + nk * j.n_bands].reshape(nk, j.n_bands).copy()))
        o += nk * j.n_bands
        k += nk
    stats = BatchStats(st.steps, st.iterations, st.solves, st.seconds, st.timed_ms, st.timed_launches,
                       st.timed_bytes, st.timed_flops, tuple(st.eigh_calls), tuple(st.eigh_sweeps),
                       tuple(st.eigh_max_sweeps), tuple(st.eigh_launch_max_sum), st.host_wait_seconds,
                       st.host_plan_seconds, st.rr_problems, st.svqb_fallbacks, st.rank_drops,
                       st.timed_flops_executed, st.alg_bytes)
    return out, stats


def run_with_options(backend: HipBackend, job: BandStructureJob, options: Optional[RunOptions] = None):
    """bandstructure::run_with_options for one job (bandstructure.rs:345-351)."""
    res, _ = run_batch(backend, [job], options, slots=1)
    return res[0]


def dielectric_batch(backend: HipBackend, geoms: Sequence[Geometry2D], nx: int, ny: int, smoothing: bool = True):
    """Dielectric2D::from_geometry for many geometries at once, on the GPU.

    Returns (eps[G, N], inv_eps[G, N], tensor[G, N, 4], kernel milliseconds).
    """
    ng = len(geoms)
    if ng == 0:
        raise ValueError("no geometries")
    arr = (L.Geometry * ng)(*[g.to_c() for g in geoms])
    n = nx * ny
    eps = np.zeros((ng, n))
    inv = np.zeros((ng, n))
    ten = np.zeros((ng, n, 4))
    ms = ctypes.c_double(0.0)
    L.check(L.lib().bh_dielectric_batch(backend.handle, arr, ng, nx, ny, int(smoothing), eps.ctypes.data_as(L.c_dp),
                                        inv.ctypes.data_as(L.c_dp), ten.ctypes.data_as(L.c_dp), ctypes.byref(ms)),
            "bh_dielectric_batch")
    return eps, inv, ten, ms.value


@dataclass
class OperatorDiagnostics:
    """One job's one-time operator diagnostics (bandstructure.rs:437-498)."""
    k_index: int               # first non-Gamma k-point, -1 when the path has none
    precond: int               # 0 none, 1 Fourier-diagonal, 2 transverse projection
    lambda_max: float
    lambda_min: float          # smallest clamped |k+G|^2
    kappa: float               # kappa(A)
    pm_lambda_max: float
    pm_kappa: float            # kappa(M^-1 A)
    self_adjoint_err: float

    @property
    def self_adjoint_status(self) -> str:
        """The reference's label for the self-adjointness error (bandstructure.rs:471-477)."""
        e = self.self_adjoint_err
        return "exact" if e < 1e-12 else ("good" if e < 1e-6 else "WARN")

    @property
    def reduction(self) -> float:
        """kappa(A) / kappa(M^-1 A), the factor the reference logs."""
        return self.kappa / self.pm_kappa if self.pm_kappa else float("nan")


def operator_diagnostics(backend: HipBackend, jobs: Sequence[BandStructureJob], options: Optional[RunOptions] = None,
                         power_iterations: int = 15):
    """The reference's one-time per-job operator diagnostics, batched on the GPU.

    For each job, at its first non-Gamma k-point: the Theta operator's
    self-adjointness error and the power-iteration estimates of kappa(A) and
    kappa(M^-1 A) (maxwell.rs:382-582, run by run_core at
    bandstructure.rs:437-498).  Returns ([OperatorDiagnostics], device ms).
    """
    opts = options or RunOptions()
    jobs = list(jobs)
    if not jobs:
        return [], 0.0
    arr = (L.Job * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = j.to_c(opts)
    out = (L.Diagnostics * len(jobs))()
    ms = ctypes.c_double(0.0)
    L.check(L.lib().bh_operator_diagnostics(backend.handle, arr, len(jobs), power_iterations, out, ctypes.byref(ms)),
            "bh_operator_diagnostics")
    res = [OperatorDiagnostics(d.k_index, d.precond, d.lambda_max, d.lambda_min, d.kappa, d.pm_lambda_max,
                               d.pm_kappa, d.self_adjoint_err) for d in out]
    return res, ms.value


def dielectric(geom: Geometry2D, nx: int, ny: int, smoothing: bool = True):
    """Dielectric2D::from_geometry (dielectric.rs:146-210): (eps, inv_eps, tensor[N,4])."""
    g = geom.to_c()
    n = nx * ny
    eps = np.zeros(n)
    inv = np.zeros(n)
    ten = np.zeros((n, 4))
    L.check(L.lib().bh_dielectric(ctypes.byref(g), nx, ny, int(smoothing), eps.ctypes.data_as(L.c_dp),
                                  inv.ctypes.data_as(L.c_dp), ten.ctypes.data_as(L.c_dp)), "bh_dielectric")
    return eps, inv, ten
