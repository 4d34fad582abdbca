"""The learned start: a CNN that predicts the initial LOBPCG block of a solve.

The network (bh_cnn_*, include/blaze2d_hip.h; kernels in csrc/cnn.hip) reads
the eps grid (eps / 16 and 1 / eps, sampled on the solver's own grid or
resampled to the network's) and an embedding of (k1, k2, polarisation,
lattice), and predicts n_vectors fields as coefficients on the plane waves
|n1|, |n2| <= pw_half.  The solver takes them as warm vectors (bh_job
warm_pw: the warm-start ingestion of initialize_block,
crates/core/src/eigensolver/initialization.rs:248-309), so a predicted start
only changes where LOBPCG begins: convergence is still judged by the unchanged
residual tolerance.

Training data are the solver's own converged subspaces: run_batch(...,
pw_half=h) returns every solve's eigenvectors projected on those plane waves.
The loss, 1 - tr(P_Y P_U) / n_vectors with P the projector on the predicted
(Y) or converged (U) vectors, does not depend on the basis either block is
written in (phases, rotations inside degenerate bands).

Fallback: ``solve`` runs every solve from its predicted start and, in the
same batch, one iteration of each from the default start, and compares the
two starts' first-iteration preconditioned residuals ||M^-1 r_i||_B (the norm
over the bands).  Solves whose predicted start is the worse one are solved
again from the default start, which gives their results.  The preconditioned
residual is the measure because the predicted fields carry no plane waves
beyond |n| <= pw_half: their plain residual A x - lambda B x keeps the
high-|G| part of B x (the jumps of eps) whatever the network predicts, while
the preconditioner, which scales it by ~1 / |k + G|^2, measures what LOBPCG has
left to correct.
"""
from __future__ import annotations

import ctypes
import dataclasses
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from .backend import HipBackend
from .bandstructure import BandStructureJob, RunOptions, dielectric_batch, run_batch
from .sweep import sweep_jobs

CH = 32
EMB_IN = 4

c_fp = ctypes.POINTER(ctypes.c_float)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(c_fp)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(L.c_dp)


class CnnNet:
    """One network on one device (bh_cnn_create)."""

    def __init__(self, backend: HipBackend, height: int = 64, width: int = 64, n_vectors: int = 12,
                 pw_half: int = 4, max_batch: int = 64, seed: int = 0):
        self.backend = backend
        self.height, self.width = height, width
        self.n_vectors, self.pw_half, self.max_batch = n_vectors, pw_half, max_batch
        self.ng = (2 * pw_half + 1) ** 2
        h = ctypes.c_void_p()
        L.check(L.lib().bh_cnn_create(backend.handle, height, width, n_vectors, pw_half, max_batch, seed,
                                      ctypes.byref(h)), "bh_cnn_create")
        self.handle = h
        self.n_params = int(L.lib().bh_cnn_param_count(h))

    def close(self) -> None:
        if self.handle:
            L.lib().bh_cnn_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        L.check(L.lib().bh_cnn_get_params(self.handle, _fp(out)), "bh_cnn_get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        L.check(L.lib().bh_cnn_set_params(self.handle, _fp(p)), "bh_cnn_set_params")

    def grads(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        L.check(L.lib().bh_cnn_get_grads(self.handle, _fp(out)), "bh_cnn_get_grads")
        return out

    def _inputs(self, grid, embed):
        g = np.ascontiguousarray(grid, dtype=np.float32)
        e = np.ascontiguousarray(embed, dtype=np.float32)
        if g.ndim != 4 or g.shape[1:] != (2, self.height, self.width):
            raise ValueError(f"grid must be [B, 2, {self.height}, {self.width}]")
        if e.shape != (g.shape[0], EMB_IN):
            raise ValueError("embed must be [B, 4]")
        return g, e

    def predict(self, grid: np.ndarray, embed: np.ndarray) -> np.ndarray:
        """[B, n_vectors, ng] complex coefficients (in chunks of max_batch)."""
        g, e = self._inputs(grid, embed)
        B = g.shape[0]
        out = np.zeros((B, self.n_vectors, self.ng), dtype=np.complex128)
        for s in range(0, B, self.max_batch):
            n = min(self.max_batch, B - s)
            o = np.zeros((n, self.n_vectors, self.ng), dtype=np.complex128)
            L.check(L.lib().bh_cnn_predict(self.handle, n, _fp(g[s:s + n]), _fp(e[s:s + n]), _dp(o.view(np.float64))),
                    "bh_cnn_predict")
            out[s:s + n] = o
        return out

    def predict_jobs(self, jobs: Sequence[BandStructureJob], k_index: int = 0) -> np.ndarray:
        """Predictions for the jobs' k-point k_index with the grid input built
        on the device: each geometry's smoothed eps evaluated on the network's
        own grid (for a job on that grid, the solver's eps)."""
        embed = job_embedding(jobs, k_index)
        B = len(jobs)
        out = np.zeros((B, self.n_vectors, self.ng), dtype=np.complex128)
        for s in range(0, B, self.max_batch):
            n = min(self.max_batch, B - s)
            geoms = (L.Geometry * n)(*[j.geom.to_c() for j in jobs[s:s + n]])
            o = np.zeros((n, self.n_vectors, self.ng), dtype=np.complex128)
            L.check(L.lib().bh_cnn_predict_geometries(self.handle, n, geoms, 1, _fp(embed[s:s + n]),
                                                      _dp(o.view(np.float64))), "bh_cnn_predict_geometries")
            out[s:s + n] = o
        return out

    def use_graphs(self, enable: bool) -> None:
        """Replay each training step's device work from a HIP graph (the
        default) or launch its kernels one by one; the results are identical."""
        L.check(L.lib().bh_cnn_set_graphs(self.handle, 1 if enable else 0), "bh_cnn_set_graphs")

    def train_step(self, grid: np.ndarray, embed: np.ndarray, targets: np.ndarray, lr: float) -> float:
        """Forward, projector loss, backward and (lr > 0) one Adam step; returns the batch loss."""
        g, e = self._inputs(grid, embed)
        t = np.ascontiguousarray(targets, dtype=np.complex128)
        if t.shape != (g.shape[0], self.n_vectors, self.ng):
            raise ValueError("targets must be [B, n_vectors, ng] complex")
        loss = ctypes.c_double(0.0)
        L.check(L.lib().bh_cnn_train_step(self.handle, g.shape[0], _fp(g), _fp(e), _dp(t.view(np.float64)), lr,
                                          ctypes.byref(loss)), "bh_cnn_train_step")
        return loss.value


def _resample(eps: np.ndarray, ny: int, nx: int, height: int, width: int) -> np.ndarray:
    """Nearest-node resampling of a periodic [ny, nx] grid to [height, width]."""
    if (ny, nx) == (height, width):
        return eps
    iy = (np.arange(height) * ny) // height
    ix = (np.arange(width) * nx) // width
    return eps[np.ix_(iy, ix)]


def job_embedding(jobs: Sequence[BandStructureJob], k_index: int = 0) -> np.ndarray:
    """[B, 4] float32: k1, k2 (fractional) of k-point k_index, polarisation
    (0 TM, 1 TE), lattice (0 square, 1 hexagonal)."""
    embed = np.zeros((len(jobs), EMB_IN), dtype=np.float32)
    for i, j in enumerate(jobs):
        k = j.k_path[k_index]
        hexagonal = abs(j.geom.lattice.a2[0]) > 1e-12
        embed[i] = (k[0], k[1], 0.0 if j.pol.upper() == "TM" else 1.0, 1.0 if hexagonal else 0.0)
    return embed


def network_inputs(backend: HipBackend, jobs: Sequence[BandStructureJob], height: int = 64, width: int = 64,
                   k_index: int = 0):
    """(grid [B, 2, height, width], embed [B, 4]) for the jobs' k-point k_index.

    The grid holds eps / 16 and 1 / eps of the smoothed dielectric on the job's
    own grid (resampled to the network's when they differ)."""
    B = len(jobs)
    grid = np.zeros((B, 2, height, width), dtype=np.float32)
    # one dielectric batch per grid size
    by_grid = {}
    for i, j in enumerate(jobs):
        by_grid.setdefault((j.nx, j.ny), []).append(i)
    for (nx, ny), idx in by_grid.items():
        eps, _, _, _ = dielectric_batch(backend, [jobs[i].geom for i in idx], nx, ny)
        for r, i in enumerate(idx):
            e = _resample(eps[r].reshape(ny, nx), ny, nx, height, width)
            grid[i, 0] = e / 16.0
            grid[i, 1] = 1.0 / e
    return grid, job_embedding(jobs, k_index)


@dataclass
class TrainingSet:
    grid: np.ndarray      # [G, 2, H, W] per band-structure job
    embed: np.ndarray     # [S, 4] per sample
    job: np.ndarray       # [S] index into grid
    targets: np.ndarray   # [S, n_vectors, ng] complex
    solve_seconds: float = 0.0


def training_set(backend: HipBackend, n_geometries: int, seed: int, grid: int = 64, n_bands: int = 12,
                 pw_half: int = 4, tol: float = 1e-7, k_per_path: Optional[int] = 3, segments_per_leg: int = 10,
                 slots: int = 256, k_seed: int = 7) -> TrainingSet:
    """Converged subspaces of the sweep's geometries (both polarisations)
    projected on the plane waves |n1|, |n2| <= pw_half.

    k_per_path: that many k-points per geometry and polarisation, drawn
    uniformly without repetition from the closed path (one-k-point solves from
    the default start); None: the whole path minus its copied closing Gamma
    (the warm-started chain)."""
    base = sweep_jobs(n_geometries, seed=seed, grid=grid, n_bands=n_bands, tol=tol,
                      segments_per_leg=segments_per_leg)
    if k_per_path is None:
        jobs = base
    else:
        rng = np.random.default_rng(k_seed)
        jobs = []
        for j in base:
            nk = j.k_path.shape[0] - 1
            for k in np.sort(rng.choice(nk, size=min(k_per_path, nk), replace=False)):
                jobs.append(BandStructureJob(j.geom, j.nx, j.ny, j.pol, j.k_path[k:k + 1], n_bands=j.n_bands,
                                             tol=j.tol))
    t0 = time.perf_counter()
    res, _ = run_batch(backend, jobs, RunOptions(exit_rule=1), slots=min(slots, len(jobs)), pw_half=pw_half)
    secs = time.perf_counter() - t0
    g, _ = network_inputs(backend, jobs, grid, grid)
    emb, jid, tg = [], [], []
    for i, (j, r) in enumerate(zip(jobs, res)):
        nk = j.k_path.shape[0] - 1 if k_per_path is None else 1
        hexagonal = abs(j.geom.lattice.a2[0]) > 1e-12
        pol = 0.0 if j.pol.upper() == "TM" else 1.0
        for k in range(nk):
            if not r.converged[k]:
                continue
            emb.append((j.k_path[k][0], j.k_path[k][1], pol, 1.0 if hexagonal else 0.0))
            jid.append(i)
            tg.append(r.pw_coefs[k])
    return TrainingSet(g, np.array(emb, dtype=np.float32), np.array(jid), np.array(tg), secs)


def train(net: CnnNet, data: TrainingSet, epochs: int, batch: int = 64, lr: float = 1e-3, seed: int = 0,
          min_lr: float = 5e-5, log=None) -> List[float]:
    """Adam over shuffled mini-batches with a cosine learning-rate decay;
    returns the mean loss of each epoch."""
    if batch > net.max_batch:
        raise ValueError("batch exceeds the network's max_batch")
    rng = np.random.default_rng(seed)
    S = data.embed.shape[0]
    steps = epochs * ((S + batch - 1) // batch)
    step = 0
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(S)
        tot = 0.0
        for s in range(0, S, batch):
            idx = perm[s:s + batch]
            cur = min_lr + 0.5 * (lr - min_lr) * (1 + np.cos(np.pi * step / max(1, steps)))
            tot += net.train_step(data.grid[data.job[idx]], data.embed[idx], data.targets[idx], cur) * len(idx)
            step += 1
        hist.append(tot / S)
        if log:
            log(f"epoch {ep + 1}/{epochs}: mean loss {hist[-1]:.4f}")
    return hist


@dataclass
class LearnedSolve:
    results: list                 # BandStructureResult per job
    used_prediction: np.ndarray   # bool per job: its result comes from the predicted start
    probe_default: np.ndarray     # ||M^-1 r||_B at the first iteration, default start
    probe_predicted: np.ndarray   # the same, predicted start
    iterations_spent: int         # every LOBPCG iteration run: probes and discarded solves included
    seconds: float                # wall time: inputs, inference, solves, probes and fallbacks
    inference_seconds: float


def keep_prediction(prec_predicted: np.ndarray, prec_default: np.ndarray, margin: float = 1.0) -> np.ndarray:
    """True where the predicted start is kept: its first-iteration
    preconditioned residual norm is not worse than margin times the default
    start's (margin 1: not worse; NaN compares as worse)."""
    pp = np.asarray(prec_predicted, dtype=float)
    pd = np.asarray(prec_default, dtype=float)
    return pp <= margin * pd


def solve(backend: HipBackend, net: CnnNet, jobs: Sequence[BandStructureJob], options: Optional[RunOptions] = None,
          slots: int = 512, groups: int = 2, fallback: bool = True, margin: float = 1.0,
          precision: int = 0) -> LearnedSolve:
    """Solves one-k-point jobs from the network's predicted start, with the
    default start as the fallback wherever the predicted block's first
    preconditioned residual is worse (see keep_prediction)."""
    opts = options or RunOptions()
    jobs = list(jobs)
    for j in jobs:
        if j.k_path.shape[0] != 1:
            raise ValueError("the learned start solves one-k-point jobs")
        if j.n_bands < net.n_vectors:
            raise ValueError("jobs need at least as many bands as the network predicts")
    t0 = time.perf_counter()
    coefs = net.predict_jobs(jobs)
    ti = time.perf_counter() - t0
    warm = [dataclasses.replace(j, warm_pw=c, pw_half=net.pw_half) for j, c in zip(jobs, coefs)]
    n = len(jobs)
    run = warm + ([dataclasses.replace(j, max_iter=1) for j in jobs] if fallback else [])
    res, _ = run_batch(backend, run, opts, slots=min(slots, len(run)), groups=groups, precision=precision)
    spent = int(sum(int(r.iterations.sum()) for r in res))
    pp = np.array([np.linalg.norm(r.first_prec_residuals[0]) for r in res[:n]])
    pd = np.full(n, np.inf)
    use = np.ones(n, dtype=bool)
    res = list(res)
    if fallback:
        pd = np.array([np.linalg.norm(r.first_prec_residuals[0]) for r in res[n:]])
        use = keep_prediction(pp, pd, margin)
        redo = [i for i in range(n) if not use[i]]
        if redo:
            again, _ = run_batch(backend, [jobs[i] for i in redo], opts, slots=min(slots, len(redo)), groups=groups,
                                 precision=precision)
            spent += int(sum(int(r.iterations.sum()) for r in again))
            for i, r in zip(redo, again):
                res[i] = r
    t1 = time.perf_counter()
    return LearnedSolve(res[:n], use, pd, pp, spent, t1 - t0, ti)


def mean_loss(net: CnnNet, grid: np.ndarray, embed: np.ndarray, targets: np.ndarray) -> float:
    """Mean projector loss of the network's predictions (the GPU loss kernel;
    no parameter update)."""
    tot = 0.0
    B = grid.shape[0]
    for s in range(0, B, net.max_batch):
        n = min(net.max_batch, B - s)
        tot += net.train_step(grid[s:s + n], embed[s:s + n], targets[s:s + n], 0.0) * n
    return tot / B


def held_out_jobs(n_geometries: int, seed: int, grid: int = 64, n_bands: int = 12, tol: float = 1e-7,
                  k_seed: int = 1) -> List[BandStructureJob]:
    """One-k-point jobs on held-out geometries: each geometry and polarisation
    at one k-point drawn uniformly from its closed path."""
    base = sweep_jobs(n_geometries, seed=seed, grid=grid, n_bands=n_bands, tol=tol)
    rng = np.random.default_rng(k_seed)
    out = []
    for j in base:
        k = int(rng.integers(0, j.k_path.shape[0] - 1))
        out.append(BandStructureJob(j.geom, j.nx, j.ny, j.pol, j.k_path[k:k + 1], n_bands=j.n_bands, tol=j.tol))
    return out
