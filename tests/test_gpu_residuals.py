"""Residual B-norm diagnostics (eigensolver.rs:744-779) against the oracle.

The reference's solve() computes ||r_i||_B / max(|lambda_i|, 1e-10) for the
block's bands on its log iterations (1-5, 10, 20, 50, 100, then every 50;
eigensolver.rs:251-259, 1729-1737), after the residuals are deflated.  The
engine reports, per k-point, those of the last log iteration of each solve
(the exit iteration included when it is one), and that iteration.

The engine forms each iteration's residuals for the bands active before that
iteration's soft-locking decision and deflates them against the bands locked
before it, as the reference does (only the search directions drop the newly
locked bands), so its iterates follow the oracle's to rounding: the
iteration index must match and the relative residuals agree to 1e-10
(absolute; they are O(1e-4..1)) on the solves from the deterministic start
(the first k-point of a path, one-k-point jobs), and to 1e-9 on warm-started
ones, whose start carries the previous solves' rounding (up to 1e-10 seen).
"""
import numpy as np
import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd import bandstructure as B
from blaze2d_amd.sweep import res64_sweep
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


def oracle_job(j, rule):
    lat = "square" if abs(j.geom.lattice.a2[0]) < 1e-12 else "hex"
    return O.Job(lattice=lat, eps_bg=j.geom.eps_bg,
                 atoms=[(a.pos[0], a.pos[1], a.radius, a.eps_inside) for a in j.geom.atoms],
                 nx=j.nx, ny=j.ny, pol=j.pol, n_bands=j.n_bands, tol=j.tol, kpath=j.k_path, exit_rule=rule)


def check(res, ref, label, cold):
    """cold: bool mask of the k-points whose solves start from the deterministic start."""
    om, it, cv, tot, rel, rit = ref
    assert np.array_equal(res.iterations, it), (label, res.iterations, it)
    assert np.array_equal(res.residual_iteration, rit), label
    # every solve reports a log iteration: the first iteration always is one
    assert (res.residual_iteration >= 1).all() and (res.residual_iteration <= res.iterations).all()
    d = np.abs(res.relative_residuals - rel)
    warm = ~cold
    print(f"{label}: cold-start k-points {cold.sum()}, max abs diff {d[cold].max():.2e} (values up to "
          f"{rel[cold].max():.2e}); warm-started {warm.sum()}, max abs diff "
          f"{d[warm].max() if warm.any() else 0.0:.2e}")
    assert d[cold].max() <= 1e-10, label
    assert d.max() <= 1e-9, label


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_residual_norms_cfg1(be, pol):
    # the whole 31-point path: the first Gamma solve (and its copy closing the
    # path) from the deterministic start, the rest warm-started
    kp = B.k_path("square", 10)
    job = B.BandStructureJob(B.Geometry2D(B.Lattice2D.square(), 13.0, [B.BasisAtom((0.5, 0.5), 0.3, 1.0)]),
                             24, 24, pol, kp, n_bands=8, tol=1e-7)
    (res,), _ = B.run_batch(be, [job], slots=1)
    cold = np.zeros(kp.shape[0], dtype=bool)
    cold[[0, -1]] = True
    check(res, O.bandstructure_residuals(oracle_job(job, 0)), f"cfg1 {pol}", cold)


@pytest.mark.parametrize("rule", [0, 1])
def test_residual_norms_cold_start_k_points(be, rule):
    # one-k-point jobs (every solve from the deterministic start) on the first
    # two sweep geometries, TM and TE, at points along their paths
    base = res64_sweep(2)
    jobs = []
    for j in base:
        for k in (3, 10, 17, 24):
            jobs.append(B.BandStructureJob(j.geom, j.nx, j.ny, j.pol, j.k_path[k:k + 1], n_bands=j.n_bands,
                                           tol=j.tol))
    res, st = B.run_batch(be, jobs, B.RunOptions(exit_rule=rule), slots=len(jobs))
    for j, r in zip(jobs, res):
        check(r, O.bandstructure_residuals(oracle_job(j, rule)),
              f"{j.pol} k=({j.k_path[0][0]:.3f},{j.k_path[0][1]:.3f}) rule {rule}", np.ones(1, dtype=bool))
