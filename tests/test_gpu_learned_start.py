"""The learned-start plumbing: plane-wave projections out, warm vectors in.

* Projections: in a uniform medium every eigenvector of Theta is a plane wave
  exp(i (k+G) r) (the fields are its periodic part exp(i G r)); at a k-point
  without ties in |k+G| the n-th eigenvector projects onto the n-th nearest G
  alone (to its convergence error), with magnitude 1/sqrt(eps N)
  (B-normalised, B = eps for TM, 1 for TE).
* First-iteration norms, closed form: one warm vector, a block of one, one
  iteration.  The block is x / ||x||_B, theta = x^H A x / x^H B x, and the
  engine's relative residual of log iteration 1 is ||A x - theta B x||_B /
  theta (eigensolver.rs:744-779) and its first-iteration preconditioned norm
  is ||M^-1 (A x - theta B x)||_B, with A, B and M^-1 applied by the oracle
  (1e-10 relative).
* Warm vectors: a second solve started from the first solve's projected
  eigenvectors (initialize_block with warm vectors, initialization.rs:248-309)
  converges to the same frequencies (1e-8, tol 1e-7 all bands converged) in
  fewer iterations than the default start.
"""
import numpy as np
import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd import bandstructure as B
from blaze2d_amd.sweep import sweep_jobs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_projection_of_plane_wave_eigenvectors(be, pol):
    n, h, eps, nb = 32, 3, 4.0, 6
    k = (0.137, 0.061)  # generic: no ties among the lowest |k+G|
    lat = B.Lattice2D.square()
    job = B.BandStructureJob(B.Geometry2D(lat, eps, []), n, n, pol, np.array([k]), n_bands=nb, tol=1e-10)
    (res,), _ = B.run_batch(be, [job], B.RunOptions(exit_rule=1), slots=1, pw_half=h)
    b1, b2 = lat.reciprocal()
    gs = [(n1, n2) for n2 in range(-h, h + 1) for n1 in range(-h, h + 1)]
    kc = np.array([k[0] * b1[0] + k[1] * b2[0], k[0] * b1[1] + k[1] * b2[1]])
    q2 = [np.sum((kc + n1 * np.array(b1) + n2 * np.array(b2)) ** 2) for n1, n2 in gs]
    order = np.argsort(q2)
    amp = 1.0 / np.sqrt((eps if pol == "TM" else 1.0) * n * n)
    c = res.pw_coefs[0]  # [band, g]
    for b in range(nb):
        g = order[b]
        assert abs(abs(c[b, g]) - amp) <= 1e-9 * amp, (b, abs(c[b, g]), amp)
        # the other plane waves carry only the eigenvector's convergence error
        # (solved to a 1e-10 eigenvalue change: ~1e-8 of the amplitude)
        rest = np.delete(np.abs(c[b]), g)
        assert rest.max() <= 1e-6 * amp
        assert abs(res.bands[0, b] - np.sqrt(q2[g] / eps)) <= 1e-9 * np.sqrt(q2[g] / eps)


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_warm_start_from_projections(be, pol):
    # one-k-point jobs of two sweep geometries: solve from the default start,
    # project the converged eigenvectors on |n1|, |n2| <= 8, then solve again
    # started from those vectors
    base = sweep_jobs(2, seed=3, grid=64, n_bands=12, tol=1e-7, pols=(pol,))
    jobs = [B.BandStructureJob(j.geom, 64, 64, pol, j.k_path[i:i + 1], n_bands=12, tol=1e-7)
            for j in base for i in (4, 13)]
    cold, _ = B.run_batch(be, jobs, B.RunOptions(exit_rule=1), slots=len(jobs), pw_half=8)
    warm_jobs = [B.BandStructureJob(j.geom, 64, 64, pol, j.k_path, n_bands=12, tol=1e-7, warm_pw=r.pw_coefs[0],
                                    pw_half=8) for j, r in zip(jobs, cold)]
    warm, _ = B.run_batch(be, warm_jobs, B.RunOptions(exit_rule=1), slots=len(jobs))
    for j, c, w in zip(jobs, cold, warm):
        rel = np.abs(np.sort(w.bands[0]) - np.sort(c.bands[0])) / np.sort(c.bands[0])
        print(f"{pol} k={j.k_path[0]}: iterations default start {c.iterations[0]}, warm {w.iterations[0]}; "
              f"max rel omega diff {rel.max():.2e}")
        assert c.converged.all() and w.converged.all()
        assert rel.max() <= 1e-6
        assert w.iterations[0] < c.iterations[0]


def _synth(coefs, h, nx, ny):
    iy, ix = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    x = np.zeros((ny, nx), complex)
    for n2 in range(-h, h + 1):
        for n1 in range(-h, h + 1):
            x += coefs[(n2 + h) * (2 * h + 1) + n1 + h] * np.exp(2j * np.pi * (n1 * ix / nx + n2 * iy / ny))
    return x.ravel()


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_first_iteration_norms_closed_form(be, pol):
    r = np.random.default_rng(11)
    h, n, k = 2, 16, (0.21, 0.07)
    coefs = (r.standard_normal(25) + 1j * r.standard_normal(25)) / (1 + np.arange(25))
    geom = B.Geometry2D(B.Lattice2D.square(), 9.0, [B.BasisAtom((0.5, 0.5), 0.3, 1.5)])
    job = B.BandStructureJob(geom, n, n, pol, np.array([k]), n_bands=1, tol=1e-7, max_iter=1, block_size=1,
                             warm_pw=coefs[None, :], pw_half=h)
    (res,), _ = B.run_batch(be, [job], B.RunOptions(exit_rule=1), slots=1)
    oj = O.Job(lattice="square", eps_bg=9.0, atoms=[(0.5, 0.5, 0.3, 1.5)], nx=n, ny=n, pol=pol, n_bands=1,
               block_size=1, max_iter=1, tol=1e-7, kpath=np.array([k]))
    x = _synth(coefs, h, n, n)[None, :]
    ax = O.operator_apply(oj, k, "A", x)[0]
    bx = O.operator_apply(oj, k, "B", x)[0]
    x = x[0]
    nb = np.vdot(x, bx).real
    theta = np.vdot(x, ax).real / nb
    rv = (ax - theta * bx) / np.sqrt(nb)
    brv = O.operator_apply(oj, k, "B", rv[None, :])[0]
    rel = np.sqrt(np.vdot(rv, brv).real) / abs(theta)
    pr = O.operator_apply(oj, k, "P", rv[None, :])[0]
    bpr = O.operator_apply(oj, k, "B", pr[None, :])[0]
    prec = np.sqrt(np.vdot(pr, bpr).real)
    print(f"{pol}: relative residual {res.relative_residuals[0, 0]:.12e} (closed form {rel:.12e}), "
          f"preconditioned {res.first_prec_residuals[0, 0]:.12e} (closed form {prec:.12e})")
    assert res.iterations[0] == 1 and res.residual_iteration[0] == 1
    assert abs(res.relative_residuals[0, 0] - rel) <= 1e-10 * rel
    assert abs(res.first_prec_residuals[0, 0] - prec) <= 1e-10 * prec
