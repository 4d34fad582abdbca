"""GPU parity on every grid / band-count shape of the synthetic sweep.

The batched-sweep workload draws grids from {32, 48, 64} and band counts from
{8, 12, 16}.  Each shape below runs seeded sweep geometries (TM and TE) through
the batched engine under both exit rules and compares every k-point's sorted
frequencies with the FP64 oracle run under the same rule, <= 1e-6 relative.
The 16-band shapes exercise the 48-column subspace kernels (dual Gram with 16
column blocks, Rayleigh-Ritz of order up to 48, 32-output Ritz update); the
48 x 48 shapes exercise the radix-3 FFT passes.

Under exit rule 1 every band has met the tolerance and every band must agree.
Under exit rule 0 (the reference's solve() accounting) a solve may end while
its top bands have not converged, and where such a band stops depends on the
search directions of its last iterations: on the 48 x 48 16-band TM case the
oracle's own top band sits 1.1e-4 from its converged value.  So under rule 0
a band is held to 1e-6 where the oracle's value is within 1e-7 of a tightly
converged oracle run (tol 1e-11, all bands converged), and otherwise to a
tenth of the oracle's own distance from that converged value.

The oracle runs (about 200 s of CPU in total) go through a thread pool: the
ctypes calls release the GIL, so they run in parallel on the host cores.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd import bandstructure as B
from blaze2d_amd.sweep import sweep_jobs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL_OMEGA = 1e-6

# (grid, bands, seed): one sweep geometry per shape, in both polarisations
SHAPES = [(32, 8, 47), (48, 12, 67), (48, 16, 71), (64, 16, 87)]
RULES = (0, 1)


TIGHT = "tight"  # oracle at tol 1e-11 with every band converged


def oracle_job(j, rule):
    lat = "square" if abs(j.geom.lattice.a2[0]) < 1e-12 else "hex"
    tight = rule == TIGHT
    return O.Job(lattice=lat, eps_bg=j.geom.eps_bg,
                 atoms=[(a.pos[0], a.pos[1], a.radius, a.eps_inside) for a in j.geom.atoms],
                 nx=j.nx, ny=j.ny, pol=j.pol, n_bands=j.n_bands, tol=1e-11 if tight else j.tol, kpath=j.k_path,
                 exit_rule=1 if tight else rule)


def rel_diff(a, b):
    """Per-entry relative difference of sorted frequencies (absolute below 1e-8)."""
    a, b = np.sort(a, axis=1), np.sort(b, axis=1)
    m = b > 1e-8
    e = np.abs(a - b)
    e[m] /= b[m]
    return e


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


@pytest.fixture(scope="module")
def cases():
    """{(grid, bands, rule): (jobs, oracle results)}, the oracle run in parallel."""
    jobs = {(g, nb): sweep_jobs(1, seed=seed, grid=g, n_bands=nb, tol=1e-7) for g, nb, seed in SHAPES}
    work = [(key, rule, i) for key in jobs for rule in RULES + (TIGHT,) for i in range(len(jobs[key]))]
    threads = max(1, min(16, os.cpu_count() or 1, len(work)))
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        refs = list(ex.map(lambda w: O.bandstructure(oracle_job(jobs[w[0]][w[2]], w[1])), work))
    out = {}
    for (key, rule, i), r in zip(work, refs):
        out.setdefault((key[0], key[1], rule), (jobs[key], {}))[1][i] = r
    return out


def compare(res, ref, tight, rule, label):
    om, it, cv, tot = ref
    assert cv.all(), label
    assert res.converged.all(), label
    rel = rel_diff(res.bands, om)
    if rule == 1:
        bound = np.full_like(rel, RTOL_OMEGA)
    else:
        own = rel_diff(om, tight[0])  # the oracle's distance from its converged value
        bound = np.where(own <= 1e-7, RTOL_OMEGA, np.maximum(RTOL_OMEGA, 0.1 * own))
    k, b = np.unravel_index(np.argmax(rel - bound), rel.shape)
    print(f"{label}: max rel omega diff {rel.max():.3e}; tightest margin at k {k} band {b}: {rel[k, b]:.3e} "
          f"<= {bound[k, b]:.3e}; iterations gpu {res.iterations.sum()} oracle {it.sum()}")
    assert (rel <= bound).all(), (label, k, b, rel[k, b], bound[k, b])


@pytest.mark.parametrize("rule", RULES)
@pytest.mark.parametrize("grid,bands,seed", SHAPES)
def test_sweep_shape_matches_oracle(be, cases, grid, bands, seed, rule):
    jobs, refs = cases[(grid, bands, rule)]
    tight = cases[(grid, bands, TIGHT)][1]
    res, st = B.run_batch(be, jobs, B.RunOptions(exit_rule=rule), slots=len(jobs))
    assert st.solves == sum(j.k_path.shape[0] - 1 for j in jobs)
    print(f"{grid}x{grid} {bands} bands rule {rule}: {st.rr_problems} Rayleigh-Ritz problems, "
          f"{st.svqb_fallbacks} on the SVQB path, {st.rank_drops} with a dropped direction")
    for i, (j, r) in enumerate(zip(jobs, res)):
        compare(r, refs[i], tight[i], rule, f"{grid}x{grid} {bands} bands {j.pol} exit rule {rule}")
