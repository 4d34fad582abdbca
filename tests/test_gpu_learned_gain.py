"""What the trained network gains: one-k-point solves from the predicted start.

A network (64 x 64 input, 12 vectors on the 81 plane waves |n1|, |n2| <= 4)
is trained on the GPU on the solver's own converged subspaces of 800 sweep
geometries (3 k-points per geometry and polarisation, 4800 samples), then
used on 48 held-out geometries (TM and TE, one k-point each):
  * every solve converges and the mean LOBPCG iteration count drops by at
    least 10% against the default start;
  * frequencies agree with a tight (tol 1e-10) default-start solve to 1e-6
    relative;
  * four held-out solves (two TM, two TE) follow the FP64 oracle started from
    the same predicted coefficients (oracle_bandstructure_warm): the same
    iteration count and frequencies to 1e-9, and lie within 1e-6 of the
    oracle's tight default-start solve;
  * the fallback path: with margin 0 every solve falls back and the results
    are those of the default start (same iterations, frequencies to 1e-12);
  * the large grid: a 32-vector network trained at 64 x 64 (600 geometries,
    32 bands) starts the 256 x 256, 32-band solves of the large combined case
    (eps evaluated on the network's 64 x 64 grid, the prediction synthesised
    on 256 x 256): every solve converges within 1e-6 of the oracle fixture
    (tests/golden/cfg5_oracle.json) in fewer iterations in total.
"""
import json
import os
import dataclasses

import numpy as np
import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd import learned as LS
from blaze2d_amd.bandstructure import BandStructureJob, RunOptions, run_batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu

OPTS = RunOptions(exit_rule=1)


@pytest.fixture(scope="module")
def trained():
    be = HipBackend(0)
    net = LS.CnnNet(be, 64, 64, 12, 4, max_batch=128, seed=2)
    data = LS.training_set(be, 800, seed=321, k_per_path=3)
    hist = LS.train(net, data, 12, batch=64, lr=1e-3, seed=1)
    print(f"trained on {data.embed.shape[0]} samples: loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    held = LS.held_out_jobs(48, seed=777)
    yield be, net, held
    net.close()
    be.close()


def _rel(a, b):
    a, b = np.sort(a), np.sort(b)
    m = b > 1e-3  # the Gamma zero mode has no relative error
    return float(np.max(np.abs(a[m] - b[m]) / b[m]))


def test_learned_start_reduces_iterations(trained):
    be, net, held = trained
    base, _ = run_batch(be, held, OPTS, slots=len(held))
    ls = LS.solve(be, net, held, OPTS, slots=len(held))
    it0 = np.array([r.iterations[0] for r in base])
    it1 = np.array([r.iterations[0] for r in ls.results])
    print(f"{len(held)} held-out solves: iterations default {it0.mean():.2f}, learned {it1.mean():.2f}; "
          f"fallback rate {1 - ls.used_prediction.mean():.3f}; iterations spent incl. probes "
          f"{ls.iterations_spent / len(held):.2f}")
    assert all(r.converged.all() for r in base) and all(r.converged.all() for r in ls.results)
    assert it1.mean() <= 0.9 * it0.mean()
    tight = [dataclasses.replace(j, tol=1e-10) for j in held]
    ref, _ = run_batch(be, tight, OPTS, slots=len(held))
    err = max(_rel(r.bands[0], t.bands[0]) for r, t in zip(ls.results, ref))
    err0 = max(_rel(r.bands[0], t.bands[0]) for r, t in zip(base, ref))
    print(f"max rel omega error vs tol 1e-10: learned start {err:.2e}, default start {err0:.2e}")
    assert err <= 1e-6


def _oracle_job(j, tol):
    lat = "square" if abs(j.geom.lattice.a2[0]) < 1e-12 else "hex"
    return O.Job(lattice=lat, eps_bg=j.geom.eps_bg,
                 atoms=[(a.pos[0], a.pos[1], a.radius, a.eps_inside) for a in j.geom.atoms],
                 nx=j.nx, ny=j.ny, pol=j.pol, n_bands=j.n_bands, tol=tol, kpath=j.k_path, exit_rule=1)


def test_learned_start_matches_oracle_from_the_same_start(trained):
    # the oracle started from the same predicted coefficients follows the same
    # LOBPCG trajectory: same iteration count, frequencies to 1e-9; and both
    # are within 1e-6 of the oracle's tight (tol 1e-10) default-start solve
    be, net, held = trained
    jobs = [j for j in held if j.pol == "TM"][:2] + [j for j in held if j.pol == "TE"][:2]
    coefs = net.predict_jobs(jobs)
    ls = LS.solve(be, net, jobs, OPTS, slots=len(jobs))
    assert ls.used_prediction.all()
    for j, c, r in zip(jobs, coefs, ls.results):
        om, it, cv, _ = O.bandstructure_warm(_oracle_job(j, j.tol), c, net.pw_half)
        omt, itt, cvt, _ = O.bandstructure(_oracle_job(j, 1e-10))
        same = _rel(r.bands[0], om[0])
        tight = _rel(r.bands[0], omt[0])
        print(f"{j.pol} k={j.k_path[0]}: learned start {r.iterations[0]} iterations, oracle from the same start "
              f"{it[0]}; max rel omega diff {same:.2e}; vs tight oracle {tight:.2e}")
        assert cv.all() and cvt.all() and r.converged.all()
        assert r.iterations[0] == it[0]
        assert same <= 1e-9
        assert tight <= 1e-6


def test_fallback_gives_default_start_results(trained):
    be, net, held = trained
    jobs = held[:16]
    base, _ = run_batch(be, jobs, OPTS, slots=len(jobs))
    ls = LS.solve(be, net, jobs, OPTS, slots=len(jobs), margin=0.0)
    assert not ls.used_prediction.any()
    for r, b in zip(ls.results, base):
        assert r.iterations[0] == b.iterations[0]
        assert np.abs(r.bands - b.bands).max() <= 1e-12 * np.abs(b.bands).max()
    # the decision itself
    assert list(LS.keep_prediction([1.0, 2.0, np.nan], [1.5, 1.5, 1.0])) == [True, False, False]


def test_learned_start_on_the_large_grid():
    be = HipBackend(0)
    net = LS.CnnNet(be, 64, 64, 32, 4, max_batch=128, seed=3)
    try:
        data = LS.training_set(be, 600, seed=4321, n_bands=32, k_per_path=3)
        hist = LS.train(net, data, 12, batch=64, lr=1e-3, seed=1)
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg5_oracle.json")) as f:
            fx = json.load(f)
        geom = _large_geometry(fx)
        cases = [c for c in fx["cases"] if c["grid"] == 256]
        jobs = [BandStructureJob(geom, 256, 256, c["pol"], np.array([c["k"]]), n_bands=32, tol=fx["tol"],
                                 max_iter=fx["max_iter"]) for c in cases]
        ls = LS.solve(be, net, jobs, OPTS, slots=len(jobs))
        tot_l = tot_d = 0
        for c, r in zip(cases, ls.results):
            err = _rel(r.bands[0], np.array(c["omegas"]))
            print(f"256x256 32 bands {c['pol']} k={c['k']}: learned start {r.iterations[0]} iterations "
                  f"(default start {c['iterations']}), max rel omega diff vs oracle {err:.2e}")
            assert r.converged.all()
            assert err <= 1e-6
            tot_l += int(r.iterations[0])
            tot_d += int(c["iterations"])
        print(f"trained on {data.embed.shape[0]} samples (loss {hist[0]:.4f} -> {hist[-1]:.4f}); iterations "
              f"{tot_l} learned vs {tot_d} default; fallback rate {1 - ls.used_prediction.mean():.2f}")
        assert tot_l < tot_d
    finally:
        net.close()
        be.close()


def _large_geometry(fx):
    from blaze2d_amd import bandstructure as B
    atoms = [B.BasisAtom((a[0], a[1]), a[2], a[3]) for a in fx["atoms"]]
    return B.Geometry2D(B.Lattice2D.square(), fx["eps_bg"], atoms)
