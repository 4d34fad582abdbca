"""Large and rectangular grids: the multi-pass FFT operators and the engine on them.

Grids beyond one workgroup's LDS (256 x 256, 512 x 512) and rectangular grids
run their FFTs as column passes and row passes through HBM.  Checked here:
  * plain 2D FFTs against numpy (forward unnormalised, inverse scaled 1/N);
  * Theta (TM, TE), mass and preconditioners (FD, TP) against the FP64
    oracle on random fields, 1e-11 relative;
  * band structures on a rectangular grid against the live oracle, and the
    512 x 512 TM 16-band solve at M of the large combined case against the
    oracle's frequencies in tests/golden/cfg5_oracle.json (made by
    tests/golden/make_cfg5_oracle.py), 1e-6 relative, every band converged.
"""
import json
import os

import numpy as np
import pytest

from blaze2d_amd import HipBackend, ThetaOperator
from blaze2d_amd import bandstructure as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
RTOL_OP = 1e-11


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


def rnd(shape, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


@pytest.mark.parametrize("nx,ny", [(96, 96), (128, 64), (48, 32), (24, 48), (256, 256), (512, 512)])
def test_fft2d_multipass_matches_numpy(be, nx, ny):
    cnt = 2
    a = rnd((cnt, ny, nx), nx + ny)
    f = be.alloc_field(nx, ny, cnt)
    f.upload(a.reshape(cnt, -1))
    be.forward_fft_2d(f)
    got = f.download().reshape(cnt, ny, nx)
    want = np.fft.fft2(a)
    assert np.abs(got - want).max() <= RTOL_OP * np.abs(want).max()
    be.inverse_fft_2d(f)
    back = f.download().reshape(cnt, ny, nx)
    assert np.abs(back - a).max() <= RTOL_OP * np.abs(a).max()


CASES = [
    ("square", 128, 64, "TM", [(0.5, 0.5, 0.3, 1.0)], 13.0, (0.2, 0.1)),
    ("square", 96, 96, "TE", [(0.5, 0.5, 0.3, 1.0)], 13.0, (0.0, 0.0)),
    ("hex", 48, 32, "TE", [(0.5, 0.5, 0.3, 1.0)], 13.0, (0.5, 0.0)),
    ("square", 256, 256, "TE", [(0.0, 0.0, 0.2, 1.0), (0.5, 0.5, 0.12, 1.0)], 12.0, (0.5, 0.5)),
    ("square", 256, 256, "TM", [(0.0, 0.0, 0.2, 1.0), (0.5, 0.5, 0.12, 1.0)], 12.0, (0.0, 0.0)),
    ("hex", 192, 384, "TE", [(0.5, 0.5, 0.25, 2.0), (0.1, 0.9, 0.12, 1.0)], 8.0, (1 / 3, 1 / 3)),
    ("square", 512, 512, "TM", [(0.0, 0.0, 0.2, 1.0), (0.5, 0.5, 0.12, 1.0)], 12.0, (0.5, 0.5)),
]


@pytest.mark.parametrize("lattice,nx,ny,pol,atoms,eps_bg,k", CASES)
def test_operators_multipass_match_oracle(be, lattice, nx, ny, pol, atoms, eps_bg, k):
    lat = B.Lattice2D.square() if lattice == "square" else B.Lattice2D.hexagonal()
    g = B.Geometry2D(lat, eps_bg, [B.BasisAtom((a[0], a[1]), a[2], a[3]) for a in atoms])
    eps, inv, ten = B.dielectric(g, nx, ny)
    b1, b2 = lat.reciprocal()
    bloch = (k[0] * b1[0] + k[1] * b2[0], k[0] * b1[1] + k[1] * b2[1])
    op = ThetaOperator(be, nx, ny, pol, b1, b2, bloch, eps, ten)
    cnt = 3
    x = rnd((cnt, nx * ny), nx + ny + len(atoms))
    fin, fout = be.alloc_field(nx, ny, cnt), be.alloc_field(nx, ny, cnt)
    fin.upload(x)
    job = O.Job(lattice=lattice, eps_bg=eps_bg, atoms=atoms, nx=nx, ny=ny, pol=pol)
    op.apply(fin, fout)
    want = O.operator_apply(job, k, "A", x)
    assert np.abs(fout.download() - want).max() <= RTOL_OP * np.abs(want).max()
    op.apply_mass(fin, fout)
    want = O.operator_apply(job, k, "B", x)
    assert np.abs(fout.download() - want).max() <= RTOL_OP * np.abs(want).max()
    fin.upload(x)
    op.precondition(fin)
    want = O.operator_apply(job, k, "P", x)
    assert np.abs(fin.download() - want).max() <= RTOL_OP * np.abs(want).max()


def _rel(got, ref):
    got, ref = np.sort(got, axis=1), np.sort(ref, axis=1)
    m = ref > 1e-8
    e = np.abs(got - ref)
    e[m] /= ref[m]
    return float(e.max())


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_bandstructure_rectangular_grid(be, pol):
    # square lattice sampled on a 96 x 64 grid, 8 bands, Gamma-X-M-Gamma at 10
    # points, all bands converged, against the oracle on the same grid
    kp = B.k_path("square", 3)
    atoms = [(0.5, 0.5, 0.3, 1.0)]
    job = B.BandStructureJob(B.Geometry2D(B.Lattice2D.square(), 13.0, [B.BasisAtom((0.5, 0.5), 0.3, 1.0)]),
                             96, 64, pol, kp, n_bands=8, tol=1e-7)
    (res,), st = B.run_batch(be, [job], B.RunOptions(exit_rule=1), slots=1)
    om, it, cv, tot = O.bandstructure(O.Job(lattice="square", eps_bg=13.0, atoms=atoms, nx=96, ny=64, pol=pol,
                                            n_bands=8, tol=1e-7, kpath=kp, exit_rule=1))
    err = _rel(res.bands, om)
    print(f"96x64 {pol}: max rel omega diff {err:.3e}, iterations gpu {res.iterations.sum()} oracle {it.sum()}")
    assert res.converged.all() and cv.all()
    assert err <= 1e-6


def cfg5_cases(grid, n_bands):
    with open(os.path.join(HERE, "golden", "cfg5_oracle.json")) as f:
        fx = json.load(f)
    atoms = [B.BasisAtom((a[0], a[1]), a[2], a[3]) for a in fx["atoms"]]
    geom = B.Geometry2D(B.Lattice2D.square(), fx["eps_bg"], atoms)
    out = []
    for c in fx["cases"]:
        if c["grid"] == grid and c["n_bands"] == n_bands:
            job = B.BandStructureJob(geom, grid, grid, c["pol"], np.array([c["k"]]), n_bands=n_bands, tol=fx["tol"],
                                     max_iter=fx["max_iter"])
            out.append((job, c))
    return out


def test_cfg5_512_tm_16_bands_at_m(be):
    cases = cfg5_cases(512, 16)
    assert len(cases) == 1
    (job, ref), = cases
    (res,), st = B.run_batch(be, [job], B.RunOptions(exit_rule=1), slots=1)
    err = _rel(res.bands, np.array([ref["omegas"]]))
    print(f"512x512 TM 16 bands at M: max rel omega diff {err:.3e}, iterations gpu {res.iterations[0]} "
          f"oracle {ref['iterations']}")
    assert res.converged.all()
    assert err <= 1e-6


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_bandstructure_32_bands(be, pol, rule):
    # 32 bands: a 96-column subspace (tiled Grams, wide Rayleigh-Ritz, the
    # row-chunked deflation against up to 33 vectors), on a 32 x 32 grid of the
    # large case's geometry along a 7-point path, against the live oracle
    kp = B.k_path("square", 2)
    atoms = [(0.0, 0.0, 0.2, 1.0), (0.5, 0.5, 0.12, 1.0)]
    geom = B.Geometry2D(B.Lattice2D.square(), 12.0, [B.BasisAtom((a[0], a[1]), a[2], a[3]) for a in atoms])
    job = B.BandStructureJob(geom, 32, 32, pol, kp, n_bands=32, tol=1e-7)
    (res,), st = B.run_batch(be, [job], B.RunOptions(exit_rule=rule), slots=1)
    om, it, cv, tot = O.bandstructure(O.Job(lattice="square", eps_bg=12.0, atoms=atoms, nx=32, ny=32, pol=pol,
                                            n_bands=32, tol=1e-7, kpath=kp, exit_rule=rule))
    err = _rel(res.bands, om)
    print(f"32x32 32 bands {pol} rule {rule}: max rel omega diff {err:.3e}, iterations gpu "
          f"{res.iterations.sum()} oracle {it.sum()}; {st.rr_problems} Rayleigh-Ritz problems, "
          f"{st.svqb_fallbacks} on the SVQB path")
    assert res.converged.all() and cv.all()
    assert err <= 1e-6


def test_cfg5_256_32_bands_tm_te(be):
    # the large combined case at 256 x 256, 32 bands, TM and TE at three
    # k-points, batched over k-points in one run, against the oracle fixture
    cases = cfg5_cases(256, 32)
    assert len(cases) == 6
    jobs = [c[0] for c in cases]
    res, st = B.run_batch(be, jobs, B.RunOptions(exit_rule=1), slots=len(jobs))
    for (job, ref), r in zip(cases, res):
        err = _rel(r.bands, np.array([ref["omegas"]]))
        print(f"256x256 32 bands {ref['pol']} k={ref['k']}: max rel omega diff {err:.3e}, iterations gpu "
              f"{r.iterations[0]} oracle {ref['iterations']}")
        assert r.converged.all()
        assert err <= 1e-6
