"""GPU dielectric tables (bh_dielectric_batch) against the FP64 oracle, the
reference's exported eps grids and the host path, through the C ABI.

The per-node arithmetic is the oracle's (dielectric.rs:146-210,
analytic_geometry.rs:183-541) evaluated on the GPU, so the tables must agree
to rounding: 1e-13 absolute on eps (values up to 16) and on the inverse
tables."""
import json
import os

import numpy as np
import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd import bandstructure as B
from blaze2d_amd.sweep import res64_sweep
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ATOL = 1e-13
GOLD = os.path.join(os.path.dirname(__file__), "golden")

CASES = [
    ("square", 24, [(0.5, 0.5, 0.3, 1.0)], 13.0),
    ("hex", 32, [(0.5, 0.5, 0.3, 1.0)], 13.0),
    ("square", 48, [(0.5, 0.5, 0.2, 3.5), (0.1, 0.8, 0.15, 1.5)], 9.0),
    ("hex", 64, [(0.5, 0.5, 0.41, 2.0), (0.05, 0.07, 0.12, 1.0)], 15.0),
]


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


def _geom(lattice, atoms, eps_bg):
    lat = B.Lattice2D.square() if lattice == "square" else B.Lattice2D.hexagonal()
    return B.Geometry2D(lat, eps_bg, [B.BasisAtom((a[0], a[1]), a[2], a[3]) for a in atoms])


@pytest.mark.parametrize("lattice,res,atoms,eps_bg", CASES)
@pytest.mark.parametrize("smoothing", [True, False])
def test_gpu_dielectric_matches_oracle(be, lattice, res, atoms, eps_bg, smoothing):
    eps, inv, ten, _ = B.dielectric_batch(be, [_geom(lattice, atoms, eps_bg)], res, res, smoothing)
    oe, oi, ot = O.dielectric(O.Job(lattice=lattice, eps_bg=eps_bg, atoms=atoms, nx=res, ny=res,
                                    smoothing=smoothing))
    assert np.abs(eps[0] - oe).max() < ATOL
    assert np.abs(inv[0] - oi).max() < ATOL
    assert np.abs(ten[0] - ot).max() < ATOL


def test_gpu_dielectric_matches_reference_epsilon_grids(be):
    # the reference CLI's own smoothed-eps exports (square, eps_bg 13, hole eps 1, r 0.3)
    grids = json.load(open(os.path.join(GOLD, "epsilon_grids.json")))
    for res, d in grids.items():
        res = int(res)
        g = _geom("square", [(0.5, 0.5, d["radius"], d["epsilon_hole"])], d["epsilon_background"])
        eps, _, _, _ = B.dielectric_batch(be, [g], res, res, True)
        ref = np.array(d["grid"], dtype=float).reshape(-1)
        assert np.abs(eps[0] - ref).max() < 1e-9, res


def test_gpu_dielectric_batch_matches_host_on_sweep(be):
    # a batch of sweep geometries (square and hex, one or two inclusions) at
    # 64 x 64, against the host tables one geometry at a time
    geoms = [j.geom for j in res64_sweep(48)[::2]]
    eps, inv, ten, ms = B.dielectric_batch(be, geoms, 64, 64, True)
    assert ms > 0.0
    for k, g in enumerate(geoms):
        he, hi, ht = B.dielectric(g, 64, 64, True)
        assert np.abs(eps[k] - he).max() < ATOL, k
        assert np.abs(inv[k] - hi).max() < ATOL, k
        assert np.abs(ten[k] - ht).max() < ATOL, k


def test_gpu_dielectric_large_grid(be):
    # 256 x 256 (the large combined case's grid): the host path as reference,
    # plus invariants that hold for any grid: eps within [min, max] of the
    # materials, <1/eps> >= 1/eps (harmonic <= arithmetic), symmetric tensor
    g = _geom("square", [(0.0, 0.0, 0.2, 1.0), (0.5, 0.5, 0.12, 1.0)], 12.0)
    eps, inv, ten, _ = B.dielectric_batch(be, [g], 256, 256, True)
    he, hi, ht = B.dielectric(g, 256, 256, True)
    assert np.abs(eps[0] - he).max() < ATOL
    assert np.abs(ten[0] - ht).max() < ATOL
    assert eps.min() >= 1.0 - 1e-12 and eps.max() <= 12.0 + 1e-12
    assert np.all(inv[0] >= 1.0 / eps[0] - 1e-14)
    assert np.array_equal(ten[0][:, 1], ten[0][:, 2])
