"""The C-ABI library loads and exports every entry point declared in
include/blaze2d_hip.h; host-only entry points agree with the oracle."""
import os
import re

import numpy as np
import pytest

from blaze2d_amd import _lib as L
from blaze2d_amd import bandstructure as B
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "blaze2d_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.EXPORTS)
    assert b"gfx950" in lib.bh_version()


def test_errors_are_reported_not_crashing():
    import ctypes
    lib = L.lib()
    out = ctypes.c_void_p()
    # a bad device index fails with a message (no GPU here, or index out of range)
    st = lib.bh_context_create(99, ctypes.byref(out))
    assert st != 0
    assert len(lib.bh_last_error()) > 0


@pytest.mark.parametrize("lattice,res,atoms,eps_bg", [
    ("square", 24, [(0.5, 0.5, 0.3, 1.0)], 13.0),
    ("hex", 32, [(0.5, 0.5, 0.3, 1.0)], 13.0),
    ("square", 48, [(0.5, 0.5, 0.2, 3.5), (0.1, 0.8, 0.15, 1.5)], 9.0),
    ("hex", 64, [(0.5, 0.5, 0.41, 2.0), (0.05, 0.07, 0.12, 1.0)], 15.0),
])
@pytest.mark.parametrize("smoothing", [True, False])
def test_product_dielectric_matches_oracle(lattice, res, atoms, eps_bg, smoothing):
    lat = B.Lattice2D.square() if lattice == "square" else B.Lattice2D.hexagonal()
    g = B.Geometry2D(lat, eps_bg, [B.BasisAtom((a[0], a[1]), a[2], a[3]) for a in atoms])
    eps, inv, ten = B.dielectric(g, res, res, smoothing)
    j = O.Job(lattice=lattice, eps_bg=eps_bg, atoms=atoms, nx=res, ny=res, smoothing=smoothing)
    oe, oi, ot = O.dielectric(j)
    assert np.abs(eps - oe).max() < 1e-13
    assert np.abs(inv - oi).max() < 1e-13
    assert np.abs(ten - ot).max() < 1e-13


def test_kpath_matches_oracle():
    for lat in ("square", "hex"):
        for seg in (1, 10, 20):
            assert np.array_equal(B.k_path(lat, seg), O.kpath(lat, seg))
    assert B.k_path("square", 10).shape == (31, 2)
