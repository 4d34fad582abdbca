"""Seeded synthetic geometry sweeps (the batched band-structure workloads).

Generator (one seed, numpy PCG64):
  lattice ~ U{square, hex}; eps_bg ~ U[2, 16]; inclusions ~ U{1, 2};
  inclusion 0 at (0.5, 0.5) with eps ~ U[1, 4], r ~ U[0.1, 0.45] (units of a);
  inclusion 1 (if any) with eps ~ U[1, 4], r ~ U[0.1, 0.45] at a uniform
  fractional position, redrawn until the Cartesian nearest-image distance to
  inclusion 0 (3 x 3 images, as crates/core/src/geometry.rs:74-99) exceeds
  r0 + r1; after 64 failed positions r1 is redrawn, and after 16 redraws the
  geometry keeps a single inclusion.
  grid ~ U{32, 48, 64} and bands ~ U{8, 12, 16} unless fixed by the caller;
  each geometry is solved in TM and TE on its 31-point closed path
  (square Gamma-X-M-Gamma, hex Gamma-M-K-Gamma, 10 points per leg).
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np

from .bandstructure import BandStructureJob, BasisAtom, Geometry2D, Lattice2D, k_path

DEFAULT_SEED = 0xB1A2E2D


def _nearest_image_distance(lat: Lattice2D, p0, p1) -> float:
    best = np.inf
    for n1 in (-1, 0, 1):
        for n2 in (-1, 0, 1):
            df = (p1[0] - p0[0] + n1, p1[1] - p0[1] + n2)
            x = lat.a1[0] * df[0] + lat.a2[0] * df[1]
            y = lat.a1[1] * df[0] + lat.a2[1] * df[1]
            best = min(best, float(np.hypot(x, y)))
    return best


def geometries(n: int, seed: int = DEFAULT_SEED):
    """Yields (lattice_name, Geometry2D, rng) for n seeded geometries."""
    rng = np.random.Generator(np.random.PCG64(seed))
    for _ in range(n):
        name = "square" if rng.integers(0, 2) == 0 else "hex"
        lat = Lattice2D.square() if name == "square" else Lattice2D.hexagonal()
        eps_bg = float(rng.uniform(2.0, 16.0))
        n_incl = int(rng.integers(1, 3))
        r0 = float(rng.uniform(0.1, 0.45))
        e0 = float(rng.uniform(1.0, 4.0))
        atoms = [BasisAtom((0.5, 0.5), r0, e0)]
        if n_incl == 2:
            e1 = float(rng.uniform(1.0, 4.0))
            r1 = float(rng.uniform(0.1, 0.45))
            placed = False
            for _redraw in range(16):
                for _try in range(64):
                    pos = (float(rng.uniform(0.0, 1.0)), float(rng.uniform(0.0, 1.0)))
                    if _nearest_image_distance(lat, (0.5, 0.5), pos) > r0 + r1:
                        placed = True
                        break
                if placed:
                    break
                r1 = float(rng.uniform(0.1, 0.45))
            if placed:
                atoms.append(BasisAtom(pos, r1, e1))
        yield name, Geometry2D(lat, eps_bg, atoms), rng


def sweep_jobs(n_geometries: int, seed: int = DEFAULT_SEED, grid: Optional[int] = None,
               n_bands: Optional[int] = None, tol: float = 1e-7, pols=("TM", "TE"),
               segments_per_leg: int = 10) -> List[BandStructureJob]:
    """Band-structure jobs of the sweep: every geometry in each polarisation."""
    jobs = []
    for name, geom, rng in geometries(n_geometries, seed):
        g = grid if grid is not None else int(rng.choice([32, 48, 64]))
        nb = n_bands if n_bands is not None else int(rng.choice([8, 12, 16]))
        kp = k_path(name, segments_per_leg)
        for pol in pols:
            jobs.append(BandStructureJob(geom, g, g, pol, kp, n_bands=nb, tol=tol))
    return jobs


def res64_sweep(n_geometries: int, seed: int = DEFAULT_SEED, tol: float = 1e-7) -> List[BandStructureJob]:
    """The headline workload: the sweep generator fixed at 64 x 64 and 12 bands."""
    return sweep_jobs(n_geometries, seed, grid=64, n_bands=12, tol=tol)


LARGE_ATOMS = [((0.0, 0.0), 0.2, 1.0), ((0.5, 0.5), 0.12, 1.0)]


def large_case_geometry() -> Geometry2D:
    """The large combined case: square lattice, eps_bg = 12, eps = 1 rods of
    r = 0.2 a at (0, 0) and r = 0.12 a at (0.5, 0.5)."""
    return Geometry2D(Lattice2D.square(), 12.0, [BasisAtom(p, r, e) for p, r, e in LARGE_ATOMS])


def large_case_jobs(grid: int = 256, n_bands: int = 32, segments_per_leg: int = 21, tol: float = 1e-7,
                    pols=("TE", "TM")) -> List[BandStructureJob]:
    """The large combined case batched over k-points: one job per (polarisation,
    k-point) of the Gamma-X-M-Gamma path (21 points per leg: 64 points, the
    closing Gamma a copy of the first, so 63 solves per polarisation), each
    solved from the deterministic start."""
    geom = large_case_geometry()
    kp = k_path("square", segments_per_leg)[:-1]
    return [BandStructureJob(geom, grid, grid, pol, kp[i:i + 1], n_bands=n_bands, tol=tol, max_iter=400)
            for pol in pols for i in range(kp.shape[0])]


def large_case_m_point(grid: int = 512, n_bands: int = 16, tol: float = 1e-7) -> BandStructureJob:
    """The 512 x 512 TM 16-band solve at M = (1/2, 1/2) of the large case."""
    return BandStructureJob(large_case_geometry(), grid, grid, "TM", np.array([[0.5, 0.5]]), n_bands=n_bands,
                            tol=tol, max_iter=400)
