"""The bulk driver's contract on the GPU path.

A bulk configuration (the reference's ``blaze-bulk --config`` TOML) is
expanded into band-structure jobs in the reference's order, the jobs run
through the batched engine (``run_batch``), and the results are written as
the reference's per-job CSV files.  ``benchmark`` replays
``BulkDriver::benchmark`` (real solves, no file output, jobs/s).

Reference behaviour restated here:

* configuration fields and defaults: ``bulk-driver-core/src/config.rs``
  (``BulkConfig`` :1043-1120, ``DefaultsConfig`` :606-618, ``BaseGeometry``
  :736-781, ``effective_*`` :1167-1183, ``RangeSpec::values`` :209-219,
  ``SweepDimension::from_spec`` :461-494);
* expansion order: ``bulk-driver-core/src/expansion.rs`` (``[[sweeps]]``,
  first sweep outermost, :230-332; parameter application :334-470; legacy
  ``[ranges]`` eps_bg -> resolution -> polarization -> lattice_type -> atoms,
  :506-610 and :679-773; job construction :776-871);
* the per-job run: ``RunOptions::default()`` with ``bulk.disable_band_tracking``
  (``bulk-driver/src/driver.rs:929-945``);
* CSV: ``bulk-driver/src/output.rs:108-190`` (``{prefix}_{index:06}.csv``,
  ``JobParams::to_columns`` at ``expansion.rs:172-206``, omega / 2 pi);
* k-path distances: ``drivers/bandstructure.rs:933-949`` (fractional).

Not restated: the operator-data solver, selective output mode, streaming and
batched I/O, adaptive threading (the GPU engine replaces the per-job thread
pool).  A job the GPU path does not support (a grid side outside 8, 16, 24,
32, 48, 64, 96, 128, 192, 256, 384 and 512, a block wider than 96 columns,
more than 64 atoms) raises an error.
"""
from __future__ import annotations

import argparse
import os
import re
import sys
import time
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .bandstructure import (BandStructureJob, BandStructureResult, BasisAtom, Geometry2D, Lattice2D, RunOptions,
                            run_batch)

_ATOM_PATH = re.compile(r"^atom(\d+)\.(pos_x|pos_y|radius|eps_inside)$")

SQUARE_PATH = [(0.0, 0.0), (0.5, 0.0), (0.5, 0.5), (0.0, 0.0)]
HEX_PATH = [(0.0, 0.0), (0.5, 0.0), (1.0 / 3.0, 1.0 / 3.0), (0.0, 0.0)]
RECT_PATH = [(0.0, 0.0), (0.5, 0.0), (0.5, 0.5), (0.0, 0.5), (0.0, 0.0)]


def load_config(source: str) -> Dict[str, Any]:
    """Parses a bulk TOML file (a path) or TOML text."""
    import tomli

    if os.path.exists(source):
        with open(source, "rb") as f:
            return tomli.load(f)
    return tomli.loads(source)


def range_values(vmin: float, vmax: float, step: float) -> List[float]:
    """RangeSpec::values: accumulate from min while v <= max + step * 1e-9.
    A step <= 0 never reaches max and is rejected."""
    if not step > 0:
        raise ValueError(f"range step must be positive (got {step})")
    out = []
    v = vmin
    while v <= vmax + step * 1e-9:
        out.append(v)
        v += step
    return out


def densify(nodes: Sequence[Tuple[float, float]], segments_per_leg: int) -> List[Tuple[float, float]]:
    """brillouin.rs densify_path: linear steps, corners not repeated."""
    if len(nodes) <= 1:
        return [tuple(n) for n in nodes]
    seg = max(int(segments_per_leg), 1)
    path = [tuple(nodes[0])]
    for a, b in zip(nodes[:-1], nodes[1:]):
        for step in range(1, seg + 1):
            t = step / seg
            p = ((1.0 - t) * a[0] + t * b[0], (1.0 - t) * a[1] + t * b[1])
            if path[-1] != p:
                path.append(p)
    return path


def path_distances(k_path: np.ndarray) -> np.ndarray:
    """compute_k_path_distances: cumulative Euclidean steps of the fractional path."""
    if len(k_path) == 0:
        return np.zeros(0)
    d = np.sqrt(np.sum(np.diff(k_path, axis=0) ** 2, axis=1))
    return np.concatenate([[0.0], np.cumsum(d)])


def rust_float(x: float) -> str:
    """An f64 as Rust's `{}` prints it: shortest round-trip digits, positional,
    no trailing `.0` ("1", "0.25", "0.0000001")."""
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "inf" if x > 0 else "-inf"
    return np.format_float_positional(float(x), trim="-")


@dataclass
class AtomParams:
    index: int
    pos: Tuple[float, float]
    radius: float
    eps_inside: float


@dataclass
class JobParams:
    eps_bg: float
    resolution: int
    polarization: str
    lattice_type: Optional[str]
    atoms: List[AtomParams]
    sweep_values: List[Tuple[str, Any]] = field(default_factory=list)

    def to_columns(self) -> List[Tuple[str, str]]:
        """JobParams::to_columns (expansion.rs:172-206)."""
        cols = [("eps_bg", f"{self.eps_bg:.6f}"), ("resolution", str(self.resolution)),
                ("polarization", self.polarization)]
        if self.lattice_type is not None:
            cols.append(("lattice_type", self.lattice_type))
        for a in self.atoms:
            p = f"atom{a.index}"
            cols += [(f"{p}_pos_x", f"{a.pos[0]:.6f}"), (f"{p}_pos_y", f"{a.pos[1]:.6f}"),
                     (f"{p}_radius", f"{a.radius:.6f}"), (f"{p}_eps_inside", f"{a.eps_inside:.6f}")]
        return cols


@dataclass
class ExpandedJob:
    index: int
    job: BandStructureJob
    params: JobParams


def _lattice_type_name(s: str) -> str:
    s = s.lower()
    return s if s in ("square", "rectangular", "triangular", "hexagonal") else "square"


def _pol(s: str, fallback: str) -> str:
    s = str(s).upper()
    return s if s in ("TM", "TE") else fallback


class BulkConfig:
    """The fields of bulk-driver-core's BulkConfig that Maxwell jobs read."""

    def __init__(self, raw: Dict[str, Any]):
        self.raw = raw
        solver = raw.get("solver", {})
        self.solver_type = str(solver.get("type", "maxwell")).lower()
        self.precision = str(solver.get("precision", "f64")).lower()
        self.bulk = raw.get("bulk", {})
        self.sweeps = list(raw.get("sweeps", []))
        self.defaults = raw.get("defaults", {})
        self.geometry = raw.get("geometry")
        self.polarization = _pol(raw.get("polarization", "TM"), "TM")  # deprecated top-level field
        grid = raw.get("grid")
        if grid is None:
            raise ValueError("bulk config: [grid] is required")
        self.grid_nx = int(grid["nx"])
        self.grid_ny = int(grid.get("ny", grid["nx"]))
        self.path = raw.get("path")
        self.k_path = [tuple(k) for k in raw.get("k_path", [])]
        eig = raw.get("eigensolver", {})
        self.n_bands = int(eig.get("n_bands", 8))
        self.max_iter = int(eig.get("max_iter", 200))
        self.tol = float(eig.get("tol", 1e-6))
        self.block_size = int(eig.get("block_size", 0))
        self.ranges = raw.get("ranges", {})
        out = raw.get("output", {})
        self.output_mode = str(out.get("mode", "full")).lower()
        self.output_dir = str(out.get("directory", "./bulk_output"))
        self.output_prefix = str(out.get("prefix", "job"))
        self.validate()

    def validate(self) -> None:
        """BulkConfig::validate (config.rs:1215-1300) with SweepSpec::validate
        (:338-376) and validate_parameter_path (:528-577); raises ValueError."""
        if self.sweeps:
            seen = set()
            for sw in self.sweeps:
                name = sw.get("parameter")
                if name in seen:
                    raise ValueError(f"duplicate sweep parameter '{name}'")
                seen.add(name)
                has_range = all(sw.get(k) is not None for k in ("min", "max", "step"))
                has_values = sw.get("values") is not None
                if not has_range and not has_values:
                    raise ValueError(f"sweep '{name}' must specify either (min, max, step) or values")
                if has_range and has_values:
                    raise ValueError(f"sweep '{name}' cannot have both range and values")
                if has_range:
                    if float(sw["step"]) <= 0.0:
                        raise ValueError(f"sweep '{name}' step must be positive")
                    if float(sw["min"]) > float(sw["max"]):
                        raise ValueError(f"sweep '{name}' min ({sw['min']}) > max ({sw['max']})")
                if name not in ("eps_bg", "resolution", "polarization", "lattice_type") and \
                        not _ATOM_PATH.match(str(name)):
                    raise ValueError(f"invalid sweep parameter path '{name}'")
            idx = [int(m.group(1)) for m in (_ATOM_PATH.match(str(sw.get("parameter"))) for sw in self.sweeps) if m]
            if idx and max(idx) > 0 and not _geom_atoms(self.effective_geometry()):
                raise ValueError(f"sweep references atom{max(idx)} but no atoms defined in geometry")
        if self.solver_type != "maxwell":
            return
        if self.effective_geometry() is None:
            raise ValueError("Maxwell solver requires geometry (use [geometry] or [defaults.geometry])")
        if not self.sweeps:
            e = self.ranges.get("eps_bg")
            if e is not None:
                if float(e["step"]) <= 0.0:
                    raise ValueError("eps_bg step must be positive")
                if float(e["min"]) > float(e["max"]):
                    raise ValueError("eps_bg min > max")
            r = self.ranges.get("resolution")
            if r is not None:
                if float(r["step"]) < 1.0:
                    raise ValueError("resolution step must be >= 1")
                if float(r["min"]) < 4.0:
                    raise ValueError("resolution min must be >= 4")

    # effective_geometry / effective_polarization / effective_resolution (config.rs:1167-1183)
    def effective_geometry(self) -> Optional[Dict[str, Any]]:
        return self.defaults.get("geometry") or self.geometry

    def effective_polarization(self) -> str:
        return _pol(self.defaults["polarization"], self.polarization) if "polarization" in self.defaults \
            else self.polarization

    def effective_resolution(self) -> int:
        return int(self.defaults.get("resolution", self.grid_nx))


def _geom_eps_bg(g: Optional[Dict[str, Any]]) -> float:
    return float(g.get("eps_bg", 12.0)) if g is not None else 12.0  # default_eps_bg


def _geom_atoms(g: Optional[Dict[str, Any]]) -> List[Tuple[float, float, float, float]]:
    if g is None:
        return []
    return [(float(a["pos"][0]), float(a["pos"][1]), float(a["radius"]), float(a.get("eps_inside", 1.0)))
            for a in g.get("atoms", [])]


def _build_lattice(cfg: BulkConfig, lattice_type: Optional[str]) -> Lattice2D:
    """build_lattice (expansion.rs:850-871) over the [geometry] lattice."""
    base = (cfg.geometry or {}).get("lattice", {})
    if "a1" in base and "a2" in base:
        return Lattice2D(tuple(map(float, base["a1"])), tuple(map(float, base["a2"])))
    a = float(base.get("a", 1.0))
    lt = lattice_type or (_lattice_type_name(base["type"]) if "type" in base else "square")
    if lt == "square":
        return Lattice2D.square(a)
    if lt == "rectangular":
        return Lattice2D((a, 0.0), (0.0, float(base.get("b", a * 1.5))))
    return Lattice2D.hexagonal(a)


def _k_path(cfg: BulkConfig) -> np.ndarray:
    """The k-path of create_maxwell_job (expansion.rs:812-837)."""
    if cfg.k_path:
        pts = cfg.k_path
    elif cfg.path is not None:
        spec = cfg.path
        if spec.get("k_path"):
            pts = [tuple(k) for k in spec["k_path"]]
        elif "preset" in spec:
            seg = int(spec.get("segments_per_leg", 8))
            preset = str(spec["preset"]).lower()
            nodes = {"square": SQUARE_PATH, "hexagonal": HEX_PATH, "triangular": HEX_PATH,
                     "rectangular": RECT_PATH}[preset]
            pts = densify(nodes, seg)
        else:
            pts = densify(SQUARE_PATH, 12)
    else:
        pts = densify(SQUARE_PATH, 12)
    return np.array(pts, dtype=np.float64).reshape(-1, 2)


def _make_job(cfg: BulkConfig, eps_bg: float, resolution: int, pol: str, lattice_type: Optional[str],
              atoms: Sequence[Tuple[float, float, float, float]]) -> BandStructureJob:
    """create_maxwell_job (expansion.rs:776-848)."""
    if cfg.geometry is None:
        raise ValueError("bulk config: geometry required for the Maxwell solver")
    lat = _build_lattice(cfg, lattice_type)
    geom = Geometry2D(lat, eps_bg, [BasisAtom((a[0], a[1]), a[2], a[3]) for a in atoms])
    return BandStructureJob(geom, resolution, resolution, pol, _k_path(cfg), n_bands=cfg.n_bands, tol=cfg.tol,
                            max_iter=cfg.max_iter, block_size=cfg.block_size)


def _sweep_values(spec: Dict[str, Any]) -> List[Any]:
    """SweepDimension::from_spec: discrete values, or a range (resolution rounded)."""
    name = spec["parameter"]
    if "values" in spec and spec["values"] is not None:
        return list(spec["values"])
    if all(k in spec for k in ("min", "max", "step")):
        vals = range_values(float(spec["min"]), float(spec["max"]), float(spec["step"]))
        return [int(round(v)) for v in vals] if name == "resolution" else vals
    raise ValueError(f"sweep '{name}' has no valid range or values")


def _expand_ordered(cfg: BulkConfig) -> List[ExpandedJob]:
    names = [s["parameter"] for s in cfg.sweeps]
    if len(set(names)) != len(names):
        raise ValueError("duplicate sweep parameter")
    dims = [(s["parameter"], _sweep_values(s)) for s in cfg.sweeps]
    geo = cfg.effective_geometry()
    jobs = []
    # first sweep outermost: the last index varies fastest (generate_ordered_indices)
    for index, combo in enumerate(np.ndindex(*[max(len(v), 1) for _, v in dims])):
        sv = [(name, vals[i]) for (name, vals), i in zip(dims, combo)]
        eps_bg = _geom_eps_bg(geo)
        resolution = cfg.effective_resolution()
        pol = cfg.effective_polarization()
        base_lat = (geo or {}).get("lattice", {})
        lattice_type = _lattice_type_name(base_lat["type"]) if "type" in base_lat else None
        atom = {i: list(a) for i, a in enumerate(_geom_atoms(geo))}
        for name, val in sv:
            if name == "eps_bg" and isinstance(val, (int, float)):
                eps_bg = float(val)
            elif name == "resolution" and isinstance(val, (int, float)):
                resolution = int(val)
            elif name == "polarization" and isinstance(val, str):
                pol = _pol(val, pol)
            elif name == "lattice_type" and isinstance(val, str):
                lattice_type = _lattice_type_name(val)
            else:
                m = _ATOM_PATH.match(name)
                if m and isinstance(val, (int, float)):
                    e = atom.setdefault(int(m.group(1)), [0.5, 0.5, 0.3, 1.0])
                    e[("pos_x", "pos_y", "radius", "eps_inside").index(m.group(2))] = float(val)
        top = max(atom) if atom else 0
        atoms = [tuple(atom.get(i, [0.5, 0.5, 0.3, 1.0])) for i in range(top + 1)]
        job = _make_job(cfg, eps_bg, resolution, pol, lattice_type, atoms)
        params = JobParams(eps_bg, resolution, pol, lattice_type,
                           [AtomParams(i, (a[0], a[1]), a[2], a[3]) for i, a in enumerate(atoms)], sv)
        jobs.append(ExpandedJob(index, job, params))
    return jobs


def _expand_legacy(cfg: BulkConfig) -> List[ExpandedJob]:
    if cfg.geometry is None:
        raise ValueError("bulk config: geometry required for the Maxwell solver")
    r = cfg.ranges

    def rng(key):
        spec = r.get(key)
        return None if spec is None else range_values(float(spec["min"]), float(spec["max"]), float(spec["step"]))

    # an axis given but empty (a range with min > max, an empty list) yields no
    # jobs, as in the reference; only an absent axis takes the base value
    eps_vals = rng("eps_bg") if "eps_bg" in r else [_geom_eps_bg(cfg.geometry)]
    res_vals = [int(round(v)) for v in rng("resolution")] if "resolution" in r else [cfg.grid_nx]
    pol_vals = [_pol(p, "TM") for p in r["polarization"]] if "polarization" in r else [cfg.polarization]
    lat_vals = [_lattice_type_name(t) for t in r["lattice_type"]] if "lattice_type" in r else [None]
    atom_ranges = r.get("atoms", [])
    per_atom = []
    for i, base in enumerate(_geom_atoms(cfg.geometry)):
        ar = atom_ranges[i] if i < len(atom_ranges) else None
        if ar is None:
            per_atom.append([base])
            continue

        def axis(key, default):
            s = ar.get(key)
            return range_values(float(s["min"]), float(s["max"]), float(s["step"])) if s else [default]

        per_atom.append([(px, py, rad, ep) for rad in axis("radius", base[2]) for px in axis("pos_x", base[0])
                         for py in axis("pos_y", base[1]) for ep in axis("eps_inside", base[3])])
    combos: List[List[Tuple[float, float, float, float]]] = [[]]
    for opts in per_atom:
        combos = [c + [o] for c in combos for o in opts]
    jobs = []
    for eps_bg in eps_vals:
        for res in res_vals:
            for pol in pol_vals:
                for lt in lat_vals:
                    for atoms in combos:
                        job = _make_job(cfg, eps_bg, res, pol, lt, atoms)
                        params = JobParams(eps_bg, res, pol, lt,
                                           [AtomParams(i, (a[0], a[1]), a[2], a[3]) for i, a in enumerate(atoms)])
                        jobs.append(ExpandedJob(len(jobs), job, params))
    return jobs


def expand_jobs(config) -> List[ExpandedJob]:
    """expand_jobs (expansion.rs:230-241) for the Maxwell solver."""
    cfg = config if isinstance(config, BulkConfig) else BulkConfig(config)
    if cfg.solver_type != "maxwell":
        raise ValueError(f"solver type '{cfg.solver_type}' is not on the GPU path (Maxwell only)")
    if cfg.sweeps:
        return _expand_ordered(cfg)
    # no [[sweeps]]: the legacy loop (a single job when no [ranges] are given)
    return _expand_legacy(cfg)


def run_jobs(backend, config, jobs: Optional[List[ExpandedJob]] = None, slots: int = 1024, groups: int = 4):
    """Runs every expanded job on the GPU in one call (one engine per grid
    size, side by side).  Returns (results in job order, engine seconds)."""
    cfg = config if isinstance(config, BulkConfig) else BulkConfig(config)
    jobs = expand_jobs(cfg) if jobs is None else jobs
    if not jobs:
        return [], 0.0
    opts = RunOptions(disable_band_tracking=bool(cfg.bulk.get("disable_band_tracking", False)))
    precision = 1 if cfg.precision == "f32" else 0
    grids = {(ej.job.nx, ej.job.ny) for ej in jobs}
    res, st = run_batch(backend, [ej.job for ej in jobs], opts, slots=min(slots, len(jobs)), precision=precision,
                        groups=groups, engines=len(grids))
    return list(res), st.seconds


def write_full_csv(directory: str, prefix: str, ej: ExpandedJob, result: BandStructureResult) -> str:
    """write_full_maxwell (output.rs:108-190): one file per job, omega / 2 pi."""
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, f"{prefix}_{ej.index:06d}.csv")
    cols = ej.params.to_columns()
    bands = result.bands
    max_bands = bands.shape[1] if bands.ndim == 2 else 0
    dist = path_distances(np.asarray(result.k_path, dtype=np.float64))
    with open(path, "w") as f:
        f.write("job_index" + "".join(f",{n}" for n, _ in cols) + ",k_index,kx,ky,k_distance")
        f.write("".join(f",band{i}" for i in range(1, max_bands + 1)) + "\n")
        vals = "".join(f",{v}" for _, v in cols)
        for k, (kp, row) in enumerate(zip(result.k_path, bands)):
            f.write(f"{ej.index}{vals},{k},{rust_float(kp[0])},{rust_float(kp[1])},{rust_float(dist[k])}")
            f.write("".join("," + rust_float(w / (2.0 * np.pi)) for w in row) + "\n")
    return path


def run(backend, config, output_dir: Optional[str] = None, slots: int = 1024, groups: int = 4):
    """BulkDriver::run for Maxwell jobs in full output mode: solve, then write the CSVs."""
    cfg = config if isinstance(config, BulkConfig) else BulkConfig(config)
    if cfg.output_mode != "full":
        raise ValueError(f"output mode '{cfg.output_mode}' is not built (full mode only)")
    jobs = expand_jobs(cfg)
    results, seconds = run_jobs(backend, cfg, jobs, slots, groups)
    out_dir = output_dir or cfg.output_dir
    paths = [write_full_csv(out_dir, cfg.output_prefix, ej, r) for ej, r in zip(jobs, results)]
    return jobs, results, paths, seconds


def benchmark(backend, config, slots: int = 1024, groups: int = 4) -> Dict[str, Any]:
    """BulkDriver::benchmark (driver.rs:1104-1257): real solves, no file output."""
    cfg = config if isinstance(config, BulkConfig) else BulkConfig(config)
    jobs = expand_jobs(cfg)
    t = time.perf_counter()
    results, _ = run_jobs(backend, cfg, jobs, slots, groups)
    total = time.perf_counter() - t
    done = sum(1 for r in results if r is not None)
    print(f"✓ {done} jobs benchmarked in {total:.2f}s ({done / total:.1f} jobs/s)")
    return {"jobs": done, "seconds": total, "jobs_per_s": done / total}


def main(argv: Optional[Sequence[str]] = None) -> int:
    """`blaze-bulk --config FILE [--benchmark] [--dry-run] [--output DIR]` (bulk-driver/src/main.rs:31-90)."""
    ap = argparse.ArgumentParser(prog="python -m blaze2d_amd.bulk")
    ap.add_argument("--config", "-c", required=True)
    ap.add_argument("--benchmark", action="store_true", help="real solves, no file output")
    ap.add_argument("--dry-run", action="store_true", help="count the jobs without running")
    ap.add_argument("--output", "-o", default=None, help="output directory (overrides [output].directory)")
    ap.add_argument("--slots", type=int, default=1024)
    args = ap.parse_args(argv)
    cfg = BulkConfig(load_config(args.config))
    jobs = expand_jobs(cfg)
    if args.dry_run:
        print(f"{len(jobs)} jobs")
        return 0
    from .backend import HipBackend

    be = HipBackend(0)
    try:
        if args.benchmark:
            benchmark(be, cfg, slots=args.slots)
        else:
            _, _, paths, seconds = run(be, cfg, args.output, slots=args.slots)
            print(f"{len(paths)} jobs written to {args.output or cfg.output_dir} ({seconds:.2f}s engine time)")
    finally:
        be.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
