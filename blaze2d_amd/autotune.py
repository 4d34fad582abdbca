"""Launch-shape autotuner for the batched band solve.

Picks, for a workload, the run options that the hand sweeps of this backend
found to matter: how many engines run side by side (`engines`), how many slot
groups each pipelines (`groups`), and from which width FP64 Gram products use
the matrix cores (`gram_mfma_min`).  It times short runs of a sample of the
caller's own jobs (every j-th job, the slots scaled by the same fraction), on
the device the jobs will run on, checks that every candidate returns the same
frequencies (to 1e-6 relative; the candidates change only the order of
floating-point sums), and keeps the fastest.  Results can be cached in a JSON
file keyed by the workload's shape (grids, band counts, polarisations,
precision, job-count bucket).

Usage:
    best = autotune.tune(backend, jobs, slots=2048)
    run_batch(backend, jobs, slots=2048, **best.options)
"""
from __future__ import annotations

import itertools
import json
import os
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from .backend import HipBackend
from .bandstructure import BandStructureJob, RunOptions, run_batch

DEFAULT_SPACE = {"engines": (1, 2, 3), "groups": (2, 4), "gram_mfma_min": (17, 33, 129)}


@dataclass
class Trial:
    options: Dict[str, int]
    seconds: float
    solves_per_s: float
    iterations: int


@dataclass
class TuneResult:
    options: Dict[str, int]           # the fastest candidate (run_batch keyword arguments)
    trials: List[Trial] = field(default_factory=list)
    sample_jobs: int = 0
    max_rel_freq_spread: float = 0.0  # largest difference between candidates' frequencies
    cached: bool = False


def signature(jobs: Sequence[BandStructureJob], precision: int, slots: int) -> str:
    """The workload's shape: what the tuned choice depends on."""
    grids = sorted({(j.nx, j.ny) for j in jobs})
    bands = sorted({j.n_bands for j in jobs})
    pols = sorted({j.pol.upper() for j in jobs})
    bucket = 1 << max(0, int(np.ceil(np.log2(max(len(jobs), 1)))))
    return json.dumps({"grids": grids, "bands": bands, "pols": pols, "precision": precision, "slots": slots,
                       "jobs": bucket}, sort_keys=True)


def _solves(results) -> int:
    return int(sum(int(r.converged[:-1].sum()) if r.converged.shape[0] > 1 else int(r.converged.sum())
                   for r in results))


def tune(backend: HipBackend, jobs: Sequence[BandStructureJob], slots: int = 2048, precision: int = 0,
         options: Optional[RunOptions] = None, space: Optional[Dict[str, Sequence[int]]] = None,
         sample: int = 512, cache_path: Optional[str] = None, repeats: int = 1) -> TuneResult:
    """Times every combination in `space` on a sample of `jobs` and returns the
    fastest.  With cache_path, a stored result for the same workload shape is
    returned without timing, and a new result is stored."""
    jobs = list(jobs)
    if not jobs:
        raise ValueError("no jobs to tune on")
    key = signature(jobs, precision, slots)
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            return TuneResult(options=cache[key]["options"], sample_jobs=cache[key]["sample_jobs"], cached=True)
    space = dict(DEFAULT_SPACE if space is None else space)
    stride = max(1, len(jobs) // max(1, sample))
    samp = jobs[::stride]
    s_slots = max(1, min(len(samp), int(round(slots * len(samp) / len(jobs)))))
    names = sorted(space)
    cands = [dict(zip(names, vals)) for vals in itertools.product(*(space[n] for n in names))]
    # first launches of every kernel shape out of the way
    run_batch(backend, samp[:2], options, slots=2, precision=precision)
    trials, ref = [], None
    spread = 0.0
    for c in cands:
        best_t = None
        for _ in range(max(1, repeats)):
            t = time.perf_counter()
            res, st = run_batch(backend, samp, options, slots=s_slots, precision=precision, **c)
            dt = time.perf_counter() - t
            best_t = dt if best_t is None else min(best_t, dt)
        if ref is None:
            ref = res
        else:
            for a, b in zip(res, ref):
                sa, sb = np.sort(a.bands, axis=1), np.sort(b.bands, axis=1)
                m = sb > 1e-8
                if m.any():
                    spread = max(spread, float(np.max(np.abs(sa[m] - sb[m]) / sb[m])))
        trials.append(Trial(c, best_t, _solves(res) / best_t, int(st.iterations)))
    if spread > 1e-6:
        raise RuntimeError(f"autotune: candidates disagree by {spread:.2e} (relative frequency)")
    best = max(trials, key=lambda t: t.solves_per_s)
    out = TuneResult(options=dict(best.options), trials=trials, sample_jobs=len(samp), max_rel_freq_spread=spread)
    if cache_path:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        cache[key] = {"options": out.options, "sample_jobs": out.sample_jobs,
                      "trials": [{"options": t.options, "solves_per_s": round(t.solves_per_s, 2)} for t in trials]}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return out


def main(argv=None) -> int:
    import argparse
    from .sweep import res64_sweep, sweep_jobs
    ap = argparse.ArgumentParser(description="tune the batched band solve's launch shape on a sweep")
    ap.add_argument("--geoms", type=int, default=512)
    ap.add_argument("--mixed", action="store_true", help="the mixed-grid sweep instead of the res-64 one")
    ap.add_argument("--slots", type=int, default=2048)
    ap.add_argument("--sample", type=int, default=512)
    ap.add_argument("--cache", default=None)
    a = ap.parse_args(argv)
    jobs = sweep_jobs(a.geoms, tol=1e-7) if a.mixed else res64_sweep(a.geoms, tol=1e-7)
    be = HipBackend(0)
    r = tune(be, jobs, slots=a.slots, sample=a.sample, cache_path=a.cache, options=RunOptions(exit_rule=1))
    for t in sorted(r.trials, key=lambda t: -t.solves_per_s):
        print(f"{t.options}: {t.solves_per_s:.0f} solves/s ({t.seconds:.3f} s)")
    print("best:", r.options)
    be.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
