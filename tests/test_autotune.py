"""The launch-shape autotuner's host logic (CPU): the workload signature and
the cache path, which returns a stored choice without touching a device."""
import json

import numpy as np

from blaze2d_amd import autotune as A
from blaze2d_amd import bandstructure as B


def _jobs(n, grid, nb, pol):
    g = B.Geometry2D(B.Lattice2D.square(), 12.0, [B.BasisAtom((0.5, 0.5), 0.2, 1.0)])
    return [B.BandStructureJob(g, grid, grid, pol, B.k_path("square", 2), n_bands=nb) for _ in range(n)]


def test_signature_depends_on_shape_not_values():
    a = A.signature(_jobs(100, 64, 12, "TM") + _jobs(100, 64, 12, "TE"), 0, 2048)
    b = A.signature(_jobs(120, 64, 12, "TE") + _jobs(80, 64, 12, "TM"), 0, 2048)
    assert a == b  # same grids, bands, polarisations, job-count bucket
    assert a != A.signature(_jobs(200, 48, 12, "TM"), 0, 2048)
    assert a != A.signature(_jobs(200, 64, 12, "TM") + _jobs(200, 64, 12, "TE"), 0, 2048)  # bucket 256 -> 512
    assert a != A.signature(_jobs(100, 64, 12, "TM") + _jobs(100, 64, 12, "TE"), 2, 2048)


def test_cached_choice_is_returned_without_a_device(tmp_path):
    jobs = _jobs(10, 32, 8, "TM")
    path = tmp_path / "tune.json"
    key = A.signature(jobs, 0, 64)
    path.write_text(json.dumps({key: {"options": {"engines": 2, "groups": 4, "gram_mfma_min": 33},
                                     "sample_jobs": 10}}))
    r = A.tune(None, jobs, slots=64, cache_path=str(path))
    assert r.cached and r.options == {"engines": 2, "groups": 4, "gram_mfma_min": 33}
