"""The autotuner on the GPU: every candidate returns the same frequencies and
the fastest is chosen and cached."""
import json

import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd import autotune as A
from blaze2d_amd.bandstructure import RunOptions
from blaze2d_amd.sweep import res64_sweep

pytestmark = pytest.mark.gpu


def test_tune_a_small_sweep(tmp_path):
    be = HipBackend(0)
    try:
        jobs = res64_sweep(32, tol=1e-7)
        space = {"engines": (1, 2), "groups": (2, 4), "gram_mfma_min": (17, 129)}
        path = str(tmp_path / "tune.json")
        r = A.tune(be, jobs, slots=64, space=space, sample=64, cache_path=path, options=RunOptions(exit_rule=1))
        for t in sorted(r.trials, key=lambda t: -t.solves_per_s):
            print(f"{t.options}: {t.solves_per_s:.0f} solves/s")
        assert len(r.trials) == 8 and not r.cached
        assert r.max_rel_freq_spread <= 1e-6
        assert r.options == max(r.trials, key=lambda t: t.solves_per_s).options
        again = A.tune(be, jobs, slots=64, space=space, sample=64, cache_path=path)
        assert again.cached and again.options == r.options
        assert A.signature(jobs, 0, 64) in json.load(open(path))
    finally:
        be.close()
