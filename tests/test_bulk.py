"""Bulk-driver contract: job expansion order, ranges and CSV output
(bulk-driver-core/src/expansion.rs, config.rs; bulk-driver/src/output.rs).

The expansion cases restate the reference's own unit tests
(expansion.rs:877-1040) and its example configurations (examples/*.toml,
read here as fixtures under tests/golden/bulk/).  The GPU test runs a small
sweep end to end and checks the CSV values against the oracle."""
import csv
import os

import numpy as np
import pytest

from blaze2d_amd import bulk

GOLD = os.path.join(os.path.dirname(__file__), "golden", "bulk")

BASE = """
[bulk]
[solver]
type = "maxwell"
[geometry]
eps_bg = 12.0
[geometry.lattice]
type = "square"
a = 1.0
[[geometry.atoms]]
pos = [0.5, 0.5]
radius = 0.3
eps_inside = 1.0
[grid]
nx = 32
ny = 32
lx = 1.0
ly = 1.0
[path]
preset = "square"
segments_per_leg = 12
"""


def _cfg(extra=""):
    return bulk.BulkConfig(bulk.load_config(BASE + extra))


# --- the reference's expansion tests (expansion.rs:877-1040) ----------------
def test_expand_single_job():
    jobs = bulk.expand_jobs(_cfg())
    assert len(jobs) == 1 and jobs[0].index == 0
    assert jobs[0].job.k_path.shape == (37, 2)  # 3 legs x 12 + 1


def test_expand_eps_range_legacy():
    assert len(bulk.expand_jobs(_cfg("[ranges]\neps_bg = {min = 10.0, max = 12.0, step = 1.0}\n"))) == 3


def test_expand_multiple_ranges_legacy():
    jobs = bulk.expand_jobs(_cfg('[ranges]\neps_bg = {min = 10.0, max = 12.0, step = 1.0}\n'
                                 'polarization = ["TM", "TE"]\n'))
    assert len(jobs) == 6
    # hard-coded legacy order: eps_bg outer, polarization inner
    assert [(j.params.eps_bg, j.params.polarization) for j in jobs[:3]] == [(10.0, "TM"), (10.0, "TE"), (11.0, "TM")]


def test_expand_empty_legacy_axis_gives_no_jobs():
    # an empty list is an axis of length zero, not an absent axis
    assert bulk.expand_jobs(_cfg("[ranges]\npolarization = []\n")) == []


@pytest.mark.parametrize("extra,msg", [
    # BulkConfig::validate (config.rs:1215-1300): legacy ranges
    ("[ranges]\neps_bg = {min = 12.0, max = 10.0, step = 1.0}\n", "eps_bg min > max"),
    ("[ranges]\neps_bg = {min = 10.0, max = 12.0, step = 0.0}\n", "eps_bg step"),
    ("[ranges]\neps_bg = {min = 10.0, max = 12.0, step = -1.0}\n", "eps_bg step"),
    ("[ranges]\nresolution = {min = 2.0, max = 32.0, step = 8.0}\n", "resolution min"),
    ("[ranges]\nresolution = {min = 16.0, max = 32.0, step = 0.5}\n", "resolution step"),
    # SweepSpec::validate (config.rs:338-376)
    ('[[sweeps]]\nparameter = "eps_bg"\nmin = 10.0\nmax = 12.0\nstep = 0.0\n', "step must be positive"),
    ('[[sweeps]]\nparameter = "eps_bg"\nmin = 12.0\nmax = 10.0\nstep = 1.0\n', "min"),
    ('[[sweeps]]\nparameter = "eps_bg"\n', "either"),
    ('[[sweeps]]\nparameter = "eps_bg"\nmin = 1.0\nmax = 2.0\nstep = 1.0\nvalues = [1.0]\n', "both"),
    ('[[sweeps]]\nparameter = "eps_bg"\nvalues = [1.0]\n[[sweeps]]\nparameter = "eps_bg"\nvalues = [2.0]\n',
     "duplicate"),
    ('[[sweeps]]\nparameter = "atomX.radius"\nvalues = [0.2]\n', "invalid sweep parameter"),
    ('[[sweeps]]\nparameter = "atom0.colour"\nvalues = [0.2]\n', "invalid sweep parameter"),
])
def test_invalid_configs_are_rejected(extra, msg):
    with pytest.raises(ValueError, match=msg):
        _cfg(extra)


def test_range_values_rejects_non_positive_step():
    for step in (0.0, -0.5):
        with pytest.raises(ValueError):
            bulk.range_values(0.0, 1.0, step)


def test_expand_ordered_sweeps():
    jobs = bulk.expand_jobs(_cfg('[[sweeps]]\nparameter = "atom0.radius"\nmin = 0.2\nmax = 0.3\nstep = 0.1\n'
                                 '[[sweeps]]\nparameter = "eps_bg"\nmin = 10.0\nmax = 12.0\nstep = 1.0\n'))
    assert len(jobs) == 6
    got = [(round(j.params.atoms[0].radius, 9), j.params.eps_bg) for j in jobs]
    assert got[:4] == [(0.2, 10.0), (0.2, 11.0), (0.2, 12.0), (0.3, 10.0)]
    assert [j.index for j in jobs] == list(range(6))


def test_expand_ordered_discrete_values():
    jobs = bulk.expand_jobs(_cfg('[[sweeps]]\nparameter = "polarization"\nvalues = ["TM", "TE"]\n'
                                 '[[sweeps]]\nparameter = "eps_bg"\nmin = 10.0\nmax = 11.0\nstep = 1.0\n'))
    assert [(j.params.polarization, j.params.eps_bg) for j in jobs] == [("TM", 10.0), ("TM", 11.0), ("TE", 10.0),
                                                                        ("TE", 11.0)]


def test_duplicate_sweep_rejected():
    with pytest.raises(ValueError):
        bulk.expand_jobs(_cfg('[[sweeps]]\nparameter = "eps_bg"\nvalues = [1.0]\n'
                              '[[sweeps]]\nparameter = "eps_bg"\nvalues = [2.0]\n'))


def test_range_values_accumulate_like_the_reference():
    # RangeSpec::values adds the step (no multiplication), keeping the last
    # value within step * 1e-9 of max
    v = bulk.range_values(0.10, 0.30, 0.05)
    assert len(v) == 5 and v[1] == 0.1 + 0.05 and v[-1] == 0.1 + 0.05 + 0.05 + 0.05 + 0.05
    assert bulk.range_values(0.2, 0.4, 0.05)[-1] <= 0.4 + 0.05 * 1e-9
    assert bulk.range_values(1.0, 0.5, 0.1) == []


# --- the reference's example configurations --------------------------------
@pytest.mark.parametrize("name,count", [("bulk_simple_sweep.toml", 10), ("bulk_benchmark_100.toml", 100),
                                        ("bulk_parameter_sweep.toml", 50), ("two_atom_sweep.toml", 243)])
def test_example_configs_expand(name, count):
    cfg = bulk.BulkConfig(bulk.load_config(os.path.join(GOLD, name)))
    jobs = bulk.expand_jobs(cfg)
    assert len(jobs) == count
    assert len({tuple(c for c in j.params.to_columns()) for j in jobs}) == count  # all distinct


def test_simple_sweep_order_and_defaults():
    cfg = bulk.BulkConfig(bulk.load_config(os.path.join(GOLD, "bulk_simple_sweep.toml")))
    jobs = bulk.expand_jobs(cfg)
    # radius outermost, polarization innermost; [defaults] has no eps_bg field,
    # so eps_bg comes from [geometry] (absent there: the 12.0 default)
    assert [j.params.polarization for j in jobs[:4]] == ["TM", "TE", "TM", "TE"]
    assert [round(j.params.atoms[0].radius, 9) for j in jobs[::2]] == [0.2, 0.25, 0.3, 0.35, 0.4]
    assert all(j.params.eps_bg == 12.0 and j.params.resolution == 32 for j in jobs)
    assert jobs[0].job.k_path.shape == (31, 2) and jobs[0].job.n_bands == 8 and jobs[0].job.tol == 1e-4


def test_benchmark_100_lattice_and_default_path():
    cfg = bulk.BulkConfig(bulk.load_config(os.path.join(GOLD, "bulk_benchmark_100.toml")))
    jobs = bulk.expand_jobs(cfg)
    assert [j.params.lattice_type for j in (jobs[0], jobs[50])] == ["square", "triangular"]
    hexa = jobs[50].job.geom.lattice
    assert np.allclose(hexa.a2, (0.5, np.sqrt(3) / 2))
    # no [path]: the default square path with 12 segments per leg for every lattice
    assert all(j.job.k_path.shape == (37, 2) for j in jobs)
    assert jobs[0].job.n_bands == 10


# --- CSV output (output.rs:108-190) -----------------------------------------
def test_rust_float_format():
    assert [bulk.rust_float(x) for x in (1.0, 0.0, 0.25, 1e-7, 0.1 + 0.2, 2.5e20)] == \
        ["1", "0", "0.25", "0.0000001", "0.30000000000000004", "250000000000000000000"]


def test_full_csv_layout(tmp_path):
    cfg = _cfg('[[sweeps]]\nparameter = "eps_bg"\nvalues = [10.0]\n')
    (ej,) = bulk.expand_jobs(cfg)
    kp = np.array([[0.0, 0.0], [0.25, 0.0], [0.5, 0.0]])
    bands = np.array([[0.0, 2 * np.pi], [np.pi, 3 * np.pi], [0.5 * np.pi, 4 * np.pi]])
    from blaze2d_amd.bandstructure import BandStructureResult
    res = BandStructureResult(kp, bands, np.zeros(3, int), np.ones(3, int))
    path = bulk.write_full_csv(str(tmp_path), "job", ej, res)
    assert os.path.basename(path) == "job_000000.csv"
    lines = open(path).read().splitlines()
    assert lines[0] == ("job_index,eps_bg,resolution,polarization,lattice_type,atom0_pos_x,atom0_pos_y,"
                        "atom0_radius,atom0_eps_inside,k_index,kx,ky,k_distance,band1,band2")
    assert lines[1] == "0,10.000000,32,TM,square,0.500000,0.500000,0.300000,1.000000,0,0,0,0,0,1"
    assert lines[2] == "0,10.000000,32,TM,square,0.500000,0.500000,0.300000,1.000000,1,0.25,0,0.25,0.5,1.5"
    assert lines[3].endswith(",2,0.5,0,0.5,0.25,2")


def test_dry_run_cli(capsys):
    assert bulk.main(["--config", os.path.join(GOLD, "bulk_benchmark_100.toml"), "--dry-run"]) == 0
    assert capsys.readouterr().out.strip() == "100 jobs"


# --- end to end on the GPU ---------------------------------------------------
@pytest.mark.gpu
def test_bulk_run_matches_oracle(tmp_path):
    from blaze2d_amd import HipBackend
    from oracle import oracle as O

    cfg = bulk.BulkConfig(bulk.load_config(BASE.replace("nx = 32", "nx = 24").replace("ny = 32", "ny = 24")
                                           .replace("segments_per_leg = 12", "segments_per_leg = 4") +
                                           '[eigensolver]\nn_bands = 6\ntol = 1e-8\n'
                                           '[[sweeps]]\nparameter = "polarization"\nvalues = ["TM", "TE"]\n'
                                           '[[sweeps]]\nparameter = "atom0.radius"\nvalues = [0.2, 0.35]\n'))
    be = HipBackend(0)
    try:
        jobs, results, paths, seconds = bulk.run(be, cfg, str(tmp_path))
        stats = bulk.benchmark(be, cfg)
    finally:
        be.close()
    assert len(paths) == 4 and seconds > 0 and stats["jobs"] == 4
    for ej, p in zip(jobs, paths):
        rows = list(csv.DictReader(open(p)))
        f = np.array([[float(r[f"band{b}"]) for b in range(1, 7)] for r in rows])
        a = ej.params.atoms[0]
        oj = O.Job(lattice="square", eps_bg=ej.params.eps_bg, atoms=[(a.pos[0], a.pos[1], a.radius, a.eps_inside)],
                   nx=24, ny=24, pol=ej.params.polarization, n_bands=6, tol=1e-8, kpath=ej.job.k_path)
        om, _, cv, _ = O.bandstructure(oj)
        want = om / (2 * np.pi)
        m = want > 1e-6
        rel = np.abs(f - want)[m] / want[m]
        assert rel.max() < 1e-6, (ej.index, rel.max())
        assert np.abs(f[~m]).max() < 1e-6
