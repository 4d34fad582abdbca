"""One-time operator diagnostics on the GPU (bh_operator_diagnostics) against
the oracle (bandstructure.rs:437-498, maxwell.rs:382-582), the closed form of
the TM power iteration, and batch invariance, through the C ABI.

Tolerances: the power-iteration Rayleigh quotients are sums over N nodes
accumulated in a different order on the GPU, so they agree to 1e-10
relative; the self-adjointness errors are both rounding-level (< 1e-13)."""
import numpy as np
import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd import bandstructure as B
from blaze2d_amd.sweep import res64_sweep
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-10


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


def _jobs(lattice, pol, res, atoms, eps_bg, kp, smoothing=True):
    lat = B.Lattice2D.square() if lattice == "square" else B.Lattice2D.hexagonal()
    g = B.Geometry2D(lat, eps_bg, [B.BasisAtom((a[0], a[1]), a[2], a[3]) for a in atoms])
    gj = B.BandStructureJob(g, res, res, pol, np.asarray(kp, float), smoothing=smoothing)
    oj = O.Job(lattice=lattice, eps_bg=eps_bg, atoms=atoms, nx=res, ny=res, pol=pol, smoothing=smoothing,
               kpath=np.asarray(kp, float))
    return gj, oj


CASES = [
    ("square", "TM", 24, [(0.5, 0.5, 0.3, 1.0)], 13.0),
    ("square", "TE", 32, [(0.5, 0.5, 0.3, 1.0)], 13.0),
    ("hex", "TM", 32, [(0.5, 0.5, 0.35, 1.0)], 12.0),
    ("hex", "TE", 24, [(0.5, 0.5, 0.2, 3.5), (0.1, 0.8, 0.15, 1.5)], 9.0),
]


@pytest.mark.parametrize("lattice,pol,res,atoms,eps_bg", CASES)
def test_diagnostics_match_oracle(be, lattice, pol, res, atoms, eps_bg):
    kp = [[0.0, 0.0], [0.15, 0.05], [0.3, 0.1]]
    gj, oj = _jobs(lattice, pol, res, atoms, eps_bg, kp)
    (d,), ms = B.operator_diagnostics(be, [gj])
    o = O.operator_diagnostics(oj)
    assert ms > 0 and d.k_index == o["k_index"] == 1 and d.precond == o["precond"]
    for key in ("lambda_max", "pm_lambda_max", "kappa", "pm_kappa"):
        assert abs(getattr(d, key) - o[key]) <= RTOL * abs(o[key]), key
    assert abs(d.lambda_min - o["lambda_min"]) <= 1e-14 * o["lambda_min"]
    assert d.self_adjoint_err < 1e-13 and o["self_adjoint_err"] < 1e-13
    assert d.self_adjoint_status == "exact"


def test_diagnostics_tm_closed_form_64(be):
    # Theta_TM is diagonal in Fourier space: the power iteration's quotient
    # after 15 steps is sum w d^29 / sum w d^28, w = |F x|^2, d = |k+G|^2
    kp = [[0.2, 0.1]]
    gj, oj = _jobs("square", "TM", 64, [(0.5, 0.5, 0.3, 1.0)], 13.0, kp)
    (d,), _ = B.operator_diagnostics(be, [gj])
    x, _ = O.diagnostic_start_vectors(64 * 64)
    w = np.abs(np.fft.fft2(x.reshape(64, 64))).reshape(-1) ** 2
    dd = O.kpg_sq(oj, kp[0])
    r = dd / dd.max()
    want = dd.max() * np.sum(w * r ** 29) / np.sum(w * r ** 28)
    assert abs(d.lambda_max - want) <= RTOL * want
    assert d.lambda_min == pytest.approx(dd.min(), rel=1e-14)


def test_diagnostics_gamma_only_path(be):
    gj, _ = _jobs("square", "TE", 24, [(0.5, 0.5, 0.3, 1.0)], 13.0, [[0.0, 0.0], [0.0, 0.0]])
    (d,), _ = B.operator_diagnostics(be, [gj])
    assert d.k_index == -1 and d.lambda_max == 0.0 and d.kappa == 0.0


def test_diagnostics_mixed_batch_matches_single_calls(be):
    # polarisations, preconditioners (auto / none / the other one) and
    # smoothing mixed in one call: each job as if run alone
    kp = [[0.0, 0.0], [0.1, 0.0], [0.2, 0.2]]
    specs = [("square", "TM", True, -1), ("hex", "TE", True, -1), ("square", "TE", False, -1),
             ("hex", "TM", False, 0), ("square", "TE", True, 1), ("hex", "TM", True, 2)]
    jobs = []
    for lat, pol, sm, pc in specs:
        gj, _ = _jobs(lat, pol, 24, [(0.5, 0.5, 0.3, 1.0)], 11.0, kp, smoothing=sm)
        jobs.append((gj, pc))
    batch = []
    for gj, pc in jobs:  # RunOptions.precond applies to a whole call: one call per setting
        (d,), _ = B.operator_diagnostics(be, [gj], B.RunOptions(precond=pc))
        batch.append(d)
    by_pc = {}
    for i, (gj, pc) in enumerate(jobs):
        by_pc.setdefault(pc, []).append(i)
    for pc, idx in by_pc.items():
        ds, _ = B.operator_diagnostics(be, [jobs[i][0] for i in idx], B.RunOptions(precond=pc))
        for i, d in zip(idx, ds):
            assert d == batch[i]
    assert batch[3].precond == 0 and batch[3].pm_kappa == 0.0
    assert batch[4].precond == 1 and batch[5].precond == 2


def test_diagnostics_sweep_sample(be):
    # 256 sweep band structures (TM and TE, square and hex, 64 x 64) in one call
    jobs = res64_sweep(128, tol=1e-7)
    ds, ms = B.operator_diagnostics(be, jobs)
    assert len(ds) == len(jobs) and ms > 0
    for d, j in zip(ds, jobs):
        assert d.k_index == 1  # the closed path starts at Gamma
        assert d.self_adjoint_status == "exact"
        assert np.isfinite(d.kappa) and d.kappa > 1
        assert d.pm_kappa != 0 and abs(d.pm_kappa) < d.kappa
    # two of them against the oracle at full size
    for i in (0, 1):
        j = jobs[i]
        lat = "square" if abs(j.geom.lattice.a2[0]) < 1e-12 else "hex"
        oj = O.Job(lattice=lat, eps_bg=j.geom.eps_bg,
                   atoms=[(a.pos[0], a.pos[1], a.radius, a.eps_inside) for a in j.geom.atoms],
                   nx=64, ny=64, pol=j.pol, kpath=j.k_path)
        o = O.operator_diagnostics(oj)
        assert abs(ds[i].lambda_max - o["lambda_max"]) <= RTOL * o["lambda_max"]
        assert abs(ds[i].pm_lambda_max - o["pm_lambda_max"]) <= RTOL * abs(o["pm_lambda_max"])
    print(f"diagnostics: {len(jobs)} jobs in {ms:.2f} ms of device time")
