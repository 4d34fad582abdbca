"""GPU parity of the k-point transitions (band tracking, subspace prediction,
extrapolation, block initialisation), which run on the device.

Band tracking (band_tracking.rs:294-700) decides the order in which each
k-point's frequencies are recorded, so these tests compare the frequencies
unsorted, band by band, with the FP64 oracle run under the same switches.  On
these paths tracking reorders the bands at most k-points (asserted), so a
wrong assignment would show.  The switches cover the kernel's other paths:
no prediction (the tracked eigenvectors start the next solve), no
extrapolation, and no tracking (ascending order)."""
import numpy as np
import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd import bandstructure as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL_OMEGA = 1e-6  # band frequency parity (relative, per band, in recorded order)


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


def _job(lat, pol, nb):
    L = B.Lattice2D.square() if lat == "square" else B.Lattice2D.hexagonal()
    kp = B.k_path(lat, 6)
    job = B.BandStructureJob(B.Geometry2D(L, 13.0, [B.BasisAtom((0.5, 0.5), 0.3, 1.0)]), 32, 32, pol, kp,
                             n_bands=nb, tol=1e-7)
    return job, kp


def _oracle(lat, pol, nb, kp, prediction=True, extrapolation=True, tracking=True):
    oj = O.Job(lattice=lat, eps_bg=13.0, atoms=[(0.5, 0.5, 0.3, 1.0)], nx=32, ny=32, pol=pol, n_bands=nb,
               tol=1e-7, kpath=kp, subspace_prediction=prediction, extrapolation=extrapolation,
               band_tracking=tracking)
    return O.bandstructure(oj)


def _compare_recorded(res, om, label):
    got, ref = np.asarray(res.bands), np.asarray(om)
    assert got.shape == ref.shape, label
    assert res.converged.all(), label
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-8)
    print(f"{label}: max rel omega diff in recorded order {rel.max():.3e}")
    assert rel.max() <= RTOL_OMEGA, (label, rel.max(), np.unravel_index(rel.argmax(), rel.shape))


@pytest.mark.parametrize("lat,pol,nb", [("square", "TE", 10), ("hex", "TE", 10), ("square", "TM", 12),
                                        ("hex", "TM", 12)])
def test_tracked_band_order_matches_oracle(be, lat, pol, nb):
    job, kp = _job(lat, pol, nb)
    om, it, cv, tot = _oracle(lat, pol, nb, kp)
    reordered = int((np.abs(np.sort(om, axis=1) - om) > 1e-12).any(axis=1).sum())
    assert reordered >= 5, f"tracking reorders only {reordered} k-points: the test would not exercise it"
    res = B.run_with_options(be, job)
    _compare_recorded(res, om, f"tracked {lat} {pol} {nb}")


@pytest.mark.parametrize("variant", ["no_prediction", "no_extrapolation", "no_tracking", "none"])
def test_transition_switches_match_oracle(be, variant):
    lat, pol, nb = "hex", "TE", 10
    job, kp = _job(lat, pol, nb)
    pred = variant not in ("no_prediction", "none")
    extrap = variant not in ("no_extrapolation", "none")
    track = variant not in ("no_tracking", "none")
    om, it, cv, tot = _oracle(lat, pol, nb, kp, prediction=pred, extrapolation=extrap, tracking=track)
    if not track:  # ascending order without tracking
        assert np.array_equal(np.sort(om, axis=1), om)
    opts = B.RunOptions(use_subspace_prediction=pred, use_extrapolation=extrap, disable_band_tracking=not track)
    res = B.run_with_options(be, job, opts)
    _compare_recorded(res, om, f"{variant} {lat} {pol} {nb}")


def test_batched_transitions_match_single_runs(be):
    # the four jobs in one batched run (transitions of several slots per
    # launch, slots reused by later jobs) against their own single runs
    specs = [("square", "TE", 10), ("hex", "TE", 10), ("square", "TM", 10), ("hex", "TM", 10)]
    jobs = [_job(*s)[0] for s in specs]
    batch, _ = B.run_batch(be, jobs, slots=2)
    for j, s, r in zip(jobs, specs, batch):
        single = B.run_with_options(be, j)
        assert np.abs(np.asarray(r.bands) - np.asarray(single.bands)).max() <= 1e-9 * np.abs(single.bands).max(), s
