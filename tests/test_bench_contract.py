"""bench.py's host-side contract: one GPU only, and its comparison helpers."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_rel_err_sorts_each_k_point_and_is_relative():
    ref = np.array([[1.0, 2.0, 4.0], [0.0, 1.0, 3.0]])
    got = np.array([[2.0, 1.0, 4.0 * (1 + 1e-7)], [1e-9, 3.0, 1.0]])  # permuted within rows
    e = bench.rel_err(got, ref)
    assert abs(e - 1e-7) < 1e-12  # the zero mode contributes 1e-9 absolute


def test_solves_do_not_count_the_closing_gamma():
    class J:
        def __init__(self, nk):
            self.k_path = np.zeros((nk, 2))

    assert bench.solves_of([J(31), J(61)]) == 30 + 60


def test_worst_diff_skips_unconverged_k_points():
    class R:
        def __init__(self, bands, conv):
            self.bands = np.array(bands)
            self.converged = np.array(conv)

    a = [R([[1.0, 2.0], [1.0, 3.0]], [1, 0])]
    b = [R([[1.0, 2.002], [1.0, 2.0]], [1, 1])]
    w = bench.worst_diff(a, b)
    assert w["k_index"] == 0 and w["band"] == 1 and abs(w["rel_diff"] - 0.002 / 2.002) < 1e-12
