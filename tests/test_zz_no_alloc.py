"""bh_gram_noalloc, the Gram entry point that allocates nothing, against bh_gram.

Both run the same kernels with the same row chunking, so the results must be
identical bit for bit.  Two calls in a row, the second smaller, check that the
shared per-device scratch carries nothing from one call into the next.
"""
import ctypes

import numpy as np
import pytest

from blaze2d_amd import HipBackend
from blaze2d_amd import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


def rnd(shape, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


def gram_both(be, n, p, q, seed, precision=0):
    X, Y = rnd((p, n * n), seed), rnd((q, n * n), seed + 1)
    fx, fy = be.alloc_field(n, n, p, precision=precision), be.alloc_field(n, n, q, precision=precision)
    fx.upload(X)
    fy.upload(Y)
    ref = be.gram_matrix(fx, fy)
    got = np.zeros((p, q), dtype=np.complex128)
    L.check(L.lib().bh_gram_noalloc(be.handle, fx.handle, 0, p, fy.handle, 0, q,
                                    got.view(np.float64).ctypes.data_as(L.c_dp)), "bh_gram_noalloc")
    want = X.conj() @ Y.T
    return got, ref, want


@pytest.mark.parametrize("precision", [0, 1])
def test_gram_noalloc_matches_gram(be, precision):
    tol = 1e-11 if precision == 0 else 1e-5
    # a large call, then a smaller one on another grid: the second must not see
    # the first's partial sums
    for n, p, q, seed in ((48, 48, 37, 3), (16, 5, 3, 9)):
        got, ref, want = gram_both(be, n, p, q, seed, precision)
        assert np.array_equal(got, ref), (n, p, q)
        assert np.abs(got - want).max() <= tol * np.abs(want).max()


def test_gram_noalloc_rejects_oversize(be):
    f = be.alloc_field(8, 8, 49)
    out = np.zeros(2 * 49 * 49)
    st = L.lib().bh_gram_noalloc(be.handle, f.handle, 0, 49, f.handle, 0, 49, out.ctypes.data_as(L.c_dp))
    assert st != 0 and b"48 columns" in L.lib().bh_last_error()
