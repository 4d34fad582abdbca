"""The batched Rayleigh-Ritz kernel (bh_rayleigh_ritz) against the oracle's SVQB.

The kernel B-orthonormalises the subspace by Cholesky QR only when that
provably keeps every direction SVQB would keep (normalization.rs:281-544 drops
Gram eigenvalues below 1e-12 of the largest); otherwise it runs SVQB itself.
Three subspaces exercise the three outcomes:
  * well conditioned: Cholesky path, full rank;
  * a Kahan-type block whose scaled Cholesky pivots all exceed 1e-8 while its
    smallest Gram eigenvalue is far below 1e-12 of the largest: the pivot test
    alone would take the Cholesky path; the norm bound sends it to SVQB, which
    drops the direction as the reference does;
  * an exactly repeated column: a zero pivot, SVQB, one direction dropped.
Reference: the oracle's svqb_orthonormalize restatement on the vectors
themselves, then the projected operator diagonalised by numpy.
"""
import numpy as np
import pytest

from blaze2d_amd import HipBackend
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    b = HipBackend(0)
    yield b
    b.close()


def make_case(kind, n, length, seed):
    r = np.random.default_rng(seed)
    q, _ = np.linalg.qr(r.standard_normal((length, n)) + 1j * r.standard_normal((length, n)))
    w = r.uniform(1.0, 3.0, length)        # B: positive diagonal (a TM mass)
    d = r.uniform(0.5, 40.0, length)       # A: positive diagonal operator
    if kind == "random":
        z = q @ (np.eye(n) + 0.3 * (r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))))
    elif kind == "kahan":
        # Z = Q R with B-orthonormal Q, so Z^H B Z = R^H R, and R unit upper
        # triangular with -1 above the diagonal (condition ~ 2^n): the scaled
        # Cholesky pivots are 1/(j+1) while lambda_min / lambda_max ~ 1e-15 at n = 24
        z = (q / np.sqrt(w)[:, None]) @ (np.eye(n) - np.triu(np.ones((n, n)), 1))
    elif kind == "kahan_wide":
        # the n = 24 Kahan block beside a well-conditioned block: the wide
        # kernel (order above 48) must also take the SVQB path and drop one direction
        rb = np.eye(n, dtype=complex)
        rb[:24, :24] -= np.triu(np.ones((24, 24)), 1)
        rb[24:, 24:] += 0.3 * (r.standard_normal((n - 24, n - 24)) + 1j * r.standard_normal((n - 24, n - 24)))
        z = (q / np.sqrt(w)[:, None]) @ rb
    else:  # repeated column
        z = q @ (np.eye(n) + 0.3 * (r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))))
        z[:, n - 1] = z[:, 3]
    return z, w, d


def scaled_gram_facts(g1):
    """(smallest Cholesky pivot, lambda_min / lambda_max) of the unit-diagonal Gram."""
    s = 1.0 / np.sqrt(np.real(np.diag(g1)))
    gs = g1 * s[:, None] * s[None, :]
    n = gs.shape[0]
    a = gs.copy()
    piv = []
    for j in range(n):  # plain right-looking Cholesky, pivots d_j
        dj = np.real(a[j, j])
        piv.append(dj)
        if dj <= 0:
            break
        a[j + 1:, j] /= np.sqrt(dj)
        a[j + 1:, j + 1:] -= np.outer(a[j + 1:, j], a[j + 1:, j].conj())
    ev = np.linalg.eigvalsh(gs)
    return min(piv), ev[0] / ev[-1]


@pytest.mark.parametrize("kind,n,expect_svqb,expect_rank", [
    ("random", 36, 0, 36),
    ("kahan", 24, 1, 23),
    ("repeat", 30, 1, 29),
    # orders above 48: the wide kernel (T, eigenvectors and G2 T in device workspace)
    ("random", 96, 0, 96),
    ("random", 61, 0, 61),
    ("kahan_wide", 80, 1, 79),
    ("repeat", 90, 1, 89),
])
def test_rayleigh_ritz_matches_oracle_svqb(be, kind, n, expect_svqb, expect_rank):
    nsplit = 12 if n <= 48 else 32
    z, w, d = make_case(kind, n, 600, 11 + n)
    g1 = z.conj().T @ (w[:, None] * z)
    g2 = z.conj().T @ (d[:, None] * z)
    pmin, ratio = scaled_gram_facts(g1)
    if kind in ("kahan", "kahan_wide"):  # the case the pivot test alone gets wrong
        assert pmin > 1e-8 and ratio < 1e-12, (pmin, ratio)
    lam, coef, info = be.rayleigh_ritz(g1, g2, nsplit)
    rank, svqb, cx, cw = (int(v) for v in info[0])
    assert svqb == expect_svqb and rank == expect_rank, info[0]
    # oracle: svqb_orthonormalize on the vectors, then Q^H A Q
    xq, bxq, orank = O.svqb(z.T, (w[:, None] * z).T)
    assert orank == rank
    qv = xq[:orank]
    m = qv.conj() @ (d[None, :] * qv).T
    ref = np.linalg.eigvalsh(0.5 * (m + m.conj().T))
    sc = np.abs(ref).max()
    assert np.abs(lam[0, :rank] - ref).max() <= 1e-10 * sc
    assert cx == min(rank, nsplit) and cw == min(2 * nsplit, rank) - cx
    # the returned coefficients give B-orthonormal Ritz vectors with those Ritz values
    x = z @ coef[0][:, :cx + cw]
    gb = x.conj().T @ (w[:, None] * x)
    ga = x.conj().T @ (d[:, None] * x)
    assert np.abs(gb - np.eye(cx + cw)).max() <= 1e-9
    assert np.abs(ga - np.diag(lam[0, :cx + cw])).max() <= 1e-9 * sc
    assert np.all(coef[0][:, cx + cw:] == 0)
