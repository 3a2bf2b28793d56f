"""The oracle's banded LU factorisation and solve (LU_Decompose_band / LU_Solve_Band,
Linear_Solve.F90:1041-1086: LAPACK dgbtrf and dgbtrs 'N' with kl = ku = (ldab - 1) / 3),
restated as the unblocked dgbtf2 and reference-BLAS dger / dtbsv order, pinned against LAPACK
itself (scipy.linalg.lapack, the LAPACK numpy ships here: the library the compiled reference
links): identical pivots and info, factors and solutions equal to rounding; against the dense
LU of the same matrix (the same pivots: partial pivoting never looks outside the band); and
against what the algebra fixes (A x = b).  The host band storage of the package
(Band_Load_Single, with the interleave of linked variables) is pinned by A x computed from the
band against the full matrix."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from oracle import oracle as O


def banded(n, kl, ku, rng, dominance=0.0):
    a = rng.standard_normal((n, n)) + dominance * np.eye(n)
    i, j = np.indices((n, n))
    a[(i - j > kl) | (j - i > ku)] = 0.0
    return a


def storage(a, kl, ku):
    import rayleigh_amd as ra
    return ra.band_storage(a, kl, ku)


@pytest.mark.parametrize("n,kl,ku", [(1, 0, 0), (2, 1, 1), (7, 2, 2), (40, 3, 3), (64, 11, 11), (200, 9, 9),
                                     (130, 33, 33), (50, 4, 7)])
def test_band_lu_matches_lapack(n, kl, ku):
    rng = np.random.default_rng(n * 100 + kl)
    a = banded(n, kl, ku, rng)
    ab = storage(a, kl, ku)
    f, piv, info = O.band_lu_decompose(ab, kl, ku)
    lu_l, piv_l, info_l = lapack.dgbtrf(ab, kl, ku)
    assert info == info_l == 0
    assert np.array_equal(piv - 1, piv_l)
    assert np.abs(f - lu_l).max() <= 1e-13 * max(1.0, np.abs(lu_l).max()) * max(1, kl)
    b = rng.standard_normal((n, 3))
    x = O.band_lu_solve(f, kl, ku, piv, b)
    x_l, info_s = lapack.dgbtrs(lu_l, kl, ku, b, piv_l)
    assert info_s == 0
    assert np.abs(x - x_l).max() <= 1e-11 * np.abs(x_l).max() * n
    assert np.abs(a @ x - b).max() <= 1e-11 * np.abs(b).max() * n * np.linalg.cond(a)
    # dense partial pivoting on the same matrix picks the same rows
    _, piv_d, _ = O.lu_decompose(a)
    assert np.array_equal(piv_d, piv)


def test_band_singular_info_matches_lapack():
    rng = np.random.default_rng(4)
    a = banded(12, 2, 2, rng)
    a[:, 5] = 0.0
    ab = storage(a, 2, 2)
    _, piv, info = O.band_lu_decompose(ab, 2, 2)
    _, piv_l, info_l = lapack.dgbtrf(ab, 2, 2)
    assert info == info_l == 6 and np.array_equal(piv - 1, piv_l)


def test_band_storage_and_interleave():
    import rayleigh_amd as ra
    rng = np.random.default_rng(8)
    nd, nl, kl = 10, 2, 5
    # two coupled variables, each banded in its own block: interleaving makes the whole banded
    blocks = [[banded(nd, 2, 2, rng) for _ in range(nl)] for _ in range(nl)]
    a = np.block(blocks)
    p = ra.band_interleave(nd, nl)
    assert sorted(p) == list(range(nd * nl)) and p[1] == nd and p[2] == 1
    ab = ra.band_storage(a, kl, kl, nlinks=nl)
    x = rng.standard_normal(nd * nl)
    ap = a[np.ix_(p, p)]
    i, j = np.indices(ap.shape)
    assert not np.any(ap[np.abs(i - j) > kl])  # banded after the interleave
    # A x from the band storage equals the full product
    y = np.zeros(nd * nl)
    for c in range(nd * nl):
        for r in range(max(0, c - kl), min(nd * nl, c + kl + 1)):
            y[r] += ab[2 * kl + r - c, c] * x[c]
    assert np.allclose(y, ap @ x, rtol=0, atol=1e-12)
    f, piv, info = O.band_lu_decompose(ab, kl, kl)
    b = rng.standard_normal(nd * nl)
    xb = O.band_lu_solve(f, kl, kl, piv, b[p])
    xs = np.empty_like(xb)
    xs[p] = xb
    assert info == 0 and np.allclose(a @ xs, b, rtol=0, atol=1e-10)
