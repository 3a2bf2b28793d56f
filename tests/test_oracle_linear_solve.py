"""The oracle's dense LU factorisation and solve (LU_Decompose_full / LU_Solve_full,
Linear_Solve.F90:1034-1107: LAPACK dgetrf and dgetrs 'N', which the reference links)
restated as the unblocked dgetf2 algorithm and reference-BLAS triangular solves, pinned
against LAPACK itself (scipy.linalg.lapack, the OpenBLAS 0.3.29 LAPACK that numpy ships
here, the same library the compiled reference links): identical pivots, L\\U and solutions
to rounding, identical info on singular matrices; and against what the algebra fixes
(P A = L U, A x = b)."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from oracle import oracle as O


def radial_like(n, rng):
    """A Chebyshev-collocation-like operator: identity minus a dense smooth block, two
    boundary rows, rows of very different scale (what the implicit solve factors)."""
    x = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    a = np.eye(n) - 1e-3 * np.outer(1 + x ** 2, np.cos(np.arange(n))) * rng.uniform(0.5, 2, n)
    a[0, :] = 1.0
    a[-1, :] = np.cos(np.arange(n) * np.pi)
    return a


@pytest.mark.parametrize("n", [1, 2, 5, 64, 200])
@pytest.mark.parametrize("kind", ["random", "radial"])
def test_lu_matches_lapack(n, kind):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) if kind == "random" else radial_like(n, rng)
    lu, piv, info = O.lu_decompose(a)
    lu_l, piv_l, info_l = lapack.dgetrf(a)
    assert info == info_l == 0
    assert np.array_equal(piv - 1, piv_l)
    assert np.abs(lu - lu_l).max() <= 1e-13 * max(1.0, np.abs(lu_l).max())
    # P A = L U
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    pa = a.copy()
    for k, p in enumerate(piv - 1):
        pa[[k, p]] = pa[[p, k]]
    assert np.abs(L @ U - pa).max() <= 1e-13 * np.abs(a).max() * n
    b = rng.standard_normal((n, 3))
    x = O.lu_solve(lu, piv, b)
    x_l, info_s = lapack.dgetrs(lu_l, piv_l, b)
    assert info_s == 0
    assert np.abs(x - x_l).max() <= 1e-12 * np.abs(x_l).max() * n
    assert np.abs(a @ x - b).max() <= 1e-12 * np.abs(b).max() * n * np.linalg.cond(a)


def test_singular_info_matches_lapack():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((6, 6))
    a[:, 3] = 0.0  # zero columns stay exactly zero under elimination: zero pivots at 4 and 6
    a[:, 5] = 0.0
    lu, piv, info = O.lu_decompose(a)
    lu_l, piv_l, info_l = lapack.dgetrf(a)
    assert info == info_l > 0
    assert np.array_equal(piv - 1, piv_l)


def test_planted_pivot_rule_fails():
    """Choosing the last (not the first) row among equal |a_ik| gives other pivots than
    LAPACK on a matrix with ties: the pivot check above would catch that rule."""
    a = np.array([[1.0, 2.0, 0.0], [-1.0, 0.5, 1.0], [1.0, -3.0, 2.0]])
    lu, piv, _ = O.lu_decompose(a)
    _, piv_l, _ = lapack.dgetrf(a)
    assert np.array_equal(piv - 1, piv_l) and piv[0] == 1  # ties |1| = |-1| = |1|: the first row


def test_nan_diagonal_pivot_matches_lapack():
    """A NaN a_kk with finite entries below it: idamax ranks nothing above the NaN, so
    LAPACK keeps row k; the oracle agrees up to that column (after it the NaN spreads over
    the trailing block and comparisons of NaNs decide nothing)."""
    n, k = 24, 3
    rng = np.random.default_rng(5)
    a = rng.standard_normal((n, n)) + 4 * np.eye(n)
    a[k, k] = np.nan
    _, piv, _ = O.lu_decompose(a)
    _, piv_l, _ = lapack.dgetrf(a)
    assert np.array_equal(piv[: k + 1] - 1, piv_l[: k + 1]) and piv[k] == k + 1
