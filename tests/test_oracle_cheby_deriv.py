"""The oracle's Chebyshev derivative recursion (Cheby_Deriv_Buffer_4D / d_by_dr_cp and
Cheby_Deriv_OutOf_Buffer, Chebyshev_Polynomials.F90:560-730) pinned against:
  * the compiled reference's own outputs (golden fixtures), bit for bit;
  * the compiled reference live, where it is built here, bit for bit, which also fixes
    that its recursion step rounds the product and the sum separately (a fused
    multiply-add differs in the last bits);
  * closed forms: d/dr of exp(-r) through the Chebyshev transforms (the reference's
    Test_Cheby case) and of a cubic, whose coefficients' derivative is exact;
  * a brute-force derivative of the Chebyshev series T_k(x(r)) on tiny cases.
"""
import numpy as np
import pytest

from oracle import oracle as O

FIXTURES = ["cheby_deriv_n48.npz", "cheby_deriv_n33.npz", "cheby_deriv_n128.npz"]


@pytest.mark.parametrize("name", FIXTURES)
def test_deriv_vs_reference_fixture(golden, name):
    g = golden(name)
    n, rmin, rmax = int(g["n"]), float(g["rmin"]), float(g["rmax"])
    assert np.array_equal(O.cheby_grid(n, rmin, rmax), g["grid"])
    for d in (1, 2, 3):
        assert np.array_equal(O.cheby_deriv(n, rmin, rmax, g["c_in"], d), g[f"d{d}"]), d
        assert np.array_equal(O.cheby_deriv(n, rmin, rmax, g["c_exp"], d), g[f"exp_d{d}"]), d


@pytest.mark.parametrize("name", FIXTURES)
def test_fused_step_is_not_the_reference(golden, name):
    """A planted change (the step as one fused multiply-add) fails the fixture: the
    fixture check discriminates the rounding of the recursion."""
    g = golden(name)
    n, rmin, rmax = int(g["n"]), float(g["rmin"]), float(g["rmax"])
    fused = O.cheby_deriv(n, rmin, rmax, g["c_in"], 2, fused=True)
    assert not np.array_equal(fused, g["d2"])
    assert np.abs(fused - g["d2"]).max() < 1e-14 * np.abs(g["d2"]).max()


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("n,rmin,rmax", [(2, 0.0, 1.0), (3, 0.2, 0.9), (17, 1.0, 3.0), (64, 0.35, 1.0), (255, 0.5, 1.5)])
def test_deriv_vs_compiled_reference(n, rmin, rmax):
    R = O.Reference()
    R.cheby_init(n, rmin, rmax)
    rng = np.random.default_rng(n)
    c = rng.standard_normal(n * 5)
    for d in (1, 2, 3, 4):
        want = R.cheby_deriv(c, d)
        assert np.array_equal(O.cheby_deriv(n, rmin, rmax, c, d), want), d
        assert np.array_equal(R.cheby_deriv(c, d, nobuffer=True), want), d


def spectral_deriv_tol(n, rmin, rmax, d, fmax):
    """Rounding bound of a d-th spectral derivative on n points: the transforms' noise in
    the top coefficients is amplified by about (n^2 deriv_scaling)^d."""
    return 10 * 2.2e-16 * (n * n * O.cheby_deriv_scaling(n, rmin, rmax)) ** d * fmax


@pytest.mark.parametrize("n", [16, 20, 24])
def test_deriv_of_exp_closed_form(n):
    """f = exp(-r): the k-th derivative is (-1)^k f, through to_spectral, the
    recursion and from_spectral (Test_Cheby's check of d_by_dr_cp)."""
    rmin, rmax = 0.35, 1.0
    r = O.cheby_grid(n, rmin, rmax)
    c = O.cheby_to_spectral(n, np.exp(-r))
    for d in (1, 2, 3):
        f = O.cheby_from_spectral(n, O.cheby_deriv(n, rmin, rmax, c, d))
        tol = spectral_deriv_tol(n, rmin, rmax, d, 1.0) + 1e-13
        np.testing.assert_allclose(f, (-1) ** d * np.exp(-r), rtol=0, atol=tol)


def test_deriv_of_cubic_is_exact():
    """f = r^3 on [0.5, 2]: f' = 3 r^2, f'' = 6 r, f''' = 6, f'''' = 0."""
    n, rmin, rmax = 16, 0.5, 2.0
    r = O.cheby_grid(n, rmin, rmax)
    c = O.cheby_to_spectral(n, r ** 3)
    want = {1: 3 * r ** 2, 2: 6 * r, 3: 6 + 0 * r, 4: 0 * r}
    for d, w in want.items():
        f = O.cheby_from_spectral(n, O.cheby_deriv(n, rmin, rmax, c, d))
        np.testing.assert_allclose(f, w, rtol=0, atol=spectral_deriv_tol(n, rmin, rmax, d, 8.0))


def test_deriv_brute_force_series():
    """On a tiny grid, differentiate the series sum' c_k T_k(x) (x the affine image of r)
    directly with numpy's Chebyshev class and compare coefficient-wise after the same
    normalisation: the reference's c_0 carries the half weight of the synthesis."""
    n, rmin, rmax = 7, 0.25, 1.25
    rng = np.random.default_rng(3)
    c = rng.standard_normal(n)
    d = O.cheby_deriv(n, rmin, rmax, c, 1)
    ds = O.cheby_deriv_scaling(n, rmin, rmax)
    # series with the half weight on c_0 (from_spectral's DCT-III convention)
    ser = np.polynomial.chebyshev.Chebyshev(np.concatenate([[c[0] / 2], c[1:]]))
    der = ser.deriv() * ds
    got = np.concatenate([[d[0] / 2], d[1:]])
    np.testing.assert_allclose(got[: n - 1], der.coef, rtol=0, atol=1e-13 * np.abs(der.coef).max())
    assert d[n - 1] == 0.0
