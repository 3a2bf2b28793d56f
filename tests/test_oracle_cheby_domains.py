"""The oracle's multi-domain Chebyshev grid (Initialize_Cheby_Grid with ndomains > 1 and the
domain loops of To/From_Spectral4D, Cheby_Deriv_Buffer_4D / _OutOf_Buffer and dealias_buffer4d:
Chebyshev_Polynomials.F90:65-207,353-730), pinned against:
  * the compiled reference live (where it is built here): the grid, the integration weights,
    deriv_scaling and rda bit for bit, with and without dealias_by; the derivatives of orders
    1-4 (both forms) and the dealias bit for bit; the transforms within the tolerance of the
    one-domain oracle against the reference's DGEMM (its summation order is the BLAS's);
  * what the mathematics fixes: a one-domain grid through the domain code is the one-domain
    oracle; every domain is an independent expansion, so a polynomial of degree < npoly in r
    round-trips on every domain and its derivative comes out on every domain's points, which
    pins each domain's map and scaling (a planted exchange of two domains' scalings fails).
From_Spectral4D of the reference reads f_temp(hoff + n_x) (:544-546) past the end of f_temp
(1:n_x) for a domain with an odd number of points after the first; those grids are compared
with the reference only through To_Spectral, the derivatives and the dealias.
"""
import numpy as np
import pytest

from oracle import oracle as O

# (npoly, bounds, dealias_by), ascending as the reference takes them; the domains are
# stored outermost first, so odd npoly appear only in the last entry here except in ODD_INNER
CASES = [
    ([16, 24, 32], [0.35, 0.5, 0.8, 1.0], None),
    ([12, 12, 12, 12], [1.0, 2.0, 3.0, 4.0, 5.0], [3, 4, 0, 11]),
    ([20, 33], [0.4, 0.7, 1.3], [5, 0]),
    ([64, 48], [0.35, 0.6, 1.0], None),
]
ODD_INNER = ([21, 33, 17], [0.4, 0.7, 1.3, 1.5], None)
TOL = 8e-15  # of max |result|, as the one-domain oracle against the reference's DGEMM


def _ref_or_skip():
    if not O.ref_available():
        pytest.skip("compiled reference not built here")
    return O.Reference()


@pytest.mark.parametrize("npoly,bounds,db", CASES + [ODD_INNER])
def test_domains_grid_and_derivatives_vs_reference(npoly, bounds, db):
    R = _ref_or_skip()
    g = O.cheby_domains(npoly, bounds, db)
    grid, w = R.cheby_init_domains(npoly, bounds, db)
    ds, rda = R.cheby_grid_arrays()
    assert np.array_equal(grid, g["grid"]) and np.array_equal(w, g["weights"])
    assert np.array_equal(ds, g["dscale"]) and list(rda) == g["rda"]
    c = np.random.default_rng(sum(npoly)).standard_normal(g["ntotal"] * 9)
    for k in (1, 2, 3, 4):
        assert np.array_equal(O.cheby_domains_deriv(g, c, k), R.cheby_deriv(c, k)), k
        assert np.array_equal(O.cheby_domains_deriv(g, c, k), R.cheby_deriv(c, k, nobuffer=True)), k
    assert np.array_equal(O.cheby_domains_dealias(g, c), R.cheby_dealias(c))
    ref = R.cheby_to_spectral(c)
    assert np.abs(O.cheby_domains_to_spectral(g, c) - ref).max() < TOL * np.abs(ref).max()


@pytest.mark.parametrize("npoly,bounds,db", CASES)
def test_domains_from_spectral_vs_reference(npoly, bounds, db):
    R = _ref_or_skip()
    g = O.cheby_domains(npoly, bounds, db)
    R.cheby_init_domains(npoly, bounds, db)
    c = np.random.default_rng(len(npoly)).standard_normal(g["ntotal"] * 9)
    ref = R.cheby_from_spectral(c)
    assert np.abs(O.cheby_domains_from_spectral(g, c) - ref).max() < TOL * np.abs(ref).max()


def test_one_domain_is_the_one_domain_oracle():
    n, lo, hi = 48, 0.35, 1.0
    g = O.cheby_domains([n], [lo, hi])
    assert np.array_equal(g["grid"], O.cheby_grid(n, lo, hi))
    assert np.array_equal(g["weights"], O.cheby_weights(n, lo, hi))
    c = np.random.default_rng(0).standard_normal(n * 5)
    assert np.array_equal(O.cheby_domains_to_spectral(g, c), O.cheby_to_spectral(n, c))
    assert np.array_equal(O.cheby_domains_deriv(g, c, 2), O.cheby_deriv(n, lo, hi, c, 2))
    assert np.array_equal(O.cheby_domains_dealias(g, c), O.cheby_dealias(n, c))


def _poly(r):
    return 0.3 + r * (1.1 + r * (-0.7 + r * (0.2 + 0.05 * r)))  # degree 4


def _dpoly(r):
    return 1.1 + r * (-1.4 + r * (0.6 + 0.2 * r))


@pytest.mark.parametrize("npoly,bounds,db", CASES + [ODD_INNER])
def test_polynomials_per_domain(npoly, bounds, db):
    """p(r) of degree 4 < every npoly: round trip and d/dr on every domain's points."""
    g = O.cheby_domains(npoly, bounds, db)
    f = _poly(g["grid"])
    c = O.cheby_domains_to_spectral(g, f)
    assert np.abs(O.cheby_domains_from_spectral(g, c) - f).max() < 1e-13
    d = O.cheby_domains_from_spectral(g, O.cheby_domains_deriv(g, c, 1))
    dp = _dpoly(g["grid"])
    assert np.abs(d - dp).max() < 1e-14 * max(npoly) ** 2 * np.abs(dp).max()  # rounding grows ~n^2 at the ends
    # every domain's coefficients past degree 4 vanish
    for n, o in zip(g["n"], g["hoff"]):
        assert np.abs(c[o + 5:o + n]).max() < 1e-13
    # a planted exchange of the first two domains' scalings breaks the derivative
    bad = dict(g, lo=[g["lo"][1], g["lo"][0]] + g["lo"][2:], hi=[g["hi"][1], g["hi"][0]] + g["hi"][2:])
    if bad["hi"][0] - bad["lo"][0] != g["hi"][0] - g["lo"][0]:
        d_bad = O.cheby_domains_from_spectral(g, O.cheby_domains_deriv(bad, c, 1))
        assert np.abs(d_bad - _dpoly(g["grid"])).max() > 1e-6


def test_grid_order_and_bounds():
    npoly, bounds = [16, 24, 32], [0.35, 0.5, 0.8, 1.0]
    g = O.cheby_domains(npoly, bounds)
    assert g["n"] == [32, 24, 16] and g["hoff"] == [0, 32, 56]
    # descending radii, outermost domain first; a domain's end points are its bounds, so
    # adjacent domains share the radius of their interface
    for n, o, lo, hi in zip(g["n"], g["hoff"], g["lo"], g["hi"]):
        r = g["grid"][o:o + n]
        assert np.all(np.diff(r) < 0) and r[0] == hi and r[-1] == lo
