"""The oracle's angular spectral derivatives (Spectral_Derivatives.F90: d_by_dtheta,
d_by_sdtheta, d_by_dphi on the rmcontainer4d layout) pinned against:
  * the compiled reference's outputs (golden fixtures), bit for bit, for the buffer
    forms and the 4dbuff2arr forms, with full, balanced and partial m lists (m = l_max
    included);
  * the compiled reference live where it is built here;
  * the textbook relations, evaluated with scipy's associated Legendre functions at
    arbitrary colatitudes: for orthonormal harmonics
      sin(t) d/dt Y_lm            = l c_{l+1,m} Y_{l+1,m} - (l+1) c_{lm} Y_{l-1,m},
      1/sin(t) d/dt (sin^2 t Y_lm) = 2 cos(t) Y_lm + sin(t) d/dt Y_lm,
      d/dphi (A e^{i m phi})        = i m A e^{i m phi},
    with c_{lm} = sqrt((l^2 - m^2) / (4 l^2 - 1)) (a series with A(l_max) = 0 has no
    l_max + 1 part to drop).
"""
import math

import numpy as np
import pytest
import scipy.special

from oracle import oracle as O

FIXTURES = ["angular_L12_ascending.npz", "angular_L41_balanced.npz", "angular_L9_partial.npz"]


@pytest.mark.parametrize("name", FIXTURES)
def test_angular_vs_reference_fixture(golden, name):
    g = golden(name)
    L, mv, nr = int(g["lmax"]), g["mvals"], int(g["nr"])
    for kind in ("dtheta", "sdtheta", "dphi"):
        got = O.angular_deriv(kind, L, mv, nr, g["buf"], 2, 0, out=g["buf"].copy(), nf_out=2, fout=1)
        assert np.array_equal(got, g[f"{kind}_buffer"]), kind
        assert np.array_equal(O.angular_deriv(kind, L, mv, nr, g["buf"], 2, 1), g[f"{kind}_arr"]), kind


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("L,kind_m", [(0, "ascending"), (1, "ascending"), (30, "balanced"), (17, "partial")])
def test_angular_vs_compiled_reference(L, kind_m):
    mv = {"ascending": np.arange(L + 1, dtype=np.int32), "balanced": O.m_order_balanced(L),
          "partial": np.array([17, 2, 9, 0], dtype=np.int32)}[kind_m]
    nr, nf = 4, 3
    R = O.Reference()
    R.angular_init(L, mv, nr)
    buf = np.random.default_rng(L).standard_normal(O.spectral_size(L, mv, nr, nf))
    for kind in ("dtheta", "sdtheta", "dphi"):
        want = R.angular_deriv(kind, buf, nf, 3, 1)
        got = O.angular_deriv(kind, L, mv, nr, buf, nf, 2, out=buf.copy(), nf_out=nf, fout=0)
        assert np.array_equal(got, want), kind
    # d_by_dphi in place (from_ind = to_ind) is the reference's in-place branch
    want = R.angular_deriv("dphi", buf, nf, 2, 2)
    got = O.angular_deriv("dphi", L, mv, nr, buf, nf, 1, out=buf.copy(), nf_out=nf, fout=1)
    assert np.array_equal(got, want)


def ylm(l, m, t):
    """Orthonormal Y_lm(t, 0) with the Condon-Shortley phase (scipy's lpmv)."""
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.exp(math.lgamma(l - m + 1) - math.lgamma(l + m + 1)))
    return norm * scipy.special.lpmv(m, l, np.cos(t))


@pytest.mark.parametrize("m", [0, 1, 4])
def test_theta_derivatives_match_calculus(m):
    L, nr = 14, 1
    t = np.linspace(0.2, 2.9, 23)
    a = np.random.default_rng(m).standard_normal(L - m + 1)
    a[-1] = 0.0  # no l_max part, so nothing beyond l_max is dropped
    buf = np.concatenate([a, np.zeros(L - m + 1)])  # one radius, (re, im) with im = 0
    mv = np.array([m], dtype=np.int32)

    def series(c):
        return sum(c[l - m] * ylm(l, m, t) for l in range(m, L + 1))

    h = 1e-5
    f = series(a)
    df = (sum(a[l - m] * ylm(l, m, t + h) for l in range(m, L + 1)) -
          sum(a[l - m] * ylm(l, m, t - h) for l in range(m, L + 1))) / (2 * h)
    b = O.angular_deriv("dtheta", L, mv, nr, buf, 1, 0)[: L - m + 1]
    np.testing.assert_allclose(series(b), np.sin(t) * df, rtol=0, atol=1e-7 * np.abs(f).max() * L)
    s = O.angular_deriv("sdtheta", L, mv, nr, buf, 1, 0)[: L - m + 1]
    np.testing.assert_allclose(series(s), 2 * np.cos(t) * f + np.sin(t) * df, rtol=0, atol=1e-7 * np.abs(f).max() * L)


def test_coefficients_are_the_textbook_ones():
    for l, m in [(0, 0), (1, 0), (5, 3), (40, 17), (511, 200)]:
        c = lambda ll: math.sqrt(max(ll * ll - m * m, 0) / (4 * ll * ll - 1)) if ll > 0 else 0.0
        d1, d2 = O.angular_coefs(l, m, "dtheta")
        s1, s2 = O.angular_coefs(l, m, "sdtheta")
        assert math.isclose(d1, (l - 1) * c(l), rel_tol=1e-15, abs_tol=1e-300)
        assert math.isclose(d2, -(l + 2) * c(l + 1), rel_tol=1e-15, abs_tol=1e-300)
        assert math.isclose(s1, (l + 1) * c(l), rel_tol=1e-15, abs_tol=1e-300)
        assert math.isclose(s2, -l * c(l + 1), rel_tol=1e-15, abs_tol=1e-300)


def test_phi_derivative_is_i_m():
    L, nr = 6, 2
    mv = np.array([0, 3, 6], dtype=np.int32)
    buf = np.random.default_rng(2).standard_normal(O.spectral_size(L, mv, nr, 1))
    out = O.angular_deriv("dphi", L, mv, nr, buf, 1, 0)
    offs = O.spectral_offsets(L, mv, nr, 1)
    for mp, m in enumerate(mv):
        nl = L - int(m) + 1
        A = buf[offs[mp]: offs[mp + 1]].reshape(2, nr, nl)
        B = out[offs[mp]: offs[mp + 1]].reshape(2, nr, nl)
        z = 1j * int(m) * (A[0] + 1j * A[1])
        assert np.array_equal(B[0], z.real + 0.0) and np.array_equal(B[1], z.imag + 0.0)
