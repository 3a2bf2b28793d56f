"""Pins the C restatement's FFT, row-permute and Chebyshev parts.  CPU only.

FFT: FFTW (the reference's FFT library) is not part of the reference tree, so the
restated r2c/c2r are pinned by numpy.fft (pocketfft), a brute-force DFT and
Parseval.  Chebyshev: the reference's own tables and transforms (golden fixtures
from the compiled reference), scipy's DCT-II/III and the round trip."""
import numpy as np
import pytest
import scipy.fft

from oracle import oracle as O


def half_complex_lines(rng, nphi, nlines, mmax=None):
    x = np.zeros((nlines, nphi + 2))
    h = nphi // 2
    top = h if mmax is None else mmax
    x[:, : 2 * (top + 1)] = rng.standard_normal((nlines, 2 * (top + 1)))
    return x


@pytest.mark.parametrize("nphi", [6, 12, 96, 192, 768])
def test_c2r_matches_numpy(nphi):
    rng = np.random.default_rng(nphi)
    x = half_complex_lines(rng, nphi, 5)
    y = O.fft_c2r(nphi, x).reshape(5, nphi + 2)
    z = x[:, 0::2] + 1j * x[:, 1::2]
    ref = np.fft.irfft(z, n=nphi, axis=1) * nphi  # unnormalised, Im of m=0 and m=n/2 dropped
    np.testing.assert_allclose(y[:, :nphi], ref, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("nphi", [6, 12, 96, 192, 768])
def test_r2c_matches_numpy(nphi):
    rng = np.random.default_rng(nphi + 1)
    x = np.zeros((4, nphi + 2))
    x[:, :nphi] = rng.standard_normal((4, nphi))
    y = O.fft_r2c(nphi, x).reshape(4, nphi + 2)
    ref = np.fft.rfft(x[:, :nphi], axis=1)
    np.testing.assert_allclose(y[:, 0::2], ref.real, atol=1e-13 * nphi)
    np.testing.assert_allclose(y[:, 1::2], ref.imag, atol=1e-13 * nphi)
    # the _RSC variant divides x(3:) by n and x(1) by 2n, x(2) untouched
    yr = O.fft_r2c(nphi, x, rescale=True).reshape(4, nphi + 2)
    np.testing.assert_allclose(yr[:, 2:], y[:, 2:] / nphi, rtol=1e-15)
    np.testing.assert_allclose(yr[:, 0], y[:, 0] / (2 * nphi), rtol=1e-15)
    assert np.array_equal(yr[:, 1], y[:, 1])


def test_brute_force_dft_and_parseval():
    nphi = 18
    rng = np.random.default_rng(0)
    x = np.zeros(nphi + 2)
    x[:nphi] = rng.standard_normal(nphi)
    y = O.fft_r2c(nphi, x)
    k = np.arange(nphi)
    for m in range(nphi // 2 + 1):
        v = np.sum(x[:nphi] * np.exp(-2j * np.pi * m * k / nphi))
        assert abs(y[2 * m] - v.real) < 1e-13 and abs(y[2 * m + 1] - v.imag) < 1e-13
    X = y[0::2] + 1j * y[1::2]
    e_spec = (abs(X[0]) ** 2 + abs(X[-1]) ** 2 + 2 * np.sum(np.abs(X[1:-1]) ** 2)) / nphi
    assert abs(e_spec - np.sum(x[:nphi] ** 2)) < 1e-12
    back = O.fft_c2r(nphi, y)
    np.testing.assert_allclose(back[:nphi], x[:nphi] * nphi, atol=1e-12)


def test_permutes_index_maps_and_inverse():
    nt, L, nr, nf = 8, 5, 3, 2
    nphi = 2 * nt
    mv = O.m_order_balanced(L)
    nrhs = 2 * nr * nf
    p2 = np.arange(nt * nrhs * len(mv), dtype=np.float64) + 1.0
    p3 = O.permute_2a3a(nt, L, mv, nr, nf, nphi, p2).reshape(nf, nt, nr, nphi + 2)
    P2 = p2.reshape(len(mv), nrhs, nt)
    for mp, m in enumerate(mv):
        for f in range(nf):
            for r in range(nr):
                assert np.array_equal(p3[f, :, r, 2 * m], P2[mp, r + 2 * nr * f])
                assert np.array_equal(p3[f, :, r, 2 * m + 1], P2[mp, r + nr + 2 * nr * f])
    assert np.all(p3[..., 2 * (L + 1):] == 0)  # m > l_max zero-filled (implicit de-aliasing)
    back = O.permute_3b2b(nt, L, mv, nr, nf, nphi, p3.ravel())
    assert np.array_equal(back, p2)
    # 3b->2b reads only m <= l_max: garbage above is truncated
    p3g = p3.copy()
    p3g[..., 2 * (L + 1):] = 7.0
    assert np.array_equal(O.permute_3b2b(nt, L, mv, nr, nf, nphi, p3g.ravel()), p2)


@pytest.mark.parametrize("name", ["cheby_n48.npz", "cheby_n33.npz"])
def test_cheby_vs_reference(golden, name):
    g = golden(name)
    n = int(g["n"])
    assert np.array_equal(O.cheby_grid(n, float(g["rmin"]), float(g["rmax"])), g["grid"])
    ce, co = O.cheby_tables(n)
    assert np.array_equal(ce, g["ceven"]) and np.array_equal(co[: g["codd"].size], g["codd"][: co.size])
    c = O.cheby_to_spectral(n, g["f_in"])
    assert np.abs(c - g["c_out"]).max() < 1e-15 * np.abs(g["c_out"]).max() * 8
    f = O.cheby_from_spectral(n, g["c_out"])
    assert np.abs(f - g["f_rt"]).max() < 1e-15 * np.abs(g["f_rt"]).max() * 8


@pytest.mark.parametrize("n", [16, 48, 128])
def test_cheby_matches_scipy_dct(n):
    rng = np.random.default_rng(n)
    f = rng.standard_normal((3, n))
    c = O.cheby_to_spectral(n, f.ravel()).reshape(3, n)
    ref = scipy.fft.dct(f, type=2, axis=1) / n  # (2/N) sum f(x_i) T_k(x_i) on the zeros grid
    np.testing.assert_allclose(c, ref, atol=1e-12)
    # synthesis: f_i = sum' c_k T_k(x_i) (half weight on c_0) = DCT-III
    f2 = O.cheby_from_spectral(n, ref.ravel()).reshape(3, n)
    np.testing.assert_allclose(f2, f, atol=1e-12)


def test_cheby_dealias():
    n = 48
    c = np.ones(n * 2)
    d = O.cheby_dealias(n, c).reshape(2, n)
    rda = (2 * n) // 3 + 1
    assert np.all(d[:, : rda - 1] == 1) and np.all(d[:, rda - 1:] == 0)


@pytest.mark.parametrize("n", [33, 48, 128, 256])
def test_cheby_exact_sums_vs_table_form(n):
    """The long-double sums on the reference's angles (the check for the DCT form) agree with
    the table form, itself pinned to the compiled reference above, to the table form's own
    rounding (rounded k theta_i and cos in every entry: up to about 1e-14 of max |c|); the
    angles are the accumulated zeros grid, within 1e-13 of (i + 1/2) pi / n."""
    rng = np.random.default_rng(n)
    f = rng.standard_normal(8 * n)
    c = O.cheby_to_spectral_exact(n, f)
    assert np.abs(c - O.cheby_to_spectral(n, f)).max() < 2e-14 * np.abs(c).max()
    back = O.cheby_from_spectral_exact(n, c)
    assert np.abs(back - O.cheby_from_spectral(n, c)).max() < 2e-14 * np.abs(back).max()
    th = O.cheby_angles_longdouble(n)
    exact = (np.arange(n) + 0.5) * np.pi / n
    assert np.abs(th.astype(np.float64) - exact).max() < 1e-13
    assert th[0] == np.float64(np.pi / n * 0.5)  # the first angle is dtheta / 2 itself


@pytest.mark.parametrize("n", [16, 64, 256])
def test_cheby_exact_sums_on_exact_grid_are_dct(n):
    """On the exact zeros-grid angles the long-double sums are scipy's DCT-II / DCT-III (a
    wrong scale, a dropped half weight on c_0 or a mirrored angle fails this)."""
    rng = np.random.default_rng(n + 1)
    f = rng.standard_normal((3, n))
    th = (np.arange(n, dtype=np.longdouble) + np.longdouble(0.5)) * np.longdouble(
        "3.14159265358979323846264338327950288") / n
    c = O.cheby_to_spectral_exact(n, f.ravel(), theta=th).reshape(3, n)
    ref = scipy.fft.dct(f, type=2, axis=1) / n
    assert np.abs(c - ref).max() < 1e-15 * np.abs(ref).max()
    f2 = O.cheby_from_spectral_exact(n, ref.ravel(), theta=th).reshape(3, n)
    assert np.abs(f2 - f).max() < 1e-14 * np.abs(f).max()
