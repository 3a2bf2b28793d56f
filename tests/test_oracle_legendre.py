"""Pins the C restatement's Legendre part (nodes, P_lm tables, StP, PtS) against
the reference's own outputs (golden fixtures written by the compiled reference),
closed forms and identities.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O

CASES = ["legendre_L41_nt64.npz", "legendre_L63_nt96.npz", "legendre_L63_partial.npz",
         "legendre_L511_nt768_sampled.npz"]


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name", CASES)
def test_nodes_bit_exact_vs_reference(golden, name):
    g = golden(name)
    xh, xl, wh, wl = O.gauss_legendre(int(g["nt"]))
    # Find_Colocation's Real*16 values, head and tail, reproduced exactly
    assert np.array_equal(xh, g["x_hi"]) and np.array_equal(xl, g["x_lo"])
    assert np.array_equal(wh, g["w_hi"]) and np.array_equal(wl, g["w_lo"])


@pytest.mark.parametrize("name", CASES)
def test_tables_bit_exact_vs_reference(golden, name):
    g = golden(name)
    if int(g["nt"]) >= 768:
        pytest.importorskip("numpy")
    t = O.legendre_tables(int(g["nt"]), int(g["lmax"]), g["mvals"])
    for key, arr in zip(("p_even", "p_odd", "ip_even", "ip_odd"), t):
        if key in g:
            assert np.array_equal(arr, g[key]), key
        else:
            assert np.array_equal(arr[g[key + "_idx"]], g[key + "_val"]), key


def test_nodes_against_numpy_and_invariants():
    for n in (8, 64, 96, 384):
        xh, xl, wh, wl = O.gauss_legendre(n)
        x, w = np.polynomial.legendre.leggauss(n)
        assert np.abs(xh - x).max() < 1e-15
        # the reference's weight uses the derivative at the previous Newton iterate:
        # weights agree with the exact ones only to ~1e-14
        assert np.abs(wh - w).max() < 5e-14
        assert abs((wh + wl).sum() - 2.0) < 1e-12
        assert np.array_equal(xh, -xh[::-1]) and np.array_equal(wh, wh[::-1])
        assert np.all(np.diff(xh) > 0) and xh[0] < 0  # ascending from the south pole


def test_tables_closed_forms():
    """Y_00, Y_10, Y_11, Y_20, Y_21, Y_22 (Legendre_Transforms.F90:73-78; Test_SHT.F90:83)."""
    nt, L = 16, 3
    mv = np.arange(L + 1, dtype=np.int32)
    pe, po, ie, io = O.legendre_tables(nt, L, mv)
    x = O.gauss_legendre(nt)[0][: nt // 2]
    s = np.sqrt(1 - x * x)
    nt2 = nt // 2

    def p(l, m):
        # de-normalise p_lm_even/odd (STP norm 1 for m = 0, 1/2 otherwise)
        off_e = sum(O.n_even(L, mm) for mm in range(m)) * nt2
        off_o = sum(O.n_odd(L, mm) for mm in range(m)) * nt2
        k = (l - m) // 2
        if (l - m) % 2 == 0:
            ne = O.n_even(L, m)
            v = pe[off_e: off_e + ne * nt2].reshape(nt2, ne)[:, k]
        else:
            no = O.n_odd(L, m)
            v = po[off_o: off_o + no * nt2].reshape(nt2, no)[:, k]
        return v / (1.0 if m == 0 else 0.5)

    pi = np.pi
    np.testing.assert_allclose(p(0, 0), np.sqrt(1 / (4 * pi)), rtol=4e-15)
    np.testing.assert_allclose(p(1, 0), np.sqrt(3 / (4 * pi)) * x, rtol=4e-15)
    np.testing.assert_allclose(p(1, 1), -np.sqrt(3 / (8 * pi)) * s, rtol=4e-15)
    np.testing.assert_allclose(p(2, 0), np.sqrt(5 / (16 * pi)) * (3 * x * x - 1), rtol=1e-14)
    np.testing.assert_allclose(p(2, 1), np.sqrt(5 / (24 * pi)) * (-3 * x * s), rtol=4e-15)
    np.testing.assert_allclose(p(2, 2), np.sqrt(5 / (96 * pi)) * 3 * s * s, rtol=4e-15)
    np.testing.assert_allclose(p(3, 3), -np.sqrt(35 / (64 * pi)) * s ** 3, rtol=1e-14)


def test_tables_match_scipy_lpmv():
    from math import factorial

    from scipy.special import lpmv
    nt, L = 48, 31
    mv = np.arange(L + 1, dtype=np.int32)
    pe, po, ie, io = O.legendre_tables(nt, L, mv)
    x = O.gauss_legendre(nt)[0][: nt // 2]
    nt2 = nt // 2
    oe = oo = 0
    for m in mv:
        ne, no = O.n_even(L, m), O.n_odd(L, m)
        E = pe[oe: oe + ne * nt2].reshape(nt2, ne)
        Od = po[oo: oo + no * nt2].reshape(nt2, no)
        norm = 1.0 if m == 0 else 0.5
        for l in range(m, L + 1):
            # lpmv carries the Condon-Shortley phase; orthonormal with int P^2 dx = 1/(2 pi)
            ref = lpmv(m, l, x) * np.sqrt((2 * l + 1) / (4 * np.pi) * factorial(l - m) / factorial(l + m))
            got = (E[:, (l - m) // 2] if (l - m) % 2 == 0 else Od[:, (l - m) // 2]) / norm
            assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
        oe += ne * nt2
        oo += no * nt2


def test_discrete_orthogonality():
    nt, L = 64, 41
    mv = O.m_order_balanced(L)
    pe, po, ie, io = O.legendre_tables(nt, L, mv)
    nt2 = nt // 2
    oe = oo = 0
    for m in mv:
        stp, ptsn = (1.0, 1.0 / (2 * nt)) if m == 0 else (0.5, 1.0 / nt)
        for (tab, itab, n, off) in ((pe, ie, O.n_even(L, m), oe), (po, io, O.n_odd(L, m), oo)):
            if n == 0:
                continue
            P = tab[off: off + n * nt2].reshape(nt2, n)          # [j][k]
            IP = itab[off: off + n * nt2].reshape(n, nt2)        # [k][j]
            G = IP @ P * 2.0 / (stp * ptsn)                        # both hemispheres
            np.testing.assert_allclose(G, np.eye(n), atol=1e-12)
        oe += O.n_even(L, m) * nt2
        oo += O.n_odd(L, m) * nt2


@pytest.mark.parametrize("name", ["legendre_L41_nt64.npz", "legendre_L63_nt96.npz", "legendre_L63_partial.npz"])
def test_stp_pts_vs_reference(golden, name):
    g = golden(name)
    nt, L, mv, nr, nf = int(g["nt"]), int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    t = O.legendre_tables(nt, L, mv)
    s = O.synthetic_spectral(L, mv, nr, nf, int(g["seed"]))
    assert np.array_equal(s, g["s_in"])  # the fixture's input is the seeded (l+1)^-2 distribution
    p = O.stp(nt, L, mv, nr, nf, t, s)
    assert rel(p, g["p_out"]) < 1e-14
    s2 = O.pts(nt, L, mv, nr, nf, t, g["p_out"])
    assert rel(s2, g["s_rt"]) < 1e-14


def test_planted_reflection_mistake_is_caught(golden):
    """A StP that reflects the odd part with the wrong sign must fail the golden check."""
    g = golden("legendre_L63_nt96.npz")
    nt, L, mv, nr, nf = int(g["nt"]), int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    pe, po, ie, io = O.legendre_tables(nt, L, mv)
    bad = O.stp(nt, L, mv, nr, nf, (pe, -po, ie, io), g["s_in"])  # E - O north, E + O south
    assert rel(bad, g["p_out"]) > 1e-3


def test_round_trip_identity():
    """PtS(StP(a)) = a / N_phi for band-limited a (STP.PTS = 1/(2 N_theta))."""
    nt, L, nr, nf = 96, 63, 3, 2
    mv = O.m_order_balanced(L)
    t = O.legendre_tables(nt, L, mv)
    s = O.synthetic_spectral(L, mv, nr, nf, 3)
    s2 = O.pts(nt, L, mv, nr, nf, t, O.stp(nt, L, mv, nr, nf, t, s))
    assert rel(s2 * 2 * nt, s) < 1e-13


def test_single_mode_known_answer():
    """Test_SHT.F90:47-200: (l,m) = (2,1) with amplitudes (2f+1) r and -(3f^2+2) r^2."""
    nt, L, nr, nf = 64, 41, 4, 3
    mv = O.m_order_balanced(L)
    t = O.legendre_tables(nt, L, mv)
    radius = np.linspace(0.5, 1.5, nr)
    off = O.spectral_offsets(L, mv, nr, nf)
    s = np.zeros(off[-1])
    mp = int(np.where(mv == 1)[0][0])
    blk = s[off[mp]: off[mp + 1]].reshape(nf, 2, nr, L)  # l = 1..41
    for f in range(1, nf + 1):
        blk[f - 1, 0, :, 2 - 1] = (2 * f + 1) * radius * 2.0      # x 1/to_phys_norm
        blk[f - 1, 1, :, 2 - 1] = -(3 * f * f + 2) * radius ** 2 * 2.0
    p = O.stp(nt, L, mv, nr, nf, t, s).reshape(len(mv), 2 * nr * nf, nt)
    x = O.gauss_legendre(nt)[0]
    true = np.sqrt(5 / (24 * np.pi)) * (-3 * x * np.sqrt(1 - x * x))
    for f in range(1, nf + 1):
        for r in range(nr):
            np.testing.assert_allclose(p[mp, r + 2 * nr * (f - 1)], (2 * f + 1) * radius[r] * true, atol=1e-13)
            np.testing.assert_allclose(p[mp, r + nr + 2 * nr * (f - 1)], -(3 * f * f + 2) * radius[r] ** 2 * true,
                                       atol=1e-13)
    others = np.delete(p, mp, axis=0)
    assert np.abs(others).max() == 0.0
    s2 = O.pts(nt, L, mv, nr, nf, t, p.ravel()) * nt  # x 1/to_spec_norm (m > 0)
    assert np.abs(s2 - s / 2.0).max() < 1e-10  # the reference's own tolerance
