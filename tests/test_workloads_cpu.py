"""The sub-problem maps the full-size GPU tests check through: a sub-problem cut from
a full input and pushed through the oracle equals the same cut of the oracle's
full output, bit for bit, for every stage of the transform stack (small sizes)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import workloads as W

L, NT, NR, NF = 15, 24, 5, 3
NPHI = 2 * NT
MV = O.m_order_balanced(L)
MPS = [0, 3, 5, len(MV) - 1]
RS = [1, 4]
FS = [0, 2]


@pytest.fixture(scope="module")
def full():
    tabs = O.legendre_tables(NT, L, MV)
    s = O.synthetic_spectral(L, MV, NR, NF, seed=7)
    p = O.stp(NT, L, MV, NR, NF, tabs, s)
    return tabs, s, p


def test_spectral_offsets_match_oracle():
    assert np.array_equal(W.spectral_offsets(L, MV, NR, NF), O.spectral_offsets(L, MV, NR, NF))


def test_stp_pts_sub_problem(full):
    _, s, p = full
    sub_m = MV[MPS]
    tabs_sub = O.legendre_tables(NT, L, sub_m)
    s_sub = s[W.spectral_index(L, MV, NR, NF, MPS, RS, FS)]
    p_sub = O.stp(NT, L, sub_m, len(RS), len(FS), tabs_sub, s_sub)
    assert np.array_equal(p_sub, p[W.hybrid_index(NT, NR, NF, MPS, RS, FS)])
    tabs = full[0]
    s_full = O.pts(NT, L, MV, NR, NF, tabs, p)
    s_sub2 = O.pts(NT, L, sub_m, len(RS), len(FS), tabs_sub, p[W.hybrid_index(NT, NR, NF, MPS, RS, FS)])
    assert np.array_equal(s_sub2, s_full[W.spectral_index(L, MV, NR, NF, MPS, RS, FS)])


def test_permute_fft_sub_problem(full):
    _, _, p = full
    p3 = O.fft_c2r(NPHI, O.permute_2a3a(NT, L, MV, NR, NF, NPHI, p))
    sub = p[W.hybrid_index(NT, NR, NF, range(len(MV)), RS, FS)]
    p3_sub = O.fft_c2r(NPHI, O.permute_2a3a(NT, L, MV, len(RS), len(FS), NPHI, sub))
    assert np.array_equal(p3_sub, p3[W.grid_index(NT, NR, NF, NPHI, RS, FS)])
    p2b = O.permute_3b2b(NT, L, MV, NR, NF, NPHI, O.fft_r2c(NPHI, p3))
    p2b_sub = O.permute_3b2b(NT, L, MV, len(RS), len(FS), NPHI,
                             O.fft_r2c(NPHI, p3[W.grid_index(NT, NR, NF, NPHI, RS, FS)]))
    assert np.array_equal(p2b_sub, p2b[W.hybrid_index(NT, NR, NF, range(len(MV)), RS, FS)])
    # theta rows of the grid: whole lines
    th = [2, 11, 23]
    g = W.grid_index(NT, NR, NF, NPHI, RS, FS, thetas=th).reshape(len(FS), len(th), len(RS), NPHI + 2)
    full_lines = p3.reshape(NF, NT, NR, NPHI + 2)
    assert np.array_equal(p3[g], full_lines[np.ix_(FS, th, RS, np.arange(NPHI + 2))])


def test_radial_sub_problem(full):
    _, s, _ = full
    nlm = int(O.lm_values(L, MV)[0].size)
    p1 = O.permute_2b1b(L, MV, NR, NF, s)
    s_sub = s[W.spectral_index(L, MV, NR, NF, range(len(MV)), RS, FS)]
    p1_sub = O.permute_2b1b(L, MV, len(RS), len(FS), s_sub)
    assert np.array_equal(p1_sub, p1[W.radial_index(nlm, NR, NF, RS, FS)])
    back = O.permute_1a2a(L, MV, len(RS), len(FS), p1[W.radial_index(nlm, NR, NF, RS, FS)])
    assert np.array_equal(back, s_sub)


def test_column_index_cheby():
    n, ncols = 12, 9
    rng = np.random.default_rng(3)
    f = rng.standard_normal(n * ncols)
    c = O.cheby_to_spectral(n, f)
    cols = [0, 4, 8]
    assert np.array_equal(O.cheby_to_spectral(n, f[W.column_index(n, cols)]), c[W.column_index(n, cols)])


def test_batch_problems_grid_rules():
    """l_max + 1 a multiple of 4 so that N_theta = 3 (l_max + 1) / 2 is even
    (ProblemSize.F90:443,719); l_max in [31, 1023]; N_r in {32, 64, 96}; 1-8 fields."""
    probs = W.batch_problems(3)
    assert len(probs) == 64
    for lmax, nt, nr, nf in probs:
        assert (lmax + 1) % 4 == 0 and 31 <= lmax <= 1023
        assert nt == 3 * (lmax + 1) // 2 and nt % 2 == 0
        assert nr in (32, 64, 96) and 1 <= nf <= 8
