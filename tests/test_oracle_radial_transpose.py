"""The oracle's one-rank radial transposes (Transpose_2b1b / Transpose_1a2a,
Spherical_Buffer.F90:705-932) and the (l, m) pair list they follow (LM_Load_Balance,
Load_Balance.F90:311-395).  The reference's buffers need its MPI parallel framework,
so they are not compiled here; the oracle is pinned by what the reference states and
the mathematics fixes:
  * the pair list holds every (l, m) with m <= l of the owned m exactly once, l slowest
    from the smallest m, m in the owned (m_values) order within one l;
  * for the default one-rank m order the list is the closed form
    [(0,0), (1,0), (1,1), (2,0), ...] and has (l_max+1)(l_max+2)/2 entries;
  * 2b->1b moves every coefficient to exactly one radial slot (a permutation of unique
    values), each radial column holds the N_r radii of one (l, m, imi, field) in
    order, and 1a->2a undoes it.
"""
import numpy as np
import pytest

from oracle import oracle as O

M_SETS = {
    "ascending": lambda L: np.arange(L + 1, dtype=np.int32),
    "balanced": lambda L: O.m_order_balanced(L),
    # a non-contiguous m list in no particular order (one rank of a row)
    "partial": lambda L: np.array(list(dict.fromkeys(m for m in (L, 0, L // 2, 3, 1) if m <= L)), dtype=np.int32),
}


@pytest.mark.parametrize("kind", sorted(M_SETS))
@pytest.mark.parametrize("L", [0, 1, 7, 40])
def test_lm_list_structure(kind, L):
    mv = M_SETS[kind](L)
    ls, mps = O.lm_values(L, mv)
    ms = mv[mps]
    pairs = set(zip(ls.tolist(), ms.tolist()))
    want = {(l, int(m)) for m in mv for l in range(int(m), L + 1)}
    assert pairs == want and len(ls) == len(want)
    assert np.all(np.diff(ls) >= 0)
    assert ls[0] == mv.min() and ls[-1] == L
    for l in np.unique(ls):
        sel = mps[ls == l]
        assert np.all(np.diff(sel) > 0)  # m in m_values order within one l


def test_lm_list_closed_form_ascending():
    L = 5
    ls, mps = O.lm_values(L, np.arange(L + 1, dtype=np.int32))
    want = [(l, m) for l in range(L + 1) for m in range(l + 1)]
    assert list(zip(ls.tolist(), mps.tolist())) == want
    assert len(ls) == (L + 1) * (L + 2) // 2


@pytest.mark.parametrize("kind", sorted(M_SETS))
def test_2b1b_is_the_stated_permutation(kind):
    L, nr, nf = 9, 5, 3
    mv = M_SETS[kind](L)
    n = O.spectral_size(L, mv, nr, nf)
    s = np.arange(1, n + 1, dtype=np.float64)  # unique values
    p1 = O.permute_2b1b(L, mv, nr, nf, s)
    assert np.array_equal(np.sort(p1), s)
    ls, mps = O.lm_values(L, mv)
    offs = O.spectral_offsets(L, mv, nr, nf)
    P = p1.reshape(nf, len(ls), 2, nr)
    for k, (l, mp) in enumerate(zip(ls, mps)):
        nl = L - int(mv[mp]) + 1
        blk = s[offs[mp]: offs[mp] + nl * nr * 2 * nf].reshape(nf, 2, nr, nl)
        assert np.array_equal(P[:, k], blk[:, :, :, l - int(mv[mp])])
    assert np.array_equal(O.permute_1a2a(L, mv, nr, nf, p1), s)
