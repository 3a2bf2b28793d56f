"""GPU parity of the batched banded LU factorisation and solve (raysht_lu_decompose_band /
raysht_lu_solve_band: LU_Decompose_band / LU_Solve_Band, Linear_Solve.F90:1041-1086): bit for
bit against the oracle (dgbtf2 / dgbtrs in the same order and rounding) and within rounding of
LAPACK, for band widths from 0 to 96 (kl = ku as the reference sets them, and kl != ku), sizes
around the 256-thread workgroup, a linked (interleaved) system, right-hand-side counts that
differ per matrix (zero included), a singular matrix, and the rejected calls."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


def banded(n, kl, ku, rng):
    a = rng.standard_normal((n, n))
    i, j = np.indices((n, n))
    a[(i - j > kl) | (j - i > ku)] = 0.0
    return a


# bands whose (kl + 1) x (kl + ku + 1) window fits one workgroup's LDS run the LDS-window
# factorisation; (260, 100, 100) and (240, 120, 90) are too wide for it and run the
# global-memory one
@pytest.mark.parametrize("n,kl,ku,nb", [(1, 0, 0, 2), (5, 1, 1, 3), (40, 3, 3, 4), (128, 11, 11, 3),
                                        (256, 32, 32, 2), (300, 96, 96, 1), (77, 4, 9, 2), (64, 9, 2, 2),
                                        (700, 20, 20, 2), (260, 100, 100, 1), (240, 120, 90, 1)])
def test_band_lu_vs_oracle(ra, n, kl, ku, nb):
    import torch
    rng = np.random.default_rng(n + kl)
    mats = [banded(n, kl, ku, rng) for _ in range(nb)]
    abs_ = [ra.band_storage(a, kl, ku) for a in mats]
    ldab = abs_[0].shape[0]
    for ab in abs_:  # garbage in the fill-in rows: the factorisation clears what LAPACK clears
        ab[:kl, :] = np.nan
    d = torch.from_numpy(np.concatenate([ab.ravel(order="F") for ab in abs_])).cuda()
    piv = torch.zeros(nb * n, dtype=torch.int32, device="cuda")
    info = torch.full((nb,), -7, dtype=torch.int32, device="cuda")
    ra.lu_decompose_band(d, piv, info, n, kl, ku)
    f_g = d.cpu().numpy().reshape(nb, n, ldab)
    piv_g, info_g = piv.cpu().numpy().reshape(nb, n), info.cpu().numpy()
    ncols = [(b * 5 + 3) % 37 for b in range(nb)]
    if nb > 2:
        ncols[1] = 0
    offs = np.concatenate([[0], np.cumsum([c * n + 3 for c in ncols])]).astype(np.int64)
    rhs_np = rng.standard_normal(int(offs[-1]) + 2)
    rhs = torch.from_numpy(rhs_np.copy()).cuda()
    ra.lu_solve_band(d, piv, rhs, torch.from_numpy(offs[:-1]).cuda(), torch.tensor(ncols, dtype=torch.int32).cuda(),
                     n, kl, max(max(ncols), 1), ku=ku)
    got, want_all = rhs.cpu().numpy(), rhs_np.copy()
    for b, (a, ab) in enumerate(zip(mats, abs_)):
        f, p, inf = O.band_lu_decompose(ab, kl, ku)
        assert np.array_equal(f_g[b].T, f, equal_nan=True), (n, kl, b)
        assert np.array_equal(piv_g[b], p) and info_g[b] == inf == 0
        _, p_l, _ = lapack.dgbtrf(np.nan_to_num(ab), kl, ku)
        assert np.array_equal(p - 1, p_l)
        if ncols[b]:
            bb = rhs_np[offs[b]: offs[b] + ncols[b] * n].reshape(ncols[b], n).T
            x = O.band_lu_solve(f, kl, ku, p, bb)
            want_all[offs[b]: offs[b] + ncols[b] * n] = x.T.ravel()
            assert np.abs(a @ x - bb).max() <= 1e-10 * n * np.abs(bb).max() * max(1.0, np.linalg.cond(a) / 1e3)
    assert np.array_equal(got, want_all)  # the gaps between right-hand sides are untouched


def test_band_linked_system(ra):
    """Two coupled variables interleaved as Band_Load_Single / Band_Arrange_RHS do it."""
    import torch
    rng = np.random.default_rng(21)
    nd, nl, kl = 48, 2, 7
    a = np.block([[banded(nd, 3, 3, rng) + 4 * np.eye(nd), banded(nd, 2, 2, rng)],
                  [banded(nd, 2, 2, rng), banded(nd, 3, 3, rng) + 4 * np.eye(nd)]])
    ab = ra.band_storage(a, kl, kl, nlinks=nl)
    p = ra.band_interleave(nd, nl)
    d = torch.from_numpy(ab.ravel(order="F").copy()).cuda()
    piv = torch.zeros(nd * nl, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ra.lu_decompose_band(d, piv, info, nd * nl, kl)
    b = rng.standard_normal((nd * nl, 4))
    rhs = torch.from_numpy(b[p].T.ravel().copy()).cuda()
    ra.lu_solve_band(d, piv, rhs, torch.zeros(1, dtype=torch.int64).cuda(), torch.tensor([4], dtype=torch.int32).cuda(),
                     nd * nl, kl, 4)
    xb = rhs.cpu().numpy().reshape(4, nd * nl).T
    x = np.empty_like(xb)
    x[p] = xb
    assert int(info.item()) == 0 and np.abs(a @ x - b).max() < 1e-10
    f, pv, _ = O.band_lu_decompose(ab, kl, kl)
    assert np.array_equal(xb, O.band_lu_solve(f, kl, kl, pv, b[p]))


def test_band_singular_info(ra):
    import torch
    rng = np.random.default_rng(1)
    a = banded(30, 2, 2, rng)
    a[:, 9] = 0.0
    ab = ra.band_storage(a, 2, 2)
    d = torch.from_numpy(np.concatenate([ab.ravel(order="F"), ra.band_storage(np.eye(30), 2, 2).ravel(order="F")])).cuda()
    piv = torch.zeros(60, dtype=torch.int32, device="cuda")
    info = torch.zeros(2, dtype=torch.int32, device="cuda")
    ra.lu_decompose_band(d, piv, info, 30, 2)
    _, _, inf = O.band_lu_decompose(ab, 2, 2)
    assert info.cpu().numpy().tolist() == [inf, 0] and inf == 10


def test_band_rejects_bad_calls(ra):
    import torch
    from rayleigh_amd._lib import RayshtError
    piv = torch.zeros(8, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(RayshtError):  # ldab < 2 kl + ku + 1
        ra.lu_decompose_band(torch.zeros(8 * 4, dtype=torch.float64, device="cuda"), piv, info, 8, 2, ldab=4)
    with pytest.raises(RayshtError):  # kl + ku > 1024
        ra.lu_decompose_band(torch.zeros(8 * 1540, dtype=torch.float64, device="cuda"), piv, info, 8, 513, ku=513,
                             ldab=1540)
    with pytest.raises(TypeError):
        ra.lu_decompose_band(torch.zeros(8 * 7, dtype=torch.float64, device="cuda"), piv.long(), info, 8, 2)
