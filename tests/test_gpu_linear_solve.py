"""GPU parity of the batched dense LU factorisation and solve (raysht_lu_decompose_full /
raysht_lu_solve_full: LU_Decompose_full / LU_Solve_full, Linear_Solve.F90:1034-1107):
bit for bit against the oracle (the same unblocked algorithm and rounding), and within
rounding of LAPACK, for sizes around the 256-thread workgroup and the 16-column chunks,
right-hand-side counts that differ per matrix (zero included), singular matrices, and the
rejected calls."""
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


def batch(n, nb, seed):
    rng = np.random.default_rng(seed)
    mats = [rng.standard_normal((n, n)) + (n / 4) * np.eye(n) * rng.uniform(-1, 1) for _ in range(nb)]
    return mats, rng


def to_dev(mats):
    import torch
    return torch.from_numpy(np.concatenate([np.asfortranarray(m).ravel(order="F") for m in mats])).cuda()


@pytest.mark.parametrize("n,nb", [(1, 3), (2, 2), (17, 5), (128, 4), (256, 3), (300, 2), (513, 1)])
def test_lu_decompose_and_solve_vs_oracle(ra, n, nb):
    import torch
    mats, rng = batch(n, nb, n)
    d = to_dev(mats)
    piv = torch.zeros(nb * n, dtype=torch.int32, device="cuda")
    info = torch.full((nb,), -7, dtype=torch.int32, device="cuda")
    ra.lu_decompose_full(d, piv, info, n)
    lu_g = d.cpu().numpy().reshape(nb, n, n)  # [b][col][row]
    piv_g, info_g = piv.cpu().numpy().reshape(nb, n), info.cpu().numpy()
    ncols = [(b * 7 + 1) % 41 for b in range(nb)]  # 1, 8, 15, 22, ... (0 for some sizes)
    if nb > 2:
        ncols[1] = 0
    offs = np.concatenate([[0], np.cumsum([c * n + 5 for c in ncols])]).astype(np.int64)  # gaps of 5
    rhs_np = rng.standard_normal(int(offs[-1]) + 3)
    rhs = torch.from_numpy(rhs_np.copy()).cuda()
    ra.lu_solve_full(d, piv, rhs, torch.from_numpy(offs[:-1]).cuda(), torch.tensor(ncols, dtype=torch.int32).cuda(),
                     n, max(ncols))
    got = rhs.cpu().numpy()
    want_all = rhs_np.copy()
    for b, a in enumerate(mats):
        lu, p, inf = O.lu_decompose(a)
        assert np.array_equal(lu_g[b].T, lu), (n, b)
        assert np.array_equal(piv_g[b], p) and info_g[b] == inf
        _, p_l, _ = lapack.dgetrf(a)
        assert np.array_equal(p - 1, p_l)
        if ncols[b]:
            bb = rhs_np[offs[b]: offs[b] + ncols[b] * n].reshape(ncols[b], n).T
            x = O.lu_solve(lu, p, bb)
            want_all[offs[b]: offs[b] + ncols[b] * n] = x.T.ravel()
            assert np.abs(a @ x - bb).max() <= 1e-11 * n * np.abs(bb).max() * max(1.0, np.linalg.cond(a) / 1e3)
    assert np.array_equal(got, want_all)  # the gaps between right-hand sides are untouched


def test_lu_singular_info(ra):
    import torch
    n = 40
    a = np.random.default_rng(1).standard_normal((n, n))
    a[:, 7] = 0.0
    d = to_dev([a, np.eye(n)])
    piv = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    info = torch.zeros(2, dtype=torch.int32, device="cuda")
    ra.lu_decompose_full(d, piv, info, n)
    _, _, inf = O.lu_decompose(a)
    assert info.cpu().numpy().tolist() == [inf, 0] and inf == 8


@pytest.mark.parametrize("n", [24, 100])  # the unblocked kernel, and the blocked one (n > 32)
def test_lu_nan_diagonal_keeps_row(ra, n):
    """idamax never ranks an entry above a NaN a_kk, so dgetrf keeps row k as the pivot
    there (and passes over NaNs below the diagonal); the factors then match the oracle."""
    import torch
    rng = np.random.default_rng(5)
    a = rng.standard_normal((n, n)) + 4 * np.eye(n)
    a[3, 3] = np.nan        # NaN pivot candidate with finite entries below it
    b = rng.standard_normal((n, n)) + 4 * np.eye(n)
    b[9, 2] = np.nan        # NaN below a finite diagonal
    d = to_dev([a, b])
    piv = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    info = torch.zeros(2, dtype=torch.int32, device="cuda")
    ra.lu_decompose_full(d, piv, info, n)
    lu_g = d.cpu().numpy().reshape(2, n, n)
    piv_g = piv.cpu().numpy().reshape(2, n)
    for k, m in enumerate((a, b)):
        lu, p, _ = O.lu_decompose(m)
        assert np.array_equal(piv_g[k], p), k
        assert np.array_equal(lu_g[k].T, lu, equal_nan=True), k
    assert piv_g[0][3] == 4  # row k kept (1-based)


def test_lu_rejects_bad_calls(ra):
    import torch
    from rayleigh_amd._lib import RayshtError
    d = torch.zeros(4, dtype=torch.float64, device="cuda")
    piv = torch.zeros(4, dtype=torch.int32, device="cuda")
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    with pytest.raises(RayshtError):
        ra.lu_decompose_full(torch.zeros(1025 * 1025, dtype=torch.float64, device="cuda"),
                             torch.zeros(1025, dtype=torch.int32, device="cuda"), info, 1025)
    with pytest.raises(TypeError):
        ra.lu_decompose_full(d, piv.long(), info, 2)
    with pytest.raises(ValueError):
        ra.lu_decompose_full(d[:3], piv, info, 2)
