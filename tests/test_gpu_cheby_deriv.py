"""GPU parity of the radial derivative recursion (raysht_cheb_deriv: d_by_dr_cp /
d_by_dr_nobuffer, Chebyshev_Polynomials.F90:560-730): bit for bit against the compiled
reference's outputs (golden fixtures) and against the oracle on ragged and large cases,
orders 1-4; the reference's buffer form (ind -> dind) and the rejected calls."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import dev, empty, host

pytestmark = pytest.mark.gpu

FIXTURES = ["cheby_deriv_n48.npz", "cheby_deriv_n33.npz", "cheby_deriv_n128.npz"]


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


@pytest.mark.parametrize("name", FIXTURES)
def test_deriv_vs_reference_fixture(ra, golden, name):
    g = golden(name)
    n = int(g["n"])
    grid = ra.ChebyGrid(n, float(g["rmin"]), float(g["rmax"]))
    assert np.array_equal(grid.grid, g["grid"])
    for src, key in (("c_in", "d"), ("c_exp", "exp_d")):
        c = dev(g[src])
        for d in (1, 2, 3):
            out = empty(c.numel())
            grid.d_by_dr(c, out, d)
            assert np.array_equal(host(out), g[f"{key}{d}"]), (name, src, d)


@pytest.mark.parametrize("n,ncols", [(2, 3), (3, 1), (16, 256), (17, 257), (33, 255), (128, 1000), (255, 513), (512, 300)])
def test_deriv_vs_oracle_ragged(ra, n, ncols):
    """Column counts around the 256-column workgroup, radial sizes around the
    16-coefficient chunk, the largest grid the library takes (512)."""
    rmin, rmax = 0.35, 1.0
    grid = ra.ChebyGrid(n, rmin, rmax)
    c_np = np.random.default_rng(n * 1000 + ncols).standard_normal(n * ncols)
    c = dev(c_np)
    for d in (1, 2, 3, 4):
        out = empty(c.numel())
        out.fill_(float("nan"))
        grid.d_by_dr(c, out, d)
        assert np.array_equal(host(out), O.cheby_deriv(n, rmin, rmax, c_np, d)), (n, ncols, d)


def test_deriv_stack_size_sampled(ra):
    """The transform-stack size (N_r = 128, 2 x 33024 x 8 columns of l_max = 255):
    every column checked against the oracle, and a linearity property on the whole."""
    n, rmin, rmax = 128, 0.35 / 0.65, 1.0 / 0.65
    ncols = 2 * 33024 * 8
    grid = ra.ChebyGrid(n, rmin, rmax)
    rng = np.random.default_rng(7)
    a_np = rng.standard_normal(n * ncols)
    a = dev(a_np)
    out = empty(a.numel())
    grid.d_by_dr(a, out, 2)
    got = host(out)
    assert np.array_equal(got, O.cheby_deriv(n, rmin, rmax, a_np, 2))
    # d(2a) = 2 d(a) exactly (scaling by 2 commutes with every rounding)
    out2 = empty(a.numel())
    grid.d_by_dr(a * 2.0, out2, 2)
    assert np.array_equal(host(out2), 2.0 * got)


def test_d_by_dr_cp_buffer_form(ra):
    """buffer(0:n-1, n2, n3, nvar) in the reference is (nvar, n3, n2, n) here; d_by_dr_cp(1, 3, buf, 2)
    writes only variable 3."""
    import torch
    n, n2, n3, nvar = 48, 3, 2, 4
    rmin, rmax = 0.35, 1.0
    grid = ra.ChebyGrid(n, rmin, rmax)
    b_np = np.random.default_rng(11).standard_normal((nvar, n3, n2, n))
    buf = torch.from_numpy(b_np.copy()).cuda()
    grid.d_by_dr_cp(1, 3, buf, 2)
    got = buf.cpu().numpy()
    for v in (0, 1, 3):
        assert np.array_equal(got[v], b_np[v])
    assert np.array_equal(got[2].ravel(), O.cheby_deriv(n, rmin, rmax, b_np[0].ravel(), 2))


def test_deriv_rejects_bad_calls(ra):
    from rayleigh_amd._lib import RayshtError
    n = 16
    grid = ra.ChebyGrid(n, 0.5, 1.0)
    c = dev(np.ones(n * 4))
    out = empty(c.numel())
    with pytest.raises(RayshtError):
        grid.d_by_dr(c, c, 1)  # in place
    with pytest.raises(RayshtError):
        grid.d_by_dr(c[n:], c[: 3 * n], 1)  # overlapping
    with pytest.raises(RayshtError):
        grid.d_by_dr(c, out, 0)
    with pytest.raises(RayshtError):
        grid.d_by_dr(c, out, 5)
    with pytest.raises(ValueError):
        grid.d_by_dr(c, empty(n * 3), 1)
    with pytest.raises(ValueError):
        grid.d_by_dr_cp(1, 1, c.view(4, n), 1)
    grid.d_by_dr(c[:0], out[:0], 1)  # no columns: nothing to do
