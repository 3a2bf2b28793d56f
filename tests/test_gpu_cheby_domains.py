"""GPU parity of multi-domain Chebyshev grids (raysht_cheb_create_domains; Initialize_Cheby_Grid
with ndomains > 1 and the domain loops of Chebyshev_Polynomials.F90:353-730): the grid, the
integration weights and deriv_scaling bit for bit against the oracle (itself bit-exact with
the compiled reference); To/From_Spectral within the table form's tolerance of the oracle;
the derivatives of orders 1-4 and the dealias bit for bit; polynomials per domain; a one-domain
grid through this entry point equal to raysht_cheb_create's; the rejected calls."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import dev, host

pytestmark = pytest.mark.gpu

CASES = [
    ([16, 24, 32], [0.35, 0.5, 0.8, 1.0], None),
    ([12, 12, 12, 12], [1.0, 2.0, 3.0, 4.0, 5.0], [3, 4, 0, 11]),
    ([20, 33], [0.4, 0.7, 1.3], [5, 0]),
    ([64, 48], [0.35, 0.6, 1.0], None),
    ([21, 33, 17], [0.4, 0.7, 1.3, 1.5], None),
    ([256, 128, 200], [0.2, 0.5, 0.9, 1.4], [60, 0, 41]),
]
TOL = 1e-14  # of max |result|: the table form against the oracle's sums


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


@pytest.mark.parametrize("npoly,bounds,db", CASES)
def test_domains_vs_oracle(ra, npoly, bounds, db):
    import torch
    g = O.cheby_domains(npoly, bounds, db)
    cg = ra.ChebyGrid.domains(npoly, bounds, db)
    assert cg.n == g["ntotal"]
    assert np.array_equal(cg.grid, g["grid"])
    x, th, w, ds = cg.grid_data()
    assert np.array_equal(w, g["weights"]) and np.array_equal(ds, g["dscale"])
    ncols = 777
    f = np.random.default_rng(sum(npoly)).standard_normal(g["ntotal"] * ncols)
    out = torch.empty(f.size, dtype=torch.float64, device="cuda")
    cg.to_spectral(dev(f), out)
    ref = O.cheby_domains_to_spectral(g, f)
    assert np.abs(host(out) - ref).max() < TOL * np.abs(ref).max()
    cg.from_spectral(dev(f), out)
    ref = O.cheby_domains_from_spectral(g, f)
    assert np.abs(host(out) - ref).max() < TOL * np.abs(ref).max()
    for k in (1, 2, 3, 4):
        cg.d_by_dr(dev(f), out, dorder=k)
        assert np.array_equal(host(out), O.cheby_domains_deriv(g, f, k)), k
    buf = dev(f)
    cg.dealias_buffer(buf)
    assert np.array_equal(host(buf), O.cheby_domains_dealias(g, f))


def test_domains_polynomial_round_trip(ra):
    import torch
    npoly, bounds = [24, 40, 32], [0.3, 0.55, 0.8, 1.1]
    cg = ra.ChebyGrid.domains(npoly, bounds)
    r = cg.grid
    f = np.tile(0.3 + r * (1.1 + r * (-0.7 + 0.2 * r)), 5)
    c = torch.empty(f.size, dtype=torch.float64, device="cuda")
    back = torch.empty_like(c)
    cg.to_spectral(dev(f), c)
    cg.from_spectral(c, back)
    assert np.abs(host(back) - f).max() < 1e-13
    cg.d_by_dr(c, back, dorder=1)
    cg.from_spectral(back, c)
    dp = np.tile(1.1 + r * (-1.4 + 0.6 * r), 5)
    assert np.abs(host(c) - dp).max() < 1e-14 * max(npoly) ** 2 * np.abs(dp).max()


def test_one_domain_through_domains_entry(ra):
    import torch
    n, lo, hi = 48, 0.35, 1.0
    a, b = ra.ChebyGrid(n, lo, hi), ra.ChebyGrid.domains([n], [lo, hi])
    assert np.array_equal(a.grid, b.grid)
    assert all(np.array_equal(u, v) for u, v in zip(a.grid_data(), b.grid_data()))
    f = dev(np.random.default_rng(3).standard_normal(n * 100))
    oa, ob = torch.empty_like(f), torch.empty_like(f)
    a.to_spectral(f, oa)
    b.to_spectral(f, ob)
    assert torch.equal(oa, ob)


def test_domains_rejected(ra):
    from rayleigh_amd._lib import RayshtError
    with pytest.raises(RayshtError):
        ra.ChebyGrid.domains([16, 300], [0.1, 0.5, 1.0])  # several domains: at most 256 points each
    with pytest.raises(RayshtError):
        ra.ChebyGrid.domains([16, 16], [0.5, 0.4, 1.0])  # bounds not ascending
    with pytest.raises(RayshtError):
        ra.ChebyGrid.domains([16, 1], [0.1, 0.5, 1.0])
    cg = ra.ChebyGrid.domains([16, 32], [0.35, 0.6, 1.0])
    with pytest.raises(RayshtError):
        cg.set_form("dct")
    with pytest.raises(RayshtError):
        cg.tables()
