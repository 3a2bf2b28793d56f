"""GPU parity of the angular spectral derivatives (raysht_d_by_dtheta / raysht_d_by_dphi:
Spectral_Derivatives.F90's d_by_dtheta, d_by_sdtheta, d_by_dphi): bit for bit against the
compiled reference's outputs (golden fixtures) and against the oracle for full, balanced
and partial m lists (m = l_max included), buffer and separate-array forms, in-place
d_by_dphi; a large case by linearity; the rejected calls."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import dev, empty, host

pytestmark = pytest.mark.gpu

KINDS = ("dtheta", "sdtheta", "dphi")


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


@pytest.mark.parametrize("name", ["angular_L12_ascending.npz", "angular_L41_balanced.npz", "angular_L9_partial.npz"])
def test_angular_vs_reference_fixture(ra, golden, name):
    g = golden(name)
    L, mv, nr = int(g["lmax"]), g["mvals"], int(g["nr"])
    plan = ra.LegendrePlan(2 * (L + 1), L, mv, nr, 2)
    for kind in KINDS:
        buf = dev(g["buf"])
        plan.angular_derivative(kind, buf, 0, buf, 1)  # field 1 -> field 2 of the same buffer
        assert np.array_equal(host(buf), g[f"{kind}_buffer"]), kind
        arr = empty(plan.spectral_size // 2)
        plan.angular_derivative(kind, dev(g["buf"]), 1, arr, 0)  # field 2 -> a one-field array
        assert np.array_equal(host(arr), g[f"{kind}_arr"]), kind


@pytest.mark.parametrize("L,nr,nf,kind_m", [(0, 2, 1, "ascending"), (1, 3, 2, "ascending"), (63, 48, 4, "balanced"),
                                            (41, 65, 3, "partial"), (255, 16, 2, "balanced")])
def test_angular_vs_oracle(ra, L, nr, nf, kind_m):
    mv = {"ascending": np.arange(L + 1, dtype=np.int32), "balanced": O.m_order_balanced(L),
          "partial": np.array([41, 0, 20, 3, 1], dtype=np.int32)}[kind_m]
    plan = ra.LegendrePlan(2 * (L + 1), L, mv, nr, nf)
    a_np = np.random.default_rng(L + nr).standard_normal(plan.spectral_size)
    fin, fout = nf - 1, 0
    for kind in KINDS:
        want = O.angular_deriv(kind, L, mv, nr, a_np, nf, fin, out=np.full(plan.spectral_size, np.nan), nf_out=nf,
                               fout=fout)
        out = empty(plan.spectral_size)
        plan.angular_derivative(kind, dev(a_np), fin, out, fout)
        got = host(out)
        sel = ~np.isnan(want)  # only field fout is written
        assert np.array_equal(np.isnan(got), ~sel) and np.array_equal(got[sel], want[sel]), kind


def test_dphi_in_place(ra):
    L, nr, nf = 31, 5, 3
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(64, L, mv, nr, nf)
    a_np = np.random.default_rng(3).standard_normal(plan.spectral_size)
    a = dev(a_np)
    plan.angular_derivative("dphi", a, 1, a, 1)
    want = O.angular_deriv("dphi", L, mv, nr, a_np, nf, 1, out=a_np.copy(), nf_out=nf, fout=1)
    assert np.array_equal(host(a), want)


def test_angular_large_linearity(ra):
    """l_max = 511, N_r = 192, 8 fields (the split-precision workload's spectral size):
    against the oracle on field 3, and derivative(2a) = 2 derivative(a) bit for bit."""
    import torch
    L, nr, nf = 511, 192, 8
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(768, L, mv, nr, nf)
    g = torch.Generator(device="cuda").manual_seed(9)
    a = torch.randn(plan.spectral_size, dtype=torch.float64, device="cuda", generator=g)
    out = torch.empty(plan.spectral_size // nf, dtype=torch.float64, device="cuda")
    out2 = torch.empty_like(out)
    for kind in KINDS:
        plan.angular_derivative(kind, a, 3, out, 0)
        plan.angular_derivative(kind, a * 2.0, 3, out2, 0)
        assert torch.equal(out2, 2.0 * out), kind
        if kind == "sdtheta":
            want = O.angular_deriv(kind, L, mv, nr, a.cpu().numpy(), nf, 3)
            assert np.array_equal(host(out), want)


def test_angular_rejects_bad_calls(ra):
    from rayleigh_amd._lib import RayshtError
    plan = ra.LegendrePlan(16, 7, np.arange(8, dtype=np.int32), 3, 2)
    a = empty(plan.spectral_size)
    with pytest.raises(RayshtError):
        plan.angular_derivative("dtheta", a, 1, a, 1)  # l stencil in place
    with pytest.raises(RayshtError):
        plan.angular_derivative("sdtheta", a, 2, a, 0)  # no field 2
    with pytest.raises(ValueError):
        plan.angular_derivative("dtheta", a, 0, empty(5), 0)
    cols = ra.LegendrePlan(16, 7, np.arange(8, dtype=np.int32), 0, 0, n_cols=6)
    with pytest.raises(ValueError):
        cols.angular_derivative("dphi", empty(cols.spectral_size), 0, empty(cols.spectral_size), 0)
