"""GPU parity of rlm_spaceb's fused combination and the l(l+1) / radial scalings
(raysht_rlm_spaceb_hydro, raysht_lm_scale, raysht_zero_lmax: Sphere_Hybrid_Space.F90:127-214,
:244, :70-72): bit for bit against the compiled reference's outputs (golden fixtures) and
against the oracle for l_max from 0 to 2049 (one to 33 chunks of 64 values per column),
m = l_max blocks, W, P, Z in any field order; in-place and cross-field scalings; a large
buffer by exact linearity; the rejected calls."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import dev, empty, host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


@pytest.mark.parametrize("name", ["rlm_spaceb_L12_ascending.npz", "rlm_spaceb_L41_partial.npz"])
def test_rlm_spaceb_vs_reference_fixture(ra, golden, name):
    g = golden(name)
    L, mv, nr, nf = int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    plan = ra.LegendrePlan(2 * (L + 1), L, mv, nr, nf)
    buf = dev(g["buf"])
    plan.rlm_spaceb_hydro(buf, *(int(x) for x in g["fields"]))
    assert np.array_equal(host(buf), g["out"])


@pytest.mark.parametrize("L,mv,nr,nf,fields", [
    (0, [0], 2, 3, (0, 1, 2)),
    (1, [1, 0], 3, 3, (2, 0, 1)),
    (63, "balanced", 20, 4, (1, 2, 3)),
    (255, "balanced", 5, 5, (4, 1, 0)),
    (300, [300, 0, 44, 1, 299], 3, 3, (0, 2, 1)),
    (1023, [0, 1, 511, 1023, 1022], 2, 4, (3, 1, 2)),
    (1500, [0, 700, 1500], 1, 3, (1, 0, 2)),
    (2049, [2049, 0, 1985, 1000], 2, 3, (2, 1, 0)),
])
def test_rlm_spaceb_vs_oracle(ra, L, mv, nr, nf, fields):
    mv = O.m_order_balanced(L) if isinstance(mv, str) else np.array(mv, np.int32)
    plan = ra.LegendrePlan(2 * (L + 1), L, mv, nr, nf)
    a = np.random.default_rng(L + nf).standard_normal(plan.spectral_size)
    buf = dev(a)
    plan.rlm_spaceb_hydro(buf, *fields)
    assert np.array_equal(host(buf), O.rlm_spaceb(L, mv, nr, a, nf, *fields)), (L, fields)


def test_rlm_spaceb_large_linearity(ra):
    """l_max = 511, N_r = 192, 8 fields: rlm(2a) = 2 rlm(a) bit for bit, and field 5 (not W,
    P or Z) keeps all but its l_max mode."""
    import torch
    L, nr, nf = 511, 192, 8
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(768, L, mv, nr, nf)
    g = torch.Generator(device="cuda").manual_seed(4)
    a = torch.randn(plan.spectral_size, dtype=torch.float64, device="cuda", generator=g)
    b, b2 = a.clone(), a * 2.0
    plan.rlm_spaceb_hydro(b, 1, 6, 7)
    plan.rlm_spaceb_hydro(b2, 1, 6, 7)
    assert torch.equal(b2, 2.0 * b)
    f5, m5 = O.zero_lmax(L, mv, nr, a.cpu().numpy(), nf), _field_mask(L, mv, nr, nf, 5)
    assert np.array_equal(host(b)[m5], f5[m5])


def _field_mask(L, mv, nr, nf, f):
    return np.concatenate([np.repeat(np.arange(nf), (L - m + 1) * 2 * nr) == f for m in mv])


@pytest.mark.parametrize("with_l,with_r", [(True, True), (True, False), (False, True), (False, False)])
def test_lm_scale_vs_oracle(ra, with_l, with_r):
    import torch
    L, nr, nf = 95, 7, 3
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(192, L, mv, nr, nf)
    rng = np.random.default_rng(11)
    a_np = rng.standard_normal(plan.spectral_size)
    lfac = ra.l_l_plus1(L) if with_l else None
    rfac = 1.0 / rng.uniform(0.35, 1.0, nr) ** 2 if with_r else None
    td = (lambda x: None if x is None else torch.from_numpy(x).cuda())
    out = empty(plan.spectral_size // nf)
    plan.lm_scale(dev(a_np), 2, out, 0, td(lfac), td(rfac))
    assert np.array_equal(host(out), O.lm_scale(L, mv, nr, a_np, nf, 2, lfac, rfac))
    # in place, and field 0 -> field 1 of the same buffer
    a = dev(a_np)
    plan.lm_scale(a, 1, a, 1, td(lfac), td(rfac))
    plan.lm_scale(a, 0, a, 2, td(ra.over_l_l_plus1(L)), None)
    want = O.lm_scale(L, mv, nr, a_np, nf, 1, lfac, rfac, out=a_np.copy(), nf_out=nf, fout=1)
    want = O.lm_scale(L, mv, nr, want, nf, 0, ra.over_l_l_plus1(L), None, out=want, nf_out=nf, fout=2)
    assert np.array_equal(host(a), want)


def test_zero_lmax_vs_oracle(ra):
    L, nr, nf = 40, 9, 3
    mv = np.array([40, 0, 13, 2], np.int32)
    plan = ra.LegendrePlan(82, L, mv, nr, nf)
    a_np = np.random.default_rng(2).standard_normal(plan.spectral_size)
    a = dev(a_np)
    plan.zero_lmax(a)
    assert np.array_equal(host(a), O.zero_lmax(L, mv, nr, a_np, nf))


def test_tables_match_reference_definition(ra):
    ol, ll = ra.over_l_l_plus1(12), ra.l_l_plus1(12)
    assert ol[0] == 1.0 and ll[0] == 0.0 and ll[12] == 156.0 and ol[3] == 1.0 / 12.0


def test_rlm_rejects_bad_calls(ra):
    from rayleigh_amd._lib import RayshtError
    plan = ra.LegendrePlan(16, 7, np.arange(8, dtype=np.int32), 3, 3)
    a = empty(plan.spectral_size)
    with pytest.raises(RayshtError):
        plan.rlm_spaceb_hydro(a, 0, 0, 1)  # W and P the same field
    with pytest.raises(RayshtError):
        plan.rlm_spaceb_hydro(a, 0, 1, 3)  # no field 3
    with pytest.raises(RayshtError):
        plan.rlm_spaceb_hydro(empty(plan.spectral_size // 3 * 2), 0, 1, 2)  # two fields
    with pytest.raises(ValueError):
        plan.lm_scale(a, 0, a, 1, empty(7), None)  # lfac needs l_max + 1 values
    with pytest.raises(RayshtError):
        plan.lm_scale(a, 0, a, 9, None, None)  # no field 9
