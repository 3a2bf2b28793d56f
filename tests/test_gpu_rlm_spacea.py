"""GPU parity of the fused rlm_spacea passes (raysht_rlm_spacea_hydro: Velocity_Components +
Velocity_Derivatives + d_by_dtheta(tvar, dtdt); raysht_compute_bandcurlb: Compute_BandCurlB;
Physics/Sphere_Hybrid_Space.F90:80,235-460) and rlm_spaceb's Adjust_Emf (raysht_adjust_emf,
:557-602): bit for bit against the compiled reference's
outputs (golden fixtures) and against the oracle for l_max from 0 to 2049 (one to 33 chunks of
64 values per column), m = l_max blocks, roles in any field order, with and without the
pseudo-incompressible term; a large buffer by exact linearity; the rejected calls.  The
output-iteration passes (raysht_hybrid_output_initial / _final: Hybrid_Output_Initial / _Final,
:224-233,437-552) bit for bit against the fixtures and the oracle, with and without magnetism."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import dev, host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


def _dev_profiles(prof):
    return {k: (None if v is None else dev(v)) for k, v in prof.items()}


@pytest.mark.parametrize("name", ["rlm_spacea_L12_ascending.npz", "rlm_spacea_L41_partial.npz"])
def test_spacea_vs_reference_fixture(ra, golden, name):
    g = golden(name)
    L, mv, nr, nf = int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    plan = ra.LegendrePlan(2 * (L + 1), L, mv, nr, nf)
    prof = {k: (g[k] if k in g else None) for k in O.HYDRO_PROFILES}
    buf = dev(g["buf"])
    plan.rlm_spacea_hydro(buf, g["hydro"], _dev_profiles(prof))
    assert np.array_equal(host(buf), g["hydro_out"])
    buf = dev(g["buf"])
    plan.compute_bandcurlb(buf, g["magnetic"], dev(g["one_over_rsq"]))
    assert np.array_equal(host(buf), g["magnetic_out"])
    buf = dev(g["buf"])
    plan.adjust_emf(buf, *(int(x) for x in g["emf"]))
    assert np.array_equal(host(buf), g["emf_out"])


@pytest.mark.parametrize("L,mv,nr,pseudo", [
    (0, [0], 2, True),
    (1, [1, 0], 3, False),
    (63, "balanced", 9, True),
    (64, [64, 0, 63, 1], 3, False),
    (255, "balanced", 3, True),
    (300, [300, 0, 44, 1, 299], 2, False),
    (1023, [0, 1, 511, 1023, 1022], 2, True),
    (2049, [2049, 0, 1985, 1000], 1, False),
])
def test_spacea_vs_oracle(ra, L, mv, nr, pseudo):
    mv = O.m_order_balanced(L) if isinstance(mv, str) else np.array(mv, np.int32)
    nf = 18
    rng = np.random.default_rng(L * 3 + nr)
    perm = rng.permutation(nf)
    hydro, magnetic, emf = perm[:10], perm[10:16], (int(perm[16]), int(perm[17]))
    plan = ra.LegendrePlan(2 * (L + 1), L, mv, nr, nf)
    a = rng.standard_normal(plan.spectral_size)
    prof = {k: rng.uniform(0.5, 2.0, nr) for k in O.HYDRO_PROFILES}
    if not pseudo:
        prof["dsdr_over_cp"] = None
    o2 = rng.uniform(0.5, 2.0, nr)
    buf = dev(a)
    plan.rlm_spacea_hydro(buf, hydro, _dev_profiles(prof))
    assert np.array_equal(host(buf), O.rlm_spacea_hydro(L, mv, nr, a, nf, hydro, prof)), (L, "hydro")
    buf = dev(a)
    plan.compute_bandcurlb(buf, magnetic, dev(o2))
    assert np.array_equal(host(buf), O.compute_bandcurlb(L, mv, nr, a, nf, magnetic, o2)), (L, "bandcurlb")
    buf = dev(a)
    plan.adjust_emf(buf, *emf)
    assert np.array_equal(host(buf), O.adjust_emf(L, mv, nr, a, nf, *emf)), (L, "adjust_emf")


def test_spacea_large_linearity(ra):
    """l_max = 511, N_r = 64, 16 fields: pass(2a) = 2 pass(a) bit for bit for both passes, and
    the fields outside the roles are untouched."""
    import torch
    L, nr, nf = 511, 64, 16
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(768, L, mv, nr, nf)
    gen = torch.Generator(device="cuda").manual_seed(11)
    a = torch.randn(plan.spectral_size, dtype=torch.float64, device="cuda", generator=gen)
    hydro, magnetic = [3, 0, 7, 12, 1, 9, 4, 15, 2, 10], [5, 13, 8, 14, 6, 11]
    prof = {k: torch.rand(nr, dtype=torch.float64, device="cuda", generator=gen) + 0.5 for k in O.HYDRO_PROFILES}
    o2 = torch.rand(nr, dtype=torch.float64, device="cuda", generator=gen) + 0.5
    per_field = plan.spectral_size // nf
    field = torch.cat([torch.arange(nf, device="cuda").repeat_interleave((L - m + 1) * 2 * nr) for m in mv])
    assert field.numel() == nf * per_field
    b1, b2 = a.clone(), 2.0 * a
    plan.rlm_spacea_hydro(b1, hydro, prof)
    plan.rlm_spacea_hydro(b2, hydro, prof)
    assert torch.equal(b2, 2.0 * b1)
    keep = ~torch.isin(field, torch.tensor(hydro, device="cuda"))
    assert torch.equal(b1[keep], a[keep])
    c1, c2 = a.clone(), 2.0 * a
    plan.compute_bandcurlb(c1, magnetic, o2)
    plan.compute_bandcurlb(c2, magnetic, o2)
    assert torch.equal(c2, 2.0 * c1)
    keep = ~torch.isin(field, torch.tensor(magnetic, device="cuda"))
    assert torch.equal(c1[keep], a[keep])


def test_spacea_rejected_calls(ra):
    import torch
    L, nr, nf = 15, 2, 10
    plan = ra.LegendrePlan(32, L, np.arange(L + 1, dtype=np.int32), nr, nf)
    buf = torch.zeros(plan.spectral_size, dtype=torch.float64, device="cuda")
    prof = {k: torch.ones(nr, dtype=torch.float64, device="cuda") for k in O.HYDRO_PROFILES}
    with pytest.raises(RuntimeError):
        plan.rlm_spacea_hydro(buf, [0, 1, 2, 3, 4, 5, 6, 7, 8, 8], prof)  # a field in two roles
    with pytest.raises(RuntimeError):
        plan.rlm_spacea_hydro(buf, [0, 1, 2, 3, 4, 5, 6, 7, 8, 10], prof)  # no field 10
    with pytest.raises(ValueError):
        plan.rlm_spacea_hydro(buf, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9], dict(prof, dlnrho=None))
    with pytest.raises(RuntimeError):
        plan.compute_bandcurlb(buf, [0, 1, 2, 3, 4, 0], prof["dlnrho"])
    with pytest.raises(RuntimeError):
        plan.compute_bandcurlb(buf, [0, 1, 2, 3, 4, -1], prof["dlnrho"])
    with pytest.raises(RuntimeError):
        plan.adjust_emf(buf, 3, 3)
    with pytest.raises(RuntimeError):
        plan.adjust_emf(buf, 3, 10)


def _ho_run(ra, plan, s, co, vi, vf, prof, magnetism):
    """raysht_hybrid_output_initial (magnetism) then raysht_hybrid_output_final on device copies"""
    import torch
    s, c = dev(s), dev(co)
    if not magnetism:
        plan.hybrid_output_final(s, c, vf)
        return host(c)
    tmp = torch.full((2 * s.numel() // plan.n_fields,), float("nan"), dtype=torch.float64, device="cuda")
    o_r, o_rsq, t_r = (dev(x) for x in prof)
    plan.hybrid_output_initial(s, c, tmp, vi, o_rsq)
    plan.hybrid_output_final(s, c, vf, tmp, o_r, o_rsq, t_r)
    return host(c)


@pytest.mark.parametrize("name", ["rlm_spacea_L12_ascending.npz", "rlm_spacea_L41_partial.npz"])
def test_hybrid_output_vs_reference_fixture(ra, golden, name):
    g = golden(name)
    L, mv, nr, nf = int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    plan = ra.LegendrePlan(2 * (L + 1), L, mv, nr, nf)
    prof = (g["ho_one_over_r"], g["ho_one_over_rsq"], g["ho_two_over_r"])
    for mag, key in ((True, "ho_out"), (False, "ho_out_hydro")):
        out = _ho_run(ra, plan, g["buf"], g["ho_co"], g["ho_vi"], g["ho_vf"], prof, mag)
        assert np.array_equal(out, g[key]), key


@pytest.mark.parametrize("L,mv,nr", [
    (0, [0], 2),
    (1, [1, 0], 3),
    (63, "balanced", 5),
    (64, [64, 0, 63, 1], 3),
    (300, [300, 0, 44, 1, 299], 2),
    (2049, [2049, 0, 1985, 1000], 1),
])
def test_hybrid_output_vs_oracle(ra, L, mv, nr):
    mv = O.m_order_balanced(L) if isinstance(mv, str) else np.array(mv, np.int32)
    nf, nf_co = 10, 9
    rng = np.random.default_rng(L * 5 + nr)
    plan = ra.LegendrePlan(2 * (L + 1), L, mv, nr, nf)
    s = rng.standard_normal(plan.spectral_size)
    co = rng.standard_normal(plan.spectral_size // nf * nf_co)
    prof = [rng.uniform(0.5, 2.0, nr) for _ in range(3)]
    ps, pc = rng.permutation(nf), rng.permutation(nf_co)
    vi = (ps[0], ps[1], ps[2], ps[3], pc[0], pc[1])
    vf = (ps[4], ps[5], ps[6], ps[7], pc[2], pc[3], pc[4], pc[1], pc[5])
    for mag in (True, False):
        want = O.hybrid_output(L, mv, nr, s, nf, co, nf_co, vi, vf, *prof, magnetism=mag)
        assert np.array_equal(_ho_run(ra, plan, s, co, vi, vf, prof, mag), want), (L, mag)


def test_hybrid_output_rejected_calls(ra):
    import torch
    L, nr, nf = 15, 2, 10
    plan = ra.LegendrePlan(32, L, np.arange(L + 1, dtype=np.int32), nr, nf)
    per = plan.spectral_size // nf
    s = torch.zeros(plan.spectral_size, dtype=torch.float64, device="cuda")
    co = torch.zeros(7 * per, dtype=torch.float64, device="cuda")
    tmp = torch.zeros(2 * per, dtype=torch.float64, device="cuda")
    o = torch.ones(nr, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError):
        plan.hybrid_output_initial(s, co, tmp, [0, 1, 2, 2, 0, 1], o)  # an s field in two roles
    with pytest.raises(RuntimeError):
        plan.hybrid_output_initial(s, co, tmp, [0, 1, 2, 3, 0, 7], o)  # co has no field 7
    with pytest.raises(RuntimeError):
        plan.hybrid_output_initial(s, s[:7 * per], tmp, [0, 1, 2, 3, 0, 1], o)  # co overlaps s
    with pytest.raises(ValueError):
        plan.hybrid_output_initial(s, co, co[:3 * per], [0, 1, 2, 3, 0, 1], o)  # tmp of 3 fields
    with pytest.raises(RuntimeError):
        plan.hybrid_output_final(s, co, [0, 1, 2, 3, 0, 1, 2, 3, 3], tmp, o, o, o)  # a co field in two roles
    with pytest.raises(RuntimeError):
        plan.hybrid_output_final(s, co, [10, 1, 2, 3, 0, 1, 2, 3, 4])  # s has no field 10
    with pytest.raises(ValueError):
        plan.hybrid_output_final(s, co, [0, 1, 2, 3, 0, 1, 2, 3, 4], tmp)  # profiles missing
