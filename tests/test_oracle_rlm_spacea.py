"""The oracle's rlm_spacea steps: Velocity_Components + Velocity_Derivatives + d_by_dtheta(tvar,
dtdt) and Compute_BandCurlB (Physics/Sphere_Hybrid_Space.F90:80,235-460), and rlm_spaceb's
Adjust_Emf (:557-602), pinned against:
  * the compiled reference's outputs (golden fixtures), bit for bit: the harness runs the
    reference's own d_by_dtheta / d_by_dphi in the routines' order with the same array
    statements (the physics module itself needs the whole model state to build);
  * the compiled reference live where it is built here (l_max = 0 and 1, m = l_max blocks,
    roles in any field order, with and without the pseudo-incompressible term);
  * what the algebra fixes: both passes are linear in the buffer (exactly, for a scale by 2),
    fields outside the roles are untouched, and planted exchanges of two roles fail the
    fixture checks.
"""
import numpy as np
import pytest

from oracle import oracle as O

FIXTURES = ["rlm_spacea_L12_ascending.npz", "rlm_spacea_L41_partial.npz"]


def _profiles(g):
    return {k: (g[k] if k in g else None) for k in O.HYDRO_PROFILES}


@pytest.mark.parametrize("name", FIXTURES)
def test_spacea_hydro_vs_reference_fixture(golden, name):
    g = golden(name)
    L, mv, nr, nf = int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    out = O.rlm_spacea_hydro(L, mv, nr, g["buf"], nf, g["hydro"], _profiles(g))
    assert np.array_equal(out, g["hydro_out"])


@pytest.mark.parametrize("name", FIXTURES)
def test_bandcurlb_vs_reference_fixture(golden, name):
    g = golden(name)
    L, mv, nr, nf = int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    out = O.compute_bandcurlb(L, mv, nr, g["buf"], nf, g["magnetic"], g["one_over_rsq"])
    assert np.array_equal(out, g["magnetic_out"])


@pytest.mark.parametrize("name", FIXTURES)
def test_adjust_emf_vs_reference_fixture(golden, name):
    g = golden(name)
    L, mv, nr, nf = int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    et, ep = (int(x) for x in g["emf"])
    assert np.array_equal(O.adjust_emf(L, mv, nr, g["buf"], nf, et, ep), g["emf_out"])
    assert not np.array_equal(O.adjust_emf(L, mv, nr, g["buf"], nf, ep, et), g["emf_out"])  # roles exchanged


@pytest.mark.parametrize("swap", [(1, 2), (3, 4), (6, 7)])
def test_planted_hydro_role_exchange_fails(golden, swap):
    g = golden(FIXTURES[0])
    L, mv, nr, nf = int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    v = list(g["hydro"])
    v[swap[0]], v[swap[1]] = v[swap[1]], v[swap[0]]
    assert not np.array_equal(O.rlm_spacea_hydro(L, mv, nr, g["buf"], nf, v, _profiles(g)), g["hydro_out"])


@pytest.mark.parametrize("swap", [(0, 1), (2, 3), (3, 4)])
def test_planted_bandcurlb_role_exchange_fails(golden, swap):
    g = golden(FIXTURES[1])
    L, mv, nr, nf = int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    v = list(g["magnetic"])
    v[swap[0]], v[swap[1]] = v[swap[1]], v[swap[0]]
    assert not np.array_equal(O.compute_bandcurlb(L, mv, nr, g["buf"], nf, v, g["one_over_rsq"]),
                              g["magnetic_out"])


def test_pseudo_incompressible_term_counts(golden):
    """Dropping dsdr_over_cp changes exactly the dvrdr field (written over dWdr)."""
    g = golden(FIXTURES[0])
    L, mv, nr, nf = int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    prof = _profiles(g)
    assert prof["dsdr_over_cp"] is not None
    a = O.rlm_spacea_hydro(L, mv, nr, g["buf"], nf, g["hydro"], prof)
    b = O.rlm_spacea_hydro(L, mv, nr, g["buf"], nf, g["hydro"], dict(prof, dsdr_over_cp=None))
    field = _field_of(L, mv, nr, nf)
    diff = a != b
    assert diff.any() and np.all(field[diff] == int(g["hydro"][1]))


def _field_of(L, mv, nr, nf):
    return np.concatenate([np.repeat(np.arange(nf), (L - m + 1) * 2 * nr) for m in mv])


def _random_profiles(rng, nr, pseudo):
    prof = {k: rng.uniform(0.5, 2.0, nr) for k in O.HYDRO_PROFILES}
    if not pseudo:
        prof["dsdr_over_cp"] = None
    return prof


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("L,mv,nr,pseudo", [
    (0, [0], 2, True),
    (1, [1, 0], 3, False),
    (30, "balanced", 2, True),
    (17, [17, 2, 9, 0], 3, False),
])
def test_spacea_vs_compiled_reference(L, mv, nr, pseudo):
    mv = O.m_order_balanced(L) if isinstance(mv, str) else np.array(mv, np.int32)
    nf = 18
    rng = np.random.default_rng(L * 13 + nr)
    perm = rng.permutation(nf)
    hydro, magnetic = perm[:10], perm[10:16]
    R = O.Reference()
    R.angular_init(L, mv, nr)
    buf = rng.standard_normal(O.spectral_size(L, mv, nr, nf))
    prof = _random_profiles(rng, nr, pseudo)
    o2 = rng.uniform(0.5, 2.0, nr)
    assert np.array_equal(O.rlm_spacea_hydro(L, mv, nr, buf, nf, hydro, prof), R.rlm_spacea_hydro(buf, nf, hydro, prof))
    assert np.array_equal(O.compute_bandcurlb(L, mv, nr, buf, nf, magnetic, o2),
                          R.compute_bandcurlb(buf, nf, magnetic, o2))
    et, ep = int(perm[16]), int(perm[17])
    assert np.array_equal(O.adjust_emf(L, mv, nr, buf, nf, et, ep), R.adjust_emf(buf, nf, et, ep))


def test_spacea_linear_and_other_fields_untouched():
    L, mv, nr, nf = 23, O.m_order_balanced(23), 4, 19
    rng = np.random.default_rng(6)
    a = rng.standard_normal(O.spectral_size(L, mv, nr, nf))
    b = rng.standard_normal(a.size)
    perm = rng.permutation(nf)
    hydro, magnetic = perm[:10], perm[10:16]
    prof = _random_profiles(rng, nr, True)
    o2 = rng.uniform(0.5, 2.0, nr)
    field = _field_of(L, mv, nr, nf)
    emf = (int(perm[16]), int(perm[17]))
    for fn, roles in ((lambda x: O.rlm_spacea_hydro(L, mv, nr, x, nf, hydro, prof), hydro),
                      (lambda x: O.compute_bandcurlb(L, mv, nr, x, nf, magnetic, o2), magnetic),
                      (lambda x: O.adjust_emf(L, mv, nr, x, nf, *emf), emf)):
        ra, rb = fn(a), fn(b)
        assert np.array_equal(fn(2.0 * a), 2.0 * ra)
        scale = np.abs(ra).max() + np.abs(rb).max()  # l(l+1) factors make values up to ~1e3
        assert np.allclose(fn(a + b), ra + rb, rtol=0, atol=1e-14 * scale)
        keep = ~np.isin(field, roles)
        assert np.array_equal(ra[keep], a[keep])


def _ho_args(g):
    return (int(g["lmax"]), g["mvals"], int(g["nr"]), g["buf"], int(g["nf"]), g["ho_co"], g["ho_co"].size //
            (g["buf"].size // int(g["nf"])), g["ho_vi"], g["ho_vf"], g["ho_one_over_r"], g["ho_one_over_rsq"],
            g["ho_two_over_r"])


@pytest.mark.parametrize("name", FIXTURES)
def test_hybrid_output_vs_reference_fixture(golden, name):
    """Hybrid_Output_Initial's magnetic part then Hybrid_Output_Final up to its transform
    (Sphere_Hybrid_Space.F90:224-233,437-552), with and without magnetism."""
    g = golden(name)
    assert np.array_equal(O.hybrid_output(*_ho_args(g), magnetism=True), g["ho_out"])
    assert np.array_equal(O.hybrid_output(*_ho_args(g), magnetism=False), g["ho_out_hydro"])


@pytest.mark.parametrize("which,swap", [("vi", (1, 2)), ("vi", (2, 3)), ("vf", (2, 3)), ("vf", (5, 6))])
def test_planted_output_role_exchange_fails(golden, which, swap):
    g = golden(FIXTURES[1])
    args = list(_ho_args(g))
    k = 7 if which == "vi" else 8
    v = list(args[k])
    v[swap[0]], v[swap[1]] = v[swap[1]], v[swap[0]]
    args[k] = v
    assert not np.array_equal(O.hybrid_output(*args, magnetism=True), g["ho_out"])


def test_hybrid_output_hydro_is_dtheta_and_lmax_zeroing():
    """Without magnetism the pass is the l_max zeroing of every co field and d_by_dtheta of
    pvar into dpdt_cb (both pinned on their own against the reference); with it, the fields
    outside the roles are only l_max-zeroed, and the pass is linear in (s, co) jointly."""
    L, mv, nr, nf, nf_co = 19, O.m_order_balanced(19), 3, 9, 8
    rng = np.random.default_rng(21)
    s = rng.standard_normal(O.spectral_size(L, mv, nr, nf))
    co = rng.standard_normal(O.spectral_size(L, mv, nr, nf_co))
    prof = [rng.uniform(0.5, 2.0, nr) for _ in range(3)]
    vi, vf = (4, 5, 6, 7, 1, 6), (2, 3, 4, 5, 0, 2, 3, 6, 7)
    out = O.hybrid_output(L, mv, nr, s, nf, co, nf_co, vi, vf, *prof, magnetism=False)
    want = O.zero_lmax(L, mv, nr, co, nf_co)
    O.angular_deriv("dtheta", L, mv, nr, s, nf, vf[0], out=want, nf_out=nf_co, fout=vf[4])
    assert np.array_equal(out, want)
    out = O.hybrid_output(L, mv, nr, s, nf, co, nf_co, vi, vf, *prof, magnetism=True)
    field = _field_of(L, mv, nr, nf_co)
    keep = ~np.isin(field, [vi[4], *vf[4:]])
    assert np.array_equal(out[keep], O.zero_lmax(L, mv, nr, co, nf_co)[keep])
    assert np.array_equal(O.hybrid_output(L, mv, nr, 2.0 * s, nf, 2.0 * co, nf_co, vi, vf, *prof), 2.0 * out)


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("L,mv,nr", [(0, [0], 2), (1, [1, 0], 3), (17, [17, 2, 9, 0], 2), (33, "balanced", 2)])
def test_hybrid_output_vs_compiled_reference(L, mv, nr):
    mv = O.m_order_balanced(L) if isinstance(mv, str) else np.array(mv, np.int32)
    nf, nf_co = 10, 9
    rng = np.random.default_rng(L * 7 + nr)
    R = O.Reference()
    R.angular_init(L, mv, nr)
    s = rng.standard_normal(O.spectral_size(L, mv, nr, nf))
    co = rng.standard_normal(O.spectral_size(L, mv, nr, nf_co))
    prof = [rng.uniform(0.5, 2.0, nr) for _ in range(3)]
    ps, pc = rng.permutation(nf), rng.permutation(nf_co)
    vi = (ps[0], ps[1], ps[2], ps[3], pc[0], pc[1])
    vf = (ps[4], ps[5], ps[6], ps[7], pc[2], pc[3], pc[4], pc[1], pc[5])
    for mag in (True, False):
        assert np.array_equal(O.hybrid_output(L, mv, nr, s, nf, co, nf_co, vi, vf, *prof, magnetism=mag),
                              R.hybrid_output(s, nf, co, nf_co, vi, vf, *prof, magnetism=mag)), mag
