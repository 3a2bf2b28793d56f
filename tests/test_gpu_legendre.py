"""GPU parity of the Legendre path (device tables, StP, PtS) against the oracle and
the reference's golden fixtures."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import dev, empty, host, rel_max, ulp_diff

pytestmark = pytest.mark.gpu

# FP64 path: agreement with the reference's FP64 transforms (identical tables;
# only the summation order of the contractions differs)
TOL_FP64 = 1e-13


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


@pytest.mark.parametrize("nt,lmax,order", [(96, 63, "balanced"), (64, 41, "balanced"), (20, 12, "ascending"),
                                           (384, 255, "balanced")])
def test_nodes_and_tables_match_oracle(ra, nt, lmax, order):
    mv = O.m_order_balanced(lmax) if order == "balanced" else np.arange(lmax + 1, dtype=np.int32)
    plan = ra.LegendrePlan(nt, lmax, mv, 1, 1)
    for a, b in zip(plan.nodes(), O.gauss_legendre(nt)):
        assert np.array_equal(a, b)
    ref = O.legendre_tables(nt, lmax, mv)
    for w, (got, want) in enumerate(zip(plan.tables(), ref)):
        assert got.shape == want.shape
        # double-double on the device vs quad on the host, both rounded once to double
        assert ulp_diff(got, want) <= 1, f"table {w}"
        # the two agree to ~1e-31 absolute; a double can still round differently where the
        # value is tiny next to that (near a zero of P_lm): allow a few such 1-ulp cases
        assert np.count_nonzero(got != want) <= max(2, want.size // 100_000), f"table {w}"


def test_high_l_tables_match_reference_samples(ra, golden):
    """l_max = 511 on 768 nodes: P_mm underflows double near the poles; the device's
    extended-exponent recursion must still reproduce the reference's quad values."""
    g = golden("legendre_L511_nt768_sampled.npz")
    plan = ra.LegendrePlan(int(g["nt"]), int(g["lmax"]), g["mvals"], 1, 1)
    for w, key in enumerate(("p_even", "p_odd", "ip_even", "ip_odd")):
        got = plan.table(w)[g[key + "_idx"]]
        assert ulp_diff(got, g[key + "_val"]) <= 1, key


@pytest.mark.parametrize("name", ["legendre_L41_nt64.npz", "legendre_L63_nt96.npz", "legendre_L63_partial.npz"])
def test_stp_pts_vs_reference_fixtures(ra, golden, name):
    g = golden(name)
    nt, L, mv, nr, nf = int(g["nt"]), int(g["lmax"]), g["mvals"], int(g["nr"]), int(g["nf"])
    plan = ra.LegendrePlan(nt, L, mv, nr, nf)
    s = dev(g["s_in"])
    p = empty(plan.hybrid_size)
    plan.stp(s, p)
    assert rel_max(host(p), g["p_out"]) < TOL_FP64
    s2 = empty(plan.spectral_size)
    plan.pts(dev(g["p_out"]), s2)
    assert rel_max(host(s2), g["s_rt"]) < TOL_FP64


@pytest.mark.parametrize("nt,lmax,nr,nf,order", [
    (96, 63, 48, 4, "balanced"),    # the CPU-reference configuration (l_max=63, N_r=48, 4 fields)
    (64, 41, 5, 3, "balanced"),     # ragged column count (30)
    (20, 12, 3, 1, "ascending"),    # ragged theta tile (10 southern nodes)
    (2, 0, 2, 1, "ascending"),      # l_max = 0: one m, one l
    (4, 1, 1, 1, "ascending"),      # l_max = 1
])
def test_stp_pts_vs_oracle(ra, nt, lmax, nr, nf, order):
    mv = O.m_order_balanced(lmax) if order == "balanced" else np.arange(lmax + 1, dtype=np.int32)
    plan = ra.LegendrePlan(nt, lmax, mv, nr, nf)
    tabs = O.legendre_tables(nt, lmax, mv)
    s_np = O.synthetic_spectral(lmax, mv, nr, nf, seed=nt + lmax)
    p_ref = O.stp(nt, lmax, mv, nr, nf, tabs, s_np)
    p = empty(plan.hybrid_size)
    plan.legendre_transform(dev(s_np), p)
    pg = host(p)
    nrhs = 2 * nr * nf
    # per field (the metric's max|d|/max|ref|)
    P, Pr = pg.reshape(len(mv), nf, 2 * nr, nt), p_ref.reshape(len(mv), nf, 2 * nr, nt)
    for f in range(nf):
        assert rel_max(P[:, f], Pr[:, f]) < TOL_FP64
    s_ref = O.pts(nt, lmax, mv, nr, nf, tabs, p_ref)
    s2 = empty(plan.spectral_size)
    plan.legendre_transform(dev(p_ref), s2)
    assert rel_max(host(s2), s_ref) < TOL_FP64
    assert nrhs == plan.nrhs


def test_partial_m_list(ra):
    nt, L, nr, nf = 96, 63, 3, 2
    mv = np.array([17, 0, 63, 5, 40], dtype=np.int32)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf)
    tabs = O.legendre_tables(nt, L, mv)
    s_np = O.synthetic_spectral(L, mv, nr, nf, seed=9)
    p = empty(plan.hybrid_size)
    plan.stp(dev(s_np), p)
    assert rel_max(host(p), O.stp(nt, L, mv, nr, nf, tabs, s_np)) < TOL_FP64


def test_single_mode_known_answer(ra):
    """Test_SHT.F90:47-200 (l, m) = (2, 1) against sqrt(5/(24 pi)) (-3 cos sin)."""
    nt, L, nr, nf = 64, 41, 4, 3
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf)
    radius = np.linspace(0.5, 1.5, nr)
    off = O.spectral_offsets(L, mv, nr, nf)
    s = np.zeros(off[-1])
    mp = int(np.where(mv == 1)[0][0])
    blk = s[off[mp]: off[mp + 1]].reshape(nf, 2, nr, L)
    for f in range(1, nf + 1):
        blk[f - 1, 0, :, 1] = (2 * f + 1) * radius * 2.0
        blk[f - 1, 1, :, 1] = -(3 * f * f + 2) * radius ** 2 * 2.0
    p = empty(plan.hybrid_size)
    plan.stp(dev(s), p)
    P = host(p).reshape(len(mv), 2 * nr * nf, nt)
    x = plan.nodes()[0]
    true = np.sqrt(5 / (24 * np.pi)) * (-3 * x * np.sqrt(1 - x * x))
    for f in range(1, nf + 1):
        for r in range(nr):
            assert np.abs(P[mp, r + 2 * nr * (f - 1)] - (2 * f + 1) * radius[r] * true).max() < 1e-10
    assert np.abs(np.delete(P, mp, axis=0)).max() < 1e-10
    s2 = empty(plan.spectral_size)
    plan.pts(p, s2)
    assert np.abs(host(s2) * nt - s / 2.0).max() < 1e-10


def test_white_noise_round_trips(ra):
    """Test_SHT.F90:328-447: all modes 1 (m=0 imaginary 0, l_max zeroed), 10 loops."""
    nt, L, nr, nf = 64, 41, 2, 3
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf)
    off = O.spectral_offsets(L, mv, nr, nf)
    s = np.zeros(off[-1])
    for mp, m in enumerate(mv):
        blk = s[off[mp]: off[mp + 1]].reshape(nf, 2, nr, L - m + 1)
        blk[:, :, :, : L - m] = 1.0
        if m == 0:
            blk[:, 1] = 0.0
    to_phys = np.where(np.repeat(mv, [(L - m + 1) * 2 * nr * nf for m in mv]) == 0, 1.0, 0.5)
    to_spec = np.where(np.repeat(mv, [(L - m + 1) * 2 * nr * nf for m in mv]) == 0, 1.0 / (2 * nt), 1.0 / nt)
    lmask = np.concatenate([np.tile(np.r_[np.ones(L - m), 0.0], 2 * nr * nf) for m in mv])
    cur = s.copy()
    p = empty(plan.hybrid_size)
    s2 = empty(plan.spectral_size)
    for _ in range(10):
        plan.stp(dev(cur / to_phys), p)
        plan.pts(p, s2)
        cur = host(s2) * lmask / to_spec
    assert np.abs(cur - s).max() / 10 < 1e-10


def test_round_trip_identity_full_size(ra):
    """PtS(StP(a)) = a / N_phi at l_max = 511, N_theta = 768, N_r = 192, 8 fields."""
    import torch
    nt, L, nr, nf = 768, 511, 192, 8
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf)
    g = torch.Generator(device="cuda").manual_seed(5)
    s = torch.randn(plan.spectral_size, dtype=torch.float64, device="cuda", generator=g)
    p = torch.empty(plan.hybrid_size, dtype=torch.float64, device="cuda")
    s2 = torch.empty_like(s)
    plan.stp(s, p)
    plan.pts(p, s2)
    torch.cuda.synchronize()
    err = float(((s2 * (2 * nt) - s).abs().max() / s.abs().max()).item())
    # unit-variance coefficients at every l (no decay): the identity holds to the accuracy of
    # the reference's Gauss weights (about 1e-14 relative each); the reference's own tolerance is 1e-10
    assert err < 1e-10


@pytest.mark.parametrize("ncols", [7, 1, 130])
def test_two_d_column_plans(ra, ncols):
    """StP_2d_dgp / PtS_2d_dgpv2: data(m:l_max, 1:ncols) blocks with any column count."""
    nt, L = 96, 63
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(nt, L, mv, 0, 0, n_cols=ncols)
    tabs = O.legendre_tables(nt, L, mv)
    rng = np.random.default_rng(ncols)
    s = rng.standard_normal(plan.spectral_size)
    p = empty(plan.hybrid_size)
    plan.stp(dev(s), p)
    pr = O.stp_cols(nt, L, mv, ncols, tabs, s)
    assert rel_max(host(p), pr) < TOL_FP64
    s2 = empty(plan.spectral_size)
    plan.pts(dev(pr), s2)
    assert rel_max(host(s2), O.pts_cols(nt, L, mv, ncols, tabs, pr)) < TOL_FP64
    with pytest.raises(Exception):
        plan.reform_2a3a(p, empty((2 * nt + 2) * nt), 2 * nt)  # no radial structure to permute


def test_host_array_entry_points(ra):
    """raysht_legendre_stp_host / _pts_host: per-m host pointers, as the Fortran layer passes them."""
    import ctypes
    from rayleigh_amd import _lib
    nt, L, nr, nf = 64, 41, 3, 2
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf)
    tabs = O.legendre_tables(nt, L, mv)
    s = O.synthetic_spectral(L, mv, nr, nf, seed=4)
    off = O.spectral_offsets(L, mv, nr, nf)
    blocks = [np.ascontiguousarray(s[off[i]:off[i + 1]]) for i in range(len(mv))]
    ptrs = (ctypes.c_void_p * len(mv))(*[b.ctypes.data for b in blocks])
    p = np.zeros(plan.hybrid_size)
    _lib.check(_lib.lib().raysht_legendre_stp_host(plan.handle, ptrs, p.ctypes.data_as(ctypes.c_void_p)), "stp_host")
    pr = O.stp(nt, L, mv, nr, nf, tabs, s)
    assert rel_max(p, pr) < TOL_FP64
    outs = [np.zeros_like(b) for b in blocks]
    optrs = (ctypes.c_void_p * len(mv))(*[b.ctypes.data for b in outs])
    _lib.check(_lib.lib().raysht_legendre_pts_host(plan.handle, pr.ctypes.data_as(ctypes.c_void_p), optrs), "pts_host")
    assert rel_max(np.concatenate(outs), O.pts(nt, L, mv, nr, nf, tabs, pr)) < TOL_FP64


@pytest.mark.parametrize("nt", [64, 96, 768])
def test_nodes_binary128_match_oracle(ra, nt):
    """raysht_plan_nodes_quad: the Real*16 coloc / gl_weights the reference stores, byte
    for byte as the oracle forms them (the oracle's are the compiled reference's)."""
    import ctypes
    from rayleigh_amd import _lib
    lmax = (2 * nt) // 3 - 1
    plan = ra.LegendrePlan(nt, lmax, np.arange(lmax + 1, dtype=np.int32), 1, 1)
    x = np.zeros((nt, 16), np.uint8)
    w = np.zeros((nt, 16), np.uint8)
    assert _lib.lib().raysht_plan_nodes_quad(plan.handle, x.ctypes.data_as(ctypes.c_void_p),
                                             w.ctypes.data_as(ctypes.c_void_p)) == 0
    xo, wo = O.gauss_legendre_binary128(nt)
    assert np.array_equal(x, xo) and np.array_equal(w, wo)
