"""GPU parity of the phi FFTs, the row permutes and the Chebyshev transforms."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import dev, empty, host, rel_max

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


@pytest.mark.parametrize("nphi", [6, 12, 18, 36, 48, 72, 96, 192, 384, 768, 1536, 3072, 6144,
                                  10, 14, 70, 210, 336, 480, 960, 1920, 1050, 686])
def test_fft_c2r_r2c_vs_oracle(ra, nphi):
    rng = np.random.default_rng(nphi)
    nl = 37
    x = rng.standard_normal((nl, nphi + 2))
    want = O.fft_c2r(nphi, x.ravel())
    t = dev(x)
    ra.fft_to_physical(t, nphi)
    got = host(t).ravel()
    assert rel_max(got, want) < 1e-14 * np.log2(nphi) * 4
    y = np.zeros((nl, nphi + 2))
    y[:, :nphi] = rng.standard_normal((nl, nphi))
    want = O.fft_r2c(nphi, y.ravel())
    t = dev(y)
    ra.fft_to_spectral(t, nphi)
    assert rel_max(host(t).ravel(), want) < 1e-14 * np.log2(nphi) * 4
    want = O.fft_r2c(nphi, y.ravel(), rescale=True)
    t = dev(y)
    ra.fft_to_spectral_rsc(t, nphi)
    assert rel_max(host(t).ravel(), want) < 1e-14 * np.log2(nphi) * 4


def test_fft_unsupported_length_is_an_error(ra):
    """Lengths with a prime factor above 7 are refused, not computed wrongly."""
    from rayleigh_amd._lib import RayshtError
    t = dev(np.zeros((2, 24)))
    with pytest.raises(RayshtError):
        ra.fft_to_physical(t, 22)


def test_fft_round_trip_large_batch(ra):
    import torch
    nphi, nlines = 768, 384 * 16
    g = torch.Generator(device="cuda").manual_seed(1)
    x = torch.zeros(nlines, nphi + 2, dtype=torch.float64, device="cuda")
    x[:, :nphi] = torch.randn(nlines, nphi, dtype=torch.float64, device="cuda", generator=g)
    x0 = x.clone()
    ra.fft_to_spectral(x.view(-1), nphi)
    ra.fft_to_physical(x.view(-1), nphi)
    torch.cuda.synchronize()
    err = ((x[:, :nphi] / nphi - x0[:, :nphi]).abs().max() / x0.abs().max()).item()
    assert err < 1e-14


@pytest.mark.parametrize("nt,lmax,nr,nf", [(12, 7, 3, 2), (96, 63, 4, 2)])
def test_permutes_vs_oracle(ra, nt, lmax, nr, nf):
    nphi = 2 * nt
    mv = O.m_order_balanced(lmax)
    plan = ra.LegendrePlan(nt, lmax, mv, nr, nf)
    rng = np.random.default_rng(nt)
    p2 = rng.standard_normal(plan.hybrid_size)
    want = O.permute_2a3a(nt, lmax, mv, nr, nf, nphi, p2)
    p3 = empty(want.size)
    plan.reform_2a3a(dev(p2), p3, nphi)
    assert np.array_equal(host(p3), want)
    p3b = rng.standard_normal(want.size)
    want2 = O.permute_3b2b(nt, lmax, mv, nr, nf, nphi, p3b)
    p2b = empty(plan.hybrid_size)
    plan.reform_3b2b(dev(p3b), p2b, nphi)
    assert np.array_equal(host(p2b), want2)


def test_full_stack_loop(ra):
    """Test_SHT.F90:449-575 (test_LT_FFT): StP -> 2a3a -> c2r -> r2c -> 3b2b -> PtS is the
    identity on band-limited data, and the fused entry points equal the separate steps."""
    nt, L, nr, nf = 96, 63, 4, 5
    nphi = 2 * nt
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf)
    off = O.spectral_offsets(L, mv, nr, nf)
    s = np.zeros(off[-1])
    for mp, m in enumerate(mv):
        blk = s[off[mp]: off[mp + 1]].reshape(nf, 2, nr, L - m + 1)
        for f in range(1, nf + 1):
            for r in range(1, nr + 1):
                blk[f - 1, 0, r - 1, : L - m] = (np.arange(m, L) + 1.0) * (2 * f + 3) * r
                blk[f - 1, 1, r - 1, : L - m] = (m + 1.0) * (4 * f + 5) * r * r
        # per-m normalisation; the m = l_max block is all zero (only l = l_max, which is zeroed)
        n1, n2 = np.abs(blk[:, 0]).max(), np.abs(blk[:, 1]).max()
        if n1 > 0:
            blk[:, 0] /= n1
        if n2 > 0:
            blk[:, 1] /= n2
        if m == 0:
            blk[:, 1] = 0.0
    n3 = (nphi + 2) * nr * nt * nf
    w = empty(plan.hybrid_size)
    p3 = empty(n3)
    plan.spectral_to_physical(dev(s), w, p3, nphi)
    # the same through the separate entry points
    p2a = empty(plan.hybrid_size)
    plan.stp(dev(s), p2a)
    p3s = empty(n3)
    plan.reform_2a3a(p2a, p3s, nphi)
    ra.fft_to_physical(p3s, nphi)
    assert np.array_equal(host(p3), host(p3s))
    # against the oracle's stack
    tabs = O.legendre_tables(nt, L, mv)
    p3o = O.fft_c2r(nphi, O.permute_2a3a(nt, L, mv, nr, nf, nphi, O.stp(nt, L, mv, nr, nf, tabs, s)))
    assert rel_max(host(p3), p3o) < 1e-13
    s2 = empty(plan.spectral_size)
    plan.physical_to_spectral(p3, w, s2, nphi)
    s2o = O.pts(nt, L, mv, nr, nf, tabs, O.permute_3b2b(nt, L, mv, nr, nf, nphi, O.fft_r2c(nphi, p3o)))
    assert rel_max(host(s2), s2o) < 1e-13
    assert np.abs(host(s2) - s).max() < 1e-10  # Test_SHT.F90:34 passing tolerance


@pytest.mark.parametrize("n", [48, 33, 128, 256])
def test_cheby_vs_oracle(ra, n):
    """The table form (Chebyshev_Polynomials.F90:375-558 as written) at <= 1e-14."""
    grid = ra.ChebyGrid(n, 0.35, 1.0)
    grid.set_form("matrix")
    assert np.array_equal(grid.grid, O.cheby_grid(n, 0.35, 1.0))
    ce, co = grid.tables()
    oe, oo = O.cheby_tables(n)
    assert np.array_equal(ce, oe) and np.array_equal(co, oo[: co.size])
    rng = np.random.default_rng(n)
    ncols = 1000
    f = rng.standard_normal(n * ncols)
    c = empty(f.size)
    grid.to_spectral(dev(f), c)
    want = O.cheby_to_spectral(n, f)
    assert rel_max(host(c), want) < 1e-14
    f2 = empty(f.size)
    grid.from_spectral(dev(want), f2)
    assert rel_max(host(f2), O.cheby_from_spectral(n, want)) < 1e-14
    d = dev(want)
    grid.dealias_buffer(d)
    assert np.array_equal(host(d), O.cheby_dealias(n, want))


@pytest.mark.parametrize("n", [300, 512, 777, 1024])
def test_cheby_long_columns_vs_oracle(ra, n):
    """n > 256: the LDS-staged table form (one wave per 16 columns, up to 131 KB of LDS at
    n = 1024) against the oracle; the sums' rounding grows with their n / 2 terms."""
    grid = ra.ChebyGrid(n, 0.35, 1.0)
    assert np.array_equal(grid.grid, O.cheby_grid(n, 0.35, 1.0))
    rng = np.random.default_rng(n)
    ncols = 300
    f = rng.standard_normal(n * ncols)
    c = empty(f.size)
    grid.to_spectral(dev(f), c)
    want = O.cheby_to_spectral(n, f)
    tol = 1e-14 * n / 256
    assert rel_max(host(c), want) < tol
    f2 = empty(f.size)
    grid.from_spectral(dev(want), f2)
    assert rel_max(host(f2), O.cheby_from_spectral(n, want)) < tol
    d = dev(want)
    grid.dealias_buffer(d)
    assert np.array_equal(host(d), O.cheby_dealias(n, want))
    d2 = empty(f.size)
    grid.d_by_dr(dev(want), d2, dorder=2)
    assert np.array_equal(host(d2), O.cheby_deriv(n, 0.35, 1.0, want, 2))


DCT_TOL = 3e-14  # against the table form's oracle (the table form's own rounding is ~1e-14)


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256])
def test_cheby_dct_form(ra, n):
    """The DCT form (the default for n = 128 and 256) against the table form's oracle, and against
    the exact sums on the reference's own angles (long double), which it meets to 1e-15
    where the table form, whose cosines are rounded doubles of rounded k theta_i, is up to
    about 1e-14 away; a partial last tile (1003 columns) writes nothing past the buffer."""
    import torch
    grid = ra.ChebyGrid(n, 0.35, 1.0)
    grid.set_form("dct")
    rng = np.random.default_rng(7 * n)
    ncols = 1003
    f = rng.standard_normal(n * ncols)
    sub = slice(0, n * 48)
    buf = torch.full((n * ncols + 2 * n,), float("nan"), dtype=torch.float64, device="cuda")
    c = buf[: n * ncols]
    grid.to_spectral(dev(f), c)
    got = host(c)
    assert np.isnan(host(buf[n * ncols:])).all()
    assert rel_max(got, O.cheby_to_spectral(n, f)) < DCT_TOL
    assert rel_max(got[sub], O.cheby_to_spectral_exact(n, f[sub])) < 1e-15
    buf.fill_(float("nan"))
    grid.from_spectral(dev(got), c)
    back = host(c)
    assert np.isnan(host(buf[n * ncols:])).all()
    assert rel_max(back, O.cheby_from_spectral(n, got)) < DCT_TOL
    assert rel_max(back[sub], O.cheby_from_spectral_exact(n, got[sub])) < 1e-15
    # the round trip: the reference's own pair returns f to 3.6e-13 at n = 256 (its tables'
    # angles are not exactly the zeros grid's, so the pair is not exactly orthogonal)
    assert rel_max(back, f) < 1e-12


@pytest.mark.parametrize("n", [64, 256])
def test_cheby_dct_unaligned_and_smooth(ra, n):
    """Columns that start 8-byte aligned take the table form; a smooth radial profile
    (exp(-r)) keeps its rapidly decaying coefficients to the same tolerance."""
    import torch
    grid = ra.ChebyGrid(n, 0.35, 1.0)
    grid.set_form("dct")
    r = O.cheby_grid(n, 0.35, 1.0)
    f = np.concatenate([np.exp(-r * (1 + 0.1 * k)) for k in range(40)])
    want = O.cheby_to_spectral(n, f)
    buf = torch.zeros(f.size + 1, dtype=torch.float64, device="cuda")
    buf[1:].copy_(torch.from_numpy(f))
    out = torch.zeros(f.size + 1, dtype=torch.float64, device="cuda")
    grid.to_spectral(buf[1:], out[1:])
    assert rel_max(host(out[1:]), want) < 1e-14
    c = empty(f.size)
    grid.to_spectral(dev(f), c)
    assert rel_max(host(c), want) < DCT_TOL
    with pytest.raises(Exception):
        ra.ChebyGrid(48, 0.35, 1.0).set_form("dct")


def test_cheby_vs_reference_fixture(ra, golden):
    for name in ("cheby_n48.npz", "cheby_n33.npz"):
        g = golden(name)
        n = int(g["n"])
        grid = ra.ChebyGrid(n, float(g["rmin"]), float(g["rmax"]))
        c = empty(g["f_in"].size)
        grid.to_spectral(dev(g["f_in"]), c)
        assert rel_max(host(c), g["c_out"]) < 1e-14
        f = empty(g["c_out"].size)
        grid.from_spectral(dev(g["c_out"]), f)
        assert rel_max(host(f), g["f_rt"]) < 1e-14


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_round_trip_graph_replays(ra, precision):
    """The round trip recorded into a hipGraph: every replay (with eager work and
    synchronisations between replays) equals the eager round trip bit for bit, and the
    round trip is the identity N_phi-scaled on band-limited data."""
    import torch
    nt, L, nr, nf = 96, 63, 6, 3
    nphi = 2 * nt
    mv = O.m_order_balanced(L)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf, precision=precision)
    s_np = O.synthetic_spectral(L, mv, nr, nf, seed=21)
    n3 = (nphi + 2) * nr * nt * nf
    s_in, w, p3 = dev(s_np), empty(plan.hybrid_size), empty(n3)
    eager = empty(plan.spectral_size)
    plan.spectral_to_physical(s_in, w, p3, nphi)
    plan.physical_to_spectral(p3, w, eager, nphi)
    want = host(eager)
    # the round trip: StP's 1/2 (m > 0) and PtS's 1/N_theta-type norms give the identity
    assert rel_max(want, s_np) < 1e-11
    s_out = empty(plan.spectral_size)
    g = plan.round_trip_graph(nphi, s_in, w, p3, s_out)
    for rep in range(4):
        s_out.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(host(s_out), want), f"replay {rep}"
        # eager work between the replays, on the same buffers
        plan.stp(s_in, w)
        torch.cuda.synchronize()


@pytest.mark.parametrize("n,rmin,rmax", [(48, 0.35, 1.0), (33, 0.5, 1.7), (256, 0.35, 1.0)])
def test_cheb_grid_data_vs_oracle(ra, n, rmin, rmax):
    """raysht_cheb_grid_data: Cheby_Grid%init's integration weights and components as the
    oracle (bit-exact with the compiled reference) forms them."""
    import ctypes
    from rayleigh_amd import _lib
    grid = ra.ChebyGrid(n, rmin, rmax)
    x, th, w, ds = (np.zeros(n) for _ in range(4))
    P = ctypes.c_void_p
    assert _lib.lib().raysht_cheb_grid_data(grid._h, x.ctypes.data_as(P), th.ctypes.data_as(P),
                                            w.ctypes.data_as(P), ds.ctypes.data_as(P)) == 0
    assert np.array_equal(w, O.cheby_weights(n, rmin, rmax))
    assert np.array_equal(grid.grid, O.cheby_grid(n, rmin, rmax))
    assert np.abs(np.cos(th) - x).max() <= 2.3e-16 and np.all(np.diff(x) < 0)
    assert np.all(ds == O.cheby_deriv_scaling(n, rmin, rmax))


def test_permutes_host_forms_equal_device_forms(ra):
    """The host-array permutes (what the Fortran Reform_* call) give the device forms' bits."""
    import ctypes
    from rayleigh_amd import _lib
    P = ctypes.c_void_p
    nt, lmax, nr, nf = 24, 15, 3, 2
    nphi = 2 * nt
    mv = O.m_order_balanced(lmax)
    plan = ra.LegendrePlan(nt, lmax, mv, nr, nf)
    lib = _lib.lib()
    rng = np.random.default_rng(9)
    p2 = rng.standard_normal(plan.hybrid_size)
    n3 = (nphi + 2) * nr * nt * nf
    p3 = np.zeros(n3)
    assert lib.raysht_permute_2a3a_host(plan.handle, nphi, p2.ctypes.data_as(P), p3.ctypes.data_as(P)) == 0
    t3 = empty(n3)
    plan.reform_2a3a(dev(p2), t3, nphi)
    assert np.array_equal(p3, host(t3))
    p3b = rng.standard_normal(n3)
    p2b = np.zeros(plan.hybrid_size)
    assert lib.raysht_permute_3b2b_host(plan.handle, nphi, p3b.ctypes.data_as(P), p2b.ctypes.data_as(P)) == 0
    t2 = empty(plan.hybrid_size)
    plan.reform_3b2b(dev(p3b), t2, nphi)
    assert np.array_equal(p2b, host(t2))
    # radial: blocks through a pointer array
    s = rng.standard_normal(plan.spectral_size)
    off = O.spectral_offsets(lmax, mv, nr, nf)
    ptrs = (P * len(mv))(*[s.ctypes.data + 8 * int(off[i]) for i in range(len(mv))])
    p1 = np.zeros(plan.spectral_size)
    assert lib.raysht_permute_2b1b_host(plan.handle, ptrs, p1.ctypes.data_as(P)) == 0
    t1 = empty(plan.spectral_size)
    plan.transpose_2b1b(dev(s), t1)
    assert np.array_equal(p1, host(t1))
    s2 = np.zeros(plan.spectral_size)
    ptrs2 = (P * len(mv))(*[s2.ctypes.data + 8 * int(off[i]) for i in range(len(mv))])
    assert lib.raysht_permute_1a2a_host(plan.handle, p1.ctypes.data_as(P), ptrs2) == 0
    assert np.array_equal(s2, s)


@pytest.mark.parametrize("nt,lmax,nr,nf,nphi,msub", [
    (384, 255, 3, 2, 768, False),   # N = 384: four lines per workgroup
    (98, 63, 2, 3, 192, False),     # N = 96: 98 theta, a partial last theta block
    (64, 41, 3, 1, 128, True),      # a plan without every m (the others read and write nothing)
    (24, 1023, 1, 1, 3072, False),  # N = 1536: two lines per workgroup
    (40, 19, 2, 2, 60, False),      # N = 30: not register-resident, the separate kernels
])
def test_fused_transposes_equal_separate(ra, nt, lmax, nr, nf, nphi, msub):
    """spectral_to_physical / physical_to_spectral (the 3b -> 2b transpose inside the r2c
    kernel where the length allows) equal StP -> 2a3a -> c2r and r2c -> 3b2b -> PtS bit
    for bit."""
    mv = np.arange(0, lmax + 1, 2 if msub else 1, dtype=np.int32)
    mv = mv[np.random.default_rng(nt).permutation(len(mv))]
    plan = ra.LegendrePlan(nt, lmax, mv, nr, nf)
    s = dev(O.synthetic_spectral(lmax, mv, nr, nf, seed=nt + lmax))
    n3 = (nphi + 2) * nr * nt * nf
    w = empty(plan.hybrid_size)
    p3 = empty(n3)
    plan.spectral_to_physical(s, w, p3, nphi)
    grid = host(p3)  # physical_to_spectral may transform p3 in place (lengths not fused)
    p2a = empty(plan.hybrid_size)
    plan.stp(s, p2a)
    p3s = empty(n3)
    plan.reform_2a3a(p2a, p3s, nphi)
    ra.fft_to_physical(p3s, nphi)
    assert np.array_equal(host(p3), host(p3s))
    # physical -> spectral on the same grid: the fused form against the separate steps
    s2 = empty(plan.spectral_size)
    plan.physical_to_spectral(p3, w, s2, nphi)
    ra.fft_to_spectral(p3s, nphi)
    p2b = empty(plan.hybrid_size)
    plan.reform_3b2b(p3s, p2b, nphi)
    s2s = empty(plan.spectral_size)
    plan.pts(p2b, s2s)
    assert np.array_equal(host(s2), host(s2s))
    # the reform + FFT calls on their own
    p3r = empty(n3)
    plan.reform_to_physical(p2a, p3r, nphi)
    assert np.array_equal(host(p3r), grid)
    p2r = empty(plan.hybrid_size)
    plan.reform_to_spectral(p3r, p2r, nphi)
    assert np.array_equal(host(p2r), host(p2b))
