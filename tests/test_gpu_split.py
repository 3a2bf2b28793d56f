"""GPU checks of the split-precision Legendre forms (BF16X6, BF16X3 and the same digits
on INT8 MFMA: INT8X6, INT8X3) against the FP64 oracle and the device FP64 path, at the
error tiers the forms are built for: X6 <= 1e-12 and X3 <= 1e-6 of max |result| (RMS
<= 1e-6).  The X3 forms' StP keeps 4 x 4 digits with s + t <= 3 (10 pairs): with 3 x 3
(s + t <= 2) the dropped level s + t = 3 (weight 2^-24 of the operand scales, summed over
up to 384 K terms) left 2e-5 - 4e-5 at l_max = 511 on spectra decaying as (l+1)^-2; the
folded PtS data do not decay and 3 x 3 stays near 2e-7."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import dev, empty, host, rel_max

pytestmark = pytest.mark.gpu

TOL = {"bf16x6": 1e-12, "bf16x3": 1e-6, "int8x6": 1e-12, "int8x3": 1e-6}
TOL_RMS = {"bf16x3": 1e-6, "int8x3": 1e-6}
FORMS = ["bf16x6", "bf16x3", "int8x6", "int8x3"]


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2) / max(np.mean(b ** 2), 1e-300)))


def check(form, got, want):
    err = rel_max(got, want)
    assert err <= TOL[form], f"{form}: max rel {err:.3e}"
    if form in TOL_RMS:
        assert rel_rms(got, want) <= TOL_RMS[form], f"{form}: rms rel {rel_rms(got, want):.3e}"
    return err


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("nt,lmax,nr,nf,order", [
    (96, 63, 48, 4, "balanced"),    # l_max=63, N_r=48, 4 fields
    (64, 41, 5, 3, "balanced"),     # 30 columns: partial column tile
    (20, 12, 3, 1, "ascending"),    # 10 southern nodes: partial theta block
    (2, 0, 2, 1, "ascending"),      # l_max = 0
    (4, 1, 1, 1, "ascending"),      # l_max = 1
])
def test_split_vs_oracle(ra, form, nt, lmax, nr, nf, order):
    mv = O.m_order_balanced(lmax) if order == "balanced" else np.arange(lmax + 1, dtype=np.int32)
    plan = ra.LegendrePlan(nt, lmax, mv, nr, nf, precision=form)
    tabs = O.legendre_tables(nt, lmax, mv)
    s_np = O.synthetic_spectral(lmax, mv, nr, nf, seed=3 * nt + lmax)
    p_ref = O.stp(nt, lmax, mv, nr, nf, tabs, s_np)
    p = empty(plan.hybrid_size)
    plan.stp(dev(s_np), p)
    check(form, host(p), p_ref)
    s_ref = O.pts(nt, lmax, mv, nr, nf, tabs, p_ref)
    s = empty(plan.spectral_size)
    plan.pts(dev(p_ref), s)
    check(form, host(s), s_ref)


@pytest.mark.parametrize("form", FORMS)
def test_split_long_contractions(ra, form):
    """l_max = 1023 on 1536 nodes: StP sums 512 terms and PtS 768, so the FP32 level
    sums are moved into FP64 two and three times."""
    nt, lmax, nr, nf = 1536, 1023, 2, 1
    mv = np.array([0, 1, 700, 1023], dtype=np.int32)
    plan = ra.LegendrePlan(nt, lmax, mv, nr, nf, precision=form)
    tabs = O.legendre_tables(nt, lmax, mv)
    s_np = O.synthetic_spectral(lmax, mv, nr, nf, seed=11)
    p_ref = O.stp(nt, lmax, mv, nr, nf, tabs, s_np)
    p = empty(plan.hybrid_size)
    plan.stp(dev(s_np), p)
    check(form, host(p), p_ref)
    s_ref = O.pts(nt, lmax, mv, nr, nf, tabs, p_ref)
    s = empty(plan.spectral_size)
    plan.pts(dev(p_ref), s)
    check(form, host(s), s_ref)


@pytest.mark.parametrize("form", FORMS)
def test_split_cols_plan(ra, form):
    """The rmcontainer pair (anonymous columns) in the split forms."""
    nt, lmax, ncols = 48, 30, 7
    mv = np.array([3, 0, 30, 17], dtype=np.int32)
    plan = ra.LegendrePlan(nt, lmax, mv, 0, 0, precision=form, n_cols=ncols)
    tabs = O.legendre_tables(nt, lmax, mv)
    rng = np.random.default_rng(5)
    s_np = rng.standard_normal(plan.spectral_size)
    p_ref = O.stp_cols(nt, lmax, mv, ncols, tabs, s_np)
    p = empty(plan.hybrid_size)
    plan.stp(dev(s_np), p)
    check(form, host(p), p_ref)
    s_ref = O.pts_cols(nt, lmax, mv, ncols, tabs, p_ref)
    s = empty(plan.spectral_size)
    plan.pts(dev(p_ref), s)
    check(form, host(s), s_ref)


def test_split_full_size_vs_device_fp64(ra):
    """l_max = 511, N_theta = 768, N_r = 192, 8 fields (3072 columns) on four m-blocks
    of very different lengths, against the FP64 path (itself within 1e-13 of the
    reference); per-field errors as well as the whole output."""
    import torch
    nt, lmax, nr, nf = 768, 511, 192, 8
    mv = np.array([0, 1, 256, 511], dtype=np.int32)
    s_np = O.synthetic_spectral(lmax, mv, nr, nf, seed=2)
    s = dev(s_np)
    ref = ra.LegendrePlan(nt, lmax, mv, nr, nf)
    p64 = empty(ref.hybrid_size)
    ref.stp(s, p64)
    s64 = empty(ref.spectral_size)
    ref.pts(p64, s64)
    for form in FORMS:
        plan = ra.LegendrePlan(nt, lmax, mv, nr, nf, precision=form)
        p = empty(plan.hybrid_size)
        plan.stp(s, p)
        got, want = host(p), host(p64)
        check(form, got, want)
        # per field: hybrid block mp holds columns c = r + imi*nr + 2*nr*f
        g4 = got.reshape(len(mv), nf, 2 * nr, nt)
        w4 = want.reshape(len(mv), nf, 2 * nr, nt)
        for f in range(nf):
            assert rel_max(g4[:, f], w4[:, f]) <= TOL[form], (form, f)
        s2 = empty(plan.spectral_size)
        plan.pts(p64, s2)
        check(form, host(s2), host(s64))
        del plan
        torch.cuda.empty_cache()


@pytest.mark.parametrize("form", ["bf16x6", "int8x6"])
def test_split_round_trip_identity(ra, form):
    """N_phi PtS(StP(s)) = s for band-limited s (N_phi = 2 N_theta): the X6 forms keep the
    round trip near the FP64 path's level (~1e-13), within a 1e-11 bound."""
    nt, lmax, nr, nf = 192, 127, 8, 2
    mv = O.m_order_balanced(lmax)
    plan = ra.LegendrePlan(nt, lmax, mv, nr, nf, precision=form)
    s_np = O.synthetic_spectral(lmax, mv, nr, nf, seed=9)
    p = empty(plan.hybrid_size)
    plan.stp(dev(s_np), p)
    s2 = empty(plan.spectral_size)
    plan.pts(p, s2)
    assert rel_max(host(s2) * (2 * nt), s_np) < 1e-11


def test_int8_refuses_long_contractions(ra):
    """The INT8 forms keep exact int32 level sums up to 16384 terms per contraction."""
    from rayleigh_amd._lib import RayshtError
    with pytest.raises(RayshtError):
        ra.LegendrePlan(32770, 3, np.arange(4, dtype=np.int32), 1, 1, precision="int8x6")
