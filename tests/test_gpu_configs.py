"""The benchmark configurations at full size on the GPU.

* l_max = 255 stack (N_theta = 384, N_phi = 768, N_r = 128, 8 fields) and l_max = 1023
  stack (N_theta = 1536, N_phi = 3072, N_r = 256, 5 fields): Chebyshev From_Spectral ->
  Transpose_1a2a -> StP -> 2a->3a -> c2r, then r2c -> 3b->2b -> PtS -> 2b->1b ->
  Chebyshev To_Spectral, chained on the same device buffers.  Every stage is checked
  against the oracle on a sub-problem cut from its own (device) input -- sampled
  m-blocks, radii, fields or radial columns, since the stages are independent per
  m-block and per column -- and the whole chain must return its input
  (Test_SHT.F90:449-575 test_LT_FFT, with the radial stages around it).
* the same l_max = 1023 Legendre round trip in the split forms against FP64.
* 64 independent problems (l_max 31-1023) in one grouped launch per direction: FP64
  bit-equal to one plan at a time, sampled problems against the oracle; BF16X6 within
  1e-12 of FP64 per problem.

Tolerances: every stage <= 1e-13 max|d| / max|ref| against the oracle (the FP64 bar);
the chain's round trip <= 1e-12 (decaying coefficients; the reference's own test
passes at 1e-10, Test_SHT.F90:34); split X6 forms <= 1e-12 against FP64.
"""
import gc

import numpy as np
import pytest

from oracle import oracle as O
from tests import workloads as W

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-13
ROUND_TRIP_TOL = 1e-12
X6_TOL = 1e-12


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


def _free():
    import torch
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _empty(n):
    import torch
    return torch.empty(int(n), dtype=torch.float64, device="cuda")


def _sample_mps(mv, lmax):
    """m-block positions of m = 0, 1, l_max/2, l_max - 1, l_max (the longest, a short and
    the one-coefficient block)."""
    want = [0, 1, lmax // 2, lmax - 1, lmax]
    return [int(np.nonzero(mv == m)[0][0]) for m in want]


def run_stack(ra, cfg, seed, r_sample):
    """The chained stack at full size with a stage-by-stage oracle check; returns the
    round-trip error of the Chebyshev coefficients."""
    import torch
    L, nt, nphi, nr, nf = cfg["lmax"], cfg["nt"], cfg["nphi"], cfg["nr"], cfg["nf"]
    print(f"\n[config stack] l_max={L} N_theta={nt} N_phi={nphi} N_r={nr} N_f={nf}", flush=True)
    mv = ra.m_values_balanced(L)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf)
    grid = ra.ChebyGrid(nr, W.RMIN, W.RMAX)
    l_lm, mp_lm = plan.lm_values()
    nlm = len(l_lm)
    rng = np.random.default_rng(seed)
    mps = _sample_mps(mv, L)
    rs = np.sort(rng.choice(nr, r_sample, replace=False)) if r_sample < nr else np.arange(nr)
    rs2 = np.sort(rng.choice(nr, 2, replace=False))
    fs2 = np.array([0, nf - 1])
    cols = np.sort(rng.choice(2 * nlm * nf, 4096, replace=False))
    tabs_sub = O.legendre_tables(nt, L, mv[mps])

    c = W.cheb_coeffs_device(L, l_lm, mv[mp_lm], nr, nf, seed)
    f1 = torch.empty_like(c)
    grid.from_spectral(c, f1)
    ci = W.column_index(nr, cols)
    err = W.rel_max(W.gather(f1, ci), O.cheby_from_spectral(nr, W.gather(c, ci)))
    assert err <= FP64_TOL, f"From_Spectral {err}"

    s = _empty(plan.spectral_size)
    plan.transpose_1a2a(f1, s)
    got = W.gather(s, W.spectral_index(L, mv, nr, nf, range(len(mv)), rs2, fs2))
    want = O.permute_1a2a(L, mv, 2, 2, W.gather(f1, W.radial_index(nlm, nr, nf, rs2, fs2)))
    assert np.array_equal(got, want), "Transpose_1a2a"

    p2 = _empty(plan.hybrid_size)
    plan.stp(s, p2)
    want = O.stp(nt, L, mv[mps], len(rs), nf, tabs_sub, W.gather(s, W.spectral_index(L, mv, nr, nf, mps, rs)))
    err = W.rel_max(W.gather(p2, W.hybrid_index(nt, nr, nf, mps, rs)), want)
    assert err <= FP64_TOL, f"StP {err}"

    p3 = _empty((nphi + 2) * nr * nt * nf)
    plan.reform_2a3a(p2, p3, nphi)
    ra.fft_to_physical(p3, nphi)
    gi = W.grid_index(nt, nr, nf, nphi, rs2, fs2)
    phys = W.gather(p3, gi)
    want = O.fft_c2r(nphi, O.permute_2a3a(nt, L, mv, 2, 2, nphi,
                                          W.gather(p2, W.hybrid_index(nt, nr, nf, range(len(mv)), rs2, fs2))))
    err = W.rel_max(phys, want)
    assert err <= FP64_TOL, f"2a3a + c2r {err}"

    ra.fft_to_spectral(p3, nphi)
    plan.reform_3b2b(p3, p2, nphi)
    del p3
    want = O.permute_3b2b(nt, L, mv, 2, 2, nphi, O.fft_r2c(nphi, phys))
    err = W.rel_max(W.gather(p2, W.hybrid_index(nt, nr, nf, range(len(mv)), rs2, fs2)), want)
    assert err <= FP64_TOL, f"r2c + 3b2b {err}"

    plan.pts(p2, s)
    want = O.pts(nt, L, mv[mps], len(rs), nf, tabs_sub, W.gather(p2, W.hybrid_index(nt, nr, nf, mps, rs)))
    err = W.rel_max(W.gather(s, W.spectral_index(L, mv, nr, nf, mps, rs)), want)
    assert err <= FP64_TOL, f"PtS {err}"
    del p2

    plan.transpose_2b1b(s, f1)
    got = W.gather(f1, W.radial_index(nlm, nr, nf, rs2, fs2))
    want = O.permute_2b1b(L, mv, 2, 2, W.gather(s, W.spectral_index(L, mv, nr, nf, range(len(mv)), rs2, fs2)))
    assert np.array_equal(got, want), "Transpose_2b1b"

    c2 = torch.empty_like(c)
    grid.to_spectral(f1, c2)
    err = W.rel_max(W.gather(c2, ci), O.cheby_to_spectral(nr, W.gather(f1, ci)))
    assert err <= FP64_TOL, f"To_Spectral {err}"
    rt = W.rel_max_dev(c2, c)
    print(f"[config stack] l_max={L}: every stage within {FP64_TOL:g} of the oracle; round trip {rt:.3e}",
          flush=True)
    return rt


def test_stack_lmax255_full_size(ra):
    _free()
    rt = run_stack(ra, W.CONFIG1, seed=1, r_sample=W.CONFIG1["nr"])
    assert rt <= ROUND_TRIP_TOL
    _free()


def test_stack_lmax1023_full_size(ra):
    _free()
    rt = run_stack(ra, W.CONFIG4, seed=4, r_sample=12)
    assert rt <= ROUND_TRIP_TOL
    _free()


@pytest.mark.parametrize("precision", ["bf16x6", "int8x6"])
def test_split_lmax1023_full_size(ra, precision):
    """StP and PtS of an X6 split form at l_max = 1023 (N_theta = 1536, N_r = 256, 5 fields)
    against the FP64 path on the same inputs, and the split round trip N_phi PtS(StP(a)) = a."""
    import torch
    _free()
    cfg = W.CONFIG4
    L, nt, nr, nf = cfg["lmax"], cfg["nt"], cfg["nr"], cfg["nf"]
    print(f"\n[config split] {precision} l_max={L}", flush=True)
    mv = ra.m_values_balanced(L)
    s = W.spectral_device(L, mv, nr, nf, seed=44)
    ref = ra.LegendrePlan(nt, L, mv, nr, nf)
    p64 = _empty(ref.hybrid_size)
    ref.stp(s, p64)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf, precision=precision)
    px = _empty(plan.hybrid_size)
    plan.stp(s, px)
    e_stp = W.rel_max_dev(px, p64)
    del px
    s64 = torch.empty_like(s)
    ref.pts(p64, s64)
    sx = torch.empty_like(s)
    plan.pts(p64, sx)
    e_pts = W.rel_max_dev(sx, s64)
    # the split round trip on its own StP output
    plan.stp(s, p64)
    plan.pts(p64, sx)
    rt = float(((sx * (2 * nt) - s).abs().max() / s.abs().max()).item())
    print(f"[config split] {precision}: StP {e_stp:.3e}, PtS {e_pts:.3e}, round trip {rt:.3e}", flush=True)
    assert e_stp <= X6_TOL and e_pts <= X6_TOL
    assert rt <= ROUND_TRIP_TOL
    del ref, plan, s, p64, s64, sx
    _free()


def _batch_inputs(ra, probs, precision):
    plans, ins = [], []
    for i, (L, nt, nr, nf) in enumerate(probs):
        mv = np.arange(L + 1, dtype=np.int32)
        plans.append(ra.LegendrePlan(nt, L, mv, nr, nf, precision=precision))
        ins.append(W.spectral_device(L, mv, nr, nf, seed=500 + i))
    return plans, ins


def test_batch64_fp64(ra):
    """The 64-problem grouped launch (l_max log-uniform in [31, 1023]): every problem's
    output equals its own plan's launch bit for bit, and three problems (the largest, the
    smallest, one between) agree with the oracle on sampled m-blocks and radii."""
    import torch
    _free()
    probs = W.batch_problems(3)
    print(f"\n[config batch] fp64: {len(probs)} problems", flush=True)
    plans, s = _batch_inputs(ra, probs, "fp64")
    batch = ra.LegendreBatch(plans)
    p = [_empty(pl.hybrid_size) for pl in plans]
    batch.stp(s, p)
    so = [_empty(pl.spectral_size) for pl in plans]
    batch.pts(p, so)
    for i, pl in enumerate(plans):
        t = _empty(pl.hybrid_size)
        pl.stp(s[i], t)
        assert torch.equal(t, p[i]), f"StP problem {i} {probs[i]}"
        t2 = _empty(pl.spectral_size)
        pl.pts(p[i], t2)
        assert torch.equal(t2, so[i]), f"PtS problem {i} {probs[i]}"
        del t, t2
        if i % 16 == 15:
            print(f"[config batch] fp64: {i + 1} problems equal to their own plans", flush=True)
    sizes = [q[0] for q in probs]
    rng = np.random.default_rng(33)
    for i in {int(np.argmax(sizes)), int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2])}:
        L, nt, nr, nf = probs[i]
        mv = np.arange(L + 1, dtype=np.int32)
        mps = _sample_mps(mv, L)
        rs = np.sort(rng.choice(nr, min(nr, 8), replace=False))
        tabs = O.legendre_tables(nt, L, mv[mps])
        want = O.stp(nt, L, mv[mps], len(rs), nf, tabs, W.gather(s[i], W.spectral_index(L, mv, nr, nf, mps, rs)))
        err = W.rel_max(W.gather(p[i], W.hybrid_index(nt, nr, nf, mps, rs)), want)
        assert err <= FP64_TOL, f"StP problem {i} {probs[i]}: {err}"
        want = O.pts(nt, L, mv[mps], len(rs), nf, tabs, W.gather(p[i], W.hybrid_index(nt, nr, nf, mps, rs)))
        err = W.rel_max(W.gather(so[i], W.spectral_index(L, mv, nr, nf, mps, rs)), want)
        assert err <= FP64_TOL, f"PtS problem {i} {probs[i]}: {err}"
    del batch, plans, s, p, so
    _free()


def test_batch64_bf16x6(ra):
    """The 64-problem grouped launch in BF16X6: within 1e-12 of each problem's FP64 plan
    (StP on the same spectral input, PtS on the same hybrid input), and bit-equal to the
    BF16X6 plans launched one at a time."""
    import torch
    _free()
    probs = W.batch_problems(3)
    print(f"\n[config batch] bf16x6: {len(probs)} problems", flush=True)
    plans, s = _batch_inputs(ra, probs, "bf16x6")
    batch = ra.LegendreBatch(plans)
    p = [_empty(pl.hybrid_size) for pl in plans]
    batch.stp(s, p)
    so = [_empty(pl.spectral_size) for pl in plans]
    batch.pts(p, so)
    worst = (0.0, 0.0)
    for i, (L, nt, nr, nf) in enumerate(probs):
        mv = np.arange(L + 1, dtype=np.int32)
        t = _empty(plans[i].hybrid_size)
        plans[i].stp(s[i], t)
        assert torch.equal(t, p[i]), f"StP problem {i} {probs[i]}"
        t2 = _empty(plans[i].spectral_size)
        plans[i].pts(p[i], t2)
        assert torch.equal(t2, so[i]), f"PtS problem {i} {probs[i]}"
        ref = ra.LegendrePlan(nt, L, mv, nr, nf)
        ref.stp(s[i], t)
        e1 = W.rel_max_dev(p[i], t)
        ref.pts(p[i], t2)
        e2 = W.rel_max_dev(so[i], t2)
        worst = (max(worst[0], e1), max(worst[1], e2))
        assert e1 <= X6_TOL and e2 <= X6_TOL, f"problem {i} {probs[i]}: StP {e1}, PtS {e2}"
        del ref, t, t2
        if i % 16 == 15:
            print(f"[config batch] bf16x6: {i + 1} problems checked", flush=True)
    print(f"[config batch] bf16x6 worst StP {worst[0]:.3e}, PtS {worst[1]:.3e}", flush=True)
    del batch, plans, s, p, so
    _free()
