"""Headline benchmark: Legendre round trip (StP then PtS) at l_max = 511.

Workload: l_max = 511, N_theta = 768 Gauss-Legendre nodes, N_r = 192, 8 fields,
all 512 m in Rayleigh's balanced one-rank order; spectral inputs with the
coefficient distribution N(0,1) (l+1)^-2 sin(pi r_i / N_r) with m = 0 real, synthetic,
resident in HBM before the timed region.  One step = StP_4d_dgp2 + PtS_4d_dgpv3
over every m (the HIP library's grouped launches).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--forms fp64,bf16x6,bf16x3]

Prints one JSON line.  value = FP64-equivalent algorithmic TFLOP/s of the round
trip of the first form, 2 * F_LT / step time with F_LT = N_theta N_r N_f (L+1)(L+2).
Every split form is also timed and its max / RMS / per-field error against the FP64
path on the same input is reported under "forms".
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LMAX, NTHETA, NR, NF = 511, 768, 192, 8
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 matrix, dense
PEAK_BF16_TFLOPS = 2500.0    # MI355X BF16 matrix, dense
PEAK_INT8_TOPS = 5000.0      # MI355X INT8 matrix, dense
METRIC = "SHT round-trip ms + TFLOP/s at l_max=511: FP64 vs split-BF16 at 1e-12 rel err"
# slice pairs (one full-size MFMA product each) per direction of the split forms
SPLIT_PAIRS = {"bf16x6": {"stp": 27, "pts": 21}, "bf16x3": {"stp": 10, "pts": 10},
               "int8x6": {"stp": 27, "pts": 21}, "int8x3": {"stp": 10, "pts": 10}}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_ms(fn, steps: int, warmup: int) -> float:
    """Mean HIP-event time of fn over steps calls after max(warmup, 1) untimed ones."""
    import torch
    for _ in range(max(warmup, 1)):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def flops_per_direction(lmax=LMAX, nt=NTHETA, nr=NR, nf=NF) -> float:
    return float(nt) * nr * nf * (lmax + 1) * (lmax + 2)


def synthetic_spectral_device(plan, mvals, seed: int):
    """Coefficients N(0,1) (l+1)^-2 sin(pi r_i / N_r) (m = 0 real), generated on the device."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    s = torch.empty(plan.spectral_size, dtype=torch.float64, device="cuda")
    rfac = torch.sin(torch.pi * torch.arange(1, NR + 1, dtype=torch.float64, device="cuda") / NR)
    off = 0
    for m in mvals:
        m = int(m)
        nl = LMAX - m + 1
        lfac = (torch.arange(m, LMAX + 1, dtype=torch.float64, device="cuda") + 1.0) ** -2
        blk = torch.randn((NF, 2, NR, nl), dtype=torch.float64, device="cuda", generator=g)
        blk *= lfac.view(1, 1, 1, nl)
        blk *= rfac.view(1, 1, NR, 1)
        if m == 0:
            blk[:, 1] = 0.0
        s[off: off + blk.numel()] = blk.reshape(-1)
        off += blk.numel()
    return s


def time_form(ra, form: str, mvals, steps: int, warmup: int):
    import torch
    log(f"form {form}: creating plan (l_max={LMAX}, N_theta={NTHETA}, N_r={NR}, N_f={NF})")
    t0 = time.time()
    plan = ra.LegendrePlan(NTHETA, LMAX, mvals, NR, NF, precision=form)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    s = synthetic_spectral_device(plan, mvals, seed=2024)
    p2 = torch.empty(plan.hybrid_size, dtype=torch.float64, device="cuda")
    s2 = torch.empty_like(s)
    log(f"form {form}: plan ready in {setup_s:.2f} s; warmup {warmup}")
    for _ in range(warmup):
        plan.stp(s, p2)
        plan.pts(p2, s2)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    log(f"form {form}: timing {steps} steps")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(steps):
        ev[k][0].record()
        plan.stp(s, p2)
        ev[k][1].record()
        plan.pts(p2, s2)
        ev[k][2].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    stp_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    pts_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    ms_per_step = wall * 1e3 / steps
    # the round-trip identity on the full problem: N_phi * PtS(StP(a)) = a
    rt = float(((s2 * (2 * NTHETA) - s).abs().max() / s.abs().max()).item())
    log(f"form {form}: {ms_per_step:.3f} ms/step (StP {stp_ms:.3f} ms, PtS {pts_ms:.3f} ms), round trip {rt:.2e}")
    return dict(plan=plan, s=s, p2=p2, s2=s2, setup_s=setup_s, ms_per_step=ms_per_step, stp_ms=stp_ms,
                pts_ms=pts_ms, round_trip_rel_err=rt)


def errors_vs_fp64(ra, res, ref, mvals):
    """Max and RMS relative error of a split form against the FP64 path on the same
    input, for StP (same spectral input) and PtS (run on the FP64 hybrid output), over
    the whole output and per field."""
    import torch
    plan = res["plan"]
    p_form = res["p2"]
    s_form = torch.empty_like(res["s2"])
    plan.pts(ref["p2"], s_form)
    torch.cuda.synchronize()

    def rel(a, b):
        d = (a - b)
        return float(d.abs().max() / b.abs().max()), float(torch.sqrt((d * d).mean() / (b * b).mean()))

    stp_max, stp_rms = rel(p_form, ref["p2"])
    pts_max, pts_rms = rel(s_form, ref["s2"])
    # per field: hybrid block mp holds columns c = r + imi*N_r + 2*N_r*f; spectral block
    # mp is data(m:l_max, N_r, 2, N_f)
    nm = len(mvals)
    dp = (p_form - ref["p2"]).view(nm, NF, 2 * NR * NTHETA).abs().amax(dim=(0, 2))
    rp = ref["p2"].view(nm, NF, 2 * NR * NTHETA).abs().amax(dim=(0, 2))
    ds_f = torch.zeros(NF, dtype=torch.float64, device="cuda")
    rs_f = torch.zeros(NF, dtype=torch.float64, device="cuda")
    off = 0
    for m in mvals:
        n = (LMAX - int(m) + 1) * 2 * NR * NF
        blk_d = (s_form[off:off + n] - ref["s2"][off:off + n]).view(NF, -1).abs().amax(dim=1)
        blk_r = ref["s2"][off:off + n].view(NF, -1).abs().amax(dim=1)
        ds_f = torch.maximum(ds_f, blk_d)
        rs_f = torch.maximum(rs_f, blk_r)
        off += n
    field_stp = (dp / rp).cpu().numpy().tolist()
    field_pts = (ds_f / rs_f).cpu().numpy().tolist()
    del s_form
    return {"stp_max": stp_max, "stp_rms": stp_rms, "pts_max": pts_max, "pts_rms": pts_rms,
            "stp_max_per_field": field_stp, "pts_max_per_field": field_pts}


def reference_check_and_baseline(res, mvals, sample_m: int):
    """Runs the reference's own FP64 StP/PtS (compiled from its sources into
    oracle/_ref, when present; else the C restatement) on a sample of the m-blocks:
    the error of the device results on those blocks, and the CPU timing."""
    import torch
    from oracle import oracle as O
    plan = res["plan"]
    pos = {int(m): i for i, m in enumerate(mvals)}
    sample = np.unique(np.linspace(0, LMAX, sample_m).round().astype(np.int32))
    mps = [pos[int(m)] for m in sample]
    off = O.spectral_offsets(LMAX, mvals, NR, NF)
    hyb = NTHETA * 2 * NR * NF
    s_all = res["s"]
    s_in = np.concatenate([s_all[off[mp]: off[mp + 1]].cpu().numpy() for mp in mps])
    p_dev = np.concatenate([res["p2"][mp * hyb: (mp + 1) * hyb].cpu().numpy() for mp in mps])
    s_dev = np.concatenate([res["s2"][off[mp]: off[mp + 1]].cpu().numpy() for mp in mps])
    flops = sum(2.0 * 2 * (NTHETA // 2) * 2 * NR * NF * (LMAX - int(m) + 1) for m in sample)
    threads = min(16, len(os.sched_getaffinity(0)))
    one = None
    if O.ref_available():
        kind = "reference"
        from threadpoolctl import threadpool_limits
        R = O.Reference()
        log(f"cpu baseline: compiled reference, {len(sample)} m-blocks, {threads} threads (setup)")
        R.legendre_init(NTHETA, LMAX, sample)
        times = []
        with threadpool_limits(limits=threads):
            for _ in range(3):
                t0 = time.perf_counter()
                p_ref = R.stp(NR, NF, s_in)
                s_ref = R.pts(NR, NF, p_dev)
                times.append(time.perf_counter() - t0)
        # the same path on one core: every fourth block of the sample, median of two runs
        sub = sample[::4]
        R.legendre_init(NTHETA, LMAX, sub)
        s_sub = np.concatenate([s_all[off[pos[int(m)]]: off[pos[int(m)] + 1]].cpu().numpy() for m in sub])
        p_sub = np.concatenate([res["p2"][pos[int(m)] * hyb: (pos[int(m)] + 1) * hyb].cpu().numpy() for m in sub])
        f_sub = sum(2.0 * 2 * (NTHETA // 2) * 2 * NR * NF * (LMAX - int(m) + 1) for m in sub)
        t1 = []
        with threadpool_limits(limits=1):
            for _ in range(2):
                t0 = time.perf_counter()
                R.stp(NR, NF, s_sub)
                R.pts(NR, NF, p_sub)
                t1.append(time.perf_counter() - t0)
        one = {"value": f_sub / float(np.median(t1)) / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": kind,
               "sample": f"StP+PtS on {len(sub)} of 512 m-blocks (m = {int(sub[0])}..{int(sub[-1])}); "
                         f"{f_sub / 1e9:.1f} GFLOP in {float(np.median(t1)):.3f} s (median of 2)"}
        log(f"cpu baseline, 1 core: {one['value']:.4f} TFLOP/s")
    else:
        kind = "port"
        threads = 1
        sample = sample[:: max(1, len(sample) // 6)]
        mps = [pos[int(m)] for m in sample]
        s_in = np.concatenate([s_all[off[mp]: off[mp + 1]].cpu().numpy() for mp in mps])
        p_dev = np.concatenate([res["p2"][mp * hyb: (mp + 1) * hyb].cpu().numpy() for mp in mps])
        s_dev = np.concatenate([res["s2"][off[mp]: off[mp + 1]].cpu().numpy() for mp in mps])
        flops = sum(2.0 * 2 * (NTHETA // 2) * 2 * NR * NF * (LMAX - int(m) + 1) for m in sample)
        log(f"cpu baseline: C restatement, {len(sample)} m-blocks, 1 thread")
        tabs = O.legendre_tables(NTHETA, LMAX, sample)
        t0 = time.perf_counter()
        p_ref = O.stp(NTHETA, LMAX, sample, NR, NF, tabs, s_in)
        s_ref = O.pts(NTHETA, LMAX, sample, NR, NF, tabs, p_dev)
        times = [time.perf_counter() - t0]
    t = float(np.median(times))
    err_stp = float(np.abs(p_dev - p_ref).max() / np.abs(p_ref).max())
    err_pts = float(np.abs(s_dev - s_ref).max() / np.abs(s_ref).max())
    del torch
    cpu = {"value": flops / t / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": kind,
           "sample": f"StP+PtS on {len(sample)} of 512 m-blocks (m = {int(sample[0])}..{int(sample[-1])}) at "
                     f"N_theta={NTHETA}, N_r={NR}, 8 fields; {flops / 1e9:.1f} GFLOP in {t:.3f} s; "
                     f"full round trip at this rate = {2 * flops_per_direction() / (flops / t) * 1e3:.0f} ms"}
    log(f"cpu baseline: {cpu['value']:.4f} TFLOP/s; device vs {kind}: StP {err_stp:.2e}, PtS {err_pts:.2e}")
    cpu["one_core"] = one
    return cpu, err_stp, err_pts


def cpu_reference_extras(L, mv, nr, nf, nlm, n_lu, ncols, out):
    """The reference's CPU path beside the device figures of the stack's stages, each on a
    bounded sample scaled to the full size: the compiled reference routines (oracle/_ref,
    one core; the harness's copies into and out of the reference's containers are
    included) for StP / PtS, the Chebyshev pair, the radial and angular derivatives and
    rlm_spaceb; numpy's pocketfft for the phi FFTs (one core: FFTW, the reference's
    library, is not available); LAPACK dgetrf / dgetrs / dgbtrf / dgbtrs for the implicit
    solve (the library the reference's solve calls, numpy's OpenBLAS here, 16 threads)."""
    import scipy.linalg.lapack as lapack
    from threadpoolctl import threadpool_limits
    from oracle import oracle as O
    res = {}
    rng = np.random.default_rng(0)
    if O.ref_available():
        R = O.Reference()
        # the Legendre pair on a sample of m-blocks and the Chebyshev pair on a sample of
        # columns, one core (the reference's own DGEMM calls), scaled to the full size
        from threadpoolctl import threadpool_limits
        nt = 3 * (L + 1) // 2
        sample = np.unique(np.linspace(0, L, 8).round().astype(np.int32))
        R.legendre_init(nt, L, sample)
        s_in = rng.standard_normal(O.spectral_size(L, sample, nr, nf))
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            p_out = R.stp(nr, nf, s_in)
            t_stp = time.perf_counter() - t0
            t0 = time.perf_counter()
            R.pts(nr, nf, p_out)
            t_pts = time.perf_counter() - t0
        frac_m = sum(L - int(m) + 1 for m in sample) / ((L + 1) * (L + 2) / 2)
        for key, t in (("stp", t_stp), ("pts", t_pts)):
            res[key] = {"ms_full_size": t / frac_m * 1e3, "cores": 1, "kind": "reference",
                        "sample": f"{len(sample)} m-blocks (m = {int(sample[0])}..{int(sample[-1])}), "
                                  f"{frac_m * 100:.1f} % of the coefficients"}
        R.cheby_init(nr, 0.35, 1.0)
        ccols = max(1, 2 * nlm * nf // 64)
        f = rng.standard_normal(nr * ccols)
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            c = R.cheby_to_spectral(f)
            t_to = time.perf_counter() - t0
            t0 = time.perf_counter()
            R.cheby_from_spectral(c)
            t_from = time.perf_counter() - t0
        for key, t in (("cheb_to_spectral", t_to), ("cheb_from_spectral", t_from)):
            res[key] = {"ms_full_size": t * (2 * nlm * nf / ccols) * 1e3, "cores": 1, "kind": "reference",
                        "sample": f"{ccols} of {2 * nlm * nf} columns of {nr}"}
        cols = 2 * nlm * nf // 32
        R.cheby_init(nr, 0.35, 1.0)
        c = rng.standard_normal(nr * cols)
        t0 = time.perf_counter()
        R.cheby_deriv(c, 1)
        t = time.perf_counter() - t0
        res["d_by_dr_cp"] = {"ms_full_size": t * 32 * 1e3, "cores": 1, "kind": "reference",
                             "sample": f"order 1 on {cols} of {2 * nlm * nf} columns of {nr}"}
        R.angular_init(L, mv, nr)
        buf = rng.standard_normal(O.spectral_size(L, mv, nr, 2))
        t0 = time.perf_counter()
        R.angular_deriv("dtheta", buf, 2, 1, 0)
        t = time.perf_counter() - t0
        res["d_by_dtheta"] = {"ms_full_size": t * 1e3, "cores": 1, "kind": "reference",
                              "sample": "one field, the full size (a two-field buffer passed in)"}
        buf = rng.standard_normal(O.spectral_size(L, mv, nr, 3))
        t0 = time.perf_counter()
        R.rlm_spaceb(buf, 3, 1, 2, 3)
        t = time.perf_counter() - t0
        res["rlm_spaceb"] = {"ms_full_size": t * 1e3, "cores": 1, "kind": "reference",
                             "sample": "W, P, Z at the full size (a three-field buffer; the reference's "
                                       "d_by_sdtheta / d_by_dphi and rlm_spaceb's array statements)"}
        # rlm_spacea's hydro steps and Compute_BandCurlB on a 16-field buffer, a radial sample
        nrs = min(nr, 8)
        R.angular_init(L, mv, nrs)
        buf = rng.standard_normal(O.spectral_size(L, mv, nrs, 16))
        prof = {k: rng.uniform(0.5, 2.0, nrs) for k in O.HYDRO_PROFILES}
        for key, call in (("rlm_spacea_hydro", lambda: R.rlm_spacea_hydro(buf, 16, list(range(10)), prof)),
                          ("compute_bandcurlb", lambda: R.compute_bandcurlb(buf, 16, list(range(10, 16)),
                                                                            prof["over_rhorsq"])),
                          ("adjust_emf", lambda: R.adjust_emf(buf, 16, 10, 11))):
            t0 = time.perf_counter()
            call()
            t = time.perf_counter() - t0
            res[key] = {"ms_full_size": t * (nr / nrs) * 1e3, "cores": 1, "kind": "reference",
                        "sample": f"{nrs} of {nr} radii of a 16-field buffer, scaled (the reference's "
                                  "d_by_dtheta / d_by_dphi and the routines' array statements)"}
        co = rng.standard_normal(O.spectral_size(L, mv, nrs, 9))
        rr = [rng.uniform(0.5, 2.0, nrs) for _ in range(3)]
        t0 = time.perf_counter()
        R.hybrid_output(buf, 16, co, 9, [11, 12, 13, 14, 2, 6], [3, 10, 11, 12, 0, 3, 5, 6, 1], *rr)
        t = time.perf_counter() - t0
        res["hybrid_output"] = {"ms_full_size": t * (nr / nrs) * 1e3, "cores": 1, "kind": "reference",
                                "sample": f"{nrs} of {nr} radii, magnetic: Hybrid_Output_Initial's magnetic "
                                          "part and Hybrid_Output_Final up to its transform, scaled"}
        R.angular_init(L, mv, nr)
    # the phi FFTs: numpy's pocketfft (FFTW, the reference's library, is not available) on
    # a sample of lines, one core, scaled to every (r, theta, field) line
    nphi = 3 * (L + 1)
    nlines = nr * (3 * (L + 1) // 2) * nf
    lines = rng.standard_normal((4096, nphi))
    t0 = time.perf_counter()
    spec = np.fft.rfft(lines, axis=1)
    t_r2c = time.perf_counter() - t0
    t0 = time.perf_counter()
    np.fft.irfft(spec, n=nphi, axis=1)
    t_c2r = time.perf_counter() - t0
    for key, t in (("fft_r2c", t_r2c), ("fft_c2r", t_c2r)):
        res[key] = {"ms_full_size": t * nlines / 4096 * 1e3, "cores": 1, "kind": "pocketfft (numpy.fft)",
                    "sample": f"4096 of {nlines} lines of {nphi}"}
    threads = min(16, len(os.sched_getaffinity(0)))
    k = 16
    mats = [rng.standard_normal((n_lu, n_lu)) + n_lu * np.eye(n_lu) for _ in range(k)]
    sel = np.linspace(0, len(ncols) - 1, k).round().astype(int)
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        facs = [lapack.dgetrf(a) for a in mats]
        t_f = time.perf_counter() - t0
        rhs = [rng.standard_normal((n_lu, ncols[i])) for i in sel]
        t0 = time.perf_counter()
        for (lu, piv, _), b in zip(facs, rhs):
            lapack.dgetrs(lu, piv, b)
        t_s = time.perf_counter() - t0
    frac = sum(ncols[i] for i in sel) / sum(ncols)
    kl = min(3 * (nr // 4), n_lu - 1)
    bands = [rng.standard_normal((3 * kl + 1, n_lu)) for _ in range(k)]
    for ab in bands:
        ab[2 * kl] += 2 * kl
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        bfacs = [lapack.dgbtrf(ab, kl, kl) for ab in bands]
        tb_f = time.perf_counter() - t0
        t0 = time.perf_counter()
        for (lu, piv, _), b in zip(bfacs, rhs):
            lapack.dgbtrs(lu, kl, kl, b, piv)
        tb_s = time.perf_counter() - t0
    res["lu_decompose_band"] = {"ms_full_size": tb_f * len(ncols) / k * 1e3, "cores": threads, "kind": "LAPACK",
                                "sample": f"dgbtrf on {k} of {len(ncols)} band matrices, n = {n_lu}, kl = ku = {kl}"}
    res["lu_solve_band"] = {"ms_full_size": tb_s / frac * 1e3, "cores": threads, "kind": "LAPACK",
                            "sample": f"dgbtrs on {k} of {len(ncols)} band matrices and their right-hand sides"}
    res["lu_decompose_full"] = {"ms_full_size": t_f * len(ncols) / k * 1e3, "cores": threads, "kind": "LAPACK",
                                "sample": f"dgetrf on {k} of {len(ncols)} matrices of n = {n_lu}"}
    res["lu_solve_full"] = {"ms_full_size": t_s / frac * 1e3, "cores": threads, "kind": "LAPACK",
                            "sample": f"dgetrs on {k} of {len(ncols)} matrices and their right-hand sides"}
    for key, v in res.items():
        log(f"cpu reference {key}: {v['ms_full_size']:.1f} ms at full size ({v['sample']})")
    return res


STACKS = {
    # l_max, N_theta, N_phi, N_r, N_f
    "stack": (255, 384, 768, 128, 8),     # full FP64 stack at l_max = 255
    "large": (1023, 1536, 3072, 256, 5),  # large one-GPU load at l_max = 1023
}


def radial_coeffs_device(lmax, l_of_lm, m_of_lm, nr, nf, seed: int):
    """Chebyshev coefficients in the radial layout p1[f][lm][imi][n] with the configuration's
    distribution: (N(0,1) + i N(0,1)) (l+1)^-2 (n+1)^-3, m = 0 real, on the device."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    nlm = len(l_of_lm)
    c = torch.randn((nf, nlm, 2, nr), dtype=torch.float64, device="cuda", generator=g)
    c *= ((torch.as_tensor(np.asarray(l_of_lm), dtype=torch.float64, device="cuda") + 1.0) ** -2).view(1, nlm, 1, 1)
    c *= ((torch.arange(nr, dtype=torch.float64, device="cuda") + 1.0) ** -3).view(1, 1, 1, nr)
    c[:, torch.as_tensor(np.asarray(m_of_lm) == 0, device="cuda"), 1, :] = 0.0
    return c.reshape(-1)


def stack_bench(ra, steps: int, warmup: int, workload: str = "stack"):
    """The transform stack (default l_max = 255, N_theta = 384, N_phi = 768, N_r = 128,
    8 fields; "large": l_max = 1023, N_theta = 1536, N_phi = 3072, N_r = 256, 5 fields),
    chained on the same device buffers as one time step runs it: Chebyshev From_Spectral ->
    Transpose_1a2a -> StP -> 2a->3a -> c2r, then (r2c + 3b->2b in one kernel) -> PtS -> Transpose_2b1b ->
    Chebyshev To_Spectral, each stage timed with HIP events (GB/s of the bandwidth-bound
    stages, TFLOP/s of the contractions), on Chebyshev coefficients with the configuration's
    distribution; the chain must return its input (N_phi = 2 N_theta, so the unnormalised
    FFT pair and the weighted PtS cancel).  Then the radial derivative recursion, the
    angular derivatives and the implicit solve on the same data."""
    import torch
    L, nt, nphi, nr, nf = STACKS[workload]
    mv = ra.m_values_balanced(L)
    plan = ra.LegendrePlan(nt, L, mv, nr, nf)
    cheb = ra.ChebyGrid(nr, 0.35, 1.0)
    cheb_form = "dct" if nr in (128, 256) else "matrix"  # the library's default for n_r
    l_lm, mp_lm = plan.lm_values()
    nlm = len(l_lm)
    c0 = radial_coeffs_device(L, l_lm, mv[mp_lm], nr, nf, seed=3)
    rad = torch.empty_like(c0)
    rad2 = torch.empty_like(c0)
    c2 = torch.empty_like(c0)
    s = torch.empty(plan.spectral_size, dtype=torch.float64, device="cuda")
    p2 = torch.empty(plan.hybrid_size, dtype=torch.float64, device="cuda")
    n3 = (nphi + 2) * nr * nt * nf
    p3 = torch.empty(n3, dtype=torch.float64, device="cuda")
    s2 = torch.empty_like(s)
    nb_rad = 16.0 * c0.numel()
    stages = [
        ("cheb_from_spectral", lambda: cheb.from_spectral(c0, rad), nb_rad),
        ("transpose_1a2a", lambda: plan.transpose_1a2a(rad, s), 16.0 * plan.spectral_size),
        ("stp", lambda: plan.stp(s, p2), None),
        # the 2a->3a reform then c2r (two kernels): bytes of one pass, hybrid read, padded grid written
        ("reform_c2r", lambda: plan.reform_to_physical(p2, p3, nphi), 8.0 * (plan.hybrid_size + n3)),
        # r2c and the 3b->2b reform in one pass: padded grid read, hybrid written
        ("r2c_reform", lambda: plan.reform_to_spectral(p3, p2, nphi), 8.0 * (plan.hybrid_size + n3)),
        ("pts", lambda: plan.pts(p2, s2), None),
        ("transpose_2b1b", lambda: plan.transpose_2b1b(s2, rad2), 16.0 * plan.spectral_size),
        ("cheb_to_spectral", lambda: cheb.to_spectral(rad2, c2), nb_rad),
    ]
    for _ in range(warmup):
        for _, fn, _ in stages:
            fn()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(stages) + 1)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record()
        for i, (_, fn, _) in enumerate(stages):
            fn()
            ev[k][i + 1].record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    rt = float(((c2 - c0).abs().max() / c0.abs().max()).item())
    log(f"stack {workload}: chained round trip {wall * 1e3:.3f} ms, returns its input to {rt:.2e}")
    if not rt < 1e-11:
        raise RuntimeError(f"stack {workload}: the chained round trip does not return its input ({rt:.3e})")
    F = flops_per_direction(L, nt, nr, nf)
    out = {"workload": workload, "l_max": L, "n_theta": nt, "n_phi": nphi, "n_r": nr, "n_fields": nf,
           "ms_per_round_trip": wall * 1e3, "round_trip_rel_err": rt, "stages": {}}
    for i, (name, _, nbytes) in enumerate(stages):
        ms = float(np.mean([e[i].elapsed_time(e[i + 1]) for e in ev]))
        rec = {"ms": ms}
        if nbytes is None:
            rec["tflops"] = F / (ms * 1e-3) / 1e12
            rec["frac_fp64_peak"] = rec["tflops"] / PEAK_FP64_TFLOPS
        else:
            rec["gbps"] = nbytes / (ms * 1e-3) / 1e9
            rec["frac_hbm_peak"] = rec["gbps"] / 8000.0
        if name.startswith("cheb_"):
            rec["form"] = cheb_form
            if cheb_form == "matrix":
                # the table form is FP64-MFMA work: N_r^2 flop per radial column (two parity
                # halves of (N_r / 2)^2 multiply-adds)
                flop = float(nr) * nr * (c0.numel() // nr)
                rec["tflops"] = flop / (ms * 1e-3) / 1e12
                rec["frac_fp64_peak"] = rec["tflops"] / PEAK_FP64_TFLOPS
        out["stages"][name] = rec
    # both Chebyshev forms on the chain's buffers, and what a plain streaming kernel (torch's
    # elementwise multiply, one read and one write of the same bytes) reaches on them
    forms = {}
    for form in (("dct", "matrix") if cheb_form == "dct" else ("matrix",)):
        cheb.set_form(form)
        rec = {}
        for name, fn in (("to_spectral", lambda: cheb.to_spectral(rad2, c2)),
                         ("from_spectral", lambda: cheb.from_spectral(c0, rad))):
            ms = time_ms(fn, steps, warmup)
            rec[name] = {"ms": ms, "gbps": nb_rad / (ms * 1e-3) / 1e9, "frac_hbm_peak": nb_rad / (ms * 1e-3) / 8e12}
        forms[form] = rec
    cheb.set_form(cheb_form)
    ms = time_ms(lambda: torch.mul(c0, 1.0000001, out=rad), steps, warmup)
    forms["stream_ceiling"] = {"ms": ms, "gbps": nb_rad / (ms * 1e-3) / 1e9, "frac_hbm_peak": nb_rad / (ms * 1e-3) / 8e12}
    out["cheb_forms"] = forms
    rad = c0  # the later figures run on the chain's own data: s (spectral), c0 (Chebyshev coefficients)
    # the one-rank radial transposes around the implicit solve (Transpose_2b1b after PtS,
    # Transpose_1a2a before StP): 16 B per spectral coefficient (read + write)
    p1 = torch.empty_like(s)
    nb = 16.0 * plan.spectral_size
    out["radial_transposes"] = {}
    for name, fn in (("transpose_2b1b", lambda: plan.transpose_2b1b(s2, p1)),
                     ("transpose_1a2a", lambda: plan.transpose_1a2a(p1, s2))):
        for _ in range(max(warmup, 1)):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out["radial_transposes"][name] = {"ms": ms, "gbps": nb / (ms * 1e-3) / 1e9,
                                          "frac_hbm_peak": nb / (ms * 1e-3) / 8e12}
    del p1
    # the angular derivatives (d_by_dtheta, d_by_sdtheta, d_by_dphi) of one field of the
    # spectral data into a one-field array: 16 B per coefficient of the field
    per_field = plan.spectral_size // nf
    f1 = torch.empty(per_field, dtype=torch.float64, device="cuda")
    nb = 16.0 * per_field
    out["angular_derivatives"] = {}
    for kind in ("dtheta", "sdtheta", "dphi"):
        fn = lambda: plan.angular_derivative(kind, s, 0, f1, 0)
        for _ in range(max(warmup, 1)):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out["angular_derivatives"][kind] = {"ms": ms, "gbps": nb / (ms * 1e-3) / 1e9,
                                            "frac_hbm_peak": nb / (ms * 1e-3) / 8e12}
    del f1
    # rlm_spaceb's combination after PtS, in place on W, P, Z (fields 1, 2, 3) of a copy of the
    # 8-field spectral data: three fields read and written (48 B per coefficient of a field)
    # and l = l_max of the other five fields zeroed
    w8 = s.clone()
    nb = 48.0 * per_field + 8.0 * (nf - 3) * 2 * nr * len(mv)
    fn = lambda: plan.rlm_spaceb_hydro(w8, 1, 2, 3)
    for _ in range(max(warmup, 1)):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    out["rlm_spaceb_hydro"] = {"ms": ms, "gbps": nb / (ms * 1e-3) / 1e9, "frac_hbm_peak": nb / (ms * 1e-3) / 8e12}
    del w8
    # rlm_spacea's passes before StP on a 16-field buffer of the same spectral layout: the
    # hydro pass (Velocity_Components + Velocity_Derivatives + dtdt: 6 fields read, 9 written,
    # 120 B per coefficient of a field) on fields 0-9, Compute_BandCurlB (5 read, 6 written,
    # 88 B) on fields 10-15, and rlm_spaceb's Adjust_Emf (2 fields read and written, 32 B) on
    # fields 10-11; repeated in place on the same buffer (the time of these passes does not
    # depend on the values)
    b16 = torch.randn(16 * per_field, dtype=torch.float64, device="cuda")
    prof = {k: torch.rand(nr, dtype=torch.float64, device="cuda") + 0.5 for k in plan.SPACEA_PROFILES}
    out["rlm_spacea"] = {}
    for name, fn, nb in (("hydro", lambda: plan.rlm_spacea_hydro(b16, list(range(10)), prof), 120.0 * per_field),
                         ("bandcurlb", lambda: plan.compute_bandcurlb(b16, list(range(10, 16)), prof["over_rhorsq"]),
                          88.0 * per_field),
                         ("adjust_emf", lambda: plan.adjust_emf(b16, 10, 11), 32.0 * per_field)):
        for _ in range(max(warmup, 1)):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out["rlm_spacea"][name] = {"ms": ms, "gbps": nb / (ms * 1e-3) / 1e9, "frac_hbm_peak": nb / (ms * 1e-3) / 8e12}
    # the output iteration's passes (Hybrid_Output_Initial / _Final up to their transforms) from
    # the 16-field buffer into a 9-field co-buffer and 2 work fields: the initial pass reads 4
    # fields and writes 4 (64 B per coefficient of a field); the final pass, magnetic, reads 4
    # of s, the 2 work fields and dbrdr_cb and writes 5 (96 B), hydro only 1 and 1 (16 B)
    co = torch.randn(9 * per_field, dtype=torch.float64, device="cuda")
    tmp = torch.randn(2 * per_field, dtype=torch.float64, device="cuda")
    o_r, o_rsq, t_r = prof["over_rhor"], prof["over_rhorsq"], prof["two_over_r"]
    vi, vf = [11, 12, 13, 14, 2, 6], [3, 10, 11, 12, 0, 3, 5, 6, 1]
    out["output_iteration"] = {}
    for name, fn, nb in (("initial", lambda: plan.hybrid_output_initial(b16, co, tmp, vi, o_rsq), 64.0 * per_field),
                         ("final", lambda: plan.hybrid_output_final(b16, co, vf, tmp, o_r, o_rsq, t_r),
                          96.0 * per_field),
                         ("final_hydro", lambda: plan.hybrid_output_final(b16, co, vf), 16.0 * per_field)):
        for _ in range(max(warmup, 1)):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out["output_iteration"][name] = {"ms": ms, "gbps": nb / (ms * 1e-3) / 1e9,
                                         "frac_hbm_peak": nb / (ms * 1e-3) / 8e12}
    del b16, co, tmp
    # one field scaled by l(l+1) and a radial profile (rlm_spacea's l_l_plus1 * f * Over_RhoRSQ):
    # 16 B per coefficient
    lfac = torch.from_numpy(ra.l_l_plus1(L)).cuda()
    rfac = torch.linspace(1.0, 8.0, nr, dtype=torch.float64, device="cuda")
    f1 = torch.empty(per_field, dtype=torch.float64, device="cuda")
    fn = lambda: plan.lm_scale(s, 0, f1, 0, lfac, rfac)
    for _ in range(max(warmup, 1)):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nb = 16.0 * per_field
    out["lm_scale"] = {"ms": ms, "gbps": nb / (ms * 1e-3) / 1e9, "frac_hbm_peak": nb / (ms * 1e-3) / 8e12}
    del f1
    # the implicit solve's dense path for the coupled W-P system (n = 2 N_r) of every l:
    # LU_Decompose_full on l_max + 1 matrices, then LU_Solve_full on 2 (l + 1) right-hand
    # sides per l (re, im of every m <= l); 2 n^3 / 3 and 2 n^2 per column flop
    n_lu, nb = 2 * nr, L + 1
    gen = torch.Generator(device="cuda").manual_seed(4)
    mats0 = torch.randn(nb * n_lu * n_lu, dtype=torch.float64, device="cuda", generator=gen)
    mats0.view(nb, n_lu, n_lu).diagonal(dim1=1, dim2=2).add_(float(n_lu))
    mats = torch.empty_like(mats0)
    piv = torch.empty(nb * n_lu, dtype=torch.int32, device="cuda")
    info = torch.empty(nb, dtype=torch.int32, device="cuda")
    ncols = [2 * (l + 1) for l in range(nb)]
    offs = np.concatenate([[0], np.cumsum([c * n_lu for c in ncols])]).astype(np.int64)
    rhs0 = torch.randn(int(offs[-1]), dtype=torch.float64, device="cuda", generator=gen)
    rhs = torch.empty_like(rhs0)
    d_offs = torch.from_numpy(offs[:-1]).cuda()
    d_ncols = torch.tensor(ncols, dtype=torch.int32, device="cuda")
    lu_rec = {}
    for name, prep, fn, flop in (
            ("lu_decompose_full", lambda: mats.copy_(mats0), lambda: ra.lu_decompose_full(mats, piv, info, n_lu),
             nb * 2.0 * n_lu ** 3 / 3),
            ("lu_solve_full", lambda: rhs.copy_(rhs0),
             lambda: ra.lu_solve_full(mats, piv, rhs, d_offs, d_ncols, n_lu, max(ncols)),
             2.0 * n_lu * n_lu * sum(ncols))):
        ts = []
        for it in range(max(warmup, 1) + steps):
            prep()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            if it >= max(warmup, 1):
                ts.append(e0.elapsed_time(e1))
        ms = float(np.mean(ts))
        lu_rec[name] = {"ms": ms, "gflops": flop / (ms * 1e-3) / 1e9,
                        "frac_fp64_peak": flop / (ms * 1e-3) / (PEAK_FP64_TFLOPS * 1e12)}
    out["implicit_solve"] = {"n": n_lu, "matrices": nb, **lu_rec}
    del mats0, mats, rhs0, rhs
    # the band path (band_solve) for the same W-P system on a grid of 4 Chebyshev domains of
    # N_r / 4 polynomials: kl = ku = 3 N_r / 4 (Band_Arrange's 3 max_npoly for the linked
    # equation), band storage ldab = 3 kl + 1; ~2 n kl (kl + ku) flop per factorisation and
    # 2 n (2 kl + ku) per right-hand side
    kl = min(3 * (nr // 4), n_lu - 1)
    ldab = 3 * kl + 1
    band0 = torch.randn(nb * ldab * n_lu, dtype=torch.float64, device="cuda", generator=gen)
    band0.view(nb, n_lu, ldab)[:, :, 2 * kl].add_(float(2 * kl))  # a strong diagonal
    band = torch.empty_like(band0)
    rhs0 = torch.randn(int(offs[-1]), dtype=torch.float64, device="cuda", generator=gen)
    rhs = torch.empty_like(rhs0)
    band_rec = {}
    for name, prep, fn, flop in (
            ("lu_decompose_band", lambda: band.copy_(band0), lambda: ra.lu_decompose_band(band, piv, info, n_lu, kl),
             nb * 2.0 * n_lu * kl * 2 * kl),
            ("lu_solve_band", lambda: rhs.copy_(rhs0),
             lambda: ra.lu_solve_band(band, piv, rhs, d_offs, d_ncols, n_lu, kl, max(ncols)),
             2.0 * n_lu * 3 * kl * sum(ncols))):
        ts = []
        for it in range(max(warmup, 1) + steps):
            prep()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            if it >= max(warmup, 1):
                ts.append(e0.elapsed_time(e1))
        ms = float(np.mean(ts))
        band_rec[name] = {"ms": ms, "gflops": flop / (ms * 1e-3) / 1e9,
                          "frac_fp64_peak": flop / (ms * 1e-3) / (PEAK_FP64_TFLOPS * 1e12)}
    out["implicit_solve_band"] = {"n": n_lu, "kl": kl, "ku": kl, "matrices": nb, **band_rec}
    del band0, band, rhs0, rhs
    out["cpu_reference"] = cpu_reference_extras(L, mv, nr, nf, nlm, n_lu, ncols, out)
    # the radial derivative recursion (d_by_dr_cp) of orders 1-3 on the spectral radial
    # buffer, as Post_Solve applies it: 16 B per coefficient (read + write)
    rad3 = torch.empty_like(rad)
    nbytes = 16.0 * rad.numel()
    out["radial_derivatives"] = {}
    for d in (1, 2, 3):
        for _ in range(max(warmup, 1)):
            cheb.d_by_dr(rad, rad3, d)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            cheb.d_by_dr(rad, rad3, d)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out["radial_derivatives"][f"order{d}"] = {"ms": ms, "gbps": nbytes / (ms * 1e-3) / 1e9,
                                                  "frac_hbm_peak": nbytes / (ms * 1e-3) / 8e12}
    del rad3
    # the Legendre + FFT round trip replayed from one hipGraph (no Chebyshev stages)
    g_rt = plan.round_trip_graph(nphi, s, p2, p3, s2)
    for _ in range(warmup):
        g_rt.launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        g_rt.launch()
    e1.record()
    torch.cuda.synchronize()
    out["graph_round_trip_ms"] = e0.elapsed_time(e1) / steps
    out["eager_round_trip_ms"] = sum(out["stages"][k]["ms"] for k in ("stp", "reform_c2r", "r2c_reform", "pts"))
    # the separate kernels the fused stages replace, on the same buffers (FFTs in place)
    sep = {}
    for name, fn, nbytes in (("permute_2a3a", lambda: plan.reform_2a3a(p2, p3, nphi), 8.0 * (plan.hybrid_size + n3)),
                             ("fft_c2r", lambda: ra.fft_to_physical(p3, nphi), 16.0 * n3),
                             ("fft_r2c", lambda: ra.fft_to_spectral(p3, nphi), 16.0 * n3),
                             ("permute_3b2b", lambda: plan.reform_3b2b(p3, p2, nphi), 8.0 * 2 * plan.hybrid_size)):
        for _ in range(max(warmup, 1)):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        sep[name] = {"ms": ms, "gbps": nbytes / (ms * 1e-3) / 1e9, "frac_hbm_peak": nbytes / (ms * 1e-3) / 8e12}
    out["separate_transpose_fft"] = sep
    print(json.dumps(out), flush=True)


def batch_problems(seed: int = 3, n: int = 64):
    """64 independent transforms: l_max log-uniform in [31, 1023] with l_max + 1 a
    multiple of 4 (so that N_theta = 3 (l_max + 1) / 2 is even), N_r from {32, 64, 96},
    1-8 fields; every m of each problem."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(32), np.log(1024), n)
    l1 = np.clip(4 * np.round(np.exp(u) / 4), 32, 1024).astype(int)
    nr = rng.choice([32, 64, 96], n)
    nf = rng.integers(1, 9, n)
    return [(int(a - 1), int(3 * a // 2), int(b), int(c)) for a, b, c in zip(l1, nr, nf)]


def spectral_device(plan, lmax, nr, nf, seed: int):
    """Coefficients N(0,1) (l+1)^-2 sin(pi r_i / N_r) (m = 0 real) for all m in ascending
    order, on the device."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    s = torch.empty(plan.spectral_size, dtype=torch.float64, device="cuda")
    rfac = torch.sin(torch.pi * torch.arange(1, nr + 1, dtype=torch.float64, device="cuda") / nr)
    lf = (torch.arange(0, lmax + 1, dtype=torch.float64, device="cuda") + 1.0) ** -2
    off = 0
    for m in range(lmax + 1):
        nl = lmax - m + 1
        blk = torch.randn((nf, 2, nr, nl), dtype=torch.float64, device="cuda", generator=g)
        blk *= lf[m:].view(1, 1, 1, nl)
        blk *= rfac.view(1, 1, nr, 1)
        if m == 0:
            blk[:, 1] = 0.0
        s[off: off + blk.numel()] = blk.reshape(-1)
        off += blk.numel()
    return s


def batch_bench(ra, steps: int, warmup: int, forms):
    """64 problems of different l_max in one grouped launch per direction, FP64 and the
    split forms side by side: eager grouped launches (HIP events per direction), the
    grouped round trip replayed from a hipGraph, and the same 64 problems launched one
    plan at a time.  The FP64 batch must equal the serial FP64 runs bit for bit; each
    split form's per-problem max / RMS error is taken against them (StP on the same
    spectral input, PtS on the FP64 hybrid output)."""
    import torch
    probs = batch_problems()
    F = sum(flops_per_direction(L, nt, nr, nf) for L, nt, nr, nf in probs)
    log(f"batch: {len(probs)} problems, l_max {min(p[0] for p in probs)}..{max(p[0] for p in probs)}, "
        f"F_LT {F:.3e} per direction")

    def ev():
        return torch.cuda.Event(enable_timing=True)

    def make_plans(form):
        t0 = time.time()
        plans = [ra.LegendrePlan(nt, L, np.arange(L + 1, dtype=np.int32), nr, nf, precision=form)
                 for L, nt, nr, nf in probs]
        torch.cuda.synchronize()
        return plans, time.time() - t0

    def run(form, plans, s_in, p_in_pts=None):
        batch = ra.LegendreBatch(plans)
        p2 = [torch.empty(p.hybrid_size, dtype=torch.float64, device="cuda") for p in plans]
        s2 = [torch.empty(p.spectral_size, dtype=torch.float64, device="cuda") for p in plans]
        tune = batch.autotune(s_in, p2, s2, reps=2)
        log(f"batch {form}: schedule candidates (ms) {tune}, chosen {batch.schedule()}")
        for _ in range(warmup):
            batch.stp(s_in, p2)
            batch.pts(p2, s2)
        torch.cuda.synchronize()
        e = [[ev(), ev(), ev()] for _ in range(steps)]
        for k in range(steps):
            e[k][0].record()
            batch.stp(s_in, p2)
            e[k][1].record()
            batch.pts(p2, s2)
            e[k][2].record()
        torch.cuda.synchronize()
        stp_ms = float(np.mean([x[0].elapsed_time(x[1]) for x in e]))
        pts_ms = float(np.mean([x[1].elapsed_time(x[2]) for x in e]))
        g = batch.round_trip_graph(s_in, p2, s2)
        for _ in range(warmup):
            g.launch()
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record()
        for _ in range(steps):
            g.launch()
        b.record()
        torch.cuda.synchronize()
        graph_ms = a.elapsed_time(b) / steps
        # one plan at a time
        a.record()
        for _ in range(steps):
            for i, p in enumerate(plans):
                p.stp(s_in[i], p2[i])
                p.pts(p2[i], s2[i])
        b.record()
        torch.cuda.synchronize()
        serial_ms = a.elapsed_time(b) / steps
        del g
        rec = {"autotune_ms": tune, "schedule": batch.schedule(),
               "stp_ms": stp_ms, "pts_ms": pts_ms, "graph_round_trip_ms": graph_ms,
               "serial_round_trip_ms": serial_ms, "tflops_graph": 2 * F / (graph_ms * 1e-3) / 1e12}
        if form == "fp64":
            rec["frac_fp64_peak"] = rec["tflops_graph"] / PEAK_FP64_TFLOPS
        log(f"batch {form}: StP {stp_ms:.2f} ms, PtS {pts_ms:.2f} ms, graph round trip {graph_ms:.2f} ms, "
            f"serial plans {serial_ms:.2f} ms")
        # the batch's own final outputs (from the serial pass, identical by construction
        # for FP64: checked below) and PtS of the FP64 hybrid input for the error table
        batch.stp(s_in, p2)
        if p_in_pts is not None:
            batch.pts(p_in_pts, s2)
        else:
            batch.pts(p2, s2)
        torch.cuda.synchronize()
        return rec, p2, s2

    log("batch fp64: creating plans")
    plans, setup_s = make_plans("fp64")
    s_in = [spectral_device(p, L, nr, nf, seed=500 + i) for i, (p, (L, nt, nr, nf)) in enumerate(zip(plans, probs))]
    rec64, p64, s64 = run("fp64", plans, s_in)
    rec64["setup_s"] = setup_s
    # grouped FP64 == one plan at a time, bit for bit
    exact = True
    for i, p in enumerate(plans):
        t = torch.empty(p.hybrid_size, dtype=torch.float64, device="cuda")
        p.stp(s_in[i], t)
        exact &= bool(torch.equal(t, p64[i]))
        u = torch.empty(p.spectral_size, dtype=torch.float64, device="cuda")
        p.pts(p64[i], u)
        exact &= bool(torch.equal(u, s64[i]))
        del t, u
    rec64["bit_exact_vs_serial"] = exact
    out = {"workload": "batch", "n_problems": len(probs), "problems": probs, "flops_per_direction": F,
           "forms": {"fp64": rec64}}
    del plans
    torch.cuda.empty_cache()
    for form in forms:
        if form == "fp64":
            continue
        log(f"batch {form}: creating plans")
        plans, setup_s = make_plans(form)
        rec, p2, s2 = run(form, plans, s_in, p_in_pts=p64)
        rec["setup_s"] = setup_s
        errs = []
        for i in range(len(probs)):
            def rel(x, y):
                d = x - y
                return float(d.abs().max() / y.abs().max()), float(torch.sqrt((d * d).mean() / (y * y).mean()))
            sm, sr = rel(p2[i], p64[i])
            pm, pr = rel(s2[i], s64[i])
            errs.append({"stp_max": sm, "stp_rms": sr, "pts_max": pm, "pts_rms": pr})
        rec["max_err"] = max(max(e["stp_max"], e["pts_max"]) for e in errs)
        rec["per_problem_err"] = errs
        log(f"batch {form}: worst per-problem max rel error vs FP64 {rec['max_err']:.2e}")
        out["forms"][form] = rec
        del plans, p2, s2
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (profiles/legendre_fp64_lmax511.traffic.json: FETCH_SIZE and WRITE_SIZE in KiB
    per dispatch, each scaled by the same counter on a 1 GiB device copy), or None."""
    path = os.path.join(ROOT, "profiles", "legendre_fp64_lmax511.traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)[kernel]
    except (OSError, KeyError, ValueError):
        return None
    return (rec["fetch_kib"] * rec["fetch_factor"] + rec["write_kib"] * rec["write_factor"]) * 1024.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--forms", default="fp64,bf16x6,bf16x3,int8x6,int8x3",
                    help="Legendre forms to run; the first is the headline, fp64 is always run as the error reference")
    ap.add_argument("--cpu-sample-m", type=int, default=48)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="legendre", choices=["legendre", "stack", "large", "batch"],
                    help="legendre: the l_max=511 round trip (headline); stack: l_max=255 full stack per "
                         "stage; large: l_max=1023, N_phi=3072, N_r=256, 5 fields full stack per stage; "
                         "batch: 64 problems of l_max 31..1023 in grouped launches")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 1 GiB device-to-device copy (scales profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import rayleigh_amd as ra
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    if args.workload in STACKS:
        stack_bench(ra, args.steps, args.warmup, args.workload)
        return
    if args.workload == "batch":
        batch_bench(ra, args.steps, args.warmup, [f.strip() for f in args.forms.split(",") if f.strip()])
        return
    mvals = ra.m_values_balanced(LMAX)
    F = flops_per_direction()
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    run_order = ["fp64"] + [f for f in forms if f != "fp64"]
    results = {}
    for form in run_order:
        results[form] = time_form(ra, form, mvals, args.steps, args.warmup)
        if form != "fp64":
            results[form]["err_vs_fp64"] = errors_vs_fp64(ra, results[form], results["fp64"], mvals)
            e = results[form]["err_vs_fp64"]
            log(f"form {form}: vs FP64 StP max {e['stp_max']:.2e} rms {e['stp_rms']:.2e}, "
                f"PtS max {e['pts_max']:.2e} rms {e['pts_rms']:.2e}")
            # keep only what the report needs
            for key in ("plan", "s", "p2", "s2"):
                results[form].pop(key, None)
            torch.cuda.empty_cache()
    if args.copy_calibration:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written by one copy kernel")
        del a, b
    head = forms[0]
    r = results[head]
    cpu, err_stp, err_pts = (None, None, None)
    if not args.no_cpu_baseline:
        cpu, err_stp, err_pts = reference_check_and_baseline(results["fp64"], mvals, args.cpu_sample_m)
    dom_ms = max(r["stp_ms"], r["pts_ms"])
    dom = "stp" if r["stp_ms"] >= r["pts_ms"] else "pts"
    if head == "fp64":
        peak, kernel, achieved = PEAK_FP64_TFLOPS, f"{dom}_fp64_kernel", F / (dom_ms * 1e-3) / 1e12
    else:
        # BF16 / INT8 MFMA work actually issued: one F-sized product per kept slice pair
        pairs = SPLIT_PAIRS[head][dom]
        peak = PEAK_INT8_TOPS if head.startswith("int8") else PEAK_BF16_TFLOPS
        kernel, achieved = f"{dom}_split_kernel", pairs * F / (dom_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": 2 * F / (r["ms_per_step"] * 1e-3) / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else ("i8" if head.startswith("int8") else "bf16"),
        "data": "synthetic",
        "config": {"workload": "Legendre round trip StP+PtS, l_max=511, N_theta=768, N_r=192, 8 fields, "
                               "all m (balanced order)", "l_max": LMAX, "n_theta": NTHETA, "n_r": NR,
                   "n_fields": NF, "form": head},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak,
                     "unit": "TOP/s" if head.startswith("int8") else "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(kernel),
                     "kernel": kernel, "kernel_ms": dom_ms},
        "cpu_baseline": cpu,
        "rel_err_vs_reference": {"stp": err_stp, "pts": err_pts},
        "round_trip_rel_err": r["round_trip_rel_err"],
        "forms": {f: {"ms_per_step": v["ms_per_step"], "stp_ms": v["stp_ms"], "pts_ms": v["pts_ms"],
                      "tflops": 2 * F / (v["ms_per_step"] * 1e-3) / 1e12,
                      "round_trip_rel_err": v["round_trip_rel_err"], "setup_s": v["setup_s"],
                      "err_vs_fp64": v.get("err_vs_fp64")}
                  for f, v in results.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
