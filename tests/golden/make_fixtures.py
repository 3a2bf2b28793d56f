"""Writes the golden fixtures in this folder from the COMPILED REFERENCE.

Run in a container that holds the reference tree, after `make -C oracle ref`
has compiled the reference's own Legendre_Polynomials, Legendre_Transforms and
Chebyshev_Polynomials modules (with the bind(C) harness oracle/ref_harness.F90):

    python tests/golden/make_fixtures.py

Every array in the .npz files is an input chosen here or an output of the
reference's routines on that input; nothing is computed by the library's code
except the seeded synthetic inputs (oracle.synthetic_spectral, the (l+1)^-2
distribution) and the balanced m order (M_Balance_v0 on one rank).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def legendre_case(name, nt, lmax, mvals, nr, nf, seed, sample_tables=None):
    R = O.Reference()
    R.legendre_init(nt, lmax, mvals)
    xh, xl, wh, wl = R.nodes()
    pe, po, ie, io = R.tables()
    rec = dict(nt=nt, lmax=lmax, mvals=np.asarray(mvals, np.int32), nr=nr, nf=nf, seed=seed,
               x_hi=xh, x_lo=xl, w_hi=wh, w_lo=wl)
    if sample_tables is None:
        rec.update(p_even=pe, p_odd=po, ip_even=ie, ip_odd=io)
    else:
        rng = np.random.default_rng(1000 + seed)
        for key, arr in (("p_even", pe), ("p_odd", po), ("ip_even", ie), ("ip_odd", io)):
            idx = np.sort(rng.choice(arr.size, size=min(sample_tables, arr.size), replace=False))
            rec[key + "_idx"] = idx.astype(np.int64)
            rec[key + "_val"] = arr[idx]
    if nr > 0:
        s = O.synthetic_spectral(lmax, mvals, nr, nf, seed)
        p = R.stp(nr, nf, s)
        s2 = R.pts(nr, nf, p)
        rec.update(s_in=s, p_out=p, s_rt=s2)
    np.savez_compressed(os.path.join(OUT, name), **rec)
    print("wrote", name)


def cheby_case(name, n, ncols, seed):
    R = O.Reference()
    g, w = R.cheby_init(n, 0.35, 1.0)
    ce, co = R.cheby_tables()
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(n * ncols)
    c = R.cheby_to_spectral(f)
    f2 = R.cheby_from_spectral(c)
    np.savez_compressed(os.path.join(OUT, name), n=n, ncols=ncols, seed=seed, rmin=0.35, rmax=1.0,
                        grid=g, ceven=ce, codd=co, f_in=f, c_out=c, f_rt=f2)
    print("wrote", name)


def cheby_deriv_case(name, n, ncols, seed, rmin, rmax):
    """d_by_dr_cp (Cheby_Deriv_Buffer_4D) of orders 1-3 on seeded coefficients, and on the
    coefficients of exp(-r) (the reference's Test_Cheby case: derivatives -f, f, -f)."""
    R = O.Reference()
    g, w = R.cheby_init(n, rmin, rmax)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(n * ncols)
    ce = R.cheby_to_spectral(np.exp(-g))
    rec = dict(n=n, ncols=ncols, seed=seed, rmin=rmin, rmax=rmax, grid=g, c_in=c, c_exp=ce)
    for d in (1, 2, 3):
        rec[f"d{d}"] = R.cheby_deriv(c, d)
        rec[f"exp_d{d}"] = R.cheby_deriv(ce, d)
    np.savez_compressed(os.path.join(OUT, name), **rec)
    print("wrote", name)


def angular_case(name, lmax, mvals, nr, seed):
    """d_by_dtheta, d_by_sdtheta and d_by_dphi (Spectral_Derivatives.F90) of a seeded
    two-field buffer: the buffer forms (field 1 -> field 2, 1-based) and the 4dbuff2arr
    forms (field 2 -> a one-field array)."""
    R = O.Reference()
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    R.angular_init(lmax, mv, nr)
    rng = np.random.default_rng(seed)
    buf = rng.standard_normal(O.spectral_size(lmax, mv, nr, 2))
    rec = dict(lmax=lmax, mvals=mv, nr=nr, buf=buf)
    for kind in ("dtheta", "sdtheta", "dphi"):
        rec[f"{kind}_buffer"] = R.angular_deriv(kind, buf, 2, 1, 2)
        rec[f"{kind}_arr"] = R.angular_deriv(kind, buf, 2, 2, 0)
    np.savez_compressed(os.path.join(OUT, name), **rec)
    print("wrote", name)


def rlm_case(name, lmax, mvals, nr, nf, fields, seed):
    """rlm_spaceb's hydro steps (Sphere_Hybrid_Space.F90:127-214) through the reference's
    d_by_sdtheta / d_by_dphi on a seeded nf-field buffer; fields = 0-based (W, P, Z)."""
    R = O.Reference()
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    R.angular_init(lmax, mv, nr)
    buf = np.random.default_rng(seed).standard_normal(O.spectral_size(lmax, mv, nr, nf))
    wv, pv, zv = fields
    out = R.rlm_spaceb(buf, nf, wv + 1, pv + 1, zv + 1)
    np.savez_compressed(os.path.join(OUT, name), lmax=lmax, mvals=mv, nr=nr, nf=nf, fields=np.array(fields, np.int32),
                        buf=buf, out=out)
    print("wrote", name)


def rlm_spacea_case(name, lmax, mvals, nr, nf, hydro, magnetic, emf, pseudo, seed):
    """rlm_spacea's Velocity_Components + Velocity_Derivatives + d_by_dtheta(tvar, dtdt),
    Compute_BandCurlB and rlm_spaceb's Adjust_Emf (Sphere_Hybrid_Space.F90:80,235-460,557-602)
    through the reference's own derivatives on a seeded nf-field buffer with seeded radial
    profiles; hydro = 0-based (W, dWdr, d2Wdr2, Z, dZdr, T, vtheta, vphi, dvtdr, dvpdr),
    magnetic = 0-based (C, A, dCdr, dAdr, d2Cdr2, curlbr), emf = 0-based (emftheta, emfphi)."""
    R = O.Reference()
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    R.angular_init(lmax, mv, nr)
    rng = np.random.default_rng(seed)
    buf = rng.standard_normal(O.spectral_size(lmax, mv, nr, nf))
    prof = {k: rng.uniform(0.5, 2.0, nr) for k in O.HYDRO_PROFILES}
    if not pseudo:
        prof["dsdr_over_cp"] = None
    one_over_rsq = rng.uniform(0.5, 2.0, nr)
    rec = dict(lmax=lmax, mvals=mv, nr=nr, nf=nf, buf=buf, hydro=np.array(hydro, np.int32),
               magnetic=np.array(magnetic, np.int32), one_over_rsq=one_over_rsq, pseudo=int(pseudo),
               hydro_out=R.rlm_spacea_hydro(buf, nf, hydro, prof),
               magnetic_out=R.compute_bandcurlb(buf, nf, magnetic, one_over_rsq),
               emf=np.array(emf, np.int32), emf_out=R.adjust_emf(buf, nf, *emf))
    # the output-iteration steps (Hybrid_Output_Initial / _Final) into a seeded co-buffer of 7
    # fields, with and without magnetism; vi, vf 0-based as oracle.hybrid_output takes them
    nf_co = 7
    co = rng.standard_normal(O.spectral_size(lmax, mv, nr, nf_co))
    ho_prof = [rng.uniform(0.5, 2.0, nr) for _ in range(3)]
    vi, vf = (magnetic[1], magnetic[2], magnetic[4], magnetic[3], 2, 6), (hydro[3], magnetic[0], magnetic[1],
                                                                          magnetic[2], 0, 3, 5, 6, 1)
    rec.update(ho_co=co, ho_vi=np.array(vi, np.int32), ho_vf=np.array(vf, np.int32), ho_one_over_r=ho_prof[0],
               ho_one_over_rsq=ho_prof[1], ho_two_over_r=ho_prof[2],
               ho_out=R.hybrid_output(buf, nf, co, nf_co, vi, vf, *ho_prof, magnetism=True),
               ho_out_hydro=R.hybrid_output(buf, nf, co, nf_co, vi, vf, *ho_prof, magnetism=False))
    for k in O.HYDRO_PROFILES:
        if prof[k] is not None:
            rec[k] = prof[k]
    np.savez_compressed(os.path.join(OUT, name), **rec)
    print("wrote", name)


def grid_case(name, thetas, chebs):
    """Initialize_Legendre's Real*16 nodes and weights as stored (16 bytes per value) for
    several N_theta, and Initialize_Cheby_Grid's radial grid and integration weights for
    several (n, rmin, rmax)."""
    R = O.Reference()
    rec = {}
    for nt in thetas:
        R.legendre_init(nt, (2 * nt) // 3 - 1, np.arange(1, dtype=np.int32))
        x16, w16 = R.nodes_binary128()
        rec[f"coloc_q_{nt}"] = x16
        rec[f"gl_weights_q_{nt}"] = w16
    for i, (n, rmin, rmax) in enumerate(chebs):
        g, w = R.cheby_init(n, rmin, rmax)
        rec[f"cheb_{i}"] = np.array([n, rmin, rmax])
        rec[f"cheb_grid_{i}"] = g
        rec[f"cheb_weights_{i}"] = w
    np.savez_compressed(os.path.join(OUT, name), **rec)
    print("wrote", name)


GRID_THETAS = (64, 96, 768)
GRID_CHEBS = ((48, 0.35, 1.0), (33, 0.5, 1.7), (128, 0.35 / 0.65, 1.0 / 0.65), (256, 0.35, 1.0))


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("compile the reference checker first: make -C oracle ref")
    # the reference's own unit-test grid (tests/unit_tests/SHT/main_input: N_theta=64 -> l_max=41)
    legendre_case("legendre_L41_nt64.npz", 64, 41, O.m_order_balanced(41), 2, 1, seed=41)
    # the l_max=63, N_theta=96 grid, balanced m order, reduced radial/field count
    legendre_case("legendre_L63_nt96.npz", 96, 63, O.m_order_balanced(63), 2, 1, seed=63, sample_tables=4000)
    # an arbitrary partial m list (one rank of a row: m values not contiguous)
    legendre_case("legendre_L63_partial.npz", 96, 63, np.array([17, 0, 63, 5, 40], np.int32), 3, 2, seed=5)
    # high-l tables where P_mm underflows double at the polar nodes: sampled entries
    legendre_case("legendre_L511_nt768_sampled.npz", 768, 511, np.arange(512, dtype=np.int32), 0, 0,
                  seed=511, sample_tables=6000)
    cheby_case("cheby_n48.npz", 48, 6, seed=48)
    cheby_case("cheby_n33.npz", 33, 4, seed=33)
    cheby_deriv_case("cheby_deriv_n48.npz", 48, 6, seed=480, rmin=0.35, rmax=1.0)
    cheby_deriv_case("cheby_deriv_n33.npz", 33, 5, seed=330, rmin=0.5, rmax=1.7)
    cheby_deriv_case("cheby_deriv_n128.npz", 128, 3, seed=1280, rmin=0.35 / 0.65, rmax=1.0 / 0.65)
    angular_case("angular_L12_ascending.npz", 12, np.arange(13), 3, seed=12)
    angular_case("angular_L41_balanced.npz", 41, O.m_order_balanced(41), 2, seed=41)
    angular_case("angular_L9_partial.npz", 9, [9, 0, 4, 3, 1], 5, seed=9)
    rlm_case("rlm_spaceb_L12_ascending.npz", 12, np.arange(13), 3, 5, (1, 2, 3), seed=120)
    rlm_case("rlm_spaceb_L41_partial.npz", 41, [41, 0, 17, 3, 40, 1], 4, 4, (3, 0, 2), seed=410)
    grid_case("grid_nodes_weights.npz", GRID_THETAS, GRID_CHEBS)
    SPACEA_CASES = (
        ("rlm_spacea_L12_ascending.npz", 12, np.arange(13), 3, 16, tuple(range(10)), (10, 11, 12, 13, 14, 15),
         (12, 11), True, 121),
        ("rlm_spacea_L41_partial.npz", 41, [41, 0, 17, 3, 40, 1], 4, 17, (11, 3, 0, 7, 2, 9, 1, 5, 12, 6),
         (16, 4, 13, 8, 10, 14), (4, 15), False, 411),
    )
    for args in SPACEA_CASES:
        rlm_spacea_case(*args)
