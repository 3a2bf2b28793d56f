"""ctypes wrappers for the checkers: the C restatement (liboracle.so) and, where it
was built, the compiled reference (libraysht_ref.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only to check or time the reference
path; the product never does.

Layouts are the reference's (see raysht_oracle.h):
  spectral  packed over mp in m_values order, block mp = data(m:lmax, nr, 2, nf), l fastest
  hybrid    p2[mp][c][theta] with c = r + imi*nr + 2*nr*f
  physical  p3[f][theta][r][nphi+2]
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
ORACLE_SO = os.path.join(REF_DIR, "liboracle.so")
REF_SO = os.path.join(REF_DIR, "libraysht_ref.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_long
_D = ctypes.c_double


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def build_oracle(with_reference: bool = False) -> None:
    """Compile the C restatement (and the compiled reference when asked and its sources exist)."""
    targets = ["oracle"]
    if with_reference and os.path.isdir("/root/reference/src"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        lib = ctypes.CDLL(ORACLE_SO)
        lib.orc_gauss_legendre.argtypes = [_I, _P, _P, _P, _P]
        lib.orc_table_size_even.restype = _L
        lib.orc_table_size_even.argtypes = [_I, _I, _P, _I]
        lib.orc_table_size_odd.restype = _L
        lib.orc_table_size_odd.argtypes = [_I, _I, _P, _I]
        lib.orc_legendre_tables.restype = _I
        lib.orc_legendre_tables.argtypes = [_I, _I, _P, _I, _P, _P, _P, _P]
        lib.orc_stp.argtypes = [_I, _I, _P, _I, _I, _I, _P, _P, _P, _P]
        lib.orc_pts.argtypes = [_I, _I, _P, _I, _I, _I, _P, _P, _P, _P]
        lib.orc_stp_cols.argtypes = [_I, _I, _P, _I, _I, _P, _P, _P, _P]
        lib.orc_pts_cols.argtypes = [_I, _I, _P, _I, _I, _P, _P, _P, _P]
        lib.orc_fft_c2r.argtypes = [_I, _L, _P]
        lib.orc_fft_r2c.argtypes = [_I, _L, _P, _I]
        lib.orc_permute_2a3a.argtypes = [_I, _I, _P, _I, _I, _I, _I, _P, _P]
        lib.orc_permute_3b2b.argtypes = [_I, _I, _P, _I, _I, _I, _I, _P, _P]
        lib.orc_angular_coefs.argtypes = [_I, _I, _I, _P]
        lib.orc_angular_dtheta.argtypes = [_I, _I, _P, _I, _I, _P, _I, _I, _P, _I, _I]
        lib.orc_angular_dphi.argtypes = [_I, _P, _I, _I, _P, _I, _I, _P, _I, _I]
        lib.orc_rlm_spaceb.argtypes = [_I, _P, _I, _I, _P, _I, _I, _I, _I]
        lib.orc_lm_scale.argtypes = [_I, _P, _I, _I, _P, _I, _I, _P, _I, _I, _P, _P]
        lib.orc_zero_lmax.argtypes = [_I, _P, _I, _I, _P, _I]
        lib.orc_rlm_spacea_hydro.argtypes = [_I, _P, _I, _I, _P, _I, _P, _P, _P, _P, _P, _P, _P]
        lib.orc_compute_bandcurlb.argtypes = [_I, _P, _I, _I, _P, _I, _P, _P]
        lib.orc_adjust_emf.argtypes = [_I, _P, _I, _I, _P, _I, _I, _I]
        lib.orc_hybrid_output_initial.argtypes = [_I, _P, _I, _I, _P, _I, _P, _I, _P, _P, _P]
        lib.orc_hybrid_output_final.argtypes = [_I, _P, _I, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P]
        lib.orc_lu_decompose.restype = _I
        lib.orc_lu_decompose.argtypes = [_I, _P, _P]
        lib.orc_lu_solve.argtypes = [_I, _P, _P, _I, _P]
        lib.orc_band_lu_decompose.restype = _I
        lib.orc_band_lu_decompose.argtypes = [_I, _I, _I, _P, _I, _P]
        lib.orc_band_lu_solve.argtypes = [_I, _I, _I, _P, _I, _P, _I, _P]
        lib.orc_lm_values.restype = _I
        lib.orc_lm_values.argtypes = [_I, _P, _I, _P, _P]
        lib.orc_permute_2b1b.argtypes = [_I, _P, _I, _I, _I, _P, _P]
        lib.orc_permute_1a2a.argtypes = [_I, _P, _I, _I, _I, _P, _P]
        lib.orc_cheby_grid.argtypes = [_I, _D, _D, _P]
        lib.orc_cheby_tables.argtypes = [_I, _P, _P]
        lib.orc_cheby_weights.argtypes = [_I, _D, _D, _P]
        lib.orc_cheby_weights_domain.argtypes = [_I, _D, _D, _D, _D, _P]
        lib.orc_gauss_legendre_binary128.argtypes = [_I, _P, _P]
        lib.orc_cheby_to_spectral.argtypes = [_I, _L, _P, _P, _P, _P]
        lib.orc_cheby_from_spectral.argtypes = [_I, _L, _P, _P, _P, _P]
        lib.orc_cheby_dealias.argtypes = [_I, _L, _P]
        lib.orc_cheby_deriv_scaling.restype = _D
        lib.orc_cheby_deriv_scaling.argtypes = [_I, _D, _D]
        lib.orc_cheby_deriv.argtypes = [_I, _D, _I, _L, _P, _P, _I]
        lib.orc_m_order_balanced.argtypes = [_I, _P]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_SO)
        lib.ref_legendre_init.argtypes = [_I, _I, _I, _P]
        lib.ref_legendre_nodes.argtypes = [_P, _P, _P, _P]
        lib.ref_legendre_nodes_q.argtypes = [_P, _P]
        lib.ref_legendre_tables.argtypes = [_P, _P, _P, _P]
        lib.ref_stp.argtypes = [_I, _I, _P, _P]
        lib.ref_pts.argtypes = [_I, _I, _P, _P]
        lib.ref_cheby_init.argtypes = [_I, _D, _D, _P, _P]
        lib.ref_cheby_tables.argtypes = [_P, _P]
        lib.ref_cheby_init_domains.argtypes = [_I, _P, _P, _I, _P, _I, _P, _P]
        lib.ref_cheby_grid_arrays.argtypes = [_P, _P]
        lib.ref_cheby_dealias.argtypes = [_I, _P]
        lib.ref_cheby_to_spectral.argtypes = [_I, _P, _P]
        lib.ref_cheby_from_spectral.argtypes = [_I, _P, _P]
        lib.ref_cheby_deriv.argtypes = [_I, _I, _P, _P]
        lib.ref_cheby_deriv_nobuffer.argtypes = [_I, _I, _P, _P]
        lib.ref_angular_init.argtypes = [_I, _P, _I, _I]
        lib.ref_angular_deriv.argtypes = [_I, _I, _I, _I, _P, _P]
        lib.ref_rlm_spaceb.argtypes = [_I, _I, _I, _I, _P]
        lib.ref_rlm_spacea_hydro.argtypes = [_I, _P, _I, _P, _P, _P, _P, _P, _P, _P]
        lib.ref_compute_bandcurlb.argtypes = [_I, _P, _P, _P]
        lib.ref_adjust_emf.argtypes = [_I, _I, _I, _P]
        lib.ref_hybrid_output.argtypes = [_I, _I, _P, _P, _I, _P, _P, _P, _P, _P]
        _ref = lib
    return _ref


# ---------------------------------------------------------------------------
# shapes and synthetic inputs


def n_even(lmax: int, m: int) -> int:
    return (lmax - m + 2) // 2


def n_odd(lmax: int, m: int) -> int:
    return (lmax - m + 1) // 2


def spectral_size(lmax: int, mvals, nr: int, nf: int) -> int:
    return int(sum((lmax - int(m) + 1) for m in mvals)) * nr * 2 * nf


def spectral_offsets(lmax: int, mvals, nr: int, nf: int) -> np.ndarray:
    sizes = [(lmax - int(m) + 1) * nr * 2 * nf for m in mvals]
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


def m_order_balanced(lmax: int) -> np.ndarray:
    mv = np.zeros(lmax + 1, dtype=np.int32)
    oracle_lib().orc_m_order_balanced(lmax, _ptr(mv))
    return mv


def synthetic_spectral(lmax: int, mvals, nr: int, nf: int, seed: int) -> np.ndarray:
    """Coefficients a_lm(r) = (N(0,1) + i N(0,1)) (l+1)^-2 sin(pi r_i / nr), i = 1..nr,
    split into Re/Im planes; the m = 0 imaginary plane is zero.  Packed layout."""
    rng = np.random.default_rng(seed)
    rfac = np.sin(np.pi * np.arange(1, nr + 1) / nr)
    out = np.empty(spectral_size(lmax, mvals, nr, nf))
    off = 0
    for m in mvals:
        m = int(m)
        nl = lmax - m + 1
        l = np.arange(m, lmax + 1)
        blk = rng.standard_normal((nf, 2, nr, nl))
        blk *= ((l + 1.0) ** -2)[None, None, None, :]
        blk *= rfac[None, None, :, None]
        if m == 0:
            blk[:, 1] = 0.0
        out[off:off + blk.size] = blk.ravel()
        off += blk.size
    return out


# ---------------------------------------------------------------------------
# the restatement


def gauss_legendre(n: int):
    xh, xl, wh, wl = (np.zeros(n) for _ in range(4))
    oracle_lib().orc_gauss_legendre(n, _ptr(xh), _ptr(xl), _ptr(wh), _ptr(wl))
    return xh, xl, wh, wl


def gauss_legendre_binary128(n: int):
    """The nodes and weights as the reference stores them (Real*16): (n, 16) byte arrays."""
    x, w = np.zeros((n, 16), np.uint8), np.zeros((n, 16), np.uint8)
    oracle_lib().orc_gauss_legendre_binary128(n, _ptr(x), _ptr(w))
    return x, w


def legendre_tables(nt: int, lmax: int, mvals):
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    lib = oracle_lib()
    ne = lib.orc_table_size_even(nt, lmax, _ptr(mv), len(mv))
    no = lib.orc_table_size_odd(nt, lmax, _ptr(mv), len(mv))
    pe, ie = np.zeros(ne), np.zeros(ne)
    po, io = np.zeros(max(no, 1)), np.zeros(max(no, 1))
    rc = lib.orc_legendre_tables(nt, lmax, _ptr(mv), len(mv), _ptr(pe), _ptr(po), _ptr(ie), _ptr(io))
    if rc != 0:
        raise ValueError(f"orc_legendre_tables failed ({rc})")
    return pe, po[:no], ie, io[:no]


def stp(nt, lmax, mvals, nr, nf, tables, s_in):
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    pe, po, _, _ = tables
    po = po if po.size else np.zeros(1)
    out = np.zeros(nt * 2 * nr * nf * len(mv))
    oracle_lib().orc_stp(nt, lmax, _ptr(mv), len(mv), nr, nf, _ptr(pe), _ptr(po),
                         _ptr(np.ascontiguousarray(s_in)), _ptr(out))
    return out


def pts(nt, lmax, mvals, nr, nf, tables, p_in):
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    _, _, ie, io = tables
    io = io if io.size else np.zeros(1)
    out = np.zeros(spectral_size(lmax, mv, nr, nf))
    oracle_lib().orc_pts(nt, lmax, _ptr(mv), len(mv), nr, nf, _ptr(ie), _ptr(io),
                         _ptr(np.ascontiguousarray(p_in)), _ptr(out))
    return out


def stp_cols(nt, lmax, mvals, ncols, tables, s_in):
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    pe, po, _, _ = tables
    po = po if po.size else np.zeros(1)
    out = np.zeros(nt * ncols * len(mv))
    oracle_lib().orc_stp_cols(nt, lmax, _ptr(mv), len(mv), ncols, _ptr(pe), _ptr(po),
                              _ptr(np.ascontiguousarray(s_in)), _ptr(out))
    return out


def pts_cols(nt, lmax, mvals, ncols, tables, p_in):
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    _, _, ie, io = tables
    io = io if io.size else np.zeros(1)
    out = np.zeros(int(sum(lmax - int(m) + 1 for m in mv)) * ncols)
    oracle_lib().orc_pts_cols(nt, lmax, _ptr(mv), len(mv), ncols, _ptr(ie), _ptr(io),
                              _ptr(np.ascontiguousarray(p_in)), _ptr(out))
    return out


def fft_c2r(nphi: int, x: np.ndarray) -> np.ndarray:
    y = np.array(x, dtype=np.float64, copy=True)
    oracle_lib().orc_fft_c2r(nphi, y.size // (nphi + 2), _ptr(y))
    return y


def fft_r2c(nphi: int, x: np.ndarray, rescale: bool = False) -> np.ndarray:
    y = np.array(x, dtype=np.float64, copy=True)
    oracle_lib().orc_fft_r2c(nphi, y.size // (nphi + 2), _ptr(y), int(rescale))
    return y


def permute_2a3a(nt, lmax, mvals, nr, nf, nphi, p2a):
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    out = np.empty((nphi + 2) * nr * nt * nf)
    oracle_lib().orc_permute_2a3a(nt, lmax, _ptr(mv), len(mv), nr, nf, nphi,
                                  _ptr(np.ascontiguousarray(p2a)), _ptr(out))
    return out


def permute_3b2b(nt, lmax, mvals, nr, nf, nphi, p3b):
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    out = np.zeros(nt * 2 * nr * nf * len(mv))
    oracle_lib().orc_permute_3b2b(nt, lmax, _ptr(mv), len(mv), nr, nf, nphi,
                                  _ptr(np.ascontiguousarray(p3b)), _ptr(out))
    return out


ANGULAR = {"dtheta": 0, "sdtheta": 1, "dphi": 2}


def angular_deriv(kind, lmax, mvals, nr, a, nf_in, fin, out=None, nf_out=1, fout=0):
    """Spectral_Derivatives.F90: d_by_dtheta / d_by_sdtheta / d_by_dphi of field fin (0-based)
    of the packed array a (nf_in fields) into field fout of out (nf_out fields; a fresh
    one-field array when out is None).  Returns out."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    a = np.ascontiguousarray(a, dtype=np.float64)
    if out is None:
        out = np.zeros(spectral_size(lmax, mv, nr, nf_out))
    if kind == "dphi":
        oracle_lib().orc_angular_dphi(lmax, _ptr(mv), len(mv), nr, _ptr(a), nf_in, fin, _ptr(out), nf_out, fout)
    else:
        oracle_lib().orc_angular_dtheta(ANGULAR[kind], lmax, _ptr(mv), len(mv), nr, _ptr(a), nf_in, fin,
                                        _ptr(out), nf_out, fout)
    return out


def rlm_spaceb(lmax, mvals, nr, buf, nf, wv, pv, zv):
    """Sphere_Hybrid_Space.F90:127-214 rlm_spaceb (hydro part) on a copy of the packed nf-field
    buffer (0-based W, P, Z fields); returns the new buffer."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    b = np.ascontiguousarray(buf, dtype=np.float64).copy()
    oracle_lib().orc_rlm_spaceb(lmax, _ptr(mv), len(mv), nr, _ptr(b), nf, wv, pv, zv)
    return b


# radial profiles of the rlm_spacea steps, in the order the entry points take them
HYDRO_PROFILES = ("over_rhor", "over_rhorsq", "drho_term", "two_over_r", "dlnrho", "dsdr_over_cp")


def rlm_spacea_hydro(lmax, mvals, nr, buf, nf, fields, prof):
    """Sphere_Hybrid_Space.F90:235-355 Velocity_Components + Velocity_Derivatives and :80
    d_by_dtheta(tvar, dtdt) on a copy of the packed nf-field buffer; fields = 0-based (W, dWdr,
    d2Wdr2, Z, dZdr, T, vtheta, vphi, dvtdr, dvpdr); prof = dict of the HYDRO_PROFILES arrays
    (N_r values each; dsdr_over_cp None unless pseudo-incompressible).  Returns the new buffer."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    b = np.ascontiguousarray(buf, dtype=np.float64).copy()
    v = np.ascontiguousarray(fields, dtype=np.int32)
    assert v.size == 10
    arrs = [None if prof.get(k) is None else np.ascontiguousarray(prof[k], dtype=np.float64) for k in HYDRO_PROFILES]
    oracle_lib().orc_rlm_spacea_hydro(lmax, _ptr(mv), len(mv), nr, _ptr(b), nf, _ptr(v),
                                      *[None if x is None else _ptr(x) for x in arrs])
    return b


def compute_bandcurlb(lmax, mvals, nr, buf, nf, fields, one_over_rsq):
    """Sphere_Hybrid_Space.F90:377-460 Compute_BandCurlB on a copy of the packed nf-field buffer;
    fields = 0-based (C, A, dCdr, dAdr, d2Cdr2, curlbr).  Returns the new buffer."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    b = np.ascontiguousarray(buf, dtype=np.float64).copy()
    v = np.ascontiguousarray(fields, dtype=np.int32)
    assert v.size == 6
    o2 = np.ascontiguousarray(one_over_rsq, dtype=np.float64)
    oracle_lib().orc_compute_bandcurlb(lmax, _ptr(mv), len(mv), nr, _ptr(b), nf, _ptr(v), _ptr(o2))
    return b


def adjust_emf(lmax, mvals, nr, buf, nf, et, ep):
    """Sphere_Hybrid_Space.F90:557-602 Adjust_Emf on a copy of the packed nf-field buffer (0-based
    emftheta, emfphi fields); returns the new buffer."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    b = np.ascontiguousarray(buf, dtype=np.float64).copy()
    oracle_lib().orc_adjust_emf(lmax, _ptr(mv), len(mv), nr, _ptr(b), nf, et, ep)
    return b


def hybrid_output(lmax, mvals, nr, s, nf, co, nf_co, vi, vf, one_over_r, one_over_rsq, two_over_r, magnetism=True):
    """Sphere_Hybrid_Space.F90:224-233,437-552: Hybrid_Output_Initial's magnetic part (into the
    co-buffer and the two work fields) then Hybrid_Output_Final up to its transform, on copies;
    vi = 0-based (avar, dcdr, d2cdr2, dadr, avar_cb, dbrdr_cb), vf = 0-based (pvar, br, btheta,
    bphi, dpdt_cb, dbtdr_cb, dbpdr_cb, dbrdr_cb, dbrdt_cb).  Returns the new co-buffer."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    sb = np.ascontiguousarray(s, dtype=np.float64)
    c = np.ascontiguousarray(co, dtype=np.float64).copy()
    a_i = np.ascontiguousarray(vi, dtype=np.int32)
    a_f = np.ascontiguousarray(vf, dtype=np.int32)
    rr = [np.ascontiguousarray(x, dtype=np.float64) for x in (one_over_r, one_over_rsq, two_over_r)]
    tmp = np.zeros(spectral_size(lmax, mv, nr, 2))
    if magnetism:
        oracle_lib().orc_hybrid_output_initial(lmax, _ptr(mv), len(mv), nr, _ptr(sb), nf, _ptr(c), nf_co, _ptr(tmp),
                                               _ptr(a_i), _ptr(rr[1]))
    oracle_lib().orc_hybrid_output_final(lmax, _ptr(mv), len(mv), nr, _ptr(sb), nf, _ptr(c), nf_co,
                                         _ptr(tmp) if magnetism else None, _ptr(a_f), *[_ptr(x) for x in rr])
    return c


def lm_scale(lmax, mvals, nr, a, nf_in, fin, lfac=None, rfac=None, out=None, nf_out=1, fout=0):
    """(lfac(l) a(l, r, imi)) rfac(r) of field fin into field fout of out (a fresh one-field
    array when out is None); returns out."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    a = np.ascontiguousarray(a, dtype=np.float64)
    if out is None:
        out = np.zeros(spectral_size(lmax, mv, nr, 1))
    lf = None if lfac is None else np.ascontiguousarray(lfac, dtype=np.float64)
    rf = None if rfac is None else np.ascontiguousarray(rfac, dtype=np.float64)
    oracle_lib().orc_lm_scale(lmax, _ptr(mv), len(mv), nr, _ptr(a), nf_in, fin, _ptr(out), nf_out, fout,
                              None if lf is None else _ptr(lf), None if rf is None else _ptr(rf))
    return out


def zero_lmax(lmax, mvals, nr, buf, nf):
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    b = np.ascontiguousarray(buf, dtype=np.float64).copy()
    oracle_lib().orc_zero_lmax(lmax, _ptr(mv), len(mv), nr, _ptr(b), nf)
    return b


def angular_coefs(l, m, kind):
    c = np.zeros(2)
    oracle_lib().orc_angular_coefs(l, m, ANGULAR[kind], _ptr(c))
    return c


def lu_decompose(a):
    """LU_Decompose_full (LAPACK dgetrf) of one n x n matrix given as a column-major
    (Fortran-order) array: returns (LU, 1-based pivots, info)."""
    n = a.shape[0]
    lu = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    piv = np.zeros(n, np.int32)
    info = oracle_lib().orc_lu_decompose(n, _ptr(lu), _ptr(piv))
    return lu, piv, info


def lu_solve(lu, piv, b):
    """LU_Solve_full (LAPACK dgetrs 'N'): b is n x nrhs, column-major."""
    n = lu.shape[0]
    x = np.asfortranarray(b, dtype=np.float64).copy(order="F")
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    oracle_lib().orc_lu_solve(n, _ptr(np.asfortranarray(lu)), _ptr(np.ascontiguousarray(piv, np.int32)), nrhs, _ptr(x))
    return x


def band_lu_decompose(ab, kl, ku):
    """LU_Decompose_band (LAPACK dgbtrf, unblocked dgbtf2 order) of a band matrix in LAPACK
    band storage ab(ldab, n) (column-major, ldab >= 2 kl + ku + 1): returns (factor, 1-based
    pivots, info)."""
    f = np.asfortranarray(ab, dtype=np.float64).copy(order="F")
    ldab, n = f.shape
    piv = np.zeros(n, np.int32)
    info = oracle_lib().orc_band_lu_decompose(n, kl, ku, _ptr(f), ldab, _ptr(piv))
    return f, piv, info


def band_lu_solve(fac, kl, ku, piv, b):
    """LU_Solve_Band (LAPACK dgbtrs 'N'): b is n x nrhs, column-major."""
    fac = np.asfortranarray(fac, dtype=np.float64)
    ldab, n = fac.shape
    x = np.asfortranarray(b, dtype=np.float64).copy(order="F")
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    oracle_lib().orc_band_lu_solve(n, kl, ku, _ptr(fac), ldab, _ptr(np.ascontiguousarray(piv, np.int32)), nrhs,
                                   _ptr(x))
    return x


def lm_values(lmax, mvals):
    """LM_Load_Balance's (l, mp) pair list on one radial rank (Load_Balance.F90:311-395)."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    n = oracle_lib().orc_lm_values(lmax, _ptr(mv), len(mv), None, None)
    ls, mps = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
    oracle_lib().orc_lm_values(lmax, _ptr(mv), len(mv), _ptr(ls), _ptr(mps))
    return ls[:n], mps[:n]


def permute_2b1b(lmax, mvals, nr, nf, s2b):
    """Transpose_2b1b on one rank (Spherical_Buffer.F90:705-843): p1b[f][lm][imi][r]."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    n = len(lm_values(lmax, mv)[0])
    s2b = np.ascontiguousarray(s2b, dtype=np.float64)
    out = np.zeros(nf * n * 2 * nr)
    oracle_lib().orc_permute_2b1b(lmax, _ptr(mv), len(mv), nr, nf, _ptr(s2b), _ptr(out))
    return out


def permute_1a2a(lmax, mvals, nr, nf, p1a):
    """Transpose_1a2a on one rank (Spherical_Buffer.F90:845-932): the packed spectral layout."""
    mv = np.ascontiguousarray(mvals, dtype=np.int32)
    p1a = np.ascontiguousarray(p1a, dtype=np.float64)
    out = np.zeros(spectral_size(lmax, mv, nr, nf))
    oracle_lib().orc_permute_1a2a(lmax, _ptr(mv), len(mv), nr, nf, _ptr(p1a), _ptr(out))
    return out


def cheby_grid(n: int, rmin: float, rmax: float) -> np.ndarray:
    g = np.zeros(n)
    oracle_lib().orc_cheby_grid(n, rmin, rmax, _ptr(g))
    return g


def cheby_weights(n: int, rmin: float, rmax: float) -> np.ndarray:
    """Initialize_Cheby_Grid's radial integration weights (one domain)."""
    w = np.zeros(n)
    oracle_lib().orc_cheby_weights(n, rmin, rmax, _ptr(w))
    return w


def cheby_tables(n: int):
    nx = n - n // 2
    ce, co = np.zeros(nx * nx), np.zeros(max(nx * (n // 2), 1))
    oracle_lib().orc_cheby_tables(n, _ptr(ce), _ptr(co))
    return ce, co


def cheby_to_spectral(n: int, f: np.ndarray) -> np.ndarray:
    ce, co = cheby_tables(n)
    f = np.ascontiguousarray(f, dtype=np.float64)
    c = np.zeros_like(f)
    oracle_lib().orc_cheby_to_spectral(n, f.size // n, _ptr(ce), _ptr(co), _ptr(f), _ptr(c))
    return c


# whether the compiled reference's recursion step b(i+2) + 2(i+1) a(i+1) is one fused
# multiply-add (fixed by comparing against it: tests/test_oracle_cheby_deriv.py)
CHEBY_DERIV_FUSED = False


def cheby_from_spectral(n: int, c: np.ndarray) -> np.ndarray:
    ce, co = cheby_tables(n)
    c = np.ascontiguousarray(c, dtype=np.float64)
    f = np.zeros_like(c)
    oracle_lib().orc_cheby_from_spectral(n, c.size // n, _ptr(ce), _ptr(co), _ptr(c), _ptr(f))
    return f


def cheby_angles_longdouble(n: int) -> np.ndarray:
    """The angles the reference's Gen_Tn tables use for every radius, in long double:
    theta_i accumulated in doubles by repeated addition (Chebyshev_Polynomials.F90:233-243)
    for the first n_x radii, pi - theta_(n-1-i) for the rest (the parity split of
    To/From_Spectral4D, :375-558, applies the first half's cosines to the mirror points)."""
    d = np.pi / n
    arg = d * 0.5
    th = np.zeros(n)
    for i in range(n):
        th[i] = arg
        arg = arg + d
    nx = n - n // 2
    pi_l = np.longdouble("3.14159265358979323846264338327950288")
    out = th.astype(np.longdouble)
    out[nx:] = pi_l - out[n - 1 - np.arange(nx, n)]
    return out


def cheby_to_spectral_exact(n: int, f: np.ndarray, theta=None) -> np.ndarray:
    """To_Spectral4D's sums c_k = (2/n) sum_i f_i cos(k theta_i) on the reference's angles
    (or on theta, long double, if given), with the cosines and sums in long double (the
    reference's tables round k theta_i and cos to doubles, which moves its result by up to
    about 1e-14 of max |c|)."""
    th = cheby_angles_longdouble(n) if theta is None else np.asarray(theta, dtype=np.longdouble)
    k = np.arange(n, dtype=np.longdouble)
    T = np.cos(np.outer(k, th))  # [k][i]
    f = np.asarray(f, dtype=np.float64).reshape(-1, n).astype(np.longdouble)
    return ((f @ T.T) * (np.longdouble(2) / n)).astype(np.float64).ravel()


def cheby_from_spectral_exact(n: int, c: np.ndarray, theta=None) -> np.ndarray:
    """From_Spectral4D's sums f_i = sum_k c_k cos(k theta_i) - c_0 / 2 on the reference's
    angles (or on theta), in long double."""
    th = cheby_angles_longdouble(n) if theta is None else np.asarray(theta, dtype=np.longdouble)
    k = np.arange(n, dtype=np.longdouble)
    T = np.cos(np.outer(k, th))
    c = np.asarray(c, dtype=np.float64).reshape(-1, n).astype(np.longdouble)
    return (c @ T - c[:, :1] / 2).astype(np.float64).ravel()


def cheby_deriv_scaling(n: int, rmin: float, rmax: float) -> float:
    return oracle_lib().orc_cheby_deriv_scaling(n, rmin, rmax)


def cheby_deriv(n: int, rmin: float, rmax: float, c: np.ndarray, dorder: int, fused: bool = CHEBY_DERIV_FUSED) -> np.ndarray:
    """Chebyshev_Polynomials.F90:560-730: the dorder-th radial derivative's coefficients
    of every column of n (d_by_dr_cp / d_by_dr_nobuffer), one domain on [rmin, rmax]."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    d = np.zeros_like(c)
    ds = cheby_deriv_scaling(n, rmin, rmax)
    oracle_lib().orc_cheby_deriv(n, ds, dorder, c.size // n, _ptr(c), _ptr(d), int(fused))
    return d


def cheby_domains(npoly, bounds, dealias_by=None) -> dict:
    """Initialize_Cheby_Grid (Chebyshev_Polynomials.F90:101-200) with several domains: npoly and
    bounds in ascending order of radius as the reference takes them, the domains kept outermost
    first (internal domain h is input domain nd - 1 - h); per domain its row offset, its
    1-based first dealiased row rda ((2n)/3 + 1, or n - dealias_by + 1 for 1 <= dealias_by < n),
    the one-domain map onto its bounds (grid, derivative scaling) and the weights with the
    whole grid's gmax^3 - gmin^3."""
    nd = len(npoly)
    gmin, gmax = float(bounds[0]), float(bounds[nd])
    g = dict(n=[], hoff=[], rda=[], lo=[], hi=[])
    grid, weights, dscale = [], [], []
    hoff = 0
    for h in range(nd):
        src = nd - 1 - h
        n, lo, hi = int(npoly[src]), float(bounds[src]), float(bounds[src + 1])
        rda = (2 * n) // 3 + 1
        if dealias_by is not None and 1 <= int(dealias_by[src]) < n:
            rda = n - int(dealias_by[src]) + 1
        w = np.zeros(n)
        oracle_lib().orc_cheby_weights_domain(n, lo, hi, gmin, gmax, _ptr(w))
        grid.append(cheby_grid(n, lo, hi))
        weights.append(w)
        dscale.append(np.full(n, cheby_deriv_scaling(n, lo, hi)))
        for k, v in (("n", n), ("hoff", hoff), ("rda", rda), ("lo", lo), ("hi", hi)):
            g[k].append(v)
        hoff += n
    g.update(ntotal=hoff, grid=np.concatenate(grid), weights=np.concatenate(weights), dscale=np.concatenate(dscale))
    return g


def _per_domain(g, a, fn):
    """fn(h, block) on every domain's rows of every column of length g['ntotal']"""
    a = np.asarray(a, dtype=np.float64).reshape(-1, g["ntotal"])
    out = np.zeros_like(a)
    for h, (n, o) in enumerate(zip(g["n"], g["hoff"])):
        out[:, o:o + n] = np.asarray(fn(h, np.ascontiguousarray(a[:, o:o + n]))).reshape(-1, n)
    return out.ravel()


def cheby_domains_to_spectral(g, f):
    """To_Spectral4D's loop over domains (:390-445): each domain's one-domain transform on its rows."""
    return _per_domain(g, f, lambda h, b: cheby_to_spectral(g["n"][h], b.ravel()))


def cheby_domains_from_spectral(g, c):
    """From_Spectral4D's loop over domains (:466-556), the c_0 / 2 of each domain included."""
    return _per_domain(g, c, lambda h, b: cheby_from_spectral(g["n"][h], b.ravel()))


def cheby_domains_deriv(g, c, dorder: int):
    """Cheby_Deriv_Buffer_4D over domains (:560-644): the recursion per domain, then each row
    times its domain's deriv_scaling**dorder."""
    return _per_domain(g, c, lambda h, b: cheby_deriv(g["n"][h], g["lo"][h], g["hi"][h], b.ravel(), dorder))


def cheby_domains_dealias(g, c):
    """dealias_buffer4d (:353-374): rows rda .. npoly (1-based) of every domain zeroed."""
    def zero(h, b):
        b = b.copy()
        b[:, g["rda"][h] - 1:] = 0.0
        return b
    return _per_domain(g, c, zero)


def cheby_dealias(n: int, c: np.ndarray) -> np.ndarray:
    y = np.array(c, dtype=np.float64, copy=True)
    oracle_lib().orc_cheby_dealias(n, y.size // n, _ptr(y))
    return y


# ---------------------------------------------------------------------------
# the compiled reference (present only where oracle/_ref was built from the sources)


class Reference:
    """The reference's own Math_Layer modules behind the bind(C) harness."""

    def __init__(self):
        self.lib = ref_lib()
        self.nt = self.lmax = None
        self.mvals = None

    def legendre_init(self, nt: int, lmax: int, mvals):
        self.mvals = np.ascontiguousarray(mvals, dtype=np.int32)
        self.nt, self.lmax = nt, lmax
        self.lib.ref_legendre_init(nt, lmax, len(self.mvals), _ptr(self.mvals))

    def nodes(self):
        xh, xl, wh, wl = (np.zeros(self.nt) for _ in range(4))
        self.lib.ref_legendre_nodes(_ptr(xh), _ptr(xl), _ptr(wh), _ptr(wl))
        return xh, xl, wh, wl

    def nodes_binary128(self):
        """Real*16 coloc and gl_weights as (n_theta, 16) byte arrays."""
        x, w = np.zeros((self.nt, 16), np.uint8), np.zeros((self.nt, 16), np.uint8)
        self.lib.ref_legendre_nodes_q(_ptr(x), _ptr(w))
        return x, w

    def tables(self):
        lib = oracle_lib()
        mv = self.mvals
        ne = lib.orc_table_size_even(self.nt, self.lmax, _ptr(mv), len(mv))
        no = lib.orc_table_size_odd(self.nt, self.lmax, _ptr(mv), len(mv))
        pe, ie = np.zeros(ne), np.zeros(ne)
        po, io = np.zeros(max(no, 1)), np.zeros(max(no, 1))
        self.lib.ref_legendre_tables(_ptr(pe), _ptr(po), _ptr(ie), _ptr(io))
        return pe, po[:no], ie, io[:no]

    def stp(self, nr: int, nf: int, s_in: np.ndarray) -> np.ndarray:
        out = np.zeros(self.nt * 2 * nr * nf * len(self.mvals))
        self.lib.ref_stp(nr, nf, _ptr(np.ascontiguousarray(s_in)), _ptr(out))
        return out

    def pts(self, nr: int, nf: int, p_in: np.ndarray) -> np.ndarray:
        out = np.zeros(spectral_size(self.lmax, self.mvals, nr, nf))
        self.lib.ref_pts(nr, nf, _ptr(np.ascontiguousarray(p_in)), _ptr(out))
        return out

    def cheby_init(self, n: int, rmin: float, rmax: float):
        self.ncheb = n
        g, w = np.zeros(n), np.zeros(n)
        self.lib.ref_cheby_init(n, rmin, rmax, _ptr(g), _ptr(w))
        return g, w

    def cheby_tables(self):
        n = self.ncheb
        nx = n - n // 2
        ce, co = np.zeros(nx * nx), np.zeros(max(nx * (n // 2), 1))
        self.lib.ref_cheby_tables(_ptr(ce), _ptr(co))
        return ce, co

    def cheby_to_spectral(self, f: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(f, dtype=np.float64)
        c = np.zeros_like(f)
        self.lib.ref_cheby_to_spectral(f.size // self.ncheb, _ptr(f), _ptr(c))
        return c

    def cheby_init_domains(self, npoly, bounds, dealias_by=None):
        """the reference's Cheby_Grid%init with several domains; returns (grid, weights)"""
        nd = len(npoly)
        npl = np.ascontiguousarray(npoly, dtype=np.int32)
        bd = np.ascontiguousarray(bounds, dtype=np.float64)
        db = np.ascontiguousarray(dealias_by if dealias_by is not None else np.zeros(nd), dtype=np.int32)
        self.ncheb = int(npl.sum())
        grid, w = np.zeros(self.ncheb), np.zeros(self.ncheb)
        self.lib.ref_cheby_init_domains(nd, _ptr(npl), _ptr(bd), int(dealias_by is not None), _ptr(db), self.ncheb,
                                        _ptr(grid), _ptr(w))
        self.ndom = nd
        return grid, w

    def cheby_grid_arrays(self):
        """the grid's deriv_scaling (every radius) and rda (every domain, internal order)"""
        ds, rda = np.zeros(self.ncheb), np.zeros(self.ndom, dtype=np.int32)
        self.lib.ref_cheby_grid_arrays(_ptr(ds), _ptr(rda))
        return ds, rda

    def cheby_dealias(self, c: np.ndarray) -> np.ndarray:
        b = np.ascontiguousarray(c, dtype=np.float64).copy()
        self.lib.ref_cheby_dealias(b.size // self.ncheb, _ptr(b))
        return b

    def cheby_from_spectral(self, c: np.ndarray) -> np.ndarray:
        c = np.ascontiguousarray(c, dtype=np.float64)
        f = np.zeros_like(c)
        self.lib.ref_cheby_from_spectral(c.size // self.ncheb, _ptr(c), _ptr(f))
        return f

    def angular_init(self, lmax: int, mvals, nr: int):
        self.ang = (lmax, np.ascontiguousarray(mvals, dtype=np.int32), nr)
        mv = self.ang[1]
        self.lib.ref_angular_init(len(mv), _ptr(mv), lmax, nr)

    def angular_deriv(self, kind: str, buf: np.ndarray, nf: int, fin: int, fout: int):
        """The reference's derivative of field fin (1-based) of a packed nf-field buffer:
        fout >= 1 the buffer form (returns the whole buffer after the call), fout = 0 the
        4dbuff2arr form (returns the one-field array)."""
        lmax, mv, nr = self.ang
        b = np.ascontiguousarray(buf, dtype=np.float64).copy()
        arr = np.zeros(spectral_size(lmax, mv, nr, 1))
        self.lib.ref_angular_deriv(ANGULAR[kind], nf, fin, fout, _ptr(b), _ptr(arr))
        return b if fout >= 1 else arr

    def rlm_spaceb(self, buf: np.ndarray, nf: int, wv: int, pv: int, zv: int) -> np.ndarray:
        """rlm_spaceb's hydro steps through the reference's own d_by_sdtheta / d_by_dphi on a
        copy of the buffer (fields 1-based); returns the new buffer."""
        b = np.ascontiguousarray(buf, dtype=np.float64).copy()
        self.lib.ref_rlm_spaceb(nf, wv, pv, zv, _ptr(b))
        return b

    def rlm_spacea_hydro(self, buf: np.ndarray, nf: int, fields, prof) -> np.ndarray:
        """Velocity_Components, Velocity_Derivatives and d_by_dtheta(tvar, dtdt) through the
        reference's own d_by_dtheta / d_by_dphi on a copy of the buffer (fields 0-based, as
        rlm_spacea_hydro above); returns the new buffer."""
        b = np.ascontiguousarray(buf, dtype=np.float64).copy()
        v = np.ascontiguousarray(np.asarray(fields) + 1, dtype=np.int32)
        pseudo = prof.get("dsdr_over_cp") is not None
        arrs = [np.ascontiguousarray(prof[k] if prof.get(k) is not None else np.zeros(1), dtype=np.float64)
                for k in HYDRO_PROFILES]
        self.lib.ref_rlm_spacea_hydro(nf, _ptr(v), int(pseudo), *[_ptr(x) for x in arrs], _ptr(b))
        return b

    def compute_bandcurlb(self, buf: np.ndarray, nf: int, fields, one_over_rsq) -> np.ndarray:
        """Compute_BandCurlB through the reference's own derivatives (fields 0-based)."""
        b = np.ascontiguousarray(buf, dtype=np.float64).copy()
        v = np.ascontiguousarray(np.asarray(fields) + 1, dtype=np.int32)
        o2 = np.ascontiguousarray(one_over_rsq, dtype=np.float64)
        self.lib.ref_compute_bandcurlb(nf, _ptr(v), _ptr(o2), _ptr(b))
        return b

    def adjust_emf(self, buf: np.ndarray, nf: int, et: int, ep: int) -> np.ndarray:
        """Adjust_Emf through the reference's own derivatives (fields 0-based)."""
        b = np.ascontiguousarray(buf, dtype=np.float64).copy()
        self.lib.ref_adjust_emf(nf, et + 1, ep + 1, _ptr(b))
        return b

    def hybrid_output(self, s, nf, co, nf_co, vi, vf, one_over_r, one_over_rsq, two_over_r, magnetism=True):
        """the reference's output-iteration steps (fields 0-based, as hybrid_output above)"""
        sb = np.ascontiguousarray(s, dtype=np.float64).copy()
        c = np.ascontiguousarray(co, dtype=np.float64).copy()
        a_i = np.ascontiguousarray(np.asarray(vi) + 1, dtype=np.int32)
        a_f = np.ascontiguousarray(np.asarray(vf) + 1, dtype=np.int32)
        rr = [np.ascontiguousarray(x, dtype=np.float64) for x in (one_over_r, one_over_rsq, two_over_r)]
        self.lib.ref_hybrid_output(nf, nf_co, _ptr(a_i), _ptr(a_f), int(magnetism), *[_ptr(x) for x in rr], _ptr(sb),
                                   _ptr(c))
        return c

    def cheby_deriv(self, c: np.ndarray, dorder: int, nobuffer: bool = False) -> np.ndarray:
        c = np.ascontiguousarray(c, dtype=np.float64)
        d = np.zeros_like(c)
        fn = self.lib.ref_cheby_deriv_nobuffer if nobuffer else self.lib.ref_cheby_deriv
        fn(c.size // self.ncheb, dorder, _ptr(c), _ptr(d))
        return d
