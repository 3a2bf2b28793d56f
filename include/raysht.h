/*
 * raysht -- MI355X-native spectral-transform engine for Rayleigh's per-timestep
 * transform stack.  Flat C ABI: plain pointers, sizes and int status codes.
 *
 * Every entry point names the reference interface (file:line under the
 * Rayleigh tree) whose body it replaces.  Data layouts are the reference's own
 * storage, so a caller hands over its buffers unchanged:
 *
 *   spectral  packed over mp = 0..n_m-1 in m_values order; block mp holds
 *             data(m:lmax, 1:nr, 1:2, 1:nf) with l fastest
 *             (rmcontainer4d, Structures.F90:37-39; Spherical_Buffer.F90:1459-1484)
 *   hybrid    p2[mp][c][theta], c = r + imi*nr + 2*nr*f, theta = 0..ntheta-1
 *             ascending cos(theta) from the south pole (Spherical_Buffer.F90:1391-1406)
 *   physical  p3[f][theta][r][nphi+2]: one in-place real-FFT line per (r, theta, f)
 *             (Spherical_Buffer.F90:1424-1440)
 *   radial    cheb buffer [col][n], radius fastest (Chebyshev_Polynomials.F90:375-558)
 *
 * Device pointers are HIP device memory on the current device; `stream` is a
 * hipStream_t (NULL = the default stream).  Calls are asynchronous on the
 * stream, validate their arguments before launching anything, and return
 * RAYSHT_OK or an error code (never abort).  A plan is used from one host
 * thread at a time.
 */
#ifndef RAYSHT_H
#define RAYSHT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct raysht_plan raysht_plan;
typedef struct raysht_cheb raysht_cheb;

enum raysht_status {
    RAYSHT_OK = 0,
    RAYSHT_EINVAL = 1,      /* bad size, layout or argument */
    RAYSHT_EHIP = 2,        /* the HIP runtime reported an error */
    RAYSHT_ENOMEM = 3,      /* device or host allocation failed */
    RAYSHT_EUNSUPPORTED = 4, /* valid request the library does not implement */
    RAYSHT_ENOTREADY = 5     /* a table the call may not make itself has not been made yet */
};

/* Arithmetic used by the Legendre contractions. */
enum raysht_precision {
    RAYSHT_FP64 = 0,    /* FP64 MFMA */
    RAYSHT_BF16X3 = 1,  /* 4-slice BF16 split, pairs of levels s + t <= 3, FP32 sums (about 1e-6) */
    RAYSHT_BF16X6 = 2,  /* 6-slice BF16 split, exact FP32 sums, FP64 recombination (about 1e-12) */
    RAYSHT_INT8X3 = 3,  /* the BF16X3 digits as signed bytes on INT8 MFMA, exact int32 sums */
    RAYSHT_INT8X6 = 4   /* the BF16X6 digits as signed bytes on INT8 MFMA, exact int32 sums */
};

const char *raysht_strerror(int status);
/* Library version as major*10000 + minor*100 + patch. */
int raysht_version(void);

/* Plan = Initialize_Legendre(nt, lmax, mval, parity=.true.)
 * (Legendre_Polynomials.F90:123-143): Gauss-Legendre nodes/weights and the
 * four parity P_lm tables (Compute_Plms :146-258, Parity_Resort :260-331) for
 * the given m list, resident on the device, for transforms of nr radii and
 * nfields fields.  ntheta must be even; m_values (host array) may be in any
 * order (Load_Balance.F90:126-286) but each m in 0..lmax at most once. */
int raysht_plan_create(raysht_plan **plan, int lmax, int ntheta, const int *m_values, int n_m,
                       int nr, int nfields, int precision);
/* The same plan over `ncols` anonymous columns per m instead of 2*nr*nfields: the
 * rmcontainer (2-D) pair of the generic, PtS_2d_dgpv2 / StP_2d_dgp
 * (Legendre_Transforms.F90:506-668), whose blocks are data(m:lmax, 1:ncols).
 * Such a plan has no radial/field structure, so the permutes reject it. */
int raysht_plan_create_cols(raysht_plan **plan, int lmax, int ntheta, const int *m_values, int n_m, int ncols,
                            int precision);
void raysht_plan_destroy(raysht_plan *plan);

/* Sizes (in doubles) of the buffers a plan transforms. */
size_t raysht_spectral_size(const raysht_plan *plan);
size_t raysht_hybrid_size(const raysht_plan *plan);

/* Legendre_Transform(s2a, p2a) -> StP_4d_dgp2 (Legendre_Transforms.F90:968-1074).
 * s: spectral (device), p2: hybrid (device), fully overwritten. */
int raysht_legendre_stp(raysht_plan *plan, const double *s, double *p2, void *stream);

/* Legendre_Transform(p2b, s2b) -> PtS_4d_dgpv3 (Legendre_Transforms.F90:1077-1175).
 * p2: hybrid (device), s: spectral (device), every l in m..lmax written. */
int raysht_legendre_pts(raysht_plan *plan, const double *p2, double *s, void *stream);

/* Host-array forms of the two transforms for callers whose buffers are host
 * memory (the Fortran rmcontainer4d arrays): s_m[mp] points at block mp
 * (data(m:lmax, ...) of that m, contiguous), p2 at the host hybrid array.  The
 * call stages through device memory held by the plan and returns when the
 * result is in the caller's array. */
int raysht_legendre_stp_host(raysht_plan *plan, const double *const *s_m, double *p2);
int raysht_legendre_pts_host(raysht_plan *plan, const double *p2, double *const *s_m);

/* Host copies of the setup data, for checking.  Nodes/weights as head+tail
 * doubles of the quad values (ntheta each).  Tables in the reference's stored
 * orientation, concatenated in mp order: which = 0 p_lm_even (n_e x nt/2, l
 * fastest), 1 p_lm_odd, 2 ip_lm_even (nt/2 x n_e, theta fastest), 3 ip_lm_odd. */
int raysht_plan_nodes(const raysht_plan *plan, double *x_hi, double *x_lo, double *w_hi, double *w_lo);
/* The nodes and weights as the reference stores them, Real*16 coloc and gl_weights
 * (Legendre_Polynomials.F90:28, :332-373): ntheta IEEE binary128 values each, 16 bytes
 * per value in host byte order (Fortran Real*16 / C __float128 on x86-64). */
int raysht_plan_nodes_quad(const raysht_plan *plan, void *x16, void *w16);
size_t raysht_plan_table_size(const raysht_plan *plan, int which);
int raysht_plan_table(const raysht_plan *plan, int which, double *host_out);

/* FFT_To_Physical = c2r_ip4D_fftw (Fourier_Transform.F90:142-183): in-place,
 * unnormalised complex-to-real transform of `howmany` lines of nphi+2 doubles
 * (Re/Im of m = 0..nphi/2 in, nphi reals out; Im of m=0 and m=nphi/2 ignored). */
int raysht_fft_to_physical(int nphi, long howmany, double *x, void *stream);

/* FFT_To_Spectral = r2c_ip4D_fftw (Fourier_Transform.F90:100-140), rescale = 0;
 * FFT_To_Spectral_RSC = r2c_ip4D_fftw_rsc (:56-98), rescale != 0. */
int raysht_fft_to_spectral(int nphi, long howmany, double *x, int rescale, void *stream);

/* Host-array forms of the two FFTs (staged through device memory; synchronous). */
int raysht_fft_to_physical_host(int nphi, long howmany, double *x);
int raysht_fft_to_spectral_host(int nphi, long howmany, double *x, int rescale);

/* SphericalBuffer%reform p2a -> p3a = Transpose_2a3a (Spherical_Buffer.F90:196-459)
 * on one rank: p3 zero-filled, then Re/Im of each m in m_values placed at
 * 2m, 2m+1 of every line.  nphi >= 2*(lmax+1). */
int raysht_permute_2a3a(raysht_plan *plan, int nphi, const double *p2, double *p3, void *stream);

/* SphericalBuffer%reform p3b -> p2b = Transpose_3b2b (Spherical_Buffer.F90:461-703)
 * on one rank (the kill-signal cargo is not carried). */
int raysht_permute_3b2b(raysht_plan *plan, int nphi, const double *p3, double *p2, void *stream);
/* Transpose_2b1b (Spherical_Buffer.F90:705-843) on one radial rank: p1b(r, imi, lm, f)
 * ([f][lm][imi][r], r fastest; spectral_size doubles) = s2b(mp)%data(l, r, imi, f) for
 * the lm-th (l, m) pair of LM_Load_Balance (Load_Balance.F90:311-395): l slowest from the
 * smallest m of the plan, the plan's m in m_values order within one l.  Device arrays,
 * out of place; plans made by raysht_plan_create_cols: RAYSHT_EUNSUPPORTED. */
int raysht_permute_2b1b(raysht_plan *plan, const double *s2b, double *p1b, void *stream);
/* Transpose_1a2a (:845-932) on one radial rank: the inverse, s2a(mp)%data(l, r, imi, f) =
 * p1a(r, imi, lm, f). */
int raysht_permute_1a2a(raysht_plan *plan, const double *p1a, double *s2a, void *stream);
/* Host-array forms of the four permutes (synchronous), what the Fortran Reform_2a3a /
 * Reform_3b2b / Reform_2b1b / Reform_1a2a call: p2 / p3 / p1 host arrays in the layouts
 * above, s_m[mp] the host spectral block of each m (as for raysht_legendre_stp_host). */
int raysht_permute_2a3a_host(raysht_plan *plan, int nphi, const double *p2, double *p3);
int raysht_permute_3b2b_host(raysht_plan *plan, int nphi, const double *p3, double *p2);
int raysht_permute_2b1b_host(raysht_plan *plan, const double *const *s_m, double *p1);
int raysht_permute_1a2a_host(raysht_plan *plan, const double *p1, double *const *s_m);
/* The pair list the radial transposes follow: its length (lm_count), and l_lm_values /
 * mp_lm_values (mp 0-based) into host arrays of lm_count ints. */
int raysht_lm_count(const raysht_plan *plan);
int raysht_lm_values(const raysht_plan *plan, int *l_out, int *mp_out);

/* Spectral_Derivatives.F90 on the spectral layout of a plan made by raysht_plan_create:
 * `in` and `out` are packed spectral arrays of nf_in / nf_out fields (field f of block mp =
 * data(m:l_max, 1:N_r, 1:2, f)); field fin of `in` goes to field fout of `out` (0-based).
 * kind 0: d_by_dtheta = sin(theta) d/dtheta (d_dtheta_buffer4D :224-249, d_dtheta_4dbuff2arr
 * :329-350); kind 1: d_by_sdtheta = 1/sin(theta) d(sin^2(theta) .)/dtheta (:277-298,
 * :381-401), with Initialize_Angular_Derivatives' coefficients (:49-83).  The same field of
 * the same array (in place) is RAYSHT_EINVAL for these; plans over anonymous columns are
 * RAYSHT_EUNSUPPORTED. */
int raysht_d_by_dtheta(raysht_plan *plan, int kind, const double *in, int nf_in, int fin, double *out, int nf_out,
                       int fout, void *stream);
/* d_by_dphi (d_by_dphi_rlmbuff4d :573-603, d_by_dphi_4dbuff2arr :495-509): i m times field fin,
 * (re, im) -> (-m im, m re); in place (same array, fin = fout) allowed. */
int raysht_d_by_dphi(raysht_plan *plan, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                     void *stream);
/* Host-array form of the three (synchronous), what the Fortran d_by_dtheta / d_by_sdtheta /
 * d_by_dphi generics call: in_m[mp] / out_m[mp] point at each m-block's data(m:l_max, 1:N_r,
 * 1:2, 1:nf) (kind 0 dtheta, 1 sdtheta, 2 dphi; 0-based fields); the same pointers for in
 * and out (same nf) is the buffer form, other fields of a separate out are kept. */
int raysht_angular_deriv_host(raysht_plan *plan, int kind, const double *const *in_m, int nf_in, int fin,
                              double *const *out_m, int nf_out, int fout);

/* rlm_spaceb's hydro part (Physics/Sphere_Hybrid_Space.F90:127-214, without adjust_emf) in
 * one pass, in place on the packed spectral buffer `buf` of nf fields after PtS, with wvar,
 * pvar, zvar the 0-based W, P (dW/dr) and Z fields (distinct):
 *   W <- W over_l_l_plus1(l);
 *   Z <- (d_by_dphi P - d_by_sdtheta Z) over_l_l_plus1(l);
 *   P <- (d_by_sdtheta P + d_by_dphi Z) over_l_l_plus1(l)      (P and Z as they came in);
 * l = 0 of W, P, Z set to 0 for m = 0, and l = l_max of every field set to 0.
 * over_l_l_plus1(l) = 1 / (l (l + 1)), 1 at l = 0 (ProblemSize.F90:506-513).  Results
 * are the doubles the reference's sequence of d_by_sdtheta / d_by_dphi calls and array
 * statements gives. */
int raysht_rlm_spaceb_hydro(raysht_plan *plan, double *buf, int nf, int wvar, int pvar, int zvar, void *stream);
/* rlm_spacea's Velocity_Components and Velocity_Derivatives, then its d_by_dtheta(tvar, dtdt)
 * (Physics/Sphere_Hybrid_Space.F90:80,235-355), in one pass, in place on the packed spectral
 * buffer `buf` of nf fields before StP.  fields (host array of 10 distinct 0-based fields) =
 * W, dWdr, d2Wdr2, Z, dZdr, T, vtheta, vphi, dvtdr, dvpdr; vr, dvrdr, dtdt and dvrdt are written
 * over W, dWdr, d2Wdr2 and dZdr (Fields.F90:323-326).  Radial profiles (device arrays of N_r
 * values, Hybrid_Init :45-62 and ProblemSize.F90:624): over_rhor, over_rhorsq, drho_term,
 * two_over_r (Two_Over_R), dlnrho (ref%dlnrho), dsdr_over_cp (ref%dsdr_over_cp; NULL when not
 * pseudo-incompressible).  The l_max zeroing around it in rlm_spacea (:69-72, :95-98) is
 * raysht_zero_lmax.  Results are the doubles of the reference's sequence of d_by_dtheta /
 * d_by_dphi calls and array statements. */
int raysht_rlm_spacea_hydro(raysht_plan *plan, double *buf, int nf, const int *fields, const double *over_rhor,
                            const double *over_rhorsq, const double *drho_term, const double *two_over_r,
                            const double *dlnrho, const double *dsdr_over_cp, void *stream);
/* Compute_BandCurlB (Physics/Sphere_Hybrid_Space.F90:377-460) in one pass, in place on `buf`
 * (nf fields).  fields (host array of 6 distinct 0-based fields) = C, A, dCdr, dAdr, d2Cdr2,
 * curlbr; Br, Btheta, Bphi, curlbtheta and curlbphi are written over C, A, dCdr, d2Cdr2 and
 * dAdr (Fields.F90:329-334).  one_over_rsq = OneOverRSquared (device, N_r values).  Results
 * are the doubles of the reference's sequence. */
int raysht_compute_bandcurlb(raysht_plan *plan, double *buf, int nf, const int *fields, const double *one_over_rsq,
                             void *stream);
/* rlm_spaceb's Adjust_Emf (Physics/Sphere_Hybrid_Space.F90:557-602) in one pass, in place on `buf`
 * (nf fields) after PtS: emftheta <- (d_by_sdtheta emfphi - d_by_dphi emftheta) over_l_l_plus1(l)
 * (Cvar is the emftheta field, Fields.F90:363-364), emfphi <- (d_by_dphi emfphi + d_by_sdtheta
 * emftheta) over_l_l_plus1(l), with emftheta as it came in; emftheta, emfphi 0-based and distinct.
 * It reads and writes only those two fields, so calling it before raysht_rlm_spaceb_hydro (whose
 * last step zeroes l_max) gives rlm_spaceb's magnetic order.  Results are the reference's doubles. */
int raysht_adjust_emf(raysht_plan *plan, double *buf, int nf, int emftheta, int emfphi, void *stream);
/* Hybrid_Output_Initial's magnetic part (Physics/Sphere_Hybrid_Space.F90:525-541) in one pass
 * from `s` (nf fields) into the output co-buffer `co` (nf_co fields) and `tmp` (2 fields, the
 * reference's ftemp3 and ftemp4): co[avar_cb] = (l(l+1) s[avar]) one_over_rsq, co[dbrdr_cb] =
 * s[dcdr], tmp[0] = s[d2cdr2], tmp[1] = s[dadr].  fields (host array of 6 0-based fields) =
 * avar, dcdr, d2cdr2, dadr (distinct, of s), avar_cb, dbrdr_cb (distinct, of co); s, co and tmp
 * must not overlap.  one_over_rsq = OneOverRSquared (device, N_r values).  Results are the
 * reference's doubles. */
int raysht_hybrid_output_initial(raysht_plan *plan, const double *s, int nf, double *co, int nf_co, double *tmp,
                                 const int *fields, const double *one_over_rsq, void *stream);
/* Hybrid_Output_Final (Physics/Sphere_Hybrid_Space.F90:543-552) up to its Legendre transform,
 * in one pass: every co field's l_max zeroed, then Hydro_Output_Derivatives (:224-233)
 * co[dpdt_cb] = d_by_dtheta s[pvar] and, when `tmp` (raysht_hybrid_output_initial's work fields)
 * is not NULL, BField_Derivatives (:437-498):
 *   co[dbtdr_cb] = (d_by_dtheta tmp[0] + d_by_dphi tmp[1]) one_over_r - s[btheta] one_over_rsq,
 *   co[dbpdr_cb] = (d_by_dphi tmp[0] - d_by_dtheta tmp[1]) one_over_r - s[bphi] one_over_rsq,
 *   co[dbrdr_cb] = (l(l+1) co[dbrdr_cb]) one_over_rsq - s[br] two_over_r,
 *   co[dbrdt_cb] = d_by_dtheta s[br].
 * fields (host array of 9 0-based fields) = pvar, br, btheta, bphi (of s), dpdt_cb, dbtdr_cb,
 * dbpdr_cb, dbrdr_cb, dbrdt_cb (of co); with tmp NULL only pvar and dpdt_cb are read and the
 * profiles may be NULL, otherwise each group is distinct and the profiles (device, N_r values:
 * One_Over_R, OneOverRSquared, Two_Over_R) are required.  s must not overlap co or tmp.
 * Results are the doubles of the reference's sequence. */
int raysht_hybrid_output_final(raysht_plan *plan, const double *s, int nf, double *co, int nf_co, const double *tmp,
                               const int *fields, const double *one_over_r, const double *one_over_rsq,
                               const double *two_over_r, void *stream);
/* The l(l+1) and radial scalings of rlm_spacea / rlm_spaceb (Velocity_Components,
 * Velocity_Derivatives, Compute_BandCurlB: Sphere_Hybrid_Space.F90:244,324,353,367,373) on one
 * field: out(l, r, imi) = (lfac(l) in(l, r, imi)) rfac(r), rounded in that order; lfac
 * (l_max + 1 values) and rfac (N_r values) are device arrays, either may be NULL (factor 1).
 * In place, or between fields of one buffer, allowed. */
int raysht_lm_scale(raysht_plan *plan, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                    const double *lfac, const double *rfac, void *stream);
/* "Zero out l_max mode" of rlm_spacea / rlm_spaceb (Sphere_Hybrid_Space.F90:70-72,102-104,
 * 207-209): data(l_max, :, :, :) = 0 for every m-block of a buffer of nf fields. */
int raysht_zero_lmax(raysht_plan *plan, double *buf, int nf, void *stream);

/* Implicit_Solve's dense path (Linear_Solve.F90:423-457) for many matrices at once.
 * raysht_lu_decompose_full = LU_Decompose_full (:1034-1045, LAPACK dgetrf) on nbatch n x n
 * column-major matrices stored one after another in `mats` (overwritten by L\U), pivots
 * (nbatch x n ints, 1-based, as dgetrf returns them) and info (nbatch ints: 0, or the
 * first zero pivot's 1-based index).  Device arrays; n <= 1024. */
int raysht_lu_decompose_full(int n, int nbatch, double *mats, int *pivots, int *info, void *stream);
/* LU_Solve_full (:1087-1107, dgetrs 'N') for every matrix b of the batch on its own
 * right-hand sides: nrhs[b] columns of n (column-major, leading dimension n) at
 * rhs + rhs_off[b] (rhs_off, nrhs: device arrays), overwritten by the solutions;
 * max_nrhs >= every nrhs[b] sizes the launch. */
int raysht_lu_solve_full(int n, int nbatch, const double *mats, const int *pivots, double *rhs, const long *rhs_off,
                         const int *nrhs, int max_nrhs, void *stream);
/* Implicit_Solve's band path (band_solve, Linear_Solve.F90:423-457):
 * raysht_lu_decompose_band = LU_Decompose_band (:1041-1055, LAPACK dgbtrf; the reference passes
 * kl = ku = (ldab - 1) / 3) on nbatch n x n band matrices in LAPACK band storage, ldab x n
 * column-major each (ldab >= 2 kl + ku + 1; rows 1..kl are fill-in space), stored one after
 * another; pivots (nbatch x n ints, 1-based) and info as for the dense path.  kl + ku <= 1024.
 * raysht_lu_solve_band = LU_Solve_Band (:1056-1086, dgbtrs 'N'), right-hand sides as for
 * raysht_lu_solve_full.  The interleave of linked variables (Band_Arrange_RHS) stays with the
 * caller. */
int raysht_lu_decompose_band(int n, int kl, int ku, int nbatch, double *mats, int ldab, int *pivots, int *info,
                             void *stream);
int raysht_lu_solve_band(int n, int kl, int ku, int nbatch, const double *mats, int ldab, const int *pivots,
                         double *rhs, const long *rhs_off, const int *nrhs, int max_nrhs, void *stream);
/* Host-array forms for one matrix (synchronous), what the Fortran LU_Decompose_full /
 * LU_Solve_full / LU_Decompose_band / LU_Solve_Band call: kl < 0 selects the dense path (mat
 * n x n, leading dimension lda >= n), kl >= 0 the band path (mat in band storage, lda = ldab);
 * rhs holds nrhs columns with leading dimension ldb >= n. */
int raysht_lu_decompose_host(int n, int kl, int ku, double *mat, int lda, int *pvt, int *info);
int raysht_lu_solve_host(int n, int kl, int ku, const double *mat, int lda, const int *pvt, double *rhs, int ldb,
                         int nrhs);

/* The hybrid <-> physical reform and the phi FFT in one call (SphericalBuffer%reform,
 * Spherical_Buffer.F90:196-703, with FFT_To_Physical / FFT_To_Spectral,
 * Fourier_Transform.F90:56-183): raysht_reform_to_physical = raysht_permute_2a3a then
 * raysht_fft_to_physical on the plan's lines, raysht_reform_to_spectral = raysht_fft_to_spectral
 * (unscaled) then raysht_permute_3b2b; results identical to the separate calls. */
int raysht_reform_to_physical(raysht_plan *plan, int nphi, const double *p2, double *p3, void *stream);
int raysht_reform_to_spectral(raysht_plan *plan, int nphi, double *p3, double *p2, void *stream);
/* Fused forms of the same steps (results identical to the separate calls):
 * StP then 2a->3a then c2r, and r2c then 3b->2b then PtS.  `work` is a hybrid
 * buffer the call may overwrite.  For nphi / 2 = 3 * 2^k (24 .. 3072, Rayleigh's
 * nphi = 3 (l_max + 1)) the 3b->2b transpose runs inside the r2c kernel (the padded grid
 * is read once); p3 is then left unchanged by raysht_reform_to_spectral /
 * raysht_physical_to_spectral, which otherwise transform it in place. */
int raysht_spectral_to_physical(raysht_plan *plan, int nphi, const double *s, double *work,
                                double *p3, void *stream);
int raysht_physical_to_spectral(raysht_plan *plan, int nphi, double *p3, double *work,
                                double *s, void *stream);
/* raysht_spectral_to_physical without a `work` argument and without any device
 * allocation: no counterpart in the reference; for callers that cannot allocate device
 * memory (code under hipGraph stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The hybrid intermediate lives in a library-owned static device array of
 * RAYSHT_SCRATCH_BYTES per GPU (held until the process ends); plans whose hybrid array
 * (raysht_hybrid_size doubles) is larger are refused with RAYSHT_EINVAL before anything
 * is launched.  The FFT twiddle table of nphi must already exist on the device (any
 * earlier FFT, round-trip or graph call of that nphi makes it); otherwise
 * RAYSHT_ENOTREADY.  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (other streams, other host threads) race; p3 and the plan's
 * outputs are as raysht_spectral_to_physical's, bit for bit. */
#define RAYSHT_SCRATCH_BYTES (32u << 20)
int raysht_spectral_to_physical_scratch(raysht_plan *plan, int nphi, const double *s, double *p3,
                                        void *stream);

/* The spectral -> physical -> spectral round trip (StP, 2a->3a, c2r, r2c, 3b->2b,
 * PtS: Sphere_Hybrid_Space.F90:111,143 / Sphere_Physical_Space.F90:115,215 around an
 * identity physical step) recorded once into a hipGraph on the given device buffers;
 * each launch replays all six kernels with one call.  The plan and buffers must
 * outlive the graph.  No reference counterpart: the reference launches each step
 * from the host every time step. */
typedef struct raysht_graph raysht_graph;
int raysht_graph_create_round_trip(raysht_graph **graph, raysht_plan *plan, int nphi, const double *s_in,
                                   double *work, double *p3, double *s_out);
int raysht_graph_launch(raysht_graph *graph, void *stream);
void raysht_graph_destroy(raysht_graph *graph);

/* Many independent Legendre problems in one grouped launch per direction: up to 64
 * plans of one precision (any l_max, N_theta, m set and column count) whose m-blocks
 * are scheduled together, longest contraction first, over the XCDs of the device.
 * raysht_batch_stp(b, s, p2, stream) does what raysht_legendre_stp(plans[i], s[i],
 * p2[i], stream) does for every i, in one kernel launch (results identical); likewise
 * raysht_batch_pts.  s / p2 are host arrays of n_plans device pointers.  The plans
 * must outlive the batch.  No reference counterpart: the reference transforms one
 * problem per call (Legendre_Transforms.F90:968,1077). */
typedef struct raysht_batch raysht_batch;
int raysht_batch_create(raysht_batch **batch, raysht_plan *const *plans, int n_plans);
void raysht_batch_destroy(raysht_batch *batch);
int raysht_batch_size(const raysht_batch *batch);
int raysht_batch_stp(raysht_batch *batch, const double *const *s, double *const *p2, void *stream);
int raysht_batch_pts(raysht_batch *batch, const double *const *p2, double *const *s, void *stream);
/* Schedule autotuner of a batch.  Candidates (index bits): bit 0 = no reduced-height
 * (3/4, 1/2, 1/4) last row tiles in the FP64 PtS kernel, bit 1 = no XCD grouping of m-blocks (items go out
 * longest first and the dispatcher spreads them over the XCDs); 0 is the default.
 * raysht_batch_autotune times every distinct candidate of each direction on the given
 * buffers (one warm-up and `reps` timed grouped launches on `stream`; outputs are
 * overwritten) and keeps the fastest; ms_out (optional, 8 doubles: [direction][candidate],
 * direction 0 = StP, 1 = PtS) receives the mean times, -1 where not timed.  Results of a
 * batch do not depend on its schedule.  No reference counterpart. */
int raysht_batch_autotune(raysht_batch *batch, const double *const *s, double *const *p2,
                          double *const *s_out, int reps, void *stream, double *ms_out);
int raysht_batch_schedule(const raysht_batch *batch, int direction);
int raysht_batch_set_schedule(raysht_batch *batch, int direction, int candidate);
/* The batch's Legendre round trip (grouped StP into p2[i], grouped PtS into s_out[i])
 * recorded into a hipGraph, launched with raysht_graph_launch.  No reference
 * counterpart. */
int raysht_graph_create_batch(raysht_graph **graph, raysht_batch *batch, const double *const *s_in,
                              double *const *p2, double *const *s_out);

/* Cheby_Grid%init (Chebyshev_Polynomials.F90:65-207) for one domain of n
 * zeros-grid points on [rmin, rmax]; grid (host, n) receives the radii. */
int raysht_cheb_create(raysht_cheb **cheb, int n, double rmin, double rmax, double *grid);
/* Cheby_Grid%init with ndomains domains (:65-207; ProblemSize.F90's ncheby / domain_bounds):
 * npoly[ndomains] points and bounds[ndomains + 1] in ascending order of radius, as the
 * reference takes them; dealias_by (NULL, or ndomains values in the same order) as Init's
 * optional argument (rda = npoly - dealias_by + 1 where 1 <= dealias_by < npoly).  Like the
 * reference, the grid holds the domains outermost first: a column is the total of npoly
 * values, domain h (from the outside) on rows hoff .. hoff + npoly - 1, and grid (host,
 * total) receives the radii in that order.  Every transform, derivative and dealias call
 * on such a grid works domain by domain as the reference's loops over domains do; with
 * several domains each has at most 256 points and the table form runs (RAYSHT_CHEB_DCT is
 * RAYSHT_EUNSUPPORTED, as is raysht_cheb_table).  ndomains = 1 is raysht_cheb_create. */
int raysht_cheb_create_domains(raysht_cheb **cheb, int ndomains, const int *npoly, const double *bounds,
                               const int *dealias_by, double *grid);
void raysht_cheb_destroy(raysht_cheb *cheb);
/* The rest of what Cheby_Grid%init computes for its caller and its components
 * (:161-200, :210-248), n values each (n = the total of every domain's points), any pointer
 * may be NULL: x = cos(theta) the colocation points (descending within each domain), theta
 * their angles, weights the radial integration weights (Init's integration_weights
 * argument), deriv_scaling the d/dr factor of the domain map at every point
 * (Cheby_Grid%deriv_scaling). */
int raysht_cheb_grid_data(const raysht_cheb *cheb, double *x, double *theta, double *weights,
                          double *deriv_scaling);
/* Gen_Tn tables (:251-297) as stored: which = 0 cheby_even (n_x x n_even), 1 cheby_odd. */
int raysht_cheb_table(const raysht_cheb *cheb, int which, double *host_out);

/* Cheby_Grid%to_spectral = To_Spectral4D (:375-448) over ncols columns of n. */
int raysht_cheb_to_spectral(raysht_cheb *cheb, const double *f, double *c, long ncols, void *stream);
/* Cheby_Grid%from_spectral = From_Spectral4D (:450-558). */
int raysht_cheb_from_spectral(raysht_cheb *cheb, const double *c, double *f, long ncols, void *stream);
/* How the two transforms above are computed.  RAYSHT_CHEB_DCT (available where n is a
 * power of two, 16 <= n <= 256; the default for n = 128 and 256, where it is the faster):
 * a DCT-II / DCT-III through one complex FFT of length n
 * per column, which also carries the first-order correction for the reference's angles
 * (its tables are cos(k theta_i) at theta_i accumulated by repeated addition, up to
 * 8e-15 away from (i + 1/2) pi / n at n = 256); within 1e-14 of the table form and
 * closer than it to the exact sums on those angles.  RAYSHT_CHEB_MATRIX: the
 * parity-split products with the Gen_Tn tables, the reference's own form.  Columns must
 * start 16-byte aligned for the DCT form; other pointers take the table form.  Setting
 * RAYSHT_CHEB_DCT for an n without it returns RAYSHT_EUNSUPPORTED.  No counterpart in
 * the reference (a choice of algorithm for Cheby_Grid%to_spectral / %from_spectral). */
enum raysht_cheb_form { RAYSHT_CHEB_DCT = 0, RAYSHT_CHEB_MATRIX = 1 };
int raysht_cheb_set_form(raysht_cheb *cheb, int form);
/* Host-array forms of the two Chebyshev transforms (synchronous). */
int raysht_cheb_to_spectral_host(raysht_cheb *cheb, const double *f, double *c, long ncols);
int raysht_cheb_from_spectral_host(raysht_cheb *cheb, const double *c, double *f, long ncols);
/* Cheby_Grid%dealias_buffer = dealias_buffer4d (:353-374): zero n >= 2n/3 (0-based). */
int raysht_cheb_dealias(raysht_cheb *cheb, double *c, long ncols, void *stream);
/* Cheby_Grid%d_by_dr_cp = Cheby_Deriv_Buffer_4D (:560-640) with ind /= dind, and
 * Cheby_Grid%d_by_dr_nobuffer = Cheby_Deriv_OutOf_Buffer (:642-730): d receives the
 * coefficients of the dorder-th radial derivative of every column of c (ncols columns
 * of n, device arrays).  Out of place only (overlapping c and d: RAYSHT_EINVAL; the
 * reference's in-place branch only rescales); dorder 1..4 (larger: RAYSHT_EUNSUPPORTED). */
int raysht_cheb_deriv(raysht_cheb *cheb, const double *c, double *d, long ncols, int dorder, void *stream);
/* Host-array form of raysht_cheb_deriv (synchronous); what the Fortran d_by_dr_cp calls. */
int raysht_cheb_deriv_host(raysht_cheb *cheb, const double *c, double *d, long ncols, int dorder);

#ifdef __cplusplus
}
#endif
#endif
