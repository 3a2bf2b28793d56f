/*
 * raysht oracle -- CPU restatement of Rayleigh's spectral-transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker.
 * The product (libraysht, rayleigh_amd) never links, loads or calls it.
 *
 * Every routine names the reference routine (file:line under the Rayleigh
 * tree) whose steps it restates.  Layouts follow the reference's storage:
 *
 *   spectral  s, packed over mp = 0..nm-1 in m_values order; block mp is
 *             data(m:lmax, 1:nr, 1:2, 1:nf), l fastest (Spherical_Buffer.F90:1459-1484)
 *   hybrid    p2[mp][c][theta], c = r + imi*nr + f*2*nr   (Spherical_Buffer.F90:1391-1406)
 *   physical  p3[f][theta][r][nphi+2]                     (Spherical_Buffer.F90:1424-1440)
 *   tables    per mp: p_even[k][j] (k = parity-l index, j = 0..nt/2-1) as
 *             p_lm_even(mp)%data(1:n_e,1:nt2); ip_even[j][k] as
 *             ip_lm_even(mp)%data(1:nt2,1:n_e); odd twins likewise.
 *
 * Parity status of each routine (what pins it) is in the header comment of
 * each definition in raysht_oracle.c.
 */
#ifndef RAYSHT_ORACLE_H
#define RAYSHT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Gauss-Legendre nodes (ascending from -1) and weights, as head/tail doubles. */
void orc_gauss_legendre(int n, double *xhi, double *xlo, double *whi, double *wlo);

/* Number of doubles in the packed even / odd parity tables. */
long orc_table_size_even(int nt, int lmax, const int *mvals, int nm);
long orc_table_size_odd(int nt, int lmax, const int *mvals, int nm);

/* The four parity tables for every m in mvals; returns 0 on success. */
int orc_legendre_tables(int nt, int lmax, const int *mvals, int nm,
                        double *p_even, double *p_odd, double *ip_even, double *ip_odd);

/* Spectral -> hybrid (StP_4d_dgp2). */
void orc_stp(int nt, int lmax, const int *mvals, int nm, int nr, int nf,
             const double *p_even, const double *p_odd, const double *s_in, double *p_out);

/* The same over ncols anonymous columns (StP_2d_dgp). */
void orc_stp_cols(int nt, int lmax, const int *mvals, int nm, int ncols,
                  const double *p_even, const double *p_odd, const double *s_in, double *p_out);

/* Hybrid -> spectral (PtS_4d_dgpv3). */
void orc_pts(int nt, int lmax, const int *mvals, int nm, int nr, int nf,
             const double *ip_even, const double *ip_odd, const double *p_in, double *s_out);

/* The same over ncols anonymous columns (PtS_2d_dgpv2). */
void orc_pts_cols(int nt, int lmax, const int *mvals, int nm, int ncols,
                  const double *ip_even, const double *ip_odd, const double *p_in, double *s_out);

/* In-place real FFTs along lines of length nphi+2 (FFTW half-complex layout). */
void orc_fft_c2r(int nphi, long howmany, double *x);
void orc_fft_r2c(int nphi, long howmany, double *x, int rescale);

/* Row transposes on one rank. */
void orc_permute_2a3a(int nt, int lmax, const int *mvals, int nm, int nr, int nf, int nphi,
                      const double *p2a, double *p3a);
void orc_angular_coefs(int l, int m, int which, double *c);
void orc_angular_dtheta(int which, int lmax, const int *mvals, int nm, int nr, const double *in, int nf_in, int fin,
                        double *out, int nf_out, int fout);
void orc_angular_dphi(int lmax, const int *mvals, int nm, int nr, const double *in, int nf_in, int fin, double *out,
                      int nf_out, int fout);
void orc_rlm_spaceb(int lmax, const int *mvals, int nm, int nr, double *buf, int nf, int wv, int pv, int zv);
void orc_lm_scale(int lmax, const int *mvals, int nm, int nr, const double *in, int nf_in, int fin, double *out,
                  int nf_out, int fout, const double *lfac, const double *rfac);
void orc_zero_lmax(int lmax, const int *mvals, int nm, int nr, double *buf, int nf);
void orc_rlm_spacea_hydro(int lmax, const int *mvals, int nm, int nr, double *buf, int nf, const int *v,
                          const double *over_rhor, const double *over_rhorsq, const double *drho_term,
                          const double *two_over_r, const double *dlnrho, const double *dsdr_over_cp);
void orc_adjust_emf(int lmax, const int *mvals, int nm, int nr, double *buf, int nf, int et, int ep);
void orc_hybrid_output_initial(int lmax, const int *mvals, int nm, int nr, const double *s, int nf, double *co,
                               int nf_co, double *tmp, const int *v, const double *one_over_rsq);
void orc_hybrid_output_final(int lmax, const int *mvals, int nm, int nr, const double *s, int nf, double *co,
                             int nf_co, const double *tmp, const int *v, const double *one_over_r,
                             const double *one_over_rsq, const double *two_over_r);
void orc_compute_bandcurlb(int lmax, const int *mvals, int nm, int nr, double *buf, int nf, const int *v,
                           const double *one_over_rsq);
int orc_lu_decompose(int n, double *a, int *piv);
void orc_lu_solve(int n, const double *a, const int *piv, int nrhs, double *b);
int orc_band_lu_decompose(int n, int kl, int ku, double *ab, int ldab, int *piv);
void orc_band_lu_solve(int n, int kl, int ku, const double *ab, int ldab, const int *piv, int nrhs, double *b);
int orc_lm_values(int lmax, const int *mvals, int nm, int *l_out, int *mp_out);
void orc_permute_2b1b(int lmax, const int *mvals, int nm, int nr, int nf, const double *s2b, double *p1b);
void orc_permute_1a2a(int lmax, const int *mvals, int nm, int nr, int nf, const double *p1a, double *s2a);
void orc_permute_3b2b(int nt, int lmax, const int *mvals, int nm, int nr, int nf, int nphi,
                      const double *p3b, double *p2b);

/* One-domain Chebyshev zeros grid: grid on [rmin,rmax] (descending), tables. */
void orc_cheby_grid(int n, double rmin, double rmax, double *grid);
void orc_cheby_tables(int n, double *ceven, double *codd);
void orc_cheby_weights_domain(int n, double lowerb, double upperb, double gmin, double gmax, double *w);
void orc_cheby_weights(int n, double rmin, double rmax, double *w);
void orc_gauss_legendre_binary128(int n, void *x16, void *w16);
void orc_cheby_to_spectral(int n, long ncols, const double *ceven, const double *codd,
                           const double *f, double *c);
void orc_cheby_from_spectral(int n, long ncols, const double *ceven, const double *codd,
                             const double *c, double *f);
void orc_cheby_dealias(int n, long ncols, double *c);
double orc_cheby_deriv_scaling(int n, double rmin, double rmax);
void orc_cheby_deriv(int n, double ds, int dorder, long ncols, const double *c, double *d, int fused);

/* Balanced single-rank m order of M_Balance_v0: 0, L, 1, L-1, ... */
void orc_m_order_balanced(int lmax, int *mvals);

#ifdef __cplusplus
}
#endif
#endif
