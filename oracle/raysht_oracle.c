/*
 * raysht oracle -- CPU restatement of Rayleigh's spectral-transform path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg compare the HIP library against.  It is
 * never linked into, loaded by or called from the product.
 *
 * Each function below restates one reference routine (file:line under the
 * Rayleigh tree) step by step:
 *   - Gauss-Legendre nodes/weights and the P_lm tables run in __float128,
 *     as the reference runs them in Real*16, with the same operation order,
 *     so the stored doubles are the reference's quad-then-round values.
 *   - StP / PtS / Chebyshev contractions run as plain loops in the order the
 *     reference's DGEMM calls define (inner sum over k ascending).
 *   - The FFT (FFTW is not part of the reference tree) is an exact-as-possible
 *     restatement of FFTW's unnormalised r2c / c2r: a mixed-radix DFT in
 *     long double, rounded to double once.
 *
 * What pins each function (see tests/test_oracle_*.py):
 *   orc_gauss_legendre   compiled reference (bit-exact), numpy leggauss, sum w = 2, symmetry
 *   orc_legendre_tables  compiled reference (bit-exact), closed-form Y_lm, discrete orthogonality
 *   orc_stp / orc_pts    compiled reference + fixtures, Test_SHT known answers, round-trip identity
 *   orc_fft_*            numpy.fft (pocketfft), brute-force long-double DFT, Parseval
 *   orc_permute_*        exhaustive index maps on small grids, zero-fill / truncation invariants
 *   orc_cheby_*          compiled reference (bit-exact tables), scipy DCT-II/III, round trip
 */
#include "raysht_oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

typedef __float128 quad;

/* compiler runtime integer power, the routine the reference's `tmp**(mv/2)` lowers to */
extern quad __powitf2(quad a, int b);

static const quad PI_Q = 3.1415926535897932384626433832795028841972Q;

static void split_quad(quad q, double *hi, double *lo)
{
    double h = (double)q;
    *hi = h;
    *lo = (double)(q - (quad)h);
}

/* ------------------------------------------------------------------------ */
/* Legendre_Polynomials.F90:375-393  nth_legendre                            */
static void nth_legendre(quad x, int n, quad *pn_out, quad *dpn_out)
{
    quad pn = 1.0Q, pn1 = 0.0Q, pn2;
    for (int j = 1; j <= n; ++j) {
        pn2 = pn1;
        pn1 = pn;
        pn = ((2.0Q * (quad)j - 1.0Q) * x * pn1 - ((quad)j - 1.0Q) * pn2) / (quad)j;
    }
    *pn_out = pn;
    *dpn_out = (quad)n * (x * pn - pn1) / (x * x - 1.0Q);
}

/* Legendre_Polynomials.F90:332-373  Find_Colocation on [-1, 1]:
 * Newton from cos(pi(i-1/4)/(n+1/2)) until |step| <= 3e-14; the weight of the
 * final iterate uses the derivative evaluated at the previous iterate, as the
 * reference does.  Nodes ascend from -1. */
static void gauss_legendre_q(int n, quad *x, quad *w)
{
    const quad x1 = -1.0Q, x2 = 1.0Q;
    const quad mid = 0.5Q * (x1 + x2), scl = 0.5Q * (x2 - x1);
    const quad eps = 3.0e-14Q;
    int nroots = (n + 1) / 2;
    for (int i = 1; i <= nroots; ++i) {
        quad r = cosq(PI_Q * ((quad)i - 0.25Q) / ((quad)n + 0.5Q));
        int converged = 0;
        while (!converged) {
            quad pn, dpn;
            nth_legendre(r, n, &pn, &dpn);
            quad nr = r - pn / dpn;
            quad del = fabsq(r - nr);
            r = nr;
            if (del <= eps) converged = 1;
            x[i - 1] = mid - scl * r;
            x[n - i] = mid + scl * r;
            w[i - 1] = 2.0Q * scl / ((1.0Q - r * r) * dpn * dpn);
            w[n - i] = w[i - 1];
        }
    }
}

void orc_gauss_legendre(int n, double *xhi, double *xlo, double *whi, double *wlo)
{
    quad *x = malloc(sizeof(quad) * (size_t)n), *w = malloc(sizeof(quad) * (size_t)n);
    gauss_legendre_q(n, x, w);
    for (int i = 0; i < n; ++i) {
        split_quad(x[i], &xhi[i], &xlo[i]);
        split_quad(w[i], &whi[i], &wlo[i]);
    }
    free(x);
    free(w);
}

static int n_even_of(int lmax, int m) { return (lmax - m + 2) / 2; }
static int n_odd_of(int lmax, int m) { return (lmax - m + 1) / 2; }

long orc_table_size_even(int nt, int lmax, const int *mvals, int nm)
{
    long s = 0;
    for (int i = 0; i < nm; ++i) s += (long)n_even_of(lmax, mvals[i]) * (nt / 2);
    return s;
}

long orc_table_size_odd(int nt, int lmax, const int *mvals, int nm)
{
    long s = 0;
    for (int i = 0; i < nm; ++i) s += (long)n_odd_of(lmax, mvals[i]) * (nt / 2);
    return s;
}

/* Legendre_Polynomials.F90:146-258 Compute_Plms (parity branch) followed by
 * :260-331 Parity_Resort.  P_lm carries the orthonormal spherical-harmonic
 * normalisation with the Condon-Shortley sign; only the nt/2 southern nodes are
 * evaluated.  The per-l recursion coefficients depend on (l, m) only, so they
 * are computed once per l here instead of once per node: the quad operations
 * and therefore the values are the same. */
int orc_legendre_tables(int nt, int lmax, const int *mvals, int nm,
                        double *p_even, double *p_odd, double *ip_even, double *ip_odd)
{
    if (nt <= 0 || (nt % 2) != 0 || lmax < 0) return -1;
    const int nt2 = nt / 2;
    quad *x = malloc(sizeof(quad) * (size_t)nt), *w = malloc(sizeof(quad) * (size_t)nt);
    gauss_legendre_q(nt, x, w);
    quad *plm = malloc(sizeof(quad) * (size_t)nt2 * (size_t)(lmax + 1));
    long oe = 0, oo = 0;
    for (int mp = 0; mp < nm; ++mp) {
        const int mv = mvals[mp];
        if (mv < 0 || mv > lmax) { free(x); free(w); free(plm); return -2; }
        /* compute_factorial_ratio, :406-415 */
        quad ratio = 1.0Q;
        for (int i = 1; i <= mv; ++i) ratio = ratio * powq(((quad)i - 0.5Q) / (quad)i, 0.5Q);
        quad amp = powq(((quad)mv + 0.5Q) / (2.0Q * PI_Q), 0.5Q);
        amp = amp * ratio;
        /* plm[(l-mv)*nt2 + i] */
        for (int i = 0; i < nt2; ++i) {
            quad xi = x[i];
            quad tmp = 1.0Q - xi * xi;
            quad v;
            if (mv % 2 == 1) v = -amp * powq(tmp, (quad)(mv / 2) + 0.5Q);
            else v = amp * __powitf2(tmp, mv / 2);
            plm[i] = v;
            if (mv < lmax) plm[(size_t)nt2 + i] = v * xi * powq(2.0Q * (quad)mv + 3.0Q, 0.5Q);
        }
        for (int l = mv + 2; l <= lmax; ++l) {
            quad a = (quad)((l - 1) * (l - 1) - mv * mv);
            a = a / (4.0Q * (quad)((l - 1) * (l - 1)) - 1.0Q);
            a = powq(a, 0.5Q);
            quad b = (4.0Q * (quad)l * (quad)l - 1.0Q) / (quad)(l * l - mv * mv);
            b = powq(b, 0.5Q);
            quad *pl = plm + (size_t)(l - mv) * nt2;
            const quad *p1 = pl - nt2, *p2 = pl - 2 * nt2;
            for (int i = 0; i < nt2; ++i) {
                quad t = p1[i] * x[i] - a * p2[i];
                pl[i] = t * b;
            }
        }
        /* Parity_Resort: FFT normalisation folded into the tables (:268-276) */
        const quad pts_norm = (mv == 0) ? (quad)(1.0 / (2.0 * (double)nt)) : (quad)(1.0 / (double)nt);
        const quad stp_norm = (mv == 0) ? 1.0Q : 0.5Q;
        const int ne = n_even_of(lmax, mv), no = n_odd_of(lmax, mv);
        for (int k = 0; k < ne; ++k) {
            const quad *pl = plm + (size_t)(2 * k) * nt2;
            for (int i = 0; i < nt2; ++i) {
                quad renorm = 2.0Q * PI_Q * w[i];
                quad t = pl[i] * renorm;
                ip_even[oe + (long)k * nt2 + i] = (double)(t * pts_norm);
                p_even[oe + (long)i * ne + k] = (double)(pl[i] * stp_norm);
            }
        }
        for (int k = 0; k < no; ++k) {
            const quad *pl = plm + (size_t)(2 * k + 1) * nt2;
            for (int i = 0; i < nt2; ++i) {
                quad renorm = 2.0Q * PI_Q * w[i];
                quad t = pl[i] * renorm;
                ip_odd[oo + (long)k * nt2 + i] = (double)(t * pts_norm);
                p_odd[oo + (long)i * no + k] = (double)(pl[i] * stp_norm);
            }
        }
        oe += (long)ne * nt2;
        oo += (long)no * nt2;
    }
    free(plm);
    free(x);
    free(w);
    return 0;
}

/* Legendre_Transforms.F90:968-1074  StP_4d_dgp2 (parity branch):
 * E = p_lm_even^T . s_even, O = p_lm_odd^T . s_odd (each nt2 x nrhs);
 * out(j) = E + O, out(nt+1-j) = E - O; out zeroed first. */
void orc_stp(int nt, int lmax, const int *mvals, int nm, int nr, int nf,
             const double *p_even, const double *p_odd, const double *s_in, double *p_out)
{
    orc_stp_cols(nt, lmax, mvals, nm, 2 * nr * nf, p_even, p_odd, s_in, p_out);
}

/* The same over ncols anonymous columns: StP_2d_dgp (Legendre_Transforms.F90:588-668). */
void orc_stp_cols(int nt, int lmax, const int *mvals, int nm, int ncols,
                  const double *p_even, const double *p_odd, const double *s_in, double *p_out)
{
    const int nt2 = nt / 2, nrhs = ncols;
    long so = 0, oe = 0, oo = 0;
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        const int ne = n_even_of(lmax, m), no = n_odd_of(lmax, m);
        const double *pe = p_even + oe, *po = p_odd + oo;
        double *out = p_out + (long)mp * nrhs * nt;
        for (int c = 0; c < nrhs; ++c) {
            const double *col = s_in + so + (long)c * nl;
            double *oc = out + (long)c * nt;
            for (int j = 0; j < nt2; ++j) {
                double acc = 0.0;
                for (int k = 0; k < ne; ++k) acc += pe[(long)j * ne + k] * col[2 * k];
                oc[j] = acc;
                oc[nt - 1 - j] = acc;
            }
            if (no > 0) {
                for (int j = 0; j < nt2; ++j) {
                    double acc = 0.0;
                    for (int k = 0; k < no; ++k) acc += po[(long)j * no + k] * col[2 * k + 1];
                    oc[j] = oc[j] + acc;
                    oc[nt - 1 - j] = oc[nt - 1 - j] - acc;
                }
            }
        }
        so += (long)nl * nrhs;
        oe += (long)ne * nt2;
        oo += (long)no * nt2;
    }
}

/* Legendre_Transforms.F90:1077-1175  PtS_4d_dgpv3 (parity branch):
 * feven(j) = f(j) + f(nt+1-j), fodd(j) = f(j) - f(nt+1-j), j <= nt2;
 * s_even = ip_lm_even^T . feven, s_odd = ip_lm_odd^T . fodd, scattered to l. */
void orc_pts(int nt, int lmax, const int *mvals, int nm, int nr, int nf,
             const double *ip_even, const double *ip_odd, const double *p_in, double *s_out)
{
    orc_pts_cols(nt, lmax, mvals, nm, 2 * nr * nf, ip_even, ip_odd, p_in, s_out);
}

/* The same over ncols anonymous columns: PtS_2d_dgpv2 (Legendre_Transforms.F90:506-585). */
void orc_pts_cols(int nt, int lmax, const int *mvals, int nm, int ncols,
                  const double *ip_even, const double *ip_odd, const double *p_in, double *s_out)
{
    const int nt2 = nt / 2, nrhs = ncols;
    long so = 0, oe = 0, oo = 0;
    double *fe = malloc(sizeof(double) * (size_t)nt2), *fo = malloc(sizeof(double) * (size_t)nt2);
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        const int ne = n_even_of(lmax, m), no = n_odd_of(lmax, m);
        const double *ie = ip_even + oe, *io = ip_odd + oo;
        const double *in = p_in + (long)mp * nrhs * nt;
        for (int c = 0; c < nrhs; ++c) {
            const double *ic = in + (long)c * nt;
            double *col = s_out + so + (long)c * nl;
            for (int j = 0; j < nt2; ++j) {
                fe[j] = ic[j] + ic[nt - 1 - j];
                fo[j] = ic[j] - ic[nt - 1 - j];
            }
            for (int k = 0; k < ne; ++k) {
                double acc = 0.0;
                for (int j = 0; j < nt2; ++j) acc += ie[(long)k * nt2 + j] * fe[j];
                col[2 * k] = acc;
            }
            for (int k = 0; k < no; ++k) {
                double acc = 0.0;
                for (int j = 0; j < nt2; ++j) acc += io[(long)k * nt2 + j] * fo[j];
                col[2 * k + 1] = acc;
            }
        }
        so += (long)nl * nrhs;
        oe += (long)ne * nt2;
        oo += (long)no * nt2;
    }
    free(fe);
    free(fo);
}

/* ------------------------------------------------------------------------ */
/* FFT.  Fourier_Transform.F90:100-183 call FFTW's r2c (e^{-i}) and c2r (e^{+i}),
 * unnormalised, in place on lines of nphi+2 doubles holding Re/Im of
 * m = 0..nphi/2.  FFTW itself is outside the reference tree; this restates the
 * transform it is documented to compute, in long double. */
typedef struct { long double re, im; } cld;

static void dft_rec(const cld *in, cld *out, int n, int stride, int sign, const cld *tw, int ntw)
{
    /* out[k] = sum_j in[j*stride] * exp(sign * 2 pi i j k / n); tw[t] = exp(sign 2 pi i t / ntw) */
    if (n == 1) { out[0] = in[0]; return; }
    int p = 0;
    if (n % 2 == 0) p = 2;
    else if (n % 3 == 0) p = 3;
    else {
        for (int k = 0; k < n; ++k) {
            long double sr = 0, si = 0;
            for (int j = 0; j < n; ++j) {
                long t = ((long)j * k % n) * (ntw / n);
                long double c = tw[t].re, s = tw[t].im;
                sr += in[(long)j * stride].re * c - in[(long)j * stride].im * s;
                si += in[(long)j * stride].re * s + in[(long)j * stride].im * c;
            }
            out[k].re = sr;
            out[k].im = si;
        }
        return;
    }
    const int q = n / p;
    /* sub-transforms of the p decimated sequences, stored consecutively */
    for (int r = 0; r < p; ++r) dft_rec(in + (long)r * stride, out + (long)r * q, q, stride * p, sign, tw, ntw);
    cld *tmp = malloc(sizeof(cld) * (size_t)n);
    memcpy(tmp, out, sizeof(cld) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        long double sr = 0, si = 0;
        for (int r = 0; r < p; ++r) {
            long t = ((long)r * k % n) * (ntw / n);
            const cld a = tmp[(long)r * q + (k % q)];
            sr += a.re * tw[t].re - a.im * tw[t].im;
            si += a.re * tw[t].im + a.im * tw[t].re;
        }
        out[k].re = sr;
        out[k].im = si;
    }
    free(tmp);
}

static cld *twiddles(int n, int sign)
{
    cld *tw = malloc(sizeof(cld) * (size_t)n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; ++t) {
        long double a = two_pi * (long double)t / (long double)n;
        tw[t].re = cosl(a);
        tw[t].im = (long double)sign * sinl(a);
    }
    return tw;
}

void orc_fft_c2r(int nphi, long howmany, double *x)
{
    const int n = nphi, h = nphi / 2;
    cld *tw = twiddles(n, +1);
    cld *z = malloc(sizeof(cld) * (size_t)n), *y = malloc(sizeof(cld) * (size_t)n);
    for (long b = 0; b < howmany; ++b) {
        double *line = x + b * (long)(nphi + 2);
        /* Hermitian extension; Im of m = 0 and m = n/2 do not enter a c2r */
        z[0].re = line[0];
        z[0].im = 0;
        for (int m = 1; m < h; ++m) {
            z[m].re = line[2 * m];
            z[m].im = line[2 * m + 1];
            z[n - m].re = line[2 * m];
            z[n - m].im = -(long double)line[2 * m + 1];
        }
        if (n % 2 == 0) { z[h].re = line[2 * h]; z[h].im = 0; }
        dft_rec(z, y, n, 1, +1, tw, n);
        for (int k = 0; k < n; ++k) line[k] = (double)y[k].re;
        line[n] = 0.0;
        line[n + 1] = 0.0;
    }
    free(z);
    free(y);
    free(tw);
}

/* Fourier_Transform.F90:100-140 (rescale = 0) and :56-98 (rescale != 0: the
 * _RSC variant divides x(3:) by n and x(1) by 2n, leaving x(2) untouched). */
void orc_fft_r2c(int nphi, long howmany, double *x, int rescale)
{
    const int n = nphi, h = nphi / 2;
    cld *tw = twiddles(n, -1);
    cld *z = malloc(sizeof(cld) * (size_t)n), *y = malloc(sizeof(cld) * (size_t)n);
    for (long b = 0; b < howmany; ++b) {
        double *line = x + b * (long)(nphi + 2);
        for (int k = 0; k < n; ++k) { z[k].re = line[k]; z[k].im = 0; }
        dft_rec(z, y, n, 1, -1, tw, n);
        for (int m = 0; m <= h; ++m) {
            line[2 * m] = (double)y[m].re;
            line[2 * m + 1] = (double)y[m].im;
        }
        if (rescale) {
            for (int i = 2; i < n + 2; ++i) line[i] = line[i] / (double)n;
            line[0] = line[0] / (double)(2 * n);
        }
    }
    free(z);
    free(y);
    free(tw);
}

/* ------------------------------------------------------------------------ */
/* Spherical_Buffer.F90:318-459 Transpose_2a3a_v0 on one rank: p3a is zeroed,
 * then p3a(2m+1, r, theta, f) = p2a(theta, r + 2 nr f, mp) (Re) and
 * p3a(2m+2, ...) = p2a(theta, r + nr + 2 nr f, mp) (Im). */
void orc_permute_2a3a(int nt, int lmax, const int *mvals, int nm, int nr, int nf, int nphi,
                      const double *p2a, double *p3a)
{
    (void)lmax;
    const long ld = nphi + 2, nrhs = 2L * nr * nf;
    memset(p3a, 0, sizeof(double) * (size_t)(ld * nr * nt * nf));
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp];
        for (int f = 0; f < nf; ++f)
            for (int t = 0; t < nt; ++t)
                for (int r = 0; r < nr; ++r) {
                    double *line = p3a + (((long)f * nt + t) * nr + r) * ld;
                    line[2 * m] = p2a[((long)mp * nrhs + r + 2L * nr * f) * nt + t];
                    line[2 * m + 1] = p2a[((long)mp * nrhs + r + nr + 2L * nr * f) * nt + t];
                }
    }
}

/* Spherical_Buffer.F90:583-703 Transpose_3b2b_v0 on one rank: the inverse map,
 * reading back only the m in m_values (truncation of m > lmax). */
void orc_permute_3b2b(int nt, int lmax, const int *mvals, int nm, int nr, int nf, int nphi,
                      const double *p3b, double *p2b)
{
    (void)lmax;
    const long ld = nphi + 2, nrhs = 2L * nr * nf;
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp];
        for (int f = 0; f < nf; ++f)
            for (int t = 0; t < nt; ++t)
                for (int r = 0; r < nr; ++r) {
                    const double *line = p3b + (((long)f * nt + t) * nr + r) * ld;
                    p2b[((long)mp * nrhs + r + 2L * nr * f) * nt + t] = line[2 * m];
                    p2b[((long)mp * nrhs + r + nr + 2L * nr * f) * nt + t] = line[2 * m + 1];
                }
    }
}

/* Spectral_Derivatives.F90:49-83 Initialize_Angular_Derivatives, in its operation order:
 * which = 0 deriv_coefs (sin(theta) d/dtheta), 1 sderiv_coefs (1/sin(theta) d(sin^2 .)/dtheta);
 * c1 multiplies A(l-1), c2 multiplies A(l+1). */
static void angular_coefs(int l, int m, int which, double *c1, double *c2)
{
    double num = (double)((l + m) * (l - m)), denom = (double)((2 * l + 1) * (2 * l - 1));
    double q = sqrt(num / denom);
    *c1 = which == 0 ? (double)(l - 1) * q : (double)(l + 1) * q;
    num = (double)((l + 1 + m) * (l + 1 - m));
    denom = (double)((2 * l + 1) * (2 * l + 3));
    q = sqrt(num / denom);
    *c2 = which == 0 ? -((double)(l + 2) * q) : -((double)l * q);
}

void orc_angular_coefs(int l, int m, int which, double *c)
{
    angular_coefs(l, m, which, c, c + 1);
}

/* Offset (doubles) of block mp in a packed spectral array of nf fields. */
static long field_block(int lmax, const int *mvals, int mp, int nr, int nf)
{
    long off = 0;
    for (int i = 0; i < mp; ++i) off += (long)(lmax - mvals[i] + 1) * nr * 2 * nf;
    return off;
}

/* Spectral_Derivatives.F90:224-249 d_dtheta_buffer4D / :277-298 d_sdtheta_buffer4d (field
 * fin -> field fout of one packed array) and :329-350 / :381-401 the 4dbuff2arr forms (into
 * another array): per m-block, per column of field fin,
 *   B(l) = A(l-1) c1(l) + A(l+1) c2(l)  (m < l < lmax),  B(m) = A(m+1) c2(m),
 *   B(lmax) = A(lmax-1) c1(lmax);  B(m) = 0 when m = lmax.
 * Fields are 0-based here; in and out may be the same array (other fields). */
void orc_angular_dtheta(int which, int lmax, const int *mvals, int nm, int nr, const double *in, int nf_in, int fin,
                        double *out, int nf_out, int fout)
{
    const long ncol = 2L * nr;
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        const double *a = in + field_block(lmax, mvals, mp, nr, nf_in) + (long)fin * nl * ncol;
        double *b = out + field_block(lmax, mvals, mp, nr, nf_out) + (long)fout * nl * ncol;
        for (long k = 0; k < ncol; ++k) {
            const double *ak = a + k * nl;
            double *bk = b + k * nl;
            if (m == lmax) {
                bk[0] = 0.0;
                continue;
            }
            for (int l = m + 1; l <= lmax - 1; ++l) {
                double c1, c2;
                angular_coefs(l, m, which, &c1, &c2);
                bk[l - m] = ak[l - 1 - m] * c1 + ak[l + 1 - m] * c2;
            }
            double c1, c2;
            angular_coefs(m, m, which, &c1, &c2);
            bk[0] = ak[1] * c2;
            angular_coefs(lmax, m, which, &c1, &c2);
            bk[lmax - m] = ak[lmax - 1 - m] * c1;
        }
    }
}

/* Spectral_Derivatives.F90:573-603 d_by_dphi_rlmbuff4d (field fin -> fout, in place allowed)
 * and :495-509 d_by_dphi_4dbuff2arr: B(:, r, re) = -(m A(:, r, im)), B(:, r, im) = m A(:, r, re). */
void orc_angular_dphi(int lmax, const int *mvals, int nm, int nr, const double *in, int nf_in, int fin, double *out,
                      int nf_out, int fout)
{
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        const long fs = (long)nl * nr;  /* one (imi) half of a field */
        const double *a = in + field_block(lmax, mvals, mp, nr, nf_in) + (long)fin * 2 * fs;
        double *b = out + field_block(lmax, mvals, mp, nr, nf_out) + (long)fout * 2 * fs;
        for (long i = 0; i < fs; ++i) {
            const double re = a[i], im = a[fs + i];
            b[i] = -((double)m * im);
            b[fs + i] = (double)m * re;
        }
    }
}

/* Linear_Solve.F90:1034-1045 LU_Decompose_full = LAPACK dgetrf on one n x n matrix (column
 * major, leading dimension n), restated as the unblocked right-looking algorithm of
 * dgetf2 (LAPACK 3.x, the reference links LAPACK): per column k the pivot is the first
 * row of largest |a_ik|, rows k and p are swapped over all columns, the subdiagonal is
 * scaled by the pivot's reciprocal (a division where |pivot| < the safe minimum), and
 * the trailing block is updated a_ij - a_ik a_kj (rank-1, dger order).  pivots are
 * 1-based; returns info (0, or the first k + 1 with a zero pivot). */
int orc_lu_decompose(int n, double *a, int *piv)
{
    int info = 0;
    const double sfmin = 2.2250738585072014e-308;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(a[k + (long)k * n]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(a[i + (long)k * n]) > best) {
                best = fabs(a[i + (long)k * n]);
                p = i;
            }
        piv[k] = p + 1;
        const double pv = a[p + (long)k * n];
        if (pv != 0.0) {
            if (p != k)
                for (int j = 0; j < n; ++j) {
                    const double t = a[k + (long)j * n];
                    a[k + (long)j * n] = a[p + (long)j * n];
                    a[p + (long)j * n] = t;
                }
            if (fabs(pv) >= sfmin) {
                const double r = 1.0 / pv;
                for (int i = k + 1; i < n; ++i) a[i + (long)k * n] *= r;
            } else {
                for (int i = k + 1; i < n; ++i) a[i + (long)k * n] /= pv;
            }
        } else if (info == 0) {
            info = k + 1;
        }
        for (int j = k + 1; j < n; ++j) {
            const double y = a[k + (long)j * n];
            if (y != 0.0) {
                const double t = -y;
                for (int i = k + 1; i < n; ++i) a[i + (long)j * n] += a[i + (long)k * n] * t;
            }
        }
    }
    return info;
}

/* Linear_Solve.F90:1087-1107 LU_Solve_full = LAPACK dgetrs('N'): the row interchanges of
 * piv applied to the nrhs columns of b (leading dimension n) in order, then the unit
 * lower and the upper triangular solves in the column order of the reference BLAS dtrsm
 * (column j final, then subtracted from the rows after / before it). */
void orc_lu_solve(int n, const double *a, const int *piv, int nrhs, double *b)
{
    for (int c = 0; c < nrhs; ++c) {
        double *x = b + (long)c * n;
        for (int k = 0; k < n; ++k) {
            const int p = piv[k] - 1;
            if (p != k) {
                const double t = x[k];
                x[k] = x[p];
                x[p] = t;
            }
        }
        for (int j = 0; j < n; ++j)
            if (x[j] != 0.0)
                for (int i = j + 1; i < n; ++i) x[i] -= x[j] * a[i + (long)j * n];
        for (int j = n - 1; j >= 0; --j)
            if (x[j] != 0.0) {
                x[j] /= a[j + (long)j * n];
                for (int i = 0; i < j; ++i) x[i] -= x[j] * a[i + (long)j * n];
            }
    }
}

/* Load_Balance.F90:311-395 LM_Load_Balance on one radial rank: the (l, m) pairs of the
 * m values owned, l slowest (from the smallest m to lmax), m in m_values order within an
 * l; l_out / mp_out (0-based mp) receive the list, the return value is its length. */
int orc_lm_values(int lmax, const int *mvals, int nm, int *l_out, int *mp_out)
{
    int lmin = lmax;
    for (int mp = 0; mp < nm; ++mp)
        if (mvals[mp] < lmin) lmin = mvals[mp];
    int k = 0;
    for (int l = lmin; l <= lmax; ++l)
        for (int mp = 0; mp < nm; ++mp)
            if (mvals[mp] <= l) {
                if (l_out) l_out[k] = l;
                if (mp_out) mp_out[k] = mp;
                ++k;
            }
    return k;
}

static long spectral_block_offset(int lmax, const int *mvals, int mp, int nr, int nf)
{
    long off = 0;
    for (int i = 0; i < mp; ++i) off += (long)(lmax - mvals[i] + 1) * nr * 2 * nf;
    return off;
}

/* Spherical_Buffer.F90:705-843 Transpose_2b1b on one rank (all radii, nfields = nf):
 * p1b(r, imi, lm, f) = s2b(mp)%data(l, r, imi, f) for the lm-th pair (l, mp) of
 * LM_Load_Balance's list.  p1b is [f][lm][imi][r] (r fastest), s2b the packed
 * spectral layout. */
void orc_permute_2b1b(int lmax, const int *mvals, int nm, int nr, int nf, const double *s2b, double *p1b)
{
    const int n = orc_lm_values(lmax, mvals, nm, NULL, NULL);
    int *ls = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1)), *mps = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    orc_lm_values(lmax, mvals, nm, ls, mps);
    for (int k = 0; k < n; ++k) {
        const int mp = mps[k], m = mvals[mp], nl = lmax - m + 1;
        const double *blk = s2b + spectral_block_offset(lmax, mvals, mp, nr, nf);
        for (int f = 0; f < nf; ++f)
            for (int imi = 0; imi < 2; ++imi)
                for (int r = 0; r < nr; ++r)
                    p1b[(((long)f * n + k) * 2 + imi) * nr + r] = blk[(((long)f * 2 + imi) * nr + r) * nl + (ls[k] - m)];
    }
    free(ls);
    free(mps);
}

/* Spherical_Buffer.F90:845-932 Transpose_1a2a on one rank: the inverse,
 * s2a(mp)%data(l, r, imi, f) = p1a(r, imi, lm, f). */
void orc_permute_1a2a(int lmax, const int *mvals, int nm, int nr, int nf, const double *p1a, double *s2a)
{
    const int n = orc_lm_values(lmax, mvals, nm, NULL, NULL);
    int *ls = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1)), *mps = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    orc_lm_values(lmax, mvals, nm, ls, mps);
    for (int k = 0; k < n; ++k) {
        const int mp = mps[k], m = mvals[mp], nl = lmax - m + 1;
        double *blk = s2a + spectral_block_offset(lmax, mvals, mp, nr, nf);
        for (int f = 0; f < nf; ++f)
            for (int imi = 0; imi < 2; ++imi)
                for (int r = 0; r < nr; ++r)
                    blk[(((long)f * 2 + imi) * nr + r) * nl + (ls[k] - m)] = p1a[(((long)f * n + k) * 2 + imi) * nr + r];
    }
    free(ls);
    free(mps);
}

/* ------------------------------------------------------------------------ */
/* Chebyshev_Polynomials.F90:210-248 Gen_Colocation_Points (zeros branch: the
 * angle accumulates by repeated addition) and :65-207 Initialize_Cheby_Grid's
 * affine map of one domain onto [rmin, rmax] (grid descends from rmax). */
static void cheby_theta(int n, double *theta)
{
    const double pi = 3.1415926535897932384626433832795028841972;
    double dtheta = pi / n;
    double arg = dtheta * 0.5;
    for (int i = 0; i < n; ++i) {
        theta[i] = arg;
        arg = arg + dtheta;
    }
}

void orc_cheby_grid(int n, double rmin, double rmax, double *grid)
{
    double *th = malloc(sizeof(double) * (size_t)n), *x = malloc(sizeof(double) * (size_t)n);
    cheby_theta(n, th);
    for (int i = 0; i < n; ++i) x[i] = cos(th[i]);
    double upperb = rmax, lowerb = rmin, xmin = x[n - 1], xmax = x[0];
    double scaling = (upperb - lowerb) / (xmax - xmin);
    for (int i = 0; i < n; ++i) grid[i] = (x[i] - xmin) * scaling + lowerb;
    free(th);
    free(x);
}

/* Chebyshev_Polynomials.F90:176-200 (one domain): the radial integration weights
 * int_scale r_i^2 sqrt(1 - x_i^2), int_scale = (3 pi scaling) / ((rmax^3 - rmin^3) n), the
 * first and last halved; scaling = (rmax - rmin) / (x_1 - x_n) as for the grid.  Products
 * and powers are rounded left to right (x**3 as (x x) x). */
/* The weights of one domain [lowerb, upperb] of n points of a grid spanning [gmin, gmax]
 * (Chebyshev_Polynomials.F90:176-200 inside its loop over domains: int_scale takes the whole
 * grid's gmax^3 - gmin^3, the domain's slope and n). */
void orc_cheby_weights_domain(int n, double lowerb, double upperb, double gmin, double gmax, double *w)
{
    const double pi = 3.1415926535897932384626433832795028841972;
    double *th = malloc(sizeof(double) * (size_t)n), *x = malloc(sizeof(double) * (size_t)n);
    cheby_theta(n, th);
    for (int i = 0; i < n; ++i) x[i] = cos(th[i]);
    const double xmin = x[n - 1], xmax = x[0];
    const double scaling = (upperb - lowerb) / (xmax - xmin);
    const double int_scale = (3 * pi * scaling) / ((gmax * gmax * gmax - gmin * gmin * gmin) * n);
    for (int i = 0; i < n; ++i) {
        const double r = (x[i] - xmin) * scaling + lowerb;
        w[i] = int_scale * (r * r) * sqrt(1.0 - x[i] * x[i]);
    }
    w[0] *= 0.5;
    w[n - 1] *= 0.5;
    free(th);
    free(x);
}

void orc_cheby_weights(int n, double rmin, double rmax, double *w)
{
    orc_cheby_weights_domain(n, rmin, rmax, rmin, rmax, w);
}

/* Legendre_Polynomials.F90:332-373 as stored (Real*16 coloc, gl_weights): the quad
 * values' 16 bytes each. */
void orc_gauss_legendre_binary128(int n, void *x16, void *w16)
{
    gauss_legendre_q(n, (quad *)x16, (quad *)w16);
}

/* Chebyshev_Polynomials.F90:251-297 Gen_Tn: T_k(x_r) = cos(k theta_r);
 * ceven[i][r] (r fastest, 1..n_x) = T_{2i}, codd = T_{2i+1}; the x = 0 row is
 * halved when n is odd. */
void orc_cheby_tables(int n, double *ceven, double *codd)
{
    double *th = malloc(sizeof(double) * (size_t)n);
    cheby_theta(n, th);
    const int n_odd = n / 2, n_even = n_odd + n % 2, n_x = n_even;
    for (int i = 0; i < n_even; ++i)
        for (int r = 0; r < n_x; ++r) ceven[(long)i * n_x + r] = cos((double)(2 * i) * th[r]);
    for (int i = 0; i < n_odd; ++i)
        for (int r = 0; r < n_x; ++r) codd[(long)i * n_x + r] = cos((double)(2 * i + 1) * th[r]);
    if (n_x != n_odd) {
        for (int i = 0; i < n_even; ++i) ceven[(long)i * n_x + n_x - 1] *= 0.5;
        for (int i = 0; i < n_odd; ++i) codd[(long)i * n_x + n_x - 1] *= 0.5;
    }
    free(th);
}

/* Chebyshev_Polynomials.F90:375-448 To_Spectral4D (one domain):
 * fold f(i) +- f(n+1-i), contract with alpha = 2/n, interleave even -> c(2i-1),
 * odd -> c(2i). */
void orc_cheby_to_spectral(int n, long ncols, const double *ceven, const double *codd,
                           const double *f, double *c)
{
    const int n_odd = n / 2, n_even = n_odd + n % 2, n_x = n_even;
    const double alpha = 2.0 / n;
    double *fe = malloc(sizeof(double) * (size_t)n_x), *fo = malloc(sizeof(double) * (size_t)n_x);
    for (long col = 0; col < ncols; ++col) {
        const double *fc = f + col * n;
        double *cc = c + col * n;
        for (int i = 0; i < n_x; ++i) {
            fe[i] = fc[i] + fc[n - 1 - i];
            fo[i] = fc[i] - fc[n - 1 - i];
        }
        for (int k = 0; k < n_even; ++k) {
            double acc = 0.0;
            for (int i = 0; i < n_x; ++i) acc += ceven[(long)k * n_x + i] * fe[i];
            cc[2 * k] = alpha * acc;
        }
        for (int k = 0; k < n_odd; ++k) {
            double acc = 0.0;
            for (int i = 0; i < n_x; ++i) acc += codd[(long)k * n_x + i] * fo[i];
            cc[2 * k + 1] = alpha * acc;
        }
    }
    free(fe);
    free(fo);
}

/* Chebyshev_Polynomials.F90:450-558 From_Spectral4D (one domain):
 * f = sum_n c_n T_n - c_0/2; the even part is mirrored, the odd part added with
 * +/-.  For odd n the x = 0 row (halved in the tables) is doubled back, and the
 * odd product at that row is added twice as the reference does (:537-539; that
 * product vanishes to rounding because T_odd(0) = 0). */
void orc_cheby_from_spectral(int n, long ncols, const double *ceven, const double *codd,
                             const double *c, double *f)
{
    const int n_odd = n / 2, n_even = n_odd + n % 2, n_x = n_even;
    double *fo = malloc(sizeof(double) * (size_t)n_x);
    for (long col = 0; col < ncols; ++col) {
        const double *cc = c + col * n;
        double *fc = f + col * n;
        for (int i = 0; i < n_x; ++i) {
            double acc = 0.0;
            for (int k = 0; k < n_even; ++k) acc += ceven[(long)k * n_x + i] * cc[2 * k];
            fc[i] = acc;
            fc[n - 1 - i] = acc;
        }
        if (n_even != n_odd) fc[n_x - 1] = fc[n_x - 1] * 2.0;
        for (int i = 0; i < n_x; ++i) {
            double acc = 0.0;
            for (int k = 0; k < n_odd; ++k) acc += codd[(long)k * n_x + i] * cc[2 * k + 1];
            fo[i] = acc;
        }
        for (int i = 0; i < n_odd; ++i) {
            fc[i] = fc[i] + fo[i];
            fc[n - 1 - i] = fc[n - 1 - i] - fo[i];
        }
        if (n_even != n_odd) fc[n_x - 1] = fc[n_x - 1] + fo[n_x - 1] * 2.0;
        const double c0h = cc[0] / 2.0;
        for (int i = 0; i < n; ++i) fc[i] = fc[i] - c0h;
    }
    free(fo);
}

/* Chebyshev_Polynomials.F90:353-374 dealias_buffer4d: zero n >= rda-1 (0-based),
 * rda = (2n)/3 + 1. */
void orc_cheby_dealias(int n, long ncols, double *c)
{
    const int rda = (2 * n) / 3 + 1;
    for (long col = 0; col < ncols; ++col)
        for (int i = rda - 1; i < n; ++i) c[col * n + i] = 0.0;
}

/* Chebyshev_Polynomials.F90:176-191: the one-domain derivative scaling, 1 / scaling with
 * scaling = (rmax - rmin) / (x_1 - x_n) (the zeros grid's end points). */
double orc_cheby_deriv_scaling(int n, double rmin, double rmax)
{
    double *th = malloc(sizeof(double) * (size_t)n);
    cheby_theta(n, th);
    const double xmin = cos(th[n - 1]), xmax = cos(th[0]);
    free(th);
    const double scaling = (rmax - rmin) / (xmax - xmin);
    return 1.0 / scaling;
}

/* Fortran's x**k for the small integer k of a derivative order: x, x*x, (x*x)*x,
 * (x*x)*(x*x) */
static double ipow_small(double x, int k)
{
    double r = 1.0;
    if (k == 1) return x;
    if (k == 2) return x * x;
    if (k == 3) return (x * x) * x;
    if (k == 4) return (x * x) * (x * x);
    for (int i = 0; i < k; ++i) r *= x;
    return r;
}

/* Chebyshev_Polynomials.F90:560-640 Cheby_Deriv_Buffer_4D (d_by_dr_cp, ind /= dind) and
 * :642-730 Cheby_Deriv_OutOf_Buffer (d_by_dr_nobuffer), one domain of n coefficients per
 * column: b(n-1) = 0, b(n-2) = 2(n-1) a(n-1), b(i) = b(i+2) + 2(i+1) a(i+1) for i = n-3 .. 0,
 * applied dorder times (each pass to the previous pass's result), then every entry
 * times deriv_scaling**dorder.  `fused` selects fma(2(i+1), a, b) for the step instead
 * of a rounded product and a rounded sum; the compiled reference decides which. */
void orc_cheby_deriv(int n, double ds, int dorder, long ncols, const double *c, double *d, int fused)
{
    double *cur = malloc(sizeof(double) * (size_t)n), *nxt = malloc(sizeof(double) * (size_t)n);
    const double p = ipow_small(ds, dorder);
    for (long col = 0; col < ncols; ++col) {
        for (int i = 0; i < n; ++i) cur[i] = c[col * n + i];
        for (int o = 0; o < dorder; ++o) {
            nxt[n - 1] = 0.0;
            if (n >= 2) nxt[n - 2] = 2.0 * (n - 1) * cur[n - 1];
            for (int i = n - 3; i >= 0; --i)
                nxt[i] = fused ? fma(2.0 * (i + 1), cur[i + 1], nxt[i + 2]) : nxt[i + 2] + 2.0 * (i + 1) * cur[i + 1];
            double *t = cur;
            cur = nxt;
            nxt = t;
        }
        for (int i = 0; i < n; ++i) d[col * n + i] = cur[i] * p;
    }
    free(cur);
    free(nxt);
}

/* Load_Balance.F90:126-226 M_Balance_v0 for one rank: pairs (i, n_m-1-i), then
 * the middle m when n_m is odd. */
void orc_m_order_balanced(int lmax, int *mvals)
{
    const int nmm = lmax + 1, npairs = nmm / 2;
    int k = 0;
    for (int i = 0; i < npairs; ++i) {
        mvals[k++] = i;
        mvals[k++] = nmm - 1 - i;
    }
    if (nmm % 2 == 1) mvals[k++] = nmm / 2;
}

/* over_l_l_plus1 (ProblemSize.F90:506-513): 1 / (l (l + 1)) from the double l, 1 at l = 0 */
static double over_l_l_plus1(int l)
{
    const double ell = (double)l;
    return l == 0 ? 1.0 : 1.0 / (ell * (ell + 1.0));
}

/* Sphere_Hybrid_Space.F90:127-214 rlm_spaceb, hydro part (adjust_emf left out), in place on a
 * packed buffer of nf fields (0-based W, P, Z fields wv, pv, zv), step for step:
 *   W = W ol;  t1 = d_by_sdtheta Z;  t2 = d_by_dphi P;  t1 = (t2 - t1) ol;
 *   t2 = d_by_dphi Z;  Z = d_by_sdtheta P;  P = (Z + t2) ol;  Z = t1;
 *   m = 0: l = 0 of P, W, Z = 0;  every block: l = l_max of every field = 0. */
void orc_rlm_spaceb(int lmax, const int *mvals, int nm, int nr, double *buf, int nf, int wv, int pv, int zv)
{
    const long fsz = field_block(lmax, mvals, nm, nr, 1);
    double *t1 = (double *)malloc(sizeof(double) * (fsz + 1));
    double *t2 = (double *)malloc(sizeof(double) * (fsz + 1));
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        double *w = buf + field_block(lmax, mvals, mp, nr, nf) + (long)wv * nl * 2 * nr;
        for (long k = 0; k < 2L * nr; ++k)
            for (int l = m; l <= lmax; ++l) w[k * nl + l - m] = w[k * nl + l - m] * over_l_l_plus1(l);
    }
    orc_angular_dtheta(1, lmax, mvals, nm, nr, buf, nf, zv, t1, 1, 0);
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, pv, t2, 1, 0);
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        const long o = field_block(lmax, mvals, mp, nr, 1);
        for (long k = 0; k < 2L * nr; ++k)
            for (int l = m; l <= lmax; ++l) {
                const long e = o + k * nl + l - m;
                t1[e] = (t2[e] - t1[e]) * over_l_l_plus1(l);
            }
    }
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, zv, t2, 1, 0);
    orc_angular_dtheta(1, lmax, mvals, nm, nr, buf, nf, pv, buf, nf, zv);
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        const long fs = (long)nl * 2 * nr, o = field_block(lmax, mvals, mp, nr, 1);
        double *blk = buf + field_block(lmax, mvals, mp, nr, nf);
        double *P = blk + pv * fs, *Z = blk + zv * fs, *W = blk + wv * fs;
        for (long k = 0; k < 2L * nr; ++k)
            for (int l = m; l <= lmax; ++l) {
                const long e = k * nl + l - m;
                P[e] = (Z[e] + t2[o + e]) * over_l_l_plus1(l);
            }
        for (long e = 0; e < fs; ++e) Z[e] = t1[o + e];
        if (m == 0)
            for (long k = 0; k < 2L * nr; ++k) P[k * nl] = W[k * nl] = Z[k * nl] = 0.0;
        for (long k = 0; k < 2L * nr * nf; ++k) blk[k * nl + nl - 1] = 0.0;
    }
    free(t1);
    free(t2);
}

/* out(l, r, imi) = (lfac(l) in(l, r, imi)) rfac(r) on one field (either factor may be NULL),
 * rounded in the order of l_l_plus1(m:l_max)*SBUFFA(IDX2,f)*Over_RhoRSQ(r)
 * (Sphere_Hybrid_Space.F90:244); in place allowed. */
void orc_lm_scale(int lmax, const int *mvals, int nm, int nr, const double *in, int nf_in, int fin, double *out,
                  int nf_out, int fout, const double *lfac, const double *rfac)
{
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        const long fs = (long)nl * 2 * nr;
        const double *a = in + field_block(lmax, mvals, mp, nr, nf_in) + fin * fs;
        double *b = out + field_block(lmax, mvals, mp, nr, nf_out) + fout * fs;
        for (int imi = 0; imi < 2; ++imi)
            for (int r = 0; r < nr; ++r)
                for (int l = m; l <= lmax; ++l) {
                    const long e = ((long)imi * nr + r) * nl + l - m;
                    double v = a[e];
                    if (lfac) v = lfac[l] * v;
                    if (rfac) v = v * rfac[r];
                    b[e] = v;
                }
    }
}

/* Sphere_Hybrid_Space.F90:70-72 (and :102-104, :207-209): data(l_max, :, :, :) = 0 per block */
void orc_zero_lmax(int lmax, const int *mvals, int nm, int nr, double *buf, int nf)
{
    for (int mp = 0; mp < nm; ++mp) {
        const int nl = lmax - mvals[mp] + 1;
        double *blk = buf + field_block(lmax, mvals, mp, nr, nf);
        for (long k = 0; k < 2L * nr * nf; ++k) blk[k * nl + nl - 1] = 0.0;
    }
}

/* l_l_plus1 (ProblemSize.F90:506-513): l (l + 1) from the double l */
static double l_l_plus1(int l)
{
    const double ell = (double)l;
    return ell * (ell + 1.0);
}

/* The elementwise statements of the rlm_spacea routines on one field-sized slab (every
 * m-block, radius, imi, l).  e is the index of (mp, imi, r, l) in a one-field buffer. */
typedef struct {
    int lmax, nm, nr;
    const int *mvals;
} slab;

static double *fld(const slab *g, double *buf, int nf, int f, int mp)
{
    const int nl = g->lmax - g->mvals[mp] + 1;
    return buf + field_block(g->lmax, g->mvals, mp, g->nr, nf) + (long)f * nl * 2 * g->nr;
}

/* field f = (l_l_plus1(l) field g) rfac(r)   (SBUFFA(IDX2,f) = l_l_plus1(m:l_max)*SBUFFA(IDX2,g)*rfac(r)) */
static void st_lscale(const slab *g, double *buf, int nf, int fo, int fi, const double *rfac)
{
    for (int mp = 0; mp < g->nm; ++mp) {
        const int m = g->mvals[mp], nl = g->lmax - m + 1;
        double *o = fld(g, buf, nf, fo, mp), *a = fld(g, buf, nf, fi, mp);
        for (int imi = 0; imi < 2; ++imi)
            for (int r = 0; r < g->nr; ++r)
                for (int l = m; l <= g->lmax; ++l) {
                    const long e = ((long)imi * g->nr + r) * nl + l - m;
                    o[e] = l_l_plus1(l) * a[e] * rfac[r];
                }
    }
}

/* field f = field f - field v rfac(r)   (SBUFFA(IDX2,f) = SBUFFA(IDX2,f) - SBUFFA(IDX2,v)*rfac(r)) */
static void st_sub_scaled(const slab *g, double *buf, int nf, int f, int v, const double *rfac)
{
    for (int mp = 0; mp < g->nm; ++mp) {
        const int m = g->mvals[mp], nl = g->lmax - m + 1;
        double *o = fld(g, buf, nf, f, mp), *a = fld(g, buf, nf, v, mp);
        for (int imi = 0; imi < 2; ++imi)
            for (int r = 0; r < g->nr; ++r)
                for (int i = 0; i < nl; ++i) {
                    const long e = ((long)imi * g->nr + r) * nl + i;
                    o[e] = o[e] - (rfac ? a[e] * rfac[r] : a[e]);
                }
    }
}

/* field f of buffer a = itself - field u of buffer b (nb fields) times rfac(r) */
static void st_sub_scaled_from(const slab *g, double *a, int na, int f, double *b, int nb, int u, const double *rfac)
{
    for (int mp = 0; mp < g->nm; ++mp) {
        const int m = g->mvals[mp], nl = g->lmax - m + 1;
        double *o = fld(g, a, na, f, mp), *x = fld(g, b, nb, u, mp);
        for (int imi = 0; imi < 2; ++imi)
            for (int r = 0; r < g->nr; ++r)
                for (int i = 0; i < nl; ++i) {
                    const long e = ((long)imi * g->nr + r) * nl + i;
                    o[e] = o[e] - x[e] * rfac[r];
                }
    }
}

/* one-field slab t times rfac(r) into field f   (SBUFFA(IDX2,f) = ftemp1(mp)%data(IDX2)*rfac(r)) */
static void st_store_scaled(const slab *g, double *buf, int nf, int f, const double *t, const double *rfac)
{
    for (int mp = 0; mp < g->nm; ++mp) {
        const int m = g->mvals[mp], nl = g->lmax - m + 1;
        double *o = fld(g, buf, nf, f, mp);
        const double *a = t + field_block(g->lmax, g->mvals, mp, g->nr, 1);
        for (int imi = 0; imi < 2; ++imi)
            for (int r = 0; r < g->nr; ++r)
                for (int i = 0; i < nl; ++i) {
                    const long e = ((long)imi * g->nr + r) * nl + i;
                    o[e] = rfac ? a[e] * rfac[r] : a[e];
                }
    }
}

/* t1 = t1 + sg t2 on one-field slabs (sg = +1 or -1) */
static void st_combine(long n, double *t1, const double *t2, int sg)
{
    for (long e = 0; e < n; ++e) t1[e] = sg > 0 ? t1[e] + t2[e] : t1[e] - t2[e];
}

/* Sphere_Hybrid_Space.F90:235-355 Velocity_Components then Velocity_Derivatives, then
 * rlm_spacea's d_by_dtheta(tvar, dtdt) (:80), in place on a packed buffer of nf fields, step for
 * step with two one-field temporaries.  v = 0-based (W, dWdr, d2Wdr2, Z, dZdr, T, vtheta, vphi,
 * dvtdr, dvpdr); vr, dvrdr, dtdt and dvrdt are the W, dWdr, d2Wdr2 and dZdr fields
 * (Fields.F90:323-326).  dsdr_over_cp NULL: not pseudo-incompressible. */
void orc_rlm_spacea_hydro(int lmax, const int *mvals, int nm, int nr, double *buf, int nf, const int *v,
                          const double *over_rhor, const double *over_rhorsq, const double *drho_term,
                          const double *two_over_r, const double *dlnrho, const double *dsdr_over_cp)
{
    const slab g = {lmax, nm, nr, mvals};
    const int W = v[0], dW = v[1], d2W = v[2], Z = v[3], dZ = v[4], T = v[5];
    const int vt = v[6], vp = v[7], dvt = v[8], dvp = v[9];
    const long fsz = field_block(lmax, mvals, nm, nr, 1);
    double *t1 = (double *)malloc(sizeof(double) * (fsz + 1));
    double *t2 = (double *)malloc(sizeof(double) * (fsz + 1));
    /* Velocity_Components */
    st_lscale(&g, buf, nf, W, W, over_rhorsq);
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, dW, t1, 1, 0);
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, Z, t2, 1, 0);
    st_combine(fsz, t1, t2, +1);
    st_store_scaled(&g, buf, nf, vt, t1, over_rhor);
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, dW, t1, 1, 0);
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, Z, t2, 1, 0);
    st_combine(fsz, t1, t2, -1);
    st_store_scaled(&g, buf, nf, vp, t1, over_rhor);
    /* Velocity_Derivatives */
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, d2W, t1, 1, 0);
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, dZ, t2, 1, 0);
    st_combine(fsz, t1, t2, +1);
    st_store_scaled(&g, buf, nf, dvt, t1, over_rhor);
    st_sub_scaled(&g, buf, nf, dvt, vt, drho_term);
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, d2W, t1, 1, 0);
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, dZ, t2, 1, 0);
    st_combine(fsz, t1, t2, -1);
    st_store_scaled(&g, buf, nf, dvp, t1, over_rhor);
    st_sub_scaled(&g, buf, nf, dvp, vp, drho_term);
    st_lscale(&g, buf, nf, dW, dW, over_rhorsq);
    st_sub_scaled(&g, buf, nf, dW, W, two_over_r);
    st_sub_scaled(&g, buf, nf, dW, W, dlnrho);
    if (dsdr_over_cp) st_sub_scaled(&g, buf, nf, dW, W, dsdr_over_cp);
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, W, buf, nf, dZ);
    st_lscale(&g, buf, nf, Z, Z, over_rhorsq);
    /* rlm_spacea: d_by_dtheta(wsp%s2a, tvar, dtdt) */
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, T, buf, nf, d2W);
    free(t1);
    free(t2);
}

/* Sphere_Hybrid_Space.F90:377-460 Compute_BandCurlB, in place on a packed buffer of nf fields,
 * step for step.  v = 0-based (C, A, dCdr, dAdr, d2Cdr2, curlbr); Br, Btheta, Bphi,
 * curlbtheta and curlbphi are the C, A, dCdr, d2Cdr2 and dAdr fields (Fields.F90:329-334);
 * one_over_rsq = OneOverRSquared. */
void orc_compute_bandcurlb(int lmax, const int *mvals, int nm, int nr, double *buf, int nf, const int *v,
                           const double *one_over_rsq)
{
    const slab g = {lmax, nm, nr, mvals};
    const int C = v[0], A = v[1], dC = v[2], dA = v[3], d2C = v[4], cbr = v[5];
    const long fsz = field_block(lmax, mvals, nm, nr, 1);
    double *t1 = (double *)malloc(sizeof(double) * (fsz + 1));
    double *t2 = (double *)malloc(sizeof(double) * (fsz + 1));
    st_lscale(&g, buf, nf, C, C, one_over_rsq);      /* Br */
    st_lscale(&g, buf, nf, cbr, A, one_over_rsq);    /* [curl B]_r */
    st_sub_scaled(&g, buf, nf, d2C, C, NULL);        /* d2cdr2 - Br */
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, dA, t1, 1, 0);
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, dA, t2, 1, 0);
    st_store_scaled(&g, buf, nf, dA, t2, NULL);
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, d2C, t2, 1, 0);
    {   /* curlbphi = curlbphi + ftemp2 */
        for (int mp = 0; mp < nm; ++mp) {
            const int nl = lmax - mvals[mp] + 1;
            double *o = fld(&g, buf, nf, dA, mp);
            const double *a = t2 + field_block(lmax, mvals, mp, nr, 1);
            for (long e = 0; e < 2L * nr * nl; ++e) o[e] = o[e] + a[e];
        }
    }
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, d2C, t2, 1, 0);
    st_combine(fsz, t1, t2, -1);
    st_store_scaled(&g, buf, nf, d2C, t1, NULL);     /* curlbtheta */
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, A, t1, 1, 0);
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, A, t2, 1, 0);
    st_store_scaled(&g, buf, nf, A, t2, NULL);
    orc_angular_dtheta(0, lmax, mvals, nm, nr, buf, nf, dC, t2, 1, 0);
    {   /* Btheta = A + ftemp2 */
        for (int mp = 0; mp < nm; ++mp) {
            const int nl = lmax - mvals[mp] + 1;
            double *o = fld(&g, buf, nf, A, mp);
            const double *a = t2 + field_block(lmax, mvals, mp, nr, 1);
            for (long e = 0; e < 2L * nr * nl; ++e) o[e] = o[e] + a[e];
        }
    }
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, dC, t2, 1, 0);
    {   /* Bphi = ftemp2 - ftemp1 */
        for (long e = 0; e < fsz; ++e) t2[e] = t2[e] - t1[e];
        st_store_scaled(&g, buf, nf, dC, t2, NULL);
    }
    free(t1);
    free(t2);
}

/* Sphere_Hybrid_Space.F90:557-602 Adjust_Emf, in place on a packed buffer of nf fields after
 * PtS, step for step: t1 = d_by_sdtheta emfphi; t2 = d_by_dphi emftheta; t3 = emftheta;
 * emftheta = (t1 - t2) ol (Cvar is the emftheta field, Fields.F90:363-364);
 * t2 = d_by_dphi emfphi; emfphi = t3; t1 = d_by_sdtheta emfphi; emfphi = (t2 + t1) ol;
 * ol = over_l_l_plus1.  et, ep = 0-based emftheta, emfphi. */
void orc_adjust_emf(int lmax, const int *mvals, int nm, int nr, double *buf, int nf, int et, int ep)
{
    const slab g = {lmax, nm, nr, mvals};
    const long fsz = field_block(lmax, mvals, nm, nr, 1);
    double *t1 = (double *)malloc(sizeof(double) * (fsz + 1));
    double *t2 = (double *)malloc(sizeof(double) * (fsz + 1));
    double *t3 = (double *)malloc(sizeof(double) * (fsz + 1));
    orc_angular_dtheta(1, lmax, mvals, nm, nr, buf, nf, ep, t1, 1, 0);
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, et, t2, 1, 0);
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        const long o = field_block(lmax, mvals, mp, nr, 1);
        double *T = fld(&g, buf, nf, et, mp);
        for (long e = 0; e < 2L * nr * nl; ++e) t3[o + e] = T[e];
        for (long k = 0; k < 2L * nr; ++k)
            for (int l = m; l <= lmax; ++l) {
                const long e = k * nl + l - m;
                T[e] = (t1[o + e] - t2[o + e]) * over_l_l_plus1(l);
            }
    }
    orc_angular_dphi(lmax, mvals, nm, nr, buf, nf, ep, t2, 1, 0);
    for (int mp = 0; mp < nm; ++mp) {
        const int nl = lmax - mvals[mp] + 1;
        const long o = field_block(lmax, mvals, mp, nr, 1);
        double *P = fld(&g, buf, nf, ep, mp);
        for (long e = 0; e < 2L * nr * nl; ++e) P[e] = t3[o + e];
    }
    orc_angular_dtheta(1, lmax, mvals, nm, nr, buf, nf, ep, t1, 1, 0);
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        const long o = field_block(lmax, mvals, mp, nr, 1);
        double *P = fld(&g, buf, nf, ep, mp);
        for (long k = 0; k < 2L * nr; ++k)
            for (int l = m; l <= lmax; ++l) {
                const long e = k * nl + l - m;
                P[e] = (t2[o + e] + t1[o + e]) * over_l_l_plus1(l);
            }
    }
    free(t1);
    free(t2);
    free(t3);
}

/* Sphere_Hybrid_Space.F90:525-541 Hybrid_Output_Initial, magnetic part: from the spectral buffer
 * s (nf fields) into the output co-buffer co (nf_co fields) and the two-field work buffer tmp
 * (the reference's ftemp3, ftemp4): co[acb] = l_l_plus1 s[avar] one_over_rsq, co[dbrcb] = s[dcdr],
 * tmp[0] = s[d2cdr2], tmp[1] = s[dadr].  v = 0-based (avar, dcdr, d2cdr2, dadr, avar_cb,
 * dbrdr_cb). */
void orc_hybrid_output_initial(int lmax, const int *mvals, int nm, int nr, const double *s, int nf, double *co,
                               int nf_co, double *tmp, const int *v, const double *one_over_rsq)
{
    const slab g = {lmax, nm, nr, mvals};
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp], nl = lmax - m + 1;
        double *S = (double *)s;
        const double *A = fld(&g, S, nf, v[0], mp), *DC = fld(&g, S, nf, v[1], mp);
        const double *D2C = fld(&g, S, nf, v[2], mp), *DA = fld(&g, S, nf, v[3], mp);
        double *ACB = fld(&g, co, nf_co, v[4], mp), *DBR = fld(&g, co, nf_co, v[5], mp);
        double *T3 = fld(&g, tmp, 2, 0, mp), *T4 = fld(&g, tmp, 2, 1, mp);
        for (int imi = 0; imi < 2; ++imi)
            for (int r = 0; r < nr; ++r)
                for (int l = m; l <= lmax; ++l) {
                    const long e = ((long)imi * nr + r) * nl + l - m;
                    ACB[e] = l_l_plus1(l) * A[e] * one_over_rsq[r];
                    DBR[e] = DC[e];
                    T3[e] = D2C[e];
                    T4[e] = DA[e];
                }
    }
}

/* Sphere_Hybrid_Space.F90:543-552 Hybrid_Output_Final up to its Legendre transform: co's l_max
 * zeroed; Hydro_Output_Derivatives (:224-233) co[dpdt_cb] = d_by_dtheta s[pvar]; with tmp (the
 * magnetic case) BField_Derivatives (:437-498):
 *   t1 = d_by_dtheta tmp[0] + d_by_dphi tmp[1]; co[dbtdr_cb] = t1 one_over_r - s[btheta] one_over_rsq;
 *   t1 = d_by_dphi tmp[0] - d_by_dtheta tmp[1]; co[dbpdr_cb] = t1 one_over_r - s[bphi] one_over_rsq;
 *   co[dbrdr_cb] = l_l_plus1 co[dbrdr_cb] one_over_rsq - s[br] two_over_r;
 *   co[dbrdt_cb] = d_by_dtheta s[br].
 * v = 0-based (pvar, br, btheta, bphi, dpdt_cb, dbtdr_cb, dbpdr_cb, dbrdr_cb, dbrdt_cb). */
void orc_hybrid_output_final(int lmax, const int *mvals, int nm, int nr, const double *s, int nf, double *co,
                             int nf_co, const double *tmp, const int *v, const double *one_over_r,
                             const double *one_over_rsq, const double *two_over_r)
{
    const slab g = {lmax, nm, nr, mvals};
    const long fsz = field_block(lmax, mvals, nm, nr, 1);
    double *S = (double *)s, *Tm = (double *)tmp;
    orc_zero_lmax(lmax, mvals, nm, nr, co, nf_co);
    orc_angular_dtheta(0, lmax, mvals, nm, nr, S, nf, v[0], co, nf_co, v[4]);
    if (!tmp) return;
    double *t1 = (double *)malloc(sizeof(double) * (fsz + 1));
    double *t2 = (double *)malloc(sizeof(double) * (fsz + 1));
    orc_angular_dtheta(0, lmax, mvals, nm, nr, Tm, 2, 0, t1, 1, 0);
    orc_angular_dphi(lmax, mvals, nm, nr, Tm, 2, 1, t2, 1, 0);
    st_combine(fsz, t1, t2, +1);
    st_store_scaled(&g, co, nf_co, v[5], t1, one_over_r);
    st_sub_scaled_from(&g, co, nf_co, v[5], S, nf, v[2], one_over_rsq);
    orc_angular_dphi(lmax, mvals, nm, nr, Tm, 2, 0, t1, 1, 0);
    orc_angular_dtheta(0, lmax, mvals, nm, nr, Tm, 2, 1, t2, 1, 0);
    st_combine(fsz, t1, t2, -1);
    st_store_scaled(&g, co, nf_co, v[6], t1, one_over_r);
    st_sub_scaled_from(&g, co, nf_co, v[6], S, nf, v[3], one_over_rsq);
    st_lscale(&g, co, nf_co, v[7], v[7], one_over_rsq);
    st_sub_scaled_from(&g, co, nf_co, v[7], S, nf, v[1], two_over_r);
    orc_angular_dtheta(0, lmax, mvals, nm, nr, S, nf, v[1], co, nf_co, v[8]);
    free(t1);
    free(t2);
}

/* LU_Decompose_band (Linear_Solve.F90:1041-1055): LAPACK dgbtrf on an n x n band matrix in
 * LAPACK band storage ab(ldab, n), ldab >= 2 kl + ku + 1 (the reference passes kl = ku =
 * (ldab - 1) / 3), restated as the unblocked dgbtf2 that dgbtrf runs for ku <= 64: fill-in rows
 * zeroed, per column the first largest |a| of the kl + 1 candidates as pivot, the row
 * interchange over columns j..ju, the multipliers scaled by the reciprocal of the pivot (dscal)
 * and the rank-1 update in reference-BLAS dger order (a column skipped where its y is 0;
 * a + x (-y), rounded separately).  Pivots 1-based; returns info (first zero pivot or 0). */
int orc_band_lu_decompose(int n, int kl, int ku, double *ab, int ldab, int *piv)
{
    const int kv = ku + kl;
#define AB(i, j) ab[(long)((j) - 1) * ldab + ((i) - 1)]
    int info = 0;
    for (int j = ku + 2; j <= (kv < n ? kv : n); ++j)
        for (int i = kv - j + 2; i <= kl; ++i) AB(i, j) = 0.0;
    int ju = 1;
    for (int j = 1; j <= n; ++j) {
        if (j + kv <= n)
            for (int i = 1; i <= kl; ++i) AB(i, j + kv) = 0.0;
        const int km = kl < n - j ? kl : n - j;
        int jp = 1;
        double best = fabs(AB(kv + 1, j));
        for (int i = 2; i <= km + 1; ++i)
            if (fabs(AB(kv + i, j)) > best) {
                best = fabs(AB(kv + i, j));
                jp = i;
            }
        piv[j - 1] = jp + j - 1;
        if (AB(kv + jp, j) != 0.0) {
            const int jn = j + ku + jp - 1 < n ? j + ku + jp - 1 : n;
            if (jn > ju) ju = jn;
            if (jp != 1)
                for (int c = 0; c <= ju - j; ++c) {
                    const double t = AB(kv + jp - c, j + c);
                    AB(kv + jp - c, j + c) = AB(kv + 1 - c, j + c);
                    AB(kv + 1 - c, j + c) = t;
                }
            if (km > 0) {
                const double r = 1.0 / AB(kv + 1, j);
                for (int i = 1; i <= km; ++i) AB(kv + 1 + i, j) = r * AB(kv + 1 + i, j);
                for (int c = 1; c <= ju - j; ++c) {
                    const double y = AB(kv + 1 - c, j + c);
                    if (y == 0.0) continue;
                    const double t = -y;
                    for (int i = 1; i <= km; ++i) AB(kv + 1 + i - c, j + c) = AB(kv + 1 + i - c, j + c) + AB(kv + 1 + i, j) * t;
                }
            }
        } else if (info == 0) {
            info = j;
        }
    }
#undef AB
    return info;
}

/* LU_Solve_Band (Linear_Solve.F90:1056-1086): dgbtrs 'N' with nrhs columns of b (leading
 * dimension n): L solve per column j of the factor (row interchange, then dger's b(j+i) +
 * l(i) (-b(j)), a right-hand side skipped where b(j) = 0), then per right-hand side the
 * reference-BLAS dtbsv upper non-unit solve (x(j) / u(j,j), then x(i) - x(j) u(i,j) for the
 * kl + ku rows above, skipped where x(j) = 0). */
void orc_band_lu_solve(int n, int kl, int ku, const double *ab, int ldab, const int *piv, int nrhs, double *b)
{
    const int kd = ku + kl + 1, k = kl + ku;
#define AB(i, j) ab[(long)((j) - 1) * ldab + ((i) - 1)]
#define B(i, c) b[(long)(c) * n + ((i) - 1)]
    if (kl > 0)
        for (int j = 1; j <= n - 1; ++j) {
            const int lm = kl < n - j ? kl : n - j, l = piv[j - 1];
            if (l != j)
                for (int c = 0; c < nrhs; ++c) {
                    const double t = B(l, c);
                    B(l, c) = B(j, c);
                    B(j, c) = t;
                }
            for (int c = 0; c < nrhs; ++c) {
                const double y = B(j, c);
                if (y == 0.0) continue;
                const double t = -y;
                for (int i = 1; i <= lm; ++i) B(j + i, c) = B(j + i, c) + AB(kd + i, j) * t;
            }
        }
    for (int c = 0; c < nrhs; ++c)
        for (int j = n; j >= 1; --j) {
            if (B(j, c) == 0.0) continue;
            B(j, c) = B(j, c) / AB(k + 1, j);
            const double t = B(j, c);
            const int lo = j - k > 1 ? j - k : 1;
            for (int i = j - 1; i >= lo; --i) B(i, c) = B(i, c) - t * AB(k + 1 + i - j, j);
        }
#undef AB
#undef B
}
