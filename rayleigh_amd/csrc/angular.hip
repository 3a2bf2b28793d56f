// Angular spectral derivatives on the spectral (rmcontainer4d) layout: per m-block
// data(m:l_max, r, imi, field), l fastest.
//
// Spectral_Derivatives.F90: d_by_dtheta (sin(theta) d/dtheta) and d_by_sdtheta
// (1/sin(theta) d(sin^2(theta) .)/dtheta) couple l to l - 1 and l + 1 with the per-(m, l)
// coefficients of Initialize_Angular_Derivatives (:49-83); d_by_dphi multiplies by i m.
// All three are memory-bound stencils along the contiguous l axis of every (r, imi)
// column: the l-stencils run one wave per l-column with the neighbours passed between
// lanes, the coefficients come from a small per-plan table in L2.  Products and sums are
// rounded separately, as the reference's compiled loops round them, so the results are its
// doubles.  The same file holds the rlm_spaceb combination built from these derivatives and
// the l(l+1) / radial scalings of the spectral fields.
#include <algorithm>
#include <cmath>
#include <vector>

#include "raysht_internal.h"

namespace raysht {
namespace {

constexpr int kAThreads = 256;
constexpr int kAMaxBlocksX = 4096;

// q = e / n for 0 <= e < 2^24 from a float estimate and one exact correction
__device__ __forceinline__ int div_small(int e, int n, float rn, int &rem)
{
    int q = (int)((float)e * rn);
    int i = e - q * n;
    if (i < 0) {
        --q;
        i += n;
    } else if (i >= n) {
        ++q;
        i -= n;
    }
    rem = i;
    return q;
}

// out(l) of one field from field fin; grid (column groups, m-block).  One wave per (r, imi)
// l-column, 64 values of l at a time: the l - 1 and l + 1 neighbours come from the
// adjacent lanes' loads (lanes 0 and 63 load theirs), so each value is loaded once.
// Block mp of a packed array of F fields starts at lm_base[mp] * 2 nr * F (lm_base = the
// l count of the blocks before it); field f of the block at f * nl * 2 nr from there.
__global__ __launch_bounds__(kAThreads) void dtheta_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                           const int *__restrict__ lm_base,
                                                           const double2 *__restrict__ coef, const double *__restrict__ in,
                                                           int nf_in, int fin, double *__restrict__ out, int nf_out,
                                                           int fout)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);
    if (k >= 2 * nr) return;  // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], ncol = 2L * nr, total = (long)nl * ncol;
    const double *ak = in + lb * ncol * nf_in + (long)fin * total + (long)k * nl;
    double *bk = out + lb * ncol * nf_out + (long)fout * total + (long)k * nl;
    const double2 *c = coef + lb;
    for (int base = 0; base < nl; base += 64) {
        const int i = base + lane;
        const bool inr = i < nl;
        double x = 0.0;
        double2 ci = make_double2(0.0, 0.0);
        if (inr) {
            x = ak[i];
            ci = c[i];
        }
        double xn = __shfl_down(x, 1), xp = __shfl_up(x, 1);
        if (lane == 63 && i + 1 < nl) xn = ak[i + 1];
        if (lane == 0 && i > 0) xp = ak[i - 1];
        if (!inr) continue;
        double v;
        if (nl == 1)
            v = 0.0;
        else if (i == 0)
            v = xn * ci.y;
        else if (i == nl - 1)
            v = xp * ci.x;
        else
            v = xp * ci.x + xn * ci.y;
        bk[i] = v;
    }
}

// out = i m in for one field: (re, im) -> (-(m im), m re); in place allowed
__global__ __launch_bounds__(kAThreads) void dphi_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                         const int *__restrict__ lm_base, const double *in, int nf_in,
                                                         int fin, double *out, int nf_out, int fout)
{
#pragma clang fp contract(off)
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], half = (long)nl * nr;
    const double *a = in + lb * 2L * nr * nf_in + (long)fin * 2 * half;
    double *b = out + lb * 2L * nr * nf_out + (long)fout * 2 * half;
    const double dm = (double)m;
    for (long e = blockIdx.x * (long)kAThreads + threadIdx.x; e < half; e += (long)gridDim.x * kAThreads) {
        const double re = a[e], im = a[half + e];
        b[e] = -(dm * im);
        b[half + e] = dm * re;
    }
}

// rlm_spaceb's hydro combination (Sphere_Hybrid_Space.F90:127-214), in place on fields
// W, P, Z of a buffer of nf fields, one pass:
//   W <- W / (l(l+1));  Z <- (d_by_dphi P - d_by_sdtheta Z) / (l(l+1));
//   P <- (d_by_sdtheta P + d_by_dphi Z) / (l(l+1));
// then l = 0 of W, P, Z zeroed for m = 0, and l = l_max of every field zeroed.
// One wave owns one radius of one m-block (its re and im l-columns of W, P, Z) and walks it
// 64 values of l at a time, so no other wave touches what it reads: the l + 1 neighbours
// come from the next lane (lane 63 loads the first value of the next chunk, not yet
// written), the l - 1 neighbours from the previous lane (lane 0 takes the last values of
// the previous chunk as loaded, carried in registers, since the wave has overwritten them).
__global__ __launch_bounds__(kAThreads) void rlm_spaceb_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                               const int *__restrict__ lm_base,
                                                               const double2 *__restrict__ scoef, double *buf, int nf,
                                                               int wv, int pv, int zv)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);
    if (r >= nr) return;  // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], half = (long)nl * nr, fsz = 2 * half;
    double *blk = buf + lb * 2L * nr * nf;
    double *W = blk + wv * fsz + (long)r * nl, *P = blk + pv * fsz + (long)r * nl, *Z = blk + zv * fsz + (long)r * nl;
    const double2 *c = scoef + lb;
    const double dm = (double)m;
    double cpr = 0.0, cpi = 0.0, czr = 0.0, czi = 0.0;  // l - 1 values for lane 0
    for (int base = 0; base < nl; base += 64) {
        const int i = base + lane, l = m + i;
        const bool in = i < nl;
        double wr = 0.0, wi = 0.0, pr = 0.0, pi = 0.0, zr = 0.0, zi = 0.0;
        double2 ci = make_double2(0.0, 0.0);
        if (in) {
            wr = W[i];
            wi = W[half + i];
            pr = P[i];
            pi = P[half + i];
            zr = Z[i];
            zi = Z[half + i];
            ci = c[i];
        }
        double prn = __shfl_down(pr, 1), pin = __shfl_down(pi, 1);  // l + 1
        double zrn = __shfl_down(zr, 1), zin = __shfl_down(zi, 1);
        if (lane == 63 && i + 1 < nl) {
            prn = P[i + 1];
            pin = P[half + i + 1];
            zrn = Z[i + 1];
            zin = Z[half + i + 1];
        }
        double prp = __shfl_up(pr, 1), pip = __shfl_up(pi, 1);  // l - 1
        double zrp = __shfl_up(zr, 1), zip = __shfl_up(zi, 1);
        if (lane == 0) {
            prp = cpr;
            pip = cpi;
            zrp = czr;
            zip = czi;
        }
        cpr = __shfl(pr, 63);
        cpi = __shfl(pi, 63);
        czr = __shfl(zr, 63);
        czi = __shfl(zi, 63);
        if (!in) continue;
        // d_by_sdtheta, rounded as dtheta_kernel rounds it
        double sdpr, sdpi, sdzr, sdzi;
        if (nl == 1) {
            sdpr = sdpi = sdzr = sdzi = 0.0;
        } else if (i == 0) {
            sdpr = prn * ci.y;
            sdpi = pin * ci.y;
            sdzr = zrn * ci.y;
            sdzi = zin * ci.y;
        } else if (i == nl - 1) {
            sdpr = prp * ci.x;
            sdpi = pip * ci.x;
            sdzr = zrp * ci.x;
            sdzi = zip * ci.x;
        } else {
            sdpr = prp * ci.x + prn * ci.y;
            sdpi = pip * ci.x + pin * ci.y;
            sdzr = zrp * ci.x + zrn * ci.y;
            sdzi = zip * ci.x + zin * ci.y;
        }
        const double ol = l == 0 ? 1.0 : 1.0 / ((double)l * ((double)l + 1.0));  // over_l_l_plus1
        const bool zero = l == lmax || l == 0;
        W[i] = zero ? 0.0 : wr * ol;
        W[half + i] = zero ? 0.0 : wi * ol;
        P[i] = zero ? 0.0 : (sdpr + -(dm * zi)) * ol;
        P[half + i] = zero ? 0.0 : (sdpi + dm * zr) * ol;
        Z[i] = zero ? 0.0 : (-(dm * pi) - sdzr) * ol;
        Z[half + i] = zero ? 0.0 : (dm * pr - sdzi) * ol;
        if (i == nl - 1)
            for (int f = 0; f < nf; ++f)
                if (f != wv && f != pv && f != zv) {
                    blk[f * fsz + (long)r * nl + i] = 0.0;
                    blk[f * fsz + half + (long)r * nl + i] = 0.0;
                }
    }
}

// out(l, r, imi) = (lfac(l) in(l, r, imi)) rfac(r) for one field (either factor may be
// absent); in place allowed.  The products round in this order, as the reference's
// l_l_plus1(m:l_max)*SBUFFA(IDX2, f)*Over_RhoRSQ(r) statements round them.  One wave per
// (r, imi) l-column: rfac(r) is one value per wave, lfac(m..l_max) a contiguous run.
__global__ __launch_bounds__(kAThreads) void lm_scale_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                             const int *__restrict__ lm_base, const double *in,
                                                             int nf_in, int fin, double *out, int nf_out, int fout,
                                                             const double *__restrict__ lfac,
                                                             const double *__restrict__ rfac)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);  // column r + nr imi
    if (k >= 2 * nr) return;                                            // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], total = 2L * nl * nr;
    const double *a = in + lb * 2L * nr * nf_in + (long)fin * total + (long)k * nl;
    double *b = out + lb * 2L * nr * nf_out + (long)fout * total + (long)k * nl;
    const double rf = rfac ? rfac[k >= nr ? k - nr : k] : 1.0;
    const double *lf = lfac ? lfac + m : nullptr;
    for (int i = lane; i < nl; i += 64) {
        double v = a[i];
        if (lf) v = lf[i] * v;
        if (rfac) v = v * rf;
        b[i] = v;
    }
}

// data(l_max, r, imi, f) = 0 for every m-block, radius, imi and field: one value per
// l-column, at nl (k + 1) - 1 for column k = r + nr (imi + 2 f)
__global__ __launch_bounds__(kAThreads) void zero_lmax_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                              const int *__restrict__ lm_base, double *buf, int nf)
{
    const int mp = blockIdx.y, nl = lmax - mvals[mp] + 1;
    const long ncol = 2L * nr * nf;
    double *blk = buf + lm_base[mp] * ncol;
    for (long k = blockIdx.x * (long)kAThreads + threadIdx.x; k < ncol; k += (long)gridDim.x * kAThreads)
        blk[k * nl + nl - 1] = 0.0;
}

// the checks every spectral-field entry point shares: a plan with fields and m-blocks that
// fit one launch, and one block's field below 2^24 values (div_small)
int check_spectral(const raysht_plan *p)
{
    if (p->nr < 1 || p->nf < 1) return RAYSHT_EUNSUPPORTED;  // column plans have no fields
    if (p->nm > 65535) return RAYSHT_EUNSUPPORTED;
    if ((long)(p->lmax + 1) * 2 * p->nr >= (1L << 24)) return RAYSHT_EUNSUPPORTED;
    return RAYSHT_OK;
}

int max_nl(const raysht_plan *p)
{
    int n = 0;
    for (int m : p->mvals) n = std::max(n, p->lmax - m + 1);
    return n;
}

int launch_angular(raysht_plan *p, int kind, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                   hipStream_t st)
{
    if (!p || !in || !out || kind < 0 || kind > 2) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    if (nf_in < 1 || nf_out < 1 || fin < 0 || fin >= nf_in || fout < 0 || fout >= nf_out) return RAYSHT_EINVAL;
    const long per_field = p->s_off[p->nm] / p->nf;  // sum over m-blocks of nl * 2 nr
    const long n_in = per_field * nf_in, n_out = per_field * nf_out;
    const bool same = in == out && nf_in == nf_out;
    if (same && fin == fout && kind != 2) return RAYSHT_EINVAL;  // the l stencil cannot run in place
    if (!same && in < out + n_out && out < in + n_in) return RAYSHT_EINVAL;
    const long work = kind == 2 ? (long)max_nl(p) * p->nr : 2L * p->nr * 64;  // dtheta: a wave per l-column
    const unsigned bx = (unsigned)std::min<long>((work + kAThreads - 1) / kAThreads, kind == 2 ? kAMaxBlocksX : 1L << 30);
    dim3 grid(bx, (unsigned)p->nm);
    if (kind == 2)
        hipLaunchKernelGGL(dphi_kernel, grid, dim3(kAThreads), 0, st, p->lmax, p->nr, p->d_mvals, p->d_lm_base, in,
                           nf_in, fin, out, nf_out, fout);
    else
        hipLaunchKernelGGL(dtheta_kernel, grid, dim3(kAThreads), 0, st, p->lmax, p->nr, p->d_mvals, p->d_lm_base,
                           p->d_acoef + (kind == 1 ? p->lm_total : 0), in, nf_in, fin, out, nf_out, fout);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_rlm_spaceb(raysht_plan *p, double *buf, int nf, int wv, int pv, int zv, hipStream_t st)
{
    if (!p || !buf) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    if (nf < 3 || wv < 0 || pv < 0 || zv < 0 || wv >= nf || pv >= nf || zv >= nf || wv == pv || wv == zv || pv == zv)
        return RAYSHT_EINVAL;
    const unsigned bx = (unsigned)((p->nr + kAThreads / 64 - 1) / (kAThreads / 64));  // a wave per radius
    hipLaunchKernelGGL(rlm_spaceb_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, p->d_acoef + p->lm_total, buf, nf, wv, pv, zv);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_lm_scale(raysht_plan *p, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                    const double *lfac, const double *rfac, hipStream_t st)
{
    if (!p || !in || !out) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    if (nf_in < 1 || nf_out < 1 || fin < 0 || fin >= nf_in || fout < 0 || fout >= nf_out) return RAYSHT_EINVAL;
    const long per_field = p->s_off[p->nm] / p->nf;  // elementwise: in place, or another field of one buffer
    const bool same = in == out && nf_in == nf_out;
    if (!same && in < out + per_field * nf_out && out < in + per_field * nf_in) return RAYSHT_EINVAL;
    const unsigned bx = (unsigned)((2L * p->nr + kAThreads / 64 - 1) / (kAThreads / 64));  // a wave per l-column
    hipLaunchKernelGGL(lm_scale_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, in, nf_in, fin, out, nf_out, fout, lfac, rfac);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_zero_lmax(raysht_plan *p, double *buf, int nf, hipStream_t st)
{
    if (!p || !buf || nf < 1) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    const unsigned bx = (unsigned)std::min<long>((2L * p->nr * nf + kAThreads - 1) / kAThreads, kAMaxBlocksX);
    hipLaunchKernelGGL(zero_lmax_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, buf, nf);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

}  // namespace

// Initialize_Angular_Derivatives (Spectral_Derivatives.F90:49-83) for the plan's m-blocks,
// in its operation order: (c1, c2) per (m, l) for d_by_dtheta, then for d_by_sdtheta, each
// indexed like lm_base (the l count before block mp) + l - m.
int angular_plan_init(raysht_plan *p)
{
    long total = 0;
    for (int m : p->mvals) total += p->lmax - m + 1;
    p->lm_total = total;
    std::vector<double2> c(2 * total);
    long k = 0;
    for (int mp = 0; mp < p->nm; ++mp) {
        const int m = p->mvals[mp];
        for (int l = m; l <= p->lmax; ++l, ++k) {
            const double q1 = std::sqrt((double)((l + m) * (l - m)) / (double)((2 * l + 1) * (2 * l - 1)));
            const double q2 = std::sqrt((double)((l + 1 + m) * (l + 1 - m)) / (double)((2 * l + 1) * (2 * l + 3)));
            c[k] = make_double2((double)(l - 1) * q1, -((double)(l + 2) * q2));
            c[total + k] = make_double2((double)(l + 1) * q1, -((double)l * q2));
        }
    }
    p->d_acoef = nullptr;
    if (c.empty()) return RAYSHT_OK;
    if (hipMalloc(reinterpret_cast<void **>(&p->d_acoef), c.size() * sizeof(double2)) != hipSuccess) {
        p->d_acoef = nullptr;
        return RAYSHT_ENOMEM;
    }
    RAYSHT_HIP_TRY(hipMemcpy(p->d_acoef, c.data(), c.size() * sizeof(double2), hipMemcpyHostToDevice));
    return RAYSHT_OK;
}

void angular_plan_free(raysht_plan *p)
{
    if (p->d_acoef) (void)hipFree(p->d_acoef);
    p->d_acoef = nullptr;
}

}  // namespace raysht

using namespace raysht;

extern "C" {

int raysht_d_by_dtheta(raysht_plan *p, int kind, const double *in, int nf_in, int fin, double *out, int nf_out,
                       int fout, void *stream)
{
    if (kind != 0 && kind != 1) return RAYSHT_EINVAL;
    return launch_angular(p, kind, in, nf_in, fin, out, nf_out, fout, reinterpret_cast<hipStream_t>(stream));
}

int raysht_d_by_dphi(raysht_plan *p, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                     void *stream)
{
    return launch_angular(p, 2, in, nf_in, fin, out, nf_out, fout, reinterpret_cast<hipStream_t>(stream));
}

int raysht_rlm_spaceb_hydro(raysht_plan *p, double *buf, int nf, int wvar, int pvar, int zvar, void *stream)
{
    return launch_rlm_spaceb(p, buf, nf, wvar, pvar, zvar, reinterpret_cast<hipStream_t>(stream));
}

int raysht_lm_scale(raysht_plan *p, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                    const double *lfac, const double *rfac, void *stream)
{
    return launch_lm_scale(p, in, nf_in, fin, out, nf_out, fout, lfac, rfac, reinterpret_cast<hipStream_t>(stream));
}

int raysht_zero_lmax(raysht_plan *p, double *buf, int nf, void *stream)
{
    return launch_zero_lmax(p, buf, nf, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
