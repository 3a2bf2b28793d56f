// Angular spectral derivatives on the spectral (rmcontainer4d) layout: per m-block
// data(m:l_max, r, imi, field), l fastest.
//
// Spectral_Derivatives.F90: d_by_dtheta (sin(theta) d/dtheta) and d_by_sdtheta
// (1/sin(theta) d(sin^2(theta) .)/dtheta) couple l to l - 1 and l + 1 with the per-(m, l)
// coefficients of Initialize_Angular_Derivatives (:49-83); d_by_dphi multiplies by i m.
// All three are memory-bound stencils along the contiguous l axis of every (r, imi)
// column: one thread per output value, neighbours from the same 256-byte l-run (L1), the
// coefficients from a small per-plan table in L2.  Products and sums are rounded
// separately, as the reference's compiled loops round them, so the results are its doubles.
#include <algorithm>
#include <cmath>
#include <vector>

#include "raysht_internal.h"

namespace raysht {
namespace {

constexpr int kAThreads = 256;
constexpr int kAMaxBlocksX = 4096;

// out(l) of one field from field fin; grid (element blocks, m-block)
// Block mp of a packed array of F fields starts at lm_base[mp] * 2 nr * F (lm_base = the
// l count of the blocks before it); field f of the block at f * nl * 2 nr from there.
__global__ __launch_bounds__(kAThreads) void dtheta_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                           const int *__restrict__ lm_base,
                                                           const double2 *__restrict__ coef, const double *__restrict__ in,
                                                           int nf_in, int fin, double *__restrict__ out, int nf_out,
                                                           int fout)
{
#pragma clang fp contract(off)
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], ncol = 2L * nr, total = (long)nl * ncol;
    const double *a = in + lb * ncol * nf_in + (long)fin * total;
    double *b = out + lb * ncol * nf_out + (long)fout * total;
    const double2 *c = coef + lb;
    const float rnl = 1.0f / (float)nl;
    for (long e = blockIdx.x * (long)kAThreads + threadIdx.x; e < total; e += (long)gridDim.x * kAThreads) {
        // column k = e / nl from a float estimate and one exact correction (e < 2^31)
        int k = (int)((float)e * rnl);
        int i = (int)e - k * nl;  // l - m
        if (i < 0) {
            --k;
            i += nl;
        } else if (i >= nl) {
            ++k;
            i -= nl;
        }
        const double *ak = a + (long)k * nl;
        double v;
        if (nl == 1)
            v = 0.0;
        else if (i == 0)
            v = ak[1] * c[0].y;
        else if (i == nl - 1)
            v = ak[i - 1] * c[i].x;
        else
            v = ak[i - 1] * c[i].x + ak[i + 1] * c[i].y;
        b[(long)k * nl + i] = v;
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

int launch_angular(raysht_plan *p, int kind, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                   hipStream_t st)
{
    if (!p || !in || !out || kind < 0 || kind > 2) return RAYSHT_EINVAL;
    if (p->nr < 1 || p->nf < 1) return RAYSHT_EUNSUPPORTED;  // column plans have no fields
    if (nf_in < 1 || nf_out < 1 || fin < 0 || fin >= nf_in || fout < 0 || fout >= nf_out) return RAYSHT_EINVAL;
    const long per_field = p->s_off[p->nm] / p->nf;  // sum over m-blocks of nl * 2 nr
    const long n_in = per_field * nf_in, n_out = per_field * nf_out;
    const bool same = in == out && nf_in == nf_out;
    if (same && fin == fout && kind != 2) return RAYSHT_EINVAL;  // the l stencil cannot run in place
    if (!same && in < out + n_out && out < in + n_in) return RAYSHT_EINVAL;
    if (p->nm > 65535) return RAYSHT_EUNSUPPORTED;
    if ((long)(p->lmax + 1) * 2 * p->nr >= (1L << 24)) return RAYSHT_EUNSUPPORTED;  // one block's field < 2^24 values
    const long ncol = 2L * p->nr;
    int max_nl = 0;
    for (int m : p->mvals) max_nl = std::max(max_nl, p->lmax - m + 1);
    const long work = kind == 2 ? (long)max_nl * p->nr : (long)max_nl * ncol;
    const unsigned bx = (unsigned)std::min<long>((work + kAThreads - 1) / kAThreads, kAMaxBlocksX);
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

}  // extern "C"
