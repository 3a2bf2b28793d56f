// The one-rank radial transposes between the spectral layout (the Legendre side: per
// m-block data(m:l_max, r, imi, f), l fastest) and the radial layout (the implicit
// solve and radial derivative side: p1(r, imi, lm, f), r fastest).
//
// Transpose_2b1b / Transpose_1a2a (Spherical_Buffer.F90:705-932) pack the (l, m)
// pairs of a radial group in the order LM_Load_Balance (Load_Balance.F90:311-395)
// lists them: l slowest from the smallest owned m, the owned m in m_values order
// within one l.  On one rank the all-to-all between them is the identity, so each
// transpose is a gather / scatter: every (m-block, l) row of N_r radii becomes the
// radial column of its pair.  The plan holds the pair index of every (m-block, l)
// and a list of 32-l tiles; a workgroup moves a 32 (l) x 64 (r) tile of one
// (field, imi) through LDS, reading 256-byte l-runs and writing 512-byte r-runs.
#include <algorithm>
#include <vector>

#include "raysht_internal.h"

namespace raysht {
namespace {

constexpr int kRL = 32, kRR = 64, kRThreads = 256;
static_assert(kRThreads % kRL == 0 && kRThreads % kRR == 0, "radial transpose tile");

// spectral -> radial (2b -> 1b)
__global__ __launch_bounds__(kRThreads) void permute_2b1b_kernel(int nr, int lmax, int lm_count,
                                                                 const int2 *__restrict__ tiles,
                                                                 const int *__restrict__ mvals,
                                                                 const long *__restrict__ s_off,
                                                                 const int *__restrict__ lm_base,
                                                                 const int *__restrict__ lm_slot,
                                                                 const double *__restrict__ s, double *__restrict__ p1)
{
    __shared__ double tile[kRR][kRL + 1];  // [r][l]
    const int2 t = tiles[blockIdx.x];
    const int mp = t.x, l0 = t.y;  // l0: l - m of the tile's first row
    const int nl = lmax - mvals[mp] + 1;
    const int r0 = blockIdx.y * kRR, fi = blockIdx.z, f = fi >> 1, imi = fi & 1;
    const double *src = s + s_off[mp] + (long)fi * nr * nl;
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < kRR / (kRThreads / kRL); ++i) {
        const int rr = tid / kRL + (kRThreads / kRL) * i, ll = tid % kRL, r = r0 + rr, l = l0 + ll;
        if (r < nr && l < nl) tile[rr][ll] = src[(long)r * nl + l];
    }
    __syncthreads();
    const int *slot = lm_slot + lm_base[mp];
#pragma unroll
    for (int i = 0; i < kRL / (kRThreads / kRR); ++i) {
        const int ll = tid / kRR + (kRThreads / kRR) * i, rr = tid % kRR, r = r0 + rr, l = l0 + ll;
        if (r < nr && l < nl) p1[(((long)f * lm_count + slot[l]) * 2 + imi) * nr + r] = tile[rr][ll];
    }
}

// radial -> spectral (1a -> 2a)
__global__ __launch_bounds__(kRThreads) void permute_1a2a_kernel(int nr, int lmax, int lm_count,
                                                                 const int2 *__restrict__ tiles,
                                                                 const int *__restrict__ mvals,
                                                                 const long *__restrict__ s_off,
                                                                 const int *__restrict__ lm_base,
                                                                 const int *__restrict__ lm_slot,
                                                                 const double *__restrict__ p1, double *__restrict__ s)
{
    __shared__ double tile[kRR][kRL + 1];  // [r][l]
    const int2 t = tiles[blockIdx.x];
    const int mp = t.x, l0 = t.y;
    const int nl = lmax - mvals[mp] + 1;
    const int r0 = blockIdx.y * kRR, fi = blockIdx.z, f = fi >> 1, imi = fi & 1;
    const int tid = threadIdx.x;
    const int *slot = lm_slot + lm_base[mp];
#pragma unroll
    for (int i = 0; i < kRL / (kRThreads / kRR); ++i) {
        const int ll = tid / kRR + (kRThreads / kRR) * i, rr = tid % kRR, r = r0 + rr, l = l0 + ll;
        if (r < nr && l < nl) tile[rr][ll] = p1[(((long)f * lm_count + slot[l]) * 2 + imi) * nr + r];
    }
    __syncthreads();
    double *dst = s + s_off[mp] + (long)fi * nr * nl;
#pragma unroll
    for (int i = 0; i < kRR / (kRThreads / kRL); ++i) {
        const int rr = tid / kRL + (kRThreads / kRL) * i, ll = tid % kRL, r = r0 + rr, l = l0 + ll;
        if (r < nr && l < nl) dst[(long)r * nl + l] = tile[rr][ll];
    }
}

template <typename T>
int upload(T **dst, const std::vector<T> &v)
{
    *dst = nullptr;
    if (v.empty()) return RAYSHT_OK;
    if (hipMalloc(reinterpret_cast<void **>(dst), v.size() * sizeof(T)) != hipSuccess) {
        *dst = nullptr;
        return RAYSHT_ENOMEM;
    }
    RAYSHT_HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return RAYSHT_OK;
}

int launch_radial(raysht_plan *p, bool to_radial, const double *in, double *out, hipStream_t st)
{
    if (!p || !in || !out) return RAYSHT_EINVAL;
    if (p->nr < 1 || p->nf < 1) return RAYSHT_EUNSUPPORTED;  // column plans have no radii
    const long n = p->s_off[p->nm];
    if (in < out + n && out < in + n) return RAYSHT_EINVAL;
    if (p->n_rtiles == 0) return RAYSHT_OK;
    if (2L * p->nf > 65535 || (p->nr + kRR - 1) / kRR > 65535) return RAYSHT_EUNSUPPORTED;
    dim3 grid((unsigned)p->n_rtiles, (unsigned)((p->nr + kRR - 1) / kRR), (unsigned)(2 * p->nf));
    if (to_radial)
        hipLaunchKernelGGL(permute_2b1b_kernel, grid, dim3(kRThreads), 0, st, p->nr, p->lmax, p->lm_count,
                           p->d_rtiles, p->d_mvals, p->d_s_off, p->d_lm_base, p->d_lm_slot, in, out);
    else
        hipLaunchKernelGGL(permute_1a2a_kernel, grid, dim3(kRThreads), 0, st, p->nr, p->lmax, p->lm_count,
                           p->d_rtiles, p->d_mvals, p->d_s_off, p->d_lm_base, p->d_lm_slot, in, out);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

}  // namespace

// LM_Load_Balance's pair list for the plan's m values (one radial rank), the pair
// index of every (m-block, l) and the transposes' tile list.
int radial_plan_init(raysht_plan *p)
{
    const int nm = p->nm, L = p->lmax;
    int lmin = L;
    for (int m : p->mvals) lmin = std::min(lmin, m);
    std::vector<int> base(nm + 1, 0);
    for (int mp = 0; mp < nm; ++mp) base[mp + 1] = base[mp] + (L - p->mvals[mp] + 1);
    std::vector<int> slot(base[nm]);
    p->lm_l.clear();
    p->lm_mp.clear();
    for (int l = lmin; l <= L; ++l)
        for (int mp = 0; mp < nm; ++mp) {
            const int m = p->mvals[mp];
            if (m <= l) {
                slot[base[mp] + (l - m)] = (int)p->lm_l.size();
                p->lm_l.push_back(l);
                p->lm_mp.push_back(mp);
            }
        }
    p->lm_count = (int)p->lm_l.size();
    std::vector<int2> tiles;
    for (int mp = 0; mp < nm; ++mp)
        for (int l0 = 0; l0 < L - p->mvals[mp] + 1; l0 += kRL) tiles.push_back(make_int2(mp, l0));
    p->n_rtiles = (int)tiles.size();
    int rc;
    if ((rc = upload(&p->d_lm_base, base)) || (rc = upload(&p->d_lm_slot, slot)) || (rc = upload(&p->d_rtiles, tiles)))
        return rc;
    return RAYSHT_OK;
}

void radial_plan_free(raysht_plan *p)
{
    for (void *d : {(void *)p->d_lm_base, (void *)p->d_lm_slot, (void *)p->d_rtiles})
        if (d) (void)hipFree(d);
    p->d_lm_base = p->d_lm_slot = nullptr;
    p->d_rtiles = nullptr;
}

}  // namespace raysht

using namespace raysht;

extern "C" {

int raysht_permute_2b1b(raysht_plan *p, const double *s2b, double *p1b, void *stream)
{
    return launch_radial(p, true, s2b, p1b, reinterpret_cast<hipStream_t>(stream));
}

int raysht_permute_1a2a(raysht_plan *p, const double *p1a, double *s2a, void *stream)
{
    return launch_radial(p, false, p1a, s2a, reinterpret_cast<hipStream_t>(stream));
}

int raysht_lm_count(const raysht_plan *p) { return p ? p->lm_count : 0; }

int raysht_lm_values(const raysht_plan *p, int *l_out, int *mp_out)
{
    if (!p || (p->lm_count > 0 && (!l_out || !mp_out))) return RAYSHT_EINVAL;
    std::copy(p->lm_l.begin(), p->lm_l.end(), l_out);
    std::copy(p->lm_mp.begin(), p->lm_mp.end(), mp_out);
    return RAYSHT_OK;
}

}  // extern "C"
