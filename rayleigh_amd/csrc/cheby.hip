// Chebyshev radial transforms on the zeros grid (one domain).
//
// Cheby_Grid%to_spectral / %from_spectral (Chebyshev_Polynomials.F90:375-558)
// fold each radial column f(i) +/- f(n+1-i), contract the two halves with the
// parity tables T_2k(x_i), T_2k+1(x_i) (one DGEMM per parity over all columns)
// and interleave even/odd coefficients.  A column is n <= 512 doubles, so one
// workgroup stages a block of columns in LDS, keeps the folded halves there and
// forms every output element of the block from LDS and the (L2-resident) tables.
// The tables are generated on the host exactly as Gen_Tn (:251-297) does -- angle
// by repeated addition, libm cos -- so they are the reference's doubles.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>
#include <vector>

#include "raysht_internal.h"
#include "small_dft.h"

namespace raysht {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double4_t mfma_f64(double a, double b, double4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kChebMaxN = 1024;
constexpr int kChebKB = 4;  // 16-row output blocks per accumulator group

// Both contractions are small dense products per column (K <= n/2) over a huge
// number of columns, so a workgroup stages a block of columns in LDS (folded or
// de-interleaved on the way in) and multiplies it with the parity tables on FP64
// MFMA (v_mfma_f64_16x16x4_f64; result rows interleaved by 4: row = lane/16 + 4v).
// The products are taken transposed (columns on the MFMA row axis) so each store
// instruction writes 16 consecutive radii/coefficients of a column.
// Each wave owns 16 columns; tc = 16 x waves columns per workgroup.

// To_Spectral4D: c(2k) = (2/n) sum_i T_2k(x_i) fe(i), c(2k+1) = (2/n) sum_i T_2k+1(x_i) fo(i).
// Tables tt_e/tt_o: [i][k] (k fastest), i < ix4 (n_x rounded up to 4), k < kp (rounded up to 16).
__global__ __launch_bounds__(256) void cheb_to_spectral_kernel(int n, int n_x, int n_even, int n_odd, int ix4, int kp,
                                                               long ncols, const double *__restrict__ tt_e,
                                                               const double *__restrict__ tt_o,
                                                               const double *__restrict__ f, double *__restrict__ c)
{
    extern __shared__ double lds[];
    const int tc = blockDim.x / 4;  // 16 columns per wave
    const int ld = ix4 + 1;
    double *fe = lds, *fo = lds + tc * ld;
    const long col0 = (long)blockIdx.x * tc;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lg = lane >> 4;
    for (int idx = tid; idx < tc * ix4; idx += blockDim.x) {
        const int cc = idx / ix4, i = idx % ix4;
        const long col = col0 + cc;
        double e = 0.0, o = 0.0;
        if (col < ncols && i < n_x) {
            const double a = f[col * n + i], b = f[col * n + (n - 1 - i)];
            e = a + b;
            o = a - b;
        }
        fe[cc * ld + i] = e;
        fo[cc * ld + i] = o;
    }
    __syncthreads();
    const double alpha = 2.0 / (double)n;
    for (int kb0 = 0; kb0 < kp / 16; kb0 += kChebKB) {
        double4_t acc_e[kChebKB], acc_o[kChebKB];
#pragma unroll
        for (int b = 0; b < kChebKB; ++b) {
            acc_e[b] = double4_t{0.0, 0.0, 0.0, 0.0};
            acc_o[b] = double4_t{0.0, 0.0, 0.0, 0.0};
        }
        const int nb = (kp / 16 - kb0) < kChebKB ? (kp / 16 - kb0) : kChebKB;
        for (int ii = 0; ii < ix4; ii += 4) {
            const double a_e = fe[(16 * wave + li) * ld + ii + lg];
            const double a_o = fo[(16 * wave + li) * ld + ii + lg];
            const double *te = tt_e + (long)(ii + lg) * kp + 16 * kb0 + li;
            const double *to = tt_o + (long)(ii + lg) * kp + 16 * kb0 + li;
#pragma unroll
            for (int b = 0; b < kChebKB; ++b) {
                if (b < nb) {
                    acc_e[b] = mfma_f64(a_e, te[16 * b], acc_e[b]);
                    acc_o[b] = mfma_f64(a_o, to[16 * b], acc_o[b]);
                }
            }
        }
#pragma unroll
        for (int b = 0; b < kChebKB; ++b) {
            if (b >= nb) break;
            const int k = 16 * (kb0 + b) + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long col = col0 + 16 * wave + lg + 4 * v;
                if (col < ncols) {
                    if (k < n_even) c[col * n + 2 * k] = alpha * acc_e[b][v];
                    if (k < n_odd) c[col * n + 2 * k + 1] = alpha * acc_o[b][v];
                }
            }
        }
    }
}

// From_Spectral4D: E(i) = sum_k T_2k(x_i) c(2k), O(i) = sum_k T_2k+1(x_i) c(2k+1);
// f(i) = E + O, f(n-1-i) = E - O, the x = 0 row of odd n doubled back, then - c(0)/2.
// Tables t_e/t_o: [k][i] (i fastest), k < k4 (rounded up to 4), i < ip (n_x rounded up to 16).
__global__ __launch_bounds__(256) void cheb_from_spectral_kernel(int n, int n_x, int n_even, int n_odd, int k4, int ip,
                                                                 long ncols, const double *__restrict__ t_e,
                                                                 const double *__restrict__ t_o,
                                                                 const double *__restrict__ c, double *__restrict__ f)
{
    extern __shared__ double lds[];
    const int tc = blockDim.x / 4;
    const int ld = k4 + 1;
    double *ce = lds, *co = lds + tc * ld;
    const long col0 = (long)blockIdx.x * tc;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lg = lane >> 4;
    for (int idx = tid; idx < tc * 2 * k4; idx += blockDim.x) {
        const int cc = idx / (2 * k4), q = idx % (2 * k4);
        const long col = col0 + cc;
        const int k = q >> 1;
        double v = 0.0;
        if (col < ncols && q < n) v = c[col * n + q];
        if (q & 1)
            co[cc * ld + k] = v;
        else
            ce[cc * ld + k] = v;
    }
    __syncthreads();
    const bool odd_n = n_even != n_odd;
    for (int ib0 = 0; ib0 < ip / 16; ib0 += kChebKB) {
        double4_t acc_e[kChebKB], acc_o[kChebKB];
#pragma unroll
        for (int b = 0; b < kChebKB; ++b) {
            acc_e[b] = double4_t{0.0, 0.0, 0.0, 0.0};
            acc_o[b] = double4_t{0.0, 0.0, 0.0, 0.0};
        }
        const int nb = (ip / 16 - ib0) < kChebKB ? (ip / 16 - ib0) : kChebKB;
        for (int kk = 0; kk < k4; kk += 4) {
            const double a_e = ce[(16 * wave + li) * ld + kk + lg];
            const double a_o = co[(16 * wave + li) * ld + kk + lg];
            const double *te = t_e + (long)(kk + lg) * ip + 16 * ib0 + li;
            const double *to = t_o + (long)(kk + lg) * ip + 16 * ib0 + li;
#pragma unroll
            for (int b = 0; b < kChebKB; ++b) {
                if (b < nb) {
                    acc_e[b] = mfma_f64(a_e, te[16 * b], acc_e[b]);
                    acc_o[b] = mfma_f64(a_o, to[16 * b], acc_o[b]);
                }
            }
        }
#pragma unroll
        for (int b = 0; b < kChebKB; ++b) {
            if (b >= nb) break;
            const int i = 16 * (ib0 + b) + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int cc = 16 * wave + lg + 4 * v;
                const long col = col0 + cc;
                if (col < ncols && i < n_x) {
                    const double e = acc_e[b][v], o = acc_o[b][v];
                    const double c0h = ce[cc * ld] / 2.0;
                    double *fc = f + col * n;
                    if (odd_n && i == n_x - 1) {
                        fc[i] = (e * 2.0 + o * 2.0) - c0h;
                    } else {
                        fc[i] = (e + o) - c0h;
                        fc[n - 1 - i] = (e - o) - c0h;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// n <= 256: the parity tables live in registers.  One wave per 16 output rows holds
// its rows' table entries for every K (KB steps of 4 per parity: the B operand of
// v_mfma_f64_16x16x4_f64, lane (k = 4 kk + lane/16, row = lane%16)), so a persistent
// workgroup reads the tables once and then only streams columns: tiles of 32 columns
// (32, or 16 for n > 128) are loaded into registers one tile ahead, staged in LDS
// (folded / de-interleaved) and multiplied as 16-column A operands.  Output rows ride the MFMA column axis,
// so every store walks 16 consecutive radii or coefficients of a column.


// floor(a / b) for 0 <= a < 2^15, 1 <= b <= 512 through float: the rounding error of
// (a + 0.5) * (1/b) stays below the 0.5/b distance to the next integer
__device__ __forceinline__ int small_div(int a, float inv_b) { return (int)(((float)a + 0.5f) * inv_b); }

// this lane's B operand entries of a [k][row] table (row stride ld): k = 4 kk + lane/16,
// row = row0 + lane%16; zero past kmax rows or ld columns
template <int KB>
__device__ __forceinline__ void load_table_regs(const double *__restrict__ t, int ld, int kmax, int row0, int lane,
                                                double (&reg)[KB])
{
    const int row = row0 + (lane & 15);
#pragma unroll
    for (int kk = 0; kk < KB; ++kk) {
        const int k = 4 * kk + (lane >> 4);
        reg[kk] = (k < kmax && row < ld) ? t[(long)k * ld + row] : 0.0;
    }
}

// From_Spectral4D with K = coefficients (KB * 4 >= k4), rows = radii i < n_x.  Columns are
// n apart, or ld apart (SD: one domain's rows of a grid of several).
template <int KB, int NW, int TC, bool SD>
__global__ __launch_bounds__(64 * NW) void cheb_from_spectral_reg_kernel(int n, int n_x, int n_even, int k4, int ip,
                                                                     long ncols, long ld, const double *__restrict__ t_e,
                                                                     const double *__restrict__ t_o,
                                                                     const double *__restrict__ c, double *__restrict__ f)
{
    extern __shared__ double lds[];
    constexpr int LD = 4 * KB + 1;  // odd row stride: the A-operand reads hit 16 banks
    double *ce = lds, *co = lds + TC * LD;
    constexpr int nthr = 64 * NW;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    double te[KB], to[KB];
    load_table_regs<KB>(t_e, ip, k4, 16 * wave, lane, te);
    load_table_regs<KB>(t_o, ip, k4, 16 * wave, lane, to);
    const bool odd_n = (n & 1) != 0;
    const float inv_n = 1.0f / (float)n;
    const long ntiles = (ncols + TC - 1) / TC;
    const int per = TC * n;  // doubles per tile
    constexpr int kMaxPer = TC * 2 * 4 * KB;  // per = 32 n <= 32 * 8 KB
    constexpr int kU = (kMaxPer + nthr - 1) / nthr;
    double pre[kU];
    auto fetch = [&](long tile) {
        if constexpr (SD) {
            const double *src = c + tile * TC * ld;
            const long left = ncols - tile * TC;  // columns
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int idx = tid + u * nthr;
                const int col = small_div(idx, inv_n), q = idx - col * n;
                pre[u] = (idx < per && col < left) ? src[col * ld + q] : 0.0;
            }
        } else {
            const double *src = c + tile * TC * (long)n;
            const long left = (ncols - tile * TC) * (long)n;
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int idx = tid + u * nthr;
                pre[u] = (idx < per && idx < left) ? src[idx] : 0.0;
            }
        }
    };
    long tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    for (; tile < ntiles; tile += gridDim.x) {
        __syncthreads();  // the previous tile's A reads are done
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int idx = tid + u * nthr;
            if (idx < per) {
                const int col = small_div(idx, inv_n), q = idx - col * n;
                (q & 1 ? co : ce)[col * LD + (q >> 1)] = pre[u];
            }
        }
        // zero K padding (coefficients past n in the last parity row of the table)
        for (int idx = tid; idx < TC * (4 * KB); idx += nthr) {
            const int col = idx / (4 * KB), k = idx % (4 * KB);
            if (2 * k >= n) ce[col * LD + k] = 0.0;
            if (2 * k + 1 >= n) co[col * LD + k] = 0.0;
        }
        __syncthreads();
        if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x);
        double4_t acc_e[TC / 16], acc_o[TC / 16];
#pragma unroll
        for (int cb = 0; cb < TC / 16; ++cb) acc_e[cb] = acc_o[cb] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < KB; ++kk)
#pragma unroll
            for (int cb = 0; cb < TC / 16; ++cb) {
                const int at = (16 * cb + li) * LD + 4 * kk + lg;
                acc_e[cb] = mfma_f64(ce[at], te[kk], acc_e[cb]);
                acc_o[cb] = mfma_f64(co[at], to[kk], acc_o[cb]);
            }
        const int i = 16 * wave + li;
        if (i < n_x) {
#pragma unroll
            for (int cb = 0; cb < TC / 16; ++cb)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int cl = 16 * cb + lg + 4 * v;
                    const long col = tile * TC + cl;
                    if (col < ncols) {
                        const double e = acc_e[cb][v], o = acc_o[cb][v], c0h = ce[cl * LD] / 2.0;
                        double *fc = f + col * (SD ? ld : (long)n);
                        if (odd_n && i == n_x - 1) {
                            fc[i] = (e * 2.0 + o * 2.0) - c0h;
                        } else {
                            fc[i] = (e + o) - c0h;
                            fc[n - 1 - i] = (e - o) - c0h;
                        }
                    }
                }
        }
    }
}

// To_Spectral4D with K = radii i < n_x (KB * 4 >= ix4), rows = coefficient pairs k; columns
// n or (SD) ld apart.
template <int KB, int NW, int TC, bool SD>
__global__ __launch_bounds__(64 * NW) void cheb_to_spectral_reg_kernel(int n, int n_x, int n_even, int n_odd, int ix4,
                                                                   int kp, long ncols, long ld,
                                                                   const double *__restrict__ tt_e,
                                                                   const double *__restrict__ tt_o,
                                                                   const double *__restrict__ f, double *__restrict__ c)
{
    extern __shared__ double lds[];
    constexpr int LD = 4 * KB + 1;
    double *fe = lds, *fo = lds + TC * LD;
    constexpr int nthr = 64 * NW;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    double te[KB], to[KB];
    load_table_regs<KB>(tt_e, kp, ix4, 16 * wave, lane, te);
    load_table_regs<KB>(tt_o, kp, ix4, 16 * wave, lane, to);
    const float inv_nx = 1.0f / (float)n_x;
    const long ntiles = (ncols + TC - 1) / TC;
    const int per = TC * n_x;  // folded pairs per tile
    const long cs = SD ? ld : (long)n;
    constexpr int kMaxPer = TC * 4 * KB;  // per = 32 n_x <= 32 * 4 KB
    constexpr int kU = (kMaxPer + nthr - 1) / nthr;
    double pa[kU], pb[kU];
    auto fetch = [&](long tile) {
        const double *src = f + tile * TC * cs;
        const long cols_left = ncols - tile * TC;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int idx = tid + u * nthr;
            const int col = small_div(idx, inv_nx), i = idx - col * n_x;
            const bool ok = idx < per && col < cols_left;
            pa[u] = ok ? src[(long)col * cs + i] : 0.0;
            pb[u] = ok ? src[(long)col * cs + (n - 1 - i)] : 0.0;
        }
    };
    const double alpha = 2.0 / (double)n;
    long tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    for (; tile < ntiles; tile += gridDim.x) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int idx = tid + u * nthr;
            if (idx < per) {
                const int col = small_div(idx, inv_nx), i = idx - col * n_x;
                fe[col * LD + i] = pa[u] + pb[u];
                fo[col * LD + i] = pa[u] - pb[u];
            }
        }
        for (int idx = tid; idx < TC * (4 * KB); idx += nthr) {
            const int col = idx / (4 * KB), i = idx % (4 * KB);
            if (i >= n_x) fe[col * LD + i] = fo[col * LD + i] = 0.0;
        }
        __syncthreads();
        if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x);
        double4_t acc_e[TC / 16], acc_o[TC / 16];
#pragma unroll
        for (int cb = 0; cb < TC / 16; ++cb) acc_e[cb] = acc_o[cb] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < KB; ++kk)
#pragma unroll
            for (int cb = 0; cb < TC / 16; ++cb) {
                const int at = (16 * cb + li) * LD + 4 * kk + lg;
                acc_e[cb] = mfma_f64(fe[at], te[kk], acc_e[cb]);
                acc_o[cb] = mfma_f64(fo[at], to[kk], acc_o[cb]);
            }
        const int k = 16 * wave + li;
#pragma unroll
        for (int cb = 0; cb < TC / 16; ++cb)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long col = tile * TC + 16 * cb + lg + 4 * v;
                if (col < ncols) {
                    if (k < n_even) c[col * cs + 2 * k] = alpha * acc_e[cb][v];
                    if (k < n_odd) c[col * cs + 2 * k + 1] = alpha * acc_o[cb][v];
                }
            }
    }
}

__global__ void cheb_dealias_kernel(int n, int first, long ncols, long ld, double *__restrict__ c)
{
    const int width = n - first;
    const long total = ncols * width;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long col = i / width;
        c[col * ld + first + (int)(i % width)] = 0.0;
    }
}

// Cheby_Deriv_Buffer_4D / Cheby_Deriv_OutOf_Buffer (Chebyshev_Polynomials.F90:560-730):
// per column the backward recursion b(i) = b(i+2) + 2(i+1) a(i+1) from b(n) = b(n+1) = 0,
// applied dorder times (each time to the previous result), then every entry times
// deriv_scaling**dorder.  The orders are chained in one descending pass: order o at i
// needs order o-1 at i+1, which the pass produced one step earlier, so a thread carries
// the two latest values of every order in registers and reads each input once.
// A workgroup owns 256 columns and streams them from the top coefficient down in
// chunks of 16: a chunk comes in as coalesced 128-byte column pieces (loaded into
// registers one chunk ahead) and goes to LDS (rows padded to 17 doubles), one thread
// per column runs the recursion over it in place, and the chunk leaves the same way.  Every step rounds the product and the sum
// separately, as the reference's compiled recursion does.
constexpr int kDerivCols = 256, kDerivChunk = 16, kDerivMaxOrder = 4;

__global__ __launch_bounds__(kDerivCols) void cheb_deriv_kernel(int n, long ld, long ncols, int dorder, double scale,
                                                                const double *__restrict__ c, double *__restrict__ d)
{
#pragma clang fp contract(off)
    __shared__ double tile[kDerivCols][kDerivChunk + 1];
    const int tid = threadIdx.x;
    const long col0 = (long)blockIdx.x * kDerivCols;
    const int cols = (int)min((long)kDerivCols, ncols - col0);
    double x1[kDerivMaxOrder], x2[kDerivMaxOrder];  // order o + 1 at i + 1 and at i + 2
#pragma unroll
    for (int o = 0; o < kDerivMaxOrder; ++o) x1[o] = x2[o] = 0.0;
    double a1 = 0.0;  // the input at i + 1
    // chunk values in flight: loaded one chunk ahead, while the current one is worked on
    double pf[kDerivChunk];
    auto fetch = [&](int i0) {
        const int w = min(kDerivChunk, n - i0);
#pragma unroll
        for (int q = 0; q < kDerivChunk; ++q) {
            const int v = q * kDerivCols + tid, cl = v / kDerivChunk, e = v % kDerivChunk;
            pf[q] = cl < cols && e < w ? c[(col0 + cl) * ld + i0 + e] : 0.0;
        }
    };
    fetch(((n - 1) / kDerivChunk) * kDerivChunk);
    for (int i0 = ((n - 1) / kDerivChunk) * kDerivChunk; i0 >= 0; i0 -= kDerivChunk) {
        const int w = min(kDerivChunk, n - i0);
#pragma unroll
        for (int q = 0; q < kDerivChunk; ++q) {
            const int v = q * kDerivCols + tid, cl = v / kDerivChunk, e = v % kDerivChunk;
            tile[cl][e] = pf[q];
        }
        __syncthreads();
        if (i0 > 0) fetch(i0 - kDerivChunk);
        if (tid < cols) {
            for (int e = w - 1; e >= 0; --e) {
                const double a = tile[tid][e];
                const double k2 = 2.0 * (i0 + e + 1);
                double prev = a1, out = 0.0;
#pragma unroll
                for (int o = 0; o < kDerivMaxOrder; ++o) {
                    if (o < dorder) {
                        const double nw = x2[o] + k2 * prev;  // not contracted (pragma above)
                        prev = x1[o];
                        x2[o] = x1[o];
                        x1[o] = nw;
                        out = nw;
                    }
                }
                a1 = a;
                tile[tid][e] = out * scale;
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int q = 0; q < kDerivChunk; ++q) {
            const int v = q * kDerivCols + tid, cl = v / kDerivChunk, e = v % kDerivChunk;
            if (cl < cols && e < w) d[(col0 + cl) * ld + i0 + e] = tile[cl][e];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// DCT form (power-of-two n = P G <= 256).  With theta_i = (i + 1/2) pi / n the reference's
// tables are cos(k theta'_i), theta'_i = theta_i + delta_i (the accumulated angles, mirrored
// to pi - theta'_{n-1-i} in the second half, which is what the parity split amounts to).
// To first order (k delta <= 2e-12, so the next term is below 1e-23):
//   To_Spectral:   c_k = (2/n) [sum_i f_i cos(k theta_i) - k sum_i delta_i f_i sin(k theta_i)]
//   From_Spectral: f_i = sum_k' c_k cos(k theta_i) - delta_i sum_k k c_k sin(k theta_i)
// (k' halving k = 0).  sin(k theta_i) = (-1)^i cos((n - k) theta_i), so each sine sum is a
// cosine sum of a second real sequence, and the two real transforms go through one complex
// FFT of length n with Makhoul's reordering (x_2m and x_2m+1 to v_m and v_(n-1-m)):
//   analysis:  u_i = f_i (1 + i e_i), e_i = (-1)^i delta_i S; V = FFT(v);
//              c_j = (1/n) [Re(W_j (V_j + conj V_(n-j))) + (j/S) Re(conj W_j (V_(n-j) - conj V_j))]
//   synthesis: V_j = conj W_j ((1 + j/n) c_j - i (j/n) c_(n-j)), v = unnormalised inverse FFT,
//              f_i = (Re v - g_i Im v) / 2, g_i = n (-1)^i delta_i
// with W_j = e^{-i pi j / (2n)} and S a power of two near 1 / max |delta|: both real
// channels then carry values of one magnitude, so rounding in one (about 1e-16 of |V|)
// does not reach the other multiplied by j.
//
// Length n = P G: a column belongs to G lanes of a wave (64 / G columns per wave), lane t
// holding the indices t + G s, s < P, before and after the FFT.  A P-point DFT in
// registers, the twiddle w_n^{t q}, a transpose through the wave's own LDS strip, then
// G-point DFTs (P / G per lane).  Loads and stores are 16-byte (x_2m, x_2m+1) pairs or
// G consecutive doubles per column; the partner values (v_(n-1-m), V_(n-j)) come from lane
// G - 1 - t or (G - t) mod G by lane shuffles.  Waves share no data, so the only
// synchronisation is the wave's own LDS ordering.
constexpr int kDctWaves = 4;

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// pass-1 results x[q] of lane t -> pass-2 operands y[r][t2] = X_t2[t + G r]
template <int P, int G>
__device__ __forceinline__ void dct_transpose(double *buf, int t, const dcplx (&x)[P], dcplx (&y)[P / G][G])
{
    constexpr int LD = G + 1;  // odd stride: the column reads hit distinct banks
    wave_lds_sync();           // the previous tile's reads of buf are done
#pragma unroll
    for (int q = 0; q < P; ++q) buf[q * LD + t] = x[q].x;
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < P / G; ++r)
#pragma unroll
        for (int t2 = 0; t2 < G; ++t2) y[r][t2].x = buf[(t + G * r) * LD + t2];
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < P; ++q) buf[q * LD + t] = x[q].y;
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < P / G; ++r)
#pragma unroll
        for (int t2 = 0; t2 < G; ++t2) y[r][t2].y = buf[(t + G * r) * LD + t2];
}

// Raw buffer accesses: a descriptor in SGPRs (the wave's column tile, or the tables) and a
// 32-bit lane offset; reads past the descriptor's range return 0 and writes there are
// dropped, so a partial last tile needs no branches.
typedef unsigned int dct_u2 __attribute__((ext_vector_type(2)));
typedef unsigned int dct_u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t dct_rsrc(const double *base, long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double dct_ld(__amdgpu_buffer_rsrc_t r, int off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ double2 dct_ld2(__amdgpu_buffer_rsrc_t r, int off)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ void dct_st(double v, __amdgpu_buffer_rsrc_t r, int off)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(dct_u2, v), r, off, 0, 0);
}
__device__ __forceinline__ void dct_st2(double2 v, __amdgpu_buffer_rsrc_t r, int off)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(dct_u4, v), r, off, 0, 0);
}

// a lane's base byte offset with its bits hidden from the optimiser: base + constant then
// stays an add, which the buffer instructions take as base VGPR + immediate (known-disjoint
// bits turn it into an or, and every access gets its own offset register)
__device__ __forceinline__ int lane_base(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// FFT of length P G on this lane's x (indices t + G s in, t + G u out); SG = -1 forward.
// btab (LDS): e^{-2 pi i 2^k t / n} stored [k][t]; the twiddle w^{t q} is applied as the
// product of the powers for the bits of q (at most 4 roundings, so within a few ulps),
// which reads 4 values where a table of w^{t q} would take P - 1.  (The scheduling
// barriers keep each phase's loads next to their use.)
template <int P, int G, int SG>
__device__ __forceinline__ void dct_fft(dcplx (&x)[P], double *buf, int t, const double2 *btab)
{
    constexpr int NB = P >= 16 ? 4 : (P >= 8 ? 3 : 2);
    double2 b[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) b[k] = btab[k * G + t];
    small_dft<P, SG>(x);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 1; q < P; ++q)
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if ((q >> k) & 1) x[q] = cmul(x[q], b[k].x, SG < 0 ? b[k].y : -b[k].y);
    __builtin_amdgcn_sched_barrier(0);
    dcplx y[P / G][G];
    dct_transpose<P, G>(buf, t, x, y);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < P / G; ++r) {
        small_dft<G, SG>(y[r]);
#pragma unroll
        for (int k = 0; k < G; ++k) x[r + (P / G) * k] = y[r][k];
    }
    __builtin_amdgcn_sched_barrier(0);
}

// a * e^{2 pi i e / 64} (e a compile-time constant once unrolled)
__device__ __forceinline__ dcplx rot64(dcplx a, int e)
{
    if ((e & 63) == 0) return a;
    return cmul(a, cos64(e), sin64(e));
}

// Lane exchanges inside a G-lane group (G = 4, 8, 16: a quad, half row or row of 16) as
// DPP moves on the vector ALU; as ds_bpermute they queued behind the LDS transposes.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// the value of lane G - 1 - t (row_mirror, row_half_mirror, quad_perm 3,2,1,0)
template <int G>
__device__ __forceinline__ double lane_reverse(double v)
{
    return dpp_d<G == 16 ? 0x140 : (G == 8 ? 0x141 : 0x1B)>(v);
}
// the value of lane (G - t) mod G for t > 0 (the mirror, then row_shr:1; quad_perm 0,3,2,1);
// what lane t = 0 receives is not meaningful
template <int G>
__device__ __forceinline__ double lane_partner(double v)
{
    if constexpr (G == 4)
        return dpp_d<0x6C>(v);
    else
        return dpp_d<0x111>(lane_reverse<G>(v));
}

// tab: e^{-2 pi i 2^k t / n} [4][G] and e^{-i pi t / (2n)} [G] as (cos, sin) pairs (10 G
// doubles), then floats: the analysis correction e and the synthesis correction g, each
// [k][t][4], entry 4 k + r of lane t being the value at radius 2 (t + G s) + b, 2 s + b =
// 4 k + r (the corrections move the result by 1e-12 of itself at most, so single
// precision in them is far below the double rounding).  W_j = e^{-i pi j/(2n)} for
// j = t + G u is W_t times the constant e^{-i pi u/(2P)}.  Every workgroup copies the
// table into LDS once: read per column tile from global memory it took as many vector
// memory operations as the columns themselves, and those are what the kernels wait on.
constexpr int dct_tab_doubles(int n, int G) { return 10 * G + n; }

__device__ __forceinline__ void dct_copy_table(double *ltab, const double *__restrict__ tab, int nd)
{
    for (int i = threadIdx.x; i < nd / 2; i += blockDim.x)
        reinterpret_cast<double2 *>(ltab)[i] = reinterpret_cast<const double2 *>(tab)[i];
    __syncthreads();
}
template <int P, int G>
__global__ __launch_bounds__(64 * kDctWaves) void cheb_to_spectral_dct_kernel(long ncols, const double *__restrict__ f,
                                                                           double *__restrict__ c,
                                                                           const double *__restrict__ tab,
                                                                           double inv_s)
{
    constexpr int n = P * G, CPW = 64 / G, LD = G + 1;
    extern __shared__ double lds[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int t = lane % G, cs = lane / G;
    constexpr int TD = dct_tab_doubles(n, G);
    double *buf = lds + TD + (wave * CPW + cs) * P * LD;
    dct_copy_table(lds, tab, TD);
    const double2 *btab = reinterpret_cast<const double2 *>(lds), *wtab = btab + 4 * G;
    const float4 *etab = reinterpret_cast<const float4 *>(lds + 10 * G);
    const long step = (long)gridDim.x * kDctWaves * CPW;
    for (long col0 = ((long)blockIdx.x * kDctWaves + wave) * CPW; col0 < ncols; col0 += step) {
        const long bytes = (ncols - col0 < CPW ? ncols - col0 : CPW) * n * 8;
        const __amdgpu_buffer_rsrc_t rf = dct_rsrc(f + col0 * n, bytes), rc = dct_rsrc(c + col0 * n, bytes);
        // lane offsets made inside the loop, so base + constant is folded into each access
        // (hoisted, every sum would hold a register)
        const int v16 = lane_base(cs * n * 8 + 16 * t), v8 = lane_base(cs * n * 8 + 8 * t);
        const int tj = lane_base(t);
        float ef[P];  // e at radii 2 (t + G s) and 2 (t + G s) + 1
#pragma unroll
        for (int k = 0; k < P / 4; ++k) {
            const float4 e4 = etab[k * G + t];
            ef[4 * k] = e4.x;
            ef[4 * k + 1] = e4.y;
            ef[4 * k + 2] = e4.z;
            ef[4 * k + 3] = e4.w;
        }
        dcplx x[P];
#pragma unroll
        for (int s = 0; s < P / 2; ++s) {
            const double2 v = dct_ld2(rf, v16 + 16 * G * s);  // (f_2m, f_2m+1), m = t + G s
            const double2 e{(double)ef[2 * s], (double)ef[2 * s + 1]};
            x[s] = dcplx{v.x, v.x * e.x};
            // u_(2m+1) = v_(n-1-m) belongs to lane G-1-t as its register P-1-s
            x[P - 1 - s] = dcplx{lane_reverse<G>(v.y), lane_reverse<G>(v.y * e.y)};
        }
        __builtin_amdgcn_sched_barrier(0);
        dct_fft<P, G, -1>(x, buf, t, btab);
        const double2 wt = wtab[t];  // W_t
        // outputs u and P-1-u read the same two registers (own and partner's), so taking
        // them together frees registers as the loop goes
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const int u = (k & 1) ? P - 1 - k / 2 : k / 2;
            if (k % 4 == 0) __builtin_amdgcn_sched_barrier(0);
            const dcplx pv = x[u];
            dcplx qv{lane_partner<G>(x[P - 1 - u].x), lane_partner<G>(x[P - 1 - u].y)};
            if (t == 0) qv = x[(P - u) % P];
            const dcplx wj = rot64(dcplx{wt.x, wt.y}, -u * (16 / P));  // W_j, j = t + G u
            const double2 w{wj.x, wj.y};
            const double sr = w.x * (pv.x + qv.x) - w.y * (pv.y - qv.y);
            const double co = w.x * (qv.x - pv.x) + w.y * (qv.y + pv.y);
            dct_st((sr + ((double)(tj + G * u) * inv_s) * co) * (1.0 / n), rc, v8 + 8 * G * u);
        }
    }
}

template <int P, int G>
__global__ __launch_bounds__(64 * kDctWaves) void cheb_from_spectral_dct_kernel(long ncols,
                                                                             const double *__restrict__ c,
                                                                             double *__restrict__ f,
                                                                             const double *__restrict__ tab)
{
    constexpr int n = P * G, CPW = 64 / G, LD = G + 1;
    extern __shared__ double lds[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int t = lane % G, cs = lane / G;
    constexpr int TD = dct_tab_doubles(n, G);
    double *buf = lds + TD + (wave * CPW + cs) * P * LD;
    dct_copy_table(lds, tab, TD);
    const double2 *btab = reinterpret_cast<const double2 *>(lds), *wtab = btab + 4 * G;
    const float4 *gtab = reinterpret_cast<const float4 *>(lds + 10 * G + n / 2);
    const long step = (long)gridDim.x * kDctWaves * CPW;
    for (long col0 = ((long)blockIdx.x * kDctWaves + wave) * CPW; col0 < ncols; col0 += step) {
        const long bytes = (ncols - col0 < CPW ? ncols - col0 : CPW) * n * 8;
        const __amdgpu_buffer_rsrc_t rc = dct_rsrc(c + col0 * n, bytes), rf = dct_rsrc(f + col0 * n, bytes);
        const int v8 = lane_base(cs * n * 8 + 8 * t),
                  v16 = lane_base(cs * n * 8 + 16 * t), tj = lane_base(t);
        const double2 wt = wtab[t];  // W_t
        double cv[P];  // c_j, j = t + G s
#pragma unroll
        for (int s = 0; s < P; ++s) cv[s] = dct_ld(rc, v8 + 8 * G * s);
        __builtin_amdgcn_sched_barrier(0);
        dcplx x[P];
#pragma unroll
        for (int s = 0; s < P; ++s) {
            if (s % 4 == 0) __builtin_amdgcn_sched_barrier(0);
            const int j = tj + G * s;
            const double cj = cv[s];
            // c_(n-j): lane (G - t) mod G holds it as its register P-1-s, lane 0 as its own
            // register (P - s) mod P; c_n = 0
            double cn = lane_partner<G>(cv[P - 1 - s]);
            if (t == 0) cn = s > 0 ? cv[P - s] : 0.0;
            const dcplx wj = rot64(dcplx{wt.x, wt.y}, -s * (16 / P));  // W_j; conj W_j = (w.x, -w.y)
            const double2 w{wj.x, wj.y};
            const double a = (1.0 + (double)j * (1.0 / n)) * cj, b = -((double)j * (1.0 / n)) * cn;
            x[s] = dcplx{w.x * a + w.y * b, w.x * b - w.y * a};
        }
        __builtin_amdgcn_sched_barrier(0);
        dct_fft<P, G, 1>(x, buf, t, btab);
        float gf[P];  // g at radii 2 (t + G u) and 2 (t + G u) + 1
#pragma unroll
        for (int k = 0; k < P / 4; ++k) {
            const float4 g4 = gtab[k * G + t];
            gf[4 * k] = g4.x;
            gf[4 * k + 1] = g4.y;
            gf[4 * k + 2] = g4.z;
            gf[4 * k + 3] = g4.w;
        }
#pragma unroll
        for (int u = 0; u < P / 2; ++u) {
            if (u % 4 == 0) __builtin_amdgcn_sched_barrier(0);
            const dcplx a = x[u];  // v_m, m = t + G u
            // v_(n-1-m) is lane G-1-t's register P-1-u
            const dcplx b{lane_reverse<G>(x[P - 1 - u].x), lane_reverse<G>(x[P - 1 - u].y)};
            const double2 g{(double)gf[2 * u], (double)gf[2 * u + 1]};
            dct_st2(double2{0.5 * (a.x - g.x * a.y), 0.5 * (b.x - g.y * b.y)}, rf, v16 + 16 * G * u);
        }
    }
}

}  // namespace
}  // namespace raysht

using namespace raysht;

namespace raysht {
namespace {
constexpr int kChebRegBlocks = 1024;  // persistent workgroups (about four per CU)

// register-table kernels: (KB, NW) = K steps of 4 held per parity, waves (16 rows each)
template <int KB, int NW, int TC, bool SD>
int launch_cheb_reg_sd(raysht_cheb *cb, bool to, const double *in, double *out, long ncols, long ld, hipStream_t st)
{
    const long ntiles = (ncols + TC - 1) / TC;
    const unsigned blocks = (unsigned)std::min<long>(ntiles, kChebRegBlocks);
    const size_t lds = sizeof(double) * 2 * TC * (4 * KB + 1);
    if (to)
        hipLaunchKernelGGL((cheb_to_spectral_reg_kernel<KB, NW, TC, SD>), dim3(blocks), dim3(64 * NW), lds, st, cb->n,
                           cb->n_x, cb->n_even, cb->n_odd, cb->ix4, cb->kp, ncols, ld, cb->d_tt_e, cb->d_tt_o, in,
                           out);
    else
        hipLaunchKernelGGL((cheb_from_spectral_reg_kernel<KB, NW, TC, SD>), dim3(blocks), dim3(64 * NW), lds, st, cb->n,
                           cb->n_x, cb->n_even, cb->k4, cb->ip, ncols, ld, cb->d_t_e, cb->d_t_o, in, out);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

template <int KB, int NW, int TC>
int launch_cheb_reg(raysht_cheb *cb, bool to, const double *in, double *out, long ncols, long ld, hipStream_t st)
{
    return ld == cb->n ? launch_cheb_reg_sd<KB, NW, TC, false>(cb, to, in, out, ncols, ld, st)
                       : launch_cheb_reg_sd<KB, NW, TC, true>(cb, to, in, out, ncols, ld, st);
}

// persistent workgroups: as many as are resident at once (a grid of several rounds leaves
// the last round's CUs partly idle)
template <typename K>
long resident_blocks(K kernel, int device, size_t lds)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * kDctWaves, lds) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || per_cu < 1)
        return 1024;
    return (long)per_cu * cus;
}

template <int P, int G>
int launch_cheb_dct(raysht_cheb *cb, bool to, const double *in, double *out, long ncols, hipStream_t st)
{
    constexpr int n = P * G, CPW = 64 / G;
    const long waves = (ncols + CPW - 1) / CPW;
    const size_t lds = sizeof(double) * (dct_tab_doubles(n, G) + kDctWaves * CPW * P * (G + 1));
    const long cap = to ? resident_blocks(cheb_to_spectral_dct_kernel<P, G>, cb->device, lds)
                        : resident_blocks(cheb_from_spectral_dct_kernel<P, G>, cb->device, lds);
    const unsigned blocks = (unsigned)std::min<long>((waves + kDctWaves - 1) / kDctWaves, cap);
    static_assert(n * 8 * (64 / G) < (1 << 20), "tile offsets fit 32 bits");
    if (to)
        hipLaunchKernelGGL((cheb_to_spectral_dct_kernel<P, G>), dim3(blocks), dim3(64 * kDctWaves), lds, st, ncols, in,
                           out, cb->d_dct, cb->dct_inv_s);
    else
        hipLaunchKernelGGL((cheb_from_spectral_dct_kernel<P, G>), dim3(blocks), dim3(64 * kDctWaves), lds, st, ncols,
                           in, out, cb->d_dct);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

bool dct_length(int n) { return n == 16 || n == 32 || n == 64 || n == 128 || n == 256; }
// lanes per column G of the DCT kernel for n (P = n / G registers per lane)
int dct_lanes(int n) { return n == 256 ? 16 : (n >= 64 ? 8 : 4); }

// the DCT kernels where the plan has them and the columns are 16-byte aligned
bool cheb_dct_dispatch(raysht_cheb *cb, bool to, const double *in, double *out, long ncols, hipStream_t st, int *rc)
{
    if (cb->form != RAYSHT_CHEB_DCT || !cb->d_dct) return false;
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) return false;
    switch (cb->n) {
    case 16: *rc = launch_cheb_dct<4, 4>(cb, to, in, out, ncols, st); return true;
    case 32: *rc = launch_cheb_dct<8, 4>(cb, to, in, out, ncols, st); return true;
    case 64: *rc = launch_cheb_dct<8, 8>(cb, to, in, out, ncols, st); return true;
    case 128: *rc = launch_cheb_dct<16, 8>(cb, to, in, out, ncols, st); return true;
    case 256: *rc = launch_cheb_dct<16, 16>(cb, to, in, out, ncols, st); return true;
    default: return false;
    }
}

// returns false when n is too long for the register tables (> 256); ld = column stride
bool cheb_reg_dispatch(raysht_cheb *cb, bool to, const double *in, double *out, long ncols, long ld, hipStream_t st,
                       int *rc)
{
    // K steps: to = radii (ix4), from = coefficients (k4); rows: to = kp, from = ip
    const int ksteps = (std::max(cb->ix4, cb->k4) + 3) / 4, rows = std::max(cb->kp, cb->ip);
    if (ksteps <= 8 && rows <= 32)
        *rc = launch_cheb_reg<8, 2, 32>(cb, to, in, out, ncols, ld, st);
    else if (ksteps <= 16 && rows <= 64)
        *rc = launch_cheb_reg<16, 4, 32>(cb, to, in, out, ncols, ld, st);
    else if (ksteps <= 32 && rows <= 128)
        *rc = launch_cheb_reg<32, 8, 16>(cb, to, in, out, ncols, ld, st);
    else
        return false;
    return true;
}

// a grid of several domains: each domain's table form on its rows of every column
int cheb_domains_dispatch(raysht_cheb *cb, bool to, const double *in, double *out, long ncols, hipStream_t st)
{
    for (size_t h = 0; h < cb->dom.size(); ++h) {
        int rc = RAYSHT_OK;
        if (!cheb_reg_dispatch(cb->dom[h], to, in + cb->hoff[h], out + cb->hoff[h], ncols, cb->n, st, &rc))
            return RAYSHT_EUNSUPPORTED;
        if (rc != RAYSHT_OK) return rc;
    }
    return RAYSHT_OK;
}
}  // namespace
}  // namespace raysht

namespace raysht {
namespace {
// One domain of n points: Gen_Colocation_Points' zeros (:233-243, the angle grown by repeated
// addition), Gen_Tn's parity tables (:251-297), their padded device copies and, where allowed,
// the DCT form's tables.  The map onto radii is the caller's.
int cheb_build(raysht_cheb *cb, int n, bool allow_dct)
{
    cb->n = n;
    cb->n_odd = n / 2;
    cb->n_even = cb->n_odd + n % 2;
    cb->n_x = cb->n_even;
    const double pi = 3.1415926535897932384626433832795028841972;
    std::vector<double> theta(n), x(n);
    const double dtheta = pi / n;
    double arg = dtheta * 0.5;
    for (int i = 0; i < n; ++i) {
        theta[i] = arg;
        x[i] = std::cos(arg);
        arg = arg + dtheta;
    }
    cb->x = x;
    cb->theta = theta;
    // Gen_Tn (:251-297): cheby_even(i, k) = T_{2k-2}(x_i), cheby_odd(i, k) = T_{2k-1}(x_i), i <= n_x
    cb->ceven.assign((size_t)cb->n_x * cb->n_even, 0.0);
    cb->codd.assign((size_t)cb->n_x * (cb->n_odd > 0 ? cb->n_odd : 1), 0.0);
    for (int k = 0; k < cb->n_even; ++k)
        for (int i = 0; i < cb->n_x; ++i) cb->ceven[(size_t)k * cb->n_x + i] = std::cos((double)(2 * k) * theta[i]);
    for (int k = 0; k < cb->n_odd; ++k)
        for (int i = 0; i < cb->n_x; ++i) cb->codd[(size_t)k * cb->n_x + i] = std::cos((double)(2 * k + 1) * theta[i]);
    if (cb->n_x != cb->n_odd) {
        for (int k = 0; k < cb->n_even; ++k) cb->ceven[(size_t)k * cb->n_x + cb->n_x - 1] *= 0.5;
        for (int k = 0; k < cb->n_odd; ++k) cb->codd[(size_t)k * cb->n_x + cb->n_x - 1] *= 0.5;
    }
    auto up = [](int v, int m) { return (v + m - 1) / m * m; };
    cb->ix4 = up(cb->n_x, 4);
    cb->kp = up(cb->n_even, 16);
    cb->k4 = up(cb->n_even, 4);
    cb->ip = up(cb->n_x, 16);
    std::vector<double> tt_e((size_t)cb->ix4 * cb->kp, 0.0), tt_o((size_t)cb->ix4 * cb->kp, 0.0);
    std::vector<double> t_e((size_t)cb->k4 * cb->ip, 0.0), t_o((size_t)cb->k4 * cb->ip, 0.0);
    for (int k = 0; k < cb->n_even; ++k)
        for (int i = 0; i < cb->n_x; ++i) {
            tt_e[(size_t)i * cb->kp + k] = cb->ceven[(size_t)k * cb->n_x + i];
            t_e[(size_t)k * cb->ip + i] = cb->ceven[(size_t)k * cb->n_x + i];
        }
    for (int k = 0; k < cb->n_odd; ++k)
        for (int i = 0; i < cb->n_x; ++i) {
            tt_o[(size_t)i * cb->kp + k] = cb->codd[(size_t)k * cb->n_x + i];
            t_o[(size_t)k * cb->ip + i] = cb->codd[(size_t)k * cb->n_x + i];
        }
    // DCT form: twiddles, e^{-i pi j/(2n)}, and the two correction sequences (delta mirrored
    // as the parity split uses the first half's angles for the second half)
    std::vector<double> dct;
    if (allow_dct && dct_length(n)) {
        std::vector<double> twc(n), tws(n), qc(n), qs(n), delta(n);
        cheb_dct_tables(n, theta.data(), twc.data(), tws.data(), qc.data(), qs.data(), delta.data());
        std::vector<double> dp(n);
        double dmax = 0.0;
        for (int i = 0; i < n; ++i) {
            dp[i] = i < n / 2 ? delta[i] : -delta[n - 1 - i];
            dmax = std::max(dmax, std::fabs(dp[i]));
        }
        const double scale = dmax > 0.0 ? std::ldexp(1.0, -std::ilogb(dmax)) : 1.0;  // S dmax in [1, 2)
        cb->dct_inv_s = 1.0 / scale;
        const int G = dct_lanes(n), P = n / G;
        dct.assign((size_t)dct_tab_doubles(n, G), 0.0);
        for (int t = 0; t < G; ++t) {
            for (int k = 0; k < 4; ++k) {  // e^{-2 pi i 2^k t / n}, [k][t]
                const int e = (t << k) % n;
                dct[2 * (k * G + t)] = twc[e];
                dct[2 * (k * G + t) + 1] = tws[e];
            }
            dct[8 * G + 2 * t] = qc[t];  // e^{-i pi t / (2n)}
            dct[8 * G + 2 * t + 1] = qs[t];
        }
        float *ef = reinterpret_cast<float *>(dct.data() + 10 * G), *gf = ef + n;
        for (int t = 0; t < G; ++t)
            for (int s = 0; s < P / 2; ++s)
                for (int b = 0; b < 2; ++b) {
                    const int i = 2 * (t + G * s) + b, entry = 2 * s + b;  // lane t's entry
                    const int at = ((entry / 4) * G + t) * 4 + entry % 4;   // [k][t][4]
                    const double sg = (i & 1) ? -1.0 : 1.0;
                    ef[at] = (float)(sg * dp[i] * scale);
                    gf[at] = (float)(sg * dp[i] * (double)n);
                }
    }
    // the DCT form by default where it is the faster (n = 128: 0.22 / 0.22 ms against 0.27 /
    // 0.35 ms for 526336 columns; n = 256: 4.45 / 4.40 ms against 7.32 / 7.63 ms for 5248000);
    // below that the table form's products are cheap enough to keep
    cb->form = (!dct.empty() && n >= 128) ? RAYSHT_CHEB_DCT : RAYSHT_CHEB_MATRIX;
    struct up_t {
        double **dst;
        const std::vector<double> *src;
    } ups[5] = {{&cb->d_tt_e, &tt_e}, {&cb->d_tt_o, &tt_o}, {&cb->d_t_e, &t_e}, {&cb->d_t_o, &t_o}, {&cb->d_dct, &dct}};
    for (auto &u : ups) {
        if (u.src->empty()) continue;
        const size_t bytes = u.src->size() * sizeof(double);
        if (hipMalloc(u.dst, bytes) != hipSuccess) {
            *u.dst = nullptr;
            return RAYSHT_ENOMEM;
        }
        if (hipMemcpy(*u.dst, u.src->data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return RAYSHT_EHIP;
    }
    return RAYSHT_OK;
}
}  // namespace
}  // namespace raysht

extern "C" {

int raysht_cheb_create_domains(raysht_cheb **out, int ndomains, const int *npoly, const double *bounds,
                               const int *dealias_by, double *grid)
{
    if (!out) return RAYSHT_EINVAL;
    *out = nullptr;
    if (ndomains < 1 || !npoly || !bounds) return RAYSHT_EINVAL;
    long ntotal = 0;
    for (int i = 0; i < ndomains; ++i) {
        if (npoly[i] < 2 || !(bounds[i + 1] > bounds[i])) return RAYSHT_EINVAL;
        // several domains run the register-table kernels, which take up to 256 points
        if (npoly[i] > (ndomains == 1 ? kChebMaxN : 256)) return RAYSHT_EUNSUPPORTED;
        ntotal += npoly[i];
    }
    if (ntotal > (1L << 20)) return RAYSHT_EUNSUPPORTED;
    raysht_cheb *cb = new (std::nothrow) raysht_cheb();
    if (!cb) return RAYSHT_ENOMEM;
    if (hipGetDevice(&cb->device) != hipSuccess) {
        delete cb;
        return RAYSHT_EHIP;
    }
    // Initialize_Cheby_Grid (:101-200): the domains and their bounds reversed (outermost
    // first), rda = (2 n) / 3 + 1 or n - dealias_by + 1, and per domain the affine map of
    // [x_n, x_1] onto [lowerb, upperb], the derivative scaling (the reciprocal of the map's
    // slope) and the radial integration weights int_scale r_i^2 sqrt(1 - x_i^2), halved at
    // the domain's ends, int_scale = 3 pi slope / ((gmax^3 - gmin^3) n); each rounded in the
    // order written there
    const double pi = 3.1415926535897932384626433832795028841972;
    const double gmax = bounds[ndomains], gmin = bounds[0];
    const double gcube = gmax * gmax * gmax - gmin * gmin * gmin;
    std::vector<double> r_all, x_all, th_all, w_all, ds_all;
    int hoff = 0;
    for (int h = 0; h < ndomains; ++h) {
        const int src = ndomains - 1 - h, n = npoly[src];
        const double upperb = bounds[src + 1], lowerb = bounds[src];
        raysht_cheb *d = cb;
        if (ndomains > 1) {
            d = new (std::nothrow) raysht_cheb();
            if (!d) {
                raysht_cheb_destroy(cb);
                return RAYSHT_ENOMEM;
            }
            d->device = cb->device;
            cb->dom.push_back(d);
        }
        if (int rc = cheb_build(d, n, ndomains == 1)) {
            raysht_cheb_destroy(cb);
            return rc;
        }
        int rda = (2 * n) / 3 + 1;
        if (dealias_by && dealias_by[src] >= 1 && dealias_by[src] < n) rda = n - dealias_by[src] + 1;
        const double xmin = d->x[n - 1], xmax = d->x[0];
        const double scaling = (upperb - lowerb) / (xmax - xmin);
        const double int_scale = (3.0 * pi * scaling) / (gcube * (double)n);
        d->deriv_scaling = 1.0 / scaling;
        for (int i = 0; i < n; ++i) {
            const double r = (d->x[i] - xmin) * scaling + lowerb;
            r_all.push_back(r);
            x_all.push_back(d->x[i]);
            th_all.push_back(d->theta[i]);
            w_all.push_back(int_scale * (r * r) * std::sqrt(1.0 - d->x[i] * d->x[i]));
            ds_all.push_back(d->deriv_scaling);
        }
        w_all[hoff] *= 0.5;
        w_all[hoff + n - 1] *= 0.5;
        cb->hoff.push_back(hoff);
        cb->rda.push_back(rda);
        hoff += n;
    }
    cb->n = (int)ntotal;
    cb->x = x_all;
    cb->theta = th_all;
    cb->weights = w_all;
    cb->dscale = ds_all;
    if (ndomains > 1) cb->form = RAYSHT_CHEB_MATRIX;
    if (grid)
        for (long i = 0; i < ntotal; ++i) grid[i] = r_all[i];
    *out = cb;
    return RAYSHT_OK;
}

int raysht_cheb_create(raysht_cheb **out, int n, double rmin, double rmax, double *grid)
{
    const double bounds[2] = {rmin, rmax};
    return raysht_cheb_create_domains(out, 1, &n, bounds, nullptr, grid);
}

void raysht_cheb_destroy(raysht_cheb *cb)
{
    if (!cb) return;
    for (raysht_cheb *d : cb->dom) raysht_cheb_destroy(d);
    for (double *d : {cb->d_tt_e, cb->d_tt_o, cb->d_t_e, cb->d_t_o, cb->d_dct})
        if (d) (void)hipFree(d);
    delete cb;
}

int raysht_cheb_grid_data(const raysht_cheb *cb, double *x, double *theta, double *weights, double *deriv_scaling)
{
    if (!cb) return RAYSHT_EINVAL;
    for (int i = 0; i < cb->n; ++i) {
        if (x) x[i] = cb->x[i];
        if (theta) theta[i] = cb->theta[i];
        if (weights) weights[i] = cb->weights[i];
        if (deriv_scaling) deriv_scaling[i] = cb->dscale[i];
    }
    return RAYSHT_OK;
}

int raysht_cheb_table(const raysht_cheb *cb, int which, double *host_out)
{
    if (!cb || !host_out || which < 0 || which > 1) return RAYSHT_EINVAL;
    if (!cb->dom.empty()) return RAYSHT_EUNSUPPORTED;  // one table set per domain
    const std::vector<double> &t = which == 0 ? cb->ceven : cb->codd;
    const size_t n = which == 0 ? (size_t)cb->n_x * cb->n_even : (size_t)cb->n_x * cb->n_odd;
    for (size_t i = 0; i < n; ++i) host_out[i] = t[i];
    return RAYSHT_OK;
}

namespace {
// 16 columns per wave; fewer waves per workgroup for long columns to bound the LDS tile
// (n > 256: one wave, 2 x 16 x (n / 2 + 1) doubles, 131 KB at n = 1024)
int cheb_waves(int n) { return n <= 128 ? 4 : (n <= 256 ? 2 : 1); }

// the LDS one workgroup of the grid's device may allocate
bool lds_fits(const raysht_cheb *cb, size_t bytes)
{
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, cb->device) != hipSuccess) return false;
    return bytes <= (size_t)v;
}
}  // namespace

int raysht_cheb_to_spectral(raysht_cheb *cb, const double *f, double *c, long ncols, void *stream)
{
    if (!cb || ncols < 0 || (ncols > 0 && (!f || !c))) return RAYSHT_EINVAL;
    if (ncols == 0) return RAYSHT_OK;
    int rc = RAYSHT_OK;
    if (!cb->dom.empty()) return cheb_domains_dispatch(cb, true, f, c, ncols, reinterpret_cast<hipStream_t>(stream));
    if (cheb_dct_dispatch(cb, true, f, c, ncols, reinterpret_cast<hipStream_t>(stream), &rc)) return rc;
    if (cheb_reg_dispatch(cb, true, f, c, ncols, cb->n, reinterpret_cast<hipStream_t>(stream), &rc)) return rc;
    const int waves = cheb_waves(cb->n), tc = 16 * waves;
    const long blocks = (ncols + tc - 1) / tc;
    if (blocks > 0x7fffffffL) return RAYSHT_EINVAL;
    const size_t lds = sizeof(double) * 2 * tc * (cb->ix4 + 1);
    if (!lds_fits(cb, lds)) return RAYSHT_EUNSUPPORTED;
    hipLaunchKernelGGL(cheb_to_spectral_kernel, dim3((unsigned)blocks), dim3(64 * waves), lds,
                       reinterpret_cast<hipStream_t>(stream), cb->n, cb->n_x, cb->n_even, cb->n_odd, cb->ix4, cb->kp,
                       ncols, cb->d_tt_e, cb->d_tt_o, f, c);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_cheb_from_spectral(raysht_cheb *cb, const double *c, double *f, long ncols, void *stream)
{
    if (!cb || ncols < 0 || (ncols > 0 && (!f || !c))) return RAYSHT_EINVAL;
    if (ncols == 0) return RAYSHT_OK;
    int rc = RAYSHT_OK;
    if (!cb->dom.empty()) return cheb_domains_dispatch(cb, false, c, f, ncols, reinterpret_cast<hipStream_t>(stream));
    if (cheb_dct_dispatch(cb, false, c, f, ncols, reinterpret_cast<hipStream_t>(stream), &rc)) return rc;
    if (cheb_reg_dispatch(cb, false, c, f, ncols, cb->n, reinterpret_cast<hipStream_t>(stream), &rc)) return rc;
    const int waves = cheb_waves(cb->n), tc = 16 * waves;
    const long blocks = (ncols + tc - 1) / tc;
    if (blocks > 0x7fffffffL) return RAYSHT_EINVAL;
    const size_t lds = sizeof(double) * 2 * tc * (cb->k4 + 1);
    if (!lds_fits(cb, lds)) return RAYSHT_EUNSUPPORTED;
    hipLaunchKernelGGL(cheb_from_spectral_kernel, dim3((unsigned)blocks), dim3(64 * waves), lds,
                       reinterpret_cast<hipStream_t>(stream), cb->n, cb->n_x, cb->n_even, cb->n_odd, cb->k4, cb->ip,
                       ncols, cb->d_t_e, cb->d_t_o, c, f);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_cheb_set_form(raysht_cheb *cb, int form)
{
    if (!cb || (form != RAYSHT_CHEB_DCT && form != RAYSHT_CHEB_MATRIX)) return RAYSHT_EINVAL;
    if (form == RAYSHT_CHEB_DCT && !cb->d_dct) return RAYSHT_EUNSUPPORTED;  // several domains: table form only
    cb->form = form;
    return RAYSHT_OK;
}

int raysht_cheb_deriv(raysht_cheb *cb, const double *c, double *d, long ncols, int dorder, void *stream)
{
    if (!cb || ncols < 0 || dorder < 1) return RAYSHT_EINVAL;
    if (dorder > kDerivMaxOrder) return RAYSHT_EUNSUPPORTED;
    if (ncols == 0) return RAYSHT_OK;
    if (!c || !d) return RAYSHT_EINVAL;
    // out of place only: the reference's in-place branch (ind == dind) is unimplemented
    const long n_el = ncols * cb->n;
    if (c < d + n_el && d < c + n_el) return RAYSHT_EINVAL;
    const long blocks = (ncols + kDerivCols - 1) / kDerivCols;
    if (blocks > 0x7fffffffL) return RAYSHT_EUNSUPPORTED;
    // one launch per domain (its rows of every column), each with its own scaling
    const int nd = cb->dom.empty() ? 1 : (int)cb->dom.size();
    for (int h = 0; h < nd; ++h) {
        const raysht_cheb *dm = cb->dom.empty() ? cb : cb->dom[h];
        // deriv_scaling**dorder as the compiled Fortran power forms it for small orders
        const double s = dm->deriv_scaling;
        const double scale = dorder == 1 ? s : dorder == 2 ? s * s : dorder == 3 ? (s * s) * s : (s * s) * (s * s);
        hipLaunchKernelGGL(cheb_deriv_kernel, dim3((unsigned)blocks), dim3(kDerivCols), 0,
                           reinterpret_cast<hipStream_t>(stream), dm->n, (long)cb->n, ncols, dorder, scale,
                           c + cb->hoff[h], d + cb->hoff[h]);
        RAYSHT_HIP_TRY(hipGetLastError());
    }
    return RAYSHT_OK;
}

int raysht_cheb_dealias(raysht_cheb *cb, double *c, long ncols, void *stream)
{
    if (!cb || ncols < 0 || (ncols > 0 && !c)) return RAYSHT_EINVAL;
    if (ncols == 0) return RAYSHT_OK;
    // dealias_buffer4d (:353-374): rows rda .. npoly (1-based) of every domain zeroed
    const int nd = cb->dom.empty() ? 1 : (int)cb->dom.size();
    for (int h = 0; h < nd; ++h) {
        const int n = cb->dom.empty() ? cb->n : cb->dom[h]->n, first = cb->rda[h] - 1;
        if (first >= n) continue;
        const long total = ncols * (n - first);
        const long blocks = std::min<long>((total + 255) / 256, 1L << 20);
        hipLaunchKernelGGL(cheb_dealias_kernel, dim3((unsigned)blocks), dim3(256), 0,
                           reinterpret_cast<hipStream_t>(stream), n, first, ncols, (long)cb->n, c + cb->hoff[h]);
        RAYSHT_HIP_TRY(hipGetLastError());
    }
    return RAYSHT_OK;
}

}  // extern "C"
