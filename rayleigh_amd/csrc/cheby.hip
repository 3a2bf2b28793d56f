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
#include <new>
#include <vector>

#include "raysht_internal.h"

namespace raysht {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double4_t mfma_f64(double a, double b, double4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kChebMaxN = 512;
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

// From_Spectral4D with K = coefficients (KB * 4 >= k4), rows = radii i < n_x.
template <int KB, int NW, int TC>
__global__ __launch_bounds__(64 * NW) void cheb_from_spectral_reg_kernel(int n, int n_x, int n_even, int k4, int ip,
                                                                     long ncols, const double *__restrict__ t_e,
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
        const double *src = c + tile * TC * (long)n;
        const long left = (ncols - tile * TC) * (long)n;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int idx = tid + u * nthr;
            pre[u] = (idx < per && idx < left) ? src[idx] : 0.0;
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

// To_Spectral4D with K = radii i < n_x (KB * 4 >= ix4), rows = coefficient pairs k.
template <int KB, int NW, int TC>
__global__ __launch_bounds__(64 * NW) void cheb_to_spectral_reg_kernel(int n, int n_x, int n_even, int n_odd, int ix4,
                                                                   int kp, long ncols, const double *__restrict__ tt_e,
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
    constexpr int kMaxPer = TC * 4 * KB;  // per = 32 n_x <= 32 * 4 KB
    constexpr int kU = (kMaxPer + nthr - 1) / nthr;
    double pa[kU], pb[kU];
    auto fetch = [&](long tile) {
        const double *src = f + tile * TC * (long)n;
        const long cols_left = ncols - tile * TC;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int idx = tid + u * nthr;
            const int col = small_div(idx, inv_nx), i = idx - col * n_x;
            const bool ok = idx < per && col < cols_left;
            pa[u] = ok ? src[(long)col * n + i] : 0.0;
            pb[u] = ok ? src[(long)col * n + (n - 1 - i)] : 0.0;
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
                    if (k < n_even) c[col * n + 2 * k] = alpha * acc_e[cb][v];
                    if (k < n_odd) c[col * n + 2 * k + 1] = alpha * acc_o[cb][v];
                }
            }
    }
}

__global__ void cheb_dealias_kernel(int n, int first, long ncols, double *__restrict__ c)
{
    const int width = n - first;
    const long total = ncols * width;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long col = i / width;
        c[col * n + first + (int)(i % width)] = 0.0;
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

__global__ __launch_bounds__(kDerivCols) void cheb_deriv_kernel(int n, long ncols, int dorder, double scale,
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
            pf[q] = cl < cols && e < w ? c[(col0 + cl) * n + i0 + e] : 0.0;
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
            if (cl < cols && e < w) d[(col0 + cl) * n + i0 + e] = tile[cl][e];
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace raysht

using namespace raysht;

namespace raysht {
namespace {
constexpr int kChebRegBlocks = 1024;  // persistent workgroups (about four per CU)

// register-table kernels: (KB, NW) = K steps of 4 held per parity, waves (16 rows each)
template <int KB, int NW, int TC>
int launch_cheb_reg(raysht_cheb *cb, bool to, const double *in, double *out, long ncols, hipStream_t st)
{
    const long ntiles = (ncols + TC - 1) / TC;
    const unsigned blocks = (unsigned)std::min<long>(ntiles, kChebRegBlocks);
    const size_t lds = sizeof(double) * 2 * TC * (4 * KB + 1);
    if (to)
        hipLaunchKernelGGL((cheb_to_spectral_reg_kernel<KB, NW, TC>), dim3(blocks), dim3(64 * NW), lds, st, cb->n,
                           cb->n_x, cb->n_even, cb->n_odd, cb->ix4, cb->kp, ncols, cb->d_tt_e, cb->d_tt_o, in, out);
    else
        hipLaunchKernelGGL((cheb_from_spectral_reg_kernel<KB, NW, TC>), dim3(blocks), dim3(64 * NW), lds, st, cb->n,
                           cb->n_x, cb->n_even, cb->k4, cb->ip, ncols, cb->d_t_e, cb->d_t_o, in, out);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

// returns false when n is too long for the register tables (> 256)
bool cheb_reg_dispatch(raysht_cheb *cb, bool to, const double *in, double *out, long ncols, hipStream_t st, int *rc)
{
    // K steps: to = radii (ix4), from = coefficients (k4); rows: to = kp, from = ip
    const int ksteps = (std::max(cb->ix4, cb->k4) + 3) / 4, rows = std::max(cb->kp, cb->ip);
    if (ksteps <= 8 && rows <= 32)
        *rc = launch_cheb_reg<8, 2, 32>(cb, to, in, out, ncols, st);
    else if (ksteps <= 16 && rows <= 64)
        *rc = launch_cheb_reg<16, 4, 32>(cb, to, in, out, ncols, st);
    else if (ksteps <= 32 && rows <= 128)
        *rc = launch_cheb_reg<32, 8, 16>(cb, to, in, out, ncols, st);
    else
        return false;
    return true;
}
}  // namespace
}  // namespace raysht

extern "C" {

int raysht_cheb_create(raysht_cheb **out, int n, double rmin, double rmax, double *grid)
{
    if (!out) return RAYSHT_EINVAL;
    *out = nullptr;
    if (n < 2 || n > kChebMaxN || !(rmax > rmin)) return RAYSHT_EINVAL;
    raysht_cheb *cb = new (std::nothrow) raysht_cheb();
    if (!cb) return RAYSHT_ENOMEM;
    if (hipGetDevice(&cb->device) != hipSuccess) {
        delete cb;
        return RAYSHT_EHIP;
    }
    cb->n = n;
    cb->n_odd = n / 2;
    cb->n_even = cb->n_odd + n % 2;
    cb->n_x = cb->n_even;
    // Gen_Colocation_Points (zeros branch, :233-243): the angle grows by repeated addition
    const double pi = 3.1415926535897932384626433832795028841972;
    std::vector<double> theta(n), x(n);
    const double dtheta = pi / n;
    double arg = dtheta * 0.5;
    for (int i = 0; i < n; ++i) {
        theta[i] = arg;
        x[i] = std::cos(arg);
        arg = arg + dtheta;
    }
    {
        // Initialize_Cheby_Grid (:161-200): affine map of [x_n, x_1] onto [rmin, rmax]; the
        // derivative scaling is the reciprocal of the map's slope; the radial integration
        // weights int_scale r_i^2 sqrt(1 - x_i^2), halved at both ends, with
        // int_scale = 3 pi slope / ((rmax^3 - rmin^3) n); each rounded in the order written there
        const double xmin = x[n - 1], xmax = x[0];
        const double scaling = (rmax - rmin) / (xmax - xmin);
        std::vector<double> r(n);
        for (int i = 0; i < n; ++i) r[i] = (x[i] - xmin) * scaling + rmin;
        if (grid)
            for (int i = 0; i < n; ++i) grid[i] = r[i];
        const double int_scale = (3.0 * pi * scaling) / ((rmax * rmax * rmax - rmin * rmin * rmin) * (double)n);
        cb->weights.resize(n);
        for (int i = 0; i < n; ++i) cb->weights[i] = int_scale * (r[i] * r[i]) * std::sqrt(1.0 - x[i] * x[i]);
        cb->weights[0] *= 0.5;
        cb->weights[n - 1] *= 0.5;
        cb->deriv_scaling = 1.0 / scaling;
        cb->x = x;
        cb->theta = theta;
    }
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
    struct up_t {
        double **dst;
        const std::vector<double> *src;
    } ups[4] = {{&cb->d_tt_e, &tt_e}, {&cb->d_tt_o, &tt_o}, {&cb->d_t_e, &t_e}, {&cb->d_t_o, &t_o}};
    for (auto &u : ups) {
        const size_t bytes = u.src->size() * sizeof(double);
        if (hipMalloc(u.dst, bytes) != hipSuccess) {
            *u.dst = nullptr;
            raysht_cheb_destroy(cb);
            return RAYSHT_ENOMEM;
        }
        if (hipMemcpy(*u.dst, u.src->data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
            raysht_cheb_destroy(cb);
            return RAYSHT_EHIP;
        }
    }
    *out = cb;
    return RAYSHT_OK;
}

void raysht_cheb_destroy(raysht_cheb *cb)
{
    if (!cb) return;
    for (double *d : {cb->d_tt_e, cb->d_tt_o, cb->d_t_e, cb->d_t_o})
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
        if (deriv_scaling) deriv_scaling[i] = cb->deriv_scaling;
    }
    return RAYSHT_OK;
}

int raysht_cheb_table(const raysht_cheb *cb, int which, double *host_out)
{
    if (!cb || !host_out || which < 0 || which > 1) return RAYSHT_EINVAL;
    const std::vector<double> &t = which == 0 ? cb->ceven : cb->codd;
    const size_t n = which == 0 ? (size_t)cb->n_x * cb->n_even : (size_t)cb->n_x * cb->n_odd;
    for (size_t i = 0; i < n; ++i) host_out[i] = t[i];
    return RAYSHT_OK;
}

namespace {
// 16 columns per wave; fewer waves per workgroup for long columns to bound the LDS tile
int cheb_waves(int n) { return n <= 128 ? 4 : (n <= 256 ? 2 : 1); }
}  // namespace

int raysht_cheb_to_spectral(raysht_cheb *cb, const double *f, double *c, long ncols, void *stream)
{
    if (!cb || ncols < 0 || (ncols > 0 && (!f || !c))) return RAYSHT_EINVAL;
    if (ncols == 0) return RAYSHT_OK;
    int rc = RAYSHT_OK;
    if (cheb_reg_dispatch(cb, true, f, c, ncols, reinterpret_cast<hipStream_t>(stream), &rc)) return rc;
    const int waves = cheb_waves(cb->n), tc = 16 * waves;
    const long blocks = (ncols + tc - 1) / tc;
    if (blocks > 0x7fffffffL) return RAYSHT_EINVAL;
    const size_t lds = sizeof(double) * 2 * tc * (cb->ix4 + 1);
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
    if (cheb_reg_dispatch(cb, false, c, f, ncols, reinterpret_cast<hipStream_t>(stream), &rc)) return rc;
    const int waves = cheb_waves(cb->n), tc = 16 * waves;
    const long blocks = (ncols + tc - 1) / tc;
    if (blocks > 0x7fffffffL) return RAYSHT_EINVAL;
    const size_t lds = sizeof(double) * 2 * tc * (cb->k4 + 1);
    hipLaunchKernelGGL(cheb_from_spectral_kernel, dim3((unsigned)blocks), dim3(64 * waves), lds,
                       reinterpret_cast<hipStream_t>(stream), cb->n, cb->n_x, cb->n_even, cb->n_odd, cb->k4, cb->ip,
                       ncols, cb->d_t_e, cb->d_t_o, c, f);
    RAYSHT_HIP_TRY(hipGetLastError());
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
    // deriv_scaling**dorder as the compiled Fortran power forms it for small orders
    const double s = cb->deriv_scaling;
    const double scale = dorder == 1 ? s : dorder == 2 ? s * s : dorder == 3 ? (s * s) * s : (s * s) * (s * s);
    hipLaunchKernelGGL(cheb_deriv_kernel, dim3((unsigned)blocks), dim3(kDerivCols), 0,
                       reinterpret_cast<hipStream_t>(stream), cb->n, ncols, dorder, scale, c, d);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_cheb_dealias(raysht_cheb *cb, double *c, long ncols, void *stream)
{
    if (!cb || ncols < 0 || (ncols > 0 && !c)) return RAYSHT_EINVAL;
    const int rda = (2 * cb->n) / 3 + 1;  // Initialize_Cheby_Grid :128
    const int first = rda - 1;
    if (ncols == 0 || first >= cb->n) return RAYSHT_OK;
    const long total = ncols * (cb->n - first);
    const long blocks = std::min<long>((total + 255) / 256, 1L << 20);
    hipLaunchKernelGGL(cheb_dealias_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), cb->n, first, ncols, c);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

}  // extern "C"
