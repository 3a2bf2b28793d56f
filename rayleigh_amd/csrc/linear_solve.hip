// Batched dense LU factorisation and solve for the implicit radial step.
//
// Implicit_Solve (Linear_Solve.F90:423-457) factors each mode's n x n matrix once per time
// step size (LU_Decompose_full = LAPACK dgetrf, :1034-1045) and every step solves it for
// the right-hand sides of all m of that l (LU_Solve_full = dgetrs 'N', :1087-1107).  Here
// many matrices go in one launch, one workgroup per matrix (factorisation) or per
// matrix and chunk of its right-hand sides (solve).
//
// Factorisation: the unblocked right-looking algorithm of dgetf2 -- pivot = first row of
// largest |a_ik| (a workgroup reduction), row swap, subdiagonal scaled by the pivot's
// reciprocal, rank-1 update of the trailing block -- with products and sums rounded
// separately.  The matrix stays in global memory (L2); per column the pivot row goes to
// LDS and each thread sweeps whole rows of the trailing block (coalesced across threads,
// sixteen loads in flight per thread).
// Solve: the row interchanges, the unit lower and the upper triangular solves in the
// column order of the reference BLAS dtrsm (column j final, then subtracted from the rows
// below / above it).
#include <algorithm>
#include <cmath>

#include "raysht_internal.h"

namespace raysht {
namespace {

constexpr int kLuThreads = 256;
constexpr int kLuMaxN = 1024;
constexpr int kLuBatch = 16;  // row loads in flight per thread in the factorisation sweep (4: 1.3x slower)

// Workgroup order of the solves.  Workgroups go to the 8 XCDs round-robin by linear id, so
// XCD k runs ids k, k + 8, k + 16, ...; its r-th workgroup takes column block r % ncb of
// matrix 8 (r / ncb) + k.  Each XCD then works through its own matrices (every eighth)
// one at a time, all column blocks of a matrix together, and the factor they share stays
// in that XCD's L2 instead of every XCD streaming every matrix from the last-level cache.
// The grid is ncb * (nbatch rounded up to 8) workgroups; ids past nbatch leave at once.
// Ungrouped (grouped = 0), matrix b runs as workgroups b, b + nbatch, ...: for small
// factors (dense n <= 256, 0.5 MB) measured faster, as the cache lines the waves read at
// one step spread over many matrices.
struct solve_block {
    int b, cb;
};
__device__ __forceinline__ solve_block solve_block_of(int ncb, int nbatch, int grouped)
{
    if (!grouped) return solve_block{(int)(blockIdx.x % (unsigned)nbatch), (int)(blockIdx.x / (unsigned)nbatch)};
    const int r = (int)(blockIdx.x >> 3);
    return solve_block{8 * (r / ncb) + (int)(blockIdx.x & 7), r % ncb};
}
inline long solve_grid(int ncb, int nbatch, int grouped)
{
    return (long)ncb * (grouped ? (nbatch + 7L) / 8 * 8 : (long)nbatch);
}
// a launch stays below 2^32 threads in all
inline bool solve_grid_fits(int ncb, int nbatch, int grouped, int threads)
{
    return solve_grid(ncb, nbatch, grouped) * threads < (1L << 32);
}

__global__ __launch_bounds__(kLuThreads) void lu_decompose_kernel(int n, double *__restrict__ mats,
                                                                  int *__restrict__ pivots, int *__restrict__ infos)
{
#pragma clang fp contract(off)
    __shared__ double s_val[kLuThreads];
    __shared__ int s_idx[kLuThreads];
    __shared__ double rowk[kLuMaxN];  // the pivot row, columns k + 1 .. n - 1
    double *a = mats + (long)blockIdx.x * n * n;
    int *piv = pivots + (long)blockIdx.x * n;
    const int tid = threadIdx.x;
    const double sfmin = 2.2250738585072014e-308;
    int info = 0;
    for (int k = 0; k < n; ++k) {
        // pivot: the first row of largest |a_ik|, i >= k
        double best = -1.0;
        int bi = n;
        for (int i = k + tid; i < n; i += kLuThreads) {
            const double v = fabs(a[i + (long)k * n]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        s_val[tid] = best;
        s_idx[tid] = bi;
        __syncthreads();
        for (int s = kLuThreads / 2; s > 0; s >>= 1) {
            if (tid < s) {
                const double v = s_val[tid + s];
                const int j = s_idx[tid + s];
                if (v > s_val[tid] || (v == s_val[tid] && j < s_idx[tid])) {
                    s_val[tid] = v;
                    s_idx[tid] = j;
                }
            }
            __syncthreads();
        }
        // idamax: a NaN a_kk is kept (nothing compares larger than it), as is row k of a
        // column of NaNs; NaNs below the diagonal are passed over
        const int p = (s_idx[0] >= n || isnan(a[k + (long)k * n])) ? k : s_idx[0];
        const double pv = a[p + (long)k * n];
        __syncthreads();  // every thread has read the pivot before the swap moves it
        if (tid == 0) piv[k] = p + 1;
        if (pv != 0.0) {
            if (p != k)
                for (int j = tid; j < n; j += kLuThreads) {
                    const double t = a[k + (long)j * n];
                    a[k + (long)j * n] = a[p + (long)j * n];
                    a[p + (long)j * n] = t;
                }
        } else if (info == 0) {
            info = k + 1;
        }
        __syncthreads();
        for (int j = k + 1 + tid; j < n; j += kLuThreads) rowk[j] = a[k + (long)j * n];
        __syncthreads();
        // each thread: its rows i > k -- scale a_ik by the pivot's reciprocal, then sweep the
        // row, a_ij + a_ik (-a_kj), four columns at a time (loads issued together)
        const double r = pv != 0.0 && fabs(pv) >= sfmin ? 1.0 / pv : 0.0;
        for (int i = k + 1 + tid; i < n; i += kLuThreads) {
            double lik = a[i + (long)k * n];
            if (pv != 0.0) {
                lik = fabs(pv) >= sfmin ? lik * r : lik / pv;
                a[i + (long)k * n] = lik;
            }
            double *ai = a + i;
            int j = k + 1;
            for (; j + kLuBatch - 1 < n; j += kLuBatch) {
                double v[kLuBatch];
#pragma unroll
                for (int u = 0; u < kLuBatch; ++u) v[u] = ai[(long)(j + u) * n];
#pragma unroll
                for (int u = 0; u < kLuBatch; ++u) {
                    const double y = rowk[j + u];
                    if (y != 0.0) ai[(long)(j + u) * n] = v[u] + lik * -y;
                }
            }
            for (; j < n; ++j) {
                const double y = rowk[j];
                if (y != 0.0) ai[(long)j * n] = ai[(long)j * n] + lik * -y;
            }
        }
        __syncthreads();
    }
    if (tid == 0) infos[blockIdx.x] = info;
}

// Blocked form of the same factorisation (n > kLuNB): panels of kLuNB columns.  Within a
// panel, each column step is the one above (pivot, interchange of the whole row, scaling),
// with the rank-1 update restricted to the panel's columns; the rest of the matrix then
// takes the panel's kLuNB updates at once: U12 (the panel's rows, later columns) by forward
// substitution with L11, and the trailing block by register tiles that apply the updates
// k = k0 .. k0 + kLuNB - 1 one after the other, a_ij + l_ik (-u_kj), product and sum rounded
// separately and skipped where u_kj = 0 (dger).  Every element thus sees the same
// operations, in the same order, as in the column-by-column sweep (an interchange moves a
// row's multipliers with it, so applying the panel's updates after the panel's
// interchanges is the same arithmetic), and the factors are bit-identical; but the
// trailing block is read and written once per panel instead of once per column, which
// was 0.72 GB per matrix of n = 512.
constexpr int kLuNB = 32;  // panel width
constexpr int kLuT = 64;   // trailing-update chunk: kLuT x kLuT, 4 x 4 per thread

__device__ __forceinline__ void lu_pivot_reduce(double &best, int &bi, double *s_val, int *s_idx)
{
    // first row of the largest |a|: (value, index) pairs ordered by value, then by lower index
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double v = __shfl_down(best, off, 64);
        const int j = __shfl_down(bi, off, 64);
        if (v > best || (v == best && j < bi)) {
            best = v;
            bi = j;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_val[wave] = best;
        s_idx[wave] = bi;
    }
    __syncthreads();
    best = s_val[0];
    bi = s_idx[0];
#pragma unroll
    for (int w = 1; w < kLuThreads / 64; ++w) {
        const double v = s_val[w];
        const int j = s_idx[w];
        if (v > best || (v == best && j < bi)) {
            best = v;
            bi = j;
        }
    }
}

__global__ __launch_bounds__(kLuThreads) void lu_decompose_blocked_kernel(int n, double *__restrict__ mats,
                                                                          int *__restrict__ pivots,
                                                                          int *__restrict__ infos)
{
#pragma clang fp contract(off)
    __shared__ double s_val[kLuThreads / 64];
    __shared__ int s_idx[kLuThreads / 64];
    __shared__ double rowk[kLuNB];            // the pivot row within the panel
    __shared__ double s_l[kLuNB][kLuT];       // L21 chunk [k][row]
    __shared__ double s_u[kLuNB][kLuT];       // U12 chunk [k][column]
    __shared__ double s_l11[kLuNB][kLuNB + 1];  // L11 [i][k]
    double *a = mats + (long)blockIdx.x * n * n;
    int *piv = pivots + (long)blockIdx.x * n;
    const int tid = threadIdx.x;
    const double sfmin = 2.2250738585072014e-308;
    int info = 0;
    for (int k0 = 0; k0 < n; k0 += kLuNB) {
        const int kend = min(n, k0 + kLuNB);
        // ---- the panel, column by column
        for (int k = k0; k < kend; ++k) {
            double best = -1.0;
            int bi = n;
            for (int i = k + tid; i < n; i += kLuThreads) {
                const double v = fabs(a[i + (long)k * n]);
                if (v > best) {
                    best = v;
                    bi = i;
                }
            }
            lu_pivot_reduce(best, bi, s_val, s_idx);
            // idamax: a NaN a_kk is kept, as is row k of a column of NaNs
            const int p = (bi >= n || isnan(a[k + (long)k * n])) ? k : bi;
            const double pv = a[p + (long)k * n];
            __syncthreads();  // every thread has read the pivot before the swap moves it
            if (tid == 0) piv[k] = p + 1;
            if (pv != 0.0) {
                if (p != k)
                    for (int j = tid; j < n; j += kLuThreads) {
                        const double t = a[k + (long)j * n];
                        a[k + (long)j * n] = a[p + (long)j * n];
                        a[p + (long)j * n] = t;
                    }
            } else if (info == 0) {
                info = k + 1;
            }
            __syncthreads();
            if (tid < kend - k - 1) rowk[tid] = a[k + (long)(k + 1 + tid) * n];
            __syncthreads();
            const double r = pv != 0.0 && fabs(pv) >= sfmin ? 1.0 / pv : 0.0;
            for (int i = k + 1 + tid; i < n; i += kLuThreads) {
                double lik = a[i + (long)k * n];
                if (pv != 0.0) {
                    lik = fabs(pv) >= sfmin ? lik * r : lik / pv;
                    a[i + (long)k * n] = lik;
                }
                double *ai = a + i;
                for (int j = k + 1; j < kend; ++j) {
                    const double y = rowk[j - k - 1];
                    if (y != 0.0) ai[(long)j * n] = ai[(long)j * n] + lik * -y;
                }
            }
            __syncthreads();
        }
        if (kend >= n) break;
        // ---- U12: rows k0 .. k0 + kLuNB - 1 of the later columns, forward substitution
        for (int e = tid; e < kLuNB * kLuNB; e += kLuThreads) {
            const int i = e % kLuNB, kk = e / kLuNB;
            s_l11[i][kk] = a[k0 + i + (long)(k0 + kk) * n];
        }
        __syncthreads();
        // a half wave per four columns: lane i % 32 holds row k0 + i of each, and step kk takes
        // u_kk from lane kk of its half (the row-kk value is final once steps 0 .. kk - 1 are
        // done); the four columns' shuffle chains are independent, so their latencies overlap
        for (int j0 = kend + 4 * (tid / 32); j0 < n; j0 += 4 * (kLuThreads / 32)) {
            const int i = tid % 32, half = tid & 32;
            double u[4];
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) u[cc] = j0 + cc < n ? a[k0 + i + (long)(j0 + cc) * n] : 0.0;
            for (int kk = 0; kk < kLuNB - 1; ++kk) {
                const double lik = s_l11[i][kk];
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    const double y = __shfl(u[cc], half | kk, 64);
                    const double upd = u[cc] + lik * -y;
                    u[cc] = (i > kk && y != 0.0) ? upd : u[cc];
                }
            }
#pragma unroll
            for (int cc = 0; cc < 4; ++cc)
                if (j0 + cc < n) a[k0 + i + (long)(j0 + cc) * n] = u[cc];
        }
        __syncthreads();
        // ---- the trailing block, kLuT x kLuT chunks
        const int tr = tid / 16, tc = tid % 16;
        for (int r0 = kend; r0 < n; r0 += kLuT) {
            for (int e = tid; e < kLuNB * kLuT; e += kLuThreads) {
                const int row = e % kLuT, kk = e / kLuT;
                s_l[kk][row] = r0 + row < n ? a[r0 + row + (long)(k0 + kk) * n] : 0.0;
            }
            for (int c0 = kend; c0 < n; c0 += kLuT) {
                for (int e = tid; e < kLuNB * kLuT; e += kLuThreads) {
                    const int col = e / kLuNB, kk = e % kLuNB;
                    s_u[kk][col] = c0 + col < n ? a[k0 + kk + (long)(c0 + col) * n] : 0.0;
                }
                __syncthreads();
                const int i0 = r0 + 4 * tr, j0 = c0 + 4 * tc;
                double t[4][4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr)
                        t[rr][c] = (i0 + rr < n && j0 + c < n) ? a[i0 + rr + (long)(j0 + c) * n] : 0.0;
#pragma unroll 4
                for (int kk = 0; kk < kLuNB; ++kk) {
                    double l[4], y[4];
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) l[rr] = s_l[kk][4 * tr + rr];
#pragma unroll
                    for (int c = 0; c < 4; ++c) y[c] = s_u[kk][4 * tc + c];
                    // dger skips a zero u_kj: a select, not a branch per column
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const bool nz = y[c] != 0.0;
#pragma unroll
                        for (int rr = 0; rr < 4; ++rr) {
                            const double upd = t[rr][c] + l[rr] * -y[c];
                            t[rr][c] = nz ? upd : t[rr][c];
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr)
                        if (i0 + rr < n && j0 + c < n) a[i0 + rr + (long)(j0 + c) * n] = t[rr][c];
                __syncthreads();  // the chunk's U12 and (at the row chunk's end) L21 are free again
            }
        }
        __syncthreads();
    }
    if (tid == 0) infos[blockIdx.x] = info;
}

// Solve with the right-hand sides in registers: G lanes of a wave own CPL columns, lane
// r holds rows r, r + G, ... (NT of them) of each.  The interchanges are composed into one
// permutation applied when the column is loaded; then each step j takes x_j from its
// owner lane by a lane shuffle and updates the lane's rows in registers, while the next
// column of L (or U) is loaded -- no barrier after the permutation.  Operation order and
// rounding are those of the reference BLAS order.
template <int G, int NT, int CPL>
__global__ __launch_bounds__(kLuThreads) void lu_solve_lanes_kernel(int n, int nbatch, int ncb, int grouped,
                                                                    const double *__restrict__ mats,
                                                                    const int *__restrict__ pivots,
                                                                    double *__restrict__ rhs,
                                                                    const long *__restrict__ rhs_off,
                                                                    const int *__restrict__ nrhs)
{
#pragma clang fp contract(off)
    __shared__ int perm[G * NT];
    constexpr int kCols = kLuThreads / G * CPL;
    const solve_block sb = solve_block_of(ncb, nbatch, grouped);
    if (sb.b >= nbatch) return;
    const int b = sb.b, ncol = nrhs[b], col0 = sb.cb * kCols;
    if (col0 >= ncol) return;  // uniform: the whole workgroup leaves
    const double *a = mats + (long)b * n * n;
    const int *piv = pivots + (long)b * n;
    const int tid = threadIdx.x;
    if (tid == 0) {
        for (int i = 0; i < n; ++i) perm[i] = i;
        for (int k = 0; k < n; ++k) {
            const int p = piv[k] - 1;
            if (p != k && p >= 0 && p < n) {
                const int t = perm[k];
                perm[k] = perm[p];
                perm[p] = t;
            }
        }
    }
    __syncthreads();
    const int lane = tid % G, cg = col0 + (tid / G) * CPL;
    double x[CPL][NT], cur[NT], nxt[NT];
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        const bool live = cg + q < ncol;
        const double *xg = rhs + rhs_off[b] + (long)(live ? cg + q : col0) * n;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int i = G * t + lane;
            x[q][t] = live && i < n ? xg[perm[i]] : 0.0;
        }
    }
    auto column = [&](int j, double(&v)[NT]) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int i = G * t + lane;
            v[t] = i < n ? a[i + (long)j * n] : 0.0;
        }
    };
    // unit lower: x_i -= x_j l_ij for i > j
    column(0, cur);
#pragma unroll
    for (int tb = 0; tb < NT; ++tb) {
        for (int r = 0; r < G; ++r) {
            const int j = G * tb + r;
            if (j >= n - 1) break;
            column(j + 1, nxt);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const double xj = __shfl(x[q][tb], r, G);
                if (xj != 0.0) {
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        if (G * t + lane > j) x[q][t] = x[q][t] - xj * cur[t];
                }
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) cur[t] = nxt[t];
        }
    }
    // upper: x_j /= u_jj, then x_i -= x_j u_ij for i < j; x_j is final
    column(n - 1, cur);
#pragma unroll
    for (int tb = NT - 1; tb >= 0; --tb) {
        for (int r = G - 1; r >= 0; --r) {
            const int j = G * tb + r;
            if (j >= n) continue;
            if (j > 0) column(j - 1, nxt);
            const double ujj = __shfl(cur[tb], r, G);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                double xj = __shfl(x[q][tb], r, G);
                if (xj != 0.0) {
                    xj = xj / ujj;
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        if (G * t + lane < j) x[q][t] = x[q][t] - xj * cur[t];
                }
                if (lane == r) x[q][tb] = xj;
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) cur[t] = nxt[t];
        }
    }
#pragma unroll
    for (int q = 0; q < CPL; ++q)
        if (cg + q < ncol) {
            double *xg = rhs + rhs_off[b] + (long)(cg + q) * n;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int i = G * t + lane;
                if (i < n) xg[i] = x[q][t];
            }
        }
}

// CPL right-hand sides per lane group share every loaded column of the factor (measured:
// 2 for 16-lane groups, 3 for 32- and 64-lane groups, within the register budget)
template <int G, int NT, int CPL>
int launch_lanes(int n, int nbatch, const double *mats, const int *pivots, double *rhs, const long *rhs_off,
                 const int *nrhs, int max_nrhs, hipStream_t st)
{
    constexpr int kCols = kLuThreads / G * CPL;
    const int ncb = (max_nrhs + kCols - 1) / kCols, grouped = n > 256;
    if (!solve_grid_fits(ncb, nbatch, grouped, kLuThreads)) return RAYSHT_EUNSUPPORTED;
    hipLaunchKernelGGL((lu_solve_lanes_kernel<G, NT, CPL>), dim3((unsigned)solve_grid(ncb, nbatch, grouped)),
                       dim3(kLuThreads), 0, st, n, nbatch, ncb, grouped, mats, pivots, rhs, rhs_off, nrhs);
    return RAYSHT_OK;
}

}  // namespace
}  // namespace raysht

using namespace raysht;

// Banded factorisation: LU_Decompose_band (Linear_Solve.F90:1041-1055, dgbtrf with
// kl = ku = (ldab - 1) / 3) in the unblocked order of dgbtf2, one workgroup per matrix in
// LAPACK band storage ab(ldab, n) (1-based AB(i, j) below; full element (r, c) at row
// kv + 1 + r - c of column c, kv = kl + ku).  Per column: the fill-in rows LAPACK clears, the
// pivot (first largest |a| of the kl + 1 candidates, idamax's rules for NaN), the interchange
// over columns j..ju, the multipliers scaled by the pivot's reciprocal, and the rank-1 update
// of the km x (ju - j) block, a column skipped where the pivot row's entry is 0 (dger);
// products and sums rounded separately, so the factors are the oracle's doubles.
constexpr int kBandMaxKv = 1024;
__global__ __launch_bounds__(kLuThreads) void band_lu_decompose_kernel(int n, int kl, int ku, int ldab,
                                                                       double *__restrict__ mats,
                                                                       int *__restrict__ pivots,
                                                                       int *__restrict__ infos)
{
#pragma clang fp contract(off)
    __shared__ double s_val[kLuThreads];
    __shared__ int s_idx[kLuThreads];
    __shared__ double s_y[kBandMaxKv + 1];  // the pivot row, columns j + 1 .. ju
    const int tid = threadIdx.x;
    double *ab = mats + (long)blockIdx.x * ldab * n;
    int *piv = pivots + (long)blockIdx.x * n;
    const int kv = ku + kl;
#define AB(i, j) ab[(long)((j) - 1) * ldab + ((i) - 1)]
    // dgbtf2's zeroed fill-in: rows max(1, kv - c + 2) .. kl of columns c = ku + 2 .. n
    for (long e = tid; e < (long)kl * n; e += kLuThreads) {
        const int i = 1 + (int)(e % kl), c = 1 + (int)(e / kl);
        if (c >= ku + 2 && i >= kv - c + 2) AB(i, c) = 0.0;
    }
    __syncthreads();
    int info = 0, ju = 1;
    for (int j = 1; j <= n; ++j) {
        const int km = min(kl, n - j);
        double best = -1.0;
        int bi = 1 << 30;
        for (int i = 1 + tid; i <= km + 1; i += kLuThreads) {
            const double v = fabs(AB(kv + i, j));
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        s_val[tid] = best;
        s_idx[tid] = bi;
        __syncthreads();
        for (int sh = kLuThreads / 2; sh > 0; sh >>= 1) {
            if (tid < sh) {
                const double v = s_val[tid + sh];
                const int q = s_idx[tid + sh];
                if (v > s_val[tid] || (v == s_val[tid] && q < s_idx[tid])) {
                    s_val[tid] = v;
                    s_idx[tid] = q;
                }
            }
            __syncthreads();
        }
        // idamax: a NaN first entry is kept (nothing compares larger); later NaNs are passed over
        const int jp = (s_idx[0] > km + 1 || isnan(AB(kv + 1, j))) ? 1 : s_idx[0];
        const double pv = AB(kv + jp, j);
        __syncthreads();  // every thread has the pivot before the interchange moves it
        if (tid == 0) piv[j - 1] = jp + j - 1;
        if (pv != 0.0) {
            ju = max(ju, min(j + ku + jp - 1, n));
            if (jp != 1)
                for (int c = tid; c <= ju - j; c += kLuThreads) {
                    const double t = AB(kv + jp - c, j + c);
                    AB(kv + jp - c, j + c) = AB(kv + 1 - c, j + c);
                    AB(kv + 1 - c, j + c) = t;
                }
            __syncthreads();
            if (km > 0) {
                const double r = 1.0 / AB(kv + 1, j);
                for (int i = 1 + tid; i <= km; i += kLuThreads) AB(kv + 1 + i, j) = r * AB(kv + 1 + i, j);
                for (int c = 1 + tid; c <= ju - j; c += kLuThreads) s_y[c] = AB(kv + 1 - c, j + c);
                __syncthreads();
                // element e = (i - 1) + km (c - 1), e = tid, tid + kLuThreads, ...: i and c stepped
                // without divisions, four elements per thread at a time with their operands (row
                // j and column j, which the update does not write) read before any is written
                const int w = ju - j, tot = km * w, di = kLuThreads % km, dc = kLuThreads / km;
                int i = 1 + tid % km, c = 1 + tid / km;
                constexpr int U = 4;
                for (int e = tid; e < tot; e += U * kLuThreads) {
                    int iu[U], cu[U];
                    bool on[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        on[u] = e + u * kLuThreads < tot;
                        iu[u] = on[u] ? i : 1;
                        cu[u] = on[u] ? c : 1;
                        i += di;
                        c += dc;
                        if (i > km) {
                            i -= km;
                            ++c;
                        }
                    }
                    double y[U], a[U], x[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        y[u] = s_y[cu[u]];
                        a[u] = AB(kv + 1 + iu[u] - cu[u], j + cu[u]);
                        x[u] = AB(kv + 1 + iu[u], j);
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (on[u] && y[u] != 0.0) AB(kv + 1 + iu[u] - cu[u], j + cu[u]) = a[u] + x[u] * (-y[u]);
                }
                __syncthreads();
            }
        } else if (info == 0) {
            info = j;
        }
    }
    if (tid == 0) infos[blockIdx.x] = info;
#undef AB
}

// The same factorisation with its active window in LDS.  Step j of dgbtf2 reads and writes
// only matrix rows j .. j + kl and columns j .. j + kv (the pivot search, the interchange
// over columns j .. ju <= j + kv, the multipliers and the rank-1 update), so that window,
// (kl + 1) x (kv + 1) values, lives in an LDS ring (matrix row % (kl + 1), column % (kv + 1),
// column-major) while the factorisation walks down the band: after step j, row j (U) and
// column j (L) are final and go to the band storage, and step j + 1 brings in row j + kl + 1
// (plain band entries) and column j + kv + 1 (rows j + 1 .. j + kl: dgbtf2's zeroed fill-in,
// so zeros are written, not read).  The arithmetic, its order and the pivot choice are the
// global-memory kernel's, statement for statement, so the factors and pivots are the same
// doubles; the difference is that the rank-1 updates no longer read and write the band
// storage through the caches on every step.  Used where the window and the reduction
// buffers fit one workgroup's LDS.
__global__ __launch_bounds__(kLuThreads) void band_lu_decompose_lds_kernel(int n, int kl, int ku, int ldab,
                                                                           double *__restrict__ mats,
                                                                           int *__restrict__ pivots,
                                                                           int *__restrict__ infos)
{
#pragma clang fp contract(off)
    extern __shared__ double win[];
    __shared__ double s_val[kLuThreads];
    __shared__ int s_idx[kLuThreads];
    const int tid = threadIdx.x;
    double *ab = mats + (long)blockIdx.x * ldab * n;
    int *piv = pivots + (long)blockIdx.x * n;
    const int kv = ku + kl, R = kl + 1, C = kv + 1;
#define AB(i, j) ab[(long)((j) - 1) * ldab + ((i) - 1)]
#define W(row, col) win[((col) % C) * R + ((row) % R)]
    // the first window, rows 1 .. min(R, n) x columns 1 .. min(C, n); entries more than ku
    // above the diagonal are fill-in, zero at the start
    {
        const int nrw = min(R, n), ncl = min(C, n);
        for (int e = tid; e < nrw * ncl; e += kLuThreads) {
            const int row = 1 + e % nrw, col = 1 + e / nrw;
            W(row, col) = col - row > ku ? 0.0 : AB(kv + 1 + row - col, col);
        }
    }
    __syncthreads();
    int info = 0, ju = 1;
    for (int j = 1; j <= n; ++j) {
        if (j > 1) {
            // row j + kl (columns j .. j + kv) and column j + kv (rows j .. j + kl - 1) into the
            // slots of row and column j - 1, written out at the end of the previous step
            const int rn = j + kl, cn = j + kv;
            if (rn <= n)
                for (int col = j + tid; col <= min(cn, n); col += kLuThreads) W(rn, col) = AB(kv + 1 + rn - col, col);
            if (cn <= n)
                for (int row = j + tid; row <= min(rn - 1, n); row += kLuThreads) W(row, cn) = 0.0;
            __syncthreads();
        }
        const int km = min(kl, n - j);
        double best = -1.0;
        int bi = 1 << 30;
        for (int i = 1 + tid; i <= km + 1; i += kLuThreads) {
            const double v = fabs(W(j + i - 1, j));
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        // the largest |a| with the lowest index among equals: a total order on the candidates
        // (NaNs never enter), so the shuffle and cross-wave order cannot change the choice
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double v = __shfl_xor(best, off);
            const int q = __shfl_xor(bi, off);
            if (v > best || (v == best && q < bi)) {
                best = v;
                bi = q;
            }
        }
        if ((tid & 63) == 0) {
            s_val[tid >> 6] = best;
            s_idx[tid >> 6] = bi;
        }
        __syncthreads();
        best = s_val[0];
        bi = s_idx[0];
#pragma unroll
        for (int w = 1; w < kLuThreads / 64; ++w)
            if (s_val[w] > best || (s_val[w] == best && s_idx[w] < bi)) {
                best = s_val[w];
                bi = s_idx[w];
            }
        // idamax: a NaN first entry is kept (nothing compares larger); later NaNs are passed over
        const int jp = (bi > km + 1 || isnan(W(j, j))) ? 1 : bi;
        const double pv = W(j + jp - 1, j);
        __syncthreads();  // every thread has the pivot before the interchange moves it
        if (tid == 0) piv[j - 1] = jp + j - 1;
        if (pv != 0.0) {
            ju = max(ju, min(j + ku + jp - 1, n));
            if (jp != 1)
                for (int c = tid; c <= ju - j; c += kLuThreads) {
                    const double t = W(j + jp - 1, j + c);
                    W(j + jp - 1, j + c) = W(j, j + c);
                    W(j, j + c) = t;
                }
            __syncthreads();
            if (km > 0) {
                const double r = 1.0 / W(j, j);
                for (int i = 1 + tid; i <= km; i += kLuThreads) W(j + i, j) = r * W(j + i, j);
                __syncthreads();
                // element e = (i - 1) + km (c - 1) of the update, e = tid, tid + kLuThreads, ...:
                // i, c and their ring slots stepped by kLuThreads elements without divisions
                const int w = ju - j, di = kLuThreads % km, dc = (kLuThreads / km) % C;
                int ii = tid % km, rs = (j + 1 + ii) % R, cs = (j + 1 + tid / km) % C;
                const int js = j % R, jc = (j % C) * R, tot = km * w;
                // four elements per thread at a time: their operands (row j and column j are
                // not written by the update) are read before any of the four is written
                constexpr int U = 4;
                for (int e = tid; e < tot; e += U * kLuThreads) {
                    int rsu[U], csu[U];
                    bool on[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        on[u] = e + u * kLuThreads < tot;
                        rsu[u] = on[u] ? rs : 0;
                        csu[u] = on[u] ? cs : 0;
                        ii += di;
                        rs += di;
                        if (rs >= R) rs -= R;
                        cs += dc;
                        if (ii >= km) {
                            ii -= km;
                            rs -= km;
                            if (rs < 0) rs += R;
                            ++cs;
                        }
                        if (cs >= C) cs -= C;
                    }
                    double y[U], a[U], x[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        y[u] = win[csu[u] * R + js];
                        a[u] = win[csu[u] * R + rsu[u]];
                        x[u] = win[jc + rsu[u]];
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (on[u] && y[u] != 0.0) win[csu[u] * R + rsu[u]] = a[u] + x[u] * (-y[u]);
                }
                __syncthreads();
            }
        } else if (info == 0) {
            info = j;
        }
        // row j (columns j .. j + kv) and column j (rows j + 1 .. j + kl) are final
        for (int col = j + tid; col <= min(j + kv, n); col += kLuThreads) AB(kv + 1 + j - col, col) = W(j, col);
        for (int row = j + 1 + tid; row <= min(j + kl, n); row += kLuThreads) AB(kv + 1 + row - j, j) = W(row, j);
        __syncthreads();
    }
    if (tid == 0) infos[blockIdx.x] = info;
#undef W
#undef AB
}

// Banded solve: LU_Solve_Band (Linear_Solve.F90:1056-1086, dgbtrs 'N').  One wave solves CPL
// right-hand-side columns of matrix blockIdx.y, held in registers across its lanes (row
// 64 s + lane in slot s, NS slots).  Per step j the wave reads x_j from its owner lane and
// updates only the slots the band reaches, with one coalesced load of the factor column shared
// by its CPL columns: the L solve (interchange, then dger's x(r) + l (-x(j)), skipped where
// x(j) = 0) and dtbsv's upper solve (x(j) / u(j, j), then x(r) - x(j) u(r, j) for the kl + ku
// rows above, skipped where x(j) = 0).  Factor entry (r, j) of both is band row kv + r - j.
// The slots rotate: while step j runs over the 64 rows of segment S = j / 64, slot k of a
// lane holds absolute slot (S + k) mod NS, so row j is always in slot 0 and every slot index
// in the loop bodies is a compile-time register index (a select over NS slots per access
// was most of the kernel's instructions).  Segments advance by one rotation of the NS
// slots, and NS rotations restore the natural order.
template <int NS>
__device__ __forceinline__ void slots_rotate_down(double (&x)[NS])  // slot k takes slot k + 1
{
    const double x0 = x[0];
#pragma unroll
    for (int k = 0; k + 1 < NS; ++k) x[k] = x[k + 1];
    x[NS - 1] = x0;
}

template <int NS>
__device__ __forceinline__ void slots_rotate_up(double (&x)[NS])  // slot k takes slot k - 1
{
    const double xl = x[NS - 1];
#pragma unroll
    for (int k = NS - 1; k > 0; --k) x[k] = x[k - 1];
    x[0] = xl;
}

// value of row `row`, in relative slot (row >> 6) - S (0 .. NS - 1 - S)
template <int NS>
__device__ __forceinline__ double band_row(const double (&x)[NS], int rel, int row)
{
    double v = x[0];
#pragma unroll
    for (int k = 1; k < NS; ++k)
        if (k == rel) v = x[k];
    return __shfl(v, row & 63);
}

template <int NS>
__device__ __forceinline__ void band_set(double (&x)[NS], int rel, int row, int lane, double v)
{
#pragma unroll
    for (int k = 0; k < NS; ++k)
        if (k == rel && lane == (row & 63)) x[k] = v;
}

template <int NS, int CPL>
__global__ __launch_bounds__(64) void band_lu_solve_kernel(int n, int kl, int ku, int ldab, int nbatch, int ncb,
                                                           const double *__restrict__ mats,
                                                           const int *__restrict__ pivots, double *__restrict__ rhs,
                                                           const long *__restrict__ rhs_off,
                                                           const int *__restrict__ nrhs)
{
#pragma clang fp contract(off)
    const solve_block sb = solve_block_of(ncb, nbatch, 1);
    if (sb.b >= nbatch) return;
    const int b = sb.b, lane = threadIdx.x, c0 = sb.cb * CPL;
    const int nc = min(CPL, nrhs[b] - c0);
    if (nc <= 0) return;
    const double *ab = mats + (long)b * ldab * n;
    const int *piv = pivots + (long)b * n;
    double *x0 = rhs + rhs_off[b] + (long)c0 * n;
    const int kv = kl + ku;
    double x[CPL][NS];
#pragma unroll
    for (int c = 0; c < CPL; ++c)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int r = s * 64 + lane;
            x[c][s] = (c < nc && r < n) ? x0[(long)c * n + r] : 0.0;
        }
    // L solve with the interchanges, segment by segment (row j in slot 0)
#pragma unroll 1
    for (int S = 0; S < NS; ++S) {
        const int jend = kl > 0 ? min(64 * (S + 1), n - 1) : 0;
        for (int j = 64 * S; j < jend; ++j) {
            const int jj = j & 63, lm = min(kl, n - 1 - j), l = piv[j] - 1, lrel = (l >> 6) - S;
            double y[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                if (l != j) {
                    const double xl = band_row(x[c], lrel, l), xj = __shfl(x[c][0], jj);
                    if (lane == jj) x[c][0] = xl;
                    band_set(x[c], lrel, l, lane, xj);
                }
                y[c] = __shfl(x[c][0], jj);
            }
            const int klo = ((j + 1) >> 6) - S, khi = ((j + lm) >> 6) - S;
            const double *col = ab + (long)j * ldab + kv - j;  // col[r] = factor entry (r, j)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                if (k < klo || k > khi) continue;
                const int r = (S + k) * 64 + lane;
                if (r <= j || r > j + lm) continue;
                const double a = col[r];
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (y[c] != 0.0) x[c][k] = x[c][k] + a * (-y[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < CPL; ++c) slots_rotate_down(x[c]);
    }
    // U solve, segments from the last down (row j in slot 0, rows above in slots NS - 1, ...)
#pragma unroll 1
    for (int S = NS - 1; S >= 0; --S) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) slots_rotate_up(x[c]);
        for (int j = min(64 * S + 63, n - 1); j >= 64 * S; --j) {
            const int jj = j & 63;
            const double *col = ab + (long)j * ldab + kv - j;
            const double d = col[j];
            double t[CPL];
            bool nz[CPL];  // dtbsv skips a column where x(j) is 0 before the division
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                t[c] = __shfl(x[c][0], jj);
                nz[c] = t[c] != 0.0;
                if (nz[c]) {
                    t[c] = t[c] / d;
                    if (lane == jj) x[c][0] = t[c];
                }
            }
            if (j == 0) continue;
            const int lo = max(0, j - kv), slo = lo >> 6, shi = (j - 1) >> 6;
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const int sa = k == 0 ? S : S - (NS - k);  // absolute slot of relative slot k
                if (sa < slo || sa > shi) continue;
                const int r = sa * 64 + lane;
                if (r < lo || r >= j) continue;
                const double a = col[r];
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (nz[c]) x[c][k] = x[c][k] - t[c] * a;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int r = s * 64 + lane;
            if (c < nc && r < n) x0[(long)c * n + r] = x[c][s];
        }
}

template <int NS, int CPL>
int launch_band_solve(int n, int kl, int ku, int nbatch, const double *mats, int ldab, const int *pivots,
                      double *rhs, const long *rhs_off, const int *nrhs, int max_nrhs, hipStream_t st)
{
    const int ncb = (max_nrhs + CPL - 1) / CPL;
    if (!solve_grid_fits(ncb, nbatch, 1, 64)) return RAYSHT_EUNSUPPORTED;
    hipLaunchKernelGGL((band_lu_solve_kernel<NS, CPL>), dim3((unsigned)solve_grid(ncb, nbatch, 1)), dim3(64), 0, st, n,
                       kl, ku, ldab, nbatch, ncb, mats, pivots, rhs, rhs_off, nrhs);
    return RAYSHT_OK;
}

extern "C" {

int raysht_lu_decompose_full(int n, int nbatch, double *mats, int *pivots, int *info, void *stream)
{
    if (n < 1 || nbatch < 0 || (nbatch > 0 && (!mats || !pivots || !info))) return RAYSHT_EINVAL;
    if (n > kLuMaxN) return RAYSHT_EUNSUPPORTED;
    if (nbatch == 0) return RAYSHT_OK;
    if (n > kLuNB)
        hipLaunchKernelGGL(lu_decompose_blocked_kernel, dim3((unsigned)nbatch), dim3(kLuThreads), 0,
                           reinterpret_cast<hipStream_t>(stream), n, mats, pivots, info);
    else
        hipLaunchKernelGGL(lu_decompose_kernel, dim3((unsigned)nbatch), dim3(kLuThreads), 0,
                           reinterpret_cast<hipStream_t>(stream), n, mats, pivots, info);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_lu_solve_full(int n, int nbatch, const double *mats, const int *pivots, double *rhs, const long *rhs_off,
                         const int *nrhs, int max_nrhs, void *stream)
{
    if (n < 1 || nbatch < 0 || max_nrhs < 0) return RAYSHT_EINVAL;
    if (nbatch > 0 && (!mats || !pivots || !rhs || !rhs_off || !nrhs)) return RAYSHT_EINVAL;
    if (n > kLuMaxN) return RAYSHT_EUNSUPPORTED;
    if (nbatch == 0 || max_nrhs == 0) return RAYSHT_OK;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int rc;
    if (n <= 64)
        rc = launch_lanes<16, 4, 2>(n, nbatch, mats, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    else if (n <= 128)
        rc = launch_lanes<16, 8, 2>(n, nbatch, mats, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    else if (n <= 256)
        rc = launch_lanes<16, 16, 2>(n, nbatch, mats, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    else if (n <= 512)
        rc = launch_lanes<32, 16, 3>(n, nbatch, mats, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    else
        rc = launch_lanes<64, 16, 3>(n, nbatch, mats, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}


int raysht_lu_decompose_band(int n, int kl, int ku, int nbatch, double *mats, int ldab, int *pivots, int *info,
                             void *stream)
{
    if (n < 1 || kl < 0 || ku < 0 || nbatch < 0 || ldab < 2 * kl + ku + 1) return RAYSHT_EINVAL;
    if (nbatch > 0 && (!mats || !pivots || !info)) return RAYSHT_EINVAL;
    if (kl + ku > kBandMaxKv) return RAYSHT_EUNSUPPORTED;
    if (nbatch == 0) return RAYSHT_OK;
    // the LDS-window kernel where its window and reduction buffers fit one workgroup's LDS
    const size_t win = sizeof(double) * (size_t)(kl + 1) * (kl + ku + 1);
    int dev = 0, lds_max = 0;
    RAYSHT_HIP_TRY(hipGetDevice(&dev));
    RAYSHT_HIP_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (win + kLuThreads * (sizeof(double) + sizeof(int)) <= (size_t)lds_max)
        hipLaunchKernelGGL(band_lu_decompose_lds_kernel, dim3((unsigned)nbatch), dim3(kLuThreads), win,
                           reinterpret_cast<hipStream_t>(stream), n, kl, ku, ldab, mats, pivots, info);
    else
        hipLaunchKernelGGL(band_lu_decompose_kernel, dim3((unsigned)nbatch), dim3(kLuThreads), 0,
                           reinterpret_cast<hipStream_t>(stream), n, kl, ku, ldab, mats, pivots, info);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_lu_solve_band(int n, int kl, int ku, int nbatch, const double *mats, int ldab, const int *pivots,
                         double *rhs, const long *rhs_off, const int *nrhs, int max_nrhs, void *stream)
{
    if (n < 1 || kl < 0 || ku < 0 || nbatch < 0 || max_nrhs < 0 || ldab < 2 * kl + ku + 1) return RAYSHT_EINVAL;
    if (nbatch > 0 && (!mats || !pivots || !rhs || !rhs_off || !nrhs)) return RAYSHT_EINVAL;
    if (n > kLuMaxN) return RAYSHT_EUNSUPPORTED;
    if (nbatch == 0 || max_nrhs == 0) return RAYSHT_OK;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int rc;
    if (n <= 64)
        rc = launch_band_solve<1, 8>(n, kl, ku, nbatch, mats, ldab, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    else if (n <= 128)
        rc = launch_band_solve<2, 8>(n, kl, ku, nbatch, mats, ldab, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    else if (n <= 256)
        rc = launch_band_solve<4, 4>(n, kl, ku, nbatch, mats, ldab, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    else if (n <= 512)
        rc = launch_band_solve<8, 4>(n, kl, ku, nbatch, mats, ldab, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    else
        rc = launch_band_solve<16, 2>(n, kl, ku, nbatch, mats, ldab, pivots, rhs, rhs_off, nrhs, max_nrhs, st);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

}  // extern "C"
