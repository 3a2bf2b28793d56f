// FP64-MFMA Legendre contractions: every m-block of uneven length in one grouped launch.
//
// StP (Legendre_Transforms.F90:968-1074): for each m the reference forms
//   E = p_lm_even^T . s_even   and   O = p_lm_odd^T . s_odd       (nt/2 x nrhs each)
// with one DGEMM per parity and writes out(j) = E + O, out(nt+1-j) = E - O.
// PtS (:1077-1175): folds f(j) +/- f(nt+1-j), contracts with the integration-weighted
// tables and scatters even/odd l back into the interleaved l column.
//
// Here one workgroup owns an output tile of one m-block and computes BOTH parities
// for it, so the fold / reflection and the l gather / scatter happen in registers
// and LDS instead of separate passes through HBM.  Work items for all m-blocks are
// listed on the host (largest contraction first) and launched as one grid.
//
// MFMA used: v_mfma_f64_16x16x4_f64 (wave64).  Per lane: A = A[i][k] with
// i = lane%16, k = lane/16; B = B[k][j] with j = lane%16, k = lane/16; the 16x16
// result holds D[i][j] at i = lane/16 + 4*v, j = lane%16 (v = 0..3) -- the rows
// of the FP64 result are interleaved by 4, unlike the FP32 MFMA result layout.
// Both kernels compute the transposed product (columns of the right-hand side on
// the i axis) so that the 16 consecutive lanes of a store walk the unit-stride
// axis of the output (theta for StP, l for PtS).
#include "raysht_internal.h"

namespace raysht {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double4_t mfma_f64(double a, double b, double4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kThreads = 256;

// ---------------------------------------------------------------- StP
// Workgroup tile: 128 columns (c) x 64 southern nodes (j), both parities.
// Wave w owns columns [32w, 32w+32): 2 column blocks x 4 node blocks x 2 parities
// = 16 accumulators of 16x16.  K runs over chunks of 16 parity-l values (32 l);
// the next chunk is fetched into registers while the current one is multiplied
// out of LDS.
constexpr int STP_TC = 128;
constexpr int STP_TJ = 64;
constexpr int STP_KC = 16;
constexpr int STP_BS_LD = 2 * STP_KC + 1;  // Bs row stride (doubles)
constexpr int STP_A_LD = STP_TJ + 16;      // Ae/Ao row stride (doubles)
constexpr int STP_B_PER_T = STP_TC * 2 * STP_KC / kThreads;  // 16
constexpr int STP_A_PER_T = STP_KC * STP_TJ / kThreads;      // 4 (per parity)

__global__ __launch_bounds__(kThreads, 1) void stp_fp64_kernel(
    const double *__restrict__ s, double *__restrict__ p2, const double *__restrict__ pte,
    const double *__restrict__ pto, const long *__restrict__ s_off, const long *__restrict__ te_off,
    const long *__restrict__ to_off, const int *__restrict__ mvals, const work_item *__restrict__ work, int lmax,
    int nt, int nt2, int nrhs)
{
    __shared__ double Ae[STP_KC * STP_A_LD];
    __shared__ double Ao[STP_KC * STP_A_LD];
    __shared__ double Bs[STP_TC * STP_BS_LD];

    const work_item w = work[blockIdx.x];
    if (w.mp < 0) return;  // padding item of the XCD-balanced schedule
    const int mp = w.mp, j0 = w.r0, c0 = w.c0;
    const int m = mvals[mp];
    const int nl = lmax - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const double *sb = s + s_off[mp];
    const double *pe = pte + te_off[mp];
    const double *po = pto + to_off[mp];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;

    double rb[STP_B_PER_T], rae[STP_A_PER_T], rao[STP_A_PER_T];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < STP_B_PER_T; ++i) {
            const int idx = tid + i * kThreads;
            const int cc = idx / (2 * STP_KC), q = idx % (2 * STP_KC);
            const int c = c0 + cc, l = 2 * k0 + q;
            rb[i] = (c < nrhs && l < nl) ? sb[(long)c * nl + l] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < STP_A_PER_T; ++i) {
            const int idx = tid + i * kThreads;
            const int kk = idx / STP_TJ, jj = idx % STP_TJ;
            const int k = k0 + kk, j = j0 + jj;
            const bool jin = j < nt2;
            rae[i] = (jin && k < ne) ? pe[(long)k * nt2 + j] : 0.0;
            rao[i] = (jin && k < no) ? po[(long)k * nt2 + j] : 0.0;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < STP_B_PER_T; ++i) {
            const int idx = tid + i * kThreads;
            Bs[(idx / (2 * STP_KC)) * STP_BS_LD + idx % (2 * STP_KC)] = rb[i];
        }
#pragma unroll
        for (int i = 0; i < STP_A_PER_T; ++i) {
            const int idx = tid + i * kThreads;
            Ae[(idx / STP_TJ) * STP_A_LD + idx % STP_TJ] = rae[i];
            Ao[(idx / STP_TJ) * STP_A_LD + idx % STP_TJ] = rao[i];
        }
    };

    double4_t acc_e[2][4], acc_o[2][4];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            acc_e[cb][b] = double4_t{0.0, 0.0, 0.0, 0.0};
            acc_o[cb][b] = double4_t{0.0, 0.0, 0.0, 0.0};
        }

    fetch(0);
    for (int k0 = 0; k0 < ne; k0 += STP_KC) {
        stash();
        __syncthreads();
        if (k0 + STP_KC < ne) fetch(k0 + STP_KC);
#pragma unroll
        for (int kk = 0; kk < STP_KC; kk += 4) {
            double be[4], bo[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                be[b] = Ae[(kk + lg) * STP_A_LD + 16 * b + li];
                bo[b] = Ao[(kk + lg) * STP_A_LD + 16 * b + li];
            }
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const int row = (32 * wave + 16 * cb + li) * STP_BS_LD + 2 * (kk + lg);
                const double a_e = Bs[row];
                const double a_o = Bs[row + 1];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    acc_e[cb][b] = mfma_f64(a_e, be[b], acc_e[cb][b]);
                    acc_o[cb][b] = mfma_f64(a_o, bo[b], acc_o[cb][b]);
                }
            }
        }
        __syncthreads();
    }

    double *ob = p2 + (long)mp * nrhs * nt;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = j0 + 16 * b + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = c0 + 32 * wave + 16 * cb + lg + 4 * v;
                if (c < nrhs && j < nt2) {
                    const double e = acc_e[cb][b][v], o = acc_o[cb][b][v];
                    ob[(long)c * nt + j] = e + o;
                    ob[(long)c * nt + (nt - 1 - j)] = e - o;
                }
            }
        }
}

// ---------------------------------------------------------------- PtS
// Workgroup tile: 128 columns (c) x 64 parity-l values of each parity (128 l).
// Wave w owns columns [32w, 32w+32): 2 column blocks x 4 l blocks x 2 parities.
// K runs over chunks of 16 southern nodes; the fold f(j) +/- f(nt+1-j) is formed
// on the way into LDS, and the next chunk is fetched while this one is multiplied.
constexpr int PTS_TC = 128;
constexpr int PTS_TK = 64;
constexpr int PTS_KJ = 16;
constexpr int PTS_F_LD = PTS_KJ + 1;   // Fe/Fo row stride (doubles)
constexpr int PTS_I_LD = PTS_TK + 4;   // Ie/Io row stride (doubles)
constexpr int PTS_F_PER_T = PTS_TC * PTS_KJ / kThreads;  // 8
constexpr int PTS_I_PER_T = PTS_KJ * PTS_TK / kThreads;  // 4 (per parity)

__global__ __launch_bounds__(kThreads, 1) void pts_fp64_kernel(
    const double *__restrict__ p2, double *__restrict__ s, const double *__restrict__ ipe,
    const double *__restrict__ ipo, const long *__restrict__ s_off, const long *__restrict__ te_off,
    const long *__restrict__ to_off, const int *__restrict__ mvals, const work_item *__restrict__ work, int lmax,
    int nt, int nt2, int nrhs)
{
    __shared__ double Fe[PTS_TC * PTS_F_LD];
    __shared__ double Fo[PTS_TC * PTS_F_LD];
    __shared__ double Ie[PTS_KJ * PTS_I_LD];
    __shared__ double Io[PTS_KJ * PTS_I_LD];

    const work_item w = work[blockIdx.x];
    if (w.mp < 0) return;  // padding item of the XCD-balanced schedule
    const int mp = w.mp, k0 = w.r0, c0 = w.c0;
    const int m = mvals[mp];
    const int nl = lmax - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const double *ib = p2 + (long)mp * nrhs * nt;
    const double *ie = ipe + te_off[mp];
    const double *io = ipo + to_off[mp];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;

    double rfa[PTS_F_PER_T], rfb[PTS_F_PER_T], rie[PTS_I_PER_T], rio[PTS_I_PER_T];
    auto fetch = [&](int j0) {
#pragma unroll
        for (int i = 0; i < PTS_F_PER_T; ++i) {
            const int idx = tid + i * kThreads;
            const int cc = idx / PTS_KJ, jj = idx % PTS_KJ;
            const int c = c0 + cc, j = j0 + jj;
            const bool in = c < nrhs && j < nt2;
            rfa[i] = in ? ib[(long)c * nt + j] : 0.0;
            rfb[i] = in ? ib[(long)c * nt + (nt - 1 - j)] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < PTS_I_PER_T; ++i) {
            const int idx = tid + i * kThreads;
            const int kk = idx / PTS_KJ, jj = idx % PTS_KJ;
            const int k = k0 + kk, j = j0 + jj;
            const bool jin = j < nt2;
            rie[i] = (jin && k < ne) ? ie[(long)k * nt2 + j] : 0.0;
            rio[i] = (jin && k < no) ? io[(long)k * nt2 + j] : 0.0;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < PTS_F_PER_T; ++i) {
            const int idx = tid + i * kThreads;
            const int at = (idx / PTS_KJ) * PTS_F_LD + idx % PTS_KJ;
            Fe[at] = rfa[i] + rfb[i];
            Fo[at] = rfa[i] - rfb[i];
        }
#pragma unroll
        for (int i = 0; i < PTS_I_PER_T; ++i) {
            const int idx = tid + i * kThreads;
            const int at = (idx % PTS_KJ) * PTS_I_LD + idx / PTS_KJ;
            Ie[at] = rie[i];
            Io[at] = rio[i];
        }
    };

    double4_t acc_e[2][4], acc_o[2][4];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            acc_e[cb][b] = double4_t{0.0, 0.0, 0.0, 0.0};
            acc_o[cb][b] = double4_t{0.0, 0.0, 0.0, 0.0};
        }

    fetch(0);
    for (int j0 = 0; j0 < nt2; j0 += PTS_KJ) {
        stash();
        __syncthreads();
        if (j0 + PTS_KJ < nt2) fetch(j0 + PTS_KJ);
#pragma unroll
        for (int jj = 0; jj < PTS_KJ; jj += 4) {
            double be[4], bo[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                be[b] = Ie[(jj + lg) * PTS_I_LD + 16 * b + li];
                bo[b] = Io[(jj + lg) * PTS_I_LD + 16 * b + li];
            }
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const int row = (32 * wave + 16 * cb + li) * PTS_F_LD + jj + lg;
                const double a_e = Fe[row];
                const double a_o = Fo[row];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    acc_e[cb][b] = mfma_f64(a_e, be[b], acc_e[cb][b]);
                    acc_o[cb][b] = mfma_f64(a_o, bo[b], acc_o[cb][b]);
                }
            }
        }
        __syncthreads();
    }

    double *sb = s + s_off[mp];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int k = k0 + 16 * b + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = c0 + 32 * wave + 16 * cb + lg + 4 * v;
                if (c < nrhs) {
                    if (k < ne) sb[(long)c * nl + 2 * k] = acc_e[cb][b][v];
                    if (k < no) sb[(long)c * nl + 2 * k + 1] = acc_o[cb][b][v];
                }
            }
        }
}

}  // namespace

int launch_stp_fp64(raysht_plan *p, const double *s, double *p2, hipStream_t st)
{
    if (p->n_work_stp == 0) return RAYSHT_OK;
    hipLaunchKernelGGL(stp_fp64_kernel, dim3(p->n_work_stp), dim3(kThreads), 0, st, s, p2, p->d_pt_even,
                       p->d_pt_odd, p->d_s_off, p->d_te_off, p->d_to_off, p->d_mvals, p->d_work_stp, p->lmax, p->nt,
                       p->nt2, p->nrhs);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_pts_fp64(raysht_plan *p, const double *p2, double *s, hipStream_t st)
{
    if (p->n_work_pts == 0) return RAYSHT_OK;
    hipLaunchKernelGGL(pts_fp64_kernel, dim3(p->n_work_pts), dim3(kThreads), 0, st, p2, s, p->d_ip_even,
                       p->d_ip_odd, p->d_s_off, p->d_te_off, p->d_to_off, p->d_mvals, p->d_work_pts, p->lmax, p->nt,
                       p->nt2, p->nrhs);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

// Tile geometry the host work lists are built with.
int stp_tile_rows() { return STP_TJ; }
int stp_tile_cols() { return STP_TC; }
int pts_tile_rows() { return PTS_TK; }
int pts_tile_cols() { return PTS_TC; }

}  // namespace raysht
