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
#include "uniform.h"

namespace raysht {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double4_t mfma_f64(double a, double b, double4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Tile geometry.  The defaults are the measured best on MI355X for the l_max = 511
// workload; they can be overridden at build time (-DRAYSHT_STP_THREADS=... etc.)
// when retuning.  Node / parity-l tiles are fixed at 64 (4 MFMA blocks).
#ifndef RAYSHT_STP_THREADS
#define RAYSHT_STP_THREADS 512
#endif
#ifndef RAYSHT_STP_CB
#define RAYSHT_STP_CB 1
#endif
#ifndef RAYSHT_STP_KC
#define RAYSHT_STP_KC 16
#endif
#ifndef RAYSHT_PTS_THREADS
#define RAYSHT_PTS_THREADS 512
#endif
#ifndef RAYSHT_PTS_TK
#define RAYSHT_PTS_TK 64
#endif
// minimum resident waves per SIMD the register allocation must allow: 4 = 128 VGPRs,
// two 8-wave workgroups per CU
#ifndef RAYSHT_FP64_MIN_WAVES
#define RAYSHT_FP64_MIN_WAVES 4
#endif

// Raw buffer access: loads whose byte offset falls outside [0, bytes) return 0
// without touching memory, so tile edges need no branches.  The whole offset goes in
// the VGPR operand: the SGPR offset of a buffer load is not range-checked.  A negative
// offset reads as a huge unsigned one, so it is out of range too.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const double *base, long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ double load_f64(__amdgpu_buffer_rsrc_t r, int voff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}

// ---------------------------------------------------------------- StP
// Workgroup tile: (waves x 16 x STP_CB) columns (c) x 64 southern nodes (j), both
// parities.  Wave w owns STP_CB blocks of 16 columns: STP_CB x 4 node blocks x 2
// parities of 16x16 accumulators.  K runs over chunks of STP_KC parity-l values;
// the next chunk is fetched into registers while the current one is multiplied
// out of LDS.
constexpr int STP_THREADS = RAYSHT_STP_THREADS;
constexpr int STP_CB = RAYSHT_STP_CB;  // 16-column blocks per wave
constexpr int STP_TC = (STP_THREADS / 64) * 16 * STP_CB;
constexpr int STP_TJ = 64;
constexpr int STP_KC = RAYSHT_STP_KC;
constexpr int STP_BS_LD = 2 * STP_KC + 1;  // Bs row stride (doubles)
constexpr int STP_A_LD = STP_TJ + 16;      // Ae/Ao row stride (doubles)
constexpr int STP_B_PER_T = STP_TC * 2 * STP_KC / STP_THREADS;
static_assert(STP_THREADS % 64 == 0 && STP_THREADS % (2 * STP_KC) == 0 && (STP_KC * STP_TJ) % STP_THREADS == 0 &&
                  STP_KC % 4 == 0,
              "StP tile geometry");
constexpr int STP_A_PER_T = STP_KC * STP_TJ / STP_THREADS;      // per parity

// One workgroup per work item; the hardware dispatcher balances the uneven items
// (a persistent variant that prefetched across items measured slower).
struct stp_item {
    int mp, j0, c0, nl, ne, no;
    __amdgpu_buffer_rsrc_t rs, rpe, rpo;
};

__global__ __launch_bounds__(STP_THREADS, RAYSHT_FP64_MIN_WAVES) void stp_fp64_kernel(
    const prob_desc *__restrict__ probs, const batch_io io, const work_item *__restrict__ work)
{
    __shared__ double Ae[STP_KC * STP_A_LD];
    __shared__ double Ao[STP_KC * STP_A_LD];
    __shared__ double Bs[STP_TC * STP_BS_LD];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    // Per-thread load geometry.  B: column bcc0 + (STP_THREADS / 2KC) i, l = 2 k0 + bq;
    // A: parity-l row akk0 + (STP_THREADS / 64) i, node ajj.
    const int bq = tid % (2 * STP_KC), bcc0 = tid / (2 * STP_KC);
    const int ajj = tid % STP_TJ, akk0 = tid / STP_TJ;

    const int mp = uniform(work[blockIdx.x].mp);
    if (mp < 0) return;  // padding item of the XCD-balanced schedule
    const int pi = uniform(work[blockIdx.x].pad) & 0xffff;
    const prob_desc &P = probs[pi];
    const int lmax = uniform(P.lmax), nt = uniform(P.nt), nt2 = uniform(P.nt2), nrhs = uniform(P.nrhs);
    double *p2 = uniform(io.out[pi]);
    stp_item cur;
    cur.mp = mp;
    cur.j0 = uniform(work[blockIdx.x].r0);
    cur.c0 = uniform(work[blockIdx.x].c0);
    cur.nl = lmax - uniform(uniform(P.mvals)[mp]) + 1;
    cur.ne = (cur.nl + 1) >> 1;
    cur.no = cur.nl >> 1;
    cur.rs = make_rsrc(uniform(io.in[pi]) + uniform(uniform(P.s_off)[mp]), (long)cur.nl * nrhs * 8);
    cur.rpe = make_rsrc(uniform(P.pt_even) + uniform(uniform(P.te_off)[mp]), (long)cur.ne * nt2 * 8);
    cur.rpo = make_rsrc(uniform(P.pt_odd) + uniform(uniform(P.to_off)[mp]), (long)cur.no * nt2 * 8);
    const int b_voff = ((cur.c0 + bcc0) * cur.nl + bq) * 8;
    const int b_sstep = (STP_THREADS / (2 * STP_KC)) * cur.nl * 8;
    const bool a_jin = cur.j0 + ajj < nt2;
    const int a_voff = (akk0 * nt2 + cur.j0 + ajj) * 8;
    const int a_sstep = (STP_THREADS / STP_TJ) * nt2 * 8;

    // The loads land in registers untouched; edge masks are applied when they are written
    // to LDS, so nothing waits for them before the current chunk is multiplied.
    double rb[STP_B_PER_T], rae[STP_A_PER_T], rao[STP_A_PER_T];
    auto fetch = [&](const stp_item &it, int k0) {
#pragma unroll
        for (int i = 0; i < STP_B_PER_T; ++i) rb[i] = load_f64(it.rs, b_voff + 2 * k0 * 8 + i * b_sstep);
        const int a_s0 = k0 * nt2 * 8;
#pragma unroll
        for (int i = 0; i < STP_A_PER_T; ++i) {
            rae[i] = load_f64(it.rpe, a_voff + a_s0 + i * a_sstep);
            rao[i] = load_f64(it.rpo, a_voff + a_s0 + i * a_sstep);
        }
    };
    auto stash = [&](int k0) {
        const bool lin = 2 * k0 + bq < cur.nl;  // rows past a column's last l belong to the next column
#pragma unroll
        for (int i = 0; i < STP_B_PER_T; ++i) {
            const int idx = tid + i * STP_THREADS;
            Bs[(idx / (2 * STP_KC)) * STP_BS_LD + idx % (2 * STP_KC)] = lin ? rb[i] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < STP_A_PER_T; ++i) {
            const int idx = tid + i * STP_THREADS;
            Ae[(idx / STP_TJ) * STP_A_LD + idx % STP_TJ] = a_jin ? rae[i] : 0.0;
            Ao[(idx / STP_TJ) * STP_A_LD + idx % STP_TJ] = a_jin ? rao[i] : 0.0;
        }
    };

    fetch(cur, 0);
    {
        double4_t acc_e[STP_CB][4], acc_o[STP_CB][4];
#pragma unroll
        for (int cb = 0; cb < STP_CB; ++cb)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                acc_e[cb][b] = double4_t{0.0, 0.0, 0.0, 0.0};
                acc_o[cb][b] = double4_t{0.0, 0.0, 0.0, 0.0};
            }
        // one MFMA step: parity-l rows k0 + kk .. + 3 of the chunk in LDS
        auto kstep = [&](int kk) {
            double be[4], bo[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                be[b] = Ae[(kk + lg) * STP_A_LD + 16 * b + li];
                bo[b] = Ao[(kk + lg) * STP_A_LD + 16 * b + li];
            }
#pragma unroll
            for (int cb = 0; cb < STP_CB; ++cb) {
                const int row = (16 * STP_CB * wave + 16 * cb + li) * STP_BS_LD + 2 * (kk + lg);
                const double a_e = Bs[row];
                const double a_o = Bs[row + 1];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    acc_e[cb][b] = mfma_f64(a_e, be[b], acc_e[cb][b]);
                    acc_o[cb][b] = mfma_f64(a_o, bo[b], acc_o[cb][b]);
                }
            }
        };
        for (int k0 = 0; k0 < cur.ne; k0 += STP_KC) {
            stash(k0);
            __syncthreads();
            if (k0 + STP_KC < cur.ne) fetch(cur, k0 + STP_KC);
            // the last, partial chunk stops after the pair of MFMA steps that holds its last
            // parity-l row (the rows past ne are zeros in LDS, so the steps skipped would
            // add exact zeros)
            const int kend = min(STP_KC, (cur.ne - k0 + 7) & ~7);
            for (int kk = 0; kk < kend; kk += 8) {
                kstep(kk);
                kstep(kk + 4);
            }
            __syncthreads();
        }

        double *ob = p2 + (long)cur.mp * nrhs * nt;
#pragma unroll
        for (int cb = 0; cb < STP_CB; ++cb)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = cur.j0 + 16 * b + li;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int c = cur.c0 + 16 * STP_CB * wave + 16 * cb + lg + 4 * v;
                    if (c < nrhs && j < nt2) {
                        const double e = acc_e[cb][b][v], o = acc_o[cb][b][v];
                        ob[(long)c * nt + j] = e + o;
                        ob[(long)c * nt + (nt - 1 - j)] = e - o;
                    }
                }
            }
    }
}

// ---------------------------------------------------------------- PtS
// Workgroup tile: (waves x 16) columns (c) x 64 parity-l values of each parity (128 l).
// Wave w owns 16 columns: 4 l blocks x 2 parities of 16x16 accumulators.  K runs over
// chunks of 16 southern nodes.  A wave's data columns are its own, so each lane loads its
// part of the next chunk (4 nodes and their mirrors, two 16-byte loads each) while the
// current chunk is multiplied, forms the fold f(j) +/- f(nt+1-j) and leaves it in the
// wave's own LDS strip; only the table chunk is shared by the waves and goes through a
// double-buffered LDS tile, so one barrier per chunk suffices.
constexpr int PTS_THREADS = RAYSHT_PTS_THREADS;
constexpr int PTS_TC = (PTS_THREADS / 64) * 16;
constexpr int PTS_TK = RAYSHT_PTS_TK;  // parity-l rows of a full tile, per parity
constexpr int PTS_NBR = PTS_TK / 16;   // 16-row MFMA blocks of a full tile
constexpr int PTS_KJ = 16;             // nodes per K chunk: 4 MFMA steps of 4
constexpr int PTS_I_LD = PTS_TK + 4;   // table chunk row stride (doubles): lane groups 0 / 1 on disjoint banks
constexpr int PTS_IBUF = PTS_KJ * PTS_I_LD;  // one parity's table chunk in LDS
constexpr int PTS_I_PER_T = PTS_KJ * PTS_TK / PTS_THREADS;  // table values per thread and parity
static_assert(PTS_THREADS % 64 == 0 && PTS_THREADS % PTS_KJ == 0 && (PTS_KJ * PTS_TK) % PTS_THREADS == 0 &&
                  PTS_NBR % 4 == 0,
              "PtS tile geometry");

struct pts_item {
    int mp, k0, c0, nl, ne, no;
    __amdgpu_buffer_rsrc_t rp, rie, rio;
};

__device__ __forceinline__ double2 load_f64x2(__amdgpu_buffer_rsrc_t r, int voff)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
}

// One PtS tile of NBR 16-row blocks of each parity: PTS_NBR (the full tile), or three
// quarters, a half or a quarter of it for the last rows of an m-block, which would leave
// part of a full tile idle.  LDS buffers come from the kernel so that all shapes share them.
//
// K order: MFMA step s of a chunk takes node j0 + 4 lg + s on lane group lg (any order
// works as long as both operands follow it), so that a lane's four nodes are contiguous:
// lane (li, lg) loads column c0 + 16 wave + li at nodes j0 + 4 lg .. + 3 and at their
// mirrors nt - 1 - j as two 16-byte loads each.  Rows are 16-byte aligned (nt is even) and
// a 16-byte piece never straddles a column end, so out-of-range pieces (columns past nrhs)
// read 0 as a whole.
template <int NBR>
__device__ __forceinline__ void pts_tile(const pts_item &cur, int nt, int nt2, int nrhs, double *__restrict__ sb,
                                         double *__restrict__ Ib, double *__restrict__ Fb)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    const int c = cur.c0 + 16 * wave + li;
    const int va = (c * nt + 4 * lg) * 8;           // + 8 j0
    const int vb = (c * nt + nt - 4 - 4 * lg) * 8;  // - 8 j0: mirrors of nodes j0 + 4 lg + 3 .. + 0
    // table loads: parity-l row fcc0 + (PTS_THREADS / KJ) i, node j0 + fjj
    const int fjj = tid % PTS_KJ, fcc0 = tid / PTS_KJ;
    const int i_sstep = (PTS_THREADS / PTS_KJ) * nt2 * 8;
    constexpr int NI = (PTS_I_PER_T * NBR + PTS_NBR - 1) / PTS_NBR;
    double *Fw = Fb + wave * 512;  // this wave's folded data: [parity][s][lane]

    // The loads land in registers untouched; edge masks are applied in settle(), so
    // nothing waits for them before the current chunk is multiplied.
    double2 fa0, fa1, fb0, fb1;  // raw data of the next chunk
    double rie[NI], rio[NI];     // table values of the next chunk
    auto fetch_f = [&](int j0) {
        fa0 = load_f64x2(cur.rp, va + 8 * j0);
        fa1 = load_f64x2(cur.rp, va + 8 * j0 + 16);
        fb0 = load_f64x2(cur.rp, vb - 8 * j0);
        fb1 = load_f64x2(cur.rp, vb - 8 * j0 + 16);
    };
    auto fetch_i = [&](int j0) {
        const int vi = ((cur.k0 + fcc0) * nt2 + j0 + fjj) * 8;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            rie[i] = load_f64(cur.rie, vi + i * i_sstep);
            rio[i] = load_f64(cur.rio, vi + i * i_sstep);
        }
    };
    // fold the fetched data into the wave's strip and the table into LDS buffer `buf`
    auto settle = [&](int j0, int buf) {
        const double a[4] = {fa0.x, fa0.y, fa1.x, fa1.y};
        const double b[4] = {fb1.y, fb1.x, fb0.y, fb0.x};  // mirror of node j0 + 4 lg + s
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const bool jin = j0 + 4 * lg + s < nt2;
            Fw[s * 64 + lane] = jin ? a[s] + b[s] : 0.0;
            Fw[256 + s * 64 + lane] = jin ? a[s] - b[s] : 0.0;
        }
        double *Ie = Ib + buf * (2 * PTS_IBUF), *Io = Ie + PTS_IBUF;
        const bool jin = j0 + fjj < nt2;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int idx = tid + i * PTS_THREADS;
            const int at = (idx % PTS_KJ) * PTS_I_LD + idx / PTS_KJ;
            Ie[at] = jin ? rie[i] : 0.0;
            Io[at] = jin ? rio[i] : 0.0;
        }
    };

    double4_t acc_e[NBR], acc_o[NBR];
#pragma unroll
    for (int b = 0; b < NBR; ++b) {
        acc_e[b] = double4_t{0.0, 0.0, 0.0, 0.0};
        acc_o[b] = double4_t{0.0, 0.0, 0.0, 0.0};
    }
    fetch_f(0);
    fetch_i(0);
    settle(0, 0);
    __syncthreads();
    int buf = 0;
    for (int j0 = 0; j0 < nt2; j0 += PTS_KJ) {
        // buffer buf ^ 1 and the strip were last read before the previous barrier
        const bool more = j0 + PTS_KJ < nt2;
        if (more) fetch_f(j0 + PTS_KJ);
        const double *Ie = Ib + buf * (2 * PTS_IBUF), *Io = Ie + PTS_IBUF;
#pragma unroll 2
        for (int s = 0; s < 4; ++s) {
            double be[NBR], bo[NBR];
#pragma unroll
            for (int b = 0; b < NBR; ++b) {
                be[b] = Ie[(4 * lg + s) * PTS_I_LD + 16 * b + li];
                bo[b] = Io[(4 * lg + s) * PTS_I_LD + 16 * b + li];
            }
            const double aes = Fw[s * 64 + lane], aos = Fw[256 + s * 64 + lane];
#pragma unroll
            for (int b = 0; b < NBR; ++b) {
                acc_e[b] = mfma_f64(aes, be[b], acc_e[b]);
                acc_o[b] = mfma_f64(aos, bo[b], acc_o[b]);
            }
            // the table chunk comes from L2 (every column tile reads it): fetched after the
            // third step, late enough to keep its registers free, early enough to land
            if (s == 2 && more) fetch_i(j0 + PTS_KJ);
        }
        buf ^= 1;
        if (more) settle(j0 + PTS_KJ, buf);
        __syncthreads();
    }

#pragma unroll
    for (int b = 0; b < NBR; ++b) {
        const int k = cur.k0 + 16 * b + li;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int cc = cur.c0 + 16 * wave + lg + 4 * v;
            if (cc < nrhs) {
                if (k < cur.ne) sb[(long)cc * cur.nl + 2 * k] = acc_e[b][v];
                if (k < cur.no) sb[(long)cc * cur.nl + 2 * k + 1] = acc_o[b][v];
            }
        }
    }
}

__global__ __launch_bounds__(PTS_THREADS, RAYSHT_FP64_MIN_WAVES) void pts_fp64_kernel(
    const prob_desc *__restrict__ probs, const batch_io io, const work_item *__restrict__ work)
{
    __shared__ double Ib[2 * 2 * PTS_IBUF];          // [buffer][parity] table chunks
    __shared__ double Fb[PTS_THREADS / 64 * 512];    // per wave: folded data [parity][step][lane]

    const int mp = uniform(work[blockIdx.x].mp);
    if (mp < 0) return;  // padding item of the XCD-balanced schedule
    const int pad = uniform(work[blockIdx.x].pad), pi = pad & 0xffff;
    const prob_desc &P = probs[pi];
    const int lmax = uniform(P.lmax), nt = uniform(P.nt), nt2 = uniform(P.nt2), nrhs = uniform(P.nrhs);
    double *sb = uniform(io.out[pi]) + uniform(uniform(P.s_off)[mp]);
    pts_item cur;
    cur.mp = mp;
    cur.k0 = uniform(work[blockIdx.x].r0);
    cur.c0 = uniform(work[blockIdx.x].c0);
    cur.nl = lmax - uniform(uniform(P.mvals)[mp]) + 1;
    cur.ne = (cur.nl + 1) >> 1;
    cur.no = cur.nl >> 1;
    cur.rp = make_rsrc(uniform(io.in[pi]) + (long)mp * nrhs * nt, (long)nrhs * nt * 8);
    cur.rie = make_rsrc(uniform(P.ip_even) + uniform(uniform(P.te_off)[mp]), (long)cur.ne * nt2 * 8);
    cur.rio = make_rsrc(uniform(P.ip_odd) + uniform(uniform(P.to_off)[mp]), (long)cur.no * nt2 * 8);
    switch (pad >> 16) {
    case 0: pts_tile<PTS_NBR>(cur, nt, nt2, nrhs, sb, Ib, Fb); break;
    case 1: pts_tile<PTS_NBR / 2>(cur, nt, nt2, nrhs, sb, Ib, Fb); break;
    case 3: pts_tile<3 * PTS_NBR / 4>(cur, nt, nt2, nrhs, sb, Ib, Fb); break;
    default: pts_tile<PTS_NBR / 4>(cur, nt, nt2, nrhs, sb, Ib, Fb); break;
    }
}

}  // namespace

int launch_stp_fp64_batch(const prob_desc *d_desc, const batch_io &io, const work_item *work, int n_work,
                          hipStream_t st)
{
    if (n_work == 0) return RAYSHT_OK;
    hipLaunchKernelGGL(stp_fp64_kernel, dim3(n_work), dim3(STP_THREADS), 0, st, d_desc, io, work);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_pts_fp64_batch(const prob_desc *d_desc, const batch_io &io, const work_item *work, int n_work,
                          hipStream_t st)
{
    if (n_work == 0) return RAYSHT_OK;
    hipLaunchKernelGGL(pts_fp64_kernel, dim3(n_work), dim3(PTS_THREADS), 0, st, d_desc, io, work);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_stp_fp64(raysht_plan *p, const double *s, double *p2, hipStream_t st)
{
    batch_io io{};
    io.in[0] = s;
    io.out[0] = p2;
    return launch_stp_fp64_batch(p->d_desc, io, p->d_work_stp, p->n_work_stp, st);
}

int launch_pts_fp64(raysht_plan *p, const double *p2, double *s, hipStream_t st)
{
    batch_io io{};
    io.in[0] = p2;
    io.out[0] = s;
    return launch_pts_fp64_batch(p->d_desc, io, p->d_work_pts, p->n_work_pts, st);
}

// Tile geometry the host work lists are built with.
int stp_tile_rows() { return STP_TJ; }
int stp_tile_cols() { return STP_TC; }
int pts_tile_rows() { return PTS_TK; }
int pts_tile_cols() { return PTS_TC; }

}  // namespace raysht
