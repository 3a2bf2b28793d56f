// Split-precision Legendre contractions on BF16 MFMA (the BF16X3 / BF16X6 forms).
//
// Each FP64 operand row (tables: one output row over K) or column (data: one column
// over K) is written as a fixed-point number relative to a shared power of two 2^e:
//     v = 2^e * sum_s p_s,   p_s = d_s * 2^(-7-8s),   d_s integer, |d_s| <= 128,
// i.e. 8-bit "slices", each exactly representable in BF16.  A slice product p_s q_t is
// exact in FP32 and every product of one level L = s + t is an integer multiple of
// 2^(-14-8L), so FP32 sums of one level are exact while the integer count stays below
// 2^24: with level sums moved into FP64 every 256 terms of K (kChunkK), every level
// with at most 4 slice pairs (levels 0-3) is summed exactly.  The remaining levels share
// one more FP32 sum, which rounds once per MFMA: at level >= 4 (weights <= 2^-32 of
// level 0) that costs ~1e-14; sharing from level 3 on costs ~3e-12 on a decaying
// spectrum.  The FP64 recombination adds the sums and scales by 2^(e_row + e_col).
//
// Forms (slices of the data operand x slices of the table, pairs kept):
//   BF16X6, StP: 7 x 6, pairs with s + t <= 6 (27 MFMAs per 32x32x16 block), five FP32
//     sums (levels 0, 1, 2, 3, >= 4).  Spectral columns decay with l, so the data operand
//     needs the seventh slice: the 6 x 6 triangle leaves ~2e-12 of max|result| on an
//     (l+1)^-2 spectrum, 7 x 6 ~5e-14.
//   BF16X6, PtS: 6 x 6, s + t <= 5 (21 MFMAs), four FP32 sums (levels 0, 1, 2, >= 3);
//     the folded hybrid data does not decay (~2e-14).
//   BF16X3, both: 3 x 3, s + t <= 2 (6 MFMAs), one FP32 sum: ~1e-6.
//
// Tables are split once at plan creation into "fragment-major" planes, so that a wave's
// B operand for one K step of 16 is a single coalesced 1 KiB load per slice.  The data
// operand is split inside the contraction kernel: each workgroup reads its FP64 column
// tile, finds the column exponents, and writes the slices of every K step into LDS in
// the MFMA A-operand order (double-buffered, one barrier per step).
//
// v_mfma_f32_32x32x16_bf16: lane l holds A[i = l%32][k = 8(l/32) .. +7] and
// B[k = 8(l/32) .. +7][j = l%32]; D[i][j] sits in lane l = j + 32*((i/4)%2), element
// v = 4*(i/8) + i%4.  Rows i are data columns c, columns j are table rows (theta nodes
// for StP, parity-l indices for PtS), so each 32-lane half stores 32 consecutive
// outputs of the theta-fastest hybrid layout.
#include <algorithm>
#include <vector>

#include "raysht_internal.h"

namespace raysht {

int split_plan_init(raysht_plan *p);
void split_plan_free(raysht_plan *p);
int launch_stp_split(raysht_plan *p, const double *s, double *p2, hipStream_t st);
int launch_pts_split(raysht_plan *p, const double *p2, double *s, hipStream_t st);

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4;  // one operand fragment, kept as 4 dwords

constexpr int kThreads = 256;  // 4 waves
constexpr int kTileC = 64;     // data columns per workgroup: two 32-column MFMA blocks per wave
constexpr int kTileR = 128;    // output rows per workgroup: one 32-row block per wave
constexpr int kChunkK = 256;   // K terms per FP32 level sum: 3 * 256 * 2^14 < 2^24
constexpr int kChunkSteps = kChunkK / 16;

// Slice counts and kept pairs of one form; pair q enumerates levels in increasing
// order, so the FP32 sum it feeds (group) is nondecreasing in q.
template <int NA_, int NB_, int LEV_, int NG_>
struct form {
    static constexpr int NA = NA_;  // data slices
    static constexpr int NB = NB_;  // table slices
    static constexpr int LEV = LEV_;
    static constexpr int NG = NG_;  // FP32 sums: one per level below NG - 1, the rest share the last
    static constexpr int pair(int q, bool want_s)
    {
        int n = 0;
        for (int L = 0; L <= LEV; ++L)
            for (int s = 0; s < NA; ++s) {
                const int t = L - s;
                if (t < 0 || t >= NB) continue;
                if (n == q) return want_s ? s : t;
                ++n;
            }
        return -1;
    }
    static constexpr int npairs()
    {
        int n = 0;
        for (int s = 0; s < NA; ++s)
            for (int t = 0; t < NB; ++t) n += (s + t <= LEV);
        return n;
    }
    static constexpr int group(int q)
    {
        const int L = pair(q, true) + pair(q, false);
        return L < NG - 1 ? L : NG - 1;
    }
};
using stp_x6 = form<7, 6, 6, 5>;
using pts_x6 = form<6, 6, 5, 4>;
using any_x3 = form<3, 3, 2, 1>;
static_assert(stp_x6::npairs() == 27 && pts_x6::npairs() == 21 && any_x3::npairs() == 6, "pair counts");

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// 2^e with |v| < 2^e for every v of a row or column whose max |v| is amax
__device__ __forceinline__ int scale_exponent(double amax)
{
    if (!(amax > 0.0)) return 0;
    int e = 0;
    (void)__builtin_frexp(amax, &e);
    return e;
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// BF16 bit pattern of slice s of x (the remainder relative to the scale), advancing x
__device__ __forceinline__ unsigned next_slice(double &x, int s)
{
    const double t = x * (s == 0 ? 128.0 : 256.0);
    const double d = __builtin_rint(t);
    x = t - d;
    const float f = (float)d * __builtin_ldexpf(1.0f, -7 - 8 * s);  // <= 8 significant bits: exact
    return __builtin_bit_cast(unsigned, f) >> 16;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const double *base, long bytes)
{
    const long n = bytes < 0x7fffffffL ? bytes : 0x7fffffffL;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, (int)n, 0x00020000);
}

// 16 bytes at byte offset voff (unsigned: a "negative" offset is past the end); offsets
// at or past the end read as zeros
__device__ __forceinline__ double2 load_d2(__amdgpu_buffer_rsrc_t rsrc, unsigned voff)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)voff, 0, 0));
}

// --------------------------------------------------------------- table split (setup)
// Writes the nb slices of 8 consecutive K values v[0..7] (relative to 2^e) as the
// 16-byte elements of fragment lane L in nb consecutive fragments starting at f.
__device__ __forceinline__ void put_table_group(uint4 *__restrict__ tdig, long f, int L, int nb, const double (&v)[8],
                                                int e)
{
    double x[8];
    for (int i = 0; i < 8; ++i) x[i] = __builtin_ldexp(v[i], -e);
    for (int s = 0; s < nb; ++s) {
        unsigned h[8];
        for (int i = 0; i < 8; ++i) h[i] = next_slice(x[i], s);
        tdig[((f + s) << 6) + L] = uint4{h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                                         h[6] | (h[7] << 16)};
    }
}

// StP table of (mp, parity): rows j (theta), K = parity-l index k.  Fragment
// ((jb * nks + ks) * NB + t) of the (mp, parity) region holds slice t of rows
// 32 jb .. +31, k = 16 ks .. +15: lane L, element e -> j = 32 jb + L%32,
// k = 16 ks + 8 (L/32) + e.  One wave per (j, parity, mp).
__global__ __launch_bounds__(64) void split_stp_table_kernel(int lmax, int nt2, int nb, const int *__restrict__ mvals,
                                                             const long *__restrict__ te_off,
                                                             const long *__restrict__ to_off,
                                                             const double *__restrict__ pte,
                                                             const double *__restrict__ pto,
                                                             const long *__restrict__ frag_off,
                                                             uint4 *__restrict__ tdig, int *__restrict__ texp)
{
    const int j = blockIdx.x, par = blockIdx.y, mp = blockIdx.z, lane = threadIdx.x;
    const int m = mvals[mp], nl = lmax - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const int n = par ? no : ne, nks = cdiv(n, 16), njb = cdiv(nt2, 32);
    const double *t = par ? pto + to_off[mp] : pte + te_off[mp];  // [k][j]
    double amax = 0.0;
    for (int k = lane; k < n; k += 64) amax = fmax(amax, fabs(t[(long)k * nt2 + j]));
    const int e = scale_exponent(wave_max(amax));
    if (lane == 0) texp[((long)mp * 2 + par) * nt2 + j] = e;
    const long region = frag_off[mp] + (par ? (long)njb * cdiv(ne, 16) * nb : 0);
    const int jb = j >> 5, jl = j & 31;
    for (int g = lane; g < 2 * nks; g += 64) {  // k = 8 g .. 8 g + 7
        double v[8];
        for (int i = 0; i < 8; ++i) v[i] = 8 * g + i < n ? t[(long)(8 * g + i) * nt2 + j] : 0.0;
        put_table_group(tdig, region + ((long)jb * nks + (g >> 1)) * nb, jl + 32 * (g & 1), nb, v, e);
    }
}

// PtS table of (mp, parity): rows k (parity-l, the output), K = j (theta).  Fragment
// ((kb * njs + js) * NB + t): lane L, element e -> k = 32 kb + L%32, j = 16 js + 8 (L/32) + e.
// One wave per (k, parity, mp); rows past the parity's length are skipped (their
// fragments were cleared).
__global__ __launch_bounds__(64) void split_pts_table_kernel(int lmax, int nt2, int nb, const int *__restrict__ mvals,
                                                             const long *__restrict__ te_off,
                                                             const long *__restrict__ to_off,
                                                             const double *__restrict__ ipe,
                                                             const double *__restrict__ ipo,
                                                             const long *__restrict__ frag_off,
                                                             const long *__restrict__ exp_off,
                                                             uint4 *__restrict__ tdig, int *__restrict__ texp)
{
    const int k = blockIdx.x, par = blockIdx.y, mp = blockIdx.z, lane = threadIdx.x;
    const int m = mvals[mp], nl = lmax - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const int n = par ? no : ne;
    if (k >= n) return;
    const int njs = cdiv(nt2, 16);
    const double *t = (par ? ipo + to_off[mp] : ipe + te_off[mp]) + (long)k * nt2;  // [k][j]
    double amax = 0.0;
    for (int j = lane; j < nt2; j += 64) amax = fmax(amax, fabs(t[j]));
    const int e = scale_exponent(wave_max(amax));
    if (lane == 0) texp[exp_off[mp] + (par ? ne : 0) + k] = e;
    const long region = frag_off[mp] + (par ? (long)cdiv(ne, 32) * njs * nb : 0);
    const int kb = k >> 5, kl = k & 31;
    for (int g = lane; g < 2 * njs; g += 64) {  // j = 8 g .. 8 g + 7
        double v[8];
        for (int i = 0; i < 8; ++i) v[i] = 8 * g + i < nt2 ? t[8 * g + i] : 0.0;
        put_table_group(tdig, region + ((long)kb * njs + (g >> 1)) * nb, kl + 32 * (g & 1), nb, v, e);
    }
}

// --------------------------------------------------------------- contraction core
// One K step of 16 for one 32-column block of a wave: acc[g] += sum over the form's
// pairs of A_s (data) x B_t (table block of this wave).
template <class F, int Q = 0>
__device__ __forceinline__ void mfma_pairs(f32x16 (&acc)[F::NG], const u32x4 (&a)[F::NA], const u32x4 (&b)[F::NB])
{
    if constexpr (Q < F::npairs()) {
        constexpr int s = F::pair(Q, true), t = F::pair(Q, false), g = F::group(Q);
        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a[s]), __builtin_bit_cast(bf16x8, b[t]),
                                                         acc[g], 0, 0, 0);
        mfma_pairs<F, Q + 1>(acc, a, b);
    }
}

// adds the FP32 level sums into this thread's FP64 sums in LDS (dsum[cb*16+v][tid]) and
// clears them
template <class F>
__device__ __forceinline__ void flush(f32x16 (&acc)[2][F::NG], double (*dsum)[kThreads])
{
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            double sum = 0.0;
#pragma unroll
            for (int g = F::NG - 1; g >= 0; --g) sum += (double)acc[cb][g][v];
            dsum[cb * 16 + v][threadIdx.x] += sum;
        }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int g = 0; g < F::NG; ++g) acc[cb][g] = f32x16{};
}

// this thread's FP64 sums in LDS, cleared
__device__ __forceinline__ void clear_sums(double (*dsum)[kThreads])
{
#pragma unroll
    for (int i = 0; i < 32; ++i) dsum[i][threadIdx.x] = 0.0;
}

// Writes the NA slices of 4 consecutive K values of data column ca (0..63) into the A
// fragments of one LDS buffer: column block ca/32, fragment lane ca%32 + 32 (q/2),
// elements 4 (q%2) .. +3, where q (0..3) is the quarter of the K step.
template <int NA>
__device__ __forceinline__ void put_slices(u32x4 *__restrict__ buf, int ca, int q, const double (&v)[4], int e)
{
    double x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = __builtin_ldexp(v[i], -e);
    const int cb = ca >> 5, L = (ca & 31) + 32 * (q >> 1);
    uint2 *b2 = reinterpret_cast<uint2 *>(buf);
#pragma unroll
    for (int s = 0; s < NA; ++s) {
        unsigned h[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) h[i] = next_slice(x[i], s);
        b2[(((cb * NA + s) << 6) + L) * 2 + (q & 1)] = uint2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
    }
}

template <int NA>
__device__ __forceinline__ void get_a(const u32x4 *__restrict__ buf, int cb, int lane, u32x4 (&a)[NA])
{
#pragma unroll
    for (int s = 0; s < NA; ++s) a[s] = buf[((cb * NA + s) << 6) + lane];
}

template <int NB>
__device__ __forceinline__ void get_b(const u32x4 *__restrict__ frag, int lane, bool active, u32x4 (&b)[NB])
{
    if (active) {
#pragma unroll
        for (int t = 0; t < NB; ++t) b[t] = frag[(t << 6) + lane];
    } else {
#pragma unroll
        for (int t = 0; t < NB; ++t) b[t] = u32x4{0u, 0u, 0u, 0u};
    }
}

// The K loop of one contraction: nks steps of 16, with the data of step ks delivered
// by load_a(ks, v) (4 values of this thread's column quarter) and the table fragments
// of step ks at tb + ks * NB fragments.  Adds the result into this thread's FP64 sums
// in LDS.  nks is uniform over the workgroup; every step ends with one barrier.
template <class F, class LoadA>
__device__ __forceinline__ void k_loop(const LoadA &load_a, int nks, const u32x4 *__restrict__ tb, bool active,
                                       u32x4 (*abuf)[2 * F::NA * 64], int ca, int q, int ea, int lane,
                                       double (*dsum)[kThreads])
{
    if (nks <= 0) return;
    {
        double v[4];
        load_a(0, v);
        put_slices<F::NA>(abuf[0], ca, q, v, ea);
    }
    u32x4 b[F::NB];
    get_b<F::NB>(tb, lane, active, b);
    __syncthreads();
    f32x16 acc[2][F::NG];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int g = 0; g < F::NG; ++g) acc[cb][g] = f32x16{};
    for (int kc = 0; kc < nks; kc += kChunkSteps) {
        const int kend = min(nks, kc + kChunkSteps);
        for (int ks = kc; ks < kend; ++ks) {
            double vn[4];
            u32x4 bn[F::NB];
            const bool more = ks + 1 < nks;
            if (more) {
                load_a(ks + 1, vn);
                get_b<F::NB>(tb + (((long)(ks + 1) * F::NB) << 6), lane, active, bn);
            }
            // unconditional: an inactive wave multiplies zero table fragments (a branch
            // here would make the accumulators leave the MFMA registers every step)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                u32x4 a[F::NA];
                get_a<F::NA>(abuf[ks & 1], cb, lane, a);
                mfma_pairs<F>(acc[cb], a, b);
            }
            if (more) {
                put_slices<F::NA>(abuf[(ks + 1) & 1], ca, q, vn, ea);
#pragma unroll
                for (int t = 0; t < F::NB; ++t) b[t] = bn[t];
            }
            __syncthreads();
        }
        flush<F>(acc, dsum);
    }
}

// --------------------------------------------------------------- StP
// Work item (mp, c0): 64 columns, every theta row, in tiles of 128 rows (wave w owns
// theta block jb = j0/32 + w) with both parities each.  The spectral tile is read as
// 16-byte pairs (l, l+1) and each parity pass keeps its half; the column exponents come
// from one pass over the tile.
template <class F>
__global__ __launch_bounds__(kThreads, 1) void stp_split_kernel(
    const work_item *__restrict__ work, int lmax, int nt, int nt2, int nrhs, const int *__restrict__ mvals,
    const long *__restrict__ s_off, const double *__restrict__ s, const long *__restrict__ frag_off,
    const u32x4 *__restrict__ tdig, const int *__restrict__ texp, double *__restrict__ p2)
{
    __shared__ u32x4 abuf[2][2 * F::NA * 64];
    __shared__ double sums[2][2 * 16][kThreads];  // FP64 sums of each thread, per parity
    __shared__ int cexp[kTileC];
    const work_item w = work[blockIdx.x];
    if (w.mp < 0) return;
    const int mp = w.mp, m = mvals[mp], nl = lmax - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double *sb = s + s_off[mp];
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(sb, (long)nl * nrhs * 8);

    // column exponents: the four threads of a column read it in the main loop's
    // pattern (l = 32 i + 8 q .. +7), four steps of loads in flight at a time
    const int ca = threadIdx.x >> 2, q = threadIdx.x & 3;
    const unsigned a_off = (unsigned)min(w.c0 + ca, nrhs) * nl * 8;  // column nrhs and beyond read as 0
    {
        double mx = 0.0;
        const int n32 = cdiv(nl, 32);
        for (int i0 = 0; i0 < n32; i0 += 4) {
            double2 d[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    d[u][e] = load_d2(rs, a_off + (unsigned)(32 * (i0 + u) + 8 * q + 2 * e) * 8);
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int l = 32 * (i0 + u) + 8 * q + 2 * e;
                    if (l < nl) mx = fmax(mx, fabs(d[u][e].x));
                    if (l + 1 < nl) mx = fmax(mx, fabs(d[u][e].y));
                }
        }
        mx = fmax(mx, __shfl_xor(mx, 1, 64));
        mx = fmax(mx, __shfl_xor(mx, 2, 64));
        if (q == 0) cexp[ca] = scale_exponent(mx);
    }
    __syncthreads();
    const int ea = cexp[ca];
    const int njb = cdiv(nt2, 32), nks_e = cdiv(ne, 16);
    double *ob = p2 + (long)mp * nrhs * nt;
    for (int r0 = 0; r0 < nt2; r0 += kTileR) {
        const int jb = r0 / 32 + wave;
        const bool active = jb < njb;
        const int j = jb * 32 + (lane & 31);
#pragma unroll
        for (int par = 0; par < 2; ++par) {
            clear_sums(sums[par]);
            const int K = par ? no : ne, nks = cdiv(K, 16);
            const u32x4 *tb =
                tdig + ((frag_off[mp] + (par ? (long)njb * nks_e * F::NB : 0) + (long)jb * nks * F::NB) << 6);
            // raw data of one K step: l = 32 ks + 8 q .. +7 as four (even, odd) pairs
            auto load_a = [&](int ks, double (&v)[4]) {
                const int l0 = 32 * ks + 8 * q;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double2 d = load_d2(rs, a_off + (unsigned)(l0 + 2 * i) * 8);
                    v[i] = l0 + 2 * i + par < nl ? (par ? d.y : d.x) : 0.0;
                }
            };
            k_loop<F>(load_a, nks, tb, active, abuf, ca, q, ea, lane, sums[par]);
        }
        if (active && j < nt2) {
            const int ej = texp[((long)mp * 2) * nt2 + j], oj = texp[((long)mp * 2 + 1) * nt2 + j];
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int cc = cb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), c = w.c0 + cc;
                    if (c < nrhs) {
                        const double e = __builtin_ldexp(sums[0][cb * 16 + v][threadIdx.x], cexp[cc] + ej);
                        const double o = __builtin_ldexp(sums[1][cb * 16 + v][threadIdx.x], cexp[cc] + oj);
                        ob[(long)c * nt + j] = e + o;
                        ob[(long)c * nt + (nt - 1 - j)] = e - o;
                    }
                }
        }
    }
}

// --------------------------------------------------------------- PtS
// Work item (mp, c0): 64 columns, every parity-l row of both parities, in row tiles of
// 128 (wave w owns row block kb = k0/32 + w).  The data operand is the fold
// f(j) + f(nt-1-j) (even) or f(j) - f(nt-1-j) (odd) of the hybrid columns; both folds'
// column exponents come from one pass over the tile.
template <class F>
__global__ __launch_bounds__(kThreads, 1) void pts_split_kernel(
    const work_item *__restrict__ work, int lmax, int nt, int nt2, int nrhs, const int *__restrict__ mvals,
    const long *__restrict__ s_off, const double *__restrict__ p2, const long *__restrict__ frag_off,
    const long *__restrict__ exp_off, const u32x4 *__restrict__ tdig, const int *__restrict__ texp,
    double *__restrict__ s)
{
    __shared__ u32x4 abuf[2][2 * F::NA * 64];
    __shared__ double sums[2 * 16][kThreads];  // FP64 sums of each thread
    __shared__ int cexp[2][kTileC];
    const work_item w = work[blockIdx.x];
    if (w.mp < 0) return;
    const int mp = w.mp, m = mvals[mp], nl = lmax - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double *pb = p2 + (long)mp * nrhs * nt;
    const __amdgpu_buffer_rsrc_t rp = make_rsrc(pb, (long)nrhs * nt * 8);
    const int ca = threadIdx.x >> 2, q = threadIdx.x & 3;
    const unsigned a_off = (unsigned)min(w.c0 + ca, nrhs) * nt * 8;
    const int njs = cdiv(nt2, 16);

    // raw data of one K step: f(j) and f(nt-1-j) for j = 16 js + 4 q .. +3 (nt-4-j0 < 0
    // only past the last valid j, where the offset wraps past the end or into a column
    // whose values are masked)
    auto load_raw = [&](int js, double (&a)[4], double (&b)[4]) {
        const int j0 = 16 * js + 4 * q;
        const double2 f0 = load_d2(rp, a_off + j0 * 8u), f1 = load_d2(rp, a_off + (j0 + 2) * 8u);
        const double2 g0 = load_d2(rp, a_off + (unsigned)(nt - 2 - j0) * 8u);  // f(nt-2-j0), f(nt-1-j0)
        const double2 g1 = load_d2(rp, a_off + (unsigned)(nt - 4 - j0) * 8u);  // f(nt-4-j0), f(nt-3-j0)
        a[0] = f0.x, a[1] = f0.y, a[2] = f1.x, a[3] = f1.y;
        b[0] = g0.y, b[1] = g0.x, b[2] = g1.y, b[3] = g1.x;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (j0 + i >= nt2) a[i] = b[i] = 0.0;
    };
    // exponents of both folds, four steps of loads in flight at a time
    {
        double me = 0.0, mo = 0.0;
        for (int j0 = 0; j0 < njs; j0 += 4) {
            double a[4][4], b[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u) load_raw(j0 + u, a[u], b[u]);  // steps past njs read zeros
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    me = fmax(me, fabs(a[u][i] + b[u][i]));
                    mo = fmax(mo, fabs(a[u][i] - b[u][i]));
                }
        }
        me = fmax(me, __shfl_xor(me, 1, 64));
        me = fmax(me, __shfl_xor(me, 2, 64));
        mo = fmax(mo, __shfl_xor(mo, 1, 64));
        mo = fmax(mo, __shfl_xor(mo, 2, 64));
        if (q == 0) {
            cexp[0][ca] = scale_exponent(me);
            cexp[1][ca] = scale_exponent(mo);
        }
    }
    __syncthreads();
    double *sb = s + s_off[mp];

#pragma unroll
    for (int par = 0; par < 2; ++par) {
        const int K_rows = par ? no : ne, ea = cexp[par][ca];
        auto load_a = [&](int js, double (&v)[4]) {
            double a[4], b[4];
            load_raw(js, a, b);
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] = par ? a[i] - b[i] : a[i] + b[i];
        };
        const long region = frag_off[mp] + (par ? (long)cdiv(ne, 32) * njs * F::NB : 0);
        for (int k0 = 0; k0 < K_rows; k0 += kTileR) {
            const int kb = k0 / 32 + wave;
            const bool active = kb * 32 < K_rows;
            const u32x4 *tb = tdig + ((region + (long)kb * njs * F::NB) << 6);
            clear_sums(sums);
            k_loop<F>(load_a, njs, tb, active, abuf, ca, q, ea, lane, sums);

            const int k = kb * 32 + (lane & 31);
            if (active && k < K_rows) {
                const int ek = texp[exp_off[mp] + (par ? ne : 0) + k];
#pragma unroll
                for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int cc = cb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), c = w.c0 + cc;
                        if (c < nrhs)
                            sb[(long)c * nl + 2 * k + par] =
                                __builtin_ldexp(sums[cb * 16 + v][threadIdx.x], cexp[par][cc] + ek);
                    }
            }
        }
    }
}

// --------------------------------------------------------------- host side
// XCD-balanced work lists of (m-block, column tile) items: whole m-blocks dealt to the
// eight XCDs by longest-processing-time, interleaved so that item i runs on XCD i % 8.
std::vector<work_item> split_work(const raysht_plan *p, bool stp)
{
    constexpr int kXcds = 8;
    std::vector<int> order(p->nm);
    for (int i = 0; i < p->nm; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return p->mvals[a] < p->mvals[b]; });
    std::vector<std::vector<work_item>> q(kXcds);
    std::vector<double> load(kXcds, 0.0);
    const int nct = cdiv(p->nrhs, kTileC), nrt = cdiv(p->nt2, kTileR), njs = cdiv(p->nt2, 16);
    for (int mp : order) {
        const int m = p->mvals[mp], nl = p->lmax - m + 1, ne = (nl + 1) / 2, no = nl / 2;
        // K steps per item (plus the exponent pass)
        const double steps = stp ? (double)nrt * (cdiv(ne, 16) + cdiv(no, 16)) + cdiv(nl, 32)
                                 : (double)(cdiv(ne, kTileR) + cdiv(no, kTileR)) * njs + njs;
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        for (int c = 0; c < nct; ++c) q[x].push_back(work_item{mp, 0, c * kTileC, 0});
        load[x] += steps * nct;
    }
    size_t len = 0;
    for (auto &v : q) len = std::max(len, v.size());
    std::vector<work_item> w(len * kXcds, work_item{-1, 0, 0, 0});
    for (size_t i = 0; i < len; ++i)
        for (int x = 0; x < kXcds; ++x)
            if (i < q[x].size()) w[i * kXcds + x] = q[x][i];
    return w;
}

template <typename T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return RAYSHT_ENOMEM;
    }
    return RAYSHT_OK;
}

template <typename T>
int dupload(T **p, const std::vector<T> &v)
{
    int rc = dalloc(p, v.size());
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return RAYSHT_OK;
}

bool x6(const raysht_plan *p) { return p->precision == RAYSHT_BF16X6; }

}  // namespace

int split_plan_init(raysht_plan *p)
{
    raysht_split &sp = p->split;
    const int nb_stp = x6(p) ? stp_x6::NB : any_x3::NB, nb_pts = x6(p) ? pts_x6::NB : any_x3::NB;
    const int nm = p->nm, nt2 = p->nt2, njb = cdiv(nt2, 32), njs = cdiv(nt2, 16);
    std::vector<long> fs(nm + 1, 0), fp(nm + 1, 0), ep(nm + 1, 0);
    for (int mp = 0; mp < nm; ++mp) {
        const int nl = p->lmax - p->mvals[mp] + 1, ne = (nl + 1) / 2, no = nl / 2;
        fs[mp + 1] = fs[mp] + (long)njb * (cdiv(ne, 16) + cdiv(no, 16)) * nb_stp;
        fp[mp + 1] = fp[mp] + (long)(cdiv(ne, 32) + cdiv(no, 32)) * njs * nb_pts;
        ep[mp + 1] = ep[mp] + ne + no;
    }
    const std::vector<work_item> ws = split_work(p, true), wp = split_work(p, false);
    sp.n_work_stp = (int)ws.size();
    sp.n_work_pts = (int)wp.size();
    int rc;
    if ((rc = dupload(&sp.d_frag_stp_off, fs)) || (rc = dupload(&sp.d_frag_pts_off, fp)) ||
        (rc = dupload(&sp.d_exp_pts_off, ep)) || (rc = dupload(&sp.d_work_stp, ws)) ||
        (rc = dupload(&sp.d_work_pts, wp)) || (rc = dalloc(&sp.d_tdig_stp, (size_t)fs[nm] * 64)) ||
        (rc = dalloc(&sp.d_tdig_pts, (size_t)fp[nm] * 64)) || (rc = dalloc(&sp.d_texp_stp, (size_t)nm * 2 * nt2)) ||
        (rc = dalloc(&sp.d_texp_pts, (size_t)ep[nm])))
        return rc;
    // rows past a parity's length in the last 32-row block of the PtS tables are read
    // as zeros: clear before the split kernel fills the rows that exist
    RAYSHT_HIP_TRY(hipMemset(sp.d_tdig_pts, 0, (size_t)fp[nm] * 64 * sizeof(uint4)));
    hipLaunchKernelGGL(split_stp_table_kernel, dim3(nt2, 2, nm), dim3(64), 0, nullptr, p->lmax, nt2, nb_stp,
                       p->d_mvals, p->d_te_off, p->d_to_off, p->d_pt_even, p->d_pt_odd, sp.d_frag_stp_off,
                       sp.d_tdig_stp, sp.d_texp_stp);
    RAYSHT_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(split_pts_table_kernel, dim3((p->lmax + 2) / 2, 2, nm), dim3(64), 0, nullptr, p->lmax, nt2,
                       nb_pts, p->d_mvals, p->d_te_off, p->d_to_off, p->d_ip_even, p->d_ip_odd, sp.d_frag_pts_off,
                       sp.d_exp_pts_off, sp.d_tdig_pts, sp.d_texp_pts);
    RAYSHT_HIP_TRY(hipGetLastError());
    RAYSHT_HIP_TRY(hipDeviceSynchronize());
    return RAYSHT_OK;
}

void split_plan_free(raysht_plan *p)
{
    raysht_split &sp = p->split;
    for (void *ptr : {(void *)sp.d_frag_stp_off, (void *)sp.d_frag_pts_off, (void *)sp.d_exp_pts_off,
                      (void *)sp.d_work_stp, (void *)sp.d_work_pts, (void *)sp.d_tdig_stp, (void *)sp.d_tdig_pts,
                      (void *)sp.d_texp_stp, (void *)sp.d_texp_pts})
        if (ptr) (void)hipFree(ptr);
    sp = raysht_split();
}

int launch_stp_split(raysht_plan *p, const double *s, double *p2, hipStream_t st)
{
    raysht_split &sp = p->split;
    if (sp.n_work_stp == 0) return RAYSHT_OK;
    if (x6(p))
        hipLaunchKernelGGL(stp_split_kernel<stp_x6>, dim3(sp.n_work_stp), dim3(kThreads), 0, st, sp.d_work_stp,
                           p->lmax, p->nt, p->nt2, p->nrhs, p->d_mvals, p->d_s_off, s, sp.d_frag_stp_off,
                           reinterpret_cast<const u32x4 *>(sp.d_tdig_stp), sp.d_texp_stp, p2);
    else
        hipLaunchKernelGGL(stp_split_kernel<any_x3>, dim3(sp.n_work_stp), dim3(kThreads), 0, st, sp.d_work_stp,
                           p->lmax, p->nt, p->nt2, p->nrhs, p->d_mvals, p->d_s_off, s, sp.d_frag_stp_off,
                           reinterpret_cast<const u32x4 *>(sp.d_tdig_stp), sp.d_texp_stp, p2);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_pts_split(raysht_plan *p, const double *p2, double *s, hipStream_t st)
{
    raysht_split &sp = p->split;
    if (sp.n_work_pts == 0) return RAYSHT_OK;
    if (x6(p))
        hipLaunchKernelGGL(pts_split_kernel<pts_x6>, dim3(sp.n_work_pts), dim3(kThreads), 0, st, sp.d_work_pts,
                           p->lmax, p->nt, p->nt2, p->nrhs, p->d_mvals, p->d_s_off, p2, sp.d_frag_pts_off,
                           sp.d_exp_pts_off, reinterpret_cast<const u32x4 *>(sp.d_tdig_pts), sp.d_texp_pts, s);
    else
        hipLaunchKernelGGL(pts_split_kernel<any_x3>, dim3(sp.n_work_pts), dim3(kThreads), 0, st, sp.d_work_pts,
                           p->lmax, p->nt, p->nt2, p->nrhs, p->d_mvals, p->d_s_off, p2, sp.d_frag_pts_off,
                           sp.d_exp_pts_off, reinterpret_cast<const u32x4 *>(sp.d_tdig_pts), sp.d_texp_pts, s);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

}  // namespace raysht
