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
//   BF16X3, both: 4 x 4, s + t <= 3 (10 MFMAs), two FP32 sums (level 0, levels >= 1): ~5e-7
//     of max|result| (3 x 3 with s + t <= 2 left 2e-5 - 4e-5 in StP on an (l+1)^-2 spectrum
//     and 1.6e-6 in PtS on unscaled random data, INT8 digits).
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
#include <type_traits>
#include <vector>

#include "raysht_internal.h"
#include "uniform.h"

namespace raysht {

int split_plan_init(raysht_plan *p);
void split_plan_free(raysht_plan *p);
int launch_stp_split(raysht_plan *p, const double *s, double *p2, hipStream_t st);
int launch_pts_split(raysht_plan *p, const double *p2, double *s, hipStream_t st);

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(16))) int i32x16;
typedef __attribute__((ext_vector_type(4))) int i32x4;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4;  // one operand fragment, kept as 4 dwords

// Workgroup layout: RAYSHT_SPLIT_WAVES waves, each one 32 x 32 output block, over
// RAYSHT_SPLIT_CB 32-column blocks; StP tiles are 4 theta blocks tall, PtS tiles 2
// parity-l blocks of each parity.
#ifndef RAYSHT_SPLIT_WAVES
#define RAYSHT_SPLIT_WAVES 4
#endif
#ifndef RAYSHT_SPLIT_CB
#define RAYSHT_SPLIT_CB 1
#endif
constexpr int kWaves = RAYSHT_SPLIT_WAVES;
constexpr int kCB = RAYSHT_SPLIT_CB;
constexpr int kThreads = 64 * kWaves;
constexpr int kTileC = 32 * kCB;                 // data columns per workgroup
constexpr int kJB = kWaves / kCB;                // StP theta blocks per tile
constexpr int kKB = kWaves / (2 * kCB);          // PtS parity-l blocks per tile and parity
constexpr int kTileR = 32 * kJB;                 // StP theta rows per tile
constexpr int kTileRP = 32 * kKB;                // PtS parity-l rows per tile and parity
constexpr int kMinWaves = kWaves >= 8 ? 1 : 2;   // workgroups per CU the registers must allow
constexpr int kABuf = 2 * kCB * 7 * 64;          // LDS A fragments of one step
static_assert(kThreads / 8 == kTileC && kJB * kCB == kWaves && 2 * kKB * kCB == kWaves, "split layout");
// K steps per FP32 (BF16 forms) level sum between moves into FP64: 16 steps of 16 = 256
// terms, 4 * 256 * 2^14 <= 2^24.  The INT8 forms' int32 level sums stay exact over a
// whole contraction of up to kMaxI8K terms (7 pairs * 2^14 * 16384 < 2^31); plans
// with longer contractions are refused for those forms.
constexpr int kChunkSteps = 16;
constexpr int kMaxI8K = 16384;

// Slice counts and kept pairs of one form; pair q enumerates levels in increasing
// order, so the FP32 sum it feeds (group) is nondecreasing in q.
template <int NA_, int NB_, int LEV_, int NG_, bool I8_ = false>
struct form {
    static constexpr int NA = NA_;  // data slices
    static constexpr int NB = NB_;  // table slices
    static constexpr int LEV = LEV_;
    static constexpr int NG = NG_;  // sums: one per level below NG - 1, the rest share the last
    // INT8 forms: the same 8-bit digits as signed bytes on v_mfma_i32_32x32x32_i8 (twice
    // the BF16 rate, K = 32 per instruction), one exact int32 sum per level
    static constexpr bool I8 = I8_;
    static constexpr int KS = I8 ? 32 : 16;  // K per step (one MFMA)
    static constexpr int VPT = KS / 8;       // data values per producer thread, block and step
    static constexpr int EPL = KS / 2;       // K values per fragment lane (16 bytes)
    using acc_t = typename std::conditional<I8, i32x16, f32x16>::type;
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
using stp_x3 = form<4, 4, 3, 2>;
using pts_x3 = form<4, 4, 3, 2>;
using stp_i6 = form<7, 6, 6, 7, true>;
using pts_i6 = form<6, 6, 5, 6, true>;
using stp_i3 = form<4, 4, 3, 4, true>;
using pts_i3 = form<4, 4, 3, 4, true>;
static_assert(stp_x6::npairs() == 27 && pts_x6::npairs() == 21 && stp_x3::npairs() == 10 && pts_x3::npairs() == 10,
              "pair counts");
static_assert(stp_i6::npairs() == 27 && pts_i6::npairs() == 21 && stp_i3::npairs() == 10 && pts_i3::npairs() == 10,
              "pair counts");

bool is_i8(int precision) { return precision == RAYSHT_INT8X6 || precision == RAYSHT_INT8X3; }
bool is_x6(int precision) { return precision == RAYSHT_BF16X6 || precision == RAYSHT_INT8X6; }

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// 2^e with |v| < 2^e for every v of a row or column whose max |v| is amax; the INT8
// forms take one more bit of headroom (|v| < 2^(e-1)), so that the leading digit
// stays within [-64, 64] and every digit fits a signed byte
__device__ __forceinline__ int scale_exponent(double amax, bool i8 = false)
{
    if (!(amax > 0.0)) return 0;
    int e = 0;
    (void)__builtin_frexp(amax, &e);
    return e + (i8 ? 1 : 0);
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
// 16-byte elements (8 BF16) of fragment lane L in nb consecutive fragments starting at f.
__device__ __forceinline__ void put_table_group(uint4 *__restrict__ tdig, long f, int L, int nb, const double *v,
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

// The INT8 forms' table digits: nb signed-byte digits of 16 consecutive K values
// (relative to 2^e, |v| < 2^(e-1)), from X = round(v 2^(8 nb - 1 - e)) in balanced base 256
// (the same expansion as the data operand's, rounded at digit nb), as the 16 bytes of
// fragment lane L in nb consecutive fragments starting at f.
__device__ __forceinline__ void put_table_group_i8(uint4 *__restrict__ tdig, long f, int L, int nb, const double *v,
                                                   int e)
{
    long long X[16];
    for (int i = 0; i < 16; ++i) X[i] = (long long)__builtin_rint(__builtin_ldexp(v[i], 8 * nb - 1 - e));
    for (int s = 0; s < nb; ++s) {
        unsigned w[4] = {0u, 0u, 0u, 0u};
        for (int i = 0; i < 16; ++i) {
            // digit s: (X + B) >> 8 (nb - 1 - s), low byte, minus the offset 128 (s >= 1)
            const unsigned long long off = 0x0000808080808080ull >> (8 * (7 - nb));
            const long long Y = (long long)((unsigned long long)X[i] + off) >> (8 * (nb - 1 - s));
            const unsigned byte = ((unsigned)Y & 0xffu) ^ (s ? 0x80u : 0u);
            w[i >> 2] |= byte << (8 * (i & 3));
        }
        tdig[((f + s) << 6) + L] = uint4{w[0], w[1], w[2], w[3]};
    }
}

// StP table of (mp, parity): rows j (theta), K = parity-l index k.  Fragment
// ((jb * nks + ks) * NB + t) of the (mp, parity) region holds slice t of rows
// 32 jb .. +31, k = KS ks .. +KS-1: lane L, element e -> j = 32 jb + L%32,
// k = KS ks + (KS/2) (L/32) + e (KS = 16 BF16 / 32 INT8 elements of 16 bytes per lane).
// One wave per (j, parity, mp).
template <bool I8>
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
    constexpr int KS = I8 ? 32 : 16, EPL = KS / 2;
    const int n = par ? no : ne, nks = cdiv(n, KS), njb = cdiv(nt2, 32);
    const double *t = par ? pto + to_off[mp] : pte + te_off[mp];  // [k][j]
    double amax = 0.0;
    for (int k = lane; k < n; k += 64) amax = fmax(amax, fabs(t[(long)k * nt2 + j]));
    const int e = scale_exponent(wave_max(amax), I8);
    if (lane == 0) texp[((long)mp * 2 + par) * nt2 + j] = e;
    const long region = frag_off[mp] + (par ? (long)njb * cdiv(ne, KS) * nb : 0);
    const int jb = j >> 5, jl = j & 31;
    for (int g = lane; g < 2 * nks; g += 64) {  // k = EPL g .. EPL g + EPL - 1
        double v[EPL];
        for (int i = 0; i < EPL; ++i) v[i] = EPL * g + i < n ? t[(long)(EPL * g + i) * nt2 + j] : 0.0;
        if constexpr (I8)
            put_table_group_i8(tdig, region + ((long)jb * nks + (g >> 1)) * nb, jl + 32 * (g & 1), nb, v, e);
        else
            put_table_group(tdig, region + ((long)jb * nks + (g >> 1)) * nb, jl + 32 * (g & 1), nb, v, e);
    }
}

// PtS table of (mp, parity): rows k (parity-l, the output), K = j (theta).  Fragment
// ((kb * njs + js) * NB + t): lane L, element e -> k = 32 kb + L%32, j = KS js + (KS/2) (L/32) + e.
// One wave per (k, parity, mp); rows past the parity's length are skipped (their
// fragments were cleared).
template <bool I8>
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
    constexpr int KS = I8 ? 32 : 16, EPL = KS / 2;
    const int njs = cdiv(nt2, KS);
    const double *t = (par ? ipo + to_off[mp] : ipe + te_off[mp]) + (long)k * nt2;  // [k][j]
    double amax = 0.0;
    for (int j = lane; j < nt2; j += 64) amax = fmax(amax, fabs(t[j]));
    const int e = scale_exponent(wave_max(amax), I8);
    if (lane == 0) texp[exp_off[mp] + (par ? ne : 0) + k] = e;
    const long region = frag_off[mp] + (par ? (long)cdiv(ne, 32) * njs * nb : 0);
    const int kb = k >> 5, kl = k & 31;
    for (int g = lane; g < 2 * njs; g += 64) {  // j = EPL g .. EPL g + EPL - 1
        double v[EPL];
        for (int i = 0; i < EPL; ++i) v[i] = EPL * g + i < nt2 ? t[EPL * g + i] : 0.0;
        if constexpr (I8)
            put_table_group_i8(tdig, region + ((long)kb * njs + (g >> 1)) * nb, kl + 32 * (g & 1), nb, v, e);
        else
            put_table_group(tdig, region + ((long)kb * njs + (g >> 1)) * nb, kl + 32 * (g & 1), nb, v, e);
    }
}

// floor(g / d) for 0 <= g < 2^16 and 1 <= d < 2^16 as one high multiply by r = ceil(2^32 / d):
// r d = 2^32 + eps with 0 <= eps < d, so g r / 2^32 = g / d + g eps / (d 2^32) and the added
// term stays below 1 / d (g eps < g d < 2^32), which keeps the floor exact (d = 1 apart)
struct small_div {
    unsigned r;
    int d;
    __device__ explicit small_div(int d_) : r(d_ > 1 ? (unsigned)((0x100000000ull + (unsigned)d_ - 1) / (unsigned)d_) : 0u), d(d_) {}
    __device__ __forceinline__ int quo(int g) const { return d > 1 ? (int)__umulhi((unsigned)g, r) : g; }
};

// x, hidden from the optimiser: what is computed from it inside a loop stays inside the
// loop (an epilogue that runs once per contraction inside the K loop keeps its addresses
// to itself instead of holding them in registers through every step)
__device__ __forceinline__ int opaque(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// --------------------------------------------------------------- contraction core
// One MFMA of a slice pair: acc += A_s (data) x B_t (table) for one K step
template <class F>
__device__ __forceinline__ void mfma_pair(typename F::acc_t &acc, const u32x4 &a, const u32x4 &b)
{
    if constexpr (F::I8)
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(i32x4, a), __builtin_bit_cast(i32x4, b), acc, 0,
                                                    0, 0);
    else
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                      0, 0, 0);
}

// The MFMAs of one K step for a wave's 32 x 32 block (acc[g] += sum over the form's pairs
// of A_s x B_t), with the NA slicing steps of the next K step spread evenly between them
// (slice s after pair s * npairs / NA), in source order
template <class F>
constexpr int slice_at(int q)
{
    for (int s = 0; s < F::NA; ++s)
        if (s * F::npairs() / F::NA == q) return s;
    return -1;
}

template <class F, class SL, int Q = 0>
__device__ __forceinline__ void mfma_pairs_sliced(typename F::acc_t (&acc)[F::NG], const u32x4 (&a)[F::NA],
                                                  const u32x4 (&b)[F::NB], SL &sl)
{
    if constexpr (Q < F::npairs()) {
        constexpr int s = F::pair(Q, true), t = F::pair(Q, false), g = F::group(Q), sc = slice_at<F>(Q);
        mfma_pair<F>(acc[g], a[s], b[t]);
        if constexpr (sc >= 0) sl.step(sc);
        mfma_pairs_sliced<F, SL, Q + 1>(acc, a, b, sl);
    }
}

// adds the level sums into the FP64 sums and clears them.  BF16 digits carry their weights
// 2^(-7-8s), so the FP32 level sums are values; the INT8 forms' exact int32 level-L sums S_L
// are worth S_L 2^(-14-8L).  INT8: levels L >= 4 (weight <= 2^-32 of level 0) are first
// combined in FP32, which rounds at 2^-24 of their size, i.e. at 2^-56 of the operand scales
// (the BF16 forms' shared FP32 sum of levels >= 4 rounds at the same place); levels 0-3 go to
// FP64 exactly and every scaling is by an exact power of two.
template <class F>
__device__ __forceinline__ void flush(typename F::acc_t (&acc)[F::NG], double (&dacc)[16])
{
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        if constexpr (F::I8) {
            constexpr int NX = F::NG < 4 ? F::NG : 4;  // levels summed in FP64
            double lo = (double)acc[NX - 1][v];       // in units of 2^(-14-8 (NX-1))
#pragma unroll
            for (int g = NX - 2; g >= 0; --g) lo = __builtin_fma(lo, 0x1p-8, (double)acc[g][v]);
            if constexpr (F::NG > 4) {
                float hi = (float)acc[F::NG - 1][v];  // in units of 2^(-14-8 (NG-1))
#pragma unroll
                for (int g = F::NG - 2; g >= 4; --g) hi = __builtin_fmaf(hi, 0x1p-8f, (float)acc[g][v]);
                lo = __builtin_fma((double)hi, 0x1p-32, lo);
            }
            dacc[v] += lo * 0x1p-14;
        } else {
            double sum = 0.0;
#pragma unroll
            for (int g = F::NG - 1; g >= 0; --g) sum += (double)acc[g][v];
            dacc[v] += sum;
        }
    }
#pragma unroll
    for (int g = 0; g < F::NG; ++g) acc[g] = typename F::acc_t{};
}

// The data operand's NA digits, from one integer per value.  For |v| < 2^e,
// X = round(v 2^(8 NA - 1 - e)) (|X| <= 2^(8 NA - 1)) is formed by the 1.5 2^52 rounding
// constant for NA <= 6 and, for the seven-digit data operand, from two FP64 steps (floor of
// v 2^(8 NA - 33 - e), then the rounded 32-bit remainder); it is held with 128 added to each
// of its NA - 1 low bytes: byte NA-1-s (s = 1 .. NA-1) is then d_s + 128 and the arithmetic
// top bits are d_0, so that v 2^-e = sum_s d_s 2^(-7-8s) + O(2^(-8 NA)) with every
// |d_s| <= 128 (exact in BF16).  Each digit then costs one byte conversion and one FMA
// instead of the three FP64 adds and a conversion of a slice taken off the remainder.
template <int NA>
__device__ __forceinline__ unsigned long long offset_digits(double v, int e)
{
    static_assert(NA >= 2 && NA <= 7, "digits of one 64-bit integer");
    constexpr unsigned long long kOffset = 0x0000808080808080ull >> (8 * (7 - NA));
    if constexpr (NA <= 6) {
        // |X| <= 2^47 < 2^51: adding 1.5 2^52 rounds t to an integer in the low mantissa
        // bits (ties to even), and the bit patterns of t + 1.5 2^52 and 1.5 2^52 differ by
        // exactly that integer (one binade): two FP64 operations and one 64-bit add
        const double t = __builtin_ldexp(v, 8 * NA - 1 - e);
        const long long bits = __builtin_bit_cast(long long, t + 0x1.8p52);
        return (unsigned long long)(bits - __builtin_bit_cast(long long, 0x1.8p52)) + kOffset;
    } else {
        const double t = __builtin_ldexp(v, 8 * NA - 33 - e);  // |t| < 2^(8 NA - 32)
        const double h = __builtin_floor(t);
        const double r = __builtin_fmin(__builtin_fma(t - h, 4294967296.0, 0.5), 4294967295.0);
        const unsigned long long X = ((unsigned long long)(unsigned)(int)h << 32) | (unsigned)r;
        return X + kOffset;
    }
}

// digit s of an offset_digits<NA> value, times 2^(-7-8s), as FP32 (exact)
template <int NA>
__device__ __forceinline__ float digit(unsigned long long X, int s)
{
    const float k = __builtin_ldexpf(1.0f, -7 - 8 * s);
    if (s == 0) return (float)(int)((long long)X >> (8 * (NA - 1))) * k;
    return __builtin_fmaf((float)(unsigned)((X >> (8 * (NA - 1 - s))) & 0xffu), k, -128.0f * k);
}

// two FP32 values (exact in BF16) as one dword of two BF16
__device__ __forceinline__ unsigned pack_bf16(float lo, float hi)
{
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2_t{lo, hi}), bf16x2_t));
}

// LDS slot of MFMA fragment lane L: lanes 32..63 are swapped in groups of 8, so that
// the two halves of a 16-K step (same rows, L and L + 32) land in different banks when
// one thread writes a pair of K values; the 16-lane groups of a 16-byte read still
// cover distinct banks.
__device__ __forceinline__ int frag_slot(int L) { return L ^ ((L >> 5) << 3); }

// slice s of four offset_digits<NA> values as four signed bytes: byte NA-1-s of each
// (the top byte is the two's-complement d_0; the others hold d_s + 128, flipped by the
// XOR with 0x80), gathered with byte permutes
template <int NA>
__device__ __forceinline__ unsigned pack_i8(const unsigned long long (&X)[4], int s)
{
    const int p = NA - 1 - s, b = p & 3;
    const unsigned w0 = p < 4 ? (unsigned)X[0] : (unsigned)(X[0] >> 32);
    const unsigned w1 = p < 4 ? (unsigned)X[1] : (unsigned)(X[1] >> 32);
    const unsigned w2 = p < 4 ? (unsigned)X[2] : (unsigned)(X[2] >> 32);
    const unsigned w3 = p < 4 ? (unsigned)X[3] : (unsigned)(X[3] >> 32);
    const unsigned sel = 0x0c0c0000u | ((4u + b) << 8) | (unsigned)b;  // (src1 byte b, src0 byte b, 0, 0)
    const unsigned t01 = __builtin_amdgcn_perm(w1, w0, sel), t23 = __builtin_amdgcn_perm(w3, w2, sel);
    const unsigned r = __builtin_amdgcn_perm(t23, t01, 0x05040100u);
    return s ? r ^ 0x80808080u : r;
}

// The slicing of VPT consecutive K values of one data column into the A fragments of an
// LDS buffer laid out [block][slice][slot][16 bytes] (fragment lane L = row % 32 +
// 32 * (k0 / EPL), bytes of k0 % EPL, k0 a multiple of VPT within the step), one slice per
// step() call, so that the calls can be spread between the MFMAs of the previous step.
// BF16 forms: two values, one dword of two BF16 per slice; INT8 forms: four values, one
// dword of four signed bytes.
template <class F>
struct digit_slicer {
    unsigned long long X[F::VPT];
    unsigned *p;
    __device__ __forceinline__ void init(u32x4 *__restrict__ buf, int blk, int row, int k0,
                                         const double (&v)[F::VPT], int e)
    {
#pragma unroll
        for (int i = 0; i < F::VPT; ++i) X[i] = offset_digits<F::NA>(v[i], e);
        const int lane = (row & 31) + 32 * (k0 / F::EPL);
        p = reinterpret_cast<unsigned *>(buf) + ((blk * F::NA) * 64 + frag_slot(lane)) * 4 +
            (k0 % F::EPL) / (F::I8 ? 4 : 2);
    }
    __device__ __forceinline__ void step(int s)
    {
        if constexpr (F::I8)
            p[s * 256] = pack_i8<F::NA>(X, s);
        else
            p[s * 256] = pack_bf16(digit<F::NA>(X[0], s), digit<F::NA>(X[1], s));
    }
};

template <class F, int N>
struct slicers {
    digit_slicer<F> sl[N];
    __device__ __forceinline__ void step(int s)
    {
#pragma unroll
        for (int i = 0; i < N; ++i) sl[i].step(s);
    }
};

template <int NA>
__device__ __forceinline__ void get_a(const u32x4 *__restrict__ buf, int blk, int lane, u32x4 (&a)[NA])
{
#pragma unroll
    for (int s = 0; s < NA; ++s) a[s] = buf[((blk * NA + s) << 6) + frag_slot(lane)];
}

// table fragments f .. f + NB - 1 of an m-block's region (buffer resource over the
// region: fragments past its end read as zeros)
template <int NB>
__device__ __forceinline__ void get_b(__amdgpu_buffer_rsrc_t rt, int f, int lane, u32x4 (&b)[NB])
{
#pragma unroll
    for (int t = 0; t < NB; ++t)
        b[t] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rt, (((f + t) << 6) + lane) << 4, 0, 0));
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_frag_rsrc(const u32x4 *base, long frags)
{
    const long n = frags * 1024 < 0x7fffffffL ? frags * 1024 : 0x7fffffffL;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<u32x4 *>(base), (short)0, (int)n, 0x00020000);
}

// Raw FP64 data of one K step for one thread (loaded two steps ahead of its slicing):
// N 16-byte pairs, with the bit mask of the values that exist (bit 2 i + h for d[i].x
// (h = 0) or d[i].y (h = 1); applied when the data are sliced, so that the loads are not
// waited for until then).
template <int N>
struct raw_step {
    double2 d[N];
    unsigned valid;
    __device__ __forceinline__ double val(int i, int h) const
    {
        const double x = h ? d[i].y : d[i].x;
        return (valid >> (2 * i + h)) & 1 ? x : 0.0;
    }
};

// The K loop of a stream of contractions that one workgroup runs back to back: global
// steps g = 0 .. G - 1 of F::KS terms each.  fetch(g) issues the loads of step g's raw data
// (steps past the end read masked zeros), slice(raw, g, buffer) returns the slicers that
// write its A fragments into an LDS buffer, one slice per step() call (spread between the
// MFMAs of the step before), and bfrag(g) is the first table fragment of step g (resource
// rt).  close(g) says what ends with step g: 0 nothing, 1 a run of FP32 level sums (BF16
// forms: they move into FP64), 2 a contraction (the sums move into FP64 and done(g, dacc)
// receives its FP64 sums).  Raw data run two steps ahead of the MFMAs and table fragments
// one step, across the ends of contractions, so only the first step of the stream waits for
// memory.  The caller issues that step's loads (r0, r1: raw data of steps 0 and 1, b0: table
// fragments of step 0) before its exponent pass, so that their latency overlaps the pass,
// and has the exponents in LDS (after a barrier) before the call.  Wave w multiplies A
// block ablk of the buffer by its table fragments.  The step body has no branches but the
// uniform close test (the last steps slice and load harmlessly past the end), so one wave
// keeps its matrix pipe and its vector ALU busy at once; every step ends with one barrier.
template <class F, class Raw, class Fetch, class Slice, class BFrag, class Close, class Done>
__device__ __forceinline__ void k_stream(int G, const Fetch &fetch, const Slice &slice, const BFrag &bfrag,
                                         const Close &close, const Done &done, __amdgpu_buffer_rsrc_t rt,
                                         const Raw &r0, const Raw &r1, const u32x4 (&b0)[F::NB],
                                         u32x4 (*abuf)[kABuf], int ablk, int lane)
{
    if (G <= 0) return;
    {
        auto sl = slice(r0, 0, abuf[0]);
#pragma unroll
        for (int sc = 0; sc < F::NA; ++sc) sl.step(sc);
    }
    // two register sets for the raw data and the table fragments, swapped by unrolling
    // the step loop by two (no copies between steps)
    Raw rA = r1, rB;
    u32x4 bA[F::NB], bB[F::NB];
#pragma unroll
    for (int t = 0; t < F::NB; ++t) bA[t] = b0[t];
    __syncthreads();
    typename F::acc_t acc[F::NG];
#pragma unroll
    for (int g = 0; g < F::NG; ++g) acc[g] = typename F::acc_t{};
    // FP64 sums of the BF16 forms' runs of level sums (the INT8 forms' int32 sums are exact
    // over a whole contraction, so they go straight to done())
    double dacc[F::I8 ? 1 : 16];
#pragma unroll
    for (int v = 0; v < (F::I8 ? 1 : 16); ++v) dacc[v] = 0.0;
    // step g: raw data of step g + 1 in rc, table fragments of step g in bc; loads step
    // g + 2's raw data into rn and step g + 1's fragments into bn
    auto step = [&](int g, const Raw &rc, Raw &rn, const u32x4 (&bc)[F::NB], u32x4 (&bn)[F::NB]) {
        rn = fetch(g + 2);
        get_b<F::NB>(rt, bfrag(g + 1), lane, bn);
        u32x4 a[F::NA];
        get_a<F::NA>(abuf[g & 1], ablk, lane, a);
        auto sl = slice(rc, g + 1, abuf[(g + 1) & 1]);
        mfma_pairs_sliced<F>(acc, a, bc, sl);
        __syncthreads();
        const int c = close(g);
        if constexpr (F::I8) {
            if (c == 2) {
                double sums[16];
#pragma unroll
                for (int v = 0; v < 16; ++v) sums[v] = 0.0;
                flush<F>(acc, sums);
                done(g, sums);
            }
        } else if (c != 0) {
            flush<F>(acc, dacc);
            if (c == 2) {
                done(g, dacc);
#pragma unroll
                for (int v = 0; v < 16; ++v) dacc[v] = 0.0;
            }
        }
    };
    int g = 0;
    for (; g + 1 < G; g += 2) {
        step(g, rA, rB, bA, bB);
        step(g + 1, rB, rA, bB, bA);
    }
    if (g < G) step(g, rA, rB, bA, bB);
}

// --------------------------------------------------------------- StP
// Work item (mp, c0): kTileC columns of one m-block, every theta row, both parities: the
// row tiles (kTileR theta rows each) run one after the other as one stream of contractions
// (row tile, parity), so the column exponents are found once per column tile and the K
// loop never drains between contractions.  Wave w multiplies column block w % kCB by theta
// block kJB rt + w / kCB of row tile rt.  The spectral tile is read as 16-byte pairs (l,
// l+1); each parity's steps slice their half, VPT values per thread and step (the pairs
// of a step come from L2 the second time).  The scaled even-parity sums of a row tile wait
// in LDS for its odd contraction, which writes E + O and E - O.
template <class F>
__global__ __launch_bounds__(kThreads, kMinWaves) void stp_split_kernel(
    const prob_desc *__restrict__ probs, const batch_io io, const work_item *__restrict__ work)
{
    __shared__ u32x4 abuf[2][kABuf];
    __shared__ double ebuf[16][kThreads];
    __shared__ int cexp[kTileC];
    const int mp = uniform(work[blockIdx.x].mp);
    if (mp < 0) return;
    const int pi = uniform(work[blockIdx.x].pad) & 0xffff, c0 = uniform(work[blockIdx.x].c0);
    const prob_desc &P = probs[pi];
    const int nt = uniform(P.nt), nt2 = uniform(P.nt2), nrhs = uniform(P.nrhs);
    const long *frag_off = uniform(P.frag_stp_off);
    const int *texp = uniform(P.texp_stp);
    const int m = uniform(uniform(P.mvals)[mp]), nl = uniform(P.lmax) - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double *sb = uniform(io.in[pi]) + uniform(uniform(P.s_off)[mp]);
    const long f_lo = uniform(frag_off[mp]), f_hi = uniform(frag_off[mp + 1]);
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(sb, (long)nl * nrhs * 8);
    const __amdgpu_buffer_rsrc_t rt =
        make_frag_rsrc(reinterpret_cast<const u32x4 *>(uniform(P.tdig_stp)) + (f_lo << 6), f_hi - f_lo);
    const int cb = wave % kCB, njb = cdiv(nt2, 32);
    // steps of the even and odd contractions (an odd contraction of an m-block without odd
    // l, nl = 1, multiplies zeros: its table region is empty and its data masked)
    const int nks_e = cdiv(ne, F::KS), nks_o = max(1, cdiv(no, F::KS)), nseg = nks_e + nks_o;
    const int G = cdiv(njb, kJB) * nseg;
    const small_div by_seg(nseg);
    struct pos {
        int rt, par, ks;
    };
    auto at = [&](int g) {
        pos p;
        p.rt = by_seg.quo(g);
        const int r = g - p.rt * nseg;
        p.par = r >= nks_e;
        p.ks = p.par ? r - nks_e : r;
        return p;
    };

    // producer role of the thread: column ca, l = 32 i + 4 q .. +3 of step i
    const int ca = tid >> 3, q = tid & 7;
    const unsigned a_off = (unsigned)min(c0 + ca, nrhs) * nl * 8;  // column nrhs and beyond read as 0
    using Raw = raw_step<F::VPT>;
    // step g: parity-k = KS ks + VPT q + i (i < VPT), i.e. l = 2 KS ks + 2 VPT q + 2 i + par,
    // read as the pairs (l, l + 1) of both parities (l >= nl is the next column or past
    // the block: masked to zero)
    auto fetch = [&](int g) {
        const int l0 = 2 * F::KS * at(g).ks + 2 * F::VPT * q;
        Raw r;
        r.valid = 0;
#pragma unroll
        for (int i = 0; i < F::VPT; ++i) {
            r.d[i] = load_d2(rs, a_off + (unsigned)(l0 + 2 * i) * 8);
            r.valid |= (unsigned)(l0 + 2 * i < nl) << (2 * i) | (unsigned)(l0 + 2 * i + 1 < nl) << (2 * i + 1);
        }
        return r;
    };
    auto bfrag = [&](int g) {
        const pos p = at(g);
        const int jb = p.rt * kJB + wave / kCB;
        return (p.par ? njb * nks_e * F::NB : 0) + (jb * (p.par ? nks_o : nks_e) + p.ks) * F::NB;
    };
    const Raw r0 = fetch(0), r1 = fetch(1);
    u32x4 b0[F::NB];
    get_b<F::NB>(rt, bfrag(0), lane, b0);
    {
        // the column exponents: sixteen 16-byte loads in flight per thread
        double mx = 0.0;
        const int n32 = cdiv(nl, 32);
        for (int i0 = 0; i0 < n32; i0 += 8) {
            double2 d[8][2];
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int e = 0; e < 2; ++e) d[u][e] = load_d2(rs, a_off + (unsigned)(32 * (i0 + u) + 4 * q + 2 * e) * 8);
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int l = 32 * (i0 + u) + 4 * q + 2 * e;
                    if (l < nl) mx = fmax(mx, fabs(d[u][e].x));
                    if (l + 1 < nl) mx = fmax(mx, fabs(d[u][e].y));
                }
        }
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
        if (q == 0) cexp[ca] = scale_exponent(mx, F::I8);
    }
    __syncthreads();
    const int ea = cexp[ca];
    double *ob = uniform(io.out[pi]) + (long)mp * nrhs * nt;
    auto slice = [&](const Raw &r, int g, u32x4 *buf) {
        const int par = at(g).par;
        double v[F::VPT];
#pragma unroll
        for (int i = 0; i < F::VPT; ++i) v[i] = r.val(i, par);
        slicers<F, 1> sl;
        sl.sl[0].init(buf, ca >> 5, ca, F::VPT * q, v, ea);
        return sl;
    };
    auto close = [&](int g) {
        const pos p = at(g);
        if (p.ks == (p.par ? nks_o : nks_e) - 1) return 2;
        return (!F::I8 && (p.ks + 1) % kChunkSteps == 0) ? 1 : 0;
    };
    auto done = [&](int g, const double (&dacc)[16]) {
        const pos p = at(g);
        const int ln = opaque(lane);
        const int jb = p.rt * kJB + wave / kCB, j = jb * 32 + (ln & 31);
        const bool store = jb < njb && j < nt2;
        const int ej = store ? texp[((long)mp * 2 + p.par) * nt2 + j] : 0;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int cc = cb * 32 + 8 * (v >> 2) + 4 * (ln >> 5) + (v & 3), c = c0 + cc;
            const double r = __builtin_ldexp(dacc[v], cexp[cc] + ej);
            if (p.par == 0) {
                ebuf[v][tid] = r;
            } else if (store && c < nrhs) {
                const double e = ebuf[v][tid];
                ob[(long)c * nt + j] = e + r;
                ob[(long)c * nt + (nt - 1 - j)] = e - r;
            }
        }
    };
    k_stream<F, Raw>(G, fetch, slice, bfrag, close, done, rt, r0, r1, b0, abuf, cb, lane);
}

// --------------------------------------------------------------- PtS
// Work item (mp, c0): kTileC columns of one m-block, every parity-l row: the row tiles
// (kTileRP parity-l rows of each parity) run one after the other as one stream, so the
// fold's column exponents are found once per column tile.  Wave w multiplies column block
// w % kCB of parity w / (kCB kKB) by row block kKB rt + (w / kCB) % kKB of row tile rt.  The
// data operand is the fold f(j) + f(nt-1-j) (even) or f(j) - f(nt-1-j) (odd) of the hybrid
// columns; each thread slices both folds of VPT nodes per step, and both folds' column
// exponents come from one pass over the tile.
// (INT8X3: registers for three waves per SIMD; INT8X6 spills at three)
template <class F>
__global__ __launch_bounds__(kThreads, F::I8 && F::NA < 6 ? 3 : kMinWaves) void pts_split_kernel(
    const prob_desc *__restrict__ probs, const batch_io io, const work_item *__restrict__ work)
{
    __shared__ u32x4 abuf[2][kABuf];
    __shared__ int cexp[2][kTileC];
    const int mp = uniform(work[blockIdx.x].mp);
    if (mp < 0) return;
    const int pi = uniform(work[blockIdx.x].pad) & 0xffff, c0 = uniform(work[blockIdx.x].c0);
    const prob_desc &P = probs[pi];
    const int nt = uniform(P.nt), nt2 = uniform(P.nt2), nrhs = uniform(P.nrhs);
    const long *frag_off = uniform(P.frag_pts_off);
    const int *texp = uniform(P.texp_pts);
    const int m = uniform(uniform(P.mvals)[mp]), nl = uniform(P.lmax) - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double *pb = uniform(io.in[pi]) + (long)mp * nrhs * nt;
    const __amdgpu_buffer_rsrc_t rp = make_rsrc(pb, (long)nrhs * nt * 8);
    const int ca = tid >> 3, q = tid & 7;
    const unsigned a_off = (unsigned)min(c0 + ca, nrhs) * nt * 8;
    const int njs = cdiv(nt2, F::KS), G = cdiv(ne, kTileRP) * njs;
    const small_div by_njs(njs);
    constexpr int H = F::VPT / 2;
    using Raw = raw_step<F::VPT>;
    const int par = wave / (kCB * kKB), cb = wave % kCB, K_rows = par ? no : ne;
    const int region = par ? cdiv(ne, 32) * njs * F::NB : 0;
    const long f_lo = uniform(frag_off[mp]), f_hi = uniform(frag_off[mp + 1]);
    const __amdgpu_buffer_rsrc_t rt =
        make_frag_rsrc(reinterpret_cast<const u32x4 *>(uniform(P.tdig_pts)) + (f_lo << 6), f_hi - f_lo);

    // raw data of step g (node step js = g % njs): nodes j = j0 + i, j0 = KS js + VPT q, i <
    // VPT, as the pairs d[u] = (f(j0 + 2u), f(j0 + 2u + 1)) and their mirrors d[H + u] =
    // (f(nt-1-j0-2u), f(nt-2-j0-2u)); zero for j >= nt2 (the mirrored offset is negative
    // only past the last valid j, where it wraps past the end or into a column whose values
    // are masked)
    auto fetch_js = [&](int js) {
        const int j0 = F::KS * js + F::VPT * q;
        Raw r;
        r.valid = 0;
#pragma unroll
        for (int u = 0; u < H; ++u) {
            r.d[u] = load_d2(rp, a_off + (unsigned)(j0 + 2 * u) * 8u);
            const double2 g = load_d2(rp, a_off + (unsigned)(nt - 2 - j0 - 2 * u) * 8u);  // f(nt-2-j), f(nt-1-j)
            r.d[H + u] = make_double2(g.y, g.x);
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (j0 + 2 * u + h < nt2) r.valid |= 1u << (2 * u + h) | 1u << (2 * (H + u) + h);
        }
        return r;
    };
    auto fetch = [&](int g) { return fetch_js(g - by_njs.quo(g) * njs); };
    auto bfrag = [&](int g) {
        const int rt_ = by_njs.quo(g), js = g - rt_ * njs;
        return region + ((kKB * rt_ + (wave / kCB) % kKB) * njs + js) * F::NB;
    };
    const Raw r0 = fetch(0), r1 = fetch(1);
    u32x4 b0[F::NB];
    get_b<F::NB>(rt, bfrag(0), lane, b0);
    {
        // both folds' column exponents: sixteen 16-byte loads in flight per thread
        constexpr int EB = 8 / H;
        double me = 0.0, mo = 0.0;
        for (int j0 = 0; j0 < njs; j0 += EB) {
            Raw r[EB];
#pragma unroll
            for (int u = 0; u < EB; ++u) r[u] = fetch_js(j0 + u);  // steps past njs are masked
#pragma unroll
            for (int u = 0; u < EB; ++u)
#pragma unroll
                for (int i = 0; i < F::VPT; ++i) {
                    const double a = r[u].val(i >> 1, i & 1), b = r[u].val(H + (i >> 1), i & 1);
                    me = fmax(me, fabs(a + b));
                    mo = fmax(mo, fabs(a - b));
                }
        }
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            me = fmax(me, __shfl_xor(me, off, 64));
            mo = fmax(mo, __shfl_xor(mo, off, 64));
        }
        if (q == 0) {
            cexp[0][ca] = scale_exponent(me, F::I8);
            cexp[1][ca] = scale_exponent(mo, F::I8);
        }
    }
    __syncthreads();
    const int ee = cexp[0][ca], eo = cexp[1][ca];
    double *sb = uniform(io.out[pi]) + uniform(uniform(P.s_off)[mp]);
    const long e_base = uniform(uniform(P.exp_pts_off)[mp]);
    // A blocks of a buffer: [parity][column block]
    auto slice = [&](const Raw &r, int, u32x4 *buf) {
        double ve[F::VPT], vo[F::VPT];
#pragma unroll
        for (int i = 0; i < F::VPT; ++i) {
            const double a = r.val(i >> 1, i & 1), b = r.val(H + (i >> 1), i & 1);
            ve[i] = a + b;
            vo[i] = a - b;
        }
        slicers<F, 2> sl;
        sl.sl[0].init(buf, ca >> 5, ca, F::VPT * q, ve, ee);
        sl.sl[1].init(buf, kCB + (ca >> 5), ca, F::VPT * q, vo, eo);
        return sl;
    };
    auto close = [&](int g) {
        const int js = g - by_njs.quo(g) * njs;
        if (js == njs - 1) return 2;
        return (!F::I8 && (js + 1) % kChunkSteps == 0) ? 1 : 0;
    };
    auto done = [&](int g, const double (&dacc)[16]) {
        const int ln = opaque(lane);
        const int kb = kKB * by_njs.quo(g) + (wave / kCB) % kKB, k = kb * 32 + (ln & 31);
        if (k < K_rows) {
            const int ek = texp[e_base + (par ? ne : 0) + k];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int cc = cb * 32 + 8 * (v >> 2) + 4 * (ln >> 5) + (v & 3), c = c0 + cc;
                if (c < nrhs) sb[(long)c * nl + 2 * k + par] = __builtin_ldexp(dacc[v], cexp[par][cc] + ek);
            }
        }
    };
    k_stream<F, Raw>(G, fetch, slice, bfrag, close, done, rt, r0, r1, b0, abuf, kCB * par + cb, lane);
}

// --------------------------------------------------------------- wave-specialised forms
// A 512-thread workgroup of two roles.  Waves 0-3 (consumers) only multiply: each streams
// table fragments from memory one step ahead, reads the data's A fragments from LDS and
// issues its MFMAs, and at the end of each contraction moves its level sums into FP64 and
// stores them.  Waves 4-7 (producers) only slice: they load the raw data two steps ahead,
// find the column exponents and write the A fragments of step g + 1 into a three-slot LDS
// ring while the consumers multiply step g.  One wave of each role shares a SIMD, so the
// digit arithmetic runs on the vector ALU while the other wave keeps the matrix pipe busy
// (in the one-role kernels both waves of a SIMD interleave the two, and the vector work,
// about ten instructions per MFMA, left the matrix pipe idle most of the time).  Both roles
// pass the same barriers: two before the stream, one per step.  Used for the X3 forms,
// whose registers leave room for two such workgroups per CU (two multiplying waves per
// SIMD); the X6 forms' level sums allow one, and one multiplying wave per SIMD measured
// slower than the one-role kernels' two (INT8X6 14.7 ms against 13.8 ms per round trip at
// l_max = 511; INT8X3 10.7 ms against 11.0 ms).
constexpr int kWsThreads = 512;
constexpr int kWsRing = 3;
constexpr int kWsABuf = 2 * 7 * 64;  // A fragments of one step: up to two blocks of seven digits

// The MFMAs of one K step for a wave's 32 x 32 block, every kept pair in level order
template <class F, int Q = 0>
__device__ __forceinline__ void mfma_pairs(typename F::acc_t (&acc)[F::NG], const u32x4 (&a)[F::NA],
                                           const u32x4 (&b)[F::NB])
{
    if constexpr (Q < F::npairs()) {
        constexpr int s = F::pair(Q, true), t = F::pair(Q, false), g = F::group(Q);
        mfma_pair<F>(acc[g], a[s], b[t]);
        mfma_pairs<F, Q + 1>(acc, a, b);
    }
}

// Consumer side of the stream: G steps; step g multiplies A block ablk of ring slot g % 3
// by the table fragments bfrag(g) (resource rt, loaded one step ahead), then passes the
// step's barrier; close(g) / done(g, sums) as in k_stream.
template <class F, class BFrag, class Close, class Done>
__device__ __forceinline__ void ws_consume(int G, const BFrag &bfrag, const Close &close, const Done &done,
                                           __amdgpu_buffer_rsrc_t rt, const u32x4 (&b0)[F::NB],
                                           u32x4 (*ring)[kWsABuf], int ablk, int lane)
{
    u32x4 bA[F::NB], bB[F::NB];
#pragma unroll
    for (int t = 0; t < F::NB; ++t) bA[t] = b0[t];
    __syncthreads();  // the exponents
    __syncthreads();  // step 0's A fragments
    typename F::acc_t acc[F::NG];
#pragma unroll
    for (int g = 0; g < F::NG; ++g) acc[g] = typename F::acc_t{};
    double dacc[F::I8 ? 1 : 16];
#pragma unroll
    for (int v = 0; v < (F::I8 ? 1 : 16); ++v) dacc[v] = 0.0;
    int slot = 0;
    auto step = [&](int g, const u32x4 (&bc)[F::NB], u32x4 (&bn)[F::NB]) {
        get_b<F::NB>(rt, bfrag(g + 1), lane, bn);
        u32x4 a[F::NA];
        get_a<F::NA>(ring[slot], ablk, lane, a);
        mfma_pairs<F>(acc, a, bc);
        slot = slot == kWsRing - 1 ? 0 : slot + 1;
        __syncthreads();
        const int c = close(g);
        if constexpr (F::I8) {
            if (c == 2) {
                double sums[16];
#pragma unroll
                for (int v = 0; v < 16; ++v) sums[v] = 0.0;
                flush<F>(acc, sums);
                done(g, sums);
            }
        } else if (c != 0) {
            flush<F>(acc, dacc);
            if (c == 2) {
                done(g, dacc);
#pragma unroll
                for (int v = 0; v < 16; ++v) dacc[v] = 0.0;
            }
        }
    };
    int g = 0;
    for (; g + 1 < G; g += 2) {
        step(g, bA, bB);
        step(g + 1, bB, bA);
    }
    if (g < G) step(g, bA, bB);
}

// Producer side: slices step 0 into slot 0 before the stream, then in step g loads step
// g + 2's raw data and slices step g + 1 into slot (g + 1) % 3 (the consumers read slot g % 3;
// slot (g + 1) % 3 was last read in step g - 2, before the barrier that ended step g - 1).
template <class F, class Raw, class Fetch, class Slice>
__device__ __forceinline__ void ws_produce(int G, const Fetch &fetch, const Slice &slice, const Raw &r0,
                                           const Raw &r1, u32x4 (*ring)[kWsABuf])
{
    {
        auto sl = slice(r0, 0, ring[0]);
#pragma unroll
        for (int sc = 0; sc < F::NA; ++sc) sl.step(sc);
    }
    __syncthreads();  // step 0's A fragments
    Raw rA = r1, rB;
    int slot = 1;
    auto step = [&](int g, const Raw &rc, Raw &rn) {
        rn = fetch(g + 2);
        auto sl = slice(rc, g + 1, ring[slot]);
#pragma unroll
        for (int sc = 0; sc < F::NA; ++sc) sl.step(sc);
        slot = slot == kWsRing - 1 ? 0 : slot + 1;
        __syncthreads();
    };
    int g = 0;
    for (; g + 1 < G; g += 2) {
        step(g, rA, rB);
        step(g + 1, rB, rA);
    }
    if (g < G) step(g, rA, rB);
}

// StP, wave-specialised: work item (mp, c0) as stp_split_kernel; consumer wave w owns theta
// block kJB rt + w of row tile rt, the producers are 256 threads over the 32 columns.
template <class F>
__global__ __launch_bounds__(kWsThreads, 1) void stp_split_ws_kernel(
    const prob_desc *__restrict__ probs, const batch_io io, const work_item *__restrict__ work)
{
    static_assert(kCB == 1 && kJB == 4, "four 32-row blocks of 32 columns per row tile");
    __shared__ u32x4 ring[kWsRing][kWsABuf];
    __shared__ double ebuf[16][256];
    __shared__ int cexp[kTileC];
    const int mp = uniform(work[blockIdx.x].mp);
    if (mp < 0) return;
    const int pi = uniform(work[blockIdx.x].pad) & 0xffff, c0 = uniform(work[blockIdx.x].c0);
    const prob_desc &P = probs[pi];
    const int nt = uniform(P.nt), nt2 = uniform(P.nt2), nrhs = uniform(P.nrhs);
    const int m = uniform(uniform(P.mvals)[mp]), nl = uniform(P.lmax) - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool producer = wave >= 4;
    const int njb = cdiv(nt2, 32), nks_e = cdiv(ne, F::KS), nks_o = max(1, cdiv(no, F::KS)), nseg = nks_e + nks_o;
    const int G = cdiv(njb, kJB) * nseg;
    const small_div by_seg(nseg);
    struct pos {
        int rt, par, ks;
    };
    auto at = [&](int g) {
        pos p;
        p.rt = by_seg.quo(g);
        const int r = g - p.rt * nseg;
        p.par = r >= nks_e;
        p.ks = p.par ? r - nks_e : r;
        return p;
    };
    if (producer) {
        const double *sb = uniform(io.in[pi]) + uniform(uniform(P.s_off)[mp]);
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(sb, (long)nl * nrhs * 8);
        const int ptid = tid - 256, ca = ptid >> 3, q = ptid & 7;
        const unsigned a_off = (unsigned)min(c0 + ca, nrhs) * nl * 8;  // column nrhs and beyond read as 0
        using Raw = raw_step<F::VPT>;
        auto fetch = [&](int g) {
            const int l0 = 2 * F::KS * at(g).ks + 2 * F::VPT * q;
            Raw r;
            r.valid = 0;
#pragma unroll
            for (int i = 0; i < F::VPT; ++i) {
                r.d[i] = load_d2(rs, a_off + (unsigned)(l0 + 2 * i) * 8);
                r.valid |= (unsigned)(l0 + 2 * i < nl) << (2 * i) | (unsigned)(l0 + 2 * i + 1 < nl) << (2 * i + 1);
            }
            return r;
        };
        const Raw r0 = fetch(0), r1 = fetch(1);
        double mx = 0.0;
        const int n32 = cdiv(nl, 32);
        for (int i0 = 0; i0 < n32; i0 += 8) {
            double2 d[8][2];
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int e = 0; e < 2; ++e) d[u][e] = load_d2(rs, a_off + (unsigned)(32 * (i0 + u) + 4 * q + 2 * e) * 8);
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int l = 32 * (i0 + u) + 4 * q + 2 * e;
                    if (l < nl) mx = fmax(mx, fabs(d[u][e].x));
                    if (l + 1 < nl) mx = fmax(mx, fabs(d[u][e].y));
                }
        }
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
        if (q == 0) cexp[ca] = scale_exponent(mx, F::I8);
        __syncthreads();  // the exponents
        const int ea = cexp[ca];
        auto slice = [&](const Raw &r, int g, u32x4 *buf) {
            const int par = at(g).par;
            double v[F::VPT];
#pragma unroll
            for (int i = 0; i < F::VPT; ++i) v[i] = r.val(i, par);
            slicers<F, 1> sl;
            sl.sl[0].init(buf, 0, ca, F::VPT * q, v, ea);
            return sl;
        };
        ws_produce<F, Raw>(G, fetch, slice, r0, r1, ring);
    } else {
        const long *frag_off = uniform(P.frag_stp_off);
        const int *texp = uniform(P.texp_stp);
        const long f_lo = uniform(frag_off[mp]), f_hi = uniform(frag_off[mp + 1]);
        const __amdgpu_buffer_rsrc_t rt =
            make_frag_rsrc(reinterpret_cast<const u32x4 *>(uniform(P.tdig_stp)) + (f_lo << 6), f_hi - f_lo);
        double *ob = uniform(io.out[pi]) + (long)mp * nrhs * nt;
        auto bfrag = [&](int g) {
            const pos p = at(g);
            const int jb = p.rt * kJB + wave;
            return (p.par ? njb * nks_e * F::NB : 0) + (jb * (p.par ? nks_o : nks_e) + p.ks) * F::NB;
        };
        u32x4 b0[F::NB];
        get_b<F::NB>(rt, bfrag(0), lane, b0);
        auto close = [&](int g) {
            const pos p = at(g);
            if (p.ks == (p.par ? nks_o : nks_e) - 1) return 2;
            return (!F::I8 && (p.ks + 1) % kChunkSteps == 0) ? 1 : 0;
        };
        auto done = [&](int g, const double (&dacc)[16]) {
            const pos p = at(g);
            const int ln = opaque(lane);
            const int jb = p.rt * kJB + wave, j = jb * 32 + (ln & 31);
            const bool store = jb < njb && j < nt2;
            const int ej = store ? texp[((long)mp * 2 + p.par) * nt2 + j] : 0;
            const int ct = opaque(tid);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int cc = 8 * (v >> 2) + 4 * (ln >> 5) + (v & 3), c = c0 + cc;
                const double r = __builtin_ldexp(dacc[v], cexp[cc] + ej);
                if (p.par == 0) {
                    ebuf[v][ct] = r;
                } else if (store && c < nrhs) {
                    const double e = ebuf[v][ct];
                    ob[(long)c * nt + j] = e + r;
                    ob[(long)c * nt + (nt - 1 - j)] = e - r;
                }
            }
        };
        ws_consume<F>(G, bfrag, close, done, rt, b0, ring, 0, lane);
    }
}

// PtS, wave-specialised: work item (mp, c0) as pts_split_kernel; consumer wave w owns
// parity w / 2 and row block kKB rt + w % 2 of row tile rt; the producers slice both folds.
template <class F>
__global__ __launch_bounds__(kWsThreads, 1) void pts_split_ws_kernel(
    const prob_desc *__restrict__ probs, const batch_io io, const work_item *__restrict__ work)
{
    static_assert(kCB == 1 && kKB == 2, "two 32-row blocks per parity and row tile");
    __shared__ u32x4 ring[kWsRing][kWsABuf];
    __shared__ int cexp[2][kTileC];
    const int mp = uniform(work[blockIdx.x].mp);
    if (mp < 0) return;
    const int pi = uniform(work[blockIdx.x].pad) & 0xffff, c0 = uniform(work[blockIdx.x].c0);
    const prob_desc &P = probs[pi];
    const int nt = uniform(P.nt), nt2 = uniform(P.nt2), nrhs = uniform(P.nrhs);
    const int m = uniform(uniform(P.mvals)[mp]), nl = uniform(P.lmax) - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool producer = wave >= 4;
    const int njs = cdiv(nt2, F::KS), G = cdiv(ne, kTileRP) * njs;
    const small_div by_njs(njs);
    if (producer) {
        const double *pb = uniform(io.in[pi]) + (long)mp * nrhs * nt;
        const __amdgpu_buffer_rsrc_t rp = make_rsrc(pb, (long)nrhs * nt * 8);
        const int ptid = tid - 256, ca = ptid >> 3, q = ptid & 7;
        const unsigned a_off = (unsigned)min(c0 + ca, nrhs) * nt * 8;
        constexpr int H = F::VPT / 2;
        using Raw = raw_step<F::VPT>;
        // raw data of node step js: nodes j = j0 + i, j0 = KS js + VPT q, i < VPT, as the pairs
        // d[u] = (f(j0 + 2u), f(j0 + 2u + 1)) and their mirrors d[H + u]; zero for j >= nt2
        auto fetch_js = [&](int js) {
            const int j0 = F::KS * js + F::VPT * q;
            Raw r;
            r.valid = 0;
#pragma unroll
            for (int u = 0; u < H; ++u) {
                r.d[u] = load_d2(rp, a_off + (unsigned)(j0 + 2 * u) * 8u);
                const double2 g = load_d2(rp, a_off + (unsigned)(nt - 2 - j0 - 2 * u) * 8u);
                r.d[H + u] = make_double2(g.y, g.x);
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    if (j0 + 2 * u + h < nt2) r.valid |= 1u << (2 * u + h) | 1u << (2 * (H + u) + h);
            }
            return r;
        };
        auto fetch = [&](int g) { return fetch_js(g - by_njs.quo(g) * njs); };
        const Raw r0 = fetch(0), r1 = fetch(1);
        constexpr int EB = 8 / H;
        double me = 0.0, mo = 0.0;
        for (int j0 = 0; j0 < njs; j0 += EB) {
            Raw r[EB];
#pragma unroll
            for (int u = 0; u < EB; ++u) r[u] = fetch_js(j0 + u);  // steps past njs are masked
#pragma unroll
            for (int u = 0; u < EB; ++u)
#pragma unroll
                for (int i = 0; i < F::VPT; ++i) {
                    const double a = r[u].val(i >> 1, i & 1), b = r[u].val(H + (i >> 1), i & 1);
                    me = fmax(me, fabs(a + b));
                    mo = fmax(mo, fabs(a - b));
                }
        }
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            me = fmax(me, __shfl_xor(me, off, 64));
            mo = fmax(mo, __shfl_xor(mo, off, 64));
        }
        if (q == 0) {
            cexp[0][ca] = scale_exponent(me, F::I8);
            cexp[1][ca] = scale_exponent(mo, F::I8);
        }
        __syncthreads();  // the exponents
        const int ee = cexp[0][ca], eo = cexp[1][ca];
        // A blocks of a slot: [parity]
        auto slice = [&](const Raw &r, int, u32x4 *buf) {
            double ve[F::VPT], vo[F::VPT];
#pragma unroll
            for (int i = 0; i < F::VPT; ++i) {
                const double a = r.val(i >> 1, i & 1), b = r.val(H + (i >> 1), i & 1);
                ve[i] = a + b;
                vo[i] = a - b;
            }
            slicers<F, 2> sl;
            sl.sl[0].init(buf, 0, ca, F::VPT * q, ve, ee);
            sl.sl[1].init(buf, 1, ca, F::VPT * q, vo, eo);
            return sl;
        };
        ws_produce<F, Raw>(G, fetch, slice, r0, r1, ring);
    } else {
        const int par = wave >> 1, K_rows = par ? no : ne;
        const long *frag_off = uniform(P.frag_pts_off);
        const int *texp = uniform(P.texp_pts);
        const int region = par ? cdiv(ne, 32) * njs * F::NB : 0;
        const long f_lo = uniform(frag_off[mp]), f_hi = uniform(frag_off[mp + 1]);
        const __amdgpu_buffer_rsrc_t rt =
            make_frag_rsrc(reinterpret_cast<const u32x4 *>(uniform(P.tdig_pts)) + (f_lo << 6), f_hi - f_lo);
        double *sb = uniform(io.out[pi]) + uniform(uniform(P.s_off)[mp]);
        const long e_base = uniform(uniform(P.exp_pts_off)[mp]);
        auto bfrag = [&](int g) {
            const int rt_ = by_njs.quo(g), js = g - rt_ * njs;
            return region + ((kKB * rt_ + (wave & 1)) * njs + js) * F::NB;
        };
        u32x4 b0[F::NB];
        get_b<F::NB>(rt, bfrag(0), lane, b0);
        auto close = [&](int g) {
            const int js = g - by_njs.quo(g) * njs;
            if (js == njs - 1) return 2;
            return (!F::I8 && (js + 1) % kChunkSteps == 0) ? 1 : 0;
        };
        auto done = [&](int g, const double (&dacc)[16]) {
            const int ln = opaque(lane);
            const int kb = kKB * by_njs.quo(g) + (wave & 1), k = kb * 32 + (ln & 31);
            if (k < K_rows) {
                const int ek = texp[e_base + (par ? ne : 0) + k];
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int cc = 8 * (v >> 2) + 4 * (ln >> 5) + (v & 3), c = c0 + cc;
                    if (c < nrhs) sb[(long)c * nl + 2 * k + par] = __builtin_ldexp(dacc[v], cexp[par][cc] + ek);
                }
            }
        };
        ws_consume<F>(G, bfrag, close, done, rt, b0, ring, par, lane);
    }
}

// --------------------------------------------------------------- host side
}  // namespace

// XCD-balanced work lists of (problem, m-block, column tile) items (each runs every row
// tile of its column tile): whole m-blocks dealt to the eight XCDs by
// longest-processing-time, interleaved so that item i runs on XCD i % 8.  The problem
// index travels in the `pad` field.
std::vector<work_item> split_work(const std::vector<const raysht_plan *> &plans, bool stp, sched_opts o)
{
    struct blk {
        int prob, mp;
        double item;  // K steps of one item (plus the exponent pass)
    };
    std::vector<blk> order;
    for (int pi = 0; pi < (int)plans.size(); ++pi) {
        const raysht_plan *p = plans[pi];
        const int ks = is_i8(p->precision) ? 32 : 16;
        const int nrt = cdiv(p->nt2, kTileR), njs = cdiv(p->nt2, ks);
        for (int mp = 0; mp < p->nm; ++mp) {
            const int nl = p->lmax - p->mvals[mp] + 1, ne = (nl + 1) / 2, no = nl / 2;
            const double steps = stp ? (double)nrt * (cdiv(ne, ks) + std::max(1, cdiv(no, ks))) + cdiv(nl, 256)
                                     : (double)cdiv(ne, kTileRP) * njs + cdiv(njs, 8);
            order.push_back(blk{pi, mp, steps});
        }
    }
    std::stable_sort(order.begin(), order.end(), [](const blk &a, const blk &b) { return a.item > b.item; });
    std::vector<std::vector<work_item>> blocks;
    std::vector<double> cost;
    for (const blk &b : order) {
        const raysht_plan *p = plans[b.prob];
        const int nct = cdiv(p->nrhs, kTileC);
        std::vector<work_item> items;
        for (int c = 0; c < nct; ++c) items.push_back(work_item{b.mp, 0, c * kTileC, b.prob});
        blocks.push_back(std::move(items));
        cost.push_back(b.item * nct);
    }
    return deal_to_xcds(blocks, cost, o.xcd_groups);
}

namespace {

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

}  // namespace

int split_plan_init(raysht_plan *p)
{
    raysht_split &sp = p->split;
    const bool x6 = is_x6(p->precision), i8 = is_i8(p->precision);
    if (i8 && (p->nt2 > kMaxI8K || (p->lmax + 2) / 2 > kMaxI8K)) return RAYSHT_EUNSUPPORTED;
    const int nb_stp = x6 ? stp_x6::NB : stp_x3::NB, nb_pts = x6 ? pts_x6::NB : pts_x3::NB;
    static_assert(stp_x6::NB == stp_i6::NB && pts_x6::NB == pts_i6::NB && stp_x3::NB == stp_i3::NB &&
                      pts_x3::NB == pts_i3::NB,
                  "table slices");
    const int ks = i8 ? 32 : 16;
    const int nm = p->nm, nt2 = p->nt2, njb = cdiv(nt2, 32), njs = cdiv(nt2, ks);
    std::vector<long> fs(nm + 1, 0), fp(nm + 1, 0), ep(nm + 1, 0);
    for (int mp = 0; mp < nm; ++mp) {
        const int nl = p->lmax - p->mvals[mp] + 1, ne = (nl + 1) / 2, no = nl / 2;
        fs[mp + 1] = fs[mp] + (long)njb * (cdiv(ne, ks) + cdiv(no, ks)) * nb_stp;
        fp[mp + 1] = fp[mp] + (long)(cdiv(ne, 32) + cdiv(no, 32)) * njs * nb_pts;
        ep[mp + 1] = ep[mp] + ne + no;
    }
    const std::vector<work_item> ws = split_work({p}, true), wp = split_work({p}, false);
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
    hipLaunchKernelGGL(i8 ? split_stp_table_kernel<true> : split_stp_table_kernel<false>, dim3(nt2, 2, nm), dim3(64),
                       0, nullptr, p->lmax, nt2, nb_stp, p->d_mvals, p->d_te_off, p->d_to_off, p->d_pt_even,
                       p->d_pt_odd, sp.d_frag_stp_off, sp.d_tdig_stp, sp.d_texp_stp);
    RAYSHT_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(i8 ? split_pts_table_kernel<true> : split_pts_table_kernel<false>,
                       dim3((p->lmax + 2) / 2, 2, nm), dim3(64), 0, nullptr, p->lmax, nt2, nb_pts, p->d_mvals,
                       p->d_te_off, p->d_to_off, p->d_ip_even, p->d_ip_odd, sp.d_frag_pts_off, sp.d_exp_pts_off,
                       sp.d_tdig_pts, sp.d_texp_pts);
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

int launch_stp_split_batch(int precision, const prob_desc *d_desc, const batch_io &io, const work_item *work,
                           int n_work, hipStream_t st)
{
    if (n_work == 0) return RAYSHT_OK;
    // X6 forms: the one-role kernel (two multiplying waves per SIMD); X3 forms: the
    // wave-specialised kernel (its registers leave room for two workgroups per CU)
    const dim3 g(n_work), b(kThreads), bw(kWsThreads);
    switch (precision) {
    case RAYSHT_BF16X6: hipLaunchKernelGGL(stp_split_kernel<stp_x6>, g, b, 0, st, d_desc, io, work); break;
    case RAYSHT_BF16X3: hipLaunchKernelGGL(stp_split_ws_kernel<stp_x3>, g, bw, 0, st, d_desc, io, work); break;
    case RAYSHT_INT8X6: hipLaunchKernelGGL(stp_split_kernel<stp_i6>, g, b, 0, st, d_desc, io, work); break;
    case RAYSHT_INT8X3: hipLaunchKernelGGL(stp_split_ws_kernel<stp_i3>, g, bw, 0, st, d_desc, io, work); break;
    default: return RAYSHT_EINVAL;
    }
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_pts_split_batch(int precision, const prob_desc *d_desc, const batch_io &io, const work_item *work,
                           int n_work, hipStream_t st)
{
    if (n_work == 0) return RAYSHT_OK;
    const dim3 g(n_work), b(kThreads), bw(kWsThreads);
    switch (precision) {
    case RAYSHT_BF16X6: hipLaunchKernelGGL(pts_split_kernel<pts_x6>, g, b, 0, st, d_desc, io, work); break;
    case RAYSHT_BF16X3: hipLaunchKernelGGL(pts_split_ws_kernel<pts_x3>, g, bw, 0, st, d_desc, io, work); break;
    case RAYSHT_INT8X6: hipLaunchKernelGGL(pts_split_kernel<pts_i6>, g, b, 0, st, d_desc, io, work); break;
    case RAYSHT_INT8X3: hipLaunchKernelGGL(pts_split_ws_kernel<pts_i3>, g, bw, 0, st, d_desc, io, work); break;
    default: return RAYSHT_EINVAL;
    }
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_stp_split(raysht_plan *p, const double *s, double *p2, hipStream_t st)
{
    batch_io io{};
    io.in[0] = s;
    io.out[0] = p2;
    return launch_stp_split_batch(p->precision, p->d_desc, io, p->split.d_work_stp, p->split.n_work_stp, st);
}

int launch_pts_split(raysht_plan *p, const double *p2, double *s, hipStream_t st)
{
    batch_io io{};
    io.in[0] = p2;
    io.out[0] = s;
    return launch_pts_split_batch(p->precision, p->d_desc, io, p->split.d_work_pts, p->split.n_work_pts, st);
}

}  // namespace raysht
