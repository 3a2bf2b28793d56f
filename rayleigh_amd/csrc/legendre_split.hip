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
using any_x3 = form<3, 3, 2, 1>;
using stp_i6 = form<7, 6, 6, 7, true>;
using pts_i6 = form<6, 6, 5, 6, true>;
using any_i3 = form<3, 3, 2, 3, true>;
static_assert(stp_x6::npairs() == 27 && pts_x6::npairs() == 21 && any_x3::npairs() == 6, "pair counts");
static_assert(stp_i6::npairs() == 27 && pts_i6::npairs() == 21 && any_i3::npairs() == 6, "pair counts");

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

// adds the level sums into the FP64 sums and clears them (BF16 digits carry their
// weights 2^(-7-8s); the INT8 forms' integer level L sums are scaled by 2^(-14-8L))
template <class F>
__device__ __forceinline__ void flush(typename F::acc_t (&acc)[F::NG], double (&dacc)[16])
{
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        double sum = 0.0;
#pragma unroll
        for (int g = F::NG - 1; g >= 0; --g) {
            if constexpr (F::I8)
                sum += __builtin_ldexp((double)acc[g][v], -14 - 8 * g);
            else
                sum += (double)acc[g][v];
        }
        dacc[v] += sum;
    }
#pragma unroll
    for (int g = 0; g < F::NG; ++g) acc[g] = typename F::acc_t{};
}

// The data operand's NA digits, from one integer per value.  For |v| < 2^e,
// X = round(v 2^(8 NA - 1 - e)) (|X| <= 2^(8 NA - 1)) is formed from two FP64 steps (floor of
// v 2^(8 NA - 33 - e), then the rounded 32-bit remainder) and held with 128 added to each
// of its NA - 1 low bytes: byte NA-1-s (s = 1 .. NA-1) is then d_s + 128 and the arithmetic
// top bits are d_0, so that v 2^-e = sum_s d_s 2^(-7-8s) + O(2^(-8 NA)) with every
// |d_s| <= 128 (exact in BF16).  Each digit then costs one byte conversion and one FMA
// instead of the three FP64 adds and a conversion of a slice taken off the remainder.
template <int NA>
__device__ __forceinline__ unsigned long long offset_digits(double v, int e)
{
    static_assert(NA >= 2 && NA <= 7, "digits of one 64-bit integer");
    const double t = __builtin_ldexp(v, 8 * NA - 33 - e);  // |t| < 2^(8 NA - 32)
    const double h = __builtin_floor(t);
    const double r = __builtin_fmin(__builtin_fma(t - h, 4294967296.0, 0.5), 4294967295.0);
    const unsigned long long X = ((unsigned long long)(unsigned)(int)h << 32) | (unsigned)r;
    return X + (0x0000808080808080ull >> (8 * (7 - NA)));
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

// The K loop of one contraction: nks steps of F::KS.  fetch(ks) issues the loads of step
// ks's raw data (steps past the end read masked zeros) and slice(raw, buffer) returns
// the slicers that write its A fragments into an LDS buffer, one slice per step() call
// (spread between the MFMAs of the step before); step ks's fragments are in buffer
// ks % 2 before its MFMAs.  Raw data run two steps ahead of the MFMAs and table fragments (fragment
// f0 + ks * NB of resource rt) one step, so their latency hides behind a step of
// MFMAs.  Wave w multiplies A block ablk of the buffer by its table fragments; the
// result is added into dacc (level sums moved there every kChunkSteps steps).
// The step body has no branches (the last step slices and loads harmlessly past the
// end), so one wave keeps its matrix pipe and its vector ALU busy at once.  nks is uniform
// over the workgroup; every step ends with one barrier.
template <class F, class Raw, class Fetch, class Slice>
__device__ __forceinline__ void k_loop(const Fetch &fetch, const Slice &slice, int nks, __amdgpu_buffer_rsrc_t rt,
                                       int f0, u32x4 (*abuf)[kABuf], int ablk, int lane, double (&dacc)[16])
{
    if (nks <= 0) return;
    {
        auto sl = slice(fetch(0), abuf[0]);
#pragma unroll
        for (int sc = 0; sc < F::NA; ++sc) sl.step(sc);
    }
    // two register sets for the raw data and the table fragments, swapped by unrolling
    // the step loop by two (no copies between steps)
    Raw rA = fetch(1), rB;
    u32x4 bA[F::NB], bB[F::NB];
    get_b<F::NB>(rt, f0, lane, bA);
    __syncthreads();
    typename F::acc_t acc[F::NG];
#pragma unroll
    for (int g = 0; g < F::NG; ++g) acc[g] = typename F::acc_t{};
    // step ks: raw data of step ks + 1 in rc, table fragments of step ks in bc; loads
    // step ks + 2's raw data into rn and step ks + 1's fragments into bn
    auto step = [&](int ks, const Raw &rc, Raw &rn, const u32x4 (&bc)[F::NB], u32x4 (&bn)[F::NB]) {
        rn = fetch(ks + 2);
        get_b<F::NB>(rt, f0 + (ks + 1) * F::NB, lane, bn);
        u32x4 a[F::NA];
        get_a<F::NA>(abuf[ks & 1], ablk, lane, a);
        auto sl = slice(rc, abuf[(ks + 1) & 1]);
        mfma_pairs_sliced<F>(acc, a, bc, sl);
        __syncthreads();
    };
    auto steps = [&](int kc, int kend) {
        int ks = kc;
        for (; ks + 1 < kend; ks += 2) {
            step(ks, rA, rB, bA, bB);
            step(ks + 1, rB, rA, bB, bA);
        }
        if (ks < kend) {  // odd step count: only in the last chunk (kChunkSteps is even)
            step(ks, rA, rB, bA, bB);
        }
    };
    if constexpr (F::I8) {
        // int32 level sums are exact over the whole contraction (kMaxI8K), one move into FP64
        steps(0, nks);
        flush<F>(acc, dacc);
    } else {
        for (int kc = 0; kc < nks; kc += kChunkSteps) {
            steps(kc, min(nks, kc + kChunkSteps));
            flush<F>(acc, dacc);
        }
    }
}
static_assert(kChunkSteps % 2 == 0, "chunks hold whole step pairs");

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

// The K loop of a contraction whose two operands are both digit fragments in memory: A
// fragments fa0 + ks NA .. (resource ra), B fragments fb0 + ks NB .. (resource rt), both
// one step ahead in registers (two sets, swapped by unrolling the step loop by two); no
// LDS and no barrier, so each wave runs on its own.  The last prefetch reads one step
// past the wave's fragments (inside the resource or as zeros) and is not used.
template <class F>
__device__ __forceinline__ void direct_k_loop(__amdgpu_buffer_rsrc_t ra, int fa0, __amdgpu_buffer_rsrc_t rt, int fb0,
                                              int nks, int lane, double (&dacc)[16])
{
    if (nks <= 0) return;
    u32x4 aA[F::NA], aB[F::NA], bA[F::NB], bB[F::NB];
    get_b<F::NA>(ra, fa0, lane, aA);
    get_b<F::NB>(rt, fb0, lane, bA);
    typename F::acc_t acc[F::NG];
#pragma unroll
    for (int g = 0; g < F::NG; ++g) acc[g] = typename F::acc_t{};
    auto step = [&](int ks, const u32x4 (&ac)[F::NA], u32x4 (&an)[F::NA], const u32x4 (&bc)[F::NB],
                    u32x4 (&bn)[F::NB]) {
        get_b<F::NA>(ra, fa0 + (ks + 1) * F::NA, lane, an);
        get_b<F::NB>(rt, fb0 + (ks + 1) * F::NB, lane, bn);
        mfma_pairs<F>(acc, ac, bc);
    };
    auto steps = [&](int kc, int kend) {
        int ks = kc;
        for (; ks + 1 < kend; ks += 2) {
            step(ks, aA, aB, bA, bB);
            step(ks + 1, aB, aA, bB, bA);
        }
        if (ks < kend) step(ks, aA, aB, bA, bB);  // odd step count: only in the last chunk
    };
    if constexpr (F::I8) {
        steps(0, nks);
        flush<F>(acc, dacc);
    } else {
        for (int kc = 0; kc < nks; kc += kChunkSteps) {
            steps(kc, min(nks, kc + kChunkSteps));
            flush<F>(acc, dacc);
        }
    }
}

// --------------------------------------------------------------- StP data slicing
// One workgroup of 256 threads per (problem, m-block, 32-column block): the exponent of
// each column (from all its l), then every digit fragment of the block in both parities.
// Fragment (cb nks + ks) NA + s of a parity's region (even region first, ncb nks_e NA
// fragments): lane L, element e -> column 32 cb + L%32, parity-k = KS ks + EPL (L/32) + e,
// i.e. l = 2 k + par; values past the m-block or the last column are zeros.  The digits are
// the ones the contraction's A operand needs (offset_digits / pack_*), computed once per
// column instead of once per row tile.
template <class F>
__global__ __launch_bounds__(256) void stp_slice_kernel(const prob_desc *__restrict__ probs, const batch_io io,
                                                        const work_item *__restrict__ work)
{
    __shared__ int cexp[32];
    const int mp = uniform(work[blockIdx.x].mp);
    if (mp < 0) return;
    const int pi = uniform(work[blockIdx.x].pad) & 0xffff, c0 = uniform(work[blockIdx.x].c0);
    const prob_desc &P = probs[pi];
    const int nrhs = uniform(P.nrhs);
    const int m = uniform(uniform(P.mvals)[mp]), nl = uniform(P.lmax) - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const double *sb = uniform(io.in[pi]) + uniform(uniform(P.s_off)[mp]);
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(sb, (long)nl * nrhs * 8);
    const int tid = threadIdx.x;
    {
        const int cc = tid >> 3, q = tid & 7, c = c0 + cc;
        const unsigned off = (unsigned)min(c, nrhs) * nl * 8;  // column nrhs and beyond read as 0
        double mx = 0.0;
        for (int l = q; l < nl; l += 8)
            mx = fmax(mx, fabs(__builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)(off + 8u * l), 0,
                                                                                             0))));
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
        if (q == 0) {
            const int e = scale_exponent(mx, F::I8);
            cexp[cc] = e;
            if (c < nrhs) uniform(P.scexp)[(long)mp * nrhs + c] = e;
        }
    }
    __syncthreads();
    const int lane = tid & 63, grp = tid >> 6;
    const int cc = lane & 31, h = lane >> 5, c = c0 + cc;
    const int e = cexp[cc];
    const bool cin = c < nrhs;
    const unsigned off = (unsigned)min(c, nrhs) * nl * 8;
    const int nks_e = cdiv(ne, F::KS), nks_o = cdiv(no, F::KS), cb = c0 >> 5, ncb = cdiv(nrhs, 32);
    uint4 *dig = uniform(P.sdig) + (uniform(uniform(P.sdig_off)[mp]) << 6);
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        const int nks = par ? nks_o : nks_e;
        const long region = (par ? (long)ncb * nks_e * F::NA : 0) + (long)cb * nks * F::NA;
        for (int ks = grp; ks < nks; ks += 4) {
            unsigned long long X[F::EPL];
#pragma unroll
            for (int i = 0; i < F::EPL; ++i) {
                const int l = 2 * (F::KS * ks + F::EPL * h + i) + par;
                const double v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                                                rs, (int)(off + 8u * (unsigned)l), 0, 0));
                X[i] = offset_digits<F::NA>(cin && l < nl ? v : 0.0, e);
            }
#pragma unroll
            for (int sd = 0; sd < F::NA; ++sd) {
                unsigned w[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if constexpr (F::I8) {
                        const unsigned long long x4[4] = {X[4 * q], X[4 * q + 1], X[4 * q + 2], X[4 * q + 3]};
                        w[q] = pack_i8<F::NA>(x4, sd);
                    } else {
                        w[q] = pack_bf16(digit<F::NA>(X[2 * q], sd), digit<F::NA>(X[2 * q + 1], sd));
                    }
                }
                dig[((region + (long)ks * F::NA + sd) << 6) + lane] = uint4{w[0], w[1], w[2], w[3]};
            }
        }
    }
}

// --------------------------------------------------------------- StP
// Work item (mp, r0, c0): kTileC columns, kTileR theta rows from r0, both parities, of a
// plan whose data digits the slicing pass has written.  Wave w multiplies column block
// c0/32 by theta block r0/32 + w: both operands come straight from the digit fragments
// in memory (the data fragments are shared by the item's waves and by the other row
// tiles of the column tile through the caches), so the waves need no LDS and no barrier
// in the K loop; the scaled even-parity sums wait in LDS (one slot per thread) for the
// odd pass, which writes E + O and E - O.
template <class F>
__global__ __launch_bounds__(kThreads, kMinWaves) void stp_split2_kernel(
    const prob_desc *__restrict__ probs, const batch_io io, const work_item *__restrict__ work)
{
    static_assert(kCB == 1, "one 32-column block per StP item");
    __shared__ double ebuf[16][kThreads];
    const int mp = uniform(work[blockIdx.x].mp);
    if (mp < 0) return;
    const int pi = uniform(work[blockIdx.x].pad) & 0xffff, c0 = uniform(work[blockIdx.x].c0);
    const int r0 = uniform(work[blockIdx.x].r0);
    const prob_desc &P = probs[pi];
    const int nt = uniform(P.nt), nt2 = uniform(P.nt2), nrhs = uniform(P.nrhs);
    const long *frag_off = uniform(P.frag_stp_off);
    const int *texp = uniform(P.texp_stp);
    const int m = uniform(uniform(P.mvals)[mp]), nl = uniform(P.lmax) - m + 1, ne = (nl + 1) >> 1, no = nl >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const long f_lo = uniform(frag_off[mp]), f_hi = uniform(frag_off[mp + 1]);
    const __amdgpu_buffer_rsrc_t rt =
        make_frag_rsrc(reinterpret_cast<const u32x4 *>(uniform(P.tdig_stp)) + (f_lo << 6), f_hi - f_lo);
    const long d_lo = uniform(uniform(P.sdig_off)[mp]), d_hi = uniform(uniform(P.sdig_off)[mp + 1]);
    const __amdgpu_buffer_rsrc_t ra =
        make_frag_rsrc(reinterpret_cast<const u32x4 *>(uniform(P.sdig)) + (d_lo << 6), d_hi - d_lo);
    const int *cexp = uniform(P.scexp) + (long)mp * nrhs;
    const int njb = cdiv(nt2, 32), nks_e = cdiv(ne, F::KS), nks_o = cdiv(no, F::KS);
    const int cb = c0 >> 5, ncb = cdiv(nrhs, 32);
    double *ob = uniform(io.out[pi]) + (long)mp * nrhs * nt;
    const int jb = r0 / 32 + wave;
    if (jb >= njb) return;  // a row block past the last theta (no barrier follows)
    const int j = jb * 32 + (lane & 31);
    const bool jin = j < nt2;
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        double dacc[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) dacc[v] = 0.0;
        const int nks = par ? nks_o : nks_e;
        const int fb0 = (par ? njb * nks_e * F::NB : 0) + jb * nks * F::NB;
        const int fa0 = (par ? ncb * nks_e * F::NA : 0) + cb * nks * F::NA;
        direct_k_loop<F>(ra, fa0, rt, fb0, nks, lane, dacc);
        const int ej = jin ? texp[((long)mp * 2 + par) * nt2 + j] : 0;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int cc = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), c = c0 + cc;
            const double r = __builtin_ldexp(dacc[v], (c < nrhs ? cexp[c] : 0) + ej);
            if (par == 0) {
                ebuf[v][tid] = r;
            } else if (jin && c < nrhs) {
                const double e = ebuf[v][tid];
                ob[(long)c * nt + j] = e + r;
                ob[(long)c * nt + (nt - 1 - j)] = e - r;
            }
        }
    }
}

// --------------------------------------------------------------- StP (sliced in the kernel)
// Work item (mp, r0, c0): kTileC columns, kTileR theta rows from r0, both parities.
// The row tiles of one column tile are consecutive items, so that they run together
// and share the column tile's spectral data in L2.  Wave w multiplies column block w % kCB by theta block
// j0/32 + w/kCB.  The spectral tile is read as 16-byte pairs (l, l+1); each parity pass
// slices its half, VPT values per thread and step.  The column exponents come from one
// pass over the tile; the scaled even-parity sums wait in LDS for the odd pass.
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
    const int r0 = uniform(work[blockIdx.x].r0);
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

    // producer role of the thread: column ca, l = 32 i + 4 q .. +3 of step i
    const int ca = tid >> 3, q = tid & 7;
    const unsigned a_off = (unsigned)min(c0 + ca, nrhs) * nl * 8;  // column nrhs and beyond read as 0
    {
        double mx = 0.0;
        const int n32 = cdiv(nl, 32);
        for (int i0 = 0; i0 < n32; i0 += 4) {
            double2 d[4][2];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int e = 0; e < 2; ++e) d[u][e] = load_d2(rs, a_off + (unsigned)(32 * (i0 + u) + 4 * q + 2 * e) * 8);
#pragma unroll
            for (int u = 0; u < 4; ++u)
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
    const int cb = wave % kCB, njb = cdiv(nt2, 32), nks_e = cdiv(ne, F::KS);
    using Raw = raw_step<F::VPT>;
    double *ob = uniform(io.out[pi]) + (long)mp * nrhs * nt;

    {
        const int jb = r0 / 32 + wave / kCB;
        const bool active = jb < njb;
        const int j = jb * 32 + (lane & 31);
#pragma unroll
        for (int par = 0; par < 2; ++par) {
            double dacc[16];
#pragma unroll
            for (int v = 0; v < 16; ++v) dacc[v] = 0.0;
            const int K = par ? no : ne, nks = cdiv(K, F::KS);
            const int f0 = (par ? njb * nks_e * F::NB : 0) + jb * nks * F::NB;
            // step ks: parity-k = KS ks + VPT q + i (i < VPT), i.e. l = 2 KS ks + 2 VPT q + 2 i + par,
            // read as the pairs (l, l + 1) of both parities (l >= nl is the next column or
            // past the block: masked to zero)
            auto fetch = [&](int ks) {
                const int l0 = 2 * F::KS * ks + 2 * F::VPT * q;
                Raw r;
                r.valid = 0;
#pragma unroll
                for (int i = 0; i < F::VPT; ++i) {
                    r.d[i] = load_d2(rs, a_off + (unsigned)(l0 + 2 * i) * 8);
                    r.valid |= (unsigned)(l0 + 2 * i < nl) << (2 * i) | (unsigned)(l0 + 2 * i + 1 < nl) << (2 * i + 1);
                }
                return r;
            };
            auto slice = [&](const Raw &r, u32x4 *buf) {
                double v[F::VPT];
#pragma unroll
                for (int i = 0; i < F::VPT; ++i) v[i] = r.val(i, par);
                slicers<F, 1> sl;
                sl.sl[0].init(buf, ca >> 5, ca, F::VPT * q, v, ea);
                return sl;
            };
            k_loop<F, Raw>(fetch, slice, nks, rt, f0, abuf, cb, lane, dacc);
            const bool store = active && j < nt2;
            const int ej = store ? texp[((long)mp * 2 + par) * nt2 + j] : 0;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int cc = cb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), c = c0 + cc;
                const double r = __builtin_ldexp(dacc[v], cexp[cc] + ej);
                if (par == 0) {
                    ebuf[v][tid] = r;
                } else if (store && c < nrhs) {
                    const double e = ebuf[v][tid];
                    ob[(long)c * nt + j] = e + r;
                    ob[(long)c * nt + (nt - 1 - j)] = e - r;
                }
            }
        }
    }
}

// --------------------------------------------------------------- PtS
// Work item (mp, k0, c0): kTileC columns, kTileRP parity-l rows from k0 of both
// parities (the row tiles of one column tile are consecutive items and share its hybrid
// data in L2).  Wave w multiplies column block w % kCB of parity w / (kCB kKB) by
// row block k0/32 + (w / kCB) % kKB.  The data operand is the fold f(j) + f(nt-1-j) (even) or
// f(j) - f(nt-1-j) (odd) of the hybrid columns; each thread slices both folds of VPT
// nodes per step, and both folds' column exponents come from one pass over the tile.
// (INT8 forms: registers for three waves per SIMD, measured 7 % faster than two)
template <class F>
__global__ __launch_bounds__(kThreads, F::I8 ? 3 : kMinWaves) void pts_split_kernel(
    const prob_desc *__restrict__ probs, const batch_io io, const work_item *__restrict__ work)
{
    __shared__ u32x4 abuf[2][kABuf];
    __shared__ int cexp[2][kTileC];
    const int mp = uniform(work[blockIdx.x].mp);
    if (mp < 0) return;
    const int pi = uniform(work[blockIdx.x].pad) & 0xffff, c0 = uniform(work[blockIdx.x].c0);
    const int k0 = uniform(work[blockIdx.x].r0);
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
    const int njs = cdiv(nt2, F::KS);
    constexpr int H = F::VPT / 2;
    using Raw = raw_step<F::VPT>;

    // raw data of step js: nodes j = j0 + i, j0 = KS js + VPT q, i < VPT, as the pairs
    // d[u] = (f(j0 + 2u), f(j0 + 2u + 1)) and their mirrors d[H + u] = (f(nt-1-j0-2u),
    // f(nt-2-j0-2u)); zero for j >= nt2 (the mirrored offset is negative only past the
    // last valid j, where it wraps past the end or into a column whose values are masked)
    auto fetch = [&](int js) {
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
    {
        double me = 0.0, mo = 0.0;
        for (int j0 = 0; j0 < njs; j0 += 4) {
            Raw r[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) r[u] = fetch(j0 + u);  // steps past njs are masked
#pragma unroll
            for (int u = 0; u < 4; ++u)
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
    const int par = wave / (kCB * kKB), cb = wave % kCB, K_rows = par ? no : ne;
    const int region = par ? cdiv(ne, 32) * njs * F::NB : 0;
    const long f_lo = uniform(frag_off[mp]), f_hi = uniform(frag_off[mp + 1]);
    const __amdgpu_buffer_rsrc_t rt =
        make_frag_rsrc(reinterpret_cast<const u32x4 *>(uniform(P.tdig_pts)) + (f_lo << 6), f_hi - f_lo);
    double *sb = uniform(io.out[pi]) + uniform(uniform(P.s_off)[mp]);
    const long e_base = uniform(uniform(P.exp_pts_off)[mp]);
    // A blocks of a buffer: [parity][column block]
    auto slice = [&](const Raw &r, u32x4 *buf) {
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

    {
        const int kb = k0 / 32 + (wave / kCB) % kKB;
        const bool active = kb * 32 < K_rows;
        double dacc[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) dacc[v] = 0.0;
        k_loop<F, Raw>(fetch, slice, njs, rt, region + kb * njs * F::NB, abuf, kCB * par + cb, lane, dacc);

        const int k = kb * 32 + (lane & 31);
        if (active && k < K_rows) {
            const int ek = texp[e_base + (par ? ne : 0) + k];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int cc = cb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), c = c0 + cc;
                if (c < nrhs) sb[(long)c * nl + 2 * k + par] = __builtin_ldexp(dacc[v], cexp[par][cc] + ek);
            }
        }
    }
}

// --------------------------------------------------------------- host side
}  // namespace

// XCD-balanced work lists of (problem, m-block, row tile, column tile) items: whole
// m-blocks dealt to the eight XCDs by longest-processing-time, interleaved so that item
// i runs on XCD i % 8; within a block the row tiles of one column tile are consecutive.
// The problem index travels in the `pad` field.
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
            // K steps (plus the exponent pass) of all row tiles of one column tile
            const double steps = stp ? (double)nrt * (cdiv(ne, ks) + cdiv(no, ks) + cdiv(nl, 32))
                                     : (double)cdiv(ne, kTileRP) * (njs + njs);
            order.push_back(blk{pi, mp, steps});
        }
    }
    std::stable_sort(order.begin(), order.end(), [](const blk &a, const blk &b) { return a.item > b.item; });
    std::vector<std::vector<work_item>> blocks;
    std::vector<double> cost;
    for (const blk &b : order) {
        const raysht_plan *p = plans[b.prob];
        const int nct = cdiv(p->nrhs, kTileC), ne = (p->lmax - p->mvals[b.mp] + 2) / 2;
        const int tile = stp ? kTileR : kTileRP, nrt = cdiv(stp ? p->nt2 : ne, tile);
        std::vector<work_item> items;
        for (int c = 0; c < nct; ++c)
            for (int r = 0; r < nrt; ++r) items.push_back(work_item{b.mp, r * tile, c * kTileC, b.prob});
        blocks.push_back(std::move(items));
        cost.push_back(b.item * nct);
    }
    return deal_to_xcds(blocks, cost, o.xcd_groups);
}

// The slicing pass's items: (m-block, 0, first column) for every 32-column block of every
// problem (problem index in `pad`), whole m-blocks dealt to the XCDs longest first.
std::vector<work_item> split_slice_work(const std::vector<const raysht_plan *> &plans)
{
    struct blk {
        int prob, mp, nl;
    };
    std::vector<blk> order;
    for (int pi = 0; pi < (int)plans.size(); ++pi)
        for (int mp = 0; mp < plans[pi]->nm; ++mp)
            order.push_back(blk{pi, mp, plans[pi]->lmax - plans[pi]->mvals[mp] + 1});
    std::stable_sort(order.begin(), order.end(), [](const blk &a, const blk &b) { return a.nl > b.nl; });
    std::vector<std::vector<work_item>> blocks;
    std::vector<double> cost;
    for (const blk &b : order) {
        const int ncb = cdiv(plans[b.prob]->nrhs, 32);
        std::vector<work_item> items;
        for (int c = 0; c < ncb; ++c) items.push_back(work_item{b.mp, 0, 32 * c, b.prob});
        blocks.push_back(std::move(items));
        cost.push_back((double)b.nl * ncb);
    }
    return deal_to_xcds(blocks, cost, true);
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
    const int nb_stp = x6 ? stp_x6::NB : any_x3::NB, nb_pts = x6 ? pts_x6::NB : any_x3::NB;
    static_assert(stp_x6::NB == stp_i6::NB && pts_x6::NB == pts_i6::NB && any_x3::NB == any_i3::NB, "table slices");
    const int ks = i8 ? 32 : 16;
    const int nm = p->nm, nt2 = p->nt2, njb = cdiv(nt2, 32), njs = cdiv(nt2, ks);
    const int na_stp = x6 ? stp_x6::NA : any_x3::NA, ncb = cdiv(p->nrhs, 32);
    static_assert(stp_x6::NA == stp_i6::NA && any_x3::NA == any_i3::NA, "data slices");
    std::vector<long> fs(nm + 1, 0), fp(nm + 1, 0), ep(nm + 1, 0), sd(nm + 1, 0);
    for (int mp = 0; mp < nm; ++mp) {
        const int nl = p->lmax - p->mvals[mp] + 1, ne = (nl + 1) / 2, no = nl / 2;
        fs[mp + 1] = fs[mp] + (long)njb * (cdiv(ne, ks) + cdiv(no, ks)) * nb_stp;
        fp[mp + 1] = fp[mp] + (long)(cdiv(ne, 32) + cdiv(no, 32)) * njs * nb_pts;
        ep[mp + 1] = ep[mp] + ne + no;
        sd[mp + 1] = sd[mp] + (long)ncb * (cdiv(ne, ks) + cdiv(no, ks)) * na_stp;
    }
    const std::vector<work_item> ws = split_work({p}, true), wp = split_work({p}, false), wsl = split_slice_work({p});
    sp.n_work_stp = (int)ws.size();
    sp.n_work_pts = (int)wp.size();
    sp.n_work_slice = (int)wsl.size();
    int rc;
    if ((rc = dupload(&sp.d_frag_stp_off, fs)) || (rc = dupload(&sp.d_frag_pts_off, fp)) ||
        (rc = dupload(&sp.d_exp_pts_off, ep)) || (rc = dupload(&sp.d_work_stp, ws)) ||
        (rc = dupload(&sp.d_work_pts, wp)) || (rc = dalloc(&sp.d_tdig_stp, (size_t)fs[nm] * 64)) ||
        (rc = dalloc(&sp.d_tdig_pts, (size_t)fp[nm] * 64)) || (rc = dalloc(&sp.d_texp_stp, (size_t)nm * 2 * nt2)) ||
        (rc = dalloc(&sp.d_texp_pts, (size_t)ep[nm])) || (rc = dupload(&sp.d_sdig_off, sd)) ||
        (rc = dalloc(&sp.d_sdig, (size_t)sd[nm] * 64)) || (rc = dalloc(&sp.d_scexp, (size_t)nm * p->nrhs)) ||
        (rc = dupload(&sp.d_work_slice, wsl)))
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
                      (void *)sp.d_texp_stp, (void *)sp.d_texp_pts, (void *)sp.d_sdig, (void *)sp.d_sdig_off,
                      (void *)sp.d_scexp, (void *)sp.d_work_slice})
        if (ptr) (void)hipFree(ptr);
    sp = raysht_split();
}

namespace {
// the slicing pass, then the contraction, on one stream
template <class F>
void launch_stp2(const prob_desc *d_desc, const batch_io &io, const work_item *slice_work, int n_slice,
                 const work_item *work, int n_work, hipStream_t st)
{
    if (n_slice > 0) hipLaunchKernelGGL(stp_slice_kernel<F>, dim3(n_slice), dim3(256), 0, st, d_desc, io, slice_work);
    hipLaunchKernelGGL(stp_split2_kernel<F>, dim3(n_work), dim3(kThreads), 0, st, d_desc, io, work);
}
}  // namespace

int launch_stp_split_batch(int precision, const prob_desc *d_desc, const batch_io &io, const work_item *slice_work,
                           int n_slice, const work_item *work, int n_work, hipStream_t st)
{
    if (n_work == 0) return RAYSHT_OK;
    switch (precision) {
    case RAYSHT_BF16X6: launch_stp2<stp_x6>(d_desc, io, slice_work, n_slice, work, n_work, st); break;
    case RAYSHT_BF16X3: launch_stp2<any_x3>(d_desc, io, slice_work, n_slice, work, n_work, st); break;
    case RAYSHT_INT8X6: launch_stp2<stp_i6>(d_desc, io, slice_work, n_slice, work, n_work, st); break;
    case RAYSHT_INT8X3: launch_stp2<any_i3>(d_desc, io, slice_work, n_slice, work, n_work, st); break;
    default: return RAYSHT_EINVAL;
    }
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_pts_split_batch(int precision, const prob_desc *d_desc, const batch_io &io, const work_item *work,
                           int n_work, hipStream_t st)
{
    if (n_work == 0) return RAYSHT_OK;
    const dim3 g(n_work), b(kThreads);
    switch (precision) {
    case RAYSHT_BF16X6: hipLaunchKernelGGL(pts_split_kernel<pts_x6>, g, b, 0, st, d_desc, io, work); break;
    case RAYSHT_BF16X3: hipLaunchKernelGGL(pts_split_kernel<any_x3>, g, b, 0, st, d_desc, io, work); break;
    case RAYSHT_INT8X6: hipLaunchKernelGGL(pts_split_kernel<pts_i6>, g, b, 0, st, d_desc, io, work); break;
    case RAYSHT_INT8X3: hipLaunchKernelGGL(pts_split_kernel<any_i3>, g, b, 0, st, d_desc, io, work); break;
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
    return launch_stp_split_batch(p->precision, p->d_desc, io, p->split.d_work_slice, p->split.n_work_slice,
                                  p->split.d_work_stp, p->split.n_work_stp, st);
}

int launch_pts_split(raysht_plan *p, const double *p2, double *s, hipStream_t st)
{
    batch_io io{};
    io.in[0] = p2;
    io.out[0] = s;
    return launch_pts_split_batch(p->precision, p->d_desc, io, p->split.d_work_pts, p->split.n_work_pts, st);
}

}  // namespace raysht
