// Batched in-place real FFTs along phi and the one-rank row permutes around them.
//
// FFT_To_Physical / FFT_To_Spectral (Fourier_Transform.F90:56-183) run FFTW's
// unnormalised c2r (e^{+i}) / r2c (e^{-i}) in place on lines of nphi+2 doubles
// holding Re/Im of m = 0..nphi/2.  A length-nphi real transform is done as a
// length-nphi/2 complex transform plus the usual even/odd split pass.  Rayleigh's
// nphi = 3 (l_max + 1) is 3 * 2^k: those lengths run register-resident (below); other
// even lengths 2^a 3^b run as radix 8/4/2/3 Stockham stages in LDS.
//
// Transpose_2a3a / 3b2b (Spherical_Buffer.F90:196-703) reduce on one rank to
// pure index maps between the hybrid layout p2[mp][c][theta] and the physical
// layout p3[f][theta][r][nphi+2]; the m > l_max part of every line is zero after
// 2a->3a, which is Rayleigh's implicit de-aliasing in phi.
#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "raysht_internal.h"

// The hybrid intermediate of raysht_spectral_to_physical_scratch: a file-scope device array,
// one copy per GPU, held until the process ends.  Host code never passes this name to a
// kernel or a copy (on the host it is only a stand-in for the symbol): it uses the device
// address that hipGetSymbolAddress gives for the current device (scratch_address below).
static __device__ __align__(256) double scratch[RAYSHT_SCRATCH_BYTES / sizeof(double)];

namespace raysht {
namespace {

struct cplx {
    double re, im;
};

__host__ __device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }

constexpr int kFftThreads = 256;
constexpr int kMaxHalf = 4096;       // largest complex length (nphi/2) held in LDS
constexpr int kFftLdsBudget = 65536;  // bytes of LDS per workgroup (two workgroups per CU)

// e^{sign 2 pi i t / n2}, t in [0, n2), from the table of e^{+2 pi i u / nphi}, nphi = 2 n2
__device__ __forceinline__ cplx twiddle(const double2 *__restrict__ tw, int nphi, long t, int sign)
{
    const double2 v = tw[t % nphi];
    return {v.x, sign * v.y};
}

// cos / sin of 2 pi k / 5 and 2 pi k / 7 (k = 1, 2 / 1, 2, 3), the constants of the radix-5
// and radix-7 butterflies
constexpr double kC51 = 0.3090169943749474241022934171828190588602, kS51 = 0.9510565162951535721164393333793821434783;
constexpr double kC52 = -0.8090169943749474241022934171828190588602, kS52 = 0.5877852522924731291687059546390727685978;
constexpr double kC71 = 0.6234898018587335305250048840042398106324, kS71 = 0.7818314824680298087084445266740577502325;
constexpr double kC72 = -0.2225209339563144042889025644967947594664, kS72 = 0.9749279121818236070181316829939312172327;
constexpr double kC73 = -0.9009688679024191262361023195074450511660, kS73 = 0.4338837391175581204757683328483587546099;

// y_k = sum_r v_r e^{sign 2 pi i r k / 5}: with t = v1 + v4, v2 + v3 and u = v1 - v4, v2 - v3,
// y_1,4 = v0 + c1 t1 + c2 t2 +- i sign (s1 u1 + s2 u2), y_2,3 = v0 + c2 t1 + c1 t2 +- i sign (s2 u1 - s1 u2)
__host__ __device__ __forceinline__ void dft5(const cplx (&v)[5], cplx (&y)[5], int sign)
{
    const cplx t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
    y[0] = cadd(v[0], cadd(t1, t2));
    const cplx a1 = {v[0].re + kC51 * t1.re + kC52 * t2.re, v[0].im + kC51 * t1.im + kC52 * t2.im};
    const cplx a2 = {v[0].re + kC52 * t1.re + kC51 * t2.re, v[0].im + kC52 * t1.im + kC51 * t2.im};
    const cplx b1 = {kS51 * u1.re + kS52 * u2.re, kS51 * u1.im + kS52 * u2.im};
    const cplx b2 = {kS52 * u1.re - kS51 * u2.re, kS52 * u1.im - kS51 * u2.im};
    const cplx ib1 = {-sign * b1.im, sign * b1.re}, ib2 = {-sign * b2.im, sign * b2.re};
    y[1] = cadd(a1, ib1);
    y[4] = csub(a1, ib1);
    y[2] = cadd(a2, ib2);
    y[3] = csub(a2, ib2);
}

// the same for 7: t_k = v_k + v_{7-k}, u_k = v_k - v_{7-k} (k = 1, 2, 3),
// y_j, y_{7-j} = v0 + sum_k cos(2 pi j k / 7) t_k +- i sign sum_k sin(2 pi j k / 7) u_k
__host__ __device__ __forceinline__ void dft7(const cplx (&v)[7], cplx (&y)[7], int sign)
{
    const cplx t1 = cadd(v[1], v[6]), t2 = cadd(v[2], v[5]), t3 = cadd(v[3], v[4]);
    const cplx u1 = csub(v[1], v[6]), u2 = csub(v[2], v[5]), u3 = csub(v[3], v[4]);
    y[0] = cadd(v[0], cadd(t1, cadd(t2, t3)));
    // cos / sin of 2 pi j k / 7 for j = 1, 2, 3 and k = 1, 2, 3 (jk mod 7 folded into 1..3 with signs)
    const double C[3][3] = {{kC71, kC72, kC73}, {kC72, kC73, kC71}, {kC73, kC71, kC72}};
    const double S[3][3] = {{kS71, kS72, kS73}, {kS72, -kS73, -kS71}, {kS73, -kS71, kS72}};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const cplx a = {v[0].re + C[j][0] * t1.re + C[j][1] * t2.re + C[j][2] * t3.re,
                        v[0].im + C[j][0] * t1.im + C[j][1] * t2.im + C[j][2] * t3.im};
        const cplx b = {S[j][0] * u1.re + S[j][1] * u2.re + S[j][2] * u3.re,
                        S[j][0] * u1.im + S[j][1] * u2.im + S[j][2] * u3.im};
        const cplx ib = {-sign * b.im, sign * b.re};
        y[j + 1] = cadd(a, ib);
        y[6 - j] = csub(a, ib);
    }
}

// One radix-R Stockham butterfly (R = 2, 3, 4, 5, 7 or 8, fixed at compile time) of one line.
template <int R>
__device__ __forceinline__ void stockham_butterfly(const cplx *sl, cplx *dl, int j, int nb, int ns, int nphi,
                                                   const double2 *__restrict__ tw, int sign)
{
    const int jm = j % ns;
    // table step for angle 2 pi jm / (ns R) on a length-nphi table
    const long step = (long)jm * (nphi / (ns * R));
    cplx v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        cplx a = sl[j + r * nb];
        if (r > 0 && jm > 0) a = cmul(a, twiddle(tw, nphi, step * r, sign));
        v[r] = a;
    }
    cplx y[R];
    if constexpr (R == 5) {
        dft5(v, y, sign);
    } else if constexpr (R == 7) {
        dft7(v, y, sign);
    } else if constexpr (R == 2) {
        y[0] = cadd(v[0], v[1]);
        y[1] = csub(v[0], v[1]);
    } else if constexpr (R == 8) {
        // two radix-4 DFTs on the even / odd inputs, combined with W8^k = e^{sign 2 pi i k / 8}
        cplx e[4], o[4];
        {
            const cplx a = cadd(v[0], v[4]), b = csub(v[0], v[4]);
            const cplx c = cadd(v[2], v[6]), d0 = csub(v[2], v[6]);
            const cplx d = {-sign * d0.im, sign * d0.re};
            e[0] = cadd(a, c);
            e[1] = cadd(b, d);
            e[2] = csub(a, c);
            e[3] = csub(b, d);
        }
        {
            const cplx a = cadd(v[1], v[5]), b = csub(v[1], v[5]);
            const cplx c = cadd(v[3], v[7]), d0 = csub(v[3], v[7]);
            const cplx d = {-sign * d0.im, sign * d0.re};
            o[0] = cadd(a, c);
            o[1] = cadd(b, d);
            o[2] = csub(a, c);
            o[3] = csub(b, d);
        }
        const double r2 = 0.7071067811865475244008443621048490392848;  // 1/sqrt(2)
        const cplx w1 = {r2, sign * r2}, w3 = {-r2, sign * r2};
        const cplx t0 = o[0];
        const cplx t1 = cmul(o[1], w1);
        const cplx t2 = {-sign * o[2].im, sign * o[2].re};
        const cplx t3 = cmul(o[3], w3);
        y[0] = cadd(e[0], t0);
        y[1] = cadd(e[1], t1);
        y[2] = cadd(e[2], t2);
        y[3] = cadd(e[3], t3);
        y[4] = csub(e[0], t0);
        y[5] = csub(e[1], t1);
        y[6] = csub(e[2], t2);
        y[7] = csub(e[3], t3);
    } else if constexpr (R == 4) {
        const cplx a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        const cplx c = cadd(v[1], v[3]), d0 = csub(v[1], v[3]);
        const cplx d = {-sign * d0.im, sign * d0.re};  // (sign i) * (v1 - v3)
        y[0] = cadd(a, c);
        y[1] = cadd(b, d);
        y[2] = csub(a, c);
        y[3] = csub(b, d);
    } else {
        const double h = 0.8660254037844386467637231707529361834714;  // sqrt(3)/2
        const cplx t1 = cadd(v[1], v[2]);
        const cplx t2 = {v[0].re - 0.5 * t1.re, v[0].im - 0.5 * t1.im};
        const cplx d0 = csub(v[1], v[2]);
        const cplx t3 = {-sign * h * d0.im, sign * h * d0.re};
        y[0] = cadd(v[0], t1);
        y[1] = cadd(t2, t3);
        y[2] = csub(t2, t3);
    }
    const int base = (j / ns) * ns * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) dl[base + r * ns] = y[r];
}

template <int R>
__device__ __forceinline__ void stockham_stage(const cplx *src, cplx *dst, int n2, int lines, int ns, int nphi,
                                               const double2 *__restrict__ tw, int sign)
{
    const int nb = n2 / R;
    const int total = nb * lines;
    for (int jj = threadIdx.x; jj < total; jj += blockDim.x) {
        const int line = jj / nb, j = jj - line * nb;
        stockham_butterfly<R>(src + line * n2, dst + line * n2, j, nb, ns, nphi, tw, sign);
    }
}

// `lines` independent length-n2 complex DFTs of buf[line][n2] (LDS); the radix of stage
// s is 3 bits of `code` (0: 2, 1: 4, 2: 3, 3: 8, 4: 5, 5: 7), product n2; sign = +1 (inverse,
// unnormalised) or -1 (forward).  Stockham auto-sort, ping-ponging with tmp.  Returns
// the buffer that holds the result.
__device__ cplx *stockham(cplx *buf, cplx *tmp, int n2, int lines, int nphi, int code, int nrad,
                          const double2 *__restrict__ tw, int sign)
{
    cplx *src = buf, *dst = tmp;
    int ns = 1;
    for (int s = 0; s < nrad; ++s) {
        const int c = (code >> (3 * s)) & 7;
        int R;
        if (c == 4) {
            stockham_stage<5>(src, dst, n2, lines, ns, nphi, tw, sign);
            R = 5;
        } else if (c == 5) {
            stockham_stage<7>(src, dst, n2, lines, ns, nphi, tw, sign);
            R = 7;
        } else if (c == 3) {
            stockham_stage<8>(src, dst, n2, lines, ns, nphi, tw, sign);
            R = 8;
        } else if (c == 1) {
            stockham_stage<4>(src, dst, n2, lines, ns, nphi, tw, sign);
            R = 4;
        } else if (c == 0) {
            stockham_stage<2>(src, dst, n2, lines, ns, nphi, tw, sign);
            R = 2;
        } else {
            stockham_stage<3>(src, dst, n2, lines, ns, nphi, tw, sign);
            R = 3;
        }
        __syncthreads();
        cplx *t = src;
        src = dst;
        dst = t;
        ns *= R;
    }
    return src;
}

// dynamic LDS: 2 x lines x n2 complex
__global__ __launch_bounds__(kFftThreads) void fft_c2r_kernel(int nphi, long howmany, int lines, double *__restrict__ x,
                                                             const double2 *__restrict__ tw, int code, int nrad)
{
    extern __shared__ cplx fft_lds[];
    const int n2 = nphi / 2;
    const long ld = nphi + 2;
    cplx *bufa = fft_lds, *bufb = fft_lds + (long)lines * n2;
    for (long l0 = (long)blockIdx.x * lines; l0 < howmany; l0 += (long)gridDim.x * lines) {
        const int nl = (int)((howmany - l0) < lines ? (howmany - l0) : lines);
        // Z_k = A_k + i B_k, A = X_k + conj(X_{n2-k}), B = (X_k - conj(X_{n2-k})) e^{+2 pi i k / nphi};
        // Im X_0 and Im X_{n2} do not enter a c2r.
        for (int kk = threadIdx.x; kk < nl * n2; kk += blockDim.x) {
            const int line = kk / n2, k = kk - line * n2;
            const double2 *xl = reinterpret_cast<const double2 *>(x + (l0 + line) * ld);
            const double2 xa = xl[k], xb = xl[n2 - k];
            const int km = n2 - k;
            cplx a = {xa.x, k == 0 ? 0.0 : xa.y};
            cplx b = {xb.x, km == n2 ? 0.0 : xb.y};
            const cplx bc = cconj(b);
            const cplx A = cadd(a, bc);
            const cplx B = cmul(csub(a, bc), twiddle(tw, nphi, k, +1));
            bufa[line * n2 + k] = {A.re - B.im, A.im + B.re};
        }
        __syncthreads();
        cplx *res = stockham(bufa, bufb, n2, nl, nphi, code, nrad, tw, +1);
        for (int kk = threadIdx.x; kk < nl * n2; kk += blockDim.x) {
            const int line = kk / n2, k = kk - line * n2;
            double2 *xl = reinterpret_cast<double2 *>(x + (l0 + line) * ld);
            const cplx z = res[line * n2 + k];
            xl[k] = make_double2(z.re, z.im);
        }
        for (int line = threadIdx.x; line < nl; line += blockDim.x) {
            x[(l0 + line) * ld + nphi] = 0.0;
            x[(l0 + line) * ld + nphi + 1] = 0.0;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kFftThreads) void fft_r2c_kernel(int nphi, long howmany, int lines, double *__restrict__ x,
                                                             const double2 *__restrict__ tw, int code, int nrad,
                                                             int rescale)
{
    extern __shared__ cplx fft_lds[];
    const int n2 = nphi / 2;
    const long ld = nphi + 2;
    cplx *bufa = fft_lds, *bufb = fft_lds + (long)lines * n2;
    for (long l0 = (long)blockIdx.x * lines; l0 < howmany; l0 += (long)gridDim.x * lines) {
        const int nl = (int)((howmany - l0) < lines ? (howmany - l0) : lines);
        for (int kk = threadIdx.x; kk < nl * n2; kk += blockDim.x) {
            const int line = kk / n2, k = kk - line * n2;
            const double2 v = reinterpret_cast<const double2 *>(x + (l0 + line) * ld)[k];
            bufa[line * n2 + k] = {v.x, v.y};
        }
        __syncthreads();
        cplx *res = stockham(bufa, bufb, n2, nl, nphi, code, nrad, tw, -1);
        // X_k = E_k + e^{-2 pi i k / nphi} O_k, E = (Z_k + conj Z_{n2-k}) / 2, O = (Z_k - conj Z_{n2-k}) / (2i)
        for (int kk = threadIdx.x; kk < nl * (n2 + 1); kk += blockDim.x) {
            const int line = kk / (n2 + 1), k = kk - line * (n2 + 1);
            const cplx *rl = res + line * n2;
            const cplx zk = rl[k % n2];
            const cplx zc = cconj(rl[(n2 - k) % n2]);
            const cplx E = {0.5 * (zk.re + zc.re), 0.5 * (zk.im + zc.im)};
            const cplx D = csub(zk, zc);
            const cplx O = {0.5 * D.im, -0.5 * D.re};
            const cplx X = cadd(E, cmul(O, twiddle(tw, nphi, k, -1)));
            double re = X.re, im = X.im;
            if (rescale) {
                re = (k == 0) ? re / (2.0 * (double)nphi) : re / (double)nphi;
                if (k > 0) im = im / (double)nphi;
            }
            reinterpret_cast<double2 *>(x + (l0 + line) * ld)[k] = make_double2(re, im);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Register-resident real FFTs for Rayleigh's line lengths, nphi = 3 (l_max + 1)
// with n2 = nphi / 2 = 3 * 2^k (24 .. 3072).  The complex length-n2 transform runs
// as radix-8 Stockham passes followed by one radix-3, -6 or -12 pass, all sizes and
// strides compile-time: every thread holds 24 points (kP) of one line in registers,
// n2 / 24 threads share a line, and the passes exchange through LDS with a padded
// layout (one 16-byte gap per 16 points: conflict-free ds_read/write_b128 for every
// pass's pattern).  The real-to-complex packing rides on the first pass (c2r) or a
// final LDS pass (r2c), so each line is read and written in HBM exactly once.

constexpr int kP = 24;

__host__ __device__ constexpr int fft3_n8(int n)  // radix-8 passes before the last
{
    int a = 0;
    while (n % 8 == 0 && n / 8 >= 3 && (n / 8 == 3 || n / 8 == 6 || n / 8 == 12 || n / 8 >= 24)) {
        n /= 8;
        ++a;
    }
    return a;
}
__host__ __device__ constexpr int fft3_last(int n)
{
    for (int a = fft3_n8(n); a > 0; --a) n /= 8;
    return n;
}
__host__ __device__ constexpr int pow8(int a) { return a == 0 ? 1 : 8 * pow8(a - 1); }
__host__ __device__ constexpr bool fft3_ok(int n)
{
    return n >= kP && n % kP == 0 && (fft3_last(n) == 3 || fft3_last(n) == 6 || fft3_last(n) == 12) &&
           pow8(fft3_n8(n)) * fft3_last(n) == n;
}
__host__ __device__ constexpr int fft3_pad(int i) { return i + (i >> 4); }

// in-place small DFTs, y_k = sum_n v_n e^{sign 2 pi i n k / R}
__host__ __device__ __forceinline__ cplx mul_i(cplx a, int sign) { return {-sign * a.im, sign * a.re}; }

__host__ __device__ __forceinline__ void dft3(cplx &a0, cplx &a1, cplx &a2, int sign)
{
    const double h = 0.8660254037844386467637231707529361834714;  // sqrt(3)/2
    const cplx t1 = cadd(a1, a2);
    const cplx t2 = {a0.re - 0.5 * t1.re, a0.im - 0.5 * t1.im};
    const cplx d0 = csub(a1, a2);
    const cplx t3 = {-sign * h * d0.im, sign * h * d0.re};
    a0 = cadd(a0, t1);
    a1 = cadd(t2, t3);
    a2 = csub(t2, t3);
}

__host__ __device__ __forceinline__ void dft4(cplx &a0, cplx &a1, cplx &a2, cplx &a3, int sign)
{
    const cplx a = cadd(a0, a2), b = csub(a0, a2);
    const cplx c = cadd(a1, a3), d = mul_i(csub(a1, a3), sign);
    a0 = cadd(a, c);
    a1 = cadd(b, d);
    a2 = csub(a, c);
    a3 = csub(b, d);
}

__host__ __device__ __forceinline__ void dft8(cplx (&v)[8], int sign)
{
    cplx e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3, sign);
    dft4(o0, o1, o2, o3, sign);
    const double r2 = 0.7071067811865475244008443621048490392848;  // 1/sqrt(2)
    const cplx t1 = cmul(o1, cplx{r2, sign * r2});
    const cplx t2 = mul_i(o2, sign);
    const cplx t3 = cmul(o3, cplx{-r2, sign * r2});
    v[0] = cadd(e0, o0), v[4] = csub(e0, o0);
    v[1] = cadd(e1, t1), v[5] = csub(e1, t1);
    v[2] = cadd(e2, t2), v[6] = csub(e2, t2);
    v[3] = cadd(e3, t3), v[7] = csub(e3, t3);
}

__host__ __device__ __forceinline__ void dft6(cplx (&v)[6], int sign)
{
    cplx e0 = v[0], e1 = v[2], e2 = v[4], o0 = v[1], o1 = v[3], o2 = v[5];
    dft3(e0, e1, e2, sign);
    dft3(o0, o1, o2, sign);
    const double h = 0.8660254037844386467637231707529361834714;
    const cplx t1 = cmul(o1, cplx{0.5, sign * h});   // e^{sign 2 pi i / 6}
    const cplx t2 = cmul(o2, cplx{-0.5, sign * h});  // e^{sign 4 pi i / 6}
    v[0] = cadd(e0, o0), v[3] = csub(e0, o0);
    v[1] = cadd(e1, t1), v[4] = csub(e1, t1);
    v[2] = cadd(e2, t2), v[5] = csub(e2, t2);
}

__host__ __device__ __forceinline__ void dft12(cplx (&v)[12], int sign)
{
    cplx e[6] = {v[0], v[2], v[4], v[6], v[8], v[10]}, o[6] = {v[1], v[3], v[5], v[7], v[9], v[11]};
    dft6(e, sign);
    dft6(o, sign);
    const double h = 0.8660254037844386467637231707529361834714;
    const cplx w[6] = {{1.0, 0.0},  {h, sign * 0.5},  {0.5, sign * h},
                       {0.0, 1.0 * sign}, {-0.5, sign * h}, {-h, sign * 0.5}};  // e^{sign 2 pi i k / 12}
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const cplx t = k == 0 ? o[0] : cmul(o[k], w[k]);
        v[k] = cadd(e[k], t);
        v[k + 6] = csub(e[k], t);
    }
}

template <int R>
__host__ __device__ __forceinline__ void dft_r(cplx (&v)[R], int sign)
{
    if constexpr (R == 8) dft8(v, sign);
    else if constexpr (R == 12) dft12(v, sign);
    else if constexpr (R == 6) dft6(v, sign);
    else dft3(v[0], v[1], v[2], sign);
}

// One Stockham pass of radix R with input stride NS over the thread's kP points:
// butterfly u (b = t + T u) takes in[u*R + r] = element b + r N/R, twiddles input r by
// e^{sign 2 pi i r (b mod NS) / (NS R)} (table of e^{2 pi i v / 2N}), and leaves
// out element (b / NS) NS R + b mod NS + r NS in v[u*R + r].
template <int N, int R, int NS>
__device__ __forceinline__ void fft3_pass(cplx (&v)[kP], int t, const double2 *__restrict__ tw, int sign)
{
    constexpr int T = N / kP;
#pragma unroll
    for (int u = 0; u < kP / R; ++u) {
        const int b = t + T * u, k = b & (NS - 1);
        cplx y[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            y[r] = v[u * R + r];
            if (NS > 1 && r > 0) {
                const double2 w = tw[2 * r * k * (N / (NS * R))];
                y[r] = cmul(y[r], cplx{w.x, sign * w.y});
            }
        }
        dft_r<R>(y, sign);
#pragma unroll
        for (int r = 0; r < R; ++r) v[u * R + r] = y[r];
    }
}

template <int N, int R, int NS>
__device__ __forceinline__ void fft3_store_lds(cplx *__restrict__ buf, const cplx (&v)[kP], int t)
{
    constexpr int T = N / kP;
#pragma unroll
    for (int u = 0; u < kP / R; ++u) {
        const int b = t + T * u;
        const int o = (b / NS) * NS * R + (b & (NS - 1));
#pragma unroll
        for (int r = 0; r < R; ++r) buf[fft3_pad(o + r * NS)] = v[u * R + r];
    }
}

template <int N, int R>
__device__ __forceinline__ void fft3_load_lds(const cplx *__restrict__ buf, cplx (&v)[kP], int t)
{
    constexpr int T = N / kP;
#pragma unroll
    for (int u = 0; u < kP / R; ++u)
#pragma unroll
        for (int r = 0; r < R; ++r) v[u * R + r] = buf[fft3_pad(t + T * u + r * (N / R))];
}

// the radix-8 passes 1 .. A-1 and the last pass's LDS exchange; pass 0 is done by the caller
template <int N, int P8, int A>
__device__ __forceinline__ void fft3_middle(cplx *__restrict__ buf, cplx (&v)[kP], int t,
                                            const double2 *__restrict__ tw, int sign)
{
    if constexpr (P8 < A) {
        constexpr int NS = pow8(P8);
        fft3_store_lds<N, 8, NS / 8>(buf, v, t);
        __syncthreads();
        fft3_load_lds<N, 8>(buf, v, t);
        __syncthreads();
        fft3_pass<N, 8, NS>(v, t, tw, sign);
        fft3_middle<N, P8 + 1, A>(buf, v, t, tw, sign);
    } else {
        constexpr int F = fft3_last(N), NS = pow8(A);
        if constexpr (A > 0) {
            fft3_store_lds<N, 8, NS / 8>(buf, v, t);
            __syncthreads();
            fft3_load_lds<N, F>(buf, v, t);
            __syncthreads();
        }
        fft3_pass<N, F, NS>(v, t, tw, sign);
    }
}

template <int N>
constexpr int fft3_threads()
{
    return N / kP > 64 ? N / kP : 64;
}

// c2r: z_k = A_k + i B_k (A = X_k + conj X_{N-k}, B = (X_k - conj X_{N-k}) e^{2 pi i k / 2N}),
// inverse length-N FFT, x_{2n} + i x_{2n+1} = z_n.  Modes above mmax are taken as zeros
// without being read (the same doubles as zeros read from memory): after the row
// transpose of a plan they are the m > l_max its permute need not write.
template <int N>
__global__ __launch_bounds__(fft3_threads<N>()) void fft3_c2r_kernel(long howmany, double *__restrict__ x,
                                                                      const double2 *__restrict__ tw, int mmax)
{
    constexpr int T = N / kP, LPB = fft3_threads<N>() / T, A = fft3_n8(N), F = fft3_last(N);
    constexpr int R0 = A > 0 ? 8 : F;
    constexpr int STRIDE = fft3_pad(N - 1) + 1;
    __shared__ cplx lds[LPB * STRIDE];
    const int t = threadIdx.x % T, li = threadIdx.x / T;
    const long line = (long)blockIdx.x * LPB + li;
    if (line >= howmany) return;  // whole lines: the barriers below see only complete lines
    double *xl = x + line * (2L * N + 2);
    cplx *buf = lds + li * STRIDE;
    cplx v[kP];
#pragma unroll
    for (int u = 0; u < kP / R0; ++u)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int k = t + T * u + r * (N / R0);
            const double2 z = make_double2(0.0, 0.0);
            const double2 xa = k <= mmax ? reinterpret_cast<const double2 *>(xl)[k] : z;
            const double2 xb = N - k <= mmax ? reinterpret_cast<const double2 *>(xl)[N - k] : z;
            const cplx a = {xa.x, k == 0 ? 0.0 : xa.y};
            const cplx bc = {xb.x, k == 0 ? 0.0 : -xb.y};  // conj X_{N-k}; Im X_N does not enter
            const cplx Ak = cadd(a, bc);
            const double2 w = tw[k];
            const cplx Bk = cmul(csub(a, bc), cplx{w.x, w.y});
            v[u * R0 + r] = {Ak.re - Bk.im, Ak.im + Bk.re};
        }
    fft3_pass<N, R0, 1>(v, t, tw, +1);
    fft3_middle<N, 1, A>(buf, v, t, tw, +1);
    constexpr int NSL = N / F;
#pragma unroll
    for (int u = 0; u < kP / F; ++u)
#pragma unroll
        for (int r = 0; r < F; ++r)
            reinterpret_cast<double2 *>(xl)[t + T * u + r * NSL] = make_double2(v[u * F + r].re, v[u * F + r].im);
    if (t == 0) reinterpret_cast<double2 *>(xl)[N] = make_double2(0.0, 0.0);
}

// r2c: z_n = x_{2n} + i x_{2n+1}, forward length-N FFT, then
// X_k = (Z_k + conj Z_{N-k}) / 2 - i e^{-2 pi i k / 2N} (Z_k - conj Z_{N-k}) / 2, k = 0..N.
template <int N>
__global__ __launch_bounds__(fft3_threads<N>()) void fft3_r2c_kernel(long howmany, double *__restrict__ x,
                                                                      const double2 *__restrict__ tw, int rescale)
{
    constexpr int T = N / kP, LPB = fft3_threads<N>() / T, A = fft3_n8(N), F = fft3_last(N);
    constexpr int R0 = A > 0 ? 8 : F;
    constexpr int STRIDE = fft3_pad(N - 1) + 1;
    __shared__ cplx lds[LPB * STRIDE];
    const int t = threadIdx.x % T, li = threadIdx.x / T;
    const long line = (long)blockIdx.x * LPB + li;
    if (line >= howmany) return;
    double *xl = x + line * (2L * N + 2);
    cplx *buf = lds + li * STRIDE;
    cplx v[kP];
#pragma unroll
    for (int u = 0; u < kP / R0; ++u)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const double2 z = reinterpret_cast<const double2 *>(xl)[t + T * u + r * (N / R0)];
            v[u * R0 + r] = {z.x, z.y};
        }
    fft3_pass<N, R0, 1>(v, t, tw, -1);
    fft3_middle<N, 1, A>(buf, v, t, tw, -1);
    constexpr int NSL = N / F;
    // Z_o for o = t + T u + r N/F, unpadded order through LDS for the mirrored pairs
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kP / F; ++u)
#pragma unroll
        for (int r = 0; r < F; ++r) buf[fft3_pad(t + T * u + r * NSL)] = v[u * F + r];
    __syncthreads();
#pragma unroll
    for (int u = 0; u <= kP; ++u) {
        const int k = t + T * u;
        if (u == kP && k != N) continue;  // k = N once, by thread 0
        const cplx zk = buf[fft3_pad(k % N)];
        const cplx zc = cconj(buf[fft3_pad((N - k) % N)]);
        const cplx E = {0.5 * (zk.re + zc.re), 0.5 * (zk.im + zc.im)};
        const cplx D = csub(zk, zc);
        const cplx O = {0.5 * D.im, -0.5 * D.re};
        const double2 w = tw[k % (2 * N)];
        const cplx X = cadd(E, cmul(O, cplx{w.x, -w.y}));
        double re = X.re, im = X.im;
        if (rescale) {
            // r2c_ip4D_fftw_rsc: X(k > 0) / nphi, Re X_0 / (2 nphi), Im X_0 untouched
            re = k == 0 ? re / (4.0 * N) : re / (2.0 * N);
            if (k > 0) im = im / (2.0 * N);
        }
        reinterpret_cast<double2 *>(xl)[k] = make_double2(re, im);
    }
}

// ---------------------------------------------------------------------------
// The 3b -> 2b row transpose fused into the register-resident r2c (the
// reform_to_spectral / physical_to_spectral entry points).  A workgroup transforms TB
// lines of one (r, f) at TB consecutive theta; each thread turns its Z pairs into X_k and
// X_{N-k} in place in its line's LDS buffer, and the workgroup then writes the m <= l_max
// coefficients of its TB lines to the hybrid rows p2[mp][c][theta], TB consecutive theta of
// every mode q = 2m + imi at a time.  The padded grid is read once and the separate pass
// that read it back for 3b -> 2b is gone (1.16 ms against 0.94 + 0.67 ms at l_max = 255,
// N_r = 128, 8 fields); the X arithmetic is fft3_r2c_kernel's own, so the results are
// identical.  (The mirror fusion for c2r, gathering the hybrid rows through an LDS tile,
// measured slower than the 2a -> 3a kernel plus c2r: one wave per 31 KB workgroup left
// the gather latency-bound.)  Work unit w = (f, r, theta block), theta block fastest;
// workgroup b takes unit (b % 8) per + b / 8, so consecutive theta blocks of one (r, f) run
// on one XCD and share the cache lines of the hybrid rows in its L2.
template <int N>
__host__ __device__ constexpr int hyb_lines()
{
    return N <= 384 ? fft3_threads<N>() / (N / kP) : N == 768 ? 4 : N == 1536 ? 2 : 1;
}

struct hyb_args {
    int nt, nr, nf, lmax, ntb, units;
    const int *mp_of_m;
};

__device__ __forceinline__ bool hyb_unit(const hyb_args &a, int &r, int &f, int &t0, int TB)
{
    const int per = gridDim.x >> 3, b = blockIdx.x;
    const int w = (b & 7) * per + (b >> 3);
    if (w >= a.units) return false;
    const int tb = w % a.ntb, rf = w / a.ntb;
    r = rf % a.nr;
    f = rf / a.nr;
    t0 = tb * TB;
    return true;
}

template <int N>
__global__ __launch_bounds__(hyb_lines<N>() * (N / kP), 2) void fft3_r2c_hyb_kernel(hyb_args a,
                                                                              const double *__restrict__ p3,
                                                                              double *__restrict__ p2,
                                                                              const double2 *__restrict__ tw)
{
    constexpr int T = N / kP, TB = hyb_lines<N>(), NTH = TB * T, A = fft3_n8(N), F = fft3_last(N);
    constexpr int R0 = A > 0 ? 8 : F;
    constexpr int STRIDE = fft3_pad(N - 1) + 1;
    extern __shared__ double hyb_lds[];
    int r, f, t0;
    if (!hyb_unit(a, r, f, t0, TB)) return;
    const int tid = threadIdx.x, t = tid % T, li = tid / T, th = t0 + li;
    const bool line_in = th < a.nt;
    const double *xl = p3 + (((long)f * a.nt + (line_in ? th : 0)) * a.nr + r) * (2L * N + 2);
    cplx *buf = reinterpret_cast<cplx *>(hyb_lds) + li * STRIDE;
    cplx v[kP];
#pragma unroll
    for (int u = 0; u < kP / R0; ++u)
#pragma unroll
        for (int rr = 0; rr < R0; ++rr) {
            const double2 z = reinterpret_cast<const double2 *>(xl)[t + T * u + rr * (N / R0)];
            v[u * R0 + rr] = {z.x, z.y};
        }
    fft3_pass<N, R0, 1>(v, t, tw, -1);
    fft3_middle<N, 1, A>(buf, v, t, tw, -1);
    constexpr int NSL = N / F;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kP / F; ++u)
#pragma unroll
        for (int rr = 0; rr < F; ++rr) buf[fft3_pad(t + T * u + rr * NSL)] = v[u * F + rr];
    __syncthreads();
    // X_k and X_{N-k} from Z_k and Z_{N-k}, written back in place of the two Z: each thread
    // owns the pairs (k, N - k) for k = t + T u <= N / 2, so no other thread reads those
    // slots (k = 0 and k = N / 2 pair with themselves; the Nyquist term X_N, m = N > l_max, is
    // not read back).  The arithmetic of each X is fft3_r2c_kernel's own.
    auto x_of = [&](int k, const cplx &zk, const cplx &zn) {  // zn = Z_{(N-k) % N}
        const cplx zc = cconj(zn);
        const cplx E = {0.5 * (zk.re + zc.re), 0.5 * (zk.im + zc.im)};
        const cplx D = csub(zk, zc);
        const cplx O = {0.5 * D.im, -0.5 * D.re};
        const double2 w = tw[k];
        return cadd(E, cmul(O, cplx{w.x, -w.y}));
    };
#pragma unroll
    for (int u = 0; u <= kP / 2; ++u) {
        const int k = t + T * u;
        if (k > N / 2) continue;  // u = kP / 2 only for k = N / 2 (thread 0)
        const int kn = (N - k) % N;
        const cplx zk = buf[fft3_pad(k)], zn = buf[fft3_pad(kn)];
        const cplx xk = x_of(k, zk, zn);
        if (kn != k) buf[fft3_pad(kn)] = x_of(kn, zn, zk);
        buf[fft3_pad(k)] = xk;
    }
    __syncthreads();
    // the m <= l_max coefficients of the TB lines to the hybrid rows, TB theta at a time
    const int nq = 2 * (a.lmax + 1);
    const long nrhs = 2L * a.nr * a.nf;
    const double *xs = hyb_lds;  // line i's buffer at i STRIDE complex
    // eight stores per thread at a time, their m-block indices loaded together first, so the
    // loop waits for one round of index loads per eight stores rather than one per store
    constexpr int U = 8;
    const int nidx = nq * TB;
    for (int base = tid; base < nidx; base += U * NTH) {
        int mpv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int idx = base + u * NTH;
            mpv[u] = idx < nidx ? a.mp_of_m[(idx / TB) >> 1] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int idx = base + u * NTH, q = idx / TB, i = idx - q * TB, th2 = t0 + i;
            if (mpv[u] >= 0 && th2 < a.nt)
                p2[((long)mpv[u] * nrhs + r + (long)(q & 1) * a.nr + 2L * a.nr * f) * a.nt + th2] =
                    xs[2 * ((long)i * STRIDE + fft3_pad(q >> 1)) + (q & 1)];
        }
    }
}

// The one-rank row transposes are 2-D transposes per (r, f): the hybrid side is
// (mode q = 2m+imi) x theta with theta contiguous, the physical side theta x q with q
// contiguous.  A Q (q) x 64 (theta) tile goes through LDS; both sides move 16-byte
// pairs (theta pairs on the hybrid side, (Re, Im) of one m on the physical side), so a
// hybrid row is read or written as 512 contiguous bytes and a physical row as Q * 8.
// Rows are 16-byte aligned: N_theta and N_phi + 2 are even.
constexpr int kPT = 64, kPQ = 32;

// 2a -> 3a: p3 line (f, theta, r) gets Re/Im of m at 2m, 2m+1 and zeros elsewhere.
// Tiles of Q modes (q) x 64 theta: a physical row is written as Q * 8 contiguous bytes.
// qlim: modes q >= qlim are not written (nphi + 2 writes the whole line, its zeros included)
template <int Q>
__global__ __launch_bounds__(256) void permute_2a3a_kernel(int nt, int nr, int nf, int nphi, int lmax, int qlim,
                                                           const int *__restrict__ mp_of_m,
                                                           const double *__restrict__ p2, double *__restrict__ p3)
{
    __shared__ double tile[Q][kPT + 2];  // [q][theta]
    const int t0 = blockIdx.x * kPT, q0 = blockIdx.y * Q;
    const int rf = blockIdx.z, r = rf % nr, f = rf / nr;
    const long ld = nphi + 2, nrhs = 2L * nr * nf;
    const int tid = threadIdx.x;
    const bool any_m = (q0 >> 1) <= lmax;  // uniform: the tile holds modes of the plan
    if (any_m) {
        const int pr = tid & 31;  // theta pair of a hybrid row
#pragma unroll
        for (int i = 0; i < Q / 8; ++i) {
            const int qq = (tid >> 5) + 8 * i, q = q0 + qq, m = q >> 1, t = t0 + 2 * pr;
            double2 v = make_double2(0.0, 0.0);
            if (m <= lmax && t < nt) {
                const int mp = mp_of_m[m];
                if (mp >= 0)
                    v = *reinterpret_cast<const double2 *>(
                        p2 + ((long)mp * nrhs + r + (long)(q & 1) * nr + 2L * nr * f) * nt + t);
            }
            tile[qq][2 * pr] = v.x;
            tile[qq][2 * pr + 1] = v.y;
        }
        __syncthreads();
    }
    constexpr int QP = Q / 2, TR = 256 / QP;  // (Re, Im) pairs of a physical row, rows per pass
    const int qp = tid % QP;
#pragma unroll
    for (int i = 0; i < kPT / TR; ++i) {
        const int tt = tid / QP + TR * i, t = t0 + tt, q = q0 + 2 * qp;
        if (t < nt && q < qlim) {
            const double2 v = any_m ? make_double2(tile[2 * qp][tt], tile[2 * qp + 1][tt]) : make_double2(0.0, 0.0);
            *reinterpret_cast<double2 *>(p3 + (((long)f * nt + t) * nr + r) * ld + q) = v;
        }
    }
}

// 64 modes per 2a -> 3a tile: 512-byte physical row segments; measured faster than 32 at
// l_max = 255 and level with it at l_max = 1023.
constexpr int kPQ2a3a = 64;

// 3b -> 2b: the inverse map, reading back only the m of the plan (m <= lmax).
__global__ __launch_bounds__(256) void permute_3b2b_kernel(int nt, int nr, int nf, int nphi, int lmax,
                                                           const int *__restrict__ mp_of_m,
                                                           const double *__restrict__ p3, double *__restrict__ p2)
{
    __shared__ double tile[kPQ][kPT + 2];  // [q][theta]
    const int t0 = blockIdx.x * kPT, q0 = blockIdx.y * kPQ;
    const int rf = blockIdx.z, r = rf % nr, f = rf / nr;
    const long ld = nphi + 2, nrhs = 2L * nr * nf;
    const int tid = threadIdx.x, qmax = 2 * (lmax + 1);
    const int qp = tid & 15;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int tt = (tid >> 4) + 16 * i, t = t0 + tt, q = q0 + 2 * qp;
        double2 v = make_double2(0.0, 0.0);
        if (t < nt && q < qmax) v = *reinterpret_cast<const double2 *>(p3 + (((long)f * nt + t) * nr + r) * ld + q);
        tile[2 * qp][tt] = v.x;
        tile[2 * qp + 1][tt] = v.y;
    }
    __syncthreads();
    const int pr = tid & 31;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int qq = (tid >> 5) + 8 * i, q = q0 + qq, m = q >> 1, t = t0 + 2 * pr;
        if (q < qmax && t < nt) {
            const int mp = mp_of_m[m];
            if (mp >= 0)
                *reinterpret_cast<double2 *>(p2 + ((long)mp * nrhs + r + (long)(q & 1) * nr + 2L * nr * f) * nt + t) =
                    make_double2(tile[qq][2 * pr], tile[qq][2 * pr + 1]);
        }
    }
}

struct twiddle_cache {
    std::mutex mu;
    std::map<std::pair<int, int>, double2 *> tables;
} g_tw;

// The device table of e^{+2 pi i u / nphi} of the current device, made on first use.  With
// create = false the table must exist already (RAYSHT_ENOTREADY otherwise): nothing is
// allocated or copied.
int get_twiddles(int nphi, const double2 **out, bool create = true)
{
    int dev = 0;
    RAYSHT_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_tw.mu);
    auto key = std::make_pair(dev, nphi);
    auto it = g_tw.tables.find(key);
    if (it != g_tw.tables.end()) {
        *out = it->second;
        return RAYSHT_OK;
    }
    if (!create) return RAYSHT_ENOTREADY;
    std::vector<double> c(nphi), s(nphi);
    fft_twiddles(nphi, c.data(), s.data());
    std::vector<double2> h(nphi);
    for (int i = 0; i < nphi; ++i) h[i] = make_double2(c[i], s[i]);
    double2 *d = nullptr;
    if (hipMalloc(&d, sizeof(double2) * nphi) != hipSuccess) return RAYSHT_ENOMEM;
    RAYSHT_HIP_TRY(hipMemcpy(d, h.data(), sizeof(double2) * nphi, hipMemcpyHostToDevice));
    g_tw.tables[key] = d;
    *out = d;
    return RAYSHT_OK;
}

// the register-resident kernels for n2 = 3 * 2^k, 24 <= n2 <= 3072; returns false for
// other lengths (handled by the generic LDS Stockham kernels)
template <int N>
int launch_fft3(bool c2r, long howmany, double *x, const double2 *tw, int rescale, hipStream_t st, int mmax)
{
    static_assert(fft3_ok(N), "length without a register-resident schedule");
    constexpr int threads = fft3_threads<N>(), lpb = threads / (N / kP);
    const long blocks = (howmany + lpb - 1) / lpb;
    if (blocks > 0x7fffffffL) return RAYSHT_EUNSUPPORTED;
    if (c2r)
        hipLaunchKernelGGL(fft3_c2r_kernel<N>, dim3((unsigned)blocks), dim3(threads), 0, st, howmany, x, tw, mmax);
    else
        hipLaunchKernelGGL(fft3_r2c_kernel<N>, dim3((unsigned)blocks), dim3(threads), 0, st, howmany, x, tw,
                           rescale);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

// mmax: the c2r reads modes 0 .. mmax of each line and takes the rest as zeros
bool fft3_dispatch(int n2, bool c2r, long howmany, double *x, const double2 *tw, int rescale, hipStream_t st,
                   int *rc, int mmax = 1 << 30)
{
    switch (n2) {
    case 24: *rc = launch_fft3<24>(c2r, howmany, x, tw, rescale, st, mmax); return true;
    case 48: *rc = launch_fft3<48>(c2r, howmany, x, tw, rescale, st, mmax); return true;
    case 96: *rc = launch_fft3<96>(c2r, howmany, x, tw, rescale, st, mmax); return true;
    case 192: *rc = launch_fft3<192>(c2r, howmany, x, tw, rescale, st, mmax); return true;
    case 384: *rc = launch_fft3<384>(c2r, howmany, x, tw, rescale, st, mmax); return true;
    case 768: *rc = launch_fft3<768>(c2r, howmany, x, tw, rescale, st, mmax); return true;
    case 1536: *rc = launch_fft3<1536>(c2r, howmany, x, tw, rescale, st, mmax); return true;
    case 3072: *rc = launch_fft3<3072>(c2r, howmany, x, tw, rescale, st, mmax); return true;
    default: return false;
    }
}

// whether the c2r of nphi runs the register-resident kernel, which can take the modes
// above a plan's l_max as zeros (the generic kernel reads every mode)
bool c2r_band_limited(int nphi)
{
    switch (nphi / 2) {
    case 24: case 48: case 96: case 192: case 384: case 768: case 1536: case 3072: return true;
    default: return false;
    }
}

// the fused r2c + 3b -> 2b kernel of one register-resident length (arguments checked)
template <int N>
int launch_hyb(const raysht_plan *p, const double *src, double *dst, const double2 *tw, hipStream_t st)
{
    constexpr int TB = hyb_lines<N>(), NTH = TB * (N / kP), STRIDE = fft3_pad(N - 1) + 1;
    hyb_args a;
    a.nt = p->nt;
    a.nr = p->nr;
    a.nf = p->nf;
    a.lmax = p->lmax;
    a.ntb = (p->nt + TB - 1) / TB;
    a.mp_of_m = p->d_mp_of_m;
    const long units = (long)a.ntb * p->nr * p->nf;
    if (units > 0x7fffffffL / 16) return RAYSHT_EUNSUPPORTED;
    a.units = (int)units;
    const size_t lds = (size_t)TB * STRIDE * sizeof(cplx);
    const dim3 grid((unsigned)(8 * ((units + 7) / 8))), block(NTH);
    hipLaunchKernelGGL(fft3_r2c_hyb_kernel<N>, grid, block, lds, st, a, src, dst, tw);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

// the fused kernel for the register-resident lengths; false for other lengths (handled
// by r2c and the separate transpose)
bool hyb_dispatch(int n2, const raysht_plan *p, const double *src, double *dst, const double2 *tw, hipStream_t st,
                  int *rc)
{
    switch (n2) {
    case 24: *rc = launch_hyb<24>(p, src, dst, tw, st); return true;
    case 48: *rc = launch_hyb<48>(p, src, dst, tw, st); return true;
    case 96: *rc = launch_hyb<96>(p, src, dst, tw, st); return true;
    case 192: *rc = launch_hyb<192>(p, src, dst, tw, st); return true;
    case 384: *rc = launch_hyb<384>(p, src, dst, tw, st); return true;
    case 768: *rc = launch_hyb<768>(p, src, dst, tw, st); return true;
    case 1536: *rc = launch_hyb<1536>(p, src, dst, tw, st); return true;
    case 3072: *rc = launch_hyb<3072>(p, src, dst, tw, st); return true;
    default: return false;
    }
}

// radix schedule for n2 = 2^a 3^b 5^c 7^d: radix-8 stages, then one radix-4 or radix-2
// for the remaining power of two, then the radix-3, -5 and -7 stages; packed 3 bits per
// stage (0: 2, 1: 4, 2: 3, 3: 8, 4: 5, 5: 7), at most 10 stages.
bool fft_schedule(int n2, int *code, int *nrad)
{
    int e[4] = {0, 0, 0, 0}, n = n2;
    const int primes[4] = {2, 3, 5, 7};
    for (int i = 0; i < 4; ++i)
        while (n % primes[i] == 0) {
            n /= primes[i];
            ++e[i];
        }
    if (n != 1) return false;
    int c = 0, s = 0;
    auto push = [&](int radix_code) {
        if (s < 10) c |= radix_code << (3 * s);
        ++s;
    };
    for (int i = 0; i < e[0] / 3; ++i) push(3);
    if (e[0] % 3 == 2) push(1);
    if (e[0] % 3 == 1) push(0);
    for (int i = 0; i < e[1]; ++i) push(2);
    for (int i = 0; i < e[2]; ++i) push(4);
    for (int i = 0; i < e[3]; ++i) push(5);
    if (s > 10) return false;
    *code = c;
    *nrad = s;
    return true;
}

// lines transformed together by one workgroup: as many as fit the LDS budget (at least one)
int fft_lines_per_block(int n2)
{
    const int per_line = 2 * 16 * n2;
    int lines = kFftLdsBudget / per_line;
    return lines < 1 ? 1 : (lines > 16 ? 16 : lines);
}

int check_fft_args(int nphi, long howmany, const double *x)
{
    if (nphi < 2 || (nphi & 1) || howmany < 0 || (!x && howmany > 0)) return RAYSHT_EINVAL;
    if (nphi / 2 > kMaxHalf) return RAYSHT_EUNSUPPORTED;
    int code, nrad;
    if (!fft_schedule(nphi / 2, &code, &nrad)) return RAYSHT_EUNSUPPORTED;
    return RAYSHT_OK;
}

// the c2r of `howmany` lines with the twiddle table tw of nphi (arguments checked); with
// mmax, modes above it are zeros the register-resident kernel need not read
int launch_c2r(int nphi, long howmany, double *x, const double2 *tw, hipStream_t st, int mmax = 1 << 30)
{
    int rc = RAYSHT_OK;
    if (fft3_dispatch(nphi / 2, true, howmany, x, tw, 0, st, &rc, mmax)) return rc;
    int code = 0, nrad = 0;
    fft_schedule(nphi / 2, &code, &nrad);
    const int lines = fft_lines_per_block(nphi / 2);
    const long blocks = std::min<long>((howmany + lines - 1) / lines, 1L << 20);
    const size_t lds = 2 * sizeof(double2) * (size_t)lines * (nphi / 2);
    hipLaunchKernelGGL(fft_c2r_kernel, dim3((unsigned)blocks), dim3(kFftThreads), lds, st, nphi, howmany, lines, x,
                       tw, code, nrad);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

}  // namespace

// 2a -> 3a then c2r for a plan (arguments checked): where the c2r is the register-resident
// kernel, the permute leaves the modes above l_max unwritten and the c2r takes them as
// zeros (every line is then written whole by the c2r, with the same doubles)
int permute_c2r(raysht_plan *p, int nphi, const double *p2, double *p3, const double2 *tw, hipStream_t st)
{
    const bool band = c2r_band_limited(nphi);
    const int qlim = band ? 2 * (p->lmax + 1) : nphi + 2;
    dim3 grid((p->nt + kPT - 1) / kPT, (qlim + kPQ2a3a - 1) / kPQ2a3a, p->nr * p->nf);
    hipLaunchKernelGGL(permute_2a3a_kernel<kPQ2a3a>, grid, dim3(256), 0, st, p->nt, p->nr, p->nf, nphi, p->lmax, qlim,
                       p->d_mp_of_m, p2, p3);
    RAYSHT_HIP_TRY(hipGetLastError());
    return launch_c2r(nphi, (long)p->nr * p->nt * p->nf, p3, tw, st, band ? p->lmax : 1 << 30);
}

int fft_prepare(int nphi)
{
    const double2 *tw = nullptr;
    return get_twiddles(nphi, &tw);
}

}  // namespace raysht

using namespace raysht;

extern "C" {

int raysht_fft_to_physical(int nphi, long howmany, double *x, void *stream)
{
    int rc = check_fft_args(nphi, howmany, x);
    if (rc != RAYSHT_OK || howmany == 0) return rc;
    const double2 *tw = nullptr;
    if ((rc = get_twiddles(nphi, &tw)) != RAYSHT_OK) return rc;
    return launch_c2r(nphi, howmany, x, tw, reinterpret_cast<hipStream_t>(stream));
}

int raysht_fft_to_spectral(int nphi, long howmany, double *x, int rescale, void *stream)
{
    int rc = check_fft_args(nphi, howmany, x);
    if (rc != RAYSHT_OK || howmany == 0) return rc;
    const double2 *tw = nullptr;
    if ((rc = get_twiddles(nphi, &tw)) != RAYSHT_OK) return rc;
    if (fft3_dispatch(nphi / 2, false, howmany, x, tw, rescale ? 1 : 0, reinterpret_cast<hipStream_t>(stream), &rc))
        return rc;
    int code = 0, nrad = 0;
    fft_schedule(nphi / 2, &code, &nrad);
    const int lines = fft_lines_per_block(nphi / 2);
    const long blocks = std::min<long>((howmany + lines - 1) / lines, 1L << 20);
    const size_t lds = 2 * sizeof(double2) * (size_t)lines * (nphi / 2);
    hipLaunchKernelGGL(fft_r2c_kernel, dim3((unsigned)blocks), dim3(kFftThreads), lds,
                       reinterpret_cast<hipStream_t>(stream), nphi, howmany, lines, x, tw, code, nrad,
                       rescale ? 1 : 0);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

namespace {
// the device address of `scratch` on the current device, looked up once per device
struct scratch_table {
    std::mutex mu;
    std::map<int, double *> addr;
} g_scratch;

int scratch_address(double **out)
{
    int dev = 0;
    RAYSHT_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch.mu);
    auto it = g_scratch.addr.find(dev);
    if (it == g_scratch.addr.end()) {
        void *ptr = nullptr;
        RAYSHT_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return RAYSHT_EHIP;
        it = g_scratch.addr.emplace(dev, static_cast<double *>(ptr)).first;
    }
    *out = it->second;
    return RAYSHT_OK;
}
}  // namespace

int raysht_spectral_to_physical_scratch(raysht_plan *p, int nphi, const double *s, double *p3, void *stream)
{
    if (!p || !s || !p3) return RAYSHT_EINVAL;
    int rc = check_fft_args(nphi, 1, p3);
    if (rc != RAYSHT_OK) return rc;
    if (raysht_hybrid_size(p) > RAYSHT_SCRATCH_BYTES / sizeof(double)) return RAYSHT_EINVAL;
    if (p->nr < 1 || nphi < 2 * (p->lmax + 1) || (long)p->nr * p->nf > 65535) return RAYSHT_EINVAL;
    const double2 *tw = nullptr;
    if ((rc = get_twiddles(nphi, &tw, false)) != RAYSHT_OK) return rc;
    double *work = nullptr;
    if ((rc = scratch_address(&work)) != RAYSHT_OK) return rc;
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if ((rc = raysht_legendre_stp(p, s, work, stream)) != RAYSHT_OK) return rc;
    return permute_c2r(p, nphi, work, p3, tw, st);
}

int raysht_permute_2a3a(raysht_plan *p, int nphi, const double *p2, double *p3, void *stream)
{
    if (!p || !p2 || !p3 || p->nr < 1 || nphi < 2 * (p->lmax + 1) || (nphi & 1)) return RAYSHT_EINVAL;
    if ((long)p->nr * p->nf > 65535) return RAYSHT_EUNSUPPORTED;
    dim3 grid((p->nt + kPT - 1) / kPT, (nphi + 2 + kPQ2a3a - 1) / kPQ2a3a, p->nr * p->nf);
    hipLaunchKernelGGL(permute_2a3a_kernel<kPQ2a3a>, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), p->nt, p->nr,
                       p->nf, nphi, p->lmax, nphi + 2, p->d_mp_of_m, p2, p3);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_permute_3b2b(raysht_plan *p, int nphi, const double *p3, double *p2, void *stream)
{
    if (!p || !p2 || !p3 || p->nr < 1 || nphi < 2 * (p->lmax + 1) || (nphi & 1)) return RAYSHT_EINVAL;
    if ((long)p->nr * p->nf > 65535) return RAYSHT_EUNSUPPORTED;
    dim3 grid((p->nt + kPT - 1) / kPT, (2 * (p->lmax + 1) + kPQ - 1) / kPQ, p->nr * p->nf);
    hipLaunchKernelGGL(permute_3b2b_kernel, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), p->nt, p->nr,
                       p->nf, nphi, p->lmax, p->d_mp_of_m, p3, p2);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_reform_to_physical(raysht_plan *p, int nphi, const double *p2, double *p3, void *stream)
{
    if (!p || !p2 || !p3) return RAYSHT_EINVAL;
    int rc = check_fft_args(nphi, 1, p3);
    if (rc != RAYSHT_OK) return rc;
    if (p->nr < 1 || nphi < 2 * (p->lmax + 1) || (long)p->nr * p->nf > 65535) return RAYSHT_EINVAL;
    const double2 *tw = nullptr;
    if ((rc = get_twiddles(nphi, &tw)) != RAYSHT_OK) return rc;
    return permute_c2r(p, nphi, p2, p3, tw, reinterpret_cast<hipStream_t>(stream));
}

int raysht_reform_to_spectral(raysht_plan *p, int nphi, double *p3, double *p2, void *stream)
{
    if (!p || !p2 || !p3) return RAYSHT_EINVAL;
    int rc = check_fft_args(nphi, 1, p3);
    if (rc != RAYSHT_OK) return rc;
    if (p->nr < 1 || nphi < 2 * (p->lmax + 1) || (long)p->nr * p->nf > 65535) return RAYSHT_EINVAL;
    const double2 *tw = nullptr;
    if ((rc = get_twiddles(nphi, &tw)) != RAYSHT_OK) return rc;
    if (hyb_dispatch(nphi / 2, p, p3, p2, tw, reinterpret_cast<hipStream_t>(stream), &rc)) return rc;
    if ((rc = raysht_fft_to_spectral(nphi, (long)p->nr * p->nt * p->nf, p3, 0, stream)) != RAYSHT_OK) return rc;
    return raysht_permute_3b2b(p, nphi, p3, p2, stream);
}

int raysht_spectral_to_physical(raysht_plan *p, int nphi, const double *s, double *work, double *p3, void *stream)
{
    if (!p || !s || !work || !p3) return RAYSHT_EINVAL;
    int rc = check_fft_args(nphi, 1, p3);
    if (rc != RAYSHT_OK) return rc;
    if (p->nr < 1 || nphi < 2 * (p->lmax + 1) || (long)p->nr * p->nf > 65535) return RAYSHT_EINVAL;
    if ((rc = raysht_legendre_stp(p, s, work, stream)) != RAYSHT_OK) return rc;
    return raysht_reform_to_physical(p, nphi, work, p3, stream);
}

int raysht_physical_to_spectral(raysht_plan *p, int nphi, double *p3, double *work, double *s, void *stream)
{
    if (!p || !s || !work || !p3) return RAYSHT_EINVAL;
    int rc = raysht_reform_to_spectral(p, nphi, p3, work, stream);
    if (rc != RAYSHT_OK) return rc;
    return raysht_legendre_pts(p, work, s, stream);
}

}  // extern "C"
