// Batched in-place real FFTs along phi and the one-rank row permutes around them.
//
// FFT_To_Physical / FFT_To_Spectral (Fourier_Transform.F90:56-183) run FFTW's
// unnormalised c2r (e^{+i}) / r2c (e^{-i}) in place on lines of nphi+2 doubles
// holding Re/Im of m = 0..nphi/2.  Rayleigh's nphi = 3(l_max+1) is 3 * 2^k, so a
// length-nphi real transform is done as a length-nphi/2 complex transform (radix
// 4/2/3 Stockham stages in LDS) plus the usual even/odd split pass.
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

namespace raysht {
namespace {

struct cplx {
    double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }

constexpr int kFftThreads = 256;
constexpr int kMaxHalf = 2048;  // largest complex length (nphi/2) held in LDS

// e^{sign 2 pi i t / n2}, t in [0, n2), from the table of e^{+2 pi i u / nphi}, nphi = 2 n2
__device__ __forceinline__ cplx twiddle(const double2 *__restrict__ tw, int nphi, long t, int sign)
{
    const double2 v = tw[t % nphi];
    return {v.x, sign * v.y};
}

// Length-n2 complex DFT of buf (in LDS) with radix schedule `radices` (product n2),
// sign = +1 (inverse, unnormalised) or -1 (forward).  Result is returned in the
// buffer the function returns (buf or tmp).
__device__ cplx *stockham(cplx *buf, cplx *tmp, int n2, int nphi, const int *radices, int nrad,
                          const double2 *__restrict__ tw, int sign)
{
    cplx *src = buf, *dst = tmp;
    int ns = 1;
    for (int s = 0; s < nrad; ++s) {
        const int R = radices[s];
        const int nb = n2 / R;
        for (int j = threadIdx.x; j < nb; j += blockDim.x) {
            const int jm = j % ns;
            // table step for angle 2 pi jm / (ns R) on a length-nphi table
            const long step = (long)jm * (nphi / (ns * R));
            cplx v[4];
            for (int r = 0; r < R; ++r) {
                cplx a = src[j + r * nb];
                if (r > 0 && jm > 0) a = cmul(a, twiddle(tw, nphi, step * r, sign));
                v[r] = a;
            }
            cplx y[4];
            if (R == 2) {
                y[0] = cadd(v[0], v[1]);
                y[1] = csub(v[0], v[1]);
            } else if (R == 4) {
                const cplx a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
                const cplx c = cadd(v[1], v[3]), d0 = csub(v[1], v[3]);
                const cplx d = {-sign * d0.im, sign * d0.re};  // (sign i) * (v1 - v3)
                y[0] = cadd(a, c);
                y[1] = cadd(b, d);
                y[2] = csub(a, c);
                y[3] = csub(b, d);
            } else {  // R == 3
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
            for (int r = 0; r < R; ++r) dst[base + r * ns] = y[r];
        }
        __syncthreads();
        cplx *t = src;
        src = dst;
        dst = t;
        ns *= R;
    }
    return src;
}

__global__ __launch_bounds__(kFftThreads) void fft_c2r_kernel(int nphi, long howmany, double *__restrict__ x,
                                                             const double2 *__restrict__ tw, int4 sched, int nrad)
{
    __shared__ cplx bufa[kMaxHalf];
    __shared__ cplx bufb[kMaxHalf];
    const int n2 = nphi / 2;
    const long ld = nphi + 2;
    // radix schedule decoded from the packed int4 (up to 16 stages of 2 bits + the radix-3 flag)
    int rad[16];
    int nr = 0;
    {
        int code = sched.x;
        for (int s = 0; s < nrad; ++s) {
            const int c = (code >> (2 * s)) & 3;
            rad[nr++] = c == 0 ? 2 : c == 1 ? 4 : 3;
        }
    }
    for (long line = blockIdx.x; line < howmany; line += gridDim.x) {
        double *xl = x + line * ld;
        // Z_k = A_k + i B_k, A = X_k + conj(X_{n2-k}), B = (X_k - conj(X_{n2-k})) e^{+2 pi i k / nphi};
        // Im X_0 and Im X_{n2} do not enter a c2r.
        for (int k = threadIdx.x; k < n2; k += blockDim.x) {
            cplx a = {xl[2 * k], k == 0 ? 0.0 : xl[2 * k + 1]};
            const int km = n2 - k;
            cplx b = {xl[2 * km], km == n2 ? 0.0 : xl[2 * km + 1]};
            const cplx bc = cconj(b);
            const cplx A = cadd(a, bc);
            const cplx B = cmul(csub(a, bc), twiddle(tw, nphi, k, +1));
            bufa[k] = {A.re - B.im, A.im + B.re};
        }
        __syncthreads();
        cplx *res = stockham(bufa, bufb, n2, nphi, rad, nr, tw, +1);
        for (int k = threadIdx.x; k < n2; k += blockDim.x) {
            xl[2 * k] = res[k].re;
            xl[2 * k + 1] = res[k].im;
        }
        if (threadIdx.x == 0) {
            xl[nphi] = 0.0;
            xl[nphi + 1] = 0.0;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kFftThreads) void fft_r2c_kernel(int nphi, long howmany, double *__restrict__ x,
                                                             const double2 *__restrict__ tw, int4 sched, int nrad,
                                                             int rescale)
{
    __shared__ cplx bufa[kMaxHalf];
    __shared__ cplx bufb[kMaxHalf];
    const int n2 = nphi / 2;
    const long ld = nphi + 2;
    int rad[16];
    int nr = 0;
    {
        int code = sched.x;
        for (int s = 0; s < nrad; ++s) {
            const int c = (code >> (2 * s)) & 3;
            rad[nr++] = c == 0 ? 2 : c == 1 ? 4 : 3;
        }
    }
    for (long line = blockIdx.x; line < howmany; line += gridDim.x) {
        double *xl = x + line * ld;
        for (int k = threadIdx.x; k < n2; k += blockDim.x) bufa[k] = {xl[2 * k], xl[2 * k + 1]};
        __syncthreads();
        cplx *res = stockham(bufa, bufb, n2, nphi, rad, nr, tw, -1);
        // X_k = E_k + e^{-2 pi i k / nphi} O_k, E = (Z_k + conj Z_{n2-k}) / 2, O = (Z_k - conj Z_{n2-k}) / (2i)
        for (int k = threadIdx.x; k <= n2; k += blockDim.x) {
            const cplx zk = res[k % n2];
            const cplx zc = cconj(res[(n2 - k) % n2]);
            const cplx E = {0.5 * (zk.re + zc.re), 0.5 * (zk.im + zc.im)};
            const cplx D = csub(zk, zc);
            const cplx O = {0.5 * D.im, -0.5 * D.re};
            const cplx X = cadd(E, cmul(O, twiddle(tw, nphi, k, -1)));
            double re = X.re, im = X.im;
            if (rescale) {
                re = (k == 0) ? re / (2.0 * (double)nphi) : re / (double)nphi;
                if (k > 0) im = im / (double)nphi;
            }
            xl[2 * k] = re;
            xl[2 * k + 1] = im;
        }
        __syncthreads();
    }
}

// 2a -> 3a: each thread writes one element of a physical line.
__global__ void permute_2a3a_kernel(int nt, int nr, int nf, int nphi, int lmax, const int *__restrict__ mp_of_m,
                                    const double *__restrict__ p2, double *__restrict__ p3, long total)
{
    const long ld = nphi + 2;
    const long nrhs = 2L * nr * nf;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int q = (int)(i % ld);
        const long line = i / ld;
        const int r = (int)(line % nr);
        const long rest = line / nr;
        const int t = (int)(rest % nt);
        const int f = (int)(rest / nt);
        const int m = q >> 1, imi = q & 1;
        double v = 0.0;
        if (m <= lmax) {
            const int mp = mp_of_m[m];
            if (mp >= 0) v = p2[((long)mp * nrhs + r + (long)imi * nr + 2L * nr * f) * nt + t];
        }
        p3[i] = v;
    }
}

// 3b -> 2b: each thread writes one hybrid element.
__global__ void permute_3b2b_kernel(int nt, int nr, int nf, int nphi, const int *__restrict__ mvals,
                                    const double *__restrict__ p3, double *__restrict__ p2, long total)
{
    const long ld = nphi + 2;
    const long nrhs = 2L * nr * nf;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int t = (int)(i % nt);
        const long col = i / nt;
        const int c = (int)(col % nrhs);
        const int mp = (int)(col / nrhs);
        const int r = c % nr, imi = (c / nr) & 1, f = c / (2 * nr);
        const int m = mvals[mp];
        p2[i] = p3[(((long)f * nt + t) * nr + r) * ld + 2 * m + imi];
    }
}

struct twiddle_cache {
    std::mutex mu;
    std::map<std::pair<int, int>, double2 *> tables;
} g_tw;

int get_twiddles(int nphi, const double2 **out)
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

// radix schedule for n2 = 2^a 3^b: radix-4 stages, one radix-2 if a is odd, then
// the radix-3 stages; packed 2 bits per stage (0: 2, 1: 4, 2: 3).
bool fft_schedule(int n2, int *code, int *nrad)
{
    int a = 0, b = 0, n = n2;
    while (n % 2 == 0) {
        n /= 2;
        ++a;
    }
    while (n % 3 == 0) {
        n /= 3;
        ++b;
    }
    if (n != 1) return false;
    int c = 0, s = 0;
    for (int i = 0; i < a / 2; ++i) c |= 1 << (2 * s++);
    if (a & 1) c |= 0 << (2 * s++);
    for (int i = 0; i < b; ++i) c |= 2 << (2 * s++);
    if (s > 16) return false;
    *code = c;
    *nrad = s;
    return true;
}

int check_fft_args(int nphi, long howmany, const double *x)
{
    if (nphi < 2 || (nphi & 1) || howmany < 0 || (!x && howmany > 0)) return RAYSHT_EINVAL;
    if (nphi / 2 > kMaxHalf) return RAYSHT_EUNSUPPORTED;
    int code, nrad;
    if (!fft_schedule(nphi / 2, &code, &nrad)) return RAYSHT_EUNSUPPORTED;
    return RAYSHT_OK;
}

}  // namespace
}  // namespace raysht

using namespace raysht;

extern "C" {

int raysht_fft_to_physical(int nphi, long howmany, double *x, void *stream)
{
    int rc = check_fft_args(nphi, howmany, x);
    if (rc != RAYSHT_OK || howmany == 0) return rc;
    const double2 *tw = nullptr;
    if ((rc = get_twiddles(nphi, &tw)) != RAYSHT_OK) return rc;
    int code = 0, nrad = 0;
    fft_schedule(nphi / 2, &code, &nrad);
    const long grid = howmany < 65536 * 8 ? howmany : 65536 * 8;
    hipLaunchKernelGGL(fft_c2r_kernel, dim3((unsigned)grid), dim3(kFftThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), nphi, howmany, x, tw, make_int4(code, 0, 0, 0), nrad);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_fft_to_spectral(int nphi, long howmany, double *x, int rescale, void *stream)
{
    int rc = check_fft_args(nphi, howmany, x);
    if (rc != RAYSHT_OK || howmany == 0) return rc;
    const double2 *tw = nullptr;
    if ((rc = get_twiddles(nphi, &tw)) != RAYSHT_OK) return rc;
    int code = 0, nrad = 0;
    fft_schedule(nphi / 2, &code, &nrad);
    const long grid = howmany < 65536 * 8 ? howmany : 65536 * 8;
    hipLaunchKernelGGL(fft_r2c_kernel, dim3((unsigned)grid), dim3(kFftThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), nphi, howmany, x, tw, make_int4(code, 0, 0, 0), nrad,
                       rescale ? 1 : 0);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_permute_2a3a(raysht_plan *p, int nphi, const double *p2, double *p3, void *stream)
{
    if (!p || !p2 || !p3 || p->nr < 1 || nphi < 2 * (p->lmax + 1) || (nphi & 1)) return RAYSHT_EINVAL;
    const long total = (long)(nphi + 2) * p->nr * p->nt * p->nf;
    const long blocks = std::min<long>((total + 255) / 256, 1L << 20);
    hipLaunchKernelGGL(permute_2a3a_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), p->nt, p->nr, p->nf, nphi, p->lmax, p->d_mp_of_m, p2,
                       p3, total);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_permute_3b2b(raysht_plan *p, int nphi, const double *p3, double *p2, void *stream)
{
    if (!p || !p2 || !p3 || p->nr < 1 || nphi < 2 * (p->lmax + 1) || (nphi & 1)) return RAYSHT_EINVAL;
    const long total = (long)p->nt * p->nrhs * p->nm;
    const long blocks = std::min<long>((total + 255) / 256, 1L << 20);
    hipLaunchKernelGGL(permute_3b2b_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), p->nt, p->nr, p->nf, nphi, p->d_mvals, p3, p2, total);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_spectral_to_physical(raysht_plan *p, int nphi, const double *s, double *work, double *p3, void *stream)
{
    if (!p || !s || !work || !p3) return RAYSHT_EINVAL;
    int rc = check_fft_args(nphi, 1, p3);
    if (rc != RAYSHT_OK) return rc;
    if ((rc = raysht_legendre_stp(p, s, work, stream)) != RAYSHT_OK) return rc;
    if ((rc = raysht_permute_2a3a(p, nphi, work, p3, stream)) != RAYSHT_OK) return rc;
    return raysht_fft_to_physical(nphi, (long)p->nr * p->nt * p->nf, p3, stream);
}

int raysht_physical_to_spectral(raysht_plan *p, int nphi, double *p3, double *work, double *s, void *stream)
{
    if (!p || !s || !work || !p3) return RAYSHT_EINVAL;
    int rc = check_fft_args(nphi, 1, p3);
    if (rc != RAYSHT_OK) return rc;
    if ((rc = raysht_fft_to_spectral(nphi, (long)p->nr * p->nt * p->nf, p3, 0, stream)) != RAYSHT_OK) return rc;
    if ((rc = raysht_permute_3b2b(p, nphi, p3, work, stream)) != RAYSHT_OK) return rc;
    return raysht_legendre_pts(p, work, s, stream);
}

}  // extern "C"
