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

constexpr int kChebThreads = 256;
constexpr int kChebMaxN = 512;
constexpr int kChebCols = 8;  // columns per workgroup

// dynamic LDS: kChebCols columns of n doubles for each folded half
__global__ __launch_bounds__(kChebThreads) void cheb_to_spectral_kernel(int n, int n_x, int n_even, int n_odd,
                                                                       long ncols, const double *__restrict__ ce,
                                                                       const double *__restrict__ co,
                                                                       const double *__restrict__ f,
                                                                       double *__restrict__ c)
{
    extern __shared__ double lds[];
    double *fe = lds;                   // [kChebCols][n_x]
    double *fo = lds + kChebCols * n_x;  // [kChebCols][n_x]
    const long col0 = (long)blockIdx.x * kChebCols;
    const int ncol = (int)((ncols - col0) < kChebCols ? (ncols - col0) : kChebCols);
    for (int idx = threadIdx.x; idx < ncol * n_x; idx += blockDim.x) {
        const int cc = idx / n_x, i = idx % n_x;
        const double *fc = f + (col0 + cc) * n;
        const double a = fc[i], b = fc[n - 1 - i];
        fe[cc * n_x + i] = a + b;
        fo[cc * n_x + i] = a - b;
    }
    __syncthreads();
    const double alpha = 2.0 / (double)n;
    for (int idx = threadIdx.x; idx < ncol * n; idx += blockDim.x) {
        const int cc = idx / n, q = idx % n;
        const int k = q >> 1;
        double acc = 0.0;
        if ((q & 1) == 0) {
            const double *t = ce + (long)k * n_x;
            const double *v = fe + cc * n_x;
            for (int i = 0; i < n_x; ++i) acc = __builtin_fma(t[i], v[i], acc);
        } else {
            const double *t = co + (long)k * n_x;
            const double *v = fo + cc * n_x;
            for (int i = 0; i < n_x; ++i) acc = __builtin_fma(t[i], v[i], acc);
        }
        c[(col0 + cc) * n + q] = alpha * acc;
    }
}

__global__ __launch_bounds__(kChebThreads) void cheb_from_spectral_kernel(int n, int n_x, int n_even, int n_odd,
                                                                         long ncols, const double *__restrict__ ce,
                                                                         const double *__restrict__ co,
                                                                         const double *__restrict__ c,
                                                                         double *__restrict__ f)
{
    extern __shared__ double lds[];
    double *cs = lds;  // [kChebCols][n] coefficient columns
    const long col0 = (long)blockIdx.x * kChebCols;
    const int ncol = (int)((ncols - col0) < kChebCols ? (ncols - col0) : kChebCols);
    for (int idx = threadIdx.x; idx < ncol * n; idx += blockDim.x) cs[idx] = c[col0 * n + idx];
    __syncthreads();
    const bool odd_n = (n_even != n_odd);
    for (int idx = threadIdx.x; idx < ncol * n; idx += blockDim.x) {
        const int cc = idx / n, r = idx % n;
        const double *cv = cs + cc * n;
        // node i of the half grid this output row mirrors, and the sign of the odd part
        const bool upper = r >= n_x;
        const int i = upper ? n - 1 - r : r;
        double ev = 0.0, od = 0.0;
        for (int k = 0; k < n_even; ++k) ev = __builtin_fma(ce[(long)k * n_x + i], cv[2 * k], ev);
        for (int k = 0; k < n_odd; ++k) od = __builtin_fma(co[(long)k * n_x + i], cv[2 * k + 1], od);
        double v;
        if (odd_n && i == n_x - 1) {
            // the x = 0 row: halved in the tables, doubled back; the reference adds the odd product twice
            v = ev * 2.0 + od * 2.0;
        } else {
            v = upper ? ev - od : ev + od;
        }
        f[(col0 + cc) * n + r] = v - cv[0] / 2.0;
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

}  // namespace
}  // namespace raysht

using namespace raysht;

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
    if (grid) {
        // Initialize_Cheby_Grid (:161-173): affine map of [x_n, x_1] onto [rmin, rmax]
        const double xmin = x[n - 1], xmax = x[0];
        const double scaling = (rmax - rmin) / (xmax - xmin);
        for (int i = 0; i < n; ++i) grid[i] = (x[i] - xmin) * scaling + rmin;
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
    if (hipMalloc(&cb->d_ceven, cb->ceven.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&cb->d_codd, cb->codd.size() * sizeof(double)) != hipSuccess) {
        raysht_cheb_destroy(cb);
        return RAYSHT_ENOMEM;
    }
    if (hipMemcpy(cb->d_ceven, cb->ceven.data(), cb->ceven.size() * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(cb->d_codd, cb->codd.data(), cb->codd.size() * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess) {
        raysht_cheb_destroy(cb);
        return RAYSHT_EHIP;
    }
    *out = cb;
    return RAYSHT_OK;
}

void raysht_cheb_destroy(raysht_cheb *cb)
{
    if (!cb) return;
    if (cb->d_ceven) (void)hipFree(cb->d_ceven);
    if (cb->d_codd) (void)hipFree(cb->d_codd);
    delete cb;
}

int raysht_cheb_table(const raysht_cheb *cb, int which, double *host_out)
{
    if (!cb || !host_out || which < 0 || which > 1) return RAYSHT_EINVAL;
    const std::vector<double> &t = which == 0 ? cb->ceven : cb->codd;
    const size_t n = which == 0 ? (size_t)cb->n_x * cb->n_even : (size_t)cb->n_x * cb->n_odd;
    for (size_t i = 0; i < n; ++i) host_out[i] = t[i];
    return RAYSHT_OK;
}

int raysht_cheb_to_spectral(raysht_cheb *cb, const double *f, double *c, long ncols, void *stream)
{
    if (!cb || ncols < 0 || (ncols > 0 && (!f || !c))) return RAYSHT_EINVAL;
    if (ncols == 0) return RAYSHT_OK;
    const long blocks = (ncols + kChebCols - 1) / kChebCols;
    if (blocks > 0x7fffffffL) return RAYSHT_EINVAL;
    const size_t lds = sizeof(double) * 2 * kChebCols * cb->n_x;
    hipLaunchKernelGGL(cheb_to_spectral_kernel, dim3((unsigned)blocks), dim3(kChebThreads), lds,
                       reinterpret_cast<hipStream_t>(stream), cb->n, cb->n_x, cb->n_even, cb->n_odd, ncols,
                       cb->d_ceven, cb->d_codd, f, c);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int raysht_cheb_from_spectral(raysht_cheb *cb, const double *c, double *f, long ncols, void *stream)
{
    if (!cb || ncols < 0 || (ncols > 0 && (!f || !c))) return RAYSHT_EINVAL;
    if (ncols == 0) return RAYSHT_OK;
    const long blocks = (ncols + kChebCols - 1) / kChebCols;
    if (blocks > 0x7fffffffL) return RAYSHT_EINVAL;
    const size_t lds = sizeof(double) * kChebCols * cb->n;
    hipLaunchKernelGGL(cheb_from_spectral_kernel, dim3((unsigned)blocks), dim3(kChebThreads), lds,
                       reinterpret_cast<hipStream_t>(stream), cb->n, cb->n_x, cb->n_even, cb->n_odd, ncols,
                       cb->d_ceven, cb->d_codd, c, f);
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
