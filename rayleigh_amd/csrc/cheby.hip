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
