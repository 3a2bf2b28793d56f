// Device construction of the parity-split P_lm tables.
//
// Rayleigh evaluates the orthonormal associated Legendre functions at the
// southern Gauss nodes in Real*16 and stores doubles (Legendre_Polynomials.F90:
// 146-331).  Here one thread owns one (m, node) column and runs the recursion of
// plm_recursion.h in double-double with a separate exponent, so P_mm ~ sin^m(theta),
// which underflows double for m beyond ~110 at the polar nodes, is carried until
// the recursion has grown it back into range.  Each stored value is rounded to
// double once from a value accurate far beyond double precision, i.e. it is the
// reference's quad-then-round double.
#include "raysht_internal.h"
#include "plm_recursion.h"

namespace raysht {
namespace {

__global__ __launch_bounds__(128) void plm_tables_kernel(
    int lmax, int nt, int nt2, const int *__restrict__ mvals, const long *__restrict__ te_off,
    const long *__restrict__ to_off, const double *__restrict__ x_dd, const double *__restrict__ renorm_dd,
    const double *__restrict__ amp_dd, double *__restrict__ pt_even, double *__restrict__ pt_odd,
    double *__restrict__ ip_even, double *__restrict__ ip_odd)
{
    using namespace plm;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int mp = blockIdx.y;
    if (j >= nt2) return;
    const int m = mvals[mp];
    const ddv x = {x_dd[2 * j], x_dd[2 * j + 1]};
    const ddv renorm = {renorm_dd[2 * j], renorm_dd[2 * j + 1]};
    const ddv amp = {amp_dd[2 * m], amp_dd[2 * m + 1]};
    // Parity_Resort normalisations (Legendre_Polynomials.F90:268-276); the PtS factor
    // is the double 1/(2 nt) or 1/nt widened to quad, exactly representable here.
    const double pts_norm = (m == 0) ? 1.0 / (2.0 * (double)nt) : 1.0 / (double)nt;
    const double stp_norm = (m == 0) ? 1.0 : 0.5;
    const long oe = te_off[mp], oo = to_off[mp];
    plm_column(m, lmax, x, amp, [&](int l, ddv v, int e) {
        const plm_values r = plm_store_values(v, e, renorm, stp_norm, pts_norm);
        const long k = (l - m) >> 1;
        if (((l - m) & 1) == 0) {
            pt_even[oe + k * nt2 + j] = r.p;
            ip_even[oe + k * nt2 + j] = r.ip;
        } else {
            pt_odd[oo + k * nt2 + j] = r.p;
            ip_odd[oo + k * nt2 + j] = r.ip;
        }
    });
}

}  // namespace

int launch_plm_tables(raysht_plan *p, const double *d_x, const double *d_renorm, const double *d_amp,
                      hipStream_t st)
{
    dim3 block(128);
    dim3 grid((p->nt2 + 127) / 128, p->nm);
    hipLaunchKernelGGL(plm_tables_kernel, grid, block, 0, st, p->lmax, p->nt, p->nt2, p->d_mvals, p->d_te_off,
                       p->d_to_off, d_x, d_renorm, d_amp, p->d_pt_even, p->d_pt_odd, p->d_ip_even, p->d_ip_odd);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

}  // namespace raysht
