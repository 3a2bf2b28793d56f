// Test driver: runs the library's double-double P_lm column recursion
// (rayleigh_amd/csrc/plm_recursion.h) on the CPU, seeded by the library's own
// quad-precision host setup (rayleigh_amd/csrc/host_quad.cpp), and writes the
// four parity tables in the device layout ([k][j], node fastest).  Built by
// tests/test_plm_recursion_cpu.py with g++ -ffp-contract=off.
#define __host__
#define __device__
#include <vector>

#include "plm_recursion.h"

namespace raysht {
void gauss_legendre_dd(int n, double *x_hi, double *x_lo, double *w_hi, double *w_lo, double *renorm_hi,
                       double *renorm_lo);
void plm_amplitudes_dd(int lmax, double *amp_hi, double *amp_lo);
}  // namespace raysht

extern "C" int plm_host_tables(int nt, int lmax, const int *mvals, int nm, double *pe, double *po, double *ie,
                               double *io)
{
    using namespace raysht;
    using namespace raysht::plm;
    const int nt2 = nt / 2;
    std::vector<double> xh(nt), xl(nt), wh(nt), wl(nt), rh(nt), rl(nt), ah(lmax + 1), al(lmax + 1);
    gauss_legendre_dd(nt, xh.data(), xl.data(), wh.data(), wl.data(), rh.data(), rl.data());
    plm_amplitudes_dd(lmax, ah.data(), al.data());
    long oe = 0, oo = 0;
    for (int mp = 0; mp < nm; ++mp) {
        const int m = mvals[mp];
        const double pts_norm = (m == 0) ? 1.0 / (2.0 * (double)nt) : 1.0 / (double)nt;
        const double stp_norm = (m == 0) ? 1.0 : 0.5;
        for (int j = 0; j < nt2; ++j) {
            const ddv x = {xh[j], xl[j]}, rn = {rh[j], rl[j]}, amp = {ah[m], al[m]};
            plm_column(m, lmax, x, amp, [&](int l, ddv v, int e) {
                const plm_values r = plm_store_values(v, e, rn, stp_norm, pts_norm);
                const long k = (l - m) >> 1;
                if (((l - m) & 1) == 0) {
                    pe[oe + k * nt2 + j] = r.p;
                    ie[oe + k * nt2 + j] = r.ip;
                } else {
                    po[oo + k * nt2 + j] = r.p;
                    io[oo + k * nt2 + j] = r.ip;
                }
            });
        }
        oe += (long)((lmax - m + 2) / 2) * nt2;
        oo += (long)((lmax - m + 1) / 2) * nt2;
    }
    return 0;
}
