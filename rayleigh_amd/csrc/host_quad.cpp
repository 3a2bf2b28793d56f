// Host-side quad-precision setup for the transform plans.
//
// Rayleigh builds its Gauss-Legendre grid and the l = m seeds of its P_lm
// recursion in Real*16 and stores doubles (Legendre_Polynomials.F90:146-227,
// 332-415).  This file does the O(N_theta^2 + l_max) part of that work in
// __float128 with the same steps, and hands the results to the device as
// head/tail double pairs; the O(N_theta * N_lm) recursion itself runs on the
// GPU in double-double (legendre_tables.hip).
#include <quadmath.h>

#include <cstring>
#include <vector>

namespace raysht {

namespace {

using quad = __float128;

const quad kPi = 3.1415926535897932384626433832795028841972Q;

void to_pair(quad q, double *hi, double *lo)
{
    const double h = static_cast<double>(q);
    *hi = h;
    *lo = static_cast<double>(q - static_cast<quad>(h));
}

// P_n(x) and P_n'(x) by the three-term recursion (Legendre_Polynomials.F90:375-393).
void legendre_pn(quad x, int n, quad &pn, quad &dpn)
{
    quad cur = 1.0Q, prev = 0.0Q;
    for (int j = 1; j <= n; ++j) {
        const quad before = prev;
        prev = cur;
        cur = ((2.0Q * static_cast<quad>(j) - 1.0Q) * x * prev - (static_cast<quad>(j) - 1.0Q) * before) /
              static_cast<quad>(j);
    }
    pn = cur;
    dpn = static_cast<quad>(n) * (x * cur - prev) / (x * x - 1.0Q);
}

}  // namespace

// Nodes ascend from -1 (Legendre_Polynomials.F90:332-373).  Each root is found
// by Newton's method from cos(pi (i - 1/4) / (n + 1/2)) until the step is at
// most 3e-14; the weight is formed from the derivative at the iterate before
// the last step, which is what the reference stores.
static void gauss_legendre_q(int n, quad *x, quad *w)
{
    const int half = (n + 1) / 2;
    for (int i = 1; i <= half; ++i) {
        quad root = cosq(kPi * (static_cast<quad>(i) - 0.25Q) / (static_cast<quad>(n) + 0.5Q));
        quad pn = 0, dpn = 0;
        for (;;) {
            legendre_pn(root, n, pn, dpn);
            const quad next = root - pn / dpn;
            const quad step = fabsq(root - next);
            root = next;
            if (step <= 3.0e-14Q) break;
        }
        x[i - 1] = -root;
        x[n - i] = root;
        w[i - 1] = 2.0Q / ((1.0Q - root * root) * dpn * dpn);
        w[n - i] = w[i - 1];
    }
}

void gauss_legendre_dd(int n, double *x_hi, double *x_lo, double *w_hi, double *w_lo,
                       double *renorm_hi, double *renorm_lo)
{
    std::vector<quad> x(n), w(n);
    gauss_legendre_q(n, x.data(), w.data());
    for (int i = 0; i < n; ++i) {
        to_pair(x[i], &x_hi[i], &x_lo[i]);
        to_pair(w[i], &w_hi[i], &w_lo[i]);
        to_pair(2.0Q * kPi * w[i], &renorm_hi[i], &renorm_lo[i]);
    }
}

// The same nodes and weights as the reference stores them (Real*16 coloc, gl_weights,
// Legendre_Polynomials.F90:28): IEEE binary128, 16 bytes each, host byte order.
void gauss_legendre_binary128(int n, void *x16, void *w16)
{
    std::vector<quad> x(n), w(n);
    gauss_legendre_q(n, x.data(), w.data());
    std::memcpy(x16, x.data(), sizeof(quad) * n);
    std::memcpy(w16, w.data(), sizeof(quad) * n);
}

// amp(m) = sqrt((m + 1/2) / (2 pi)) * prod_{i<=m} sqrt((i - 1/2) / i)
// (Legendre_Polynomials.F90:193-197, 406-415).  The product is accumulated
// once over m; for every m it is the same sequence of quad multiplications the
// reference repeats from i = 1.
void plm_amplitudes_dd(int lmax, double *amp_hi, double *amp_lo)
{
    quad ratio = 1.0Q;
    for (int m = 0; m <= lmax; ++m) {
        if (m > 0) ratio = ratio * powq((static_cast<quad>(m) - 0.5Q) / static_cast<quad>(m), 0.5Q);
        quad amp = powq((static_cast<quad>(m) + 0.5Q) / (2.0Q * kPi), 0.5Q);
        amp = amp * ratio;
        to_pair(amp, &amp_hi[m], &amp_lo[m]);
    }
}

// cos and sin of 2 pi k / n, k = 0..n-1, correctly rounded from quad.
void fft_twiddles(int n, double *cos_out, double *sin_out)
{
    for (int k = 0; k < n; ++k) {
        const quad a = 2.0Q * kPi * static_cast<quad>(k) / static_cast<quad>(n);
        cos_out[k] = static_cast<double>(cosq(a));
        sin_out[k] = static_cast<double>(sinq(a));
        // quarter turns exactly
        if ((4L * k) % n == 0) {
            const long q = (4L * k) / n;
            cos_out[k] = q == 0 ? 1.0 : q == 2 ? -1.0 : 0.0;
            sin_out[k] = q == 1 ? 1.0 : q == 3 ? -1.0 : 0.0;
        }
    }
}

// The DCT form of the Chebyshev transforms: e^{-2 pi i k / n} (quarter turns exact),
// e^{-i pi j / (2 n)}, and the distance of each accumulated zeros-grid angle theta[i] (the
// doubles Gen_Colocation_Points produces by repeated addition) from (i + 1/2) pi / n.
void cheb_dct_tables(int n, const double *theta, double *tw_cos, double *tw_sin, double *q_cos, double *q_sin,
                     double *delta)
{
    fft_twiddles(n, tw_cos, tw_sin);
    for (int k = 0; k < n; ++k) tw_sin[k] = -tw_sin[k];
    for (int j = 0; j < n; ++j) {
        const quad b = kPi * static_cast<quad>(j) / static_cast<quad>(2 * n);
        q_cos[j] = static_cast<double>(cosq(b));
        q_sin[j] = -static_cast<double>(sinq(b));
    }
    for (int i = 0; i < n; ++i)
        delta[i] = static_cast<double>(static_cast<quad>(theta[i]) -
                                       (static_cast<quad>(i) + 0.5Q) * kPi / static_cast<quad>(n));
}

}  // namespace raysht
