// Double-double P_lm column recursion shared by the device table builder and
// its host-side unit test.
//
// One call evaluates P_lm(x) for l = m..lmax at one node x, in double-double
// (about 106 significant bits) with a separate power-of-two exponent, and hands
// each value to `store(l, value, exponent)`.  The steps follow Rayleigh's
// Compute_Plms (Legendre_Polynomials.F90:146-227): P_mm in closed form
// (-1)^m amp(m) (1-x^2)^(m/2), P_m+1,m = P_mm x sqrt(2m+3), then
// P_l = (P_l-1 x - a_l P_l-2) b_l with a_l = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)) and
// b_l = sqrt((4l^2-1)/(l^2-m^2)).
//
// Every product and sum must be rounded separately for the error-free
// transformations to hold: translation units including this header are built
// with floating-point contraction off, and fused multiply-adds appear only
// where written explicitly.
#pragma once

#pragma clang fp contract(off)

namespace raysht {
namespace plm {

struct ddv {
    double hi, lo;
};

__host__ __device__ inline ddv two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__host__ __device__ inline ddv fast_two_sum(double a, double b)
{
    const double s = a + b;
    return {s, b - (s - a)};
}

__host__ __device__ inline ddv dd_add(ddv a, ddv b)
{
    ddv s = two_sum(a.hi, b.hi);
    ddv t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}

__host__ __device__ inline ddv dd_neg(ddv a) { return {-a.hi, -a.lo}; }

__host__ __device__ inline ddv dd_mul(ddv a, ddv b)
{
    const double p = a.hi * b.hi;
    double e = __builtin_fma(a.hi, b.hi, -p);
    e = __builtin_fma(a.hi, b.lo, e);
    e = __builtin_fma(a.lo, b.hi, e);
    return fast_two_sum(p, e);
}

__host__ __device__ inline ddv dd_mul_d(ddv a, double b)
{
    const double p = a.hi * b;
    double e = __builtin_fma(a.hi, b, -p);
    e = __builtin_fma(a.lo, b, e);
    return fast_two_sum(p, e);
}

__host__ __device__ inline ddv dd_div(ddv a, ddv b)
{
    const double q1 = a.hi / b.hi;
    ddv r = dd_add(a, dd_neg(dd_mul_d(b, q1)));
    const double q2 = r.hi / b.hi;
    r = dd_add(r, dd_neg(dd_mul_d(b, q2)));
    const double q3 = r.hi / b.hi;
    ddv q = fast_two_sum(q1, q2);
    return dd_add(q, {q3, 0.0});
}

__host__ __device__ inline ddv dd_sqrt(ddv a)
{
    if (a.hi <= 0.0) return {0.0, 0.0};
    const double x = __builtin_sqrt(a.hi);
    // one Newton correction in double-double: x + (a - x^2) / (2x)
    const double p = x * x;
    const double pe = __builtin_fma(x, x, -p);
    ddv r = dd_add(a, {-p, -pe});
    return fast_two_sum(x, r.hi / (2.0 * x));
}

// floor(log2|x|) for finite nonzero x
__host__ __device__ inline int exponent_of(double x)
{
    int e = 0;
    (void)__builtin_frexp(x, &e);
    return e - 1;
}

__host__ __device__ inline ddv dd_ldexp(ddv a, int k) { return {__builtin_ldexp(a.hi, k), __builtin_ldexp(a.lo, k)}; }

// value = (m.hi + m.lo) * 2^e
struct xdd {
    ddv m;
    int e;
};

__host__ __device__ inline xdd xdd_norm(ddv a, int e)
{
    if (a.hi == 0.0) return {{0.0, 0.0}, 0};
    const int k = exponent_of(a.hi);
    return {dd_ldexp(a, -k), e + k};
}

__host__ __device__ inline xdd xdd_mul(xdd a, xdd b) { return xdd_norm(dd_mul(a.m, b.m), a.e + b.e); }

// Rounds (m.hi + m.lo) 2^e to the nearest double (flushing below the subnormal range to zero).
__host__ __device__ inline double xdd_to_double(ddv m, int e)
{
    if (m.hi == 0.0) return 0.0;
    if (e < -2200) return 0.0;
    if (e > 1100) return m.hi > 0 ? __builtin_inf() : -__builtin_inf();
    // m.hi already is m.hi + m.lo rounded to 53 bits; the scaling is exact while normal
    return __builtin_ldexp(m.hi, e);
}

// The stored doubles of one (l, m, node): P * STP_norm and 2 pi w P * PTS_norm
// (Parity_Resort, Legendre_Polynomials.F90:268-323), each rounded once.
struct plm_values {
    double p, ip;
};

__host__ __device__ inline plm_values plm_store_values(ddv v, int e, ddv renorm, double stp_norm, double pts_norm)
{
    plm_values r;
    r.p = xdd_to_double(dd_mul_d(v, stp_norm), e);
    ddv t = dd_mul(v, renorm);
    t = dd_mul_d(t, pts_norm);
    r.ip = xdd_to_double(t, e);
    return r;
}

template <class Store>
__host__ __device__ inline void plm_column(int m, int lmax, ddv x, ddv amp, Store &&store)
{
    // P_mm by repeated squaring on an extended-exponent double-double
    const ddv one_m_x2 = dd_add({1.0, 0.0}, dd_neg(dd_mul(x, x)));
    xdd base = xdd_norm(one_m_x2, 0);
    xdd pw = {{1.0, 0.0}, 0};
    for (int q = m / 2; q > 0; q >>= 1) {
        if (q & 1) pw = xdd_mul(pw, base);
        if (q > 1) base = xdd_mul(base, base);
    }
    if (m & 1) pw = xdd_norm(dd_mul(pw.m, dd_sqrt(one_m_x2)), pw.e);
    xdd pmm = xdd_norm(dd_mul(pw.m, amp), pw.e);
    if (m & 1) pmm.m = dd_neg(pmm.m);

    // the recursion state shares one exponent: (p1, p2) = (P_l-1, P_l-2) * 2^-e
    int e = pmm.e;
    ddv p2v = {0.0, 0.0};
    ddv p1v = pmm.m;
    store(m, p1v, e);
    if (m < lmax) {
        const ddv c = dd_sqrt({2.0 * (double)m + 3.0, 0.0});
        ddv pv = dd_mul(dd_mul(p1v, x), c);
        p2v = p1v;
        p1v = pv;
        store(m + 1, p1v, e);
    }
    const double mm = (double)m * (double)m;
    for (int l = m + 2; l <= lmax; ++l) {
        const double lm1 = (double)(l - 1);
        ddv a = dd_div({lm1 * lm1 - mm, 0.0}, {4.0 * lm1 * lm1 - 1.0, 0.0});
        a = dd_sqrt(a);
        const double ld = (double)l;
        ddv b = dd_div({4.0 * ld * ld - 1.0, 0.0}, {ld * ld - mm, 0.0});
        b = dd_sqrt(b);
        ddv t = dd_add(dd_mul(p1v, x), dd_neg(dd_mul(a, p2v)));
        ddv pv = dd_mul(t, b);
        p2v = p1v;
        p1v = pv;
        store(l, p1v, e);
        // keep the pair inside a comfortable exponent window
        if (p1v.hi != 0.0) {
            const int k = exponent_of(p1v.hi);
            if (k > 256 || k < -256) {
                p1v = dd_ldexp(p1v, -k);
                p2v = dd_ldexp(p2v, -k);
                e += k;
            }
        }
    }
}

}  // namespace plm
}  // namespace raysht
