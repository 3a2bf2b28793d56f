// Register-resident DFTs of length 2, 4, 8 and 16 on complex doubles, fully unrolled:
// every index and twiddle is a compile-time constant once the loops are unrolled, so a
// length-16 transform is straight-line FMA code on the thread's registers.  Sign SG = -1
// is the forward transform X_k = sum_m x_m e^{-2 pi i m k / R}; SG = +1 the unnormalised
// inverse.  Host-callable too, so the same code is checked on the CPU.
#pragma once

namespace raysht {

struct dcplx {
    double x, y;
};

__host__ __device__ __forceinline__ dcplx operator+(dcplx a, dcplx b) { return dcplx{a.x + b.x, a.y + b.y}; }
__host__ __device__ __forceinline__ dcplx operator-(dcplx a, dcplx b) { return dcplx{a.x - b.x, a.y - b.y}; }
__host__ __device__ __forceinline__ dcplx cmul(dcplx a, double c, double s)
{
    return dcplx{a.x * c - a.y * s, a.x * s + a.y * c};
}

// cos and sin of 2 pi e / 16
__host__ __device__ constexpr double cos16(int e)
{
    constexpr double c1 = 0.92387953251128675612818318939678828682;  // cos(pi/8)
    constexpr double c2 = 0.70710678118654752440084436210484903928;  // cos(pi/4)
    constexpr double c3 = 0.38268343236508977172845998403039886676;  // sin(pi/8)
    switch (e & 15) {
    case 0: return 1.0;
    case 1: return c1;
    case 2: return c2;
    case 3: return c3;
    case 4: return 0.0;
    case 5: return -c3;
    case 6: return -c2;
    case 7: return -c1;
    case 8: return -1.0;
    case 9: return -c1;
    case 10: return -c2;
    case 11: return -c3;
    case 12: return 0.0;
    case 13: return c3;
    case 14: return c2;
    default: return c1;
    }
}
__host__ __device__ constexpr double sin16(int e) { return cos16(e - 4); }

// cos and sin of 2 pi e / 64
__host__ __device__ constexpr double cos64_first_quadrant(int r)  // r = 0..16
{
    constexpr double c[17] = {1.0,
                              0.99518472667219688624483695310947992158,
                              0.98078528040323044912618223613423903697,
                              0.95694033573220886493579788698026996948,
                              0.92387953251128675612818318939678828682,
                              0.88192126434835502971275686366038834951,
                              0.83146961230254523707878837761790575674,
                              0.77301045336273696081090660975846980097,
                              0.70710678118654752440084436210484903928,
                              0.63439328416364549821517161322549337068,
                              0.55557023301960222474283081394853287437,
                              0.47139673682599764855638762590525437766,
                              0.38268343236508977172845998403039886676,
                              0.29028467725446236763619237581739527469,
                              0.19509032201612826784828486847702224093,
                              0.098017140329560601994195563888641845861,
                              0.0};
    return c[r];
}
__host__ __device__ constexpr double cos64(int e)
{
    const int q = (e >> 4) & 3, r = e & 15;
    return q == 0 ? cos64_first_quadrant(r)
         : q == 1 ? -cos64_first_quadrant(16 - r)
         : q == 2 ? -cos64_first_quadrant(r)
                  : cos64_first_quadrant(16 - r);
}
__host__ __device__ constexpr double sin64(int e) { return cos64(e - 16); }

// a * e^{SG 2 pi i k / R} for R dividing 16; quarter turns exact
template <int SG>
__host__ __device__ __forceinline__ dcplx rot(dcplx a, int k, int R)
{
    const int e = (k * (16 / R)) & 15;
    if (e == 0) return a;
    if (e == 4) return SG > 0 ? dcplx{-a.y, a.x} : dcplx{a.y, -a.x};
    if (e == 8) return dcplx{-a.x, -a.y};
    if (e == 12) return SG > 0 ? dcplx{a.y, -a.x} : dcplx{-a.y, a.x};
    return cmul(a, cos16(e), SG * sin16(e));
}

// in-place DFT of x[0..R): radix-4 decimation in time down to lengths 4 and 2
template <int R, int SG>
__host__ __device__ __forceinline__ void small_dft(dcplx (&x)[R])
{
    static_assert(R == 1 || R == 2 || R == 4 || R == 8 || R == 16, "length 1, 2, 4, 8 or 16");
    if constexpr (R == 2) {
        const dcplx a = x[0], b = x[1];
        x[0] = a + b;
        x[1] = a - b;
    } else if constexpr (R == 4) {
        const dcplx t0 = x[0] + x[2], t1 = x[0] - x[2], t2 = x[1] + x[3], t3 = rot<SG>(x[1] - x[3], 1, 4);
        x[0] = t0 + t2;
        x[2] = t0 - t2;
        x[1] = t1 + t3;
        x[3] = t1 - t3;
    } else if constexpr (R >= 8) {
        // X[k + M q] = sum_p w_4^{p q} (w_R^{p k} Y_p[k]), Y_p = DFT_M of x[4 m + p]
        constexpr int M = R / 4;
        dcplx y[4][M];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
#pragma unroll
            for (int m = 0; m < M; ++m) y[p][m] = x[4 * m + p];
            small_dft<M, SG>(y[p]);
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
            dcplx z[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) z[p] = rot<SG>(y[p][k], p * k, R);
            small_dft<4, SG>(z);
#pragma unroll
            for (int q = 0; q < 4; ++q) x[k + M * q] = z[q];
        }
    }
}

}  // namespace raysht
