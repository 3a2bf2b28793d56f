// Angular spectral derivatives on the spectral (rmcontainer4d) layout: per m-block
// data(m:l_max, r, imi, field), l fastest.
//
// Spectral_Derivatives.F90: d_by_dtheta (sin(theta) d/dtheta) and d_by_sdtheta
// (1/sin(theta) d(sin^2(theta) .)/dtheta) couple l to l - 1 and l + 1 with the per-(m, l)
// coefficients of Initialize_Angular_Derivatives (:49-83); d_by_dphi multiplies by i m.
// All three are memory-bound stencils along the contiguous l axis of every (r, imi)
// column: the l-stencils run one wave per l-column with the neighbours passed between
// lanes, the coefficients come from a small per-plan table in L2.  Products and sums are
// rounded separately, as the reference's compiled loops round them, so the results are its
// doubles.  The same file holds the passes built from these derivatives, each one kernel
// whose waves own whole radii: rlm_spaceb's combination and its Adjust_Emf after PtS,
// rlm_spacea's velocity steps and Compute_BandCurlB before StP, the output iterations'
// Hybrid_Output_Initial / _Final into the output co-buffer, and the l(l+1) / radial
// scalings of the spectral fields.
#include <algorithm>
#include <cmath>
#include <vector>

#include "raysht_internal.h"

namespace raysht {
namespace {

constexpr int kAThreads = 256;
constexpr int kAMaxBlocksX = 4096;

// q = e / n for 0 <= e < 2^24 from a float estimate and one exact correction
__device__ __forceinline__ int div_small(int e, int n, float rn, int &rem)
{
    int q = (int)((float)e * rn);
    int i = e - q * n;
    if (i < 0) {
        --q;
        i += n;
    } else if (i >= n) {
        ++q;
        i -= n;
    }
    rem = i;
    return q;
}

// out(l) of one field from field fin; grid (column groups, m-block).  One wave per (r, imi)
// l-column, 64 values of l at a time: the l - 1 and l + 1 neighbours come from the
// adjacent lanes' loads (lanes 0 and 63 load theirs), so each value is loaded once.
// Block mp of a packed array of F fields starts at lm_base[mp] * 2 nr * F (lm_base = the
// l count of the blocks before it); field f of the block at f * nl * 2 nr from there.
__global__ __launch_bounds__(kAThreads) void dtheta_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                           const int *__restrict__ lm_base,
                                                           const double2 *__restrict__ coef, const double *__restrict__ in,
                                                           int nf_in, int fin, double *__restrict__ out, int nf_out,
                                                           int fout)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);
    if (k >= 2 * nr) return;  // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], ncol = 2L * nr, total = (long)nl * ncol;
    const double *ak = in + lb * ncol * nf_in + (long)fin * total + (long)k * nl;
    double *bk = out + lb * ncol * nf_out + (long)fout * total + (long)k * nl;
    const double2 *c = coef + lb;
    for (int base = 0; base < nl; base += 64) {
        const int i = base + lane;
        const bool inr = i < nl;
        double x = 0.0;
        double2 ci = make_double2(0.0, 0.0);
        if (inr) {
            x = ak[i];
            ci = c[i];
        }
        double xn = __shfl_down(x, 1), xp = __shfl_up(x, 1);
        if (lane == 63 && i + 1 < nl) xn = ak[i + 1];
        if (lane == 0 && i > 0) xp = ak[i - 1];
        if (!inr) continue;
        double v;
        if (nl == 1)
            v = 0.0;
        else if (i == 0)
            v = xn * ci.y;
        else if (i == nl - 1)
            v = xp * ci.x;
        else
            v = xp * ci.x + xn * ci.y;
        bk[i] = v;
    }
}

// out = i m in for one field: (re, im) -> (-(m im), m re); in place allowed
__global__ __launch_bounds__(kAThreads) void dphi_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                         const int *__restrict__ lm_base, const double *in, int nf_in,
                                                         int fin, double *out, int nf_out, int fout)
{
#pragma clang fp contract(off)
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], half = (long)nl * nr;
    const double *a = in + lb * 2L * nr * nf_in + (long)fin * 2 * half;
    double *b = out + lb * 2L * nr * nf_out + (long)fout * 2 * half;
    const double dm = (double)m;
    for (long e = blockIdx.x * (long)kAThreads + threadIdx.x; e < half; e += (long)gridDim.x * kAThreads) {
        const double re = a[e], im = a[half + e];
        b[e] = -(dm * im);
        b[half + e] = dm * re;
    }
}

// rlm_spaceb's hydro combination (Sphere_Hybrid_Space.F90:127-214), in place on fields
// W, P, Z of a buffer of nf fields, one pass:
//   W <- W / (l(l+1));  Z <- (d_by_dphi P - d_by_sdtheta Z) / (l(l+1));
//   P <- (d_by_sdtheta P + d_by_dphi Z) / (l(l+1));
// then l = 0 of W, P, Z zeroed for m = 0, and l = l_max of every field zeroed.
// One wave owns one radius of one m-block (its re and im l-columns of W, P, Z) and walks it
// 64 values of l at a time, so no other wave touches what it reads: the l + 1 neighbours
// come from the next lane (lane 63 loads the first value of the next chunk, not yet
// written), the l - 1 neighbours from the previous lane (lane 0 takes the last values of
// the previous chunk as loaded, carried in registers, since the wave has overwritten them).
__global__ __launch_bounds__(kAThreads) void rlm_spaceb_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                               const int *__restrict__ lm_base,
                                                               const double2 *__restrict__ scoef, double *buf, int nf,
                                                               int wv, int pv, int zv)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);
    if (r >= nr) return;  // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], half = (long)nl * nr, fsz = 2 * half;
    double *blk = buf + lb * 2L * nr * nf;
    double *W = blk + wv * fsz + (long)r * nl, *P = blk + pv * fsz + (long)r * nl, *Z = blk + zv * fsz + (long)r * nl;
    const double2 *c = scoef + lb;
    const double dm = (double)m;
    double cpr = 0.0, cpi = 0.0, czr = 0.0, czi = 0.0;  // l - 1 values for lane 0
    for (int base = 0; base < nl; base += 64) {
        const int i = base + lane, l = m + i;
        const bool in = i < nl;
        double wr = 0.0, wi = 0.0, pr = 0.0, pi = 0.0, zr = 0.0, zi = 0.0;
        double2 ci = make_double2(0.0, 0.0);
        if (in) {
            wr = W[i];
            wi = W[half + i];
            pr = P[i];
            pi = P[half + i];
            zr = Z[i];
            zi = Z[half + i];
            ci = c[i];
        }
        double prn = __shfl_down(pr, 1), pin = __shfl_down(pi, 1);  // l + 1
        double zrn = __shfl_down(zr, 1), zin = __shfl_down(zi, 1);
        if (lane == 63 && i + 1 < nl) {
            prn = P[i + 1];
            pin = P[half + i + 1];
            zrn = Z[i + 1];
            zin = Z[half + i + 1];
        }
        double prp = __shfl_up(pr, 1), pip = __shfl_up(pi, 1);  // l - 1
        double zrp = __shfl_up(zr, 1), zip = __shfl_up(zi, 1);
        if (lane == 0) {
            prp = cpr;
            pip = cpi;
            zrp = czr;
            zip = czi;
        }
        cpr = __shfl(pr, 63);
        cpi = __shfl(pi, 63);
        czr = __shfl(zr, 63);
        czi = __shfl(zi, 63);
        if (!in) continue;
        // d_by_sdtheta, rounded as dtheta_kernel rounds it
        double sdpr, sdpi, sdzr, sdzi;
        if (nl == 1) {
            sdpr = sdpi = sdzr = sdzi = 0.0;
        } else if (i == 0) {
            sdpr = prn * ci.y;
            sdpi = pin * ci.y;
            sdzr = zrn * ci.y;
            sdzi = zin * ci.y;
        } else if (i == nl - 1) {
            sdpr = prp * ci.x;
            sdpi = pip * ci.x;
            sdzr = zrp * ci.x;
            sdzi = zip * ci.x;
        } else {
            sdpr = prp * ci.x + prn * ci.y;
            sdpi = pip * ci.x + pin * ci.y;
            sdzr = zrp * ci.x + zrn * ci.y;
            sdzi = zip * ci.x + zin * ci.y;
        }
        const double ol = l == 0 ? 1.0 : 1.0 / ((double)l * ((double)l + 1.0));  // over_l_l_plus1
        const bool zero = l == lmax || l == 0;
        W[i] = zero ? 0.0 : wr * ol;
        W[half + i] = zero ? 0.0 : wi * ol;
        P[i] = zero ? 0.0 : (sdpr + -(dm * zi)) * ol;
        P[half + i] = zero ? 0.0 : (sdpi + dm * zr) * ol;
        Z[i] = zero ? 0.0 : (-(dm * pi) - sdzr) * ol;
        Z[half + i] = zero ? 0.0 : (dm * pr - sdzi) * ol;
        if (i == nl - 1)
            for (int f = 0; f < nf; ++f)
                if (f != wv && f != pv && f != zv) {
                    blk[f * fsz + (long)r * nl + i] = 0.0;
                    blk[f * fsz + half + (long)r * nl + i] = 0.0;
                }
    }
}

// l - 1 and l + 1 neighbours of a per-lane value (lane = position in the 64-value l chunk):
// from the adjacent lanes; lane 63 takes `next` (the first value of the next chunk, not yet
// overwritten), lane 0 `carry` (the last value of the previous chunk, kept from before the
// wave overwrote it); `carry` then becomes this chunk's last value.
struct lnb {
    double p, n;
};
__device__ __forceinline__ lnb l_neighbours(double x, double next, double &carry, int lane)
{
    // both shuffles by every lane: a lane permute reads nothing from a lane that is not
    // executing it, so the shuffle must not sit under the lane test
    lnb v;
    v.n = __shfl_down(x, 1);
    v.p = __shfl_up(x, 1);
    if (lane == 63) v.n = next;
    if (lane == 0) v.p = carry;
    carry = __shfl(x, 63);
    return v;
}

// d_by_dtheta at position i of an l-column of nl values, as dtheta_kernel rounds it
__device__ __forceinline__ double dth_at(lnb v, double2 c, int i, int nl)
{
#pragma clang fp contract(off)  // the pragma of the calling kernel does not reach an inlined callee
    if (nl == 1) return 0.0;
    if (i == 0) return v.n * c.y;
    if (i == nl - 1) return v.p * c.x;
    return v.p * c.x + v.n * c.y;
}

__device__ __forceinline__ double l_l_plus1(int l)
{
#pragma clang fp contract(off)
    const double ell = (double)l;
    return ell * (ell + 1.0);
}

// field roles of the rlm_spacea passes (0-based fields of one packed buffer)
struct spacea_hydro_fields {
    int w, dw, d2w, z, dz, t, vt, vp, dvt, dvp;
};
struct bandcurlb_fields {
    int c, a, dc, da, d2c, cbr;
};

// rlm_spacea's Velocity_Components and Velocity_Derivatives, then d_by_dtheta(tvar, dtdt)
// (Sphere_Hybrid_Space.F90:80,235-355), in place, one pass.  With L = l(l+1), D = d_by_dtheta,
// P = d_by_dphi ((re, im) -> (-(m im), m re)) and the radial factors of radius r:
//   vr = (L W) over_rhorsq;  vtheta = (D dWdr + P Z) over_rhor;  vphi = (P dWdr - D Z) over_rhor;
//   dvtdr = (D d2Wdr2 + P dZdr) over_rhor - vtheta drho_term;
//   dvpdr = (P d2Wdr2 - D dZdr) over_rhor - vphi drho_term;
//   dvrdr = (L dWdr) over_rhorsq - vr two_over_r - vr dlnrho [- vr dsdr_over_cp];
//   dvrdt = D vr;  Z = (L Z) over_rhorsq;  dtdt = D T;
// vr, dvrdr, dtdt and dvrdt replace W, dWdr, d2Wdr2 and dZdr.  Every product and sum is
// rounded where the reference's statements round it, so the results are its doubles.  One
// wave owns one radius of one m-block (both imi of every field it reads or writes), as in
// rlm_spaceb_kernel.
__global__ __launch_bounds__(kAThreads) void rlm_spacea_hydro_kernel(
    int lmax, int nr, const int *__restrict__ mvals, const int *__restrict__ lm_base, const double2 *__restrict__ dcoef,
    double *buf, int nf, spacea_hydro_fields f, const double *__restrict__ over_rhor,
    const double *__restrict__ over_rhorsq, const double *__restrict__ drho_term,
    const double *__restrict__ two_over_r, const double *__restrict__ dlnrho,
    const double *__restrict__ dsdr_over_cp)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);
    if (r >= nr) return;  // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], half = (long)nl * nr, fsz = 2 * half;
    double *blk = buf + lb * 2L * nr * nf + (long)r * nl;
    double *W = blk + f.w * fsz, *dW = blk + f.dw * fsz, *d2W = blk + f.d2w * fsz, *Z = blk + f.z * fsz;
    double *dZ = blk + f.dz * fsz, *T = blk + f.t * fsz;
    double *Vt = blk + f.vt * fsz, *Vp = blk + f.vp * fsz, *dVt = blk + f.dvt * fsz, *dVp = blk + f.dvp * fsz;
    const double2 *c = dcoef + lb;
    const double dm = (double)m, orr = over_rhor[r], orsq = over_rhorsq[r], drho = drho_term[r];
    const double tor = two_over_r[r], dl = dlnrho[r], ds = dsdr_over_cp ? dsdr_over_cp[r] : 0.0;
    // carries for lane 0: dWdr, Z, d2Wdr2, dZdr, T and vr (re, im)
    double k_dwr = 0.0, k_dwi = 0.0, k_zr = 0.0, k_zi = 0.0, k_d2r = 0.0, k_d2i = 0.0;
    double k_dzr = 0.0, k_dzi = 0.0, k_tr = 0.0, k_ti = 0.0, k_vrr = 0.0, k_vri = 0.0;
    for (int base = 0; base < nl; base += 64) {
        const int i = base + lane, l = m + i;
        const bool in = i < nl;
        double wr = 0.0, wi = 0.0, dwr = 0.0, dwi = 0.0, d2r = 0.0, d2i = 0.0;
        double zr = 0.0, zi = 0.0, dzr = 0.0, dzi = 0.0, tr = 0.0, ti = 0.0;
        double2 ci = make_double2(0.0, 0.0);
        if (in) {
            wr = W[i];
            wi = W[half + i];
            dwr = dW[i];
            dwi = dW[half + i];
            d2r = d2W[i];
            d2i = d2W[half + i];
            zr = Z[i];
            zi = Z[half + i];
            dzr = dZ[i];
            dzi = dZ[half + i];
            tr = T[i];
            ti = T[half + i];
            ci = c[i];
        }
        // the first values of the next chunk (lane 63 only)
        double n_wr = 0.0, n_wi = 0.0, n_dwr = 0.0, n_dwi = 0.0, n_d2r = 0.0, n_d2i = 0.0;
        double n_zr = 0.0, n_zi = 0.0, n_dzr = 0.0, n_dzi = 0.0, n_tr = 0.0, n_ti = 0.0;
        if (lane == 63 && i + 1 < nl) {
            n_wr = W[i + 1];
            n_wi = W[half + i + 1];
            n_dwr = dW[i + 1];
            n_dwi = dW[half + i + 1];
            n_d2r = d2W[i + 1];
            n_d2i = d2W[half + i + 1];
            n_zr = Z[i + 1];
            n_zi = Z[half + i + 1];
            n_dzr = dZ[i + 1];
            n_dzi = dZ[half + i + 1];
            n_tr = T[i + 1];
            n_ti = T[half + i + 1];
        }
        const double L = l_l_plus1(l);
        const double vrr = L * wr * orsq, vri = L * wi * orsq;
        const double Ln = l_l_plus1(l + 1);
        const lnb b_dwr = l_neighbours(dwr, n_dwr, k_dwr, lane), b_dwi = l_neighbours(dwi, n_dwi, k_dwi, lane);
        const lnb b_zr = l_neighbours(zr, n_zr, k_zr, lane), b_zi = l_neighbours(zi, n_zi, k_zi, lane);
        const lnb b_d2r = l_neighbours(d2r, n_d2r, k_d2r, lane), b_d2i = l_neighbours(d2i, n_d2i, k_d2i, lane);
        const lnb b_dzr = l_neighbours(dzr, n_dzr, k_dzr, lane), b_dzi = l_neighbours(dzi, n_dzi, k_dzi, lane);
        const lnb b_tr = l_neighbours(tr, n_tr, k_tr, lane), b_ti = l_neighbours(ti, n_ti, k_ti, lane);
        const lnb b_vrr = l_neighbours(vrr, Ln * n_wr * orsq, k_vrr, lane);
        const lnb b_vri = l_neighbours(vri, Ln * n_wi * orsq, k_vri, lane);
        if (!in) continue;
        const double vtr = (dth_at(b_dwr, ci, i, nl) + -(dm * zi)) * orr;
        const double vti = (dth_at(b_dwi, ci, i, nl) + dm * zr) * orr;
        const double vpr = (-(dm * dwi) - dth_at(b_zr, ci, i, nl)) * orr;
        const double vpi = (dm * dwr - dth_at(b_zi, ci, i, nl)) * orr;
        double dvtr = (dth_at(b_d2r, ci, i, nl) + -(dm * dzi)) * orr;
        double dvti = (dth_at(b_d2i, ci, i, nl) + dm * dzr) * orr;
        dvtr = dvtr - vtr * drho;
        dvti = dvti - vti * drho;
        double dvpr = (-(dm * d2i) - dth_at(b_dzr, ci, i, nl)) * orr;
        double dvpi = (dm * d2r - dth_at(b_dzi, ci, i, nl)) * orr;
        dvpr = dvpr - vpr * drho;
        dvpi = dvpi - vpi * drho;
        double dvrr = L * dwr * orsq, dvri = L * dwi * orsq;
        dvrr = dvrr - vrr * tor;
        dvri = dvri - vri * tor;
        dvrr = dvrr - vrr * dl;
        dvri = dvri - vri * dl;
        if (dsdr_over_cp) {
            dvrr = dvrr - vrr * ds;
            dvri = dvri - vri * ds;
        }
        W[i] = vrr;
        W[half + i] = vri;
        Vt[i] = vtr;
        Vt[half + i] = vti;
        Vp[i] = vpr;
        Vp[half + i] = vpi;
        dVt[i] = dvtr;
        dVt[half + i] = dvti;
        dVp[i] = dvpr;
        dVp[half + i] = dvpi;
        dW[i] = dvrr;
        dW[half + i] = dvri;
        dZ[i] = dth_at(b_vrr, ci, i, nl);
        dZ[half + i] = dth_at(b_vri, ci, i, nl);
        Z[i] = L * zr * orsq;
        Z[half + i] = L * zi * orsq;
        d2W[i] = dth_at(b_tr, ci, i, nl);
        d2W[half + i] = dth_at(b_ti, ci, i, nl);
    }
}

// Compute_BandCurlB (Sphere_Hybrid_Space.F90:377-460), in place, one pass.  With L, D, P as
// above and o = OneOverRSquared(r), e = d2Cdr2 - Br:
//   Br = (L C) o;  [curl B]_r = (L A) o;  curlbphi = P dAdr + D e;  curlbtheta = D dAdr - P e;
//   Btheta = P A + D dCdr;  Bphi = P dCdr - D A;
// Br, Btheta, Bphi, curlbtheta and curlbphi replace C, A, dCdr, d2Cdr2 and dAdr.  Rounded
// where the reference rounds; one wave per radius of one m-block.
__global__ __launch_bounds__(kAThreads) void bandcurlb_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                              const int *__restrict__ lm_base,
                                                              const double2 *__restrict__ dcoef, double *buf, int nf,
                                                              bandcurlb_fields f,
                                                              const double *__restrict__ one_over_rsq)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);
    if (r >= nr) return;  // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], half = (long)nl * nr, fsz = 2 * half;
    double *blk = buf + lb * 2L * nr * nf + (long)r * nl;
    double *C = blk + f.c * fsz, *A = blk + f.a * fsz, *dC = blk + f.dc * fsz, *dA = blk + f.da * fsz;
    double *d2C = blk + f.d2c * fsz, *Cbr = blk + f.cbr * fsz;
    const double2 *c = dcoef + lb;
    const double dm = (double)m, o = one_over_rsq[r];
    double k_dar = 0.0, k_dai = 0.0, k_er = 0.0, k_ei = 0.0, k_dcr = 0.0, k_dci = 0.0, k_ar = 0.0, k_ai = 0.0;
    for (int base = 0; base < nl; base += 64) {
        const int i = base + lane, l = m + i;
        const bool in = i < nl;
        double cr = 0.0, cim = 0.0, ar = 0.0, ai = 0.0, dcr = 0.0, dci = 0.0, dar = 0.0, dai = 0.0;
        double d2r = 0.0, d2i = 0.0;
        double2 ci = make_double2(0.0, 0.0);
        if (in) {
            cr = C[i];
            cim = C[half + i];
            ar = A[i];
            ai = A[half + i];
            dcr = dC[i];
            dci = dC[half + i];
            dar = dA[i];
            dai = dA[half + i];
            d2r = d2C[i];
            d2i = d2C[half + i];
            ci = c[i];
        }
        double n_cr = 0.0, n_ci = 0.0, n_ar = 0.0, n_ai = 0.0, n_dcr = 0.0, n_dci = 0.0, n_dar = 0.0, n_dai = 0.0;
        double n_d2r = 0.0, n_d2i = 0.0;
        if (lane == 63 && i + 1 < nl) {
            n_cr = C[i + 1];
            n_ci = C[half + i + 1];
            n_ar = A[i + 1];
            n_ai = A[half + i + 1];
            n_dcr = dC[i + 1];
            n_dci = dC[half + i + 1];
            n_dar = dA[i + 1];
            n_dai = dA[half + i + 1];
            n_d2r = d2C[i + 1];
            n_d2i = d2C[half + i + 1];
        }
        const double L = l_l_plus1(l), Ln = l_l_plus1(l + 1);
        const double brr = L * cr * o, bri = L * cim * o;
        const double er = d2r - brr, ei = d2i - bri;
        const lnb b_dar = l_neighbours(dar, n_dar, k_dar, lane), b_dai = l_neighbours(dai, n_dai, k_dai, lane);
        const lnb b_er = l_neighbours(er, n_d2r - Ln * n_cr * o, k_er, lane);
        const lnb b_ei = l_neighbours(ei, n_d2i - Ln * n_ci * o, k_ei, lane);
        const lnb b_dcr = l_neighbours(dcr, n_dcr, k_dcr, lane), b_dci = l_neighbours(dci, n_dci, k_dci, lane);
        const lnb b_ar = l_neighbours(ar, n_ar, k_ar, lane), b_ai = l_neighbours(ai, n_ai, k_ai, lane);
        if (!in) continue;
        C[i] = brr;
        C[half + i] = bri;
        Cbr[i] = L * ar * o;
        Cbr[half + i] = L * ai * o;
        dA[i] = -(dm * dai) + dth_at(b_er, ci, i, nl);
        dA[half + i] = dm * dar + dth_at(b_ei, ci, i, nl);
        d2C[i] = dth_at(b_dar, ci, i, nl) - -(dm * ei);
        d2C[half + i] = dth_at(b_dai, ci, i, nl) - dm * er;
        A[i] = -(dm * ai) + dth_at(b_dcr, ci, i, nl);
        A[half + i] = dm * ar + dth_at(b_dci, ci, i, nl);
        dC[i] = -(dm * dci) - dth_at(b_ar, ci, i, nl);
        dC[half + i] = dm * dcr - dth_at(b_ai, ci, i, nl);
    }
}

// rlm_spaceb's Adjust_Emf (Sphere_Hybrid_Space.F90:557-602), in place, one pass.  With S =
// d_by_sdtheta, P = d_by_dphi and ol = over_l_l_plus1(l):
//   emftheta <- (S emfphi - P emftheta) ol   (Cvar is the emftheta field);
//   emfphi   <- (P emfphi + S emftheta) ol   (emftheta as it came in).
// Rounded where the reference rounds; one wave per radius of one m-block.
__global__ __launch_bounds__(kAThreads) void adjust_emf_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                               const int *__restrict__ lm_base,
                                                               const double2 *__restrict__ scoef, double *buf, int nf,
                                                               int et, int ep)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);
    if (r >= nr) return;  // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], half = (long)nl * nr, fsz = 2 * half;
    double *blk = buf + lb * 2L * nr * nf + (long)r * nl;
    double *T = blk + et * fsz, *P = blk + ep * fsz;
    const double2 *c = scoef + lb;
    const double dm = (double)m;
    double k_tr = 0.0, k_ti = 0.0, k_pr = 0.0, k_pi = 0.0;
    for (int base = 0; base < nl; base += 64) {
        const int i = base + lane, l = m + i;
        const bool in = i < nl;
        double tr = 0.0, ti = 0.0, pr = 0.0, pi = 0.0;
        double2 ci = make_double2(0.0, 0.0);
        if (in) {
            tr = T[i];
            ti = T[half + i];
            pr = P[i];
            pi = P[half + i];
            ci = c[i];
        }
        double n_tr = 0.0, n_ti = 0.0, n_pr = 0.0, n_pi = 0.0;
        if (lane == 63 && i + 1 < nl) {
            n_tr = T[i + 1];
            n_ti = T[half + i + 1];
            n_pr = P[i + 1];
            n_pi = P[half + i + 1];
        }
        const lnb b_tr = l_neighbours(tr, n_tr, k_tr, lane), b_ti = l_neighbours(ti, n_ti, k_ti, lane);
        const lnb b_pr = l_neighbours(pr, n_pr, k_pr, lane), b_pi = l_neighbours(pi, n_pi, k_pi, lane);
        if (!in) continue;
        const double ol = l == 0 ? 1.0 : 1.0 / l_l_plus1(l);
        T[i] = (dth_at(b_pr, ci, i, nl) - -(dm * ti)) * ol;
        T[half + i] = (dth_at(b_pi, ci, i, nl) - dm * tr) * ol;
        P[i] = (-(dm * pi) + dth_at(b_tr, ci, i, nl)) * ol;
        P[half + i] = (dm * pr + dth_at(b_ti, ci, i, nl)) * ol;
    }
}

// Hybrid_Output_Initial, magnetic part (Sphere_Hybrid_Space.F90:525-541): from s into the
// output co-buffer and the two work fields (the reference's ftemp3, ftemp4), element by element:
// co[avar_cb] = (L s[avar]) one_over_rsq, co[dbrdr_cb] = s[dcdr], tmp0 = s[d2cdr2], tmp1 = s[dadr].
// One wave per (r, imi) l-column.
struct output_initial_fields {
    int avar, dcdr, d2cdr2, dadr, avar_cb, dbrdr_cb;
};
__global__ __launch_bounds__(kAThreads) void hybrid_output_initial_kernel(
    int lmax, int nr, const int *__restrict__ mvals, const int *__restrict__ lm_base, const double *__restrict__ s,
    int nf, double *__restrict__ co, int nf_co, double *__restrict__ tmp, output_initial_fields f,
    const double *__restrict__ one_over_rsq)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);  // column r + nr imi
    if (k >= 2 * nr) return;
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], fsz = 2L * nl * nr, col = (long)k * nl;
    const double *sb = s + lb * 2L * nr * nf + col;
    double *cb = co + lb * 2L * nr * nf_co + col, *t = tmp + lb * 2L * nr * 2 + col;
    const double o2 = one_over_rsq[k >= nr ? k - nr : k];
    for (int i = lane; i < nl; i += 64) {
        cb[f.avar_cb * fsz + i] = l_l_plus1(m + i) * sb[f.avar * fsz + i] * o2;
        cb[f.dbrdr_cb * fsz + i] = sb[f.dcdr * fsz + i];
        t[i] = sb[f.d2cdr2 * fsz + i];
        t[fsz + i] = sb[f.dadr * fsz + i];
    }
}

// Hybrid_Output_Final up to its transform (:543-552): co's l_max zeroed, Hydro_Output_Derivatives
// (:224-233) co[dpdt_cb] = D s[pvar], and with the work fields (magnetism) BField_Derivatives
// (:437-498), with D = d_by_dtheta, P = d_by_dphi and the radial factors of radius r:
//   co[dbtdr_cb] = (D tmp0 + P tmp1) one_over_r - s[btheta] one_over_rsq;
//   co[dbpdr_cb] = (P tmp0 - D tmp1) one_over_r - s[bphi] one_over_rsq;
//   co[dbrdr_cb] = (L co[dbrdr_cb]) one_over_rsq - s[br] two_over_r   (l_max of co[dbrdr_cb] read as 0);
//   co[dbrdt_cb] = D s[br].
// Rounded where the reference rounds; one wave per radius of one m-block (both imi), l
// neighbours as in rlm_spacea_hydro_kernel.
struct output_final_fields {
    int pvar, br, btheta, bphi, dpdt_cb, dbtdr_cb, dbpdr_cb, dbrdr_cb, dbrdt_cb;
};
// (MAG: tmp given; the hydro-only instance needs far fewer registers, so more waves stay resident)
template <bool MAG>
__global__ __launch_bounds__(kAThreads) void hybrid_output_final_kernel(
    int lmax, int nr, const int *__restrict__ mvals, const int *__restrict__ lm_base, const double2 *__restrict__ dcoef,
    const double *__restrict__ s, int nf, double *__restrict__ co, int nf_co, const double *__restrict__ tmp,
    output_final_fields f, const double *__restrict__ one_over_r, const double *__restrict__ one_over_rsq,
    const double *__restrict__ two_over_r)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);
    if (r >= nr) return;  // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], half = (long)nl * nr, fsz = 2 * half;
    const double *sb = s + lb * 2L * nr * nf + (long)r * nl;
    double *cb = co + lb * 2L * nr * nf_co + (long)r * nl;
    const double *P = sb + f.pvar * fsz, *Br = sb + f.br * fsz, *Bt = sb + f.btheta * fsz, *Bp = sb + f.bphi * fsz;
    constexpr bool mag = MAG;
    const double *T3 = mag ? tmp + lb * 2L * nr * 2 + (long)r * nl : nullptr, *T4 = mag ? T3 + fsz : nullptr;
    const double2 *c = dcoef + lb;
    const double dm = (double)m, orr = one_over_r[r], o2 = one_over_rsq[r], tor = two_over_r[r];
    double k_pr = 0.0, k_pi = 0.0, k_br = 0.0, k_bi = 0.0, k_3r = 0.0, k_3i = 0.0, k_4r = 0.0, k_4i = 0.0;
    for (int base = 0; base < nl; base += 64) {
        const int i = base + lane, l = m + i;
        const bool in = i < nl;
        double pr = 0.0, pi = 0.0, brr = 0.0, bri = 0.0, t3r = 0.0, t3i = 0.0, t4r = 0.0, t4i = 0.0;
        double2 ci = make_double2(0.0, 0.0);
        if (in) {
            pr = P[i];
            pi = P[half + i];
            ci = c[i];
            if (mag) {
                brr = Br[i];
                bri = Br[half + i];
                t3r = T3[i];
                t3i = T3[half + i];
                t4r = T4[i];
                t4i = T4[half + i];
            }
        }
        double n_pr = 0.0, n_pi = 0.0, n_br = 0.0, n_bi = 0.0, n_3r = 0.0, n_3i = 0.0, n_4r = 0.0, n_4i = 0.0;
        if (lane == 63 && i + 1 < nl) {
            n_pr = P[i + 1];
            n_pi = P[half + i + 1];
            if (mag) {
                n_br = Br[i + 1];
                n_bi = Br[half + i + 1];
                n_3r = T3[i + 1];
                n_3i = T3[half + i + 1];
                n_4r = T4[i + 1];
                n_4i = T4[half + i + 1];
            }
        }
        const lnb b_pr = l_neighbours(pr, n_pr, k_pr, lane), b_pi = l_neighbours(pi, n_pi, k_pi, lane);
        const lnb b_br = l_neighbours(brr, n_br, k_br, lane), b_bi = l_neighbours(bri, n_bi, k_bi, lane);
        const lnb b_3r = l_neighbours(t3r, n_3r, k_3r, lane), b_3i = l_neighbours(t3i, n_3i, k_3i, lane);
        const lnb b_4r = l_neighbours(t4r, n_4r, k_4r, lane), b_4i = l_neighbours(t4i, n_4i, k_4i, lane);
        if (!in) continue;
        if (i == nl - 1)  // co's l_max zeroed first; the fields computed below overwrite theirs
            for (int g = 0; g < nf_co; ++g) {
                cb[g * fsz + i] = 0.0;
                cb[g * fsz + half + i] = 0.0;
            }
        double *Dp = cb + f.dpdt_cb * fsz;
        Dp[i] = dth_at(b_pr, ci, i, nl);
        Dp[half + i] = dth_at(b_pi, ci, i, nl);
        if (!mag) continue;
        double *Dt = cb + f.dbtdr_cb * fsz, *Dq = cb + f.dbpdr_cb * fsz, *Dr = cb + f.dbrdr_cb * fsz;
        double *Dd = cb + f.dbrdt_cb * fsz;
        const double dr_r = l == lmax ? 0.0 : Dr[i], dr_i = l == lmax ? 0.0 : Dr[half + i];
        double v;
        v = (dth_at(b_3r, ci, i, nl) + -(dm * t4i)) * orr;
        Dt[i] = v - Bt[i] * o2;
        v = (dth_at(b_3i, ci, i, nl) + dm * t4r) * orr;
        Dt[half + i] = v - Bt[half + i] * o2;
        v = (-(dm * t3i) - dth_at(b_4r, ci, i, nl)) * orr;
        Dq[i] = v - Bp[i] * o2;
        v = (dm * t3r - dth_at(b_4i, ci, i, nl)) * orr;
        Dq[half + i] = v - Bp[half + i] * o2;
        const double L = l_l_plus1(l);
        v = L * dr_r * o2;
        Dr[i] = v - brr * tor;
        v = L * dr_i * o2;
        Dr[half + i] = v - bri * tor;
        Dd[i] = dth_at(b_br, ci, i, nl);
        Dd[half + i] = dth_at(b_bi, ci, i, nl);
    }
}

// out(l, r, imi) = (lfac(l) in(l, r, imi)) rfac(r) for one field (either factor may be
// absent); in place allowed.  The products round in this order, as the reference's
// l_l_plus1(m:l_max)*SBUFFA(IDX2, f)*Over_RhoRSQ(r) statements round them.  One wave per
// (r, imi) l-column: rfac(r) is one value per wave, lfac(m..l_max) a contiguous run.
__global__ __launch_bounds__(kAThreads) void lm_scale_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                             const int *__restrict__ lm_base, const double *in,
                                                             int nf_in, int fin, double *out, int nf_out, int fout,
                                                             const double *__restrict__ lfac,
                                                             const double *__restrict__ rfac)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * (kAThreads / 64) + (threadIdx.x >> 6);  // column r + nr imi
    if (k >= 2 * nr) return;                                            // whole waves
    const int mp = blockIdx.y, m = mvals[mp], nl = lmax - m + 1;
    const long lb = lm_base[mp], total = 2L * nl * nr;
    const double *a = in + lb * 2L * nr * nf_in + (long)fin * total + (long)k * nl;
    double *b = out + lb * 2L * nr * nf_out + (long)fout * total + (long)k * nl;
    const double rf = rfac ? rfac[k >= nr ? k - nr : k] : 1.0;
    const double *lf = lfac ? lfac + m : nullptr;
    for (int i = lane; i < nl; i += 64) {
        double v = a[i];
        if (lf) v = lf[i] * v;
        if (rfac) v = v * rf;
        b[i] = v;
    }
}

// data(l_max, r, imi, f) = 0 for every m-block, radius, imi and field: one value per
// l-column, at nl (k + 1) - 1 for column k = r + nr (imi + 2 f)
__global__ __launch_bounds__(kAThreads) void zero_lmax_kernel(int lmax, int nr, const int *__restrict__ mvals,
                                                              const int *__restrict__ lm_base, double *buf, int nf)
{
    const int mp = blockIdx.y, nl = lmax - mvals[mp] + 1;
    const long ncol = 2L * nr * nf;
    double *blk = buf + lm_base[mp] * ncol;
    for (long k = blockIdx.x * (long)kAThreads + threadIdx.x; k < ncol; k += (long)gridDim.x * kAThreads)
        blk[k * nl + nl - 1] = 0.0;
}

// the checks every spectral-field entry point shares: a plan with fields and m-blocks that
// fit one launch, and one block's field below 2^24 values (div_small)
int check_spectral(const raysht_plan *p)
{
    if (p->nr < 1 || p->nf < 1) return RAYSHT_EUNSUPPORTED;  // column plans have no fields
    if (p->nm > 65535) return RAYSHT_EUNSUPPORTED;
    if ((long)(p->lmax + 1) * 2 * p->nr >= (1L << 24)) return RAYSHT_EUNSUPPORTED;
    return RAYSHT_OK;
}

int max_nl(const raysht_plan *p)
{
    int n = 0;
    for (int m : p->mvals) n = std::max(n, p->lmax - m + 1);
    return n;
}

int launch_angular(raysht_plan *p, int kind, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                   hipStream_t st)
{
    if (!p || !in || !out || kind < 0 || kind > 2) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    if (nf_in < 1 || nf_out < 1 || fin < 0 || fin >= nf_in || fout < 0 || fout >= nf_out) return RAYSHT_EINVAL;
    const long per_field = p->s_off[p->nm] / p->nf;  // sum over m-blocks of nl * 2 nr
    const long n_in = per_field * nf_in, n_out = per_field * nf_out;
    const bool same = in == out && nf_in == nf_out;
    if (same && fin == fout && kind != 2) return RAYSHT_EINVAL;  // the l stencil cannot run in place
    if (!same && in < out + n_out && out < in + n_in) return RAYSHT_EINVAL;
    const long work = kind == 2 ? (long)max_nl(p) * p->nr : 2L * p->nr * 64;  // dtheta: a wave per l-column
    const unsigned bx = (unsigned)std::min<long>((work + kAThreads - 1) / kAThreads, kind == 2 ? kAMaxBlocksX : 1L << 30);
    dim3 grid(bx, (unsigned)p->nm);
    if (kind == 2)
        hipLaunchKernelGGL(dphi_kernel, grid, dim3(kAThreads), 0, st, p->lmax, p->nr, p->d_mvals, p->d_lm_base, in,
                           nf_in, fin, out, nf_out, fout);
    else
        hipLaunchKernelGGL(dtheta_kernel, grid, dim3(kAThreads), 0, st, p->lmax, p->nr, p->d_mvals, p->d_lm_base,
                           p->d_acoef + (kind == 1 ? p->lm_total : 0), in, nf_in, fin, out, nf_out, fout);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_rlm_spaceb(raysht_plan *p, double *buf, int nf, int wv, int pv, int zv, hipStream_t st)
{
    if (!p || !buf) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    if (nf < 3 || wv < 0 || pv < 0 || zv < 0 || wv >= nf || pv >= nf || zv >= nf || wv == pv || wv == zv || pv == zv)
        return RAYSHT_EINVAL;
    const unsigned bx = (unsigned)((p->nr + kAThreads / 64 - 1) / (kAThreads / 64));  // a wave per radius
    hipLaunchKernelGGL(rlm_spaceb_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, p->d_acoef + p->lm_total, buf, nf, wv, pv, zv);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

// every role a distinct field of the buffer
bool distinct_fields(const int *v, int n, int nf)
{
    for (int a = 0; a < n; ++a) {
        if (v[a] < 0 || v[a] >= nf) return false;
        for (int b = 0; b < a; ++b)
            if (v[a] == v[b]) return false;
    }
    return true;
}

int launch_spacea_hydro(raysht_plan *p, double *buf, int nf, const int *fields, const double *over_rhor,
                        const double *over_rhorsq, const double *drho_term, const double *two_over_r,
                        const double *dlnrho, const double *dsdr_over_cp, hipStream_t st)
{
    if (!p || !buf || !fields || !over_rhor || !over_rhorsq || !drho_term || !two_over_r || !dlnrho)
        return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    if (nf < 10 || !distinct_fields(fields, 10, nf)) return RAYSHT_EINVAL;
    const spacea_hydro_fields f{fields[0], fields[1], fields[2], fields[3], fields[4],
                                fields[5], fields[6], fields[7], fields[8], fields[9]};
    const unsigned bx = (unsigned)((p->nr + kAThreads / 64 - 1) / (kAThreads / 64));  // a wave per radius
    hipLaunchKernelGGL(rlm_spacea_hydro_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, p->d_acoef, buf, nf, f, over_rhor, over_rhorsq, drho_term,
                       two_over_r, dlnrho, dsdr_over_cp);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_bandcurlb(raysht_plan *p, double *buf, int nf, const int *fields, const double *one_over_rsq,
                     hipStream_t st)
{
    if (!p || !buf || !fields || !one_over_rsq) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    if (nf < 6 || !distinct_fields(fields, 6, nf)) return RAYSHT_EINVAL;
    const bandcurlb_fields f{fields[0], fields[1], fields[2], fields[3], fields[4], fields[5]};
    const unsigned bx = (unsigned)((p->nr + kAThreads / 64 - 1) / (kAThreads / 64));
    hipLaunchKernelGGL(bandcurlb_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, p->d_acoef, buf, nf, f, one_over_rsq);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_adjust_emf(raysht_plan *p, double *buf, int nf, int et, int ep, hipStream_t st)
{
    if (!p || !buf) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    const int v[2] = {et, ep};
    if (nf < 2 || !distinct_fields(v, 2, nf)) return RAYSHT_EINVAL;
    const unsigned bx = (unsigned)((p->nr + kAThreads / 64 - 1) / (kAThreads / 64));
    hipLaunchKernelGGL(adjust_emf_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, p->d_acoef + p->lm_total, buf, nf, et, ep);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

// s and co (and tmp) must not overlap: the passes read s while writing co
bool disjoint(const void *a, long na, const void *b, long nb)
{
    const char *x = static_cast<const char *>(a), *y = static_cast<const char *>(b);
    return x + na * 8 <= y || y + nb * 8 <= x;
}

int launch_output_initial(raysht_plan *p, const double *s, int nf, double *co, int nf_co, double *tmp,
                          const int *fields, const double *one_over_rsq, hipStream_t st)
{
    if (!p || !s || !co || !tmp || !fields || !one_over_rsq) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    const long per_field = p->s_off[p->nm] / p->nf;
    const int vs[4] = {fields[0], fields[1], fields[2], fields[3]}, vc[2] = {fields[4], fields[5]};
    if (nf < 4 || nf_co < 2 || !distinct_fields(vs, 4, nf) || !distinct_fields(vc, 2, nf_co)) return RAYSHT_EINVAL;
    if (!disjoint(s, per_field * nf, co, per_field * nf_co) || !disjoint(s, per_field * nf, tmp, per_field * 2) ||
        !disjoint(co, per_field * nf_co, tmp, per_field * 2))
        return RAYSHT_EINVAL;
    const output_initial_fields f{fields[0], fields[1], fields[2], fields[3], fields[4], fields[5]};
    const unsigned bx = (unsigned)((2L * p->nr + kAThreads / 64 - 1) / (kAThreads / 64));
    hipLaunchKernelGGL(hybrid_output_initial_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, s, nf, co, nf_co, tmp, f, one_over_rsq);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_output_final(raysht_plan *p, const double *s, int nf, double *co, int nf_co, const double *tmp,
                        const int *fields, const double *one_over_r, const double *one_over_rsq,
                        const double *two_over_r, hipStream_t st)
{
    if (!p || !s || !co || !fields) return RAYSHT_EINVAL;
    if (tmp && (!one_over_r || !one_over_rsq || !two_over_r)) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    const long per_field = p->s_off[p->nm] / p->nf;
    const int vs[4] = {fields[0], fields[1], fields[2], fields[3]};
    const int vc[5] = {fields[4], fields[5], fields[6], fields[7], fields[8]};
    if (nf < 1 || nf_co < 1) return RAYSHT_EINVAL;
    if (tmp ? (!distinct_fields(vs, 4, nf) || !distinct_fields(vc, 5, nf_co))
            : (vs[0] < 0 || vs[0] >= nf || vc[0] < 0 || vc[0] >= nf_co))
        return RAYSHT_EINVAL;
    if (!disjoint(s, per_field * nf, co, per_field * nf_co) || (tmp && !disjoint(co, per_field * nf_co, tmp, per_field * 2)))
        return RAYSHT_EINVAL;
    // the hydro-only form reads and writes one field each: the other indices are unused
    const output_final_fields f{fields[0], tmp ? fields[1] : fields[0], tmp ? fields[2] : fields[0],
                                tmp ? fields[3] : fields[0], fields[4], tmp ? fields[5] : fields[4],
                                tmp ? fields[6] : fields[4], tmp ? fields[7] : fields[4], tmp ? fields[8] : fields[4]};
    const unsigned bx = (unsigned)((p->nr + kAThreads / 64 - 1) / (kAThreads / 64));
    // the hydro-only instance reads no profile: s stands in for them
    hipLaunchKernelGGL(tmp ? hybrid_output_final_kernel<true> : hybrid_output_final_kernel<false>,
                       dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr, p->d_mvals, p->d_lm_base,
                       p->d_acoef, s, nf, co, nf_co, tmp, f, tmp ? one_over_r : s, tmp ? one_over_rsq : s,
                       tmp ? two_over_r : s);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_lm_scale(raysht_plan *p, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                    const double *lfac, const double *rfac, hipStream_t st)
{
    if (!p || !in || !out) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    if (nf_in < 1 || nf_out < 1 || fin < 0 || fin >= nf_in || fout < 0 || fout >= nf_out) return RAYSHT_EINVAL;
    const long per_field = p->s_off[p->nm] / p->nf;  // elementwise: in place, or another field of one buffer
    const bool same = in == out && nf_in == nf_out;
    if (!same && in < out + per_field * nf_out && out < in + per_field * nf_in) return RAYSHT_EINVAL;
    const unsigned bx = (unsigned)((2L * p->nr + kAThreads / 64 - 1) / (kAThreads / 64));  // a wave per l-column
    hipLaunchKernelGGL(lm_scale_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, in, nf_in, fin, out, nf_out, fout, lfac, rfac);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

int launch_zero_lmax(raysht_plan *p, double *buf, int nf, hipStream_t st)
{
    if (!p || !buf || nf < 1) return RAYSHT_EINVAL;
    if (int rc = check_spectral(p)) return rc;
    const unsigned bx = (unsigned)std::min<long>((2L * p->nr * nf + kAThreads - 1) / kAThreads, kAMaxBlocksX);
    hipLaunchKernelGGL(zero_lmax_kernel, dim3(bx, (unsigned)p->nm), dim3(kAThreads), 0, st, p->lmax, p->nr,
                       p->d_mvals, p->d_lm_base, buf, nf);
    RAYSHT_HIP_TRY(hipGetLastError());
    return RAYSHT_OK;
}

}  // namespace

// Initialize_Angular_Derivatives (Spectral_Derivatives.F90:49-83) for the plan's m-blocks,
// in its operation order: (c1, c2) per (m, l) for d_by_dtheta, then for d_by_sdtheta, each
// indexed like lm_base (the l count before block mp) + l - m.
int angular_plan_init(raysht_plan *p)
{
    long total = 0;
    for (int m : p->mvals) total += p->lmax - m + 1;
    p->lm_total = total;
    std::vector<double2> c(2 * total);
    long k = 0;
    for (int mp = 0; mp < p->nm; ++mp) {
        const int m = p->mvals[mp];
        for (int l = m; l <= p->lmax; ++l, ++k) {
            const double q1 = std::sqrt((double)((l + m) * (l - m)) / (double)((2 * l + 1) * (2 * l - 1)));
            const double q2 = std::sqrt((double)((l + 1 + m) * (l + 1 - m)) / (double)((2 * l + 1) * (2 * l + 3)));
            c[k] = make_double2((double)(l - 1) * q1, -((double)(l + 2) * q2));
            c[total + k] = make_double2((double)(l + 1) * q1, -((double)l * q2));
        }
    }
    p->d_acoef = nullptr;
    if (c.empty()) return RAYSHT_OK;
    if (hipMalloc(reinterpret_cast<void **>(&p->d_acoef), c.size() * sizeof(double2)) != hipSuccess) {
        p->d_acoef = nullptr;
        return RAYSHT_ENOMEM;
    }
    RAYSHT_HIP_TRY(hipMemcpy(p->d_acoef, c.data(), c.size() * sizeof(double2), hipMemcpyHostToDevice));
    return RAYSHT_OK;
}

void angular_plan_free(raysht_plan *p)
{
    if (p->d_acoef) (void)hipFree(p->d_acoef);
    p->d_acoef = nullptr;
}

}  // namespace raysht

using namespace raysht;

extern "C" {

int raysht_d_by_dtheta(raysht_plan *p, int kind, const double *in, int nf_in, int fin, double *out, int nf_out,
                       int fout, void *stream)
{
    if (kind != 0 && kind != 1) return RAYSHT_EINVAL;
    return launch_angular(p, kind, in, nf_in, fin, out, nf_out, fout, reinterpret_cast<hipStream_t>(stream));
}

int raysht_d_by_dphi(raysht_plan *p, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                     void *stream)
{
    return launch_angular(p, 2, in, nf_in, fin, out, nf_out, fout, reinterpret_cast<hipStream_t>(stream));
}

int raysht_rlm_spaceb_hydro(raysht_plan *p, double *buf, int nf, int wvar, int pvar, int zvar, void *stream)
{
    return launch_rlm_spaceb(p, buf, nf, wvar, pvar, zvar, reinterpret_cast<hipStream_t>(stream));
}

int raysht_rlm_spacea_hydro(raysht_plan *p, double *buf, int nf, const int *fields, const double *over_rhor,
                            const double *over_rhorsq, const double *drho_term, const double *two_over_r,
                            const double *dlnrho, const double *dsdr_over_cp, void *stream)
{
    return launch_spacea_hydro(p, buf, nf, fields, over_rhor, over_rhorsq, drho_term, two_over_r, dlnrho,
                               dsdr_over_cp, reinterpret_cast<hipStream_t>(stream));
}

int raysht_hybrid_output_initial(raysht_plan *p, const double *s, int nf, double *co, int nf_co, double *tmp,
                                 const int *fields, const double *one_over_rsq, void *stream)
{
    return launch_output_initial(p, s, nf, co, nf_co, tmp, fields, one_over_rsq, reinterpret_cast<hipStream_t>(stream));
}

int raysht_hybrid_output_final(raysht_plan *p, const double *s, int nf, double *co, int nf_co, const double *tmp,
                               const int *fields, const double *one_over_r, const double *one_over_rsq,
                               const double *two_over_r, void *stream)
{
    return launch_output_final(p, s, nf, co, nf_co, tmp, fields, one_over_r, one_over_rsq, two_over_r,
                               reinterpret_cast<hipStream_t>(stream));
}

int raysht_adjust_emf(raysht_plan *p, double *buf, int nf, int emftheta, int emfphi, void *stream)
{
    return launch_adjust_emf(p, buf, nf, emftheta, emfphi, reinterpret_cast<hipStream_t>(stream));
}

int raysht_compute_bandcurlb(raysht_plan *p, double *buf, int nf, const int *fields, const double *one_over_rsq,
                             void *stream)
{
    return launch_bandcurlb(p, buf, nf, fields, one_over_rsq, reinterpret_cast<hipStream_t>(stream));
}

int raysht_lm_scale(raysht_plan *p, const double *in, int nf_in, int fin, double *out, int nf_out, int fout,
                    const double *lfac, const double *rfac, void *stream)
{
    return launch_lm_scale(p, in, nf_in, fin, out, nf_out, fout, lfac, rfac, reinterpret_cast<hipStream_t>(stream));
}

int raysht_zero_lmax(raysht_plan *p, double *buf, int nf, void *stream)
{
    return launch_zero_lmax(p, buf, nf, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
