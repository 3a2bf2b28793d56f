// Internal declarations shared by the raysht translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/raysht.h"

#define RAYSHT_HIP_TRY(expr)                                  \
    do {                                                      \
        hipError_t raysht_e_ = (expr);                        \
        if (raysht_e_ != hipSuccess) {                        \
            raysht::set_last_hip_error(raysht_e_, #expr);     \
            return RAYSHT_EHIP;                               \
        }                                                     \
    } while (0)

namespace raysht {

void set_last_hip_error(hipError_t e, const char *what);

// Double-double value of a quad computed on the host (head + tail).
struct dd_pair {
    double hi, lo;
};

// One unit of grouped-GEMM work: m-block mp, row tile start, column tile start, and in
// `pad` the problem index (low 16 bits) and, for the FP64 PtS kernel, the tile shape (high
// bits: 0 = the full tile, 1 / 2 / 3 = the half- / quarter- / three-quarter-height tile for
// the last rows of an m-block).  The split kernels' items run every row tile of their
// column tile.
struct work_item {
    int mp;
    int r0;
    int c0;
    int pad;
};

// Device state of the split-precision (BF16 slice) forms: the tables pre-split into
// 1 KiB MFMA B-operand fragments (64 lanes x 8 BF16), the power-of-two scale exponent
// of every table row, and the work lists of the 64 x 128 tiles.
struct raysht_split {
    long *d_frag_stp_off = nullptr;  // first StP fragment of block mp (even, then odd region)
    long *d_frag_pts_off = nullptr;  // first PtS fragment of block mp (even, then odd region)
    long *d_exp_pts_off = nullptr;   // PtS row exponents of block mp: even rows, then odd rows
    work_item *d_work_stp = nullptr, *d_work_pts = nullptr;
    int n_work_stp = 0, n_work_pts = 0;
    uint4 *d_tdig_stp = nullptr, *d_tdig_pts = nullptr;
    int *d_texp_stp = nullptr;  // [mp][parity][j]
    int *d_texp_pts = nullptr;
};

// Everything a contraction kernel reads about one problem (one plan), in device
// memory, so that one grouped launch can cover the m-blocks of several problems: the
// `pad` field of a work item selects the problem.
struct prob_desc {
    const double *pt_even, *pt_odd, *ip_even, *ip_odd;  // FP64 tables
    const long *s_off, *te_off, *to_off;
    const int *mvals;
    // split forms (null for FP64 plans)
    const long *frag_stp_off, *frag_pts_off, *exp_pts_off;
    const uint4 *tdig_stp, *tdig_pts;
    const int *texp_stp, *texp_pts;
    int lmax, nt, nt2, nrhs;
};

// The per-call operands of a grouped launch, passed by value: input and output of each
// problem (StP: spectral in, hybrid out; PtS: hybrid in, spectral out).
constexpr int kMaxBatch = 64;
struct batch_io {
    const double *in[kMaxBatch];
    double *out[kMaxBatch];
};

}  // namespace raysht

struct raysht_plan {
    int device = 0;
    int lmax = 0, nt = 0, nt2 = 0, nm = 0, nr = 0, nf = 0, nrhs = 0, precision = 0;
    std::vector<int> mvals;
    std::vector<long> s_off;   // spectral offset of block mp (doubles), nm + 1 entries
    std::vector<long> te_off;  // even-table offset of block mp (doubles), nm + 1 entries
    std::vector<long> to_off;  // odd-table offset of block mp (doubles), nm + 1 entries
    std::vector<double> x_hi, x_lo, w_hi, w_lo;
    // device-resident setup data
    double *d_pt_even = nullptr, *d_pt_odd = nullptr;  // StP tables P*STP_norm, [k][j] j fastest
    double *d_ip_even = nullptr, *d_ip_odd = nullptr;  // PtS tables 2 pi w P PTS_norm, [k][j]
    long *d_s_off = nullptr, *d_te_off = nullptr, *d_to_off = nullptr;
    int *d_mvals = nullptr;
    int *d_mp_of_m = nullptr;  // m -> mp, -1 where the plan has no such m
    raysht::work_item *d_work_stp = nullptr, *d_work_pts = nullptr;
    int n_work_stp = 0, n_work_pts = 0;
    // staging buffers of the host-array entry points, allocated on first use
    double *d_stage_s = nullptr, *d_stage_p = nullptr;
    raysht::raysht_split split;  // used by the BF16X3 / BF16X6 forms only
    raysht::prob_desc desc{};           // this plan as one problem of a grouped launch
    raysht::prob_desc *d_desc = nullptr;
    // radial transposes: LM_Load_Balance's (l, mp) pairs, the pair index of every
    // (m-block, l) at lm_base[mp] + l - m, and the (mp, first l - m) tiles
    int lm_count = 0, n_rtiles = 0;
    std::vector<int> lm_l, lm_mp;
    int *d_lm_base = nullptr, *d_lm_slot = nullptr;
    int2 *d_rtiles = nullptr;
    // angular derivatives: (c1, c2) per (m-block, l) for d_by_dtheta, then d_by_sdtheta
    long lm_total = 0;
    double2 *d_acoef = nullptr;
};

struct raysht_cheb {
    int device = 0;
    int n = 0, n_even = 0, n_odd = 0, n_x = 0;
    std::vector<double> ceven, codd;  // [k][i] i fastest (n_x), as Gen_Tn stores them
    std::vector<double> x, theta, weights;  // colocation points, their angles, radial integration weights
    // device copies, zero-padded for the MFMA tiles: [i][k] for the analysis
    // (ix4 x kp) and [k][i] for the synthesis (k4 x ip)
    int ix4 = 0, kp = 0, k4 = 0, ip = 0;
    double *d_tt_e = nullptr, *d_tt_o = nullptr, *d_t_e = nullptr, *d_t_o = nullptr;
    double deriv_scaling = 0.0;  // d/dr of the domain map, as Initialize_Cheby_Grid rounds it
    // DCT form (power-of-two n <= 256): one device block of e^{-2 pi i k/n} [n] (double2),
    // e^{-i pi j/(2n)} [n] (double2), the analysis correction [n] and the synthesis one [n]
    double *d_dct = nullptr;
    double dct_inv_s = 1.0;  // 1 / the power of two the analysis correction is scaled by
    int form = 0;            // RAYSHT_CHEB_DCT where d_dct exists, else RAYSHT_CHEB_MATRIX
    // Every grid keeps its domains in the reference's internal order (outermost first): one
    // domain is the grid itself; several (Initialize_Cheby_Grid with ndomains > 1) are
    // table-only grids in `dom`, domain h on rows hoff[h] .. hoff[h] + dom[h]->n - 1 of each
    // column of n = the total length.  rda[h] = the domain's first dealiased row + 1
    // (1-based, as the reference's rda), dscale = deriv_scaling per row.
    std::vector<raysht_cheb *> dom;
    std::vector<int> hoff, rda;
    std::vector<double> dscale;
};

namespace raysht {

// Host-side quad-precision setup (host_quad.cpp): Gauss-Legendre nodes and
// weights as head/tail pairs, 2*pi*w_j, and the l = m amplitude of each m.
void gauss_legendre_dd(int n, double *x_hi, double *x_lo, double *w_hi, double *w_lo,
                       double *renorm_hi, double *renorm_lo);
void gauss_legendre_binary128(int n, void *x16, void *w16);
void plm_amplitudes_dd(int lmax, double *amp_hi, double *amp_lo);
void fft_twiddles(int n, double *cos_out, double *sin_out);
void cheb_dct_tables(int n, const double *theta, double *tw_cos, double *tw_sin, double *q_cos, double *q_sin,
                     double *delta);

// Radial transpose tables of a plan (radial_transpose.hip).
int radial_plan_init(raysht_plan *p);
void radial_plan_free(raysht_plan *p);
// Angular derivative coefficients of a plan (angular.hip).
int angular_plan_init(raysht_plan *p);
void angular_plan_free(raysht_plan *p);

// Device launches (each returns RAYSHT_OK or an error code).
int launch_plm_tables(raysht_plan *p, const double *d_x, const double *d_renorm, const double *d_amp,
                      hipStream_t st);
int launch_stp_fp64(raysht_plan *p, const double *s, double *p2, hipStream_t st);
int launch_pts_fp64(raysht_plan *p, const double *p2, double *s, hipStream_t st);
// grouped launches over the work items of several problems (descriptors in device memory)
int launch_stp_fp64_batch(const prob_desc *d_desc, const batch_io &io, const work_item *work, int n_work,
                          hipStream_t st);
int launch_pts_fp64_batch(const prob_desc *d_desc, const batch_io &io, const work_item *work, int n_work,
                          hipStream_t st);
int launch_stp_split_batch(int precision, const prob_desc *d_desc, const batch_io &io, const work_item *work,
                           int n_work, hipStream_t st);
int launch_pts_split_batch(int precision, const prob_desc *d_desc, const batch_io &io, const work_item *work,
                           int n_work, hipStream_t st);
// XCD-balanced work lists over the m-blocks of several problems (work_item.pad = index
// into `plans`), for the FP64 and the split kernels
// Schedule options of a work list: PtS half-height last tiles (FP64 kernels), and
// whether whole m-blocks are kept on one XCD (else the items go out in LPT order and
// the dispatcher spreads consecutive items over the XCDs).
struct sched_opts {
    bool half_tiles = true;
    bool xcd_groups = true;
};
std::vector<work_item> fp64_work(const std::vector<const raysht_plan *> &plans, bool stp, sched_opts o = {});
std::vector<work_item> split_work(const std::vector<const raysht_plan *> &plans, bool stp, sched_opts o = {});
// deals per-block item lists to the XCDs (LPT over the block costs, blocks in the given
// order) and interleaves them so that item i runs on XCD i % 8, or concatenates them
std::vector<work_item> deal_to_xcds(const std::vector<std::vector<work_item>> &blocks,
                                    const std::vector<double> &cost, bool xcd_groups);
// allocates the device twiddle table of length-nphi FFTs (done by the first FFT call
// otherwise; needed before a stream capture, which may not allocate)
int fft_prepare(int nphi);

}  // namespace raysht
