// Plan life cycle and the Legendre entry points of the C ABI.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "raysht_internal.h"

namespace raysht {

int stp_tile_rows();
int stp_tile_cols();
int pts_tile_rows();
int pts_tile_cols();
int split_plan_init(raysht_plan *p);
void split_plan_free(raysht_plan *p);
int launch_stp_split(raysht_plan *p, const double *s, double *p2, hipStream_t st);
int launch_pts_split(raysht_plan *p, const double *p2, double *s, hipStream_t st);

namespace {
thread_local std::string g_last_error;

template <typename T>
int device_upload(T **dst, const T *src, size_t n)
{
    *dst = nullptr;
    if (n == 0) return RAYSHT_OK;
    if (hipMalloc(reinterpret_cast<void **>(dst), n * sizeof(T)) != hipSuccess) {
        *dst = nullptr;
        return RAYSHT_ENOMEM;
    }
    if (src) RAYSHT_HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return RAYSHT_OK;
}

template <typename T>
void device_free(T *&ptr)
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
}

int n_even(int lmax, int m) { return (lmax - m + 2) / 2; }
int n_odd(int lmax, int m) { return (lmax - m + 1) / 2; }

}  // namespace

// Work items of one grouped launch.  Workgroups are dispatched to the 8 XCDs of the
// chip round-robin by workgroup index, each XCD with its own L2.  Whole m-blocks are
// assigned to XCDs (longest first, each to the least-loaded XCD) so that the tiles
// of one m-block, which share its P_lm table and its input columns, run on one XCD
// and hit in its L2; within an XCD the largest contractions come first.  The list
// is interleaved so that workgroup b runs item b/8 of XCD b%8, padded with empty
// items (mp = -1) where XCD queues differ in length.  Without XCD groups the blocks'
// items are concatenated in the same longest-first order.
std::vector<work_item> deal_to_xcds(const std::vector<std::vector<work_item>> &blocks,
                                    const std::vector<double> &cost, bool xcd_groups)
{
    constexpr int kXcds = 8;
    if (!xcd_groups) {
        std::vector<work_item> w;
        for (const auto &b : blocks) w.insert(w.end(), b.begin(), b.end());
        return w;
    }
    std::vector<std::vector<work_item>> q(kXcds);
    std::vector<double> load(kXcds, 0.0);
    for (size_t i = 0; i < blocks.size(); ++i) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        q[x].insert(q[x].end(), blocks[i].begin(), blocks[i].end());
        load[x] += cost[i];
    }
    size_t len = 0;
    for (auto &v : q) len = std::max(len, v.size());
    std::vector<work_item> w(len * kXcds, work_item{-1, 0, 0, 0});
    for (size_t i = 0; i < len; ++i)
        for (int x = 0; x < kXcds; ++x)
            if (i < q[x].size()) w[i * kXcds + x] = q[x][i];
    return w;
}

// The FP64 kernels' work list.  Inside an m-block the row tiles of one column tile are
// consecutive, so they find the column tile's input in L2.  The m-blocks of several
// problems share one list; `pad` holds the problem index and the tile shape.
std::vector<work_item> fp64_work(const std::vector<const raysht_plan *> &plans, bool stp, sched_opts o)
{
    const int tc = stp ? stp_tile_cols() : pts_tile_cols();
    const int tr = stp ? stp_tile_rows() : pts_tile_rows();
    struct blk {
        int prob, mp;
        double item;  // cost of one item ~ its K loop: parity-l chunks (StP), node chunks (PtS)
    };
    std::vector<blk> order;
    for (int pi = 0; pi < (int)plans.size(); ++pi) {
        const raysht_plan *p = plans[pi];
        for (int mp = 0; mp < p->nm; ++mp)
            order.push_back(blk{pi, mp, stp ? (double)n_even(p->lmax, p->mvals[mp]) : (double)p->nt2});
    }
    // longest first; one problem's blocks keep ascending m among equal costs
    std::stable_sort(order.begin(), order.end(), [&](const blk &a, const blk &b) {
        if (a.item != b.item) return a.item > b.item;
        if (a.prob != b.prob) return a.prob < b.prob;
        return plans[a.prob]->mvals[a.mp] < plans[b.prob]->mvals[b.mp];
    });
    std::vector<std::vector<work_item>> blocks;
    std::vector<double> cost;
    for (const blk &b : order) {
        const raysht_plan *p = plans[b.prob];
        const int rows = stp ? p->nt2 : n_even(p->lmax, p->mvals[b.mp]);
        const int nrt = (rows + tr - 1) / tr, nct = (p->nrhs + tc - 1) / tc;
        // PtS: a last row tile of at most a quarter, a half or three quarters of the height
        // runs as that shape (shape code 2, 1 or 3)
        const int rem = rows - (nrt - 1) * tr;
        const int shape = (stp || !o.half_tiles) ? 0
                          : rem <= tr / 4        ? 2
                          : rem <= tr / 2        ? 1
                          : rem <= 3 * tr / 4    ? 3
                                                 : 0;
        std::vector<work_item> items;
        // column tile outer, row tile inner: consecutive items share the column tile's
        // input (spectral columns for StP, hybrid columns for PtS) while it is in L2
        for (int c = 0; c < nct; ++c)
            for (int r = 0; r < nrt; ++r)
                items.push_back(work_item{b.mp, r * tr, c * tc, b.prob | ((r == nrt - 1 ? shape : 0) << 16)});
        blocks.push_back(std::move(items));
        cost.push_back(b.item * (nrt - (shape == 2 ? 0.75 : shape == 1 ? 0.5 : shape == 3 ? 0.25 : 0.0)) * nct);
    }
    return deal_to_xcds(blocks, cost, o.xcd_groups);
}

void set_last_hip_error(hipError_t e, const char *what)
{
    g_last_error = std::string(hipGetErrorString(e)) + " in " + what;
}

}  // namespace raysht

using namespace raysht;

extern "C" {

const char *raysht_strerror(int status)
{
    switch (status) {
    case RAYSHT_OK: return "ok";
    case RAYSHT_EINVAL: return "invalid argument";
    case RAYSHT_EHIP: return g_last_error.empty() ? "HIP runtime error" : g_last_error.c_str();
    case RAYSHT_ENOMEM: return "out of memory";
    case RAYSHT_EUNSUPPORTED: return "not supported";
    case RAYSHT_ENOTREADY: return "a table this call may not make has not been made yet";
    default: return "unknown status";
    }
}

int raysht_version(void) { return 100; }

void raysht_plan_destroy(raysht_plan *p)
{
    if (!p) return;
    device_free(p->d_pt_even);
    device_free(p->d_pt_odd);
    device_free(p->d_ip_even);
    device_free(p->d_ip_odd);
    device_free(p->d_s_off);
    device_free(p->d_te_off);
    device_free(p->d_to_off);
    device_free(p->d_mvals);
    device_free(p->d_mp_of_m);
    device_free(p->d_work_stp);
    device_free(p->d_work_pts);
    device_free(p->d_stage_s);
    device_free(p->d_stage_p);
    device_free(p->d_desc);
    split_plan_free(p);
    radial_plan_free(p);
    angular_plan_free(p);
    delete p;
}

}  // extern "C"

namespace {

prob_desc describe(const raysht_plan *p)
{
    prob_desc d{};
    d.pt_even = p->d_pt_even;
    d.pt_odd = p->d_pt_odd;
    d.ip_even = p->d_ip_even;
    d.ip_odd = p->d_ip_odd;
    d.s_off = p->d_s_off;
    d.te_off = p->d_te_off;
    d.to_off = p->d_to_off;
    d.mvals = p->d_mvals;
    d.frag_stp_off = p->split.d_frag_stp_off;
    d.frag_pts_off = p->split.d_frag_pts_off;
    d.exp_pts_off = p->split.d_exp_pts_off;
    d.tdig_stp = p->split.d_tdig_stp;
    d.tdig_pts = p->split.d_tdig_pts;
    d.texp_stp = p->split.d_texp_stp;
    d.texp_pts = p->split.d_texp_pts;
    d.lmax = p->lmax;
    d.nt = p->nt;
    d.nt2 = p->nt2;
    d.nrhs = p->nrhs;
    return d;
}

// nr = nfields = 0 makes a plan over `ncols` anonymous columns (the rmcontainer
// 2-D transforms); otherwise ncols = 2 nr nfields.
int plan_create(raysht_plan **out, int lmax, int ntheta, const int *m_values, int n_m, int nr, int nfields,
                long ncols, int precision)
{
    if (!out) return RAYSHT_EINVAL;
    *out = nullptr;
    if (lmax < 0 || lmax > 8191 || ntheta < 2 || (ntheta & 1) || !m_values || n_m < 1 || n_m > lmax + 1 ||
        ncols < 1 || nr < 0 || nfields < 0)
        return RAYSHT_EINVAL;
    if (precision < RAYSHT_FP64 || precision > RAYSHT_INT8X6) return RAYSHT_EINVAL;
    // the INT8 forms' exact int32 sums cover contractions of up to 16384 terms
    if ((precision == RAYSHT_INT8X3 || precision == RAYSHT_INT8X6) && (ntheta / 2 > 16384 || (lmax + 2) / 2 > 16384))
        return RAYSHT_EUNSUPPORTED;
    // every m-block of the hybrid and spectral layouts is addressed with 32-bit byte offsets
    if (ncols * (long)ntheta * 8 >= (1L << 31) || ncols * (long)(lmax + 1) * 8 >= (1L << 31)) return RAYSHT_EINVAL;
    std::vector<int> seen(lmax + 1, -1);
    for (int i = 0; i < n_m; ++i) {
        const int m = m_values[i];
        if (m < 0 || m > lmax || seen[m] >= 0) return RAYSHT_EINVAL;
        seen[m] = i;
    }

    int device = 0;
    RAYSHT_HIP_TRY(hipGetDevice(&device));
    raysht_plan *p = new (std::nothrow) raysht_plan();
    if (!p) return RAYSHT_ENOMEM;
    int rc = RAYSHT_OK;
    p->device = device;
    p->lmax = lmax;
    p->nt = ntheta;
    p->nt2 = ntheta / 2;
    p->nm = n_m;
    p->nr = nr;
    p->nf = nfields;
    p->nrhs = (int)ncols;
    p->precision = precision;
    p->mvals.assign(m_values, m_values + n_m);
    p->s_off.assign(n_m + 1, 0);
    p->te_off.assign(n_m + 1, 0);
    p->to_off.assign(n_m + 1, 0);
    for (int i = 0; i < n_m; ++i) {
        const int m = m_values[i];
        p->s_off[i + 1] = p->s_off[i] + (long)(lmax - m + 1) * p->nrhs;
        p->te_off[i + 1] = p->te_off[i] + (long)n_even(lmax, m) * p->nt2;
        p->to_off[i + 1] = p->to_off[i] + (long)n_odd(lmax, m) * p->nt2;
    }

    // quad-precision setup on the host
    p->x_hi.resize(ntheta);
    p->x_lo.resize(ntheta);
    p->w_hi.resize(ntheta);
    p->w_lo.resize(ntheta);
    std::vector<double> rn_hi(ntheta), rn_lo(ntheta), amp_hi(lmax + 1), amp_lo(lmax + 1);
    gauss_legendre_dd(ntheta, p->x_hi.data(), p->x_lo.data(), p->w_hi.data(), p->w_lo.data(), rn_hi.data(),
                      rn_lo.data());
    plm_amplitudes_dd(lmax, amp_hi.data(), amp_lo.data());
    std::vector<double> x_dd(2 * p->nt2), rn_dd(2 * p->nt2), amp_dd(2 * (lmax + 1));
    for (int j = 0; j < p->nt2; ++j) {
        x_dd[2 * j] = p->x_hi[j];
        x_dd[2 * j + 1] = p->x_lo[j];
        rn_dd[2 * j] = rn_hi[j];
        rn_dd[2 * j + 1] = rn_lo[j];
    }
    for (int m = 0; m <= lmax; ++m) {
        amp_dd[2 * m] = amp_hi[m];
        amp_dd[2 * m + 1] = amp_lo[m];
    }

    std::vector<work_item> ws = fp64_work({p}, true), wp = fp64_work({p}, false);
    p->n_work_stp = (int)ws.size();
    p->n_work_pts = (int)wp.size();

    double *d_x = nullptr, *d_rn = nullptr, *d_amp = nullptr;
    const size_t ne_tot = p->te_off[n_m], no_tot = p->to_off[n_m];
    if ((rc = device_upload(&p->d_s_off, p->s_off.data(), p->s_off.size())) ||
        (rc = device_upload(&p->d_te_off, p->te_off.data(), p->te_off.size())) ||
        (rc = device_upload(&p->d_to_off, p->to_off.data(), p->to_off.size())) ||
        (rc = device_upload(&p->d_mvals, p->mvals.data(), p->mvals.size())) ||
        (rc = device_upload(&p->d_mp_of_m, seen.data(), seen.size())) ||
        (rc = device_upload(&p->d_work_stp, ws.data(), ws.size())) ||
        (rc = device_upload(&p->d_work_pts, wp.data(), wp.size())) ||
        (rc = device_upload(&p->d_pt_even, (const double *)nullptr, ne_tot)) ||
        (rc = device_upload(&p->d_ip_even, (const double *)nullptr, ne_tot)) ||
        (rc = device_upload(&p->d_pt_odd, (const double *)nullptr, std::max<size_t>(no_tot, 1))) ||
        (rc = device_upload(&p->d_ip_odd, (const double *)nullptr, std::max<size_t>(no_tot, 1))) ||
        (rc = device_upload(&d_x, x_dd.data(), x_dd.size())) ||
        (rc = device_upload(&d_rn, rn_dd.data(), rn_dd.size())) ||
        (rc = device_upload(&d_amp, amp_dd.data(), amp_dd.size()))) {
        device_free(d_x);
        device_free(d_rn);
        device_free(d_amp);
        raysht_plan_destroy(p);
        return rc;
    }
    rc = launch_plm_tables(p, d_x, d_rn, d_amp, nullptr);
    if (rc == RAYSHT_OK && hipDeviceSynchronize() != hipSuccess) rc = RAYSHT_EHIP;
    device_free(d_x);
    device_free(d_rn);
    device_free(d_amp);
    if (rc == RAYSHT_OK && precision != RAYSHT_FP64) rc = split_plan_init(p);
    if (rc == RAYSHT_OK) rc = radial_plan_init(p);
    if (rc == RAYSHT_OK) rc = angular_plan_init(p);
    if (rc == RAYSHT_OK) {
        p->desc = describe(p);
        rc = device_upload(&p->d_desc, &p->desc, 1);
    }
    if (rc != RAYSHT_OK) {
        raysht_plan_destroy(p);
        return rc;
    }
    *out = p;
    return RAYSHT_OK;
}

}  // namespace

extern "C" {

int raysht_plan_create(raysht_plan **out, int lmax, int ntheta, const int *m_values, int n_m, int nr, int nfields,
                       int precision)
{
    if (nr < 1 || nfields < 1) {
        if (out) *out = nullptr;
        return RAYSHT_EINVAL;
    }
    return plan_create(out, lmax, ntheta, m_values, n_m, nr, nfields, 2L * nr * nfields, precision);
}

int raysht_plan_create_cols(raysht_plan **out, int lmax, int ntheta, const int *m_values, int n_m, int ncols,
                            int precision)
{
    return plan_create(out, lmax, ntheta, m_values, n_m, 0, 0, ncols, precision);
}

size_t raysht_spectral_size(const raysht_plan *p) { return p ? (size_t)p->s_off[p->nm] : 0; }

size_t raysht_hybrid_size(const raysht_plan *p) { return p ? (size_t)p->nt * p->nrhs * p->nm : 0; }

int raysht_legendre_stp(raysht_plan *p, const double *s, double *p2, void *stream)
{
    if (!p || !s || !p2) return RAYSHT_EINVAL;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (p->precision == RAYSHT_FP64) return launch_stp_fp64(p, s, p2, st);
    return launch_stp_split(p, s, p2, st);
}

int raysht_legendre_pts(raysht_plan *p, const double *p2, double *s, void *stream)
{
    if (!p || !s || !p2) return RAYSHT_EINVAL;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (p->precision == RAYSHT_FP64) return launch_pts_fp64(p, p2, s, st);
    return launch_pts_split(p, p2, s, st);
}

int raysht_plan_nodes(const raysht_plan *p, double *x_hi, double *x_lo, double *w_hi, double *w_lo)
{
    if (!p || !x_hi || !x_lo || !w_hi || !w_lo) return RAYSHT_EINVAL;
    std::memcpy(x_hi, p->x_hi.data(), sizeof(double) * p->nt);
    std::memcpy(x_lo, p->x_lo.data(), sizeof(double) * p->nt);
    std::memcpy(w_hi, p->w_hi.data(), sizeof(double) * p->nt);
    std::memcpy(w_lo, p->w_lo.data(), sizeof(double) * p->nt);
    return RAYSHT_OK;
}

int raysht_plan_nodes_quad(const raysht_plan *p, void *x16, void *w16)
{
    if (!p || !x16 || !w16) return RAYSHT_EINVAL;
    gauss_legendre_binary128(p->nt, x16, w16);
    return RAYSHT_OK;
}

size_t raysht_plan_table_size(const raysht_plan *p, int which)
{
    if (!p || which < 0 || which > 3) return 0;
    return (which % 2 == 0) ? (size_t)p->te_off[p->nm] : (size_t)p->to_off[p->nm];
}

int raysht_plan_table(const raysht_plan *p, int which, double *host_out)
{
    if (!p || !host_out || which < 0 || which > 3) return RAYSHT_EINVAL;
    const bool even = (which % 2 == 0);
    const size_t n = raysht_plan_table_size(p, which);
    if (n == 0) return RAYSHT_OK;
    const double *src = which == 0 ? p->d_pt_even : which == 1 ? p->d_pt_odd : which == 2 ? p->d_ip_even : p->d_ip_odd;
    std::vector<double> tmp(n);
    RAYSHT_HIP_TRY(hipMemcpy(tmp.data(), src, n * sizeof(double), hipMemcpyDeviceToHost));
    if (which >= 2) {
        std::memcpy(host_out, tmp.data(), n * sizeof(double));  // [k][j] = ip_lm(1:nt2, 1:n)
        return RAYSHT_OK;
    }
    // device StP tables are [k][j]; the reference stores p_lm(1:n, 1:nt2), l fastest
    const std::vector<long> &off = even ? p->te_off : p->to_off;
    for (int mp = 0; mp < p->nm; ++mp) {
        const int m = p->mvals[mp];
        const int nk = even ? n_even(p->lmax, m) : n_odd(p->lmax, m);
        const double *b = tmp.data() + off[mp];
        double *o = host_out + off[mp];
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < p->nt2; ++j) o[(long)j * nk + k] = b[(long)k * p->nt2 + j];
    }
    return RAYSHT_OK;
}

}  // extern "C"
