// Many independent Legendre problems (plans of different l_max, N_theta and column
// counts) in one grouped launch per direction: the m-blocks of every plan are dealt
// to the XCDs together, longest contraction first, and each work item names its
// problem; the kernels read that problem's tables and sizes from a descriptor array
// in device memory and its input / output pointers from the launch arguments.
#include <algorithm>
#include <new>
#include <vector>

#include "raysht_internal.h"

// Schedule candidates of a batch (index bits): bit 0 = no PtS reduced-height tiles (FP64
// kernels), bit 1 = no XCD grouping of m-blocks.  Candidate 0 is the default; the
// autotuner times the distinct candidates of each direction and keeps the fastest.
constexpr int kSchedules = 4;

struct raysht_batch {
    int device = 0;
    int precision = 0;
    std::vector<raysht_plan *> plans;
    raysht::prob_desc *d_desc = nullptr;
    raysht::work_item *d_work[2][kSchedules] = {};  // [direction: 0 StP, 1 PtS][candidate]
    int n_work[2][kSchedules] = {};
    int choice[2] = {0, 0};
};

using namespace raysht;

namespace {

void batch_free(raysht_batch *b)
{
    if (b->d_desc) (void)hipFree(b->d_desc);
    for (auto &dir : b->d_work)
        for (work_item *w : dir)
            if (w) (void)hipFree(w);
    delete b;
}

sched_opts schedule(int c) { return sched_opts{(c & 1) == 0, (c & 2) == 0}; }

// candidates that differ from a lower-numbered one are the ones worth timing
bool distinct(const raysht_batch *b, int dir, int c)
{
    const bool half_matters = b->precision == RAYSHT_FP64 && dir == 1;
    return half_matters || (c & 1) == 0;
}

int launch_dir(raysht_batch *b, int dir, int c, const batch_io &io, hipStream_t st)
{
    work_item *w = b->d_work[dir][c];
    const int n = b->n_work[dir][c];
    if (b->precision == RAYSHT_FP64)
        return dir == 0 ? launch_stp_fp64_batch(b->d_desc, io, w, n, st) : launch_pts_fp64_batch(b->d_desc, io, w, n, st);
    return dir == 0 ? launch_stp_split_batch(b->precision, b->d_desc, io, w, n, st)
                    : launch_pts_split_batch(b->precision, b->d_desc, io, w, n, st);
}

template <typename T>
int upload(T **dst, const std::vector<T> &v)
{
    *dst = nullptr;
    const size_t n = std::max<size_t>(v.size(), 1);
    if (hipMalloc(reinterpret_cast<void **>(dst), n * sizeof(T)) != hipSuccess) {
        *dst = nullptr;
        return RAYSHT_ENOMEM;
    }
    if (!v.empty()) RAYSHT_HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return RAYSHT_OK;
}

int batch_io_of(const raysht_batch *b, const double *const *in, double *const *out, batch_io &io)
{
    if (!in || !out) return RAYSHT_EINVAL;
    io = batch_io{};
    for (size_t i = 0; i < b->plans.size(); ++i) {
        if (!in[i] || !out[i]) return RAYSHT_EINVAL;
        io.in[i] = in[i];
        io.out[i] = out[i];
    }
    return RAYSHT_OK;
}

}  // namespace

extern "C" {

int raysht_batch_create(raysht_batch **out, raysht_plan *const *plans, int n_plans)
{
    if (!out) return RAYSHT_EINVAL;
    *out = nullptr;
    if (!plans || n_plans < 1 || n_plans > kMaxBatch) return RAYSHT_EINVAL;
    int device = 0;
    RAYSHT_HIP_TRY(hipGetDevice(&device));
    for (int i = 0; i < n_plans; ++i)
        if (!plans[i] || plans[i]->precision != plans[0]->precision || plans[i]->device != device)
            return RAYSHT_EINVAL;
    raysht_batch *b = new (std::nothrow) raysht_batch();
    if (!b) return RAYSHT_ENOMEM;
    b->device = device;
    b->precision = plans[0]->precision;
    b->plans.assign(plans, plans + n_plans);
    std::vector<const raysht_plan *> cp(plans, plans + n_plans);
    std::vector<prob_desc> desc(n_plans);
    for (int i = 0; i < n_plans; ++i) desc[i] = plans[i]->desc;
    const bool fp64 = b->precision == RAYSHT_FP64;
    int rc = upload(&b->d_desc, desc);
    for (int dir = 0; dir < 2 && rc == RAYSHT_OK; ++dir)
        for (int c = 0; c < kSchedules && rc == RAYSHT_OK; ++c) {
            const std::vector<work_item> w =
                fp64 ? fp64_work(cp, dir == 0, schedule(c)) : split_work(cp, dir == 0, schedule(c));
            b->n_work[dir][c] = (int)w.size();
            rc = upload(&b->d_work[dir][c], w);
        }
    if (rc != RAYSHT_OK) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return RAYSHT_OK;
}

void raysht_batch_destroy(raysht_batch *b)
{
    if (b) batch_free(b);
}

int raysht_batch_size(const raysht_batch *b) { return b ? (int)b->plans.size() : 0; }

int raysht_batch_stp(raysht_batch *b, const double *const *s, double *const *p2, void *stream)
{
    if (!b) return RAYSHT_EINVAL;
    batch_io io;
    int rc = batch_io_of(b, s, p2, io);
    if (rc != RAYSHT_OK) return rc;
    return launch_dir(b, 0, b->choice[0], io, reinterpret_cast<hipStream_t>(stream));
}

int raysht_batch_pts(raysht_batch *b, const double *const *p2, double *const *s, void *stream)
{
    if (!b) return RAYSHT_EINVAL;
    batch_io io;
    int rc = batch_io_of(b, p2, s, io);
    if (rc != RAYSHT_OK) return rc;
    return launch_dir(b, 1, b->choice[1], io, reinterpret_cast<hipStream_t>(stream));
}

int raysht_batch_autotune(raysht_batch *b, const double *const *s, double *const *p2, double *const *s_out,
                          int reps, void *stream, double *ms_out)
{
    if (!b || reps < 1) return RAYSHT_EINVAL;
    batch_io io[2];
    int rc;
    if ((rc = batch_io_of(b, s, p2, io[0])) != RAYSHT_OK || (rc = batch_io_of(b, p2, s_out, io[1])) != RAYSHT_OK)
        return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RAYSHT_HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return RAYSHT_EHIP;
    }
    rc = RAYSHT_OK;
    for (int dir = 0; dir < 2 && rc == RAYSHT_OK; ++dir) {
        float best = -1.0f;
        for (int c = 0; c < kSchedules && rc == RAYSHT_OK; ++c) {
            if (ms_out) ms_out[dir * kSchedules + c] = -1.0;
            if (!distinct(b, dir, c)) continue;
            rc = launch_dir(b, dir, c, io[dir], st);  // warm-up
            if (rc == RAYSHT_OK && hipEventRecord(e0, st) != hipSuccess) rc = RAYSHT_EHIP;
            for (int r = 0; r < reps && rc == RAYSHT_OK; ++r) rc = launch_dir(b, dir, c, io[dir], st);
            float ms = 0.0f;
            if (rc == RAYSHT_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = RAYSHT_EHIP;
            if (rc != RAYSHT_OK) break;
            ms /= reps;
            if (ms_out) ms_out[dir * kSchedules + c] = ms;
            if (best < 0.0f || ms < best) {
                best = ms;
                b->choice[dir] = c;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int raysht_batch_schedule(const raysht_batch *b, int direction)
{
    if (!b || direction < 0 || direction > 1) return -1;
    return b->choice[direction];
}

int raysht_batch_set_schedule(raysht_batch *b, int direction, int candidate)
{
    if (!b || direction < 0 || direction > 1 || candidate < 0 || candidate >= kSchedules) return RAYSHT_EINVAL;
    b->choice[direction] = candidate;
    return RAYSHT_OK;
}

}  // extern "C"
