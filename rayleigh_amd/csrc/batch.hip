// Many independent Legendre problems (plans of different l_max, N_theta and column
// counts) in one grouped launch per direction: the m-blocks of every plan are dealt
// to the XCDs together, longest contraction first, and each work item names its
// problem; the kernels read that problem's tables and sizes from a descriptor array
// in device memory and its input / output pointers from the launch arguments.
#include <algorithm>
#include <new>
#include <vector>

#include "raysht_internal.h"

struct raysht_batch {
    int device = 0;
    int precision = 0;
    std::vector<raysht_plan *> plans;
    raysht::prob_desc *d_desc = nullptr;
    raysht::work_item *d_work_stp = nullptr, *d_work_pts = nullptr;
    int n_work_stp = 0, n_work_pts = 0;
};

using namespace raysht;

namespace {

void batch_free(raysht_batch *b)
{
    for (void *ptr : {(void *)b->d_desc, (void *)b->d_work_stp, (void *)b->d_work_pts})
        if (ptr) (void)hipFree(ptr);
    delete b;
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
    const std::vector<work_item> ws = fp64 ? fp64_work(cp, true) : split_work(cp, true);
    const std::vector<work_item> wp = fp64 ? fp64_work(cp, false) : split_work(cp, false);
    b->n_work_stp = (int)ws.size();
    b->n_work_pts = (int)wp.size();
    int rc;
    if ((rc = upload(&b->d_desc, desc)) || (rc = upload(&b->d_work_stp, ws)) || (rc = upload(&b->d_work_pts, wp))) {
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
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (b->precision == RAYSHT_FP64) return launch_stp_fp64_batch(b->d_desc, io, b->d_work_stp, b->n_work_stp, st);
    return launch_stp_split_batch(b->precision, b->d_desc, io, b->d_work_stp, b->n_work_stp, st);
}

int raysht_batch_pts(raysht_batch *b, const double *const *p2, double *const *s, void *stream)
{
    if (!b) return RAYSHT_EINVAL;
    batch_io io;
    int rc = batch_io_of(b, p2, s, io);
    if (rc != RAYSHT_OK) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (b->precision == RAYSHT_FP64) return launch_pts_fp64_batch(b->d_desc, io, b->d_work_pts, b->n_work_pts, st);
    return launch_pts_split_batch(b->precision, b->d_desc, io, b->d_work_pts, b->n_work_pts, st);
}

}  // extern "C"
