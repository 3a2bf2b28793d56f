// Host-array entry points: the drop-in form for callers whose buffers live in host
// memory (Rayleigh's own rmcontainer4d arrays, reached through the Fortran
// ISO_C_BINDING module).  Each call stages its operands through device memory,
// runs the device transform and copies the result back; it returns when the
// result is in the caller's array.  Device-resident callers use the device-pointer
// entry points instead and pay no transfer.
#include <cstring>
#include <vector>

#include "raysht_internal.h"

using namespace raysht;

namespace {

int ensure_stage(raysht_plan *p)
{
    if (p->d_stage_s && p->d_stage_p) return RAYSHT_OK;
    const size_t ns = (size_t)p->s_off[p->nm] * sizeof(double);
    const size_t nh = (size_t)p->nt * p->nrhs * p->nm * sizeof(double);
    if (!p->d_stage_s && hipMalloc(&p->d_stage_s, ns) != hipSuccess) {
        p->d_stage_s = nullptr;
        return RAYSHT_ENOMEM;
    }
    if (!p->d_stage_p && hipMalloc(&p->d_stage_p, nh) != hipSuccess) {
        p->d_stage_p = nullptr;
        return RAYSHT_ENOMEM;
    }
    return RAYSHT_OK;
}

struct scoped_device_buffer {
    double *ptr = nullptr;
    ~scoped_device_buffer()
    {
        if (ptr) (void)hipFree(ptr);
    }
};

struct scoped_device_bytes {
    void *ptr = nullptr;
    ~scoped_device_bytes()
    {
        if (ptr) (void)hipFree(ptr);
    }
};

}  // namespace

extern "C" {

int raysht_legendre_stp_host(raysht_plan *p, const double *const *s_m, double *p2)
{
    if (!p || !s_m || !p2) return RAYSHT_EINVAL;
    for (int mp = 0; mp < p->nm; ++mp)
        if (!s_m[mp]) return RAYSHT_EINVAL;
    int rc = ensure_stage(p);
    if (rc != RAYSHT_OK) return rc;
    for (int mp = 0; mp < p->nm; ++mp) {
        const size_t n = (size_t)(p->s_off[mp + 1] - p->s_off[mp]) * sizeof(double);
        RAYSHT_HIP_TRY(hipMemcpy(p->d_stage_s + p->s_off[mp], s_m[mp], n, hipMemcpyHostToDevice));
    }
    if ((rc = raysht_legendre_stp(p, p->d_stage_s, p->d_stage_p, nullptr)) != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(p2, p->d_stage_p, (size_t)p->nt * p->nrhs * p->nm * sizeof(double),
                             hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

int raysht_legendre_pts_host(raysht_plan *p, const double *p2, double *const *s_m)
{
    if (!p || !s_m || !p2) return RAYSHT_EINVAL;
    for (int mp = 0; mp < p->nm; ++mp)
        if (!s_m[mp]) return RAYSHT_EINVAL;
    int rc = ensure_stage(p);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(p->d_stage_p, p2, (size_t)p->nt * p->nrhs * p->nm * sizeof(double),
                             hipMemcpyHostToDevice));
    if ((rc = raysht_legendre_pts(p, p->d_stage_p, p->d_stage_s, nullptr)) != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipDeviceSynchronize());
    for (int mp = 0; mp < p->nm; ++mp) {
        const size_t n = (size_t)(p->s_off[mp + 1] - p->s_off[mp]) * sizeof(double);
        RAYSHT_HIP_TRY(hipMemcpy(s_m[mp], p->d_stage_s + p->s_off[mp], n, hipMemcpyDeviceToHost));
    }
    return RAYSHT_OK;
}

int raysht_fft_to_physical_host(int nphi, long howmany, double *x)
{
    if (nphi < 2 || (nphi & 1) || howmany < 0 || (!x && howmany > 0)) return RAYSHT_EINVAL;
    if (howmany == 0) return RAYSHT_OK;
    const size_t bytes = (size_t)howmany * (nphi + 2) * sizeof(double);
    scoped_device_buffer d;
    if (hipMalloc(&d.ptr, bytes) != hipSuccess) {
        d.ptr = nullptr;
        return RAYSHT_ENOMEM;
    }
    RAYSHT_HIP_TRY(hipMemcpy(d.ptr, x, bytes, hipMemcpyHostToDevice));
    int rc = raysht_fft_to_physical(nphi, howmany, d.ptr, nullptr);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(x, d.ptr, bytes, hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

int raysht_fft_to_spectral_host(int nphi, long howmany, double *x, int rescale)
{
    if (nphi < 2 || (nphi & 1) || howmany < 0 || (!x && howmany > 0)) return RAYSHT_EINVAL;
    if (howmany == 0) return RAYSHT_OK;
    const size_t bytes = (size_t)howmany * (nphi + 2) * sizeof(double);
    scoped_device_buffer d;
    if (hipMalloc(&d.ptr, bytes) != hipSuccess) {
        d.ptr = nullptr;
        return RAYSHT_ENOMEM;
    }
    RAYSHT_HIP_TRY(hipMemcpy(d.ptr, x, bytes, hipMemcpyHostToDevice));
    int rc = raysht_fft_to_spectral(nphi, howmany, d.ptr, rescale, nullptr);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(x, d.ptr, bytes, hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

// SphericalBuffer%reform between the hybrid and physical layouts from host arrays (the
// Fortran Reform_2a3a / Reform_3b2b): the hybrid array goes through the plan's staging
// buffer, the physical one through a buffer of this call.
int raysht_permute_2a3a_host(raysht_plan *p, int nphi, const double *p2, double *p3)
{
    if (!p || !p2 || !p3 || nphi < 2) return RAYSHT_EINVAL;
    int rc = ensure_stage(p);
    if (rc != RAYSHT_OK) return rc;
    const size_t nh = (size_t)p->nt * p->nrhs * p->nm * sizeof(double);
    const size_t ng = (size_t)(nphi + 2) * p->nt * p->nr * p->nf * sizeof(double);
    scoped_device_buffer d;
    if (ng > 0 && hipMalloc(&d.ptr, ng) != hipSuccess) {
        d.ptr = nullptr;
        return RAYSHT_ENOMEM;
    }
    RAYSHT_HIP_TRY(hipMemcpy(p->d_stage_p, p2, nh, hipMemcpyHostToDevice));
    if ((rc = raysht_permute_2a3a(p, nphi, p->d_stage_p, d.ptr, nullptr)) != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(p3, d.ptr, ng, hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

int raysht_permute_3b2b_host(raysht_plan *p, int nphi, const double *p3, double *p2)
{
    if (!p || !p2 || !p3 || nphi < 2) return RAYSHT_EINVAL;
    int rc = ensure_stage(p);
    if (rc != RAYSHT_OK) return rc;
    const size_t nh = (size_t)p->nt * p->nrhs * p->nm * sizeof(double);
    const size_t ng = (size_t)(nphi + 2) * p->nt * p->nr * p->nf * sizeof(double);
    scoped_device_buffer d;
    if (ng > 0 && hipMalloc(&d.ptr, ng) != hipSuccess) {
        d.ptr = nullptr;
        return RAYSHT_ENOMEM;
    }
    RAYSHT_HIP_TRY(hipMemcpy(d.ptr, p3, ng, hipMemcpyHostToDevice));
    if ((rc = raysht_permute_3b2b(p, nphi, d.ptr, p->d_stage_p, nullptr)) != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(p2, p->d_stage_p, nh, hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

// SphericalBuffer%reform between the spectral (rmcontainer4d) and radial layouts from host
// arrays (the Fortran Reform_2b1b / Reform_1a2a): s_m[mp] as for the Legendre host forms,
// p1 the host radial array p1(1:nr, 1:2, 1:lm_count, 1:nf).
int raysht_permute_2b1b_host(raysht_plan *p, const double *const *s_m, double *p1)
{
    if (!p || !s_m || !p1) return RAYSHT_EINVAL;
    for (int mp = 0; mp < p->nm; ++mp)
        if (!s_m[mp]) return RAYSHT_EINVAL;
    int rc = ensure_stage(p);
    if (rc != RAYSHT_OK) return rc;
    const size_t ns = (size_t)p->s_off[p->nm] * sizeof(double);
    scoped_device_buffer d;
    if (hipMalloc(&d.ptr, ns) != hipSuccess) {
        d.ptr = nullptr;
        return RAYSHT_ENOMEM;
    }
    for (int mp = 0; mp < p->nm; ++mp) {
        const size_t n = (size_t)(p->s_off[mp + 1] - p->s_off[mp]) * sizeof(double);
        RAYSHT_HIP_TRY(hipMemcpy(p->d_stage_s + p->s_off[mp], s_m[mp], n, hipMemcpyHostToDevice));
    }
    if ((rc = raysht_permute_2b1b(p, p->d_stage_s, d.ptr, nullptr)) != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(p1, d.ptr, ns, hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

int raysht_permute_1a2a_host(raysht_plan *p, const double *p1, double *const *s_m)
{
    if (!p || !s_m || !p1) return RAYSHT_EINVAL;
    for (int mp = 0; mp < p->nm; ++mp)
        if (!s_m[mp]) return RAYSHT_EINVAL;
    int rc = ensure_stage(p);
    if (rc != RAYSHT_OK) return rc;
    const size_t ns = (size_t)p->s_off[p->nm] * sizeof(double);
    scoped_device_buffer d;
    if (hipMalloc(&d.ptr, ns) != hipSuccess) {
        d.ptr = nullptr;
        return RAYSHT_ENOMEM;
    }
    RAYSHT_HIP_TRY(hipMemcpy(d.ptr, p1, ns, hipMemcpyHostToDevice));
    if ((rc = raysht_permute_1a2a(p, d.ptr, p->d_stage_s, nullptr)) != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipDeviceSynchronize());
    for (int mp = 0; mp < p->nm; ++mp) {
        const size_t n = (size_t)(p->s_off[mp + 1] - p->s_off[mp]) * sizeof(double);
        RAYSHT_HIP_TRY(hipMemcpy(s_m[mp], p->d_stage_s + p->s_off[mp], n, hipMemcpyDeviceToHost));
    }
    return RAYSHT_OK;
}

int raysht_cheb_to_spectral_host(raysht_cheb *cb, const double *f, double *c, long ncols)
{
    if (!cb || ncols < 0 || (ncols > 0 && (!f || !c))) return RAYSHT_EINVAL;
    if (ncols == 0) return RAYSHT_OK;
    const size_t bytes = (size_t)ncols * cb->n * sizeof(double);
    scoped_device_buffer din, dout;
    if (hipMalloc(&din.ptr, bytes) != hipSuccess || hipMalloc(&dout.ptr, bytes) != hipSuccess) return RAYSHT_ENOMEM;
    RAYSHT_HIP_TRY(hipMemcpy(din.ptr, f, bytes, hipMemcpyHostToDevice));
    int rc = raysht_cheb_to_spectral(cb, din.ptr, dout.ptr, ncols, nullptr);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(c, dout.ptr, bytes, hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

int raysht_cheb_from_spectral_host(raysht_cheb *cb, const double *c, double *f, long ncols)
{
    if (!cb || ncols < 0 || (ncols > 0 && (!f || !c))) return RAYSHT_EINVAL;
    if (ncols == 0) return RAYSHT_OK;
    const size_t bytes = (size_t)ncols * cb->n * sizeof(double);
    scoped_device_buffer din, dout;
    if (hipMalloc(&din.ptr, bytes) != hipSuccess || hipMalloc(&dout.ptr, bytes) != hipSuccess) return RAYSHT_ENOMEM;
    RAYSHT_HIP_TRY(hipMemcpy(din.ptr, c, bytes, hipMemcpyHostToDevice));
    int rc = raysht_cheb_from_spectral(cb, din.ptr, dout.ptr, ncols, nullptr);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(f, dout.ptr, bytes, hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

int raysht_angular_deriv_host(raysht_plan *p, int kind, const double *const *in_m, int nf_in, int fin,
                              double *const *out_m, int nf_out, int fout)
{
    if (!p || !in_m || !out_m || kind < 0 || kind > 2 || nf_in < 1 || nf_out < 1) return RAYSHT_EINVAL;
    if (p->nr < 1 || p->nf < 1) return RAYSHT_EUNSUPPORTED;
    bool same = nf_in == nf_out;
    for (int mp = 0; mp < p->nm; ++mp) {
        if (!in_m[mp] || !out_m[mp]) return RAYSHT_EINVAL;
        same = same && in_m[mp] == out_m[mp];
    }
    const long per_field = p->s_off[p->nm] / p->nf;  // doubles of one field over all m-blocks
    const long blk_field = 2L * p->nr;               // doubles of one field per l
    scoped_device_buffer din, dout;
    if (hipMalloc(&din.ptr, (size_t)per_field * nf_in * sizeof(double)) != hipSuccess) return RAYSHT_ENOMEM;
    if (!same && hipMalloc(&dout.ptr, (size_t)per_field * nf_out * sizeof(double)) != hipSuccess)
        return RAYSHT_ENOMEM;
    long off_in = 0, off_out = 0;
    for (int mp = 0; mp < p->nm; ++mp) {
        const long nl = p->lmax - p->mvals[mp] + 1;
        RAYSHT_HIP_TRY(hipMemcpy(din.ptr + off_in, in_m[mp], (size_t)nl * blk_field * nf_in * sizeof(double),
                                 hipMemcpyHostToDevice));
        // the other fields of a separate output keep their values
        if (!same && nf_out > 1)
            RAYSHT_HIP_TRY(hipMemcpy(dout.ptr + off_out, out_m[mp], (size_t)nl * blk_field * nf_out * sizeof(double),
                                     hipMemcpyHostToDevice));
        off_in += nl * blk_field * nf_in;
        off_out += nl * blk_field * nf_out;
    }
    double *out = same ? din.ptr : dout.ptr;
    const int rc = kind == 2 ? raysht_d_by_dphi(p, din.ptr, nf_in, fin, out, nf_out, fout, nullptr)
                             : raysht_d_by_dtheta(p, kind, din.ptr, nf_in, fin, out, nf_out, fout, nullptr);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipDeviceSynchronize());
    off_out = 0;
    for (int mp = 0; mp < p->nm; ++mp) {
        const long nl = p->lmax - p->mvals[mp] + 1;
        RAYSHT_HIP_TRY(hipMemcpy(out_m[mp], out + off_out, (size_t)nl * blk_field * nf_out * sizeof(double),
                                 hipMemcpyDeviceToHost));
        off_out += nl * blk_field * nf_out;
    }
    return RAYSHT_OK;
}

int raysht_cheb_deriv_host(raysht_cheb *cb, const double *c, double *d, long ncols, int dorder)
{
    if (!cb || ncols < 0 || (ncols > 0 && (!c || !d))) return RAYSHT_EINVAL;
    if (ncols == 0) return RAYSHT_OK;
    const size_t bytes = (size_t)ncols * cb->n * sizeof(double);
    scoped_device_buffer din, dout;
    if (hipMalloc(&din.ptr, bytes) != hipSuccess || hipMalloc(&dout.ptr, bytes) != hipSuccess) return RAYSHT_ENOMEM;
    RAYSHT_HIP_TRY(hipMemcpy(din.ptr, c, bytes, hipMemcpyHostToDevice));
    int rc = raysht_cheb_deriv(cb, din.ptr, dout.ptr, ncols, dorder, nullptr);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipMemcpy(d, dout.ptr, bytes, hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

// One matrix of the implicit solve from host arrays (synchronous), what the Fortran
// LU_Decompose_full / LU_Decompose_band call: kl < 0 is the dense path (mat n x n with leading
// dimension lda >= n), kl >= 0 the band path (mat in LAPACK band storage, lda = ldab).
int raysht_lu_decompose_host(int n, int kl, int ku, double *mat, int lda, int *pvt, int *info)
{
    const bool dense = kl < 0;
    if (n < 1 || !mat || !pvt || !info || lda < (dense ? n : 1)) return RAYSHT_EINVAL;  // band: lda = ldab
    const long rows = dense ? n : lda;  // device leading dimension
    scoped_device_buffer dm;
    scoped_device_bytes dp, di;
    if (hipMalloc(&dm.ptr, (size_t)rows * n * sizeof(double)) != hipSuccess) return RAYSHT_ENOMEM;
    if (hipMalloc(&dp.ptr, (size_t)n * sizeof(int)) != hipSuccess) return RAYSHT_ENOMEM;
    if (hipMalloc(&di.ptr, sizeof(int)) != hipSuccess) return RAYSHT_ENOMEM;
    RAYSHT_HIP_TRY(hipMemcpy2D(dm.ptr, rows * sizeof(double), mat, (size_t)lda * sizeof(double), rows * sizeof(double),
                               n, hipMemcpyHostToDevice));
    const int rc = dense ? raysht_lu_decompose_full(n, 1, dm.ptr, (int *)dp.ptr, (int *)di.ptr, nullptr)
                         : raysht_lu_decompose_band(n, kl, ku, 1, dm.ptr, lda, (int *)dp.ptr, (int *)di.ptr, nullptr);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipDeviceSynchronize());
    RAYSHT_HIP_TRY(hipMemcpy2D(mat, (size_t)lda * sizeof(double), dm.ptr, rows * sizeof(double), rows * sizeof(double),
                               n, hipMemcpyDeviceToHost));
    RAYSHT_HIP_TRY(hipMemcpy(pvt, dp.ptr, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    RAYSHT_HIP_TRY(hipMemcpy(info, di.ptr, sizeof(int), hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

// LU_Solve_full / LU_Solve_Band from host arrays (synchronous): nrhs columns of n at rhs with
// leading dimension ldb >= n, overwritten by the solutions; mat and kl as for the factorisation.
int raysht_lu_solve_host(int n, int kl, int ku, const double *mat, int lda, const int *pvt, double *rhs, int ldb,
                         int nrhs)
{
    const bool dense = kl < 0;
    if (n < 1 || nrhs < 0 || !mat || !pvt || !rhs || lda < (dense ? n : 1) || ldb < n) return RAYSHT_EINVAL;
    if (nrhs == 0) return RAYSHT_OK;
    const long rows = dense ? n : lda;
    scoped_device_buffer dm, db;
    scoped_device_bytes dp, doff, dn;
    if (hipMalloc(&dm.ptr, (size_t)rows * n * sizeof(double)) != hipSuccess) return RAYSHT_ENOMEM;
    if (hipMalloc(&db.ptr, (size_t)n * nrhs * sizeof(double)) != hipSuccess) return RAYSHT_ENOMEM;
    if (hipMalloc(&dp.ptr, (size_t)n * sizeof(int)) != hipSuccess) return RAYSHT_ENOMEM;
    if (hipMalloc(&doff.ptr, sizeof(long)) != hipSuccess) return RAYSHT_ENOMEM;
    if (hipMalloc(&dn.ptr, sizeof(int)) != hipSuccess) return RAYSHT_ENOMEM;
    const long zero = 0;
    RAYSHT_HIP_TRY(hipMemcpy2D(dm.ptr, rows * sizeof(double), mat, (size_t)lda * sizeof(double), rows * sizeof(double),
                               n, hipMemcpyHostToDevice));
    RAYSHT_HIP_TRY(hipMemcpy2D(db.ptr, (size_t)n * sizeof(double), rhs, (size_t)ldb * sizeof(double),
                               (size_t)n * sizeof(double), nrhs, hipMemcpyHostToDevice));
    RAYSHT_HIP_TRY(hipMemcpy(dp.ptr, pvt, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    RAYSHT_HIP_TRY(hipMemcpy(doff.ptr, &zero, sizeof(long), hipMemcpyHostToDevice));
    RAYSHT_HIP_TRY(hipMemcpy(dn.ptr, &nrhs, sizeof(int), hipMemcpyHostToDevice));
    const int rc = dense ? raysht_lu_solve_full(n, 1, dm.ptr, (const int *)dp.ptr, db.ptr, (const long *)doff.ptr,
                                                (const int *)dn.ptr, nrhs, nullptr)
                         : raysht_lu_solve_band(n, kl, ku, 1, dm.ptr, lda, (const int *)dp.ptr, db.ptr,
                                                (const long *)doff.ptr, (const int *)dn.ptr, nrhs, nullptr);
    if (rc != RAYSHT_OK) return rc;
    RAYSHT_HIP_TRY(hipDeviceSynchronize());
    RAYSHT_HIP_TRY(hipMemcpy2D(rhs, (size_t)ldb * sizeof(double), db.ptr, (size_t)n * sizeof(double),
                               (size_t)n * sizeof(double), nrhs, hipMemcpyDeviceToHost));
    return RAYSHT_OK;
}

}  // extern "C"
