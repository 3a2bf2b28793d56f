/*
 * Checker-only link adapter for the compiled reference in oracle/_ref.
 *
 * The reference calls the Fortran BLAS symbol dgemm_ with 32-bit integers
 * (Legendre_Transforms.F90:1028,1058,1130,1150; Chebyshev_Polynomials.F90:415,
 * 432,490,524).  The only BLAS in this image is the OpenBLAS 0.3.29 that numpy
 * bundles, which exports the same routine as scipy_dgemm_64_ with 64-bit
 * integers.  This file widens the integer arguments and forwards the call;
 * it computes nothing itself.
 */
#include <stddef.h>
#include <stdint.h>

extern void scipy_dgemm_64_(const char *transa, const char *transb,
                            const int64_t *m, const int64_t *n, const int64_t *k,
                            const double *alpha, const double *a, const int64_t *lda,
                            const double *b, const int64_t *ldb, const double *beta,
                            double *c, const int64_t *ldc, size_t la, size_t lb);

void dgemm_(const char *transa, const char *transb, const int *m, const int *n,
            const int *k, const double *alpha, const double *a, const int *lda,
            const double *b, const int *ldb, const double *beta, double *c,
            const int *ldc, size_t la, size_t lb)
{
    int64_t M = *m, N = *n, K = *k, LDA = *lda, LDB = *ldb, LDC = *ldc;
    (void)la;
    (void)lb;
    scipy_dgemm_64_(transa, transb, &M, &N, &K, alpha, a, &LDA, b, &LDB, beta, c, &LDC, 1, 1);
}
