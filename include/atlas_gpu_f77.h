/*
 * atlas_gpu_f77.h -- the Fortran 77 interface of the MI355X library's
 * level-3 path, callable from Fortran (gfortran convention: lower case,
 * trailing underscore, every argument by reference, hidden CHARACTER
 * lengths ignored) or from C.  Host arrays, column-major, synchronous, in
 * place; pivots are 1-based.  Illegal arguments call XERBLA (xerbla_) with
 * the routine's name and the reference INFO value.
 */
#ifndef ATLAS_GPU_F77_H
#define ATLAS_GPU_F77_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Replaces XERBLA (AtlasBase/Antoine/atlF77blas.base:115-140): prints
 * " ** On entry to NAME   parameter number  I had an illegal value" and
 * stops (exit status 0, as Fortran's STOP).  Weak; atlg_set_xerbla_hook
 * intercepts it. */
void xerbla_(const char *srname, const int *info, size_t len);

/* Replace DGEMM, DSYRK, DTRSM, DTRMM, DSYMM, DSYR2K
 * (AtlasBase/Antoine/atlF77blas.base:1585-2060 for the argument checks,
 * AtlasBase/Antoine/atlF77wrap.base:1026-1090 for DGEMM's SYRK route). */
void dgemm_(const char *transa, const char *transb, const int *m,
            const int *n, const int *k, const double *alpha, const double *a,
            const int *lda, const double *b, const int *ldb,
            const double *beta, double *c, const int *ldc);
void dsyrk_(const char *uplo, const char *trans, const int *n, const int *k,
            const double *alpha, const double *a, const int *lda,
            const double *beta, double *c, const int *ldc);
void dtrsm_(const char *side, const char *uplo, const char *transa,
            const char *diag, const int *m, const int *n, const double *alpha,
            const double *a, const int *lda, double *b, const int *ldb);
void dtrmm_(const char *side, const char *uplo, const char *transa,
            const char *diag, const int *m, const int *n, const double *alpha,
            const double *a, const int *lda, double *b, const int *ldb);
void dsymm_(const char *side, const char *uplo, const int *m, const int *n,
            const double *alpha, const double *a, const int *lda,
            const double *b, const int *ldb, const double *beta, double *c,
            const int *ldc);
void dsyr2k_(const char *uplo, const char *trans, const int *n, const int *k,
             const double *alpha, const double *a, const int *lda,
             const double *b, const int *ldb, const double *beta, double *c,
             const int *ldc);

/* Replace DGETRF (AtlasBase/Clint/atlas-fint.base:1616), DGETRS (:1395),
 * DGESV (:1518), DPOTRF (:1169), DPOTRS (:1289), DPOSV (:1044), DTRTRI
 * (:1703), DPOTRI (:1825), DLAUUM (:2038), and LAPACK's DLASWP (1-based
 * K1..K2 and IPIV, any nonzero INCX). */
void dgetrf_(const int *m, const int *n, double *a, const int *lda, int *ipiv,
             int *info);
void dgetrs_(const char *trans, const int *n, const int *nrhs, const double *a,
             const int *lda, const int *ipiv, double *b, const int *ldb,
             int *info);
void dgesv_(const int *n, const int *nrhs, double *a, const int *lda,
            int *ipiv, double *b, const int *ldb, int *info);
void dpotrf_(const char *uplo, const int *n, double *a, const int *lda,
             int *info);
void dpotrs_(const char *uplo, const int *n, const int *nrhs, const double *a,
             const int *lda, double *b, const int *ldb, int *info);
void dposv_(const char *uplo, const int *n, const int *nrhs, double *a,
            const int *lda, double *b, const int *ldb, int *info);
void dtrtri_(const char *uplo, const char *diag, const int *n, double *a,
             const int *lda, int *info);
void dpotri_(const char *uplo, const int *n, double *a, const int *lda,
             int *info);
void dlauum_(const char *uplo, const int *n, double *a, const int *lda,
             int *info);
void dlaswp_(const int *n, double *a, const int *lda, const int *k1,
             const int *k2, const int *ipiv, const int *incx);

#ifdef __cplusplus
}
#endif
#endif
