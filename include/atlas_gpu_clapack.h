/*
 * atlas_gpu_clapack.h -- CLAPACK entry points served by the MI355X library.
 * Prototypes follow ATLAS's clapack.h (AtlasBase/Clint/atlas-clp.base:611-671).
 * Host pointers, synchronous, in place.  ipiv holds 0-based absolute row
 * indices (ATLAS's C convention).  Return value: 0, -i for an illegal i-th
 * argument (after cblas_xerbla), or the LAPACK info.
 */
#ifndef ATLAS_GPU_CLAPACK_H
#define ATLAS_GPU_CLAPACK_H
#include "atlas_gpu_cblas.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Replaces clapack_dgetrf (atlas-clp.base:140-217): PA = LU (ColMajor) or
 * AP = LU with unit-diagonal U (RowMajor).  info = first exactly-zero pivot
 * (1-based); the factorization continues past it. */
int clapack_dgetrf(const enum CBLAS_ORDER Order, const int M, const int N,
                   double *A, const int lda, int *ipiv);

/* Replaces clapack_dgetrs (atlas-clp.base:218-265). */
int clapack_dgetrs(const enum CBLAS_ORDER Order,
                   const enum CBLAS_TRANSPOSE Trans, const int N,
                   const int NRHS, const double *A, const int lda,
                   const int *ipiv, double *B, const int ldb);

/* Replaces clapack_dgesv (atlas-clp.base:62-114): getrf, then getrs when
 * info == 0. */
int clapack_dgesv(const enum CBLAS_ORDER Order, const int N, const int NRHS,
                  double *A, const int lda, int *ipiv, double *B,
                  const int ldb);

/* Replaces clapack_dpotrf (atlas-clp.base:320-362): info = first
 * non-positive leading minor (1-based); stops there. */
int clapack_dpotrf(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const int N, double *A, const int lda);

/* Replaces clapack_dpotrs (atlas-clp.base:363-410). */
int clapack_dpotrs(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const int N, const int NRHS, const double *A,
                   const int lda, double *B, const int ldb);

/* Replaces clapack_dposv (atlas-clp.base:266-315). */
int clapack_dposv(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const int N, const int NRHS, double *A, const int lda,
                  double *B, const int ldb);

/* Replaces clapack_dtrtri (atlas-clp.base:411-452, with ATL_trtri at
 * atlas-lp.base:2341-2380): A <- inv(A) for the Uplo triangle; a zero on a
 * NonUnit diagonal returns its 1-based index and leaves A unchanged.  (The
 * reference wrapper reads its error code uninitialized; here the inverse is
 * computed exactly when every argument is legal.) */
int clapack_dtrtri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const enum CBLAS_DIAG Diag, const int N, double *A,
                   const int lda);

/* Replaces clapack_dpotri (atlas-clp.base:316-362): the inverse of the SPD
 * matrix whose Cholesky factor (clapack_dpotrf) A holds, in the Uplo
 * triangle; info > 0 when the factor has a zero diagonal entry. */
int clapack_dpotri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const int N, double *A, const int lda);

/* Replaces clapack_dlauum (atlas-clp.base:316-362, ATL_lauum at
 * atlas-lp.base:221-330): Lower A <- L^T*L, Upper A <- U*U^T, in the Uplo
 * triangle (ColMajor; RowMajor is the other triangle of the same memory). */
int clapack_dlauum(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const int N, double *A, const int lda);

#ifdef __cplusplus
}
#endif
#endif
