/*
 * atlas_gpu_atl.h -- ATLAS-internal ATL_d* entry points that ATLAS's own
 * testers (l3blastst, lutst, llttst) call, served by the MI355X library.
 * Prototypes follow gen include/atlas_level3.h:89-135 and
 * include/atlas_lapack.h:350-370 (AtlasBase/Clint/atlas.base:20992,
 * atlas-lp.base:2920-3100).  ATL_CSZT is ATLAS's const pointer-sized
 * unsigned integer (atlas.base:20554-20555, 23756-23762): unsigned long on
 * LP64.  Host pointers, synchronous, no argument checking (as in ATLAS).
 */
#ifndef ATLAS_GPU_ATL_H
#define ATLAS_GPU_ATL_H
#include "atlas_gpu_cblas.h"
#ifdef __cplusplus
extern "C" {
#endif

#ifndef ATL_CSZT
#define ATL_CSZT const unsigned long
#endif

/* replaces ATL_dgemm == ATL_dammm (atlas-ammm.base:17700-17810) */
void ATL_dgemm(const enum CBLAS_TRANSPOSE TA, const enum CBLAS_TRANSPOSE TB,
               ATL_CSZT M, ATL_CSZT N, ATL_CSZT K, const double alpha,
               const double *A, ATL_CSZT lda, const double *B, ATL_CSZT ldb,
               const double beta, double *C, ATL_CSZT ldc);
/* replaces ATL_dsyrk (AtlasBase/Clint/l3micro.base:2358-2720) */
void ATL_dsyrk(const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE Trans,
               ATL_CSZT N, ATL_CSZT K, const double alpha, const double *A,
               ATL_CSZT lda, const double beta, double *C, ATL_CSZT ldc);
/* replaces ATL_dtrsm (AtlasBase/Clint/trsm-micro.base:200-420) */
void ATL_dtrsm(const enum CBLAS_SIDE Side, const enum CBLAS_UPLO Uplo,
               const enum CBLAS_TRANSPOSE TA, const enum CBLAS_DIAG Diag,
               ATL_CSZT M, ATL_CSZT N, const double alpha, const double *A,
               ATL_CSZT lda, double *B, ATL_CSZT ldb);

/* replaces ATL_dgetrf (atlas-lp.base:1619-1630) */
int ATL_dgetrf(const enum CBLAS_ORDER Order, const int M, const int N,
               double *A, const int lda, int *ipiv);
/* replaces ATL_dgetrfC (atlas-lp.base:1497-1618) */
int ATL_dgetrfC(const int M, const int N, double *A, const int lda, int *ipiv);
/* replaces ATL_dgetrfR (atlas-lp.base:1210-1297) */
int ATL_dgetrfR(const int M, const int N, double *A, const int lda, int *ipiv);
/* replaces ATL_dgetrs (atlas-lp.base:1631-1720) */
void ATL_dgetrs(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE Trans,
                const int N, const int NRHS, const double *A, const int lda,
                const int *ipiv, double *B, const int ldb);
/* replaces ATL_dlaswp (atlas-lp.base:1053-1166) */
void ATL_dlaswp(const int N, double *A, const int lda0, const int K1,
                const int K2, const int *ipiv, const int inci);
/* replaces ATL_dpotrf (atlas-lp.base:332-360) */
int ATL_dpotrf(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
               const int N, double *A, const int lda);
/* replaces ATL_dpotrfL (atlas-lp.base:809-968) and ATL_dpotrfU (:675-733) */
int ATL_dpotrfL(const int N, double *A, const int lda);
int ATL_dpotrfU(const int N, double *A, const int lda);
/* replaces ATL_dpotrs (atlas-lp.base:969-1050) */
void ATL_dpotrs(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                const int N, const int NRHS, const double *A, const int lda,
                double *B, const int ldb);

/* replaces ATL_dtrmm, ATL_dsymm, ATL_dsyr2k (prototypes gen
 * include/atlas_level3.h from AtlasBase/Clint/atlas.base:21021-21048) */
void ATL_dtrmm(const enum CBLAS_SIDE Side, const enum CBLAS_UPLO Uplo,
               const enum CBLAS_TRANSPOSE TA, const enum CBLAS_DIAG Diag,
               ATL_CSZT M, ATL_CSZT N, const double alpha, const double *A,
               ATL_CSZT lda, double *B, ATL_CSZT ldb);
void ATL_dsymm(const enum CBLAS_SIDE Side, const enum CBLAS_UPLO Uplo,
               ATL_CSZT M, ATL_CSZT N, const double alpha, const double *A,
               ATL_CSZT lda, const double *B, ATL_CSZT ldb,
               const double beta, double *C, ATL_CSZT ldc);
void ATL_dsyr2k(const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE Trans,
                ATL_CSZT N, ATL_CSZT K, const double alpha, const double *A,
                ATL_CSZT lda, const double *B, ATL_CSZT ldb,
                const double beta, double *C, ATL_CSZT ldc);
/* replaces ATL_dtrtri (atlas-lp.base:2341-2380) */
int ATL_dtrtri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
               const enum CBLAS_DIAG Diag, const int N, double *A,
               const int lda);
/* replaces ATL_dlauum (atlas-lp.base:312-330) */
void ATL_dlauum(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                const int N, double *A, const int lda);
/* the body of clapack_dpotri (atlas-clp.base:356-357: ATL_trtri, then
 * ATL_lauum when it succeeded) */
int ATL_dpotri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
               const int N, double *A, const int lda);

#ifdef __cplusplus
}
#endif
#endif
