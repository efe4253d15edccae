/*
 * atlas_gpu_cblas.h -- CBLAS level-3 entry points served by the MI355X
 * library (libatlas_gpu.so).  Names, argument order, integer widths and enum
 * values are ATLAS's own, so a program built against ATLAS's cblas.h links
 * against this library unchanged.
 *
 * Calls take HOST pointers, are synchronous and in place, exactly as in
 * ATLAS; the library stages the operands through device memory.
 */
#ifndef ATLAS_GPU_CBLAS_H
#define ATLAS_GPU_CBLAS_H
#ifdef __cplusplus
extern "C" {
#endif

/* enum values: AtlasBase/Clint/atlas.base:19795+ (gen include/cblas.h) */
#ifndef CBLAS_ENUM_DEFINED_H
#define CBLAS_ENUM_DEFINED_H
enum CBLAS_ORDER {CblasRowMajor=101, CblasColMajor=102};
enum CBLAS_TRANSPOSE {CblasNoTrans=111, CblasTrans=112, CblasConjTrans=113,
                      AtlasConj=114};
enum CBLAS_UPLO {CblasUpper=121, CblasLower=122};
enum CBLAS_DIAG {CblasNonUnit=131, CblasUnit=132};
enum CBLAS_SIDE {CblasLeft=141, CblasRight=142};
#endif

/* Replaces cblas_xerbla (AtlasBase/Clint/cblas.base:42-59): prints
 * "Parameter %d to routine %s was incorrect" and the message to stderr, then
 * exit(-1).  Weak, so a tester may supply its own definition; see also
 * atlg_set_xerbla_hook in atlas_gpu.h. */
void cblas_xerbla(int p, const char *rout, const char *form, ...);

/* Replaces cblas_dgemm (AtlasBase/Clint/cblas.base:110-236):
 * C <- alpha*op(A)*op(B) + beta*C.  Same argument checks and messages; the
 * A==B, M==N, TA!=TB, lda==ldb, beta==0 case is routed to SYRK plus a
 * reflection, as in the reference (cblas.base:221-231). */
void cblas_dgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TA,
                 const enum CBLAS_TRANSPOSE TB, const int M, const int N,
                 const int K, const double alpha, const double *A,
                 const int lda, const double *B, const int ldb,
                 const double beta, double *C, const int ldc);

/* Replaces cblas_dtrsm (AtlasBase/Clint/cblas.base:587-693):
 * B <- alpha*op(A)^-1*B (Left) or alpha*B*op(A)^-1 (Right), in place. */
void cblas_dtrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TA,
                 const enum CBLAS_DIAG Diag, const int M, const int N,
                 const double alpha, const double *A, const int lda,
                 double *B, const int ldb);

/* Replaces cblas_dsyrk (AtlasBase/Clint/cblas.base:339-449):
 * C_uplo <- alpha*op(A)*op(A)^T + beta*C_uplo; the other triangle is not
 * referenced. */
void cblas_dsyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                 const double alpha, const double *A, const int lda,
                 const double beta, double *C, const int ldc);

/* Replaces cblas_dtrmm (AtlasBase/Clint/cblas.base:587-693, the wrapper it
 * shares with TRSM): B <- alpha*op(A)*B (Left) or alpha*B*op(A) (Right),
 * in place; same checks and parameter numbers as cblas_dtrsm. */
void cblas_dtrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TA,
                 const enum CBLAS_DIAG Diag, const int M, const int N,
                 const double alpha, const double *A, const int lda,
                 double *B, const int ldb);

/* Replaces cblas_dsymm (AtlasBase/Clint/cblas.base:237-338):
 * C <- alpha*A*B + beta*C (Left) or alpha*B*A + beta*C (Right), A symmetric
 * with only the Uplo triangle referenced. */
void cblas_dsymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N,
                 const double alpha, const double *A, const int lda,
                 const double *B, const int ldb, const double beta, double *C,
                 const int ldc);

/* Replaces cblas_dsyr2k (AtlasBase/Clint/cblas.base:450-586):
 * C_uplo <- alpha*(op(A)*op(B)^T + op(B)*op(A)^T) + beta*C_uplo, op(X) = X
 * (NoTrans) or X^T (Trans); the other triangle is not referenced. */
void cblas_dsyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                  const double alpha, const double *A, const int lda,
                  const double *B, const int ldb, const double beta, double *C,
                  const int ldc);

#ifdef __cplusplus
}
#endif
#endif
