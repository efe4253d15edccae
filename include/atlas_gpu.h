/*
 * atlas_gpu.h -- device-pointer API of the MI355X library (libatlas_gpu.so).
 *
 * The cblas_/clapack_/ATL_ entry points (atlas_gpu_cblas.h,
 * atlas_gpu_clapack.h, atlas_gpu_atl.h) take host pointers, as ATLAS does.
 * The atlg_* calls below are the same operations on DEVICE pointers, queued
 * on a caller-supplied HIP stream and asynchronous with respect to the host.
 * They have no ATLAS counterpart: they are the layer the host entry points
 * are built on, exposed for callers that keep matrices resident in HBM
 * (timing harnesses, PyTorch bindings).  Matrices are column-major with
 * leading dimensions, as in ATLAS; enum arguments take ATLAS's values.
 *
 * Return codes: 0 on success; -i when the i-th argument is illegal (nothing
 * is queued); ATLG_EHIP when a HIP call failed (atlg_last_error() has the
 * message).  LAPACK info values are written to DEVICE memory (*dinfo), so
 * no call synchronizes.  ipiv arrays live in device memory as int32, 0-based
 * absolute row indices, as in ATLAS's C interface.
 *
 * Streams: `stream` is a hipStream_t passed as void*; NULL is HIP's null
 * (default) stream, as in other HIP libraries.  Workspace is owned by the
 * library per stream and reused across calls; calls on different streams
 * may run concurrently.  The host-pointer entry points use a stream of the
 * library's own.
 */
#ifndef ATLAS_GPU_H
#define ATLAS_GPU_H
#ifdef __cplusplus
extern "C" {
#endif

#define ATLG_EHIP (-1000)

typedef void *atlg_stream_t;

/* Creates the library's stream on the current HIP device (idempotent). */
int atlg_init(void);
/* Message of the last failed call on this thread (static storage). */
const char *atlg_last_error(void);
/* Makes cblas_xerbla call `hook` (with the formatted message) instead of
 * exiting; NULL restores ATLAS's print-and-exit behaviour. */
void atlg_set_xerbla_hook(void (*hook)(int p, const char *rout,
                                       const char *msg));

/* C <- alpha*op(A)*op(B) + beta*C; beta == 0 never reads C. */
int atlg_dgemm(int TA, int TB, long M, long N, long K, double alpha,
               const double *A, long lda, const double *B, long ldb,
               double beta, double *C, long ldc, atlg_stream_t stream);
/* C_uplo <- alpha*op(A)*op(A)^T + beta*C_uplo. */
int atlg_dsyrk(int Uplo, int Trans, long N, long K, double alpha,
               const double *A, long lda, double beta, double *C, long ldc,
               atlg_stream_t stream);
/* B <- alpha*op(A)^-1*B (Left) / alpha*B*op(A)^-1 (Right). */
int atlg_dtrsm(int Side, int Uplo, int TA, int Diag, long M, long N,
               double alpha, const double *A, long lda, double *B, long ldb,
               atlg_stream_t stream);
/* Row interchanges ipiv[K1..K2) applied in order (inci=1) or in reverse
 * (inci=-1) to N columns of A. */
int atlg_dlaswp(long N, double *A, long lda, int K1, int K2, const int *ipiv,
                int inci, atlg_stream_t stream);
/* PA = LU, column-major; *dinfo = first zero pivot (1-based) or 0. */
int atlg_dgetrf(long M, long N, double *A, long lda, int *ipiv, int *dinfo,
                atlg_stream_t stream);
/* Solves op(A)X = B with the factors of atlg_dgetrf. */
int atlg_dgetrs(int Trans, long N, long NRHS, const double *A, long lda,
                const int *ipiv, double *B, long ldb, atlg_stream_t stream);
/* A = LL^T (Lower) or U^T U (Upper); *dinfo = first non-positive minor. */
int atlg_dpotrf(int Uplo, long N, double *A, long lda, int *dinfo,
                atlg_stream_t stream);
/* Solves AX = B with the factor of atlg_dpotrf. */
int atlg_dpotrs(int Uplo, long N, long NRHS, const double *A, long lda,
                double *B, long ldb, atlg_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
