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

/* ---- Failure reporting of the bounded waits (no ATLAS counterpart) ----
 * Two kernels wait on other workgroups of their own launch: the getrf
 * panel (pivot-candidate mailboxes) and the few-RHS triangular solve
 * (per-block ready flags).  Every such wait is bounded; a wait that runs
 * past its bound makes the kernel exit early and ORs one of the bits below
 * into the stream's failure word (the results of that call are then
 * unspecified; getrf also writes *dinfo = INT_MIN and identity pivots for
 * the affected panel, so later launches stay in bounds).
 * atlg_stream_status: synchronizes `stream`, stores its failure word in
 *   *status (0: no failure) and clears it.
 * atlg_set_failure_hook: the host-pointer entry points (cblas_, clapack_,
 *   ATL_) check the failure word of their stream before returning and
 *   report a failure through hook(routine, bits, message); the default
 *   (NULL) prints the message and aborts, so no host call returns a partial
 *   result silently.  If a hook returns, getrf/potrf/gesv/posv return
 *   INT_MIN and the outputs are unspecified.  atlg_dsgesv (which
 *   synchronizes) returns ATLG_EWAIT instead.
 * atlg_set_wait_limit: polls (each one memory round trip) before a wait
 *   gives up, process-wide; polls <= 0 restores the defaults (2^20 for the
 *   panel, 2^22 for the solve, far beyond any wait of a resident grid). */
#define ATLG_EWAIT (-1001)
#define ATLG_FAIL_PANEL 1
#define ATLG_FAIL_TRSV 2
int atlg_stream_status(atlg_stream_t stream, int *status);
void atlg_set_failure_hook(void (*hook)(const char *rout, int bits,
                                        const char *msg));
void atlg_set_wait_limit(long polls);

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
/* atlg_dtrsm without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and results as atlg_dtrsm, except that the order of op(A)
 * (M for Left, N for Right) may be at most 8192: larger orders return
 * ATLG_ESIZE (atlg_last_error() says why) and queue nothing.  Its workspace
 * (the inverses of the 128 x 128 diagonal blocks, 8 MiB) is a static device
 * array, one per GPU, held until the process ends and shared by every call
 * on that GPU: calls that overlap on one GPU (different streams, or host
 * threads) race and must be serialized by the caller.  It never calls
 * hipMalloc or hipFree (atlg_init resolves the array's address on the
 * current device; another device's on its first call).  Few right-hand
 * sides take the blocked solve here, not atlg_dtrsm's sync-free one, so
 * results agree with atlg_dtrsm to rounding, bit for bit when NRHS > 4 or
 * the order is below 1024. */
#define ATLG_ESIZE (-1002)
int atlg_dtrsm_noalloc(int Side, int Uplo, int TA, int Diag, long M, long N,
                       double alpha, const double *A, long lda, double *B,
                       long ldb, atlg_stream_t stream);
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

/* The rest of the real level-3 BLAS and the triangular / SPD inverses,
 * column-major on device pointers (same operations as cblas_dtrmm,
 * cblas_dsymm, cblas_dsyr2k, ATL_dtrtri, ATL_dlauum and clapack_dpotri).
 * atlg_dtrtri / atlg_dpotri write info to *dinfo (device): the 1-based
 * index of the first zero on the diagonal of the triangle (A then
 * untouched), else 0.  From order 256 on, while a dense copy of the
 * triangle (order rounded up to 32) fits 4 GiB, trtri / lauum / potri work
 * on such a copy in the library's per-stream scratch (allocated on first
 * use and kept); larger orders take the in-place recursions. */
int atlg_dtrmm(int Side, int Uplo, int TA, int Diag, long M, long N,
               double alpha, const double *A, long lda, double *B, long ldb,
               atlg_stream_t stream);
int atlg_dsymm(int Side, int Uplo, long M, long N, double alpha,
               const double *A, long lda, const double *B, long ldb,
               double beta, double *C, long ldc, atlg_stream_t stream);
int atlg_dsyr2k(int Uplo, int Trans, long N, long K, double alpha,
                const double *A, long lda, const double *B, long ldb,
                double beta, double *C, long ldc, atlg_stream_t stream);
int atlg_dtrtri(int Uplo, int Diag, long N, double *A, long lda, int *dinfo,
                atlg_stream_t stream);
int atlg_dlauum(int Uplo, long N, double *A, long lda, atlg_stream_t stream);
int atlg_dpotri(int Uplo, long N, double *A, long lda, int *dinfo,
                atlg_stream_t stream);

/* ---- Adjoints of the Cholesky path (no ATLAS counterpart) -------------
 * For K = L L^T with L lower (as atlg_dpotrf with Lower leaves it):
 * atlg_dpotrf_adjoint: on entry G holds Lbar (lower triangle read, upper
 *   ignored); on exit G holds the symmetric Kbar = sym(L^-T Phi(L^T Lbar)
 *   L^-1), Phi = lower triangle with halved diagonal.
 * atlg_dtrsm_adjoint: for X = L^-1 B (N x NRHS): Xbar is overwritten by
 *   Bbar = L^-T Xbar, and -tril(Bbar X^T) is added to Lbar (if not NULL).
 * atlg_dlogdet_chol: *out (device) = log det K = 2 sum log L_ii.
 * atlg_dlogdet_chol_adjoint: adds 2 (*dbar) / L_ii to the diagonal of Lbar
 *   (dbar in device memory). */
int atlg_dpotrf_adjoint(long N, const double *L, long ldl, double *G, long ldg,
                        atlg_stream_t stream);
int atlg_dtrsm_adjoint(long N, long NRHS, const double *L, long ldl,
                       const double *X, long ldx, double *Xbar, long ldxb,
                       double *Lbar, long ldlb, atlg_stream_t stream);
int atlg_dlogdet_chol(long N, const double *L, long ldl, double *out,
                      atlg_stream_t stream);
int atlg_dlogdet_chol_adjoint(long N, const double *L, long ldl,
                              const double *dbar, double *Lbar, long ldlb,
                              atlg_stream_t stream);

/* ---- Single precision and the mixed-precision solver (the FP32 path is
 * the ATL_s instance of the same ATLAS routines: ATL_sgemm, ATL_sgetrf,
 * ATL_sgetrs; AtlasBase/Clint/atlas-lp.base:1497-1720) ------------------
 * atlg_sgemm: as atlg_dgemm in FP32 (FP32 MFMA).
 * atlg_sgemm_bf16: C = alpha*bf16(A)*bf16(B) + beta*C, NoTrans/NoTrans,
 *   operands rounded to bfloat16 (nearest even), FP32 accumulation and C;
 *   split = 1: each operand as hi + mid + lo BF16 terms (all 24 bits),
 *   the six products above FP32 resolution (six BF16 MFMAs, FP32
 *   accuracy); split = 2: hi + lo (16 bits), the three products hi*hi,
 *   hi*lo, lo*hi (about 2^-15 relative error); other values: -12.
 * atlg_sgetrf: as atlg_dgetrf in FP32; gemm_prec 0 = FP32 trailing updates,
 *   1 = BF16 trailing updates with FP32 accumulation, 2 = split BF16 (x6),
 *   3 = two-term split BF16 (x3); with 1-3 the updates of depth k >= 256
 *   take the BF16 form, the narrower ones inside panels stay FP32.
 * atlg_sgetrs: as atlg_dgetrs in FP32.
 * atlg_dsgesv: LAPACK dsgesv semantics.  Factors float(A) with atlg_sgetrf
 *   (gemm_prec = lowprec: 0 FP32, 1 BF16, 2 split BF16, 3 two-term split BF16), refines X with
 *   FP64 residuals B - A X until
 *   ||r_j||_inf < ||x_j||_inf ||A||_inf eps sqrt(N) for every column j (at
 *   most itermax corrections; itermax <= 0 means 30).  *iter (host) = the
 *   corrections used, or -2 (A overflows FP32), -3 (FP32 factor singular),
 *   -(itermax+1) (no convergence); when negative, A holds the FP64 LU
 *   factors and X the FP64 dgetrs solution.  ipiv holds the pivots of
 *   whichever factor was used.  refine 0: classical refinement (LAPACK),
 *   x += solve(r); refine 1: GMRES-IR (Carson and Higham 2018), each
 *   correction solves A d = r by FP64 GMRES left-preconditioned with the
 *   low-precision factor (at most 30 iterations, tolerance 1e-6), which converges for
 *   condition numbers well beyond 1/u of the factor; it stops early, with
 *   ITER = -(itermax+1) and the FP64 fallback, when two corrections in a
 *   row fail to halve the residual; *inner (host, may be NULL) receives
 *   the GMRES iterations.  Synchronizes the stream. */
int atlg_sgemm(int TA, int TB, long M, long N, long K, float alpha,
               const float *A, long lda, const float *B, long ldb, float beta,
               float *C, long ldc, atlg_stream_t stream);
int atlg_sgemm_bf16(long M, long N, long K, float alpha, const float *A,
                    long lda, const float *B, long ldb, float beta, float *C,
                    long ldc, int split, atlg_stream_t stream);
int atlg_sgetrf(long M, long N, float *A, long lda, int *ipiv, int *dinfo,
                int gemm_prec, atlg_stream_t stream);
int atlg_sgetrs(int Trans, long N, long NRHS, const float *A, long lda,
                const int *ipiv, float *B, long ldb, atlg_stream_t stream);
int atlg_dsgesv(long N, long NRHS, double *A, long lda, int *ipiv,
                const double *B, long ldb, double *X, long ldx, int *dinfo,
                int itermax, int lowprec, int refine, int *iter, int *inner,
                atlg_stream_t stream);

/* ---- Batched SPD solvers (no ATLAS counterpart; per system the same
 * lower Cholesky as ATL_dpotrf and the two solves of ATL_dpotrs,
 * AtlasBase/Clint/atlas-lp.base:332-360, 969-1050) ----------------------
 * nsys independent systems of uneven order in ONE persistent launch.
 * n, lda, ldb, ldx: host arrays of nsys ints; A, B, X: host arrays of nsys
 * device pointers (column-major, lower triangle of A read and written);
 * info, iter: device arrays of nsys ints.  Systems are taken longest first
 * from a device work queue.  Per system, info = 0 or the order of the first
 * leading minor that is not positive definite (that system is not solved).
 * Returns -(argument position) for an illegal argument (n[k] < 0,
 * lda[k] < max(1,n[k]), ...), or ATLG_EHIP. */
int atlg_dpotrf_batched(int nsys, const int *n, double *const *A,
                        const int *lda, int *info, atlg_stream_t stream);
int atlg_dpotrs_batched(int nsys, const int *n, double *const *A,
                        const int *lda, double *const *B, const int *ldb,
                        int nrhs, atlg_stream_t stream);
int atlg_dposv_batched(int nsys, const int *n, double *const *A,
                       const int *lda, double *const *B, const int *ldb,
                       int nrhs, int *info, atlg_stream_t stream);
int atlg_spotrf_batched(int nsys, const int *n, float *const *A,
                        const int *lda, int *info, atlg_stream_t stream);
int atlg_sposv_batched(int nsys, const int *n, float *const *A,
                       const int *lda, float *const *B, const int *ldb,
                       int nrhs, int *info, atlg_stream_t stream);
/* Mixed precision (LAPACK dsposv semantics per system): the lower triangle
 * of the FP64 A is factored in FP32, X is refined with FP64 residuals
 * b - A x until ||r||_inf < ||x||_inf ||A||_inf eps sqrt(n) for every
 * right-hand side (at most itermax corrections; itermax <= 0 means 30).
 * iter[k] = corrections used (>= 0; A and B untouched), or -2 (A overflows
 * FP32), -3 (FP32 factor failed), -(itermax+1) (no convergence): those
 * systems are factored and solved in FP64, A then holding the FP64 factor.
 * Synchronizes the stream before returning. */
int atlg_dsposv_batched(int nsys, const int *n, double *const *A,
                        const int *lda, const double *const *B,
                        const int *ldb, double *const *X, const int *ldx,
                        int nrhs, int itermax, int *info, int *iter,
                        atlg_stream_t stream);
/* Tuning of the batched solvers (prec 0 = FP64, 1 = FP32): with t16_max
 * >= 0, systems with n <= t16_max use 16-wide block columns and wider ones
 * 32; with t16_max < 0 (the default) the decision tree generated by the
 * empirical search picks the width per system; wg_per_cu persistent
 * workgroups per compute unit.  The defaults come from the search's
 * generated header; set overrides them for this process (the search itself
 * times candidates through it). */
int atlg_batched_tune_get(int prec, int *t16_max, int *wg_per_cu);
int atlg_batched_tune_set(int prec, int t16_max, int wg_per_cu);

#ifdef __cplusplus
}
#endif
#endif
