// Internal declarations shared by the MI355X level-3 library sources.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

#include "atlas_gpu.h"

namespace atlg {

// ATLAS enum values (AtlasBase/Clint/atlas.base:19795+).
enum : int {
  kRowMajor = 101, kColMajor = 102,
  kNoTrans = 111, kTrans = 112, kConjTrans = 113,
  kUpper = 121, kLower = 122,
  kNonUnit = 131, kUnit = 132,
  kLeft = 141, kRight = 142
};

inline bool is_trans(int t) { return t == kTrans || t == kConjTrans; }

// Records a failure message for atlg_last_error (thread-local).
void set_error(const std::string& msg);

// The Fortran XERBLA's report: the hook of atlg_set_xerbla_hook when one is
// installed (p = INFO), else the message on stdout and exit(0), as the
// reference XERBLA's WRITE and STOP do.
void f77_xerbla(int p, const char* rout, const char* msg);

// Thrown inside drivers on a HIP failure; entry points convert it to
// ATLG_EHIP.
struct HipFailure {
  std::string msg;
};

#define ATLG_HIP(call)                                                         \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::atlg::HipFailure{std::string(#call) + ": " +                     \
                               hipGetErrorString(e_)};                         \
  } while (0)

#define ATLG_LAUNCHED()                                                        \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess)                                                      \
      throw ::atlg::HipFailure{std::string("kernel launch: ") +                \
                               hipGetErrorString(e_)};                         \
  } while (0)

// The caller's stream (NULL is HIP's null stream).
hipStream_t resolve_stream(atlg_stream_t s);
// The library's own stream on the current device, used by the host-pointer
// entry points (created on first use, default priority, blocking).
hipStream_t library_stream();

// A second library stream on the current device (non-blocking, default
// priority) and kAuxEvents events, for work that overlaps the caller's stream
// (getrf's look-ahead panel).  Created on first use.
hipStream_t aux_stream();
constexpr int kAuxEvents = 8;
hipEvent_t aux_event(int i);  // i = 0 .. kAuxEvents-1
// A pair of streams restricted to disjoint compute units: *small gets the
// first `cus` units, *large the rest (hipExtStreamCreateWithCUMask).  Made
// once per device and unit count; false when cus <= 0 or the masked streams
// are unavailable.
bool split_streams(int cus, hipStream_t* small, hipStream_t* large);
// Held by a look-ahead driver while it queues work on the shared auxiliary
// streams and events above, so that calls from several host threads do not
// interleave their event records and waits (the device work of successive
// calls then runs in queue order on those streams).
std::unique_lock<std::mutex> aux_lock();

// Per-stream scratch memory in named slots.  A slot grows (after a stream
// synchronize) and is reused by later calls on the same stream; pointers from
// different slots never alias.
enum Slot : int {
  kSlotTrsmInv = 0,   // inverted diagonal blocks for TRSM
  kSlotLaswp = 1,     // composed row lists for laswp
  kSlotPanel = 2,     // pivot-search partials and counters for getrf panels
  kSlotInfo = 3,      // device info words used by host entry points
  kSlotStageA = 4,    // host-API staging of operands
  kSlotStageB = 5,
  kSlotStageC = 6,
  kSlotIpiv = 7,
  kSlotFactInv = 8,   // diagonal-block inverses written by factorization leaves
  kSlotAdjoint = 9,   // work matrices of the Cholesky adjoint
  kSlotBatchCtr = 10,   // work-queue counter of the batched solvers
  kSlotBatchDesc = 11,  // per-system descriptors of the batched solvers
  kSlotBatchWork = 12,  // FP32 factors and right-hand sides of the mixed solver
  kSlotMixed = 13,      // FP32 factor, residual partials and norms of dsgesv
  kSlotBf16 = 14,       // BF16 operand copies of the low-precision update
  kSlotTrsv = 15,       // ready flags of the few-RHS triangular solve
  kSlotLaLists = 16,    // per-leaf interchange lists of the look-ahead LU
  kSlotLaPerm = 17,     // row maps of the look-ahead LU's deferred interchanges
  kSlotLaTemp = 18,     // staging of one column block for those interchanges
  kSlotTrmmBlk = 19,    // dense diagonal blocks of op(A) for TRMM leaves
  kSlotSymm = 20,       // dense symmetric copy of A for SYMM
  kSlotTriFull = 21,    // dense zero-padded copy of a triangle (TRTRI, LAUUM, POTRI)
  kSlotTriTemp = 22,    // products of the doubling inverse
  kSlotAdjP = 23,       // the potrf adjoint's product matrix
  kNumSlots = 24
};
void* scratch(hipStream_t s, Slot slot, size_t bytes);

// The stream's failure word: a device int that kernels whose bounded waits
// give up OR their ATLG_FAIL_* bit into (zero otherwise).  Allocated and
// cleared on first use per stream; read and cleared by take_status().
int* status_word(hipStream_t s);
// Copies the stream's failure word to the host, waits for the stream,
// clears the word and returns its value.
int take_status(hipStream_t s);
// Poll bound of the bounded waits: the process-wide value set by
// atlg_set_wait_limit, or `dflt` when none is set.
long wait_limit(long dflt);
// Reports a failure of a host entry point through the failure hook
// (atlg_set_failure_hook); the default hook prints and aborts.
void report_failure(const char* rout, int status, const char* msg);

// ---- kernels / drivers (device pointers, column-major) --------------------

// GEMM family (S = double on the FP64 MFMA, S = float on the FP32 MFMA).
// `mode`: 0 = full C, 1 = lower triangle only (SYRK), 2 = upper triangle
// only.  `skip` (device int, may be null): when *skip != 0 the launch does
// nothing.  C may alias B when M <= 128 and beta == 0, or alias A when
// N <= 128 and beta == 0 (used by TRSM's in-place diagonal apply).
// max_wg > 0 caps the grid (a multiple of 8 workgroups loop over the
// tiles), leaving room on the device for concurrent work.
template <class S>
void gemm(hipStream_t s, int TA, int TB, long M, long N, long K, S alpha, const S* A, long lda,
          const S* B, long ldb, S beta, S* C, long ldc, int mode = 0,
          const int* skip = nullptr, int max_wg = 0);
// C = alpha op(A) op(B) + beta C for K <= 128 and one short side: N <= 128
// with A not transposed, or M <= 128 with B not transposed, by 32-wide
// strips of the long side (4x the workgroups of the 128x128 tiling).  C must
// not overlap A or B.  tri: 0 all of C, 1 only C(i,j) with i >= j, 2 only
// i <= j.  `skip` as for gemm.  Returns false (launching nothing) when the
// shape does not fit.
template <class S>
bool gemm_narrow(hipStream_t s, int TA, int TB, long M, long N, long K, S alpha, const S* A,
                 long lda, const S* B, long ldb, S beta, S* C, long ldc, int tri = 0,
                 const int* skip = nullptr);
// FP64 register-path GEMM over `batch` problems whose operands sit at fixed
// element strides (sa, sb, sc), optionally restricted per 128x128 tile to
// the K range a triangular operand with a zero other triangle can reach
// (ktrim bit 1: B lower, k >= tile column; bit 2: A^T lower, k >= tile row;
// bit 4: A lower, k < tile row end).  Requires K a multiple of 32, aligned
// operands, even strides and no aliasing of C with A or B; returns false
// (launching nothing) when the register path does not apply.
bool gemm_batched(hipStream_t s, int TA, int TB, long M, long N, long K, double alpha,
                  const double* A, long lda, long sa, const double* B, long ldb, long sb,
                  double beta, double* C, long ldc, long sc, int batch, int mode, int ktrim,
                  const int* skip);
// Dense zero-padded triangles (order rounded up to 32, leading dimension
// that order): tri_copy_to writes op(A)'s triangle (transposed in when
// upper; 1 on a unit diagonal) as a lower triangle with zeros elsewhere;
// tri_invert_full inverts such a lower triangle in place (128-block
// inverses, then doubling levels of batched, K-trimmed products).
long tri_full_order(long n);
void tri_copy_to(hipStream_t s, bool lower, bool unit, long n, const double* A, long lda,
                 double* X, const int* skip);
void tri_invert_full(hipStream_t s, int diag, long n, double* X, long np, const int* skip);
// C <- beta*C on the M x N (or triangular per mode) part; beta==0 zeroes.
template <class S>
void scale(hipStream_t s, long M, long N, S beta, S* C, long ldc, int mode = 0,
           const int* skip = nullptr);

// C(i,j) <- C(j,i) for i > j (fills the lower triangle from the upper).
void syreflect_upper(hipStream_t s, long N, double* C, long ldc);

// TRSM on the matrix cores.  Diagonal blocks of order kTrsmNB are inverted
// in one launch unless `dinv` supplies them: block b (rows/cols
// [b*kTrsmNB, (b+1)*kTrsmNB) of op(A)) at dinv + b*kTrsmNB^2, column-major
// with leading dimension kTrsmNB, holding inv(op(A)_bb) or, when
// dinv_trans != 0, its transpose.
constexpr int kTrsmNB = 128;
// Writes inv(op(A)_bb) for every kTrsmNB block of the order-na triangle in
// the layout trsm() takes as `dinv` (one launch, one workgroup per block).
template <class S>
void tri_inverse_blocks(hipStream_t s, long na, const S* A, long lda, int ta, int uplo, int diag,
                        S* Dinv, const int* skip = nullptr);
template <class S>
void trsm(hipStream_t s, int side, int uplo, int ta, int diag, long M, long N, S alpha,
          const S* A, long lda, S* B, long ldb, const int* skip = nullptr,
          const S* dinv = nullptr, int dinv_trans = 0);

// trsm() without allocation: the diagonal-block inverses live in a
// file-scope device array sized for op(A) of order <= kNoAllocMaxOrder (one
// copy per device, shared by every call on it, so calls that overlap on one
// device race); the solve is the blocked recursion.  noalloc_scratch()
// resolves the array's address on the current device (once per device).
constexpr long kNoAllocMaxOrder = 8192;
constexpr size_t kNoAllocBytes =
    (size_t)(kNoAllocMaxOrder / kTrsmNB) * kTrsmNB * kTrsmNB * sizeof(double);
template <class S>
void trsm_noalloc(hipStream_t s, int side, int uplo, int ta, int diag, long M, long N, S alpha,
                  const S* A, long lda, S* B, long ldb);
void* noalloc_scratch();

template <class S>
void laswp(hipStream_t s, long N, S* A, long lda, int K1, int K2, const int* ipiv, int inci);

// The rest of the real level-3 BLAS and the triangular / SPD inverses
// (blas3x.hip).  trmm: B <- alpha op(A) B (Left) or alpha B op(A) (Right);
// symm: C <- alpha A B + beta C (Left) or alpha B A + beta C (Right), A
// symmetric with only `uplo` stored; syr2k: C_uplo <- alpha (A B^T + B A^T)
// + beta C_uplo (NoTrans) or alpha (A^T B + B^T A) + beta C_uplo (Trans);
// trtri: A <- inv(A) in place (triangle only), *dinfo = first zero
// diagonal (1-based, NonUnit) and then A is untouched; lauum: lower
// A <- L^T L, upper A <- U U^T (triangle only); potri: the SPD inverse
// from the Cholesky factor, trtri then lauum.  `skip` (device word, may
// be null): nonzero makes every launch a no-op.
void trmm(hipStream_t s, int side, int uplo, int ta, int diag, long M, long N, double alpha,
          const double* A, long lda, double* B, long ldb, const int* skip = nullptr);
void symm(hipStream_t s, int side, int uplo, long M, long N, double alpha, const double* A,
          long lda, const double* B, long ldb, double beta, double* C, long ldc);
void syr2k(hipStream_t s, int uplo, int trans, long N, long K, double alpha, const double* A,
           long lda, const double* B, long ldb, double beta, double* C, long ldc);
void trtri(hipStream_t s, int uplo, int diag, long N, double* A, long lda, int* dinfo);
void lauum(hipStream_t s, int uplo, long N, double* A, long lda, const int* skip = nullptr);
void potri(hipStream_t s, int uplo, long N, double* A, long lda, int* dinfo);

// Precision of getrf's trailing-update GEMMs: the scalar type's own MFMA,
// or (float only) BF16 operands with FP32 accumulation, single or split
// (hi + mid + lo, six products: FP32 accuracy; hi + lo, three products:
// about 2^-15) operands.
enum GemmPrec : int { kGemmNative = 0, kGemmBf16 = 1, kGemmBf16x6 = 2, kGemmBf16x3 = 3 };
template <class S>
void getrf(hipStream_t s, long M, long N, S* A, long lda, int* ipiv, int* dinfo,
           int gemm_prec = kGemmNative);
template <class S>
void getrs(hipStream_t s, int trans, long N, long NRHS, const S* A, long lda, const int* ipiv,
           S* B, long ldb);
void potrf(hipStream_t s, int uplo, long N, double* A, long lda, int* dinfo);
void potrs(hipStream_t s, int uplo, long N, long NRHS, const double* A,
           long lda, double* B, long ldb);

// Adjoints of the Cholesky path (lower factor L, K = L L^T):
//   potrf_adjoint: G holds Lbar (lower triangle read) and receives
//     Kbar = sym(L^-T Phi(L^T Lbar) L^-1) as a full symmetric matrix;
//   trsm_adjoint: for X = L^-1 B, overwrites Xbar with Bbar = L^-T Xbar and
//     adds -tril(Bbar X^T) to Lbar (when Lbar is not null);
//   logdet_chol: *out = 2 sum log L_ii; its adjoint adds 2 dbar / L_ii to
//     the diagonal of Lbar.
void potrf_adjoint(hipStream_t s, long N, const double* L, long ldl, double* G, long ldg);
void trsm_adjoint(hipStream_t s, long N, long NRHS, const double* L, long ldl, const double* X,
                  long ldx, double* Xbar, long ldxb, double* Lbar, long ldlb);
void logdet_chol(hipStream_t s, long N, const double* L, long ldl, double* out);
void logdet_chol_adjoint(hipStream_t s, long N, const double* L, long ldl, const double* dbar,
                         double* Lbar, long ldlb);

// Batched SPD solvers (lower triangle; one persistent launch).  Host arrays
// n/lda/ldb and host arrays of device pointers A/B; `info` is a device array
// of nsys ints.  mode: 0 potrf, 1 potrf + potrs, 2 potrs with a given
// factor.  Bsrc (may be null) is copied into B (same ldb) before the solve.
// `orig` (may be null) maps list position -> index into info.
struct BatchTune {
  int t16_max;    // >= 0: systems with n <= t16_max use 16-wide block columns, 32 otherwise;
                  // < 0: the generated decision tree chooses per system
  int wg_per_cu;  // persistent workgroups per compute unit
};
BatchTune batch_tune(bool single);
void set_batch_tune(bool single, const BatchTune& t);
template <class S>
void posv_batched(hipStream_t s, int mode, int nsys, const int* n, S* const* A, const int* lda,
                  S* const* B, const int* ldb, const S* const* Bsrc, int nrhs, int* info,
                  const BatchTune& tune, const int* orig = nullptr);
// FP32 factor + FP64 refinement (LAPACK dsposv semantics per system); iter is
// a device array.  Synchronizes the stream (the fallback set is decided on
// the host).
void dsposv_batched(hipStream_t s, int nsys, const int* n, double* const* A, const int* lda,
                    const double* const* B, const int* ldb, double* const* X, const int* ldx,
                    int nrhs, int* info, int* iter, int itermax, const BatchTune& tune);

// C(f32) = alpha * bf16(A) * bf16(B) + beta * C, A M x K and B K x N
// column-major FP32 (NoTrans, NoTrans), FP32 accumulation.  terms = 1: one
// BF16 term per operand; 3: hi + mid + lo, the six products above FP32
// resolution (FP32 accuracy); 2: hi + lo, the three products hi*hi, hi*lo,
// lo*hi (about 2^-15 relative error).
void gemm_bf16(hipStream_t s, long M, long N, long K, float alpha, const float* A, long lda,
               const float* B, long ldb, float beta, float* C, long ldc, int terms = 1);

// Mixed-precision dgesv (LAPACK dsgesv semantics): FP32 LU (lowprec 1: BF16
// trailing updates, 2: three-term split BF16, 3: two-term split BF16), FP64
// residual refinement (refine 0:
// classical, 1: GMRES-IR), FP64 fallback.  *iter_out receives ITER,
// *inner_out (if not null) the GMRES iterations.  Synchronizes the stream.
int dsgesv(hipStream_t s, long N, long NRHS, double* A, long lda, int* ipiv, const double* B,
           long ldb, double* X, long ldx, int* dinfo, int itermax, int lowprec, int refine,
           int* iter_out, int* inner_out);

// zero an int on device (kernel, graph-capture safe)
void set_int(hipStream_t s, int* p, int v);

}  // namespace atlg
