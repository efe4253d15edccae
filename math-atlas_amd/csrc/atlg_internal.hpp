// Internal declarations shared by the MI355X level-3 library sources.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
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
  kNumSlots = 9
};
void* scratch(hipStream_t s, Slot slot, size_t bytes);

// ---- kernels / drivers (device pointers, column-major) --------------------

// GEMM family.  `mode`: 0 = full C, 1 = lower triangle only (SYRK),
// 2 = upper triangle only.  `skip` (device int, may be null): when *skip != 0
// the launch does nothing.  C may alias B when M <= 128 and beta == 0, or
// alias A when N <= 128 and beta == 0 (used by TRSM's in-place diagonal
// apply).
void gemm(hipStream_t s, int TA, int TB, long M, long N, long K, double alpha,
          const double* A, long lda, const double* B, long ldb, double beta,
          double* C, long ldc, int mode = 0, const int* skip = nullptr);
// C <- beta*C on the M x N (or triangular per mode) part; beta==0 zeroes.
void scale(hipStream_t s, long M, long N, double beta, double* C, long ldc,
           int mode = 0, const int* skip = nullptr);

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
void tri_inverse_blocks(hipStream_t s, long na, const double* A, long lda, int ta, int uplo,
                        int diag, double* Dinv, const int* skip = nullptr);
void trsm(hipStream_t s, int side, int uplo, int ta, int diag, long M, long N,
          double alpha, const double* A, long lda, double* B, long ldb,
          const int* skip = nullptr, const double* dinv = nullptr,
          int dinv_trans = 0);

void laswp(hipStream_t s, long N, double* A, long lda, int K1, int K2,
           const int* ipiv, int inci);


void getrf(hipStream_t s, long M, long N, double* A, long lda, int* ipiv,
           int* dinfo);
void getrs(hipStream_t s, int trans, long N, long NRHS, const double* A,
           long lda, const int* ipiv, double* B, long ldb);
void potrf(hipStream_t s, int uplo, long N, double* A, long lda, int* dinfo);
void potrs(hipStream_t s, int uplo, long N, long NRHS, const double* A,
           long lda, double* B, long ldb);

// zero an int on device (kernel, graph-capture safe)
void set_int(hipStream_t s, int* p, int v);

}  // namespace atlg
