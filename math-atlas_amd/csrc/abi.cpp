// Host-pointer C ABI: ATLAS's cblas_/clapack_/ATL_ entry points.
//
// Each call checks its arguments exactly as the reference wrapper does
// (same parameter numbers and messages), stages the host operands into
// device memory, runs the device driver on the library stream and copies the
// results back before returning, so the call is synchronous and in place
// like ATLAS's.  Calls are serialized by a process-wide mutex (ATLAS's serial
// routines are reentrant; this keeps the shared staging buffers safe).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <type_traits>

#include "atlas_gpu_atl.h"
#include "atlas_gpu_clapack.h"
#include "atlg_internal.hpp"

using namespace atlg;

namespace {

std::mutex g_api_mu;
void (*g_xerbla_hook)(int, const char*, const char*) = nullptr;

// cblas_errprn (gen interfaces/blas/C/src/cblas_errprn.c): print the message,
// return min(ierr, info).
int errprn(int ierr, int info, const char* form, ...) {
  va_list ap;
  va_start(ap, form);
  vfprintf(stderr, form, ap);
  va_end(ap);
  return std::min(ierr, info);
}

void fail_hip(const HipFailure& h, const char* rout) {
  set_error(h.msg);
  fprintf(stderr, "%s: HIP failure: %s\n", rout, h.msg.c_str());
  abort();  // never return a silent partial result
}

// A host column-major block (rows x cols, leading dimension ld) mirrored in a
// device slot with the same leading dimension.
// Host operands go in as one linear copy of the span they occupy (the gaps
// between columns come along; they lie inside the caller's array and the
// device copy ignores them): a pitched copy from pageable memory moves one
// column per transfer and was several times slower (M=N=K=1000: 35 ms for the
// call).  Results come back linearly when the columns are contiguous, else
// pitched, so the caller's gaps are never written.
struct Stage {
  double* d = nullptr;
  void in(hipStream_t s, Slot slot, const double* h, long rows, long cols, long ld, bool copy) {
    const size_t n = (rows > 0 && cols > 0) ? (size_t)ld * (cols - 1) + rows : 0;
    d = static_cast<double*>(scratch(s, slot, std::max<size_t>(n, 1) * sizeof(double)));
    if (copy && n)
      ATLG_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, s));
  }
  void out(hipStream_t s, double* h, long rows, long cols, long ld) const {
    if (rows <= 0 || cols <= 0) return;
    if (ld == rows || cols == 1)
      ATLG_HIP(hipMemcpyAsync(h, d, ((size_t)ld * (cols - 1) + rows) * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    else
      ATLG_HIP(hipMemcpy2DAsync(h, ld * sizeof(double), d, ld * sizeof(double),
                                rows * sizeof(double), cols, hipMemcpyDeviceToHost, s));
  }
};

hipStream_t lib_stream() { return library_stream(); }

// ---- host drivers on column-major data ------------------------------------

void h_gemm(int TA, int TB, long M, long N, long K, double alpha, const double* A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc) {
  if (M == 0 || N == 0) return;
  hipStream_t s = lib_stream();
  const bool at = is_trans(TA), bt = is_trans(TB);
  Stage a, b, c;
  const bool need_ab = (K > 0 && alpha != 0.0);
  if (need_ab) {
    a.in(s, kSlotStageA, A, at ? K : M, at ? M : K, lda, true);
    b.in(s, kSlotStageB, B, bt ? N : K, bt ? K : N, ldb, true);
  }
  c.in(s, kSlotStageC, C, M, N, ldc, beta != 0.0);
  gemm(s, TA, TB, M, N, K, alpha, a.d, lda, b.d, ldb, beta, c.d, ldc);
  c.out(s, C, M, N, ldc);
  ATLG_HIP(hipStreamSynchronize(s));
}

void h_syrk(int Uplo, int Trans, long N, long K, double alpha, const double* A, long lda,
            double beta, double* C, long ldc, bool reflect) {
  if (N == 0) return;
  hipStream_t s = lib_stream();
  const int mode = (Uplo == kLower) ? 1 : 2;
  Stage a, c;
  c.in(s, kSlotStageC, C, N, N, ldc, beta != 0.0 || !reflect);
  if (K > 0 && alpha != 0.0) {
    const bool tr = is_trans(Trans);
    a.in(s, kSlotStageA, A, tr ? K : N, tr ? N : K, lda, true);
    if (tr) gemm(s, kTrans, kNoTrans, N, N, K, alpha, a.d, lda, a.d, lda, beta, c.d, ldc, mode);
    else gemm(s, kNoTrans, kTrans, N, N, K, alpha, a.d, lda, a.d, lda, beta, c.d, ldc, mode);
  } else if (beta != 1.0) {
    scale(s, N, N, beta, c.d, ldc, mode);
  }
  if (reflect) syreflect_upper(s, N, c.d, ldc);
  c.out(s, C, N, N, ldc);
  ATLG_HIP(hipStreamSynchronize(s));
}

void h_trsm(int Side, int Uplo, int TA, int Diag, long M, long N, double alpha, const double* A,
            long lda, double* B, long ldb) {
  if (M == 0 || N == 0) return;
  hipStream_t s = lib_stream();
  const long na = (Side == kLeft) ? M : N;
  Stage a, b;
  a.in(s, kSlotStageA, A, na, na, lda, true);
  b.in(s, kSlotStageB, B, M, N, ldb, alpha != 0.0);
  trsm(s, Side, Uplo, TA, Diag, M, N, alpha, a.d, lda, b.d, ldb);
  b.out(s, B, M, N, ldb);
  ATLG_HIP(hipStreamSynchronize(s));
}

int* info_word(hipStream_t s) {
  return static_cast<int*>(scratch(s, kSlotInfo, 64));
}

int h_getrfC(long M, long N, double* A, long lda, int* ipiv) {
  const long MN = std::min(M, N);
  if (MN == 0) return 0;
  hipStream_t s = lib_stream();
  Stage a;
  a.in(s, kSlotStageA, A, M, N, lda, true);
  int* dpiv = static_cast<int*>(scratch(s, kSlotIpiv, sizeof(int) * MN));
  int* dinfo = info_word(s);
  getrf(s, M, N, a.d, lda, dpiv, dinfo);
  a.out(s, A, M, N, lda);
  int info = 0;
  ATLG_HIP(hipMemcpyAsync(ipiv, dpiv, sizeof(int) * MN, hipMemcpyDeviceToHost, s));
  ATLG_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  ATLG_HIP(hipStreamSynchronize(s));
  return info;
}

void h_laswp(long N, double* A, long lda, int K1, int K2, const int* ipiv, int inci) {
  if (N <= 0 || K2 <= K1) return;
  hipStream_t s = lib_stream();
  int maxrow = 0;
  for (int i = K1; i < K2; i++) {
    if (ipiv[i] < 0) {
      fprintf(stderr, "ATL_dlaswp: negative pivot index ipiv[%d] = %d\n", i, ipiv[i]);
      abort();
    }
    maxrow = std::max(maxrow, std::max(i, ipiv[i]));
  }
  const long rows = maxrow + 1;
  Stage a;
  a.in(s, kSlotStageA, A, rows, N, lda, true);
  int* dpiv = static_cast<int*>(scratch(s, kSlotIpiv, sizeof(int) * K2));
  ATLG_HIP(hipMemcpyAsync(dpiv, ipiv, sizeof(int) * K2, hipMemcpyHostToDevice, s));
  laswp(s, N, a.d, lda, K1, K2, dpiv, inci);
  a.out(s, A, rows, N, lda);
  ATLG_HIP(hipStreamSynchronize(s));
}

// ATL_dgetrs, both orders (atlas-lp.base:1631-1720).  For RowMajor the B
// array holds NRHS rows of length N (ldb >= N), i.e. the same memory as a
// column-major N x NRHS matrix; each cblas_trsm(RowMajor, Right, uplo, ...)
// of the reference is the column-major Left solve with uplo flipped.
void h_getrs(int Order, int Trans, long N, long NRHS, const double* A, long lda, const int* ipiv,
             double* B, long ldb) {
  if (N == 0 || NRHS == 0) return;
  hipStream_t s = lib_stream();
  Stage a, b;
  a.in(s, kSlotStageA, A, N, N, lda, true);
  b.in(s, kSlotStageB, B, N, NRHS, ldb, true);
  int* dpiv = static_cast<int*>(scratch(s, kSlotIpiv, sizeof(int) * N));
  ATLG_HIP(hipMemcpyAsync(dpiv, ipiv, sizeof(int) * N, hipMemcpyHostToDevice, s));
  if (Order == kColMajor) {
    getrs(s, Trans, N, NRHS, a.d, lda, dpiv, b.d, ldb);
  } else if (!is_trans(Trans)) {
    trsm(s, kLeft, kUpper, kTrans, kNonUnit, N, NRHS, 1.0, a.d, lda, b.d, ldb);
    trsm(s, kLeft, kLower, kTrans, kUnit, N, NRHS, 1.0, a.d, lda, b.d, ldb);
    laswp(s, NRHS, b.d, ldb, 0, (int)N, dpiv, -1);
  } else {
    laswp(s, NRHS, b.d, ldb, 0, (int)N, dpiv, 1);
    trsm(s, kLeft, kLower, kNoTrans, kUnit, N, NRHS, 1.0, a.d, lda, b.d, ldb);
    trsm(s, kLeft, kUpper, kNoTrans, kNonUnit, N, NRHS, 1.0, a.d, lda, b.d, ldb);
  }
  b.out(s, B, N, NRHS, ldb);
  ATLG_HIP(hipStreamSynchronize(s));
}

int h_potrf(int uplo, long N, double* A, long lda) {
  if (N == 0) return 0;
  hipStream_t s = lib_stream();
  Stage a;
  a.in(s, kSlotStageA, A, N, N, lda, true);
  int* dinfo = info_word(s);
  potrf(s, uplo, N, a.d, lda, dinfo);
  a.out(s, A, N, N, lda);
  int info = 0;
  ATLG_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  ATLG_HIP(hipStreamSynchronize(s));
  return info;
}

void h_trmm(int Side, int Uplo, int TA, int Diag, long M, long N, double alpha, const double* A,
            long lda, double* B, long ldb) {
  if (M == 0 || N == 0) return;
  hipStream_t s = lib_stream();
  const long na = (Side == kLeft) ? M : N;
  Stage a, b;
  a.in(s, kSlotStageA, A, na, na, lda, true);
  b.in(s, kSlotStageB, B, M, N, ldb, alpha != 0.0);
  trmm(s, Side, Uplo, TA, Diag, M, N, alpha, a.d, lda, b.d, ldb);
  b.out(s, B, M, N, ldb);
  ATLG_HIP(hipStreamSynchronize(s));
}

void h_symm(int Side, int Uplo, long M, long N, double alpha, const double* A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc) {
  if (M == 0 || N == 0) return;
  hipStream_t s = lib_stream();
  const long na = (Side == kLeft) ? M : N;
  Stage a, b, c;
  if (alpha != 0.0) {
    a.in(s, kSlotStageA, A, na, na, lda, true);
    b.in(s, kSlotStageB, B, M, N, ldb, true);
  }
  c.in(s, kSlotStageC, C, M, N, ldc, beta != 0.0);
  symm(s, Side, Uplo, M, N, alpha, a.d, lda, b.d, ldb, beta, c.d, ldc);
  c.out(s, C, M, N, ldc);
  ATLG_HIP(hipStreamSynchronize(s));
}

void h_syr2k(int Uplo, int Trans, long N, long K, double alpha, const double* A, long lda,
             const double* B, long ldb, double beta, double* C, long ldc) {
  if (N == 0) return;
  hipStream_t s = lib_stream();
  const bool tr = is_trans(Trans);
  Stage a, b, c;
  if (K > 0 && alpha != 0.0) {
    a.in(s, kSlotStageA, A, tr ? K : N, tr ? N : K, lda, true);
    b.in(s, kSlotStageB, B, tr ? K : N, tr ? N : K, ldb, true);
  }
  c.in(s, kSlotStageC, C, N, N, ldc, true);  // the other triangle is kept
  syr2k(s, Uplo, Trans, N, K, alpha, a.d, lda, b.d, ldb, beta, c.d, ldc);
  c.out(s, C, N, N, ldc);
  ATLG_HIP(hipStreamSynchronize(s));
}

int h_trtri(int uplo, int diag, long N, double* A, long lda) {
  if (N == 0) return 0;
  hipStream_t s = lib_stream();
  Stage a;
  a.in(s, kSlotStageA, A, N, N, lda, true);
  int* dinfo = info_word(s);
  trtri(s, uplo, diag, N, a.d, lda, dinfo);
  a.out(s, A, N, N, lda);
  int info = 0;
  ATLG_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  ATLG_HIP(hipStreamSynchronize(s));
  return info;
}

void h_lauum(int uplo, long N, double* A, long lda) {
  if (N == 0) return;
  hipStream_t s = lib_stream();
  Stage a;
  a.in(s, kSlotStageA, A, N, N, lda, true);
  lauum(s, uplo, N, a.d, lda);
  a.out(s, A, N, N, lda);
  ATLG_HIP(hipStreamSynchronize(s));
}

int h_potri(int uplo, long N, double* A, long lda) {
  if (N == 0) return 0;
  hipStream_t s = lib_stream();
  Stage a;
  a.in(s, kSlotStageA, A, N, N, lda, true);
  int* dinfo = info_word(s);
  potri(s, uplo, N, a.d, lda, dinfo);
  a.out(s, A, N, N, lda);
  int info = 0;
  ATLG_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  ATLG_HIP(hipStreamSynchronize(s));
  return info;
}

// ATL_dpotrs, both orders (atlas-lp.base:969-1050); RowMajor B as in h_getrs.
void h_potrs(int Order, int Uplo, long N, long NRHS, const double* A, long lda, double* B,
             long ldb) {
  if (N == 0 || NRHS == 0) return;
  hipStream_t s = lib_stream();
  Stage a, b;
  a.in(s, kSlotStageA, A, N, N, lda, true);
  b.in(s, kSlotStageB, B, N, NRHS, ldb, true);
  if (Order == kColMajor) {
    potrs(s, Uplo, N, NRHS, a.d, lda, b.d, ldb);
  } else if (Uplo == kUpper) {
    trsm(s, kLeft, kLower, kNoTrans, kNonUnit, N, NRHS, 1.0, a.d, lda, b.d, ldb);
    trsm(s, kLeft, kLower, kTrans, kNonUnit, N, NRHS, 1.0, a.d, lda, b.d, ldb);
  } else {
    trsm(s, kLeft, kUpper, kTrans, kNonUnit, N, NRHS, 1.0, a.d, lda, b.d, ldb);
    trsm(s, kLeft, kUpper, kNoTrans, kNonUnit, N, NRHS, 1.0, a.d, lda, b.d, ldb);
  }
  b.out(s, B, N, NRHS, ldb);
  ATLG_HIP(hipStreamSynchronize(s));
}

// After a host call's device work: reads and clears the library stream's
// failure word and reports a bounded wait that gave up through the failure
// hook.  True when there was a failure (and the hook returned).
bool check_failure(const char* rout) {
  const int st = take_status(lib_stream());
  if (st == 0) return false;
  std::string msg = "a bounded wait gave up:";
  if (st & ATLG_FAIL_PANEL) msg += " getrf panel pivot exchange;";
  if (st & ATLG_FAIL_TRSV) msg += " few-RHS triangular solve ready flags;";
  msg += " the results of this call are unspecified";
  report_failure(rout, st, msg.c_str());
  return true;
}

// Runs a host entry point under the API mutex: HIP failures abort with a
// message, and a bounded-wait failure goes to the failure hook; if the hook
// returns, routines with an int result return INT_MIN.
template <class F>
auto locked(const char* rout, F&& f) -> decltype(f()) {
  std::lock_guard<std::mutex> lk(g_api_mu);
  try {
    if constexpr (std::is_void_v<decltype(f())>) {
      f();
      check_failure(rout);
      return;
    } else {
      const auto r = f();
      if (check_failure(rout)) return INT_MIN;
      return r;
    }
  } catch (const HipFailure& h) {
    fail_hip(h, rout);
  }
  return decltype(f())();
}

}  // namespace

extern "C" {

// ---- error reporting -------------------------------------------------------

// cblas_xerbla (AtlasBase/Clint/cblas.base:42-59)
__attribute__((weak)) void cblas_xerbla(int p, const char* rout, const char* form, ...) {
  char msg[512];
  va_list ap;
  va_start(ap, form);
  vsnprintf(msg, sizeof msg, form, ap);
  va_end(ap);
  if (g_xerbla_hook) {
    g_xerbla_hook(p, rout, msg);
    return;
  }
  if (p) fprintf(stderr, "Parameter %d to routine %s was incorrect\n", p, rout);
  fputs(msg, stderr);
  exit(-1);
}

void atlg_set_xerbla_hook(void (*hook)(int, const char*, const char*)) { g_xerbla_hook = hook; }

}  // extern "C"

void atlg::f77_xerbla(int p, const char* rout, const char* msg) {
  if (g_xerbla_hook) {
    g_xerbla_hook(p, rout, msg);
    return;
  }
  fputs(msg, stdout);
  fflush(stdout);
  exit(0);
}

extern "C" {

// ---- CBLAS -----------------------------------------------------------------

// cblas_dgemm (AtlasBase/Clint/cblas.base:110-236)
void cblas_dgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TA,
                 const enum CBLAS_TRANSPOSE TB, const int M, const int N, const int K,
                 const double alpha, const double* A, const int lda, const double* B,
                 const int ldb, const double beta, double* C, const int ldc) {
  int info = 2000;
  if (M < 0) info = errprn(4, info, "M cannot be less than zero 0,; is set to %d.", M);
  if (N < 0) info = errprn(5, info, "N cannot be less than zero 0,; is set to %d.", N);
  if (K < 0) info = errprn(6, info, "K cannot be less than zero 0,; is set to %d.", K);
  if (Order == CblasRowMajor) {
    if (TA == CblasNoTrans) {
      if (lda < K || lda < 1)
        info = errprn(9, info, "lda must be >= MAX(K,1): lda=%d K=%d", lda, K);
    } else {
      if (TA != CblasConjTrans && TA != CblasTrans)
        info = errprn(2, info, "TransA must be %d, %d or %d, but is set to %d", CblasNoTrans,
                      CblasTrans, CblasConjTrans, TA);
      if (lda < M || lda < 1)
        info = errprn(9, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    }
    if (TB == CblasNoTrans) {
      if (ldb < N || ldb < 1)
        info = errprn(11, info, "ldb must be >= MAX(N,1): ldb=%d N=%d", ldb, N);
    } else {
      if (TB != CblasConjTrans && TB != CblasTrans)
        info = errprn(3, info, "TransB must be %d, %d or %d, but is set to %d", CblasNoTrans,
                      CblasTrans, CblasConjTrans, TB);
      if (ldb < K || ldb < 1)
        info = errprn(11, info, "ldb must be >= MAX(K,1): ldb=%d K=%d", ldb, K);
    }
    if (ldc < N || ldc < 1)
      info = errprn(14, info, "ldc must be >= MAX(N,1): ldc=%d N=%d", ldc, N);
  } else if (Order == CblasColMajor) {
    if (TA == CblasNoTrans) {
      if (lda < M || lda < 1)
        info = errprn(9, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    } else {
      if (TA != CblasConjTrans && TA != CblasTrans)
        info = errprn(2, info, "TransA must be %d, %d or %d, but is set to %d", CblasNoTrans,
                      CblasTrans, CblasConjTrans, TA);
      if (lda < K || lda < 1)
        info = errprn(9, info, "lda must be >= MAX(K,1): lda=%d K=%d", lda, K);
    }
    if (TB == CblasNoTrans) {
      if (ldb < K || ldb < 1)
        info = errprn(11, info, "ldb must be >= MAX(K,1): ldb=%d K=%d", ldb, K);
    } else {
      if (TB != CblasConjTrans && TB != CblasTrans)
        info = errprn(3, info, "TransB must be %d, %d or %d, but is set to %d", CblasNoTrans,
                      CblasTrans, CblasConjTrans, TB);
      if (ldb < N || ldb < 1)
        info = errprn(11, info, "ldb must be >= MAX(N,1): ldb=%d N=%d", ldb, N);
    }
    if (ldc < M || ldc < 1)
      info = errprn(14, info, "ldc must be >= MAX(M,1): ldc=%d M=%d", ldc, M);
  } else {
    info = errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                  CblasColMajor, Order);
  }
  if (info != 2000) {
    cblas_xerbla(info, "cblas_dgemm", "");
    return;
  }
  locked("cblas_dgemm", [&] {
    // A==B asking for A*A^T (or A^T*A) with beta==0: SYRK of the upper
    // triangle plus reflection, as the reference does (cblas.base:221-231)
    if (A == B && M == N && TA != TB && lda == ldb && beta == 0.0) {
      h_syrk(kUpper, (Order == CblasColMajor) ? TA : TB, N, K, alpha, A, lda, beta, C, ldc, true);
      return;
    }
    if (Order == CblasColMajor)
      h_gemm(TA, TB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    else
      h_gemm(TB, TA, N, M, K, alpha, B, ldb, A, lda, beta, C, ldc);
  });
}

}  // extern "C"

namespace {
// The argument checks cblas_dtrsm and cblas_dtrmm share (one wrapper in
// AtlasBase/Clint/cblas.base:587-693); 2000 when every argument is legal.
int trxm_checks(int Order, int Side, int Uplo, int TA, int Diag, int M, int N, int lda,
                int ldb) {
  int info = 2000;
  if (Order == CblasColMajor) {
    if (Side == CblasLeft) {
      if (lda < M || lda < 1)
        info = errprn(10, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    } else if (Side == CblasRight) {
      if (lda < N || lda < 1)
        info = errprn(10, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
    } else {
      info = errprn(2, info, "SIDE must be %d or %d, but is set to %d", CblasRight, CblasLeft,
                    Side);
    }
    if (ldb < M || ldb < 1)
      info = errprn(12, info, "ldb must be >= MAX(M,1): ldb=%d M=%d", ldb, M);
  } else if (Order == CblasRowMajor) {
    if (Side == CblasLeft) {
      if (lda < M || lda < 1)
        info = errprn(10, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    } else if (Side == CblasRight) {
      if (lda < N || lda < 1)
        info = errprn(10, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
    } else {
      info = errprn(2, info, "SIDE must be %d or %d, but is set to %d", CblasRight, CblasLeft,
                    Side);
    }
    if (ldb < N || ldb < 1)
      info = errprn(12, info, "ldb must be >= MAX(N,1): ldb=%d N=%d", ldb, N);
  } else {
    info = errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                  CblasColMajor, Order);
  }
  if (Uplo != CblasUpper && Uplo != CblasLower)
    info = errprn(3, info, "UPLO must be %d or %d, but is set to %d", CblasUpper, CblasLower,
                  Uplo);
  if (TA != CblasNoTrans && TA != CblasTrans && TA != CblasConjTrans)
    info = errprn(4, info, "TransA must be %d, %d or %d, but is set to %d", CblasNoTrans,
                  CblasTrans, CblasConjTrans, TA);
  if (Diag != CblasUnit && Diag != CblasNonUnit)
    info = errprn(5, info, "UPLO must be %d or %d, but is set to %d", CblasUnit, CblasNonUnit,
                  Diag);
  if (M < 0) info = errprn(6, info, "M cannot be less than zero; it is set to %d.", M);
  if (N < 0) info = errprn(7, info, "N cannot be less than zero; it is set to %d.", N);
  return info;
}
}  // namespace

extern "C" {

// cblas_dtrsm (AtlasBase/Clint/cblas.base:587-693)
void cblas_dtrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const double alpha,
                 const double* A, const int lda, double* B, const int ldb) {
  const int info = trxm_checks(Order, Side, Uplo, TA, Diag, M, N, lda, ldb);
  if (info != 2000) {
    cblas_xerbla(info, "cblas_dtrsm", "");
    return;
  }
  locked("cblas_dtrsm", [&] {
    if (Order == CblasColMajor)
      h_trsm(Side, Uplo, TA, Diag, M, N, alpha, A, lda, B, ldb);
    else
      h_trsm(Side == CblasLeft ? kRight : kLeft, Uplo == CblasUpper ? kLower : kUpper, TA, Diag,
             N, M, alpha, A, lda, B, ldb);
  });
}

// cblas_dtrmm (AtlasBase/Clint/cblas.base:587-693, the TRSM/TRMM wrapper)
void cblas_dtrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const double alpha,
                 const double* A, const int lda, double* B, const int ldb) {
  const int info = trxm_checks(Order, Side, Uplo, TA, Diag, M, N, lda, ldb);
  if (info != 2000) {
    cblas_xerbla(info, "cblas_dtrmm", "");
    return;
  }
  locked("cblas_dtrmm", [&] {
    if (Order == CblasColMajor)
      h_trmm(Side, Uplo, TA, Diag, M, N, alpha, A, lda, B, ldb);
    else
      h_trmm(Side == CblasLeft ? kRight : kLeft, Uplo == CblasUpper ? kLower : kUpper, TA, Diag,
             N, M, alpha, A, lda, B, ldb);
  });
}

// cblas_dsymm (AtlasBase/Clint/cblas.base:237-338)
void cblas_dsymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const double alpha,
                 const double* A, const int lda, const double* B, const int ldb,
                 const double beta, double* C, const int ldc) {
  int info = 2000;
  if (Order == CblasColMajor || Order == CblasRowMajor) {
    const bool cm = Order == CblasColMajor;
    if (Side == CblasLeft) {
      if (lda < M || lda < 1)
        info = errprn(8, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    } else if (Side == CblasRight) {
      if (lda < N || lda < 1)
        info = errprn(8, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
    } else {
      info = errprn(2, info, "SIDE must be %d or %d, but is set to %d", CblasRight, CblasLeft,
                    Side);
    }
    if (cm) {
      if (ldb < M || ldb < 1)
        info = errprn(10, info, "ldb must be >= MAX(M,1): ldb=%d M=%d", ldb, M);
      if (ldc < M || ldc < 1)
        info = errprn(13, info, "ldc must be >= MAX(M,1): ldc=%d M=%d", ldc, M);
    } else {
      if (ldb < N || ldb < 1)
        info = errprn(10, info, "ldb must be >= MAX(N,1): ldb=%d N=%d", ldb, N);
      if (ldc < N || ldc < 1)
        info = errprn(13, info, "ldc must be >= MAX(N,1): ldc=%d N=%d", ldc, N);
    }
  } else {
    info = errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                  CblasColMajor, Order);
  }
  if (Uplo != CblasUpper && Uplo != CblasLower)
    info = errprn(3, info, "UPLO must be %d or %d, but is set to %d", CblasUpper, CblasLower,
                  Uplo);
  if (M < 0) info = errprn(4, info, "M cannot be less than zero; it is set to %d.", M);
  if (N < 0) info = errprn(5, info, "N cannot be less than zero; it is set to %d.", N);
  if (info != 2000) {
    cblas_xerbla(info, "cblas_dsymm", "");
    return;
  }
  locked("cblas_dsymm", [&] {
    if (Order == CblasColMajor)
      h_symm(Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
    else
      h_symm(Side == CblasLeft ? kRight : kLeft, Uplo == CblasUpper ? kLower : kUpper, N, M,
             alpha, A, lda, B, ldb, beta, C, ldc);
  });
}

// cblas_dsyr2k (AtlasBase/Clint/cblas.base:450-586)
void cblas_dsyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const double alpha,
                  const double* A, const int lda, const double* B, const int ldb,
                  const double beta, double* C, const int ldc) {
  int info = 2000;
  if (Uplo != CblasUpper && Uplo != CblasLower)
    info = errprn(2, info, "UPLO must be %d or %d, but is set to %d", CblasUpper, CblasLower,
                  Uplo);
  if (N < 0) info = errprn(4, info, "N cannot be less than zero; it is set to %d.", N);
  if (K < 0) info = errprn(5, info, "K cannot be less than zero; it is set to %d.", K);
  if (Order == CblasColMajor || Order == CblasRowMajor) {
    // the long dimension of A and B: N for (ColMajor, NoTrans) and
    // (RowMajor, Trans), K otherwise
    const bool nt = Trans == CblasNoTrans;
    if (!nt && Trans != CblasTrans && Trans != CblasConjTrans)
      info = errprn(3, info, "Trans must be %d or %d, but is set to %d", CblasNoTrans, CblasTrans,
                    Trans);
    const bool byN = (Order == CblasColMajor) == nt;
    const int need = byN ? N : K;
    const char* nm = byN ? "N" : "K";
    if (lda < need || lda < 1)
      info = errprn(8, info, "lda must be >= MAX(%s,1): lda=%d %s=%d", nm, lda, nm, need);
    if (ldb < need || ldb < 1)
      info = errprn(10, info, "ldb must be >= MAX(%s,1): ldb=%d %s=%d", nm, ldb, nm, need);
  } else {
    info = errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                  CblasColMajor, Order);
  }
  if (ldc < N || ldc < 1)
    info = errprn(13, info, "ldc must be >= MAX(N,1): ldc=%d N=%d", ldc, N);
  if (info != 2000) {
    cblas_xerbla(info, "cblas_dsyr2k", "");
    return;
  }
  locked("cblas_dsyr2k", [&] {
    if (Order == CblasColMajor)
      h_syr2k(Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    else
      h_syr2k(Uplo == CblasUpper ? kLower : kUpper, Trans == CblasNoTrans ? kTrans : kNoTrans, N,
              K, alpha, A, lda, B, ldb, beta, C, ldc);
  });
}

// cblas_dsyrk (AtlasBase/Clint/cblas.base:339-449)
void cblas_dsyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const double alpha,
                 const double* A, const int lda, const double beta, double* C, const int ldc) {
  int info = 2000;
  if (Uplo != CblasUpper && Uplo != CblasLower)
    info = errprn(2, info, "UPLO must be %d or %d, but is set to %d", CblasUpper, CblasLower,
                  Uplo);
  if (N < 0) info = errprn(4, info, "N cannot be less than zero; it is set to %d.", N);
  if (K < 0) info = errprn(5, info, "K cannot be less than zero; it is set to %d.", K);
  if (Order == CblasColMajor) {
    if (Trans == CblasNoTrans) {
      if (lda < N || lda < 1)
        info = errprn(8, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
    } else {
      if (Trans != CblasTrans && Trans != CblasConjTrans)
        info = errprn(3, info, "Trans must be %d, %d or %d, but is set to %d", CblasNoTrans,
                      CblasTrans, CblasConjTrans, Trans);
      if (lda < K || lda < 1)
        info = errprn(8, info, "lda must be >= MAX(K,1): lda=%d K=%d", lda, K);
    }
  } else if (Order == CblasRowMajor) {
    if (Trans == CblasNoTrans) {
      if (lda < K || lda < 1)
        info = errprn(8, info, "lda must be >= MAX(K,1): lda=%d K=%d", lda, K);
    } else {
      if (Trans != CblasTrans && Trans != CblasConjTrans)
        info = errprn(3, info, "Trans must be %d, %d or %d, but is set to %d", CblasNoTrans,
                      CblasTrans, CblasConjTrans, Trans);
      if (lda < N || lda < 1)
        info = errprn(8, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
    }
  } else {
    info = errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                  CblasColMajor, Order);
  }
  if (ldc < N || ldc < 1)
    info = errprn(11, info, "ldc must be >= MAX(N,1): ldc=%d N=%d", ldc, N);
  if (info != 2000) {
    cblas_xerbla(info, "cblas_dsyrk", "");
    return;
  }
  locked("cblas_dsyrk", [&] {
    if (Order == CblasColMajor)
      h_syrk(Uplo, Trans, N, K, alpha, A, lda, beta, C, ldc, false);
    else
      h_syrk(Uplo == CblasUpper ? kLower : kUpper, Trans == CblasNoTrans ? kTrans : kNoTrans, N,
             K, alpha, A, lda, beta, C, ldc, false);
  });
}

// ---- ATL_ internal names (no argument checking, as in ATLAS) ---------------

void ATL_dgemm(const enum CBLAS_TRANSPOSE TA, const enum CBLAS_TRANSPOSE TB, ATL_CSZT M,
               ATL_CSZT N, ATL_CSZT K, const double alpha, const double* A, ATL_CSZT lda,
               const double* B, ATL_CSZT ldb, const double beta, double* C, ATL_CSZT ldc) {
  locked("ATL_dgemm", [&] {
    h_gemm(TA, TB, (long)M, (long)N, (long)K, alpha, A, (long)lda, B, (long)ldb, beta, C,
           (long)ldc);
  });
}

void ATL_dsyrk(const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE Trans, ATL_CSZT N,
               ATL_CSZT K, const double alpha, const double* A, ATL_CSZT lda, const double beta,
               double* C, ATL_CSZT ldc) {
  locked("ATL_dsyrk", [&] {
    h_syrk(Uplo, Trans, (long)N, (long)K, alpha, A, (long)lda, beta, C, (long)ldc, false);
  });
}

void ATL_dtrsm(const enum CBLAS_SIDE Side, const enum CBLAS_UPLO Uplo,
               const enum CBLAS_TRANSPOSE TA, const enum CBLAS_DIAG Diag, ATL_CSZT M, ATL_CSZT N,
               const double alpha, const double* A, ATL_CSZT lda, double* B, ATL_CSZT ldb) {
  locked("ATL_dtrsm", [&] {
    h_trsm(Side, Uplo, TA, Diag, (long)M, (long)N, alpha, A, (long)lda, B, (long)ldb);
  });
}

int ATL_dgetrfC(const int M, const int N, double* A, const int lda, int* ipiv) {
  return locked("ATL_dgetrfC", [&] { return h_getrfC(M, N, A, lda, ipiv); });
}

// Row-major LU (A = L U Pc, unit U) is the column-major LU of A^T, which is
// the same memory (gen src/lapack/ATL_getrfR.c mirrors getrfC with rows and
// columns exchanged).
int ATL_dgetrfR(const int M, const int N, double* A, const int lda, int* ipiv) {
  return locked("ATL_dgetrfR", [&] { return h_getrfC(N, M, A, lda, ipiv); });
}

int ATL_dgetrf(const enum CBLAS_ORDER Order, const int M, const int N, double* A, const int lda,
               int* ipiv) {
  return (Order == CblasColMajor) ? ATL_dgetrfC(M, N, A, lda, ipiv)
                                  : ATL_dgetrfR(M, N, A, lda, ipiv);
}

void ATL_dgetrs(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE Trans, const int N,
                const int NRHS, const double* A, const int lda, const int* ipiv, double* B,
                const int ldb) {
  locked("ATL_dgetrs", [&] { h_getrs(Order, Trans, N, NRHS, A, lda, ipiv, B, ldb); });
}

void ATL_dlaswp(const int N, double* A, const int lda0, const int K1, const int K2,
                const int* ipiv, const int inci) {
  locked("ATL_dlaswp", [&] { h_laswp(N, A, lda0, K1, K2, ipiv, inci); });
}

int ATL_dpotrfL(const int N, double* A, const int lda) {
  return locked("ATL_dpotrfL", [&] { return h_potrf(kLower, N, A, lda); });
}

int ATL_dpotrfU(const int N, double* A, const int lda) {
  return locked("ATL_dpotrfU", [&] { return h_potrf(kUpper, N, A, lda); });
}

// ATL_dpotrf (atlas-lp.base:332-360): row-major Upper is column-major Lower
// of the same memory and vice versa.
int ATL_dpotrf(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N, double* A,
               const int lda) {
  const bool lower = (Order == CblasColMajor) ? (Uplo == CblasLower) : (Uplo == CblasUpper);
  return lower ? ATL_dpotrfL(N, A, lda) : ATL_dpotrfU(N, A, lda);
}

void ATL_dpotrs(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                const int NRHS, const double* A, const int lda, double* B, const int ldb) {
  locked("ATL_dpotrs", [&] { h_potrs(Order, Uplo, N, NRHS, A, lda, B, ldb); });
}

void ATL_dtrmm(const enum CBLAS_SIDE Side, const enum CBLAS_UPLO Uplo,
               const enum CBLAS_TRANSPOSE TA, const enum CBLAS_DIAG Diag, ATL_CSZT M, ATL_CSZT N,
               const double alpha, const double* A, ATL_CSZT lda, double* B, ATL_CSZT ldb) {
  locked("ATL_dtrmm", [&] {
    h_trmm(Side, Uplo, TA, Diag, (long)M, (long)N, alpha, A, (long)lda, B, (long)ldb);
  });
}

void ATL_dsymm(const enum CBLAS_SIDE Side, const enum CBLAS_UPLO Uplo, ATL_CSZT M, ATL_CSZT N,
               const double alpha, const double* A, ATL_CSZT lda, const double* B, ATL_CSZT ldb,
               const double beta, double* C, ATL_CSZT ldc) {
  locked("ATL_dsymm", [&] {
    h_symm(Side, Uplo, (long)M, (long)N, alpha, A, (long)lda, B, (long)ldb, beta, C, (long)ldc);
  });
}

void ATL_dsyr2k(const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE Trans, ATL_CSZT N,
                ATL_CSZT K, const double alpha, const double* A, ATL_CSZT lda, const double* B,
                ATL_CSZT ldb, const double beta, double* C, ATL_CSZT ldc) {
  locked("ATL_dsyr2k", [&] {
    h_syr2k(Uplo, Trans, (long)N, (long)K, alpha, A, (long)lda, B, (long)ldb, beta, C,
            (long)ldc);
  });
}

// ATL_trtri (atlas-lp.base:2341-2380): a zero on a NonUnit diagonal returns
// its 1-based index and leaves A unchanged.  Row-major is the column-major
// inverse of the transpose, i.e. the other triangle of the same memory.
int ATL_dtrtri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
               const enum CBLAS_DIAG Diag, const int N, double* A, const int lda) {
  const bool lower = (Order == CblasColMajor) == (Uplo == CblasLower);
  return locked("ATL_dtrtri", [&] { return h_trtri(lower ? kLower : kUpper, Diag, N, A, lda); });
}

// ATL_lauum (atlas-lp.base:312-330): lower A <- L^T L, upper A <- U U^T;
// row-major is the column-major routine for the other triangle.
void ATL_dlauum(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                double* A, const int lda) {
  const bool lower = (Order == CblasColMajor) == (Uplo == CblasLower);
  locked("ATL_dlauum", [&] { h_lauum(lower ? kLower : kUpper, N, A, lda); });
}

// The SPD inverse from the Cholesky factor (clapack_dpotri's trtri + lauum,
// atlas-clp.base:316-362).
int ATL_dpotri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N, double* A,
               const int lda) {
  const bool lower = (Order == CblasColMajor) == (Uplo == CblasLower);
  return locked("ATL_dpotri", [&] { return h_potri(lower ? kLower : kUpper, N, A, lda); });
}

// ---- CLAPACK (AtlasBase/Clint/atlas-clp.base) ------------------------------

// clapack_dtrtri (atlas-clp.base:411-452).  The reference reads its error
// code uninitialized and calls ATL_trtri only when it is nonzero; here the
// inverse is computed exactly when every argument is legal.
int clapack_dtrtri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const enum CBLAS_DIAG Diag, const int N, double* A, const int lda) {
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, "clapack_dtrtri", "Order must be %d or %d, but is set to %d\n",
                 CblasRowMajor, CblasColMajor, Order);
  }
  if (Uplo != CblasUpper && Uplo != CblasLower) {
    ierr = -2;
    cblas_xerbla(2, "clapack_dtrtri", "Uplo must be %d or %d, but is set to %d\n", CblasUpper,
                 CblasLower, Uplo);
  }
  if (Diag != CblasUnit && Diag != CblasNonUnit) {
    ierr = -3;
    cblas_xerbla(3, "clapack_dtrtri", "Diag must be %d or %d, but is set to %d\n", CblasNonUnit,
                 CblasUnit, Diag);
  }
  if (N < 0) {
    ierr = -4;
    cblas_xerbla(4, "clapack_dtrtri", "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (lda < N || lda < 1) {
    ierr = -6;
    cblas_xerbla(6, "clapack_dtrtri", "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (!ierr) ierr = ATL_dtrtri(Order, Uplo, Diag, N, A, lda);
  return ierr;
}

namespace {
// The checks clapack_dpotri and clapack_dlauum share with clapack_dpotrf
// (atlas-clp.base:316-356); 0 or the negative error code.
int potri_like_checks(const char* rt, int Order, int Uplo, int N, int lda) {
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, rt, "Order must be %d or %d, but is set to %d\n", CblasRowMajor,
                 CblasColMajor, Order);
  }
  if (Uplo != CblasUpper && Uplo != CblasLower) {
    ierr = -2;
    cblas_xerbla(2, rt, "Uplo must be %d or %d, but is set to %d\n", CblasUpper, CblasLower,
                 Uplo);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, rt, "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (lda < N || lda < 1) {
    ierr = -5;
    cblas_xerbla(5, rt, "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  return ierr;
}
}  // namespace

// clapack_dpotri (atlas-clp.base:316-362): trtri, then lauum if it succeeded.
int clapack_dpotri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   double* A, const int lda) {
  int ierr = potri_like_checks("clapack_dpotri", Order, Uplo, N, lda);
  if (!ierr) ierr = ATL_dpotri(Order, Uplo, N, A, lda);
  return ierr;
}

// clapack_dlauum (atlas-clp.base:316-362)
int clapack_dlauum(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   double* A, const int lda) {
  const int ierr = potri_like_checks("clapack_dlauum", Order, Uplo, N, lda);
  if (!ierr) ATL_dlauum(Order, Uplo, N, A, lda);
  return ierr;
}

int clapack_dgetrf(const enum CBLAS_ORDER Order, const int M, const int N, double* A,
                   const int lda, int* ipiv) {
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, "clapack_dgetrf", "Order must be %d or %d, but is set to %d\n",
                 CblasRowMajor, CblasColMajor, Order);
  }
  if (M < 0) {
    ierr = -2;
    cblas_xerbla(2, "clapack_dgetrf", "M cannot be less than zero 0,; is set to %d.\n", M);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, "clapack_dgetrf", "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (Order == CblasColMajor) {
    if (lda < M || lda < 1) {
      ierr = -6;
      cblas_xerbla(6, "clapack_dgetrf", "lda must be >= MAX(M,1): lda=%d M=%d\n", lda, M);
    }
  } else {
    if (lda < N || lda < 1) {
      ierr = -6;
      cblas_xerbla(6, "clapack_dgetrf", "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
    }
  }
  if (!ierr) ierr = ATL_dgetrf(Order, M, N, A, lda, ipiv);
  return ierr;
}

int clapack_dgetrs(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE Trans, const int N,
                   const int NRHS, const double* A, const int lda, const int* ipiv, double* B,
                   const int ldb) {
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, "clapack_dgetrs", "Order must be %d or %d, but is set to %d\n",
                 CblasRowMajor, CblasColMajor, Order);
  }
  if (Trans != CblasNoTrans && Trans != CblasTrans && Trans != CblasConjTrans) {
    ierr = -2;
    cblas_xerbla(2, "clapack_dgetrs", "Trans must be %d, %d, or %d, but is set to %d\n",
                 CblasNoTrans, CblasTrans, CblasConjTrans, Trans);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, "clapack_dgetrs", "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (NRHS < 0) {
    ierr = -4;
    cblas_xerbla(4, "clapack_dgetrs", "NRHS cannot be less than zero 0,; is set to %d.\n", NRHS);
  }
  if (lda < N || lda < 1) {
    ierr = -6;
    cblas_xerbla(6, "clapack_dgetrs", "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (ldb < N || ldb < 1) {
    ierr = -9;
    cblas_xerbla(9, "clapack_dgetrs", "ldb must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (!ierr) ATL_dgetrs(Order, Trans, N, NRHS, A, lda, ipiv, B, ldb);
  return ierr;
}

int clapack_dgesv(const enum CBLAS_ORDER Order, const int N, const int NRHS, double* A,
                  const int lda, int* ipiv, double* B, const int ldb) {
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, "clapack_dgesv", "Order must be %d or %d, but is set to %d.\n",
                 CblasRowMajor, CblasColMajor, Order);
  }
  if (N < 0) {
    ierr = -2;
    cblas_xerbla(2, "clapack_dgesv", "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (NRHS < 0) {
    ierr = -3;
    cblas_xerbla(3, "clapack_dgesv", "NRHS cannot be less than zero 0,; is set to %d.\n", NRHS);
  }
  if (lda < N || lda < 1) {
    ierr = -5;
    cblas_xerbla(5, "clapack_dgesv", "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (ldb < N || ldb < 1) {
    ierr = -8;
    cblas_xerbla(8, "clapack_dgesv", "ldb must be >= MAX(N,1): ldb=%d N=%d\n", ldb, N);
  }
  if (!ierr) ierr = ATL_dgetrf(Order, N, N, A, lda, ipiv);
  if (!ierr) ATL_dgetrs(Order, CblasNoTrans, N, NRHS, A, lda, ipiv, B, ldb);
  return ierr;
}

int clapack_dpotrf(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   double* A, const int lda) {
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, "clapack_dpotrf", "Order must be %d or %d, but is set to %d\n",
                 CblasRowMajor, CblasColMajor, Order);
  }
  if (Uplo != CblasUpper && Uplo != CblasLower) {
    ierr = -2;
    cblas_xerbla(2, "clapack_dpotrf", "Uplo must be %d or %d, but is set to %d\n", CblasUpper,
                 CblasLower, Uplo);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, "clapack_dpotrf", "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (lda < N || lda < 1) {
    ierr = -5;
    cblas_xerbla(5, "clapack_dpotrf", "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (!ierr) ierr = ATL_dpotrf(Order, Uplo, N, A, lda);
  return ierr;
}

int clapack_dpotrs(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   const int NRHS, const double* A, const int lda, double* B, const int ldb) {
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, "clapack_dpotrs", "Order must be %d or %d, but is set to %d\n",
                 CblasRowMajor, CblasColMajor, Order);
  }
  if (Uplo != CblasUpper && Uplo != CblasLower) {
    ierr = -2;
    cblas_xerbla(2, "clapack_dpotrs", "Uplo must be %d or %d, but is set to %d\n", CblasUpper,
                 CblasLower, Uplo);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, "clapack_dpotrs", "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (NRHS < 0) {
    ierr = -4;
    cblas_xerbla(4, "clapack_dpotrs", "NRHS cannot be less than zero 0,; is set to %d.\n", NRHS);
  }
  if (lda < N || lda < 1) {
    ierr = -6;
    cblas_xerbla(6, "clapack_dpotrs", "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (ldb < N || ldb < 1) {
    ierr = -8;
    cblas_xerbla(8, "clapack_dpotrs", "ldb must be >= MAX(N,1): ldb=%d N=%d\n", ldb, N);
  }
  if (!ierr) ATL_dpotrs(Order, Uplo, N, NRHS, A, lda, B, ldb);
  return ierr;
}

int clapack_dposv(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                  const int NRHS, double* A, const int lda, double* B, const int ldb) {
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, "clapack_dposv", "Order must be %d or %d, but is set to %d\n", CblasRowMajor,
                 CblasColMajor, Order);
  }
  if (Uplo != CblasUpper && Uplo != CblasLower) {
    ierr = -2;
    cblas_xerbla(2, "clapack_dposv", "Uplo must be %d or %d, but is set to %d\n", CblasUpper,
                 CblasLower, Uplo);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, "clapack_dposv", "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (NRHS < 0) {
    ierr = -4;
    cblas_xerbla(4, "clapack_dposv", "NRHS cannot be less than zero 0,; is set to %d.\n", NRHS);
  }
  if (lda < N || lda < 1) {
    ierr = -6;
    cblas_xerbla(6, "clapack_dposv", "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (ldb < N || ldb < 1) {
    ierr = -8;
    cblas_xerbla(8, "clapack_dposv", "ldb must be >= MAX(N,1): ldb=%d N=%d\n", ldb, N);
  }
  if (!ierr) ierr = ATL_dpotrf(Order, Uplo, N, A, lda);
  if (!ierr) ATL_dpotrs(Order, Uplo, N, NRHS, A, lda, B, ldb);
  return ierr;
}

}  // extern "C"
