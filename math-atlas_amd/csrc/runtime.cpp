// Process-global runtime of the level-3 library: per-device stream, per-stream
// scratch slots, error reporting, and the atlg_* device-pointer entry points.
//
// ATLAS keeps no handle or context (its serial routines malloc workspace per
// call, atlas-ammm.base:17266+).  The drop-in therefore keeps an implicit
// context here: one library stream per device, created on first use, and
// scratch memory cached per (stream, slot) so that steady-state calls never
// allocate.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "atlg_internal.hpp"

namespace atlg {

namespace {
thread_local std::string g_last_error;
std::mutex g_mu;
std::map<int, hipStream_t> g_streams;  // device -> library stream

struct SlotBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
std::map<std::pair<hipStream_t, int>, SlotBuf> g_slots;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

hipStream_t resolve_stream(atlg_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

hipStream_t library_stream() {
  int dev = 0;
  ATLG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_streams.find(dev);
  if (it != g_streams.end()) return it->second;
  hipStream_t st;
  ATLG_HIP(hipStreamCreate(&st));
  g_streams[dev] = st;
  return st;
}

void* scratch(hipStream_t s, Slot slot, size_t bytes) {
  if (bytes == 0) bytes = 256;
  std::lock_guard<std::mutex> lk(g_mu);
  SlotBuf& b = g_slots[std::make_pair(s, static_cast<int>(slot))];
  if (b.bytes >= bytes) return b.p;
  if (b.p) {
    // earlier work queued on this stream may still use the old buffer
    ATLG_HIP(hipStreamSynchronize(s));
    ATLG_HIP(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes + bytes / 4;  // some headroom for slowly growing sizes
  want = (want + 4095) & ~static_cast<size_t>(4095);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    e = hipMalloc(&p, bytes);
    want = bytes;
    if (e != hipSuccess)
      throw HipFailure{std::string("hipMalloc scratch: ") + hipGetErrorString(e)};
  }
  b.p = p;
  b.bytes = want;
  return p;
}

}  // namespace atlg

using namespace atlg;

namespace {

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const HipFailure& h) {
    set_error(h.msg);
    return ATLG_EHIP;
  }
}

bool bad_trans(int t) { return t != kNoTrans && t != kTrans && t != kConjTrans; }
bool bad_uplo(int u) { return u != kUpper && u != kLower; }
bool bad_diag(int d) { return d != kUnit && d != kNonUnit; }
bool bad_side(int s) { return s != kLeft && s != kRight; }

}  // namespace

extern "C" {

int atlg_init(void) {
  return guarded([] { (void)library_stream(); });
}

const char* atlg_last_error(void) { return g_last_error.c_str(); }

int atlg_dgemm(int TA, int TB, long M, long N, long K, double alpha,
               const double* A, long lda, const double* B, long ldb,
               double beta, double* C, long ldc, atlg_stream_t stream) {
  if (bad_trans(TA)) return -1;
  if (bad_trans(TB)) return -2;
  if (M < 0) return -3;
  if (N < 0) return -4;
  if (K < 0) return -5;
  if (lda < std::max(1L, is_trans(TA) ? K : M)) return -8;
  if (ldb < std::max(1L, is_trans(TB) ? N : K)) return -10;
  if (ldc < std::max(1L, M)) return -13;
  return guarded([&] {
    hipStream_t s = resolve_stream(stream);
    gemm(s, TA, TB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  });
}

int atlg_dsyrk(int Uplo, int Trans, long N, long K, double alpha,
               const double* A, long lda, double beta, double* C, long ldc,
               atlg_stream_t stream) {
  if (bad_uplo(Uplo)) return -1;
  if (bad_trans(Trans)) return -2;
  if (N < 0) return -3;
  if (K < 0) return -4;
  if (lda < std::max(1L, is_trans(Trans) ? K : N)) return -7;
  if (ldc < std::max(1L, N)) return -10;
  return guarded([&] {
    hipStream_t s = resolve_stream(stream);
    const int mode = (Uplo == kLower) ? 1 : 2;
    if (N == 0) return;
    if (K == 0 || alpha == 0.0) {
      if (beta != 1.0) scale(s, N, N, beta, C, ldc, mode);
      return;
    }
    if (is_trans(Trans))
      gemm(s, kTrans, kNoTrans, N, N, K, alpha, A, lda, A, lda, beta, C, ldc, mode);
    else
      gemm(s, kNoTrans, kTrans, N, N, K, alpha, A, lda, A, lda, beta, C, ldc, mode);
  });
}

int atlg_dtrsm(int Side, int Uplo, int TA, int Diag, long M, long N,
               double alpha, const double* A, long lda, double* B, long ldb,
               atlg_stream_t stream) {
  if (bad_side(Side)) return -1;
  if (bad_uplo(Uplo)) return -2;
  if (bad_trans(TA)) return -3;
  if (bad_diag(Diag)) return -4;
  if (M < 0) return -5;
  if (N < 0) return -6;
  if (lda < std::max(1L, Side == kLeft ? M : N)) return -9;
  if (ldb < std::max(1L, M)) return -11;
  return guarded([&] {
    trsm(resolve_stream(stream), Side, Uplo, TA, Diag, M, N, alpha, A, lda, B, ldb);
  });
}

int atlg_dlaswp(long N, double* A, long lda, int K1, int K2, const int* ipiv,
                int inci, atlg_stream_t stream) {
  if (N < 0) return -1;
  if (lda < 1) return -3;
  if (inci != 1 && inci != -1) return -7;
  return guarded([&] { laswp(resolve_stream(stream), N, A, lda, K1, K2, ipiv, inci); });
}

int atlg_dgetrf(long M, long N, double* A, long lda, int* ipiv, int* dinfo,
                atlg_stream_t stream) {
  if (M < 0) return -1;
  if (N < 0) return -2;
  if (lda < std::max(1L, M)) return -4;
  if (!dinfo) return -6;
  return guarded([&] { getrf(resolve_stream(stream), M, N, A, lda, ipiv, dinfo); });
}

int atlg_dgetrs(int Trans, long N, long NRHS, const double* A, long lda,
                const int* ipiv, double* B, long ldb, atlg_stream_t stream) {
  if (bad_trans(Trans)) return -1;
  if (N < 0) return -2;
  if (NRHS < 0) return -3;
  if (lda < std::max(1L, N)) return -5;
  if (ldb < std::max(1L, N)) return -8;
  return guarded([&] {
    getrs(resolve_stream(stream), Trans, N, NRHS, A, lda, ipiv, B, ldb);
  });
}

int atlg_dpotrf(int Uplo, long N, double* A, long lda, int* dinfo,
                atlg_stream_t stream) {
  if (bad_uplo(Uplo)) return -1;
  if (N < 0) return -2;
  if (lda < std::max(1L, N)) return -4;
  if (!dinfo) return -5;
  return guarded([&] { potrf(resolve_stream(stream), Uplo, N, A, lda, dinfo); });
}

int atlg_dpotrs(int Uplo, long N, long NRHS, const double* A, long lda,
                double* B, long ldb, atlg_stream_t stream) {
  if (bad_uplo(Uplo)) return -1;
  if (N < 0) return -2;
  if (NRHS < 0) return -3;
  if (lda < std::max(1L, N)) return -5;
  if (ldb < std::max(1L, N)) return -7;
  return guarded([&] { potrs(resolve_stream(stream), Uplo, N, NRHS, A, lda, B, ldb); });
}

}  // extern "C"
