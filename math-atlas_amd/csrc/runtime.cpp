// Process-global runtime of the level-3 library: per-device stream, per-stream
// scratch slots, error reporting, and the atlg_* device-pointer entry points.
//
// ATLAS keeps no handle or context (its serial routines malloc workspace per
// call, atlas-ammm.base:17266+).  The drop-in therefore keeps an implicit
// context here: one library stream per device, created on first use, and
// scratch memory cached per (stream, slot) so that steady-state calls never
// allocate.
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>
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
std::map<hipStream_t, int*> g_status;  // stream -> failure word
std::atomic<long> g_wait_limit{0};
void (*g_failure_hook)(const char*, int, const char*) = nullptr;

// Orderly release of everything the runtime created: run once at process
// exit, registered after the HIP runtime has initialized (so it runs before
// the runtime's own teardown).  Waits for each device, frees the scratch
// slots and failure words, destroys the events and streams.
void teardown();
std::once_flag g_atexit_once;
void arm_teardown() {
  std::call_once(g_atexit_once, [] { std::atexit(teardown); });
}
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
  arm_teardown();
  return st;
}

namespace {
struct Aux {
  hipStream_t stream = nullptr;
  hipEvent_t ev[kAuxEvents] = {};
  // CU-masked pairs by the small side's unit count, created on first use
  // ({nullptr, nullptr}: unavailable)
  std::map<int, std::pair<hipStream_t, hipStream_t>> split;
};
std::map<int, Aux> g_aux;
Aux& aux_for_device() {
  int dev = 0;
  ATLG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  Aux& a = g_aux[dev];
  if (!a.stream) {
    ATLG_HIP(hipStreamCreateWithFlags(&a.stream, hipStreamNonBlocking));
    for (auto& e : a.ev) ATLG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    arm_teardown();
  }
  return a;
}
}  // namespace

std::mutex g_aux_mu;
std::unique_lock<std::mutex> aux_lock() { return std::unique_lock<std::mutex>(g_aux_mu); }

hipStream_t aux_stream() { return aux_for_device().stream; }
hipEvent_t aux_event(int i) { return aux_for_device().ev[(unsigned)i % kAuxEvents]; }

bool split_streams(int cus, hipStream_t* small, hipStream_t* large) {
  Aux& a = aux_for_device();
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = a.split.find(cus);
  if (it == a.split.end()) {
    int dev = 0, total = 0;
    ATLG_HIP(hipGetDevice(&dev));
    ATLG_HIP(hipDeviceGetAttribute(&total, hipDeviceAttributeMultiprocessorCount, dev));
    hipStream_t p[2] = {nullptr, nullptr};
    if (cus > 0 && cus < total) {
      const int words = (total + 31) / 32;
      std::vector<uint32_t> m0(words, 0u), m1(words, 0u);
      for (int c = 0; c < total; c++) (c < cus ? m0 : m1)[c / 32] |= 1u << (c % 32);
      if (hipExtStreamCreateWithCUMask(&p[0], words, m0.data()) != hipSuccess ||
          hipExtStreamCreateWithCUMask(&p[1], words, m1.data()) != hipSuccess) {
        (void)hipGetLastError();
        for (auto& st : p)
          if (st) (void)hipStreamDestroy(st);
        p[0] = p[1] = nullptr;
      }
    }
    it = a.split.emplace(cus, std::make_pair(p[0], p[1])).first;
  }
  if (!it->second.first) return false;
  *small = it->second.first;
  *large = it->second.second;
  return true;
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
  arm_teardown();
  return p;
}

int* status_word(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  int*& w = g_status[s];
  if (!w) {
    int* p = nullptr;
    hipError_t e = hipMalloc(&p, 256);
    if (e == hipSuccess) e = hipMemset(p, 0, 256);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      if (p) (void)hipFree(p);
      g_status.erase(s);
      throw HipFailure{std::string("failure word: ") + hipGetErrorString(e)};
    }
    w = p;
    arm_teardown();
  }
  return w;
}

int take_status(hipStream_t s) {
  int* w = status_word(s);
  int v = 0;
  ATLG_HIP(hipMemcpyAsync(&v, w, sizeof(int), hipMemcpyDeviceToHost, s));
  ATLG_HIP(hipStreamSynchronize(s));
  if (v != 0) {
    set_int(s, w, 0);
    ATLG_HIP(hipStreamSynchronize(s));
  }
  return v;
}

long wait_limit(long dflt) {
  const long v = g_wait_limit.load();
  return v > 0 ? v : dflt;
}

void report_failure(const char* rout, int status, const char* msg) {
  if (g_failure_hook) {
    g_failure_hook(rout, status, msg);
    return;
  }
  fprintf(stderr, "%s: %s (failure bits 0x%x)\n", rout, msg, status);
  abort();  // never return a silent partial result
}

namespace {
void teardown() {
  std::lock_guard<std::mutex> lk(g_mu);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess) return;
  for (int d = 0; d < ndev; d++) {
    if (hipSetDevice(d) == hipSuccess) (void)hipDeviceSynchronize();
  }
  for (auto& kv : g_slots)
    if (kv.second.p) (void)hipFree(kv.second.p);
  g_slots.clear();
  for (auto& kv : g_status)
    if (kv.second) (void)hipFree(kv.second);
  g_status.clear();
  for (auto& kv : g_aux) {
    Aux& a = kv.second;
    for (auto& sp : a.split) {
      if (sp.second.first) (void)hipStreamDestroy(sp.second.first);
      if (sp.second.second) (void)hipStreamDestroy(sp.second.second);
    }
    a.split.clear();
    for (auto& e : a.ev)
      if (e) (void)hipEventDestroy(e);
    if (a.stream) (void)hipStreamDestroy(a.stream);
    a.stream = nullptr;
  }
  g_aux.clear();
  for (auto& kv : g_streams) (void)hipStreamDestroy(kv.second);
  g_streams.clear();
  (void)hipGetLastError();
}
}  // namespace

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
  return guarded([] {
    (void)library_stream();
    (void)noalloc_scratch();  // the allocation-free TRSM's array on this device
  });
}

const char* atlg_last_error(void) { return g_last_error.c_str(); }

int atlg_stream_status(atlg_stream_t stream, int* status) {
  if (!status) return -2;
  return guarded([&] { *status = take_status(resolve_stream(stream)); });
}

void atlg_set_failure_hook(void (*hook)(const char*, int, const char*)) {
  g_failure_hook = hook;
}

void atlg_set_wait_limit(long polls) { g_wait_limit.store(polls); }

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

int atlg_dtrsm_noalloc(int Side, int Uplo, int TA, int Diag, long M, long N, double alpha,
                       const double* A, long lda, double* B, long ldb, atlg_stream_t stream) {
  if (bad_side(Side)) return -1;
  if (bad_uplo(Uplo)) return -2;
  if (bad_trans(TA)) return -3;
  if (bad_diag(Diag)) return -4;
  if (M < 0) return -5;
  if (N < 0) return -6;
  if (lda < std::max(1L, Side == kLeft ? M : N)) return -9;
  if (ldb < std::max(1L, M)) return -11;
  const long na = Side == kLeft ? M : N;
  if (na > kNoAllocMaxOrder) {
    set_error("atlg_dtrsm_noalloc: order of op(A) " + std::to_string(na) +
              " exceeds the allocation-free maximum " + std::to_string(kNoAllocMaxOrder));
    return ATLG_ESIZE;
  }
  return guarded([&] {
    trsm_noalloc(resolve_stream(stream), Side, Uplo, TA, Diag, M, N, alpha, A, lda, B, ldb);
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

int atlg_dtrmm(int Side, int Uplo, int TA, int Diag, long M, long N, double alpha,
               const double* A, long lda, double* B, long ldb, atlg_stream_t stream) {
  if (bad_side(Side)) return -1;
  if (bad_uplo(Uplo)) return -2;
  if (bad_trans(TA)) return -3;
  if (bad_diag(Diag)) return -4;
  if (M < 0) return -5;
  if (N < 0) return -6;
  if (lda < std::max(1L, Side == kLeft ? M : N)) return -9;
  if (ldb < std::max(1L, M)) return -11;
  return guarded([&] {
    trmm(resolve_stream(stream), Side, Uplo, TA, Diag, M, N, alpha, A, lda, B, ldb);
  });
}

int atlg_dsymm(int Side, int Uplo, long M, long N, double alpha, const double* A, long lda,
               const double* B, long ldb, double beta, double* C, long ldc,
               atlg_stream_t stream) {
  if (bad_side(Side)) return -1;
  if (bad_uplo(Uplo)) return -2;
  if (M < 0) return -3;
  if (N < 0) return -4;
  if (lda < std::max(1L, Side == kLeft ? M : N)) return -7;
  if (ldb < std::max(1L, M)) return -9;
  if (ldc < std::max(1L, M)) return -12;
  return guarded([&] {
    symm(resolve_stream(stream), Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
  });
}

int atlg_dsyr2k(int Uplo, int Trans, long N, long K, double alpha, const double* A, long lda,
                const double* B, long ldb, double beta, double* C, long ldc,
                atlg_stream_t stream) {
  if (bad_uplo(Uplo)) return -1;
  if (bad_trans(Trans)) return -2;
  if (N < 0) return -3;
  if (K < 0) return -4;
  if (lda < std::max(1L, is_trans(Trans) ? K : N)) return -7;
  if (ldb < std::max(1L, is_trans(Trans) ? K : N)) return -9;
  if (ldc < std::max(1L, N)) return -12;
  return guarded([&] {
    syr2k(resolve_stream(stream), Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  });
}

int atlg_dtrtri(int Uplo, int Diag, long N, double* A, long lda, int* dinfo,
                atlg_stream_t stream) {
  if (bad_uplo(Uplo)) return -1;
  if (bad_diag(Diag)) return -2;
  if (N < 0) return -3;
  if (lda < std::max(1L, N)) return -5;
  if (!dinfo) return -6;
  return guarded([&] { trtri(resolve_stream(stream), Uplo, Diag, N, A, lda, dinfo); });
}

int atlg_dlauum(int Uplo, long N, double* A, long lda, atlg_stream_t stream) {
  if (bad_uplo(Uplo)) return -1;
  if (N < 0) return -2;
  if (lda < std::max(1L, N)) return -4;
  return guarded([&] { lauum(resolve_stream(stream), Uplo, N, A, lda); });
}

int atlg_dpotri(int Uplo, long N, double* A, long lda, int* dinfo, atlg_stream_t stream) {
  if (bad_uplo(Uplo)) return -1;
  if (N < 0) return -2;
  if (lda < std::max(1L, N)) return -4;
  if (!dinfo) return -5;
  return guarded([&] { potri(resolve_stream(stream), Uplo, N, A, lda, dinfo); });
}

int atlg_dpotrf_adjoint(long N, const double* L, long ldl, double* G, long ldg,
                        atlg_stream_t stream) {
  if (N < 0) return -1;
  if (ldl < std::max(1L, N)) return -3;
  if (ldg < std::max(1L, N)) return -5;
  return guarded([&] { potrf_adjoint(resolve_stream(stream), N, L, ldl, G, ldg); });
}

int atlg_dtrsm_adjoint(long N, long NRHS, const double* L, long ldl, const double* X, long ldx,
                       double* Xbar, long ldxb, double* Lbar, long ldlb, atlg_stream_t stream) {
  if (N < 0) return -1;
  if (NRHS < 0) return -2;
  if (ldl < std::max(1L, N)) return -4;
  if (ldx < std::max(1L, N)) return -6;
  if (ldxb < std::max(1L, N)) return -8;
  if (Lbar && ldlb < std::max(1L, N)) return -10;
  return guarded([&] {
    trsm_adjoint(resolve_stream(stream), N, NRHS, L, ldl, X, ldx, Xbar, ldxb, Lbar, ldlb);
  });
}

int atlg_dlogdet_chol(long N, const double* L, long ldl, double* out, atlg_stream_t stream) {
  if (N < 0) return -1;
  if (ldl < std::max(1L, N)) return -3;
  return guarded([&] { logdet_chol(resolve_stream(stream), N, L, ldl, out); });
}

int atlg_dlogdet_chol_adjoint(long N, const double* L, long ldl, const double* dbar,
                              double* Lbar, long ldlb, atlg_stream_t stream) {
  if (N < 0) return -1;
  if (ldl < std::max(1L, N)) return -3;
  if (ldlb < std::max(1L, N)) return -6;
  return guarded([&] { logdet_chol_adjoint(resolve_stream(stream), N, L, ldl, dbar, Lbar, ldlb); });
}


int atlg_sgemm(int TA, int TB, long M, long N, long K, float alpha, const float* A, long lda,
               const float* B, long ldb, float beta, float* C, long ldc, atlg_stream_t stream) {
  if (bad_trans(TA)) return -1;
  if (bad_trans(TB)) return -2;
  if (M < 0) return -3;
  if (N < 0) return -4;
  if (K < 0) return -5;
  if (lda < std::max(1L, is_trans(TA) ? K : M)) return -8;
  if (ldb < std::max(1L, is_trans(TB) ? N : K)) return -10;
  if (ldc < std::max(1L, M)) return -13;
  return guarded([&] {
    gemm<float>(resolve_stream(stream), TA, TB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  });
}

int atlg_sgemm_bf16(long M, long N, long K, float alpha, const float* A, long lda,
                    const float* B, long ldb, float beta, float* C, long ldc, int split,
                    atlg_stream_t stream) {
  if (M < 0) return -1;
  if (N < 0) return -2;
  if (K < 0) return -3;
  if (lda < std::max(1L, M)) return -6;
  if (ldb < std::max(1L, K)) return -8;
  if (ldc < std::max(1L, M)) return -11;
  if (split < 0 || split > 2) return -12;
  return guarded([&] {
    gemm_bf16(resolve_stream(stream), M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
              split == 1 ? 3 : (split == 2 ? 2 : 1));
  });
}

int atlg_sgetrf(long M, long N, float* A, long lda, int* ipiv, int* dinfo, int gemm_prec,
                atlg_stream_t stream) {
  if (M < 0) return -1;
  if (N < 0) return -2;
  if (lda < std::max(1L, M)) return -4;
  if (!dinfo) return -6;
  if (gemm_prec < kGemmNative || gemm_prec > kGemmBf16x3) return -7;
  return guarded(
      [&] { getrf<float>(resolve_stream(stream), M, N, A, lda, ipiv, dinfo, gemm_prec); });
}

int atlg_sgetrs(int Trans, long N, long NRHS, const float* A, long lda, const int* ipiv,
                float* B, long ldb, atlg_stream_t stream) {
  if (bad_trans(Trans)) return -1;
  if (N < 0) return -2;
  if (NRHS < 0) return -3;
  if (lda < std::max(1L, N)) return -5;
  if (ldb < std::max(1L, N)) return -8;
  return guarded(
      [&] { getrs<float>(resolve_stream(stream), Trans, N, NRHS, A, lda, ipiv, B, ldb); });
}

int atlg_dsgesv(long N, long NRHS, double* A, long lda, int* ipiv, const double* B, long ldb,
                double* X, long ldx, int* dinfo, int itermax, int lowprec, int refine, int* iter,
                int* inner, atlg_stream_t stream) {
  if (N < 0) return -1;
  if (NRHS < 0) return -2;
  if (lda < std::max(1L, N)) return -4;
  if (ldb < std::max(1L, N)) return -7;
  if (ldx < std::max(1L, N)) return -9;
  if (!dinfo) return -10;
  if (lowprec < 0 || lowprec > 3) return -12;
  if (refine != 0 && refine != 1) return -13;
  if (!iter) return -14;
  int st = 0;
  const int rc = guarded([&] {
    hipStream_t s = resolve_stream(stream);
    dsgesv(s, N, NRHS, A, lda, ipiv, B, ldb, X, ldx, dinfo, itermax > 0 ? itermax : 30, lowprec,
           refine, iter, inner);
    st = take_status(s);
  });
  if (rc == 0 && st != 0) {
    set_error("atlg_dsgesv: a bounded wait gave up (failure bits " + std::to_string(st) + ")");
    return ATLG_EWAIT;
  }
  return rc;
}

namespace {
// Argument check shared by the batched entry points; returns 0 or -pos.
int check_batch(int nsys, const int* n, const void* A, const int* lda, int pos_n, int pos_A,
                int pos_lda) {
  if (nsys < 0) return -1;
  if (nsys == 0) return 0;
  if (!n) return -pos_n;
  if (!A) return -pos_A;
  if (!lda) return -pos_lda;
  for (int k = 0; k < nsys; k++) {
    if (n[k] < 0) return -pos_n;
    if (lda[k] < std::max(1, n[k])) return -pos_lda;
  }
  return 0;
}
int check_rhs(int nsys, const int* n, const void* B, const int* ldb, int pos_B, int pos_ldb) {
  if (nsys == 0) return 0;
  if (!B) return -pos_B;
  if (!ldb) return -pos_ldb;
  for (int k = 0; k < nsys; k++)
    if (ldb[k] < std::max(1, n[k])) return -pos_ldb;
  return 0;
}
}  // namespace

int atlg_dpotrf_batched(int nsys, const int* n, double* const* A, const int* lda, int* info,
                        atlg_stream_t stream) {
  if (int e = check_batch(nsys, n, A, lda, 2, 3, 4)) return e;
  if (nsys > 0 && !info) return -5;
  return guarded([&] {
    posv_batched<double>(resolve_stream(stream), 0, nsys, n, A, lda, nullptr, nullptr, nullptr,
                         0, info, batch_tune(false));
  });
}

int atlg_dpotrs_batched(int nsys, const int* n, double* const* A, const int* lda,
                        double* const* B, const int* ldb, int nrhs, atlg_stream_t stream) {
  if (int e = check_batch(nsys, n, A, lda, 2, 3, 4)) return e;
  if (int e = check_rhs(nsys, n, B, ldb, 5, 6)) return e;
  if (nrhs < 0) return -7;
  if (nrhs == 0) return 0;
  return guarded([&] {
    posv_batched<double>(resolve_stream(stream), 2, nsys, n, A, lda, B, ldb, nullptr, nrhs,
                         nullptr, batch_tune(false));
  });
}

int atlg_dposv_batched(int nsys, const int* n, double* const* A, const int* lda,
                       double* const* B, const int* ldb, int nrhs, int* info,
                       atlg_stream_t stream) {
  if (int e = check_batch(nsys, n, A, lda, 2, 3, 4)) return e;
  if (int e = check_rhs(nsys, n, B, ldb, 5, 6)) return e;
  if (nrhs < 0) return -7;
  if (nsys > 0 && !info) return -8;
  return guarded([&] {
    posv_batched<double>(resolve_stream(stream), nrhs ? 1 : 0, nsys, n, A, lda, B, ldb, nullptr,
                         nrhs, info, batch_tune(false));
  });
}

int atlg_spotrf_batched(int nsys, const int* n, float* const* A, const int* lda, int* info,
                        atlg_stream_t stream) {
  if (int e = check_batch(nsys, n, A, lda, 2, 3, 4)) return e;
  if (nsys > 0 && !info) return -5;
  return guarded([&] {
    posv_batched<float>(resolve_stream(stream), 0, nsys, n, A, lda, nullptr, nullptr, nullptr, 0,
                        info, batch_tune(true));
  });
}

int atlg_sposv_batched(int nsys, const int* n, float* const* A, const int* lda, float* const* B,
                       const int* ldb, int nrhs, int* info, atlg_stream_t stream) {
  if (int e = check_batch(nsys, n, A, lda, 2, 3, 4)) return e;
  if (int e = check_rhs(nsys, n, B, ldb, 5, 6)) return e;
  if (nrhs < 0) return -7;
  if (nsys > 0 && !info) return -8;
  return guarded([&] {
    posv_batched<float>(resolve_stream(stream), nrhs ? 1 : 0, nsys, n, A, lda, B, ldb, nullptr,
                        nrhs, info, batch_tune(true));
  });
}

int atlg_dsposv_batched(int nsys, const int* n, double* const* A, const int* lda,
                        const double* const* B, const int* ldb, double* const* X, const int* ldx,
                        int nrhs, int itermax, int* info, int* iter, atlg_stream_t stream) {
  if (int e = check_batch(nsys, n, A, lda, 2, 3, 4)) return e;
  if (int e = check_rhs(nsys, n, B, ldb, 5, 6)) return e;
  if (int e = check_rhs(nsys, n, X, ldx, 7, 8)) return e;
  if (nrhs < 1) return -9;
  if (nsys > 0 && !info) return -11;
  if (nsys > 0 && !iter) return -12;
  return guarded([&] {
    dsposv_batched(resolve_stream(stream), nsys, n, A, lda, B, ldb, X, ldx, nrhs, info, iter,
                   itermax > 0 ? itermax : 30, batch_tune(true));
  });
}

int atlg_batched_tune_get(int prec, int* t16_max, int* wg_per_cu) {
  if (prec != 0 && prec != 1) return -1;
  const BatchTune t = batch_tune(prec == 1);
  if (t16_max) *t16_max = t.t16_max;
  if (wg_per_cu) *wg_per_cu = t.wg_per_cu;
  return 0;
}

int atlg_batched_tune_set(int prec, int t16_max, int wg_per_cu) {
  if (prec != 0 && prec != 1) return -1;
  if (wg_per_cu < 1 || wg_per_cu > 16) return -3;
  set_batch_tune(prec == 1, BatchTune{t16_max, wg_per_cu});
  return 0;
}

}  // extern "C"
