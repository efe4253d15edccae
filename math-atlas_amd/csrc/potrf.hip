// Cholesky factorization and solve.
//
// Replaces ATL_dpotrf (AtlasBase/Clint/atlas-lp.base:332-360), the recursive
// ATL_dpotrfL (:909-968) / ATL_dpotrfU (:675-733) with their small closed
// forms (:809-908), and ATL_dpotrs (:969-1050).
//
// Same recursion as ATLAS: factor the leading half, TRSM the off-diagonal
// block, SYRK the trailing block, factor it.  The recursion stops at
// diagonal blocks of order <= DB, factored by one workgroup in LDS.  The
// first non-positive pivot writes info to device memory; every later launch
// of the factorization reads that word and returns at once, which gives
// ATLAS's "stop at the first failure" without a host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "atlg_internal.hpp"

namespace atlg {
namespace {

constexpr int DB = 128;
constexpr int LDS_LD = DB + 1;
constexpr int P_THREADS = 256;

// Factors the n x n (n <= DB) diagonal block at A.  lower: A = L L^T from the
// lower triangle; else A = U^T U from the upper triangle (held transposed in
// LDS, so one lower algorithm serves both).  The other triangle is never
// read or written.
__global__ void __launch_bounds__(P_THREADS)
potf2_kernel(int n, double* __restrict__ A, long lda, int lower, int* __restrict__ dinfo,
             int info_off) {
  if (*dinfo != 0) return;
  __shared__ double S[DB * LDS_LD];
  const int t = threadIdx.x;
  for (int idx = t; idx < n * n; idx += P_THREADS) {
    const int i = idx % n, j = idx / n;
    if (i >= j) S[i * LDS_LD + j] = lower ? A[i + (size_t)j * lda] : A[j + (size_t)i * lda];
  }
  __syncthreads();
  int jend = n;
  for (int j = 0; j < n; j++) {
    const double d = S[j * LDS_LD + j];
    if (!(d > 0.0)) {
      if (t == 0) *dinfo = info_off + j + 1;
      jend = j;
      break;
    }
    const double l = sqrt(d), rinv = 1.0 / l;
    __syncthreads();
    if (t == 0) S[j * LDS_LD + j] = l;
    for (int i = j + 1 + t; i < n; i += P_THREADS) S[i * LDS_LD + j] *= rinv;
    __syncthreads();
    const int m = n - j - 1;  // trailing order
    for (int e = t; e < m * m; e += P_THREADS) {
      const int i = j + 1 + e % m, k = j + 1 + e / m;
      if (k <= i) S[i * LDS_LD + k] -= S[i * LDS_LD + j] * S[k * LDS_LD + j];
    }
    __syncthreads();
  }
  __syncthreads();
  // write back what was computed (columns < jend factored; ATLAS likewise
  // leaves a partially updated block behind on failure)
  for (int idx = t; idx < n * n; idx += P_THREADS) {
    const int i = idx % n, j = idx / n;
    if (i >= j && (j < jend || jend == n)) {
      if (lower) A[i + (size_t)j * lda] = S[i * LDS_LD + j];
      else A[j + (size_t)i * lda] = S[i * LDS_LD + j];
    }
  }
}

void potrf_rec(hipStream_t s, bool lower, long N, double* A, long lda, int* dinfo, int off) {
  if (N <= 0) return;
  if (N <= DB) {
    hipLaunchKernelGGL(potf2_kernel, dim3(1), dim3(P_THREADS), 0, s, (int)N, A, lda,
                       lower ? 1 : 0, dinfo, off);
    ATLG_LAUNCHED();
    return;
  }
  long nl = ((N / 2) / DB) * DB;
  if (nl < DB) nl = DB;
  const long nr = N - nl;
  potrf_rec(s, lower, nl, A, lda, dinfo, off);
  double* A22 = A + nl + nl * lda;
  if (lower) {
    double* A21 = A + nl;
    trsm(s, kRight, kLower, kTrans, kNonUnit, nr, nl, 1.0, A, lda, A21, lda, dinfo);
    gemm(s, kNoTrans, kTrans, nr, nr, nl, -1.0, A21, lda, A21, lda, 1.0, A22, lda, 1, dinfo);
  } else {
    double* A12 = A + nl * lda;
    trsm(s, kLeft, kUpper, kTrans, kNonUnit, nl, nr, 1.0, A, lda, A12, lda, dinfo);
    gemm(s, kTrans, kNoTrans, nr, nr, nl, -1.0, A12, lda, A12, lda, 1.0, A22, lda, 2, dinfo);
  }
  potrf_rec(s, lower, nr, A22, lda, dinfo, off + (int)nl);
}

}  // namespace

void potrf(hipStream_t s, int uplo, long N, double* A, long lda, int* dinfo) {
  set_int(s, dinfo, 0);
  potrf_rec(s, uplo == kLower, N, A, lda, dinfo, 0);
}

void potrs(hipStream_t s, int uplo, long N, long NRHS, const double* A, long lda, double* B,
           long ldb) {
  if (N == 0 || NRHS == 0) return;
  if (uplo == kUpper) {
    trsm(s, kLeft, kUpper, kTrans, kNonUnit, N, NRHS, 1.0, A, lda, B, ldb);
    trsm(s, kLeft, kUpper, kNoTrans, kNonUnit, N, NRHS, 1.0, A, lda, B, ldb);
  } else {
    trsm(s, kLeft, kLower, kNoTrans, kNonUnit, N, NRHS, 1.0, A, lda, B, ldb);
    trsm(s, kLeft, kLower, kTrans, kNonUnit, N, NRHS, 1.0, A, lda, B, ldb);
  }
}

}  // namespace atlg
