// Cholesky factorization and solve.
//
// Replaces ATL_dpotrf (AtlasBase/Clint/atlas-lp.base:332-360), the recursive
// ATL_dpotrfL (:909-968) / ATL_dpotrfU (:675-733) with their small closed
// forms (:809-908), and ATL_dpotrs (:969-1050).
//
// Same recursion as ATLAS: factor the leading half, TRSM the off-diagonal
// block, SYRK the trailing block, factor it.  The recursion stops at
// diagonal blocks of order <= DB, factored by one workgroup in registers,
// which also writes the inverse of its factor; the TRSMs of the recursion
// apply those inverses as GEMMs instead of inverting again.  The
// first non-positive pivot writes info to device memory; every later launch
// of the factorization reads that word and returns at once, which gives
// ATLAS's "stop at the first failure" without a host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "atlg_internal.hpp"
#include "tri_inv.hpp"

namespace atlg {
namespace {

constexpr int DB = kTrsmNB;    // diagonal block order
constexpr int LDS_LD = kTriLD;
constexpr int P_THREADS = kTriThreads;  // 16 x 16 grid of 8 x 8 register tiles
static_assert(DB == kTriN, "leaf order must match the block inverter");

// Factors the n x n (n <= DB) diagonal block at A and writes inv(L) of the
// block to Dinv (DB x DB, column-major, zeros above the diagonal).
// lower: A = L L^T from the lower triangle; else A = U^T U from the upper
// triangle, handled as L = U^T, so one algorithm serves both.  The other
// triangle of A is never read or written.
//
// Factorization: thread (tr, tc) holds the 8x8 tile at rows 8tr.., columns
// 8tc.. of the lower triangle in registers; per column, the diagonal owner
// publishes a_jj, the owners of column j scale it by 1/sqrt(a_jj) (the
// reciprocal multiply of ATLAS's potrfL closed forms) and publish it, and
// every lower tile applies the rank-1 update.  Order beyond n is padded with
// the identity.  The factor is then inverted in LDS (lower_inverse_inplace).
__global__ void __launch_bounds__(P_THREADS)
potf2_inv_kernel(int n, double* __restrict__ A, long lda, int lower, int* __restrict__ dinfo,
                 int info_off, double* __restrict__ Dinv) {
  if (*dinfo != 0) return;
  __shared__ double S[DB * LDS_LD];
  __shared__ double col[DB];
  __shared__ double dsh;
  const int t = threadIdx.x, tr = t >> 4, tc = t & 15;
  const bool act = tr >= tc;
  double a[8][8];
#pragma unroll
  for (int r = 0; r < 8; r++)
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int i = 8 * tr + r, k = 8 * tc + c;
      double v = 0.0;
      if (act && i >= k) {
        if (i < n && k < n) v = lower ? A[i + (size_t)k * lda] : A[k + (size_t)i * lda];
        else if (i == k) v = 1.0;
      }
      a[r][c] = v;
    }

  int fail = -1;
  for (int j0 = 0; j0 < DB / 8 && fail < 0; j0++) {
#pragma unroll
    for (int jj = 0; jj < 8; jj++) {
      const int j = 8 * j0 + jj;
      if (tr == j0 && tc == j0) dsh = a[jj][jj];
      __syncthreads();
      const double d = dsh;
      if (!(d > 0.0)) {  // same value in every thread: uniform exit
        fail = j;
        break;
      }
      const double l = sqrt(d), rinv = 1.0 / l;
      if (tc == j0 && tr >= j0) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
          const int i = 8 * tr + r;
          if (i == j) {
            a[r][jj] = l;
          } else if (i > j) {
            a[r][jj] *= rinv;
            col[i] = a[r][jj];
          }
        }
      }
      __syncthreads();
      if (act && tc >= j0) {
        double ci[8], ck[8];
#pragma unroll
        for (int r = 0; r < 8; r++) ci[r] = col[8 * tr + r];
#pragma unroll
        for (int c = 0; c < 8; c++) ck[c] = col[8 * tc + c];
#pragma unroll
        for (int r = 0; r < 8; r++)
#pragma unroll
          for (int c = 0; c < 8; c++) {
            const int i = 8 * tr + r, k = 8 * tc + c;
            if (k > j && i >= k) a[r][c] -= ci[r] * ck[c];
          }
      }
    }
  }
  if (fail >= 0) {
    if (t == 0) *dinfo = info_off + fail + 1;
  }
  // write back (on failure: the partially factored block, as ATLAS leaves it)
#pragma unroll
  for (int r = 0; r < 8; r++)
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int i = 8 * tr + r, k = 8 * tc + c;
      if (act && i >= k && i < n && k < n) {
        if (lower) A[i + (size_t)k * lda] = a[r][c];
        else A[k + (size_t)i * lda] = a[r][c];
      }
    }
  if (fail >= 0 || Dinv == nullptr) return;

  // ---- inverse of the factor ----
#pragma unroll
  for (int r = 0; r < 8; r++)
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int i = 8 * tr + r, k = 8 * tc + c;
      S[i * LDS_LD + k] = (act && i >= k) ? a[r][c] : 0.0;
    }
  __syncthreads();
  lower_inverse_inplace(S);
  for (int idx = t; idx < DB * DB; idx += P_THREADS) {
    const int i = idx % DB, k = idx / DB;
    Dinv[idx] = (i >= k) ? S[i * LDS_LD + k] : 0.0;
  }
}

void potrf_rec(hipStream_t s, bool lower, long N, double* A, long lda, int* dinfo, int off,
               double* W) {
  if (N <= 0) return;
  if (N <= DB) {
    hipLaunchKernelGGL(potf2_inv_kernel, dim3(1), dim3(P_THREADS), 0, s, (int)N, A, lda,
                       lower ? 1 : 0, dinfo, off, W + (size_t)(off / DB) * DB * DB);
    ATLG_LAUNCHED();
    return;
  }
  long nl = ((N / 2) / DB) * DB;
  if (nl < DB) nl = DB;
  const long nr = N - nl;
  potrf_rec(s, lower, nl, A, lda, dinfo, off, W);
  double* A22 = A + nl + nl * lda;
  const double* Winv = W + (size_t)(off / DB) * DB * DB;  // inverses of this node's A11
  if (lower) {
    // A21 <- A21 L11^-T: op(A) = L11^T, whose diagonal-block inverses are
    // the transposes of the stored inv(L_bb)
    double* A21 = A + nl;
    trsm(s, kRight, kLower, kTrans, kNonUnit, nr, nl, 1.0, A, lda, A21, lda, dinfo, Winv, 1);
    gemm(s, kNoTrans, kTrans, nr, nr, nl, -1.0, A21, lda, A21, lda, 1.0, A22, lda, 1, dinfo);
  } else {
    // A12 <- U11^-T A12: op(A) = U11^T = L11, inverses stored as is
    double* A12 = A + nl * lda;
    trsm(s, kLeft, kUpper, kTrans, kNonUnit, nl, nr, 1.0, A, lda, A12, lda, dinfo, Winv, 0);
    gemm(s, kTrans, kNoTrans, nr, nr, nl, -1.0, A12, lda, A12, lda, 1.0, A22, lda, 2, dinfo);
  }
  potrf_rec(s, lower, nr, A22, lda, dinfo, off + (int)nl, W);
}

}  // namespace

void potrf(hipStream_t s, int uplo, long N, double* A, long lda, int* dinfo) {
  set_int(s, dinfo, 0);
  if (N <= 0) return;
  const long nblk = (N + DB - 1) / DB;
  double* W = static_cast<double*>(scratch(s, kSlotFactInv, sizeof(double) * nblk * DB * DB));
  potrf_rec(s, uplo == kLower, N, A, lda, dinfo, 0, W);
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
