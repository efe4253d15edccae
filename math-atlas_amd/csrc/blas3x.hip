// The level-3 neighbours of the hot path, on the same GEMM/TRSM kernels:
// TRMM, SYMM and SYR2K (the rest of ATLAS's real level-3 BLAS), and the
// triangular inverse, LAUUM and the SPD inverse built on them.
//
// Replaces ATL_dtrmm (AtlasBase/Clint/trmm-micro.base:1-915), ATL_dsymm and
// ATL_dsyr2k (AtlasBase/Clint/l3micro.base:4803-7058, :771), ATL_dtrtri
// (AtlasBase/Clint/atlas-lp.base:1703-2340, 2341-2380: the column-major
// recursions ATL_trtriCL/CU), ATL_dlauum (atlas-lp.base:221-330) and
// clapack_dpotri (AtlasBase/Clint/atlas-clp.base:316-362).
//
//  * TRMM: the recursive halving of TRSM's driver, run as a product: each
//    level updates one half with a GEMM from the other half's original
//    values before that half is overwritten, and a leaf (order <= 128)
//    multiplies by op(A)'s diagonal block materialized as a dense 128 x 128
//    tile (zeros outside the triangle, ones on a unit diagonal).
//  * SYMM: A's stored triangle is mirrored into a dense scratch copy, then
//    one GEMM.  SYR2K: two triangle-masked GEMMs into C.
//  * TRTRI: ATLAS's recursion (split, A21 <- A21 inv(A11), A21 <- -inv(A22)
//    A21, recurse into both halves) with the diagonal 128-blocks inverted
//    once up front (they serve every TRSM of the recursion as GEMMs) and
//    copied into A at the end.  A zero diagonal entry (NonUnit) is found
//    first, as ATL_trtri does, and then nothing is written (a device info
//    word makes every later launch a no-op).
//  * LAUUM: ATLAS's recursion (lauum(L11); A11 += L21^T L21; L21 <-
//    L22^T L21; lauum(L22)), with a leaf kernel for blocks <= 128.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "atlg_internal.hpp"

namespace atlg {
namespace {

constexpr int NB = kTrsmNB;  // leaf order / diagonal block order

// Element (i, j) of op(A).
__device__ inline double opa(const double* A, long lda, bool trans, long i, long j) {
  return trans ? A[j + i * lda] : A[i + j * lda];
}

// Dense copies of op(A)'s diagonal NB-blocks (order na): block b at
// T + b*NB*NB, leading dimension NB, zeros outside the effective triangle,
// ones on the diagonal when unit; the part beyond na is zero.
__global__ void __launch_bounds__(256)
tri_blocks_kernel(int na, const double* __restrict__ A, long lda, int trans, int lower_eff,
                  int unit, double* __restrict__ T, const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int b = blockIdx.x;
  const long o = (long)b * NB;
  const int n = (int)min((long)NB, na - o);
  double* D = T + (size_t)b * NB * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int i = idx % NB, j = idx / NB;
    double v = 0.0;
    if (i < n && j < n && (lower_eff ? i >= j : i <= j)) {
      if (i == j && unit) v = 1.0;
      else v = opa(A, lda, trans != 0, o + i, o + j);
    }
    D[idx] = v;
  }
}

struct Tm {
  hipStream_t s;
  bool left, lower_eff, trans;
  const double* A;
  long lda;
  double* B;
  long ldb;
  long M, N;
  const double* T;  // dense diagonal blocks of op(A)
  const int* skip;  // device word: nonzero makes every launch a no-op
};

inline void opa_blk(const Tm& c, long i0, long j0, const double** p, int* t) {
  if (!c.trans) {
    *p = c.A + i0 + j0 * c.lda;
    *t = kNoTrans;
  } else {
    *p = c.A + j0 + i0 * c.lda;
    *t = kTrans;
  }
}

long half_split(long n) {
  long n1 = ((n / 2) / NB) * NB;
  return n1 > 0 ? n1 : NB;
}

// B <- alpha * op(A)[r0:r0+n] * B (left) or alpha * B * op(A)[...] (right)
// on rows / columns [r0, r0 + n).
void tm_rec(const Tm& c, long r0, long n, double alpha) {
  if (n <= NB) {
    const double* T = c.T + (size_t)(r0 / NB) * NB * NB;
    if (c.left) {
      double* Bb = c.B + r0;
      gemm<double>(c.s, kNoTrans, kNoTrans, n, c.N, n, alpha, T, NB, Bb, c.ldb, 0.0, Bb, c.ldb, 0,
                   c.skip);
    } else {
      double* Bb = c.B + r0 * c.ldb;
      gemm<double>(c.s, kNoTrans, kNoTrans, c.M, n, n, alpha, Bb, c.ldb, T, NB, 0.0, Bb, c.ldb, 0,
                   c.skip);
    }
    return;
  }
  const long n1 = half_split(n), n2 = n - n1;
  const double* Ap;
  int ta;
  if (c.left) {
    double* B1 = c.B + r0;
    double* B2 = c.B + r0 + n1;
    if (c.lower_eff) {  // X2 = A21 B1 + A22 B2, X1 = A11 B1
      tm_rec(c, r0 + n1, n2, alpha);
      opa_blk(c, r0 + n1, r0, &Ap, &ta);
      gemm<double>(c.s, ta, kNoTrans, n2, c.N, n1, alpha, Ap, c.lda, B1, c.ldb, 1.0, B2, c.ldb, 0,
                   c.skip);
      tm_rec(c, r0, n1, alpha);
    } else {  // X1 = A11 B1 + A12 B2, X2 = A22 B2
      tm_rec(c, r0, n1, alpha);
      opa_blk(c, r0, r0 + n1, &Ap, &ta);
      gemm<double>(c.s, ta, kNoTrans, n1, c.N, n2, alpha, Ap, c.lda, B2, c.ldb, 1.0, B1, c.ldb, 0,
                   c.skip);
      tm_rec(c, r0 + n1, n2, alpha);
    }
  } else {
    double* B1 = c.B + r0 * c.ldb;
    double* B2 = c.B + (r0 + n1) * c.ldb;
    if (c.lower_eff) {  // X1 = B1 A11 + B2 A21, X2 = B2 A22
      tm_rec(c, r0, n1, alpha);
      opa_blk(c, r0 + n1, r0, &Ap, &ta);
      gemm<double>(c.s, kNoTrans, ta, c.M, n1, n2, alpha, B2, c.ldb, Ap, c.lda, 1.0, B1, c.ldb, 0,
                   c.skip);
      tm_rec(c, r0 + n1, n2, alpha);
    } else {  // X2 = B1 A12 + B2 A22, X1 = B1 A11
      tm_rec(c, r0 + n1, n2, alpha);
      opa_blk(c, r0, r0 + n1, &Ap, &ta);
      gemm<double>(c.s, kNoTrans, ta, c.M, n2, n1, alpha, B1, c.ldb, Ap, c.lda, 1.0, B2, c.ldb, 0,
                   c.skip);
      tm_rec(c, r0, n1, alpha);
    }
  }
}

// S = the symmetric matrix whose `lower` (else upper) triangle is stored in
// A (n x n), written densely with leading dimension n.
__global__ void sym_fill_kernel(int n, const double* __restrict__ A, long lda, int lower,
                                double* __restrict__ S) {
  const long total = (long)n * n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const bool stored = lower ? i >= j : i <= j;
    S[idx] = stored ? A[i + (size_t)j * lda] : A[j + (size_t)i * lda];
  }
}

// *dinfo = 1 + the first zero on A's diagonal (0 if none).
__global__ void diag_zero_kernel(int n, const double* __restrict__ A, long lda, int* dinfo) {
  __shared__ int first;
  if (threadIdx.x == 0) first = n;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    if (A[i + (size_t)i * lda] == 0.0) atomicMin(&first, i);
  __syncthreads();
  if (threadIdx.x == 0) *dinfo = first < n ? first + 1 : 0;
}

// Copies the diagonal-block inverses (ld NB) into A's triangle (the
// diagonal too unless unit); the other triangle is not touched.
__global__ void __launch_bounds__(256)
put_tri_blocks_kernel(int na, double* __restrict__ A, long lda, int lower, int unit,
                      const double* __restrict__ D, const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int b = blockIdx.x;
  const long o = (long)b * NB;
  const int n = (int)min((long)NB, na - o);
  const double* Db = D + (size_t)b * NB * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int i = idx % NB, j = idx / NB;
    if (i >= n || j >= n) continue;
    if (lower ? i < j : i > j) continue;
    if (i == j && unit) continue;
    A[(o + i) + (size_t)(o + j) * lda] = Db[idx];
  }
}

// In-place LAUUM of an n <= NB block: lower: A <- L^T L (lower triangle);
// upper: A <- U U^T (upper triangle).  The block goes to LDS; each thread
// computes entries of the result.
__global__ void __launch_bounds__(256)
lauum_leaf_kernel(int n, double* __restrict__ A, long lda, int lower,
                  const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double S[NB * (NB + 1)];
  constexpr int LD = NB + 1;
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int i = idx % NB, j = idx / NB;
    double v = 0.0;
    if (i < n && j < n && (lower ? i >= j : i <= j)) v = A[i + (size_t)j * lda];
    S[i * LD + j] = v;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    const int i = idx % n, j = idx / n;
    if (lower ? i < j : i > j) continue;
    double acc = 0.0;
    if (lower) {  // (L^T L)(i, j) = sum_{k >= i} L(k, i) L(k, j), i >= j
      for (int k = i; k < n; k++) acc += S[k * LD + i] * S[k * LD + j];
    } else {      // (U U^T)(i, j) = sum_{k >= j} U(i, k) U(j, k), i <= j
      for (int k = j; k < n; k++) acc += S[i * LD + k] * S[j * LD + k];
    }
    A[i + (size_t)j * lda] = acc;
  }
}

// ATL_lauumL/U's recursion (atlas-lp.base:242-315), on the device.
void lauum_rec(hipStream_t s, bool lower, long N, double* A, long lda, const int* skip) {
  if (N <= NB) {
    if (N > 0) {
      hipLaunchKernelGGL(lauum_leaf_kernel, dim3(1), dim3(256), 0, s, (int)N, A, lda,
                         lower ? 1 : 0, skip);
      ATLG_LAUNCHED();
    }
    return;
  }
  const long nl = half_split(N), nr = N - nl;
  if (lower) {
    double* G = A + nl;               // L21 (nr x nl)
    double* U1 = A + nl + nl * lda;   // L22
    lauum_rec(s, true, nl, A, lda, skip);
    gemm<double>(s, kTrans, kNoTrans, nl, nl, nr, 1.0, G, lda, G, lda, 1.0, A, lda, 1, skip);
    trmm(s, kLeft, kLower, kTrans, kNonUnit, nr, nl, 1.0, U1, lda, G, lda, skip);
    lauum_rec(s, true, nr, U1, lda, skip);
  } else {
    double* G = A + nl * lda;         // U12 (nl x nr)
    double* U1 = A + nl + nl * lda;   // U22
    lauum_rec(s, false, nl, A, lda, skip);
    gemm<double>(s, kNoTrans, kTrans, nl, nl, nr, 1.0, G, lda, G, lda, 1.0, A, lda, 2, skip);
    trmm(s, kRight, kUpper, kTrans, kNonUnit, nl, nr, 1.0, U1, lda, G, lda, skip);
    lauum_rec(s, false, nr, U1, lda, skip);
  }
}

// ATL_trtriCL/CU's recursion (atlas-lp.base:2127-2180, 1903-1960) with the
// diagonal blocks' inverses precomputed in D (block b of the whole
// triangle at D + b*NB*NB) and applied by the TRSMs.
void trtri_rec(hipStream_t s, bool lower, int diag, long off, long N, double* A0, long lda,
               const double* D, const int* skip) {
  if (N <= NB) return;  // the leaf is D's block, copied in at the end
  const long nl = half_split(N), nr = N - nl;
  double* A = A0 + off + off * lda;
  double* A22 = A + nl + nl * lda;
  const double* D11 = D + (size_t)(off / NB) * NB * NB;
  const double* D22 = D + (size_t)((off + nl) / NB) * NB * NB;
  if (lower) {
    double* A21 = A + nl;
    trsm<double>(s, kRight, kLower, kNoTrans, diag, nr, nl, 1.0, A, lda, A21, lda, skip, D11, 0);
    trsm<double>(s, kLeft, kLower, kNoTrans, diag, nr, nl, -1.0, A22, lda, A21, lda, skip, D22,
                 0);
  } else {
    double* A12 = A + nl * lda;
    trsm<double>(s, kRight, kUpper, kNoTrans, diag, nl, nr, 1.0, A22, lda, A12, lda, skip, D22,
                 0);
    trsm<double>(s, kLeft, kUpper, kNoTrans, diag, nl, nr, -1.0, A, lda, A12, lda, skip, D11, 0);
  }
  trtri_rec(s, lower, diag, off, nl, A0, lda, D, skip);
  trtri_rec(s, lower, diag, off + nl, nr, A0, lda, D, skip);
}

}  // namespace

void trmm(hipStream_t s, int side, int uplo, int ta, int diag, long M, long N, double alpha,
          const double* A, long lda, double* B, long ldb, const int* skip) {
  if (M <= 0 || N <= 0) return;
  if (alpha == 0.0) {
    scale<double>(s, M, N, 0.0, B, ldb, 0, skip);
    return;
  }
  Tm c;
  c.s = s;
  c.left = side == kLeft;
  c.trans = is_trans(ta);
  c.lower_eff = (uplo == kLower) != c.trans;
  c.A = A;
  c.lda = lda;
  c.B = B;
  c.ldb = ldb;
  c.M = M;
  c.N = N;
  c.skip = skip;
  const long na = c.left ? M : N;
  const long nblk = (na + NB - 1) / NB;
  double* T = static_cast<double*>(scratch(s, kSlotTrmmBlk, sizeof(double) * nblk * NB * NB));
  hipLaunchKernelGGL(tri_blocks_kernel, dim3((unsigned)nblk), dim3(256), 0, s, (int)na, A, lda,
                     c.trans ? 1 : 0, c.lower_eff ? 1 : 0, diag == kUnit ? 1 : 0, T, skip);
  ATLG_LAUNCHED();
  c.T = T;
  tm_rec(c, 0, na, alpha);
}

void symm(hipStream_t s, int side, int uplo, long M, long N, double alpha, const double* A,
          long lda, const double* B, long ldb, double beta, double* C, long ldc) {
  if (M <= 0 || N <= 0) return;
  const long na = side == kLeft ? M : N;
  if (alpha == 0.0) {
    scale<double>(s, M, N, beta, C, ldc);
    return;
  }
  double* S = static_cast<double*>(scratch(s, kSlotSymm, sizeof(double) * na * na));
  const int nb = (int)std::min<long>((na * na + 255) / 256, 4096);
  hipLaunchKernelGGL(sym_fill_kernel, dim3(nb), dim3(256), 0, s, (int)na, A, lda,
                     uplo == kLower ? 1 : 0, S);
  ATLG_LAUNCHED();
  if (side == kLeft)
    gemm<double>(s, kNoTrans, kNoTrans, M, N, M, alpha, S, na, B, ldb, beta, C, ldc);
  else
    gemm<double>(s, kNoTrans, kNoTrans, M, N, N, alpha, B, ldb, S, na, beta, C, ldc);
}

void syr2k(hipStream_t s, int uplo, int trans, long N, long K, double alpha, const double* A,
           long lda, const double* B, long ldb, double beta, double* C, long ldc) {
  if (N <= 0) return;
  const int mode = uplo == kLower ? 1 : 2;
  if (K == 0 || alpha == 0.0) {
    if (beta != 1.0) scale<double>(s, N, N, beta, C, ldc, mode);
    return;
  }
  if (is_trans(trans)) {  // C = alpha A^T B + alpha B^T A + beta C
    gemm<double>(s, kTrans, kNoTrans, N, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode);
    gemm<double>(s, kTrans, kNoTrans, N, N, K, alpha, B, ldb, A, lda, 1.0, C, ldc, mode);
  } else {                // C = alpha A B^T + alpha B A^T + beta C
    gemm<double>(s, kNoTrans, kTrans, N, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode);
    gemm<double>(s, kNoTrans, kTrans, N, N, K, alpha, B, ldb, A, lda, 1.0, C, ldc, mode);
  }
}

void trtri(hipStream_t s, int uplo, int diag, long N, double* A, long lda, int* dinfo) {
  set_int(s, dinfo, 0);
  if (N <= 0) return;
  const bool lower = uplo == kLower;
  if (diag == kNonUnit) {
    hipLaunchKernelGGL(diag_zero_kernel, dim3(1), dim3(256), 0, s, (int)N, A, lda, dinfo);
    ATLG_LAUNCHED();
  }
  const long nblk = (N + NB - 1) / NB;
  double* D = static_cast<double*>(scratch(s, kSlotTrsmInv, sizeof(double) * nblk * NB * NB));
  tri_inverse_blocks<double>(s, N, A, lda, kNoTrans, uplo, diag, D, dinfo);
  trtri_rec(s, lower, diag, 0, N, A, lda, D, dinfo);
  hipLaunchKernelGGL(put_tri_blocks_kernel, dim3((unsigned)nblk), dim3(256), 0, s, (int)N, A, lda,
                     lower ? 1 : 0, diag == kUnit ? 1 : 0, D, dinfo);
  ATLG_LAUNCHED();
}

void lauum(hipStream_t s, int uplo, long N, double* A, long lda, const int* skip) {
  if (N <= 0) return;
  lauum_rec(s, uplo == kLower, N, A, lda, skip);
}

// clapack_dpotri (atlas-clp.base:316-362): trtri, then lauum only when the
// factor had no zero on its diagonal (the device info word skips it).
void potri(hipStream_t s, int uplo, long N, double* A, long lda, int* dinfo) {
  trtri(s, uplo, kNonUnit, N, A, lda, dinfo);
  lauum(s, uplo, N, A, lda, dinfo);
}

}  // namespace atlg
