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

// ---- dense-scratch forms of TRTRI / LAUUM / POTRI --------------------------
// The triangle is copied into a dense scratch matrix X (order np = n rounded
// up to 32, leading dimension np) as a lower triangle with zeros everywhere
// else (an upper triangle is transposed in), so that the products below can
// take whole tiles: TRTRI by doubling, LAUUM as one triangle-masked product.
constexpr long kTriFullMin = 2 * NB;                  // smallest order that uses them
constexpr size_t kTriFullCap = size_t(4) << 30;       // largest dense copy (bytes)
constexpr int kCT = 32;                               // copy tile

// X = the lower triangle op(A) (transposed in when !lower; 1 on a unit
// diagonal), zero elsewhere, in 32x32 tiles through LDS so that both the
// reads and the writes are coalesced.
__global__ void __launch_bounds__(256)
tri_copy_in_kernel(int n, int np, const double* __restrict__ A, long lda, int lower, int unit,
                   double* __restrict__ X, const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double tl[kCT][kCT + 1];
  const int tiles = np / kCT;
  const int i0 = (blockIdx.x % tiles) * kCT, j0 = (blockIdx.x / tiles) * kCT;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < kCT; r += 8) {
    // lower: element (i0+tx, j0+r) = A(i0+tx, j0+r); upper: element
    // (i0+r, j0+tx) = A(j0+tx, i0+r), read along A's column
    double v;
    if (lower) {
      const int i = i0 + tx, j = j0 + r;
      v = (i < n && j <= i) ? A[i + (size_t)j * lda] : 0.0;
      tl[tx][r] = v;
    } else {
      const int i = i0 + r, j = j0 + tx;
      v = (i < n && j <= i) ? A[j + (size_t)i * lda] : 0.0;
      tl[r][tx] = v;
    }
  }
  __syncthreads();
  for (int r = ty; r < kCT; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    double v = tl[tx][r];
    if (unit && i == j && i < n) v = 1.0;
    X[i + (size_t)j * np] = v;
  }
}

// The lower triangle of X back into A's triangle (transposed when !lower);
// a unit diagonal is not written.
__global__ void __launch_bounds__(256)
tri_copy_out_kernel(int n, int np, const double* __restrict__ X, int lower, int unit,
                    double* __restrict__ A, long lda, const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double tl[kCT][kCT + 1];
  const int tiles = np / kCT;
  const int i0 = (blockIdx.x % tiles) * kCT, j0 = (blockIdx.x / tiles) * kCT;
  if (j0 > i0 + kCT - 1 || i0 >= n || j0 >= n) return;  // no lower element in the tile
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < kCT; r += 8) tl[tx][r] = X[(i0 + tx) + (size_t)(j0 + r) * np];
  __syncthreads();
  for (int r = ty; r < kCT; r += 8) {
    if (lower) {
      const int i = i0 + tx, j = j0 + r;
      if (i < n && j < i + (unit ? 0 : 1)) A[i + (size_t)j * lda] = tl[tx][r];
    } else {
      const int i = i0 + r, j = j0 + tx;  // A(j, i) = X(i, j)
      if (i < n && j < i + (unit ? 0 : 1)) A[j + (size_t)i * lda] = tl[r][tx];
    }
  }
}

double* tri_copy_in(hipStream_t s, bool lower, bool unit, long n, const double* A, long lda,
                    const int* skip) {
  const long np = tri_full_order(n);
  double* X = static_cast<double*>(scratch(s, kSlotTriFull, sizeof(double) * np * np));
  tri_copy_to(s, lower, unit, n, A, lda, X, skip);
  return X;
}

// A's triangle <- X^T X (lower: L^T L; upper, X = U^T: U U^T), one product
// masked to the triangle, each tile's K range starting at its row and column.
void lauum_from_full(hipStream_t s, bool lower, long n, const double* X, long np, double* A,
                     long lda, const int* skip) {
  if (!gemm_batched(s, kTrans, kNoTrans, n, n, np, 1.0, X, np, 0, X, np, 0, 0.0, A, lda, 0, 1,
                    lower ? 1 : 2, 3, skip))
    throw HipFailure{"lauum: the masked product does not apply"};
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

long tri_full_order(long n) { return (n + kCT - 1) / kCT * kCT; }

static bool tri_full_fits(long n) {
  const long np = tri_full_order(n);
  return n >= kTriFullMin && (size_t)np * np * sizeof(double) <= kTriFullCap;
}

void tri_copy_to(hipStream_t s, bool lower, bool unit, long n, const double* A, long lda,
                 double* X, const int* skip) {
  const long np = tri_full_order(n);
  const long tiles = np / kCT;
  hipLaunchKernelGGL(tri_copy_in_kernel, dim3((unsigned)(tiles * tiles)), dim3(256), 0, s, (int)n,
                     (int)np, A, lda, lower ? 1 : 0, unit ? 1 : 0, X, skip);
  ATLG_LAUNCHED();
}

// X (order n, leading dimension np, lower, zeros elsewhere) <- inv(X):
// the NB-blocks' inverses, then doubling: for blocks of order sz = 2NB,
// 4NB, ... with halves already inverted, X21 <- -X22 (X21 X11).  The full
// blocks of a level go in one batched launch per product, each product
// restricted to the K range its triangular operand can reach; a ragged
// last block takes two ordinary products (its zero padding keeps them exact).
void tri_invert_full(hipStream_t s, int diag, long n, double* X, long np, const int* skip) {
  const long nblk = (n + NB - 1) / NB;
  double* D = static_cast<double*>(scratch(s, kSlotTrsmInv, sizeof(double) * nblk * NB * NB));
  tri_inverse_blocks<double>(s, n, X, np, kNoTrans, kLower, diag, D, skip);
  hipLaunchKernelGGL(put_tri_blocks_kernel, dim3((unsigned)nblk), dim3(256), 0, s, (int)n, X, np,
                     1, 0, D, skip);
  ATLG_LAUNCHED();
  long tneed = 0;
  for (long sz = 2 * NB; sz / 2 < n; sz *= 2) tneed = std::max(tneed, std::max(n / sz, 1L) * (sz / 2) * (sz / 2));
  if (tneed == 0) return;
  double* T = static_cast<double*>(scratch(s, kSlotTriTemp, sizeof(double) * tneed));
  for (long sz = 2 * NB; sz / 2 < n; sz *= 2) {
    const long h = sz / 2, nfull = n / sz, st = sz * (np + 1);
    if (nfull > 0) {
      // T_b = X21_b X11_b (X11 lower: k >= tile column), then
      // X21_b = -X22_b T_b (X22 lower: k < tile row end)
      if (!gemm_batched(s, kNoTrans, kNoTrans, h, h, h, 1.0, X + h, np, st, X, np, st, 0.0, T, h,
                        h * h, (int)nfull, 0, 1, skip) ||
          !gemm_batched(s, kNoTrans, kNoTrans, h, h, h, -1.0, X + h + h * np, np, st, T, h, h * h,
                        0.0, X + h, np, st, (int)nfull, 0, 4, skip))
        throw HipFailure{"trtri: the batched product does not apply"};
    }
    const long o = nfull * sz, r = n - o - h;
    if (r > 0) {
      gemm<double>(s, kNoTrans, kNoTrans, r, h, h, 1.0, X + o + h + o * np, np, X + o + o * np, np,
                   0.0, T, h, 0, skip);
      gemm<double>(s, kNoTrans, kNoTrans, r, h, r, -1.0, X + (o + h) + (o + h) * np, np, T, h, 0.0,
                   X + o + h + o * np, np, 0, skip);
    }
  }
}

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
  if (tri_full_fits(N)) {
    const long np = tri_full_order(N);
    double* X = tri_copy_in(s, lower, diag == kUnit, N, A, lda, dinfo);
    tri_invert_full(s, diag, N, X, np, dinfo);
    const long tiles = np / kCT;
    hipLaunchKernelGGL(tri_copy_out_kernel, dim3((unsigned)(tiles * tiles)), dim3(256), 0, s,
                       (int)N, (int)np, X, lower ? 1 : 0, diag == kUnit ? 1 : 0, A, lda, dinfo);
    ATLG_LAUNCHED();
    return;
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
  const bool lower = uplo == kLower;
  if (tri_full_fits(N)) {
    const double* X = tri_copy_in(s, lower, false, N, A, lda, skip);
    lauum_from_full(s, lower, N, X, tri_full_order(N), A, lda, skip);
    return;
  }
  lauum_rec(s, lower, N, A, lda, skip);
}

// clapack_dpotri (atlas-clp.base:316-362): trtri, then lauum only when the
// factor had no zero on its diagonal (the device info word skips it).
void potri(hipStream_t s, int uplo, long N, double* A, long lda, int* dinfo) {
  if (N > 0 && tri_full_fits(N)) {
    // the inverse stays in the dense copy, from which the product reads
    const bool lower = uplo == kLower;
    set_int(s, dinfo, 0);
    hipLaunchKernelGGL(diag_zero_kernel, dim3(1), dim3(256), 0, s, (int)N, A, lda, dinfo);
    ATLG_LAUNCHED();
    const long np = tri_full_order(N);
    double* X = tri_copy_in(s, lower, false, N, A, lda, dinfo);
    tri_invert_full(s, kNonUnit, N, X, np, dinfo);
    lauum_from_full(s, lower, N, X, np, A, lda, dinfo);
    return;
  }
  trtri(s, uplo, kNonUnit, N, A, lda, dinfo);
  lauum(s, uplo, N, A, lda, dinfo);
}

}  // namespace atlg
