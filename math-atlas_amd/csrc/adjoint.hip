// Reverse-mode (adjoint) kernels for the Cholesky path: potrf, trsm with a
// Cholesky factor, and logdet.  They have no ATLAS counterpart; they extend
// the level-3 routines so a Gaussian log-likelihood
//   l(K, y) = -1/2 y^T K^-1 y - 1/2 log det K
// can be differentiated on the device.  With K = L L^T (lower) and
// a = L^-1 y, the pieces are
//   logdet:  d = 2 sum log L_ii            -> Lbar_ii += 2 dbar / L_ii
//   trsm:    a = L^-1 y                     -> ybar = L^-T abar,
//                                              Lbar -= tril(ybar a^T)
//   potrf:   L = chol(K)                    -> Kbar = sym(L^-T Phi(L^T Lbar) L^-1)
// where Phi takes the lower triangle with the diagonal halved and
// sym(X) = (X + X^T)/2 (Murray, "Differentiation of the Cholesky
// decomposition", 2016, its blocked matrix form).  Every product runs on the
// MFMA kernels of this library (5 n^3 / 3 flops for the potrf adjoint up to
// the orders whose dense copies fit, 3 n^3 beyond).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "atlg_internal.hpp"

namespace atlg {
namespace {

// dst = tril(src) (zero above the diagonal)
__global__ void tril_copy_kernel(long N, const double* src, long lds,
                                 double* dst, long ldd) {  // may run in place
  const long total = N * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long i = e % N, j = e / N;
    dst[i + j * ldd] = (i >= j) ? src[i + j * lds] : 0.0;
  }
}

// X <- Phi(X): zero above the diagonal, halve the diagonal
__global__ void phi_kernel(long N, double* __restrict__ X, long ldx) {
  const long total = N * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long i = e % N, j = e / N;
    double* x = X + i + j * ldx;
    if (i < j) *x = 0.0;
    else if (i == j) *x *= 0.5;
  }
}

// X (order np, the first n rows and columns in use) <- Phi(X) with every
// element outside the leading lower triangle zeroed, so that X can be a
// triangular operand of the K-trimmed products
__global__ void phi_padded_kernel(long n, long np, double* __restrict__ X) {
  const long total = np * np;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long i = e % np, j = e / np;
    double* x = X + e;
    if (i < j || i >= n) *x = 0.0;
    else if (i == j) *x *= 0.5;
  }
}

// dst = (X + X^T) / 2, full square
__global__ void sym_kernel(long N, const double* __restrict__ X, long ldx,
                           double* __restrict__ dst, long ldd) {
  const long total = N * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long i = e % N, j = e / N;
    dst[i + j * ldd] = 0.5 * (X[i + j * ldx] + X[j + i * ldx]);
  }
}

// out = 2 sum log L_ii (one workgroup, fixed summation order)
__global__ void __launch_bounds__(1024)
logdet_kernel(long N, const double* __restrict__ L, long ldl, double* __restrict__ out) {
  __shared__ double part[1024];
  double s = 0.0;
  for (long i = threadIdx.x; i < N; i += blockDim.x) s += log(L[i + i * ldl]);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = 2.0 * part[0];
}

// Lbar_ii += 2 * dbar / L_ii
__global__ void logdet_adjoint_kernel(long N, const double* __restrict__ L, long ldl,
                                      const double* __restrict__ dbar, double* __restrict__ Lbar,
                                      long ldlb) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < N) Lbar[i + i * ldlb] += 2.0 * (*dbar) / L[i + i * ldl];
}

unsigned grid_for(long total) {
  return (unsigned)std::min<long>((total + 255) / 256, 8192);
}

}  // namespace

// Kbar = sym(L^-T Phi(L^T Lbar) L^-1).
//
// Up to the size where three dense n x n copies fit (the common case): L
// and tril(Lbar) are copied into zero-padded dense matrices, then on the
// register GEMM with each 128x128 tile's K range trimmed to where its
// triangular operands can be nonzero:
//   P = L^T tril(Lbar)   (lower part only: n^3/3 flops), P <- Phi(P);
//   X = L^-1             (doubling inverse in place of the copy of L, n^3/3);
//   Q = P X              (lower times lower, lower: n^3/3);
//   R = X^T Q            (full: 2n^3/3), Kbar = sym(R).
// 5n^3/3 flops against the 3n^3 of the TRMM + two-solve form below, which
// serves larger orders.  Only L's lower triangle is read.
void potrf_adjoint(hipStream_t s, long N, const double* L, long ldl, double* G, long ldg) {
  if (N <= 0) return;
  const long np = tri_full_order(N);
  const unsigned gp = grid_for(np * np);
  if (N >= 2 * kTrsmNB && 3 * (size_t)np * np * sizeof(double) <= (size_t(6) << 30)) {
    const size_t bytes = sizeof(double) * np * np;
    double* X = static_cast<double*>(scratch(s, kSlotTriFull, bytes));
    double* M = static_cast<double*>(scratch(s, kSlotAdjoint, bytes));
    double* P = static_cast<double*>(scratch(s, kSlotAdjP, bytes));
    tri_copy_to(s, true, false, N, L, ldl, X, nullptr);
    tri_copy_to(s, true, false, N, G, ldg, M, nullptr);
    // the register path applies: K = np is a multiple of 32, the copies are
    // aligned with an even leading dimension, and every trimmed range is
    // nonempty (N <= np)
    if (!gemm_batched(s, kTrans, kNoTrans, N, N, np, 1.0, X, np, 0, M, np, 0, 0.0, P, np, 0, 1,
                      1, 3, nullptr))
      throw HipFailure{"potrf adjoint: the trimmed product does not apply"};
    hipLaunchKernelGGL(phi_padded_kernel, dim3(gp), dim3(256), 0, s, N, np, P);
    ATLG_LAUNCHED();
    tri_invert_full(s, kNonUnit, N, X, np, nullptr);
    if (!gemm_batched(s, kNoTrans, kNoTrans, N, N, np, 1.0, P, np, 0, X, np, 0, 0.0, M, np, 0, 1,
                      1, 4 | 1, nullptr) ||
        !gemm_batched(s, kTrans, kNoTrans, N, N, np, 1.0, X, np, 0, M, np, 0, 0.0, P, np, 0, 1,
                      0, 3, nullptr))
      throw HipFailure{"potrf adjoint: the trimmed product does not apply"};
    const unsigned g = grid_for(N * N);
    hipLaunchKernelGGL(sym_kernel, dim3(g), dim3(256), 0, s, N, P, np, G, ldg);
    ATLG_LAUNCHED();
    return;
  }
  // TRMM + two solves: P = L^T tril(Lbar) by TRMM, Phi(P), then the two
  // triangular solves with L's diagonal-block inverses computed once, 3 n^3
  // flops in all.
  constexpr long NB = kTrsmNB;
  const long nblk = (N + NB - 1) / NB;
  double* W = static_cast<double*>(
      scratch(s, kSlotAdjoint, sizeof(double) * ((size_t)N * N + (size_t)nblk * NB * NB)));
  double* P = W;                       // N x N, ld N
  double* D = W + (size_t)N * N;       // inv(L_bb)
  const unsigned g = grid_for(N * N);
  hipLaunchKernelGGL(tril_copy_kernel, dim3(g), dim3(256), 0, s, N, G, ldg, P, N);
  ATLG_LAUNCHED();
  trmm(s, kLeft, kLower, kTrans, kNonUnit, N, N, 1.0, L, ldl, P, N);
  hipLaunchKernelGGL(phi_kernel, dim3(g), dim3(256), 0, s, N, P, N);
  ATLG_LAUNCHED();
  tri_inverse_blocks<double>(s, N, L, ldl, kNoTrans, kLower, kNonUnit, D);
  trsm(s, kLeft, kLower, kTrans, kNonUnit, N, N, 1.0, L, ldl, P, N, nullptr, D, 1);
  trsm(s, kRight, kLower, kNoTrans, kNonUnit, N, N, 1.0, L, ldl, P, N, nullptr, D, 0);
  hipLaunchKernelGGL(sym_kernel, dim3(g), dim3(256), 0, s, N, P, N, G, ldg);
  ATLG_LAUNCHED();
}

void trsm_adjoint(hipStream_t s, long N, long NRHS, const double* L, long ldl, const double* X,
                  long ldx, double* Xbar, long ldxb, double* Lbar, long ldlb) {
  if (N <= 0 || NRHS <= 0) return;
  trsm(s, kLeft, kLower, kTrans, kNonUnit, N, NRHS, 1.0, L, ldl, Xbar, ldxb);
  if (Lbar) gemm(s, kNoTrans, kTrans, N, N, NRHS, -1.0, Xbar, ldxb, X, ldx, 1.0, Lbar, ldlb, 1);
}

void logdet_chol(hipStream_t s, long N, const double* L, long ldl, double* out) {
  hipLaunchKernelGGL(logdet_kernel, dim3(1), dim3(1024), 0, s, N, L, ldl, out);
  ATLG_LAUNCHED();
}

void logdet_chol_adjoint(hipStream_t s, long N, const double* L, long ldl, const double* dbar,
                         double* Lbar, long ldlb) {
  if (N <= 0) return;
  hipLaunchKernelGGL(logdet_adjoint_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s,
                     N, L, ldl, dbar, Lbar, ldlb);
  ATLG_LAUNCHED();
}

}  // namespace atlg
