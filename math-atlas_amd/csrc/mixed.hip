// Mixed-precision dgesv: LU in FP32 (or with BF16 trailing updates and FP32
// accumulation) on the matrix cores, FP64 residuals and corrections.
//
// There is no ATLAS counterpart; the driver follows LAPACK's dsgesv
// (Langou et al., "Exploiting the performance of 32 bit floating point
// arithmetic in obtaining 64 bit accuracy", SC'06):
//   ANRM = ||A||_inf, CTE = ANRM * eps * sqrt(n)  (eps = 2^-53)
//   A32 = float(A)  (ITER = -2 if an entry overflows FP32)
//   P L U = A32     (ITER = -3 if the FP32 factor is singular)
//   x = U^-1 L^-1 P float(b);  loop: r = b - A x (FP64);
//     stop when ||r_j||_inf < ||x_j||_inf * CTE for every column j (ITER =
//     corrections used), else x += U^-1 L^-1 P float(r), at most ITERMAX
//     corrections (ITER = -(ITERMAX+1));
//   on any negative ITER the FP64 dgetrf/dgetrs of this library solve the
//   system instead, A then holding the FP64 factors.
// The ATLAS routines on the path are the ones it calls: getrf/getrs
// (ATL_dgetrf, ATL_dgetrs; AtlasBase/Clint/atlas-lp.base:1497-1720) in FP32
// and FP64.
//
// The FP64 passes over A (norm + conversion, residual) are HBM-bound: each
// reads A once.  They split the columns into chunks so that enough
// workgroups stream A, write per-chunk partial row sums, and reduce the
// partials in a fixed order (deterministic results).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "atlg_internal.hpp"

namespace atlg {
namespace {

constexpr int kRows = 256;   // rows per workgroup (one per thread)
constexpr int kChunk = 512;  // columns per chunk

int chunks_of(long N) { return (int)((N + kChunk - 1) / kChunk); }

// partial[c][i] = sum_{j in chunk c} |A(i,j)|; A32 = float(A); *over = 1 if
// an entry exceeds FLT_MAX in magnitude.
__global__ void __launch_bounds__(kRows)
norm_convert_kernel(long N, const double* __restrict__ A, long lda, float* __restrict__ A32,
                    long lda32, double* __restrict__ partial, int* __restrict__ over) {
  const long i = (long)blockIdx.x * kRows + threadIdx.x;
  const int c = blockIdx.y;
  const long j0 = (long)c * kChunk, j1 = std::min(N, j0 + kChunk);
  if (i >= N) return;
  double s = 0.0;
  bool big = false;
  for (long j = j0; j < j1; j++) {
    const double a = A[i + j * lda];
    s += fabs(a);
    big |= fabs(a) > (double)FLT_MAX;
    A32[i + j * lda32] = (float)a;
  }
  partial[(size_t)c * N + i] = s;
  if (big) atomicOr(over, 1);
}

// partial[c][i, k] = sum_{j in chunk c} A(i,j) X(j,k) for k < nrhs (<= 4)
__global__ void __launch_bounds__(kRows)
gemv_partial_kernel(long N, int nrhs, const double* __restrict__ A, long lda,
                    const double* __restrict__ X, long ldx, double* __restrict__ partial) {
  __shared__ double xs[4][kChunk];
  const long i = (long)blockIdx.x * kRows + threadIdx.x;
  const int c = blockIdx.y;
  const long j0 = (long)c * kChunk;
  const int w = (int)std::min<long>(kChunk, N - j0);
  for (int e = threadIdx.x; e < w * nrhs; e += kRows) {
    const int k = e / w, j = e - k * w;
    xs[k][j] = X[j0 + j + (size_t)k * ldx];
  }
  __syncthreads();
  if (i >= N) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double* a = A + i + j0 * lda;
  for (int j = 0; j < w; j++) {
    const double v = a[(size_t)j * lda];
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (k < nrhs) acc[k] += v * xs[k][j];
  }
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (k < nrhs) partial[((size_t)c * 4 + k) * N + i] = acc[k];
}

// R = B - sum_c partial[c];  W32[pos[i]] = float(R[i]) (the row
// interchanges of the factor applied on the way);  per column k: atomic max
// of |r| and |x| into nrm[2k], nrm[2k+1] (non-negative doubles compare as
// their bit patterns).
__global__ void __launch_bounds__(kRows)
residual_finish_kernel(long N, int nrhs, int nchunk, const double* __restrict__ partial,
                       const double* __restrict__ B, long ldb, const double* __restrict__ X,
                       long ldx, float* __restrict__ W32, long ldw, const int* __restrict__ pos,
                       unsigned long long* __restrict__ nrm) {
  const long i = (long)blockIdx.x * kRows + threadIdx.x;
  for (int k = 0; k < nrhs; k++) {
    double r = 0.0, x = 0.0;
    if (i < N) {
      double s = 0.0;
      for (int c = 0; c < nchunk; c++) s += partial[((size_t)c * 4 + k) * N + i];
      r = B[i + (size_t)k * ldb] - s;
      x = X[i + (size_t)k * ldx];
      W32[pos[i] + (size_t)k * ldw] = (float)r;
    }
    double rm = fabs(r), xm = fabs(x);
    for (int o = 32; o > 0; o >>= 1) {
      rm = fmax(rm, __shfl_xor(rm, o, 64));
      xm = fmax(xm, __shfl_xor(xm, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
      atomicMax(&nrm[2 * k], (unsigned long long)__double_as_longlong(rm));
      atomicMax(&nrm[2 * k + 1], (unsigned long long)__double_as_longlong(xm));
    }
  }
}

// rowsum[i] = sum_c partial[c][i]; *anrm = max_i rowsum[i] (bit-pattern max)
__global__ void __launch_bounds__(kRows)
rowsum_max_kernel(long N, int nchunk, const double* __restrict__ partial,
                  unsigned long long* __restrict__ anrm) {
  const long i = (long)blockIdx.x * kRows + threadIdx.x;
  double s = 0.0;
  if (i < N)
    for (int c = 0; c < nchunk; c++) s += partial[(size_t)c * N + i];
  for (int o = 32; o > 0; o >>= 1) s = fmax(s, __shfl_xor(s, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(anrm, (unsigned long long)__double_as_longlong(s));
}

// W[pos[i]] = float(B[i]) per column
__global__ void d2s_kernel(long N, int nrhs, const double* __restrict__ B, long ldb,
                           float* __restrict__ W, long ldw, const int* __restrict__ pos) {
  const long total = N * nrhs;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long i = e % N, k = e / N;
    W[pos[i] + k * ldw] = (float)B[i + k * ldb];
  }
}

// X (=|+=) double(W)
__global__ void s2d_kernel(long N, int nrhs, const float* __restrict__ W, long ldw,
                           double* __restrict__ X, long ldx, int accumulate) {
  const long total = N * nrhs;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long i = e % N, k = e / N;
    const double w = (double)W[i + k * ldw];
    X[i + k * ldx] = accumulate ? X[i + k * ldx] + w : w;
  }
}

__global__ void zero_u64(unsigned long long* p, int n) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    p[e] = 0ull;
}

unsigned grid1(long total) { return (unsigned)std::min<long>((total + 255) / 256, 16384); }

double as_double(unsigned long long u) {
  double d;
  std::memcpy(&d, &u, sizeof d);
  return d;
}

}  // namespace

int dsgesv(hipStream_t s, long N, long NRHS, double* A, long lda, int* ipiv, const double* B,
           long ldb, double* X, long ldx, int* dinfo, int itermax, int lowprec, int* iter_out) {
  set_int(s, dinfo, 0);
  if (N == 0 || NRHS == 0) {
    *iter_out = 0;
    return 0;
  }
  const int nchunk = chunks_of(N), nr = (int)NRHS;
  const long rb = (N + kRows - 1) / kRows;
  const long ldw = N;
  // scratch: A32 (N x N) | W32 (N x NRHS) | partials | small words:
  //   [2k] max|r_k|, [2k+1] max|x_k| (k < NRHS), [2 NRHS] ||A||_inf,
  //   [2 NRHS + 1] = {overflow flag, FP32 info} as two ints
  const size_t bA32 = ((size_t)N * N * sizeof(float) + 255) & ~size_t(255);
  const size_t bW = ((size_t)N * NRHS * sizeof(float) + 255) & ~size_t(255);
  const size_t bP = ((size_t)nchunk * 4 * N * sizeof(double) + 255) & ~size_t(255);
  const int nsmall = 2 * nr + 2;
  const size_t bS = (sizeof(unsigned long long) * nsmall + 255) & ~size_t(255);
  const size_t bPos = (sizeof(int) * N + 255) & ~size_t(255);
  const long nblk = (N + kTrsmNB - 1) / kTrsmNB;
  const size_t bInv = sizeof(float) * nblk * kTrsmNB * kTrsmNB;
  char* base = static_cast<char*>(
      scratch(s, kSlotMixed, bA32 + bW + bP + bS + bPos + 2 * bInv));
  float* A32 = reinterpret_cast<float*>(base);
  float* W32 = reinterpret_cast<float*>(base + bA32);
  double* part = reinterpret_cast<double*>(base + bA32 + bW);
  unsigned long long* small = reinterpret_cast<unsigned long long*>(base + bA32 + bW + bP);
  int* dpos = reinterpret_cast<int*>(base + bA32 + bW + bP + bS);
  float* invL = reinterpret_cast<float*>(base + bA32 + bW + bP + bS + bPos);
  float* invU = invL + (size_t)nblk * kTrsmNB * kTrsmNB;
  int* flags = reinterpret_cast<int*>(small + 2 * nr + 1);  // [0] overflow, [1] FP32 info
  std::vector<unsigned long long> hs(nsmall);
  const unsigned gs = (unsigned)std::min(64, (nsmall + 255) / 256);

  // ||A||_inf, A32 = float(A), FP32 factor
  hipLaunchKernelGGL(zero_u64, dim3(gs), dim3(256), 0, s, small, nsmall);
  ATLG_LAUNCHED();
  hipLaunchKernelGGL(norm_convert_kernel, dim3((unsigned)rb, (unsigned)nchunk), dim3(kRows), 0, s,
                     N, A, lda, A32, N, part, flags);
  ATLG_LAUNCHED();
  hipLaunchKernelGGL(rowsum_max_kernel, dim3((unsigned)rb), dim3(kRows), 0, s, N, nchunk, part,
                     small + 2 * nr);
  ATLG_LAUNCHED();
  getrf<float>(s, N, N, A32, N, ipiv, flags + 1,
               lowprec == 1 ? kGemmBf16 : (lowprec == 2 ? kGemmBf16x6 : kGemmNative));
  std::vector<int> hpiv(N);
  ATLG_HIP(hipMemcpyAsync(hs.data(), small, sizeof(unsigned long long) * nsmall,
                          hipMemcpyDeviceToHost, s));
  ATLG_HIP(hipMemcpyAsync(hpiv.data(), ipiv, sizeof(int) * N, hipMemcpyDeviceToHost, s));
  ATLG_HIP(hipStreamSynchronize(s));
  const double anrm = as_double(hs[2 * nr]);
  int hflags[2];
  std::memcpy(hflags, &hs[2 * nr + 1], sizeof hflags);
  const double cte = anrm * (DBL_EPSILON * 0.5) * std::sqrt((double)N);

  int iter;
  bool done = false;
  if (hflags[0]) {
    iter = -2;
  } else if (hflags[1] != 0) {
    iter = -3;
  } else {
    // The factor's row interchanges composed once: row i of P b is row
    // perm[i] of b, so b's row i lands at pos[i] = perm^-1[i].  They are
    // applied inside the conversions below, and the triangles' diagonal-
    // block inverses are computed once for all the solves.
    {
      std::vector<int> perm(N), pos(N);
      for (long i = 0; i < N; i++) perm[i] = (int)i;
      for (long i = 0; i < N; i++) std::swap(perm[i], perm[hpiv[i]]);
      for (long i = 0; i < N; i++) pos[perm[i]] = (int)i;
      ATLG_HIP(hipMemcpyAsync(dpos, pos.data(), sizeof(int) * N, hipMemcpyHostToDevice, s));
      ATLG_HIP(hipStreamSynchronize(s));  // pos lives on this stack frame
    }
    tri_inverse_blocks<float>(s, N, A32, N, kNoTrans, kLower, kUnit, invL);
    tri_inverse_blocks<float>(s, N, A32, N, kNoTrans, kUpper, kNonUnit, invU);
    auto solve32 = [&]() {
      trsm<float>(s, kLeft, kLower, kNoTrans, kUnit, N, NRHS, 1.0f, A32, N, W32, ldw, nullptr,
                  invL, 0);
      trsm<float>(s, kLeft, kUpper, kNoTrans, kNonUnit, N, NRHS, 1.0f, A32, N, W32, ldw,
                  nullptr, invU, 0);
    };
    // x = solve32(b)
    hipLaunchKernelGGL(d2s_kernel, dim3(grid1(N * NRHS)), dim3(256), 0, s, N, nr, B, ldb, W32,
                       ldw, dpos);
    ATLG_LAUNCHED();
    solve32();
    hipLaunchKernelGGL(s2d_kernel, dim3(grid1(N * NRHS)), dim3(256), 0, s, N, nr, W32, ldw, X,
                       ldx, 0);
    ATLG_LAUNCHED();
    iter = -(itermax + 1);
    for (int k = 0; k <= itermax; k++) {
      // r = b - A x (FP64), W32 = float(r), column norms of r and x
      hipLaunchKernelGGL(zero_u64, dim3(gs), dim3(256), 0, s, small, 2 * nr);
      ATLG_LAUNCHED();
      for (int k0 = 0; k0 < nr; k0 += 4) {
        const int kk = std::min(4, nr - k0);
        hipLaunchKernelGGL(gemv_partial_kernel, dim3((unsigned)rb, (unsigned)nchunk), dim3(kRows),
                           0, s, N, kk, A, lda, X + (size_t)k0 * ldx, ldx, part);
        ATLG_LAUNCHED();
        hipLaunchKernelGGL(residual_finish_kernel, dim3((unsigned)rb), dim3(kRows), 0, s, N, kk,
                           nchunk, part, B + (size_t)k0 * ldb, ldb, X + (size_t)k0 * ldx, ldx,
                           W32 + (size_t)k0 * ldw, ldw, dpos, small + 2 * k0);
        ATLG_LAUNCHED();
      }
      ATLG_HIP(hipMemcpyAsync(hs.data(), small, sizeof(unsigned long long) * 2 * nr,
                              hipMemcpyDeviceToHost, s));
      ATLG_HIP(hipStreamSynchronize(s));
      bool ok = true;
      for (int q = 0; q < nr; q++)
        if (!(as_double(hs[2 * q]) < as_double(hs[2 * q + 1]) * cte)) ok = false;
      if (ok) {
        iter = k;
        done = true;
        break;
      }
      if (k == itermax) break;
      // x += solve32(r)
      solve32();
      hipLaunchKernelGGL(s2d_kernel, dim3(grid1(N * NRHS)), dim3(256), 0, s, N, nr, W32, ldw, X,
                         ldx, 1);
      ATLG_LAUNCHED();
    }
  }
  if (!done) {
    // FP64 fallback: A <- its LU factors, X <- A^-1 B
    ATLG_HIP(hipMemcpy2DAsync(X, sizeof(double) * ldx, B, sizeof(double) * ldb,
                              sizeof(double) * N, NRHS, hipMemcpyDeviceToDevice, s));
    getrf<double>(s, N, N, A, lda, ipiv, dinfo);
    getrs<double>(s, kNoTrans, N, NRHS, A, lda, ipiv, X, ldx);
  }
  *iter_out = iter;
  return 0;
}

}  // namespace atlg
