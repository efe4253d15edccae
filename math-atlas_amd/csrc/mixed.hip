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
                       unsigned long long* __restrict__ nrm, double* __restrict__ R64) {
  const long i = (long)blockIdx.x * kRows + threadIdx.x;
  for (int k = 0; k < nrhs; k++) {
    double r = 0.0, x = 0.0;
    if (i < N) {
      double s = 0.0;
      for (int c = 0; c < nchunk; c++) s += partial[((size_t)c * 4 + k) * N + i];
      r = B[i + (size_t)k * ldb] - s;
      x = X[i + (size_t)k * ldx];
      W32[pos[i] + (size_t)k * ldw] = (float)r;
      if (R64) R64[i + (size_t)k * N] = r;
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

// ---- GMRES-IR building blocks (one right-hand side) ----------------------

// W32[pos[i]] = float(sum_c partial[c][i])  (A v, permuted for the solves)
__global__ void __launch_bounds__(kRows)
matvec_finish_kernel(long N, int nchunk, const double* __restrict__ partial,
                     float* __restrict__ W32, const int* __restrict__ pos) {
  const long i = (long)blockIdx.x * kRows + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; c++) s += partial[(size_t)c * 4 * N + i];
  W32[pos[i]] = (float)s;
}

// part[chunk][j] = sum over this chunk's rows of V[r][j] * w[r], j < k;
// V has leading dimension ldv (k <= kMaxBasis + 1)
constexpr int kMaxBasis = 30;
constexpr int kInnerBudget = 10 * kMaxBasis;  // preconditioned products per solve, at most
__global__ void __launch_bounds__(kRows)
dots_kernel(long N, int k, const double* __restrict__ V, long ldv, const double* __restrict__ w,
            double* __restrict__ part) {
  __shared__ double red[kRows / 64][kMaxBasis + 2];
  const long r = (long)blockIdx.x * kRows + threadIdx.x;
  const double wr = (r < N) ? w[r] : 0.0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = 0; j < k; j++) {
    double v = (r < N) ? V[r + (size_t)j * ldv] * wr : 0.0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wv][j] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < k) {
    double s = 0.0;
    for (int q = 0; q < kRows / 64; q++) s += red[q][threadIdx.x];
    part[(size_t)blockIdx.x * (kMaxBasis + 2) + threadIdx.x] = s;
  }
}

// out[j] = sum over chunks of part[chunk][j] (fixed order), j < k
__global__ void sum_chunks_kernel(int nb, int k, const double* __restrict__ part,
                                  double* __restrict__ out) {
  const int j = threadIdx.x;
  if (j >= k) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += part[(size_t)b * (kMaxBasis + 2) + j];
  out[j] = s;
}

// w[r] += sign * sum_{j < k} V[r][j] * h[j]
__global__ void __launch_bounds__(kRows)
axpy_basis_kernel(long N, int k, const double* __restrict__ V, long ldv,
                  const double* __restrict__ h, double sign, double* __restrict__ w) {
  const long r = (long)blockIdx.x * kRows + threadIdx.x;
  if (r >= N) return;
  double s = 0.0;
  for (int j = 0; j < k; j++) s += V[r + (size_t)j * ldv] * h[j];
  w[r] += sign * s;
}

// dst[r] = src[r] * a
__global__ void scale_copy_kernel(long N, const double* __restrict__ src, double a,
                                  double* __restrict__ dst) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < N;
       r += (long)gridDim.x * blockDim.x)
    dst[r] = src[r] * a;
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

// Work vectors of one GMRES solve.
struct Gmres {
  double* V;     // N x (kMaxBasis + 1) Krylov basis (ld N)
  double* w;     // N
  double* part;  // per-chunk partial dot products
  double* hA;    // kMaxBasis + 2: first Gram-Schmidt pass
  double* hB;    //                 second pass
  double* nrm;   //                 |w|^2
  double* y;     //                 solution of the small least-squares problem
};

// out[0..k) = V[:, 0..k)^T w
void dev_dots(hipStream_t s, long N, int k, const double* V, long ldv, const double* w,
              double* part, double* out) {
  const unsigned nb = (unsigned)((N + kRows - 1) / kRows);
  hipLaunchKernelGGL(dots_kernel, dim3(nb), dim3(kRows), 0, s, N, k, V, ldv, w, part);
  ATLG_LAUNCHED();
  hipLaunchKernelGGL(sum_chunks_kernel, dim3(1), dim3(64), 0, s, (int)nb, k, part, out);
  ATLG_LAUNCHED();
}

// One GMRES-IR correction (Carson and Higham, SIAM J. Sci. Comput. 40(2),
// 2018): solves A d = r by GMRES on the left-preconditioned system
// M^-1 A d = M^-1 r, M = P^T L U the low-precision factor (applied in FP32
// by `solve32` on W32 after the FP64 vector is rounded and permuted),
// everything else in FP64; classical Gram-Schmidt applied twice, Givens
// rotations on the host.  Stops after kMaxBasis iterations or when the
// preconditioned residual has dropped by `tol` (a preconditioner applied in
// FP32 limits that drop to about FP32 roundoff; the outer FP64 refinement
// supplies the rest); then x += d.  Returns the iterations used.
template <class Solve>
int gmres_correct(hipStream_t s, long N, const double* A, long lda, const double* r, double* x,
                  Solve&& solve32, float* W32, const int* dpos, double* part_mv, int nchunk,
                  Gmres& g, double tol) {
  const long rb = (N + kRows - 1) / kRows;
  const unsigned gr = grid1(N);
  constexpr int m = kMaxBasis, ld = kMaxBasis + 2;
  double hbuf[3 * ld];
  auto to_host = [&](const double* src, double* dst, int n) {
    ATLG_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  };
  // z = M^-1 r, beta = |z|
  hipLaunchKernelGGL(d2s_kernel, dim3(gr), dim3(256), 0, s, N, 1, r, N, W32, N, dpos);
  ATLG_LAUNCHED();
  solve32();
  hipLaunchKernelGGL(s2d_kernel, dim3(gr), dim3(256), 0, s, N, 1, W32, N, g.w, N, 0);
  ATLG_LAUNCHED();
  dev_dots(s, N, 1, g.w, N, g.w, g.part, g.nrm);
  to_host(g.nrm, hbuf, 1);
  ATLG_HIP(hipStreamSynchronize(s));
  const double beta = std::sqrt(hbuf[0]);
  if (!(beta > 0.0) || !std::isfinite(beta)) return 0;
  hipLaunchKernelGGL(scale_copy_kernel, dim3(gr), dim3(256), 0, s, N, g.w, 1.0 / beta, g.V);
  ATLG_LAUNCHED();
  double R[m][m] = {}, cs[m] = {}, sn[m] = {}, gv[m + 1] = {};
  gv[0] = beta;
  int k = 0;
  for (int i = 0; i < m; i++) {
    double* vi = g.V + (size_t)i * N;
    // w = M^-1 A v_i
    hipLaunchKernelGGL(gemv_partial_kernel, dim3((unsigned)rb, (unsigned)nchunk), dim3(kRows), 0,
                       s, N, 1, A, lda, vi, N, part_mv);
    ATLG_LAUNCHED();
    hipLaunchKernelGGL(matvec_finish_kernel, dim3((unsigned)rb), dim3(kRows), 0, s, N, nchunk,
                       part_mv, W32, dpos);
    ATLG_LAUNCHED();
    solve32();
    hipLaunchKernelGGL(s2d_kernel, dim3(gr), dim3(256), 0, s, N, 1, W32, N, g.w, N, 0);
    ATLG_LAUNCHED();
    // orthogonalize against v_0..v_i, twice
    const int nk = i + 1;
    dev_dots(s, N, nk, g.V, N, g.w, g.part, g.hA);
    hipLaunchKernelGGL(axpy_basis_kernel, dim3((unsigned)rb), dim3(kRows), 0, s, N, nk, g.V, N,
                       g.hA, -1.0, g.w);
    ATLG_LAUNCHED();
    dev_dots(s, N, nk, g.V, N, g.w, g.part, g.hB);
    hipLaunchKernelGGL(axpy_basis_kernel, dim3((unsigned)rb), dim3(kRows), 0, s, N, nk, g.V, N,
                       g.hB, -1.0, g.w);
    ATLG_LAUNCHED();
    dev_dots(s, N, 1, g.w, N, g.w, g.part, g.nrm);
    to_host(g.hA, hbuf, 3 * ld);  // hA, hB, nrm are contiguous
    ATLG_HIP(hipStreamSynchronize(s));
    double h[m + 1];
    for (int j = 0; j < nk; j++) h[j] = hbuf[j] + hbuf[ld + j];
    const double hk = std::sqrt(hbuf[2 * ld]);
    for (int j = 0; j < i; j++) {  // earlier rotations
      const double a = h[j], b = h[j + 1];
      h[j] = cs[j] * a + sn[j] * b;
      h[j + 1] = -sn[j] * a + cs[j] * b;
    }
    const double den = std::hypot(h[i], hk);
    cs[i] = den > 0.0 ? h[i] / den : 1.0;
    sn[i] = den > 0.0 ? hk / den : 0.0;
    h[i] = den;
    gv[i + 1] = -sn[i] * gv[i];
    gv[i] = cs[i] * gv[i];
    for (int j = 0; j <= i; j++) R[j][i] = h[j];
    k = nk;
    if (!(hk > 0.0) || std::fabs(gv[i + 1]) <= tol * beta || i + 1 == m) break;
    hipLaunchKernelGGL(scale_copy_kernel, dim3(gr), dim3(256), 0, s, N, g.w, 1.0 / hk,
                       g.V + (size_t)(i + 1) * N);
    ATLG_LAUNCHED();
  }
  // y = R^-1 g, x += V y
  double yv[m];
  for (int i = k - 1; i >= 0; i--) {
    double acc = gv[i];
    for (int j = i + 1; j < k; j++) acc -= R[i][j] * yv[j];
    yv[i] = (R[i][i] != 0.0) ? acc / R[i][i] : 0.0;
  }
  ATLG_HIP(hipMemcpyAsync(g.y, yv, sizeof(double) * k, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(axpy_basis_kernel, dim3((unsigned)rb), dim3(kRows), 0, s, N, k, g.V, N, g.y,
                     1.0, x);
  ATLG_LAUNCHED();
  ATLG_HIP(hipStreamSynchronize(s));  // yv lives on this stack frame
  return k;
}

}  // namespace

int dsgesv(hipStream_t s, long N, long NRHS, double* A, long lda, int* ipiv, const double* B,
           long ldb, double* X, long ldx, int* dinfo, int itermax, int lowprec, int refine,
           int* iter_out, int* inner_out) {
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
  // GMRES-IR: FP64 residuals R64 (N x NRHS), basis, work vector, partials
  const size_t bR = refine ? ((sizeof(double) * N * NRHS + 255) & ~size_t(255)) : 0;
  const size_t bV = refine ? ((sizeof(double) * N * (kMaxBasis + 2) + 255) & ~size_t(255)) : 0;
  const size_t bGp = refine ? ((sizeof(double) * ((N + kRows - 1) / kRows) * (kMaxBasis + 2) +
                                sizeof(double) * 4 * (kMaxBasis + 2) + 255) & ~size_t(255))
                            : 0;
  char* base = static_cast<char*>(
      scratch(s, kSlotMixed, bA32 + bW + bP + bS + bPos + 2 * bInv + bR + bV + bGp));
  float* A32 = reinterpret_cast<float*>(base);
  float* W32 = reinterpret_cast<float*>(base + bA32);
  double* part = reinterpret_cast<double*>(base + bA32 + bW);
  unsigned long long* small = reinterpret_cast<unsigned long long*>(base + bA32 + bW + bP);
  int* dpos = reinterpret_cast<int*>(base + bA32 + bW + bP + bS);
  float* invL = reinterpret_cast<float*>(base + bA32 + bW + bP + bS + bPos);
  float* invU = invL + (size_t)nblk * kTrsmNB * kTrsmNB;
  char* gbase = base + bA32 + bW + bP + bS + bPos + 2 * bInv;
  double* R64 = refine ? reinterpret_cast<double*>(gbase) : nullptr;
  Gmres gw{};
  if (refine) {
    gw.V = reinterpret_cast<double*>(gbase + bR);
    gw.w = gw.V + (size_t)N * (kMaxBasis + 1);
    gw.part = reinterpret_cast<double*>(gbase + bR + bV);
    gw.hA = gw.part + ((N + kRows - 1) / kRows) * (kMaxBasis + 2);
    gw.hB = gw.hA + (kMaxBasis + 2);
    gw.nrm = gw.hB + (kMaxBasis + 2);
    gw.y = gw.nrm + (kMaxBasis + 2);
  }
  int inner = 0;
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
               lowprec == 1   ? kGemmBf16
               : lowprec == 2 ? kGemmBf16x6
               : lowprec == 3 ? kGemmBf16x3
                              : kGemmNative);
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
    auto solve32n = [&](long ncol) {
      trsm<float>(s, kLeft, kLower, kNoTrans, kUnit, N, ncol, 1.0f, A32, N, W32, ldw, nullptr,
                  invL, 0);
      trsm<float>(s, kLeft, kUpper, kNoTrans, kNonUnit, N, ncol, 1.0f, A32, N, W32, ldw,
                  nullptr, invU, 0);
    };
    auto solve32 = [&]() { solve32n(NRHS); };
    auto solve32one = [&]() { solve32n(1); };
    // x = solve32(b)
    hipLaunchKernelGGL(d2s_kernel, dim3(grid1(N * NRHS)), dim3(256), 0, s, N, nr, B, ldb, W32,
                       ldw, dpos);
    ATLG_LAUNCHED();
    solve32();
    hipLaunchKernelGGL(s2d_kernel, dim3(grid1(N * NRHS)), dim3(256), 0, s, N, nr, W32, ldw, X,
                       ldx, 0);
    ATLG_LAUNCHED();
    iter = -(itermax + 1);
    double prev_ratio = HUGE_VAL;
    int slow = 0;
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
                           W32 + (size_t)k0 * ldw, ldw, dpos, small + 2 * k0,
                           R64 ? R64 + (size_t)k0 * N : nullptr);
        ATLG_LAUNCHED();
      }
      ATLG_HIP(hipMemcpyAsync(hs.data(), small, sizeof(unsigned long long) * 2 * nr,
                              hipMemcpyDeviceToHost, s));
      ATLG_HIP(hipStreamSynchronize(s));
      bool ok = true;
      double ratio = 0.0;  // max over columns of ||r|| / (||x|| ||A|| eps sqrt(n))
      for (int q = 0; q < nr; q++) {
        const double rn = as_double(hs[2 * q]), xn = as_double(hs[2 * q + 1]);
        if (!(rn < xn * cte)) ok = false;
        ratio = std::max(ratio, (xn * cte > 0.0) ? rn / (xn * cte) : HUGE_VAL);
      }
      if (ok) {
        iter = k;
        done = true;
        break;
      }
      if (k == itermax) break;
      // GMRES-IR is restarted GMRES(kMaxBasis) whose restarts take a fresh
      // FP64 residual: it continues while the corrections make progress and
      // gives up early (same ITER code as running out of iterations) when
      // two corrections in a row have cut the residual ratio by less than
      // 10%, or after kInnerBudget preconditioned products in all
      if (refine) {
        slow = (ratio > 0.9 * prev_ratio) ? slow + 1 : 0;
        if (slow >= 2 || inner >= kInnerBudget) break;
      }
      prev_ratio = ratio;
      if (refine) {
        // x_q += GMRES(A d = r_q) for every right-hand side
        for (int q = 0; q < nr; q++)
          inner += gmres_correct(s, N, A, lda, R64 + (size_t)q * N, X + (size_t)q * ldx,
                                 solve32one, W32, dpos, part, nchunk, gw, 1e-6);
      } else {
        // x += solve32(r)
        solve32();
        hipLaunchKernelGGL(s2d_kernel, dim3(grid1(N * NRHS)), dim3(256), 0, s, N, nr, W32, ldw,
                           X, ldx, 1);
        ATLG_LAUNCHED();
      }
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
  if (inner_out) *inner_out = inner;
  return 0;
}

}  // namespace atlg
