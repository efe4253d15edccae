// In-LDS inversion of a lower-triangular block (order 32, 64 or 128) by one
// 256-thread workgroup, shared by the TRSM diagonal-block inverter and the
// Cholesky leaf kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace atlg {

constexpr int kTriN = 128;          // largest block order
constexpr int kTriLD = kTriN + 1;   // LDS pitch (elements) of a kTriN block
constexpr int kTriThreads = 256;

// One 16x16x4 matrix-core step per scalar type.  Lane l supplies A(l&15,
// l>>4) and B(l>>4, l&15); register r of lane l receives D(row(l, r), l&15).
template <class R>
struct TriMfma;
template <>
struct TriMfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <>
struct TriMfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
};

// One doubling level of order sz >= 32 (half h = sz/2 >= 16) on the matrix
// cores: for every diagonal block [[X11, 0], [L21, X22]] whose halves are
// already inverted, T = L21 X11, then X21 = -X22 T, by 16x16 output tiles
// spread over the four wavefronts.  The triangular factors' known zeros are
// skipped; blocks whose lower half lies in the identity padding (row >= n)
// are left alone (their L21 is zero).
template <class R, int NT>
__device__ inline void tri_level_mfma(R* S, int sz, int n) {
  constexpr int LD = NT + 1, NW = kTriThreads / 64;
  constexpr int MAXU = (NT * NT / 4096) > 0 ? NT * NT / 4096 : 1;  // tiles per wave
  typedef TriMfma<R> M;
  typedef typename M::acc_t acc_t;
  const int h = sz / 2, th = h / 16, tpb = th * th, nt = (NT / sz) * tpb;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lx = lane & 15, lk = lane >> 4;
  acc_t acc[MAXU];
#pragma unroll
  for (int u = 0; u < MAXU; u++) {  // T = L21 * X11 (X11 lower: k >= column)
    acc[u] = acc_t{0, 0, 0, 0};
    const int tile = wv + NW * u;
    const int q = tile / tpb, rem = tile - q * tpb, ti = rem / th, tj = rem - ti * th;
    const int o = q * sz;
    if (tile < nt && o + h < n) {
      const int ar = o + h + 16 * ti + lx, bc = o + 16 * tj + lx;
      for (int k0 = 16 * tj; k0 < h; k0 += 16) {  // four steps' operands in flight
        R av[4], bv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int k = o + k0 + 4 * q + lk;
          av[q] = S[ar * LD + k];
          bv[q] = S[k * LD + bc];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) acc[u] = M::op(av[q], bv[q], acc[u]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < MAXU; u++) {
    const int tile = wv + NW * u;
    const int q = tile / tpb, rem = tile - q * tpb, ti = rem / th, tj = rem - ti * th;
    const int o = q * sz;
    if (tile < nt && o + h < n) {
#pragma unroll
      for (int r = 0; r < 4; r++)
        S[(o + h + 16 * ti + M::row(lane, r)) * LD + o + 16 * tj + lx] = acc[u][r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < MAXU; u++) {  // X21 = -X22 * T (X22 lower: k <= row)
    acc[u] = acc_t{0, 0, 0, 0};
    const int tile = wv + NW * u;
    const int q = tile / tpb, rem = tile - q * tpb, ti = rem / th, tj = rem - ti * th;
    const int o = q * sz;
    if (tile < nt && o + h < n) {
      const int ar = o + h + 16 * ti + lx, bc = o + 16 * tj + lx;
      for (int k0 = 0; k0 < 16 * (ti + 1); k0 += 16) {
        R av[4], bv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int k = o + h + k0 + 4 * q + lk;
          av[q] = S[ar * LD + k];
          bv[q] = S[k * LD + bc];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) acc[u] = M::op(av[q], bv[q], acc[u]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < MAXU; u++) {
    const int tile = wv + NW * u;
    const int q = tile / tpb, rem = tile - q * tpb, ti = rem / th, tj = rem - ti * th;
    const int o = q * sz;
    if (tile < nt && o + h < n) {
#pragma unroll
      for (int r = 0; r < 4; r++)
        S[(o + h + 16 * ti + M::row(lane, r)) * LD + o + 16 * tj + lx] = -acc[u][r];
    }
  }
  __syncthreads();
}

// S[i*(NT+1) + k] holds L(i,k) for i >= k and 0 above the diagonal; on
// return it holds inv(L).  Order beyond the used part must be padded with
// the identity.  Method: every 8x8 diagonal tile is inverted in registers by
// one thread (forward substitution with reciprocal diagonals); order 16
// combines pairs of them with scalar products; orders 32 .. NT run on the
// matrix cores (tri_level_mfma).  Only the leading order n (<= NT) is
// general: rows and columns at or beyond n are the identity padding, so
// every off-diagonal block that starts at or beyond n stays zero and is
// skipped.  Must be called by all kTriThreads threads; ends with a barrier.
template <class R, int NT = kTriN>
__device__ inline void lower_inverse_inplace(R* S, int n = NT) {
  static_assert(NT == 32 || NT == 64 || NT == 128, "block order");
  constexpr int LD = NT + 1;
  const int t = threadIdx.x;
  if (t < NT / 8 && 8 * t < n) {
    const int o = 8 * t;
    R a[8][8], x[8][8];
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c < 8; c++) a[r][c] = (c <= r) ? S[(o + r) * LD + o + c] : R(0);
#pragma unroll
    for (int c = 0; c < 8; c++) {
#pragma unroll
      for (int r = 0; r < 8; r++) x[r][c] = R(0);
      x[c][c] = R(1) / a[c][c];
#pragma unroll
      for (int r = c + 1; r < 8; r++) {
        R acc = R(0);
#pragma unroll
        for (int k = c; k < r; k++) acc += a[r][k] * x[k][c];
        x[r][c] = -acc * (R(1) / a[r][r]);
      }
    }
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c <= r; c++) S[(o + r) * LD + o + c] = x[r][c];
  }
  __syncthreads();
  if (n > 8) {  // order 16: 8x8 halves, 64 entries per block
    constexpr int TOT = (NT / 16) * 64, CNT = (TOT + kTriThreads - 1) / kTriThreads;
    R v[CNT];
#pragma unroll
    for (int u = 0; u < CNT; u++) {  // T = L21 * X11
      const int e = t + kTriThreads * u, q = e >> 6, r = (e >> 3) & 7, c = e & 7, o = 16 * q;
      R acc = R(0);
      if (e < TOT)
        for (int k = c; k < 8; k++) acc += S[(o + 8 + r) * LD + o + k] * S[(o + k) * LD + o + c];
      v[u] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CNT; u++) {
      const int e = t + kTriThreads * u, q = e >> 6, r = (e >> 3) & 7, c = e & 7, o = 16 * q;
      if (e < TOT) S[(o + 8 + r) * LD + o + c] = v[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CNT; u++) {  // X21 = -X22 * T
      const int e = t + kTriThreads * u, q = e >> 6, r = (e >> 3) & 7, c = e & 7, o = 16 * q;
      R acc = R(0);
      if (e < TOT)
        for (int k = 0; k <= r; k++)
          acc += S[(o + 8 + r) * LD + o + 8 + k] * S[(o + 8 + k) * LD + o + c];
      v[u] = -acc;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CNT; u++) {
      const int e = t + kTriThreads * u, q = e >> 6, r = (e >> 3) & 7, c = e & 7, o = 16 * q;
      if (e < TOT) S[(o + 8 + r) * LD + o + c] = v[u];
    }
    __syncthreads();
  }
  for (int sz = 32; sz <= NT; sz *= 2) {
    if (sz / 2 >= n) break;  // the remaining levels only meet the identity padding
    tri_level_mfma<R, NT>(S, sz, n);
  }
}

}  // namespace atlg
