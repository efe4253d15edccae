// In-LDS inversion of a 128 x 128 lower-triangular block by one 256-thread
// workgroup, shared by the TRSM diagonal-block inverter and the Cholesky /
// LU leaf kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace atlg {

constexpr int kTriN = 128;          // block order
constexpr int kTriLD = kTriN + 1;   // LDS pitch (elements)
constexpr int kTriThreads = 256;

// S[i*kTriLD + k] holds L(i,k) for i >= k and 0 above the diagonal; on return
// it holds inv(L).  Order beyond the used part must be padded with the
// identity.  Method: every 8x8 diagonal tile is inverted in registers by one
// thread (forward substitution with reciprocal diagonals), then for orders
// 16, 32, 64, 128 each diagonal block [[X11, 0], [L21, X22]] with inverted
// halves gets its off-diagonal part X21 = -X22 * (L21 * X11).
// Only the leading order n (<= kTriN) is general: rows and columns at or
// beyond n are the identity padding, so every off-diagonal block that starts
// at or beyond n stays zero and is skipped.
// Must be called by all kTriThreads threads; ends with a barrier.
template <class R>
__device__ inline void lower_inverse_inplace(R* S, int n = kTriN) {
  const int t = threadIdx.x;
  if (t < kTriN / 8 && 8 * t < n) {
    const int o = 8 * t;
    R a[8][8], x[8][8];
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c < 8; c++) a[r][c] = (c <= r) ? S[(o + r) * kTriLD + o + c] : R(0);
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
      for (int c = 0; c <= r; c++) S[(o + r) * kTriLD + o + c] = x[r][c];
  }
  __syncthreads();
  for (int sz = 16; sz <= kTriN; sz *= 2) {
    const int h = sz / 2, hh = h * h;
    if (h >= n) break;  // the remaining levels only meet the identity padding
    const int cnt = (kTriN / sz) * hh / kTriThreads;  // 2, 4, 8, 16 per thread
    R v[16];
    for (int u = 0; u < cnt; u++) {  // T = L21 * X11 (X11 lower)
      const int e = t + kTriThreads * u, q = e / hh, rem = e - q * hh;
      const int r = rem / h, c = rem - r * h, o = q * sz;
      R acc = R(0);
      for (int k = c; k < h; k++) acc += S[(o + h + r) * kTriLD + o + k] * S[(o + k) * kTriLD + o + c];
      v[u] = acc;
    }
    __syncthreads();
    for (int u = 0; u < cnt; u++) {
      const int e = t + kTriThreads * u, q = e / hh, rem = e - q * hh;
      const int r = rem / h, c = rem - r * h, o = q * sz;
      S[(o + h + r) * kTriLD + o + c] = v[u];
    }
    __syncthreads();
    for (int u = 0; u < cnt; u++) {  // X21 = -X22 * T (X22 lower)
      const int e = t + kTriThreads * u, q = e / hh, rem = e - q * hh;
      const int r = rem / h, c = rem - r * h, o = q * sz;
      R acc = R(0);
      for (int k = 0; k <= r; k++)
        acc += S[(o + h + r) * kTriLD + o + h + k] * S[(o + h + k) * kTriLD + o + c];
      v[u] = -acc;
    }
    __syncthreads();
    for (int u = 0; u < cnt; u++) {
      const int e = t + kTriThreads * u, q = e / hh, rem = e - q * hh;
      const int r = rem / h, c = rem - r * h, o = q * sz;
      S[(o + h + r) * kTriLD + o + c] = v[u];
    }
    __syncthreads();
  }
}

}  // namespace atlg
