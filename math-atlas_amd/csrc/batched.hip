// Batched Cholesky factor/solve of many independent SPD systems of uneven
// order in one persistent launch.  ATLAS factors one matrix per call
// (ATL_dpotrf / ATL_dpotrs, AtlasBase/Clint/atlas-lp.base:332-360, 969-1050);
// the per-system arithmetic here is the same lower Cholesky A = L L^T and the
// same pair of triangular solves, and the mixed-precision driver follows
// LAPACK's dsposv refinement (single-precision factor, double-precision
// residual, stop when ||r||_inf < ||x||_inf * ||A||_inf * eps * sqrt(n) for
// every right-hand side, at most 30 iterations).
//
// Work distribution: the host orders the systems longest first (n^3 is the
// work) and every workgroup pulls the next system from an atomic counter
// until the queue is drained, so every workgroup exits after the last system.
// One workgroup (kBW waves) owns a system from start to finish.
//
// Per system, left-looking by block columns of width T (16 or 32, chosen per
// system on the host by the decision tree the empirical search generates):
//   U: C = A[J:n, J:J+T] - L[J:n, 0:J] L[J:J+T, 0:J]^T   (MFMA 16x16x4 tiles,
//      operands straight from global memory, lower part written back);
//   F: one wave factors the T x T diagonal block in LDS and forms its inverse;
//   T: L[J+T:n, J:J+T] = C * inv(L_JJ)^T                 (MFMA, inverse in LDS).
// The solve sweeps block rows: one wave solves the diagonal block for all
// right-hand sides with shuffles, then every thread updates the remaining
// rows (forward: row-parallel, backward: column-parallel over L^T).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "atlg_internal.hpp"
#include "atlg_batched_tune.h"
#include "tri_inv.hpp"

namespace atlg {
namespace {

constexpr int kBW = 4;          // waves per workgroup
constexpr int kBT = 64 * kBW;   // threads per workgroup
constexpr int kTmax = 32;       // widest block column
constexpr int kDP = kTmax + 1;  // LDS pitch of the diagonal block and inverse

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// The 16x16x4 MFMA for each scalar type.  D = X1 * X2 with lane l supplying
// X1[l&15][l>>4] and X2[l>>4][l&15]; lane l, register r holds
// D[row(l, r)][l & 15].
template <class S>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t op(double x1, double x2, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x1, x2, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t op(float x1, float x2, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x1, x2, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

static_assert(kBT == kTriThreads, "the in-LDS inverse takes the whole workgroup");

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of one workgroup.  `D` and `V` are typed per scalar at use; the
// mixed solver's residual uses the same space as `XS`.  `Y` holds up to four
// right-hand sides of a system of order <= kYmax / 4 while its factorization
// runs the forward substitution alongside.
constexpr int kYmax = 2048;
struct alignas(16) Smem {
  union {
    struct {
      double D[kTmax * kDP];  // diagonal block (factor)
      double V[kTmax * kDP];  // its inverse
    };
    double XS[2 * kTmax * kDP];  // residual: X[0:n, r0:r0+4] as [row][q]
  };
  double red[kBW * 4];      // reductions
  double Y[kYmax];          // a fused forward solve's right-hand sides, [row][q] for q < 4
  double bs[kTmax * 4];     // a solve's current block of solved values (first 4 right-hand sides)
  double rd[kTmax];         // a solve's reciprocal pivots of the current diagonal block
  int flag;                 // failure index + 1, 0 = fine
  int cur;                  // system taken from the queue
};

// ---- U: C = A[J:n, J:J+w] - L[J:n, 0:J] L[J:J+w, 0:J]^T (lower part) -----
// The block row L[J:J+w, 0:J] that every row tile multiplies is staged in
// LDS in chunks of kUC columns (the diagonal block's space, free between a
// block column's TRSM and the next factor); each wave takes two 16-row
// tiles per pass, so a staged chunk serves eight row tiles and the global
// loads are the row tiles' own L rows.
constexpr int kUC = 64;          // k per staged chunk
constexpr int kUP = kTmax + 1;   // LDS pitch of a staged k row
static_assert(kUC * kUP <= 2 * kTmax * kDP, "staging fits the diagonal block's space");
template <class S, int T>
__device__ void update_block(int n, S* __restrict__ A, int lda, int J, int w, S* __restrict__ Ls) {
  typedef Mfma<S> M;
  constexpr int CT = T / 16, NS = kUC * T / kBT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int RT = (n - J + 15) >> 4;
  for (int p0 = 0; p0 < RT; p0 += 2 * kBW) {
    int row[2];
    bool rv[2];
    const S* pr[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      row[h] = J + 16 * (p0 + 2 * wave + h) + li;
      rv[h] = row[h] < n;
      pr[h] = A + (rv[h] ? row[h] : 0) + (size_t)lk * lda;
    }
    typename M::acc_t acc[2][CT];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int ct = 0; ct < CT; ct++) acc[h][ct] = typename M::acc_t{0, 0, 0, 0};
    for (int k0 = 0; k0 < J; k0 += kUC) {
      const int kc = min(kUC, J - k0);  // a multiple of 4 (J is)
      S v[NS];
#pragma unroll
      for (int u = 0; u < NS; u++) {
        const int e = threadIdx.x + u * kBT, k = e / T, c = e % T;
        v[u] = (k < kc && c < w) ? A[(J + c) + (size_t)(k0 + k) * lda] : S(0);
      }
      __syncthreads();  // the previous chunk's (or phase's) readers are done
#pragma unroll
      for (int u = 0; u < NS; u++) {
        const int e = threadIdx.x + u * kBT;
        Ls[(e / T) * kUP + e % T] = v[u];
      }
      __syncthreads();
      // four k steps at a time (kc is a multiple of 16: J is)
      for (int q0 = 0; q0 < (kc >> 2); q0 += 4) {
        S x2[4][2];
#pragma unroll
        for (int dq = 0; dq < 4; dq++) {
          const size_t off = (size_t)(k0 + 4 * (q0 + dq)) * lda;
          x2[dq][0] = rv[0] ? pr[0][off] : S(0);
          x2[dq][1] = rv[1] ? pr[1][off] : S(0);
        }
#pragma unroll
        for (int dq = 0; dq < 4; dq++)
#pragma unroll
          for (int ct = 0; ct < CT; ct++) {
            const S x1 = Ls[(4 * (q0 + dq) + lk) * kUP + 16 * ct + li];
            acc[0][ct] = M::op(x1, x2[dq][0], acc[0][ct]);
            acc[1][ct] = M::op(x1, x2[dq][1], acc[1][ct]);
          }
      }
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int row0 = J + 16 * (p0 + 2 * wave + h);
#pragma unroll
      for (int ct = 0; ct < CT; ct++) {
        if (row0 + 15 < J + 16 * ct) continue;  // tile wholly above the diagonal
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int col = J + 16 * ct + M::row(lane, r);
          if (rv[h] && col < J + w && row[h] >= col) A[row[h] + (size_t)col * lda] -= acc[h][ct][r];
        }
      }
    }
  }
}

// D (pitch kDP) = the lower triangle of A[J:J+w, J:J+w], zero padded to
// T x T, by wave 0: four loads per lane in flight at a time (one at a time
// costs a memory round trip per element row when the device is busy).
template <class S, int T>
__device__ __forceinline__ void load_diag(const S* __restrict__ A, int lda, int J, int w, S* D) {
  constexpr int NF = T * T / 64, NBT = 4;
  static_assert(NF % NBT == 0, "fill batches");
  const int l = threadIdx.x & 63;
#pragma unroll 1
  for (int u0 = 0; u0 < NF; u0 += NBT) {
    S v[NBT];
#pragma unroll
    for (int u = 0; u < NBT; u++) {
      const int e = l + 64 * (u0 + u), i = e % T, c = e / T;
      v[u] = (i < w && c <= i) ? A[(J + i) + (size_t)(J + c) * lda] : S(0);
    }
#pragma unroll
    for (int u = 0; u < NBT; u++) {
      const int e = l + 64 * (u0 + u), i = e % T, c = e / T;
      D[i * kDP + c] = v[u];
    }
  }
}

// ---- F: factor the w x w diagonal block (wave 0) ---------------------------
// Lane i holds row i of the block in registers and the factorization runs
// column by column without barriers: the pivot from lane j by v_readlane,
// the multipliers scaled by the reciprocal square root (hardware estimate
// and Newton steps), column j broadcast to the lanes through a wave-private
// LDS line (a wave's LDS accesses complete in order), read back eight at a
// time.  Returns 0, or j+1 for the first non-positive pivot.  L_JJ goes back
// to A, and to V (pitch kDP) padded with the identity to kTmax for the
// in-LDS inverse that follows.
template <class S>
__device__ __forceinline__ S lane_val(S v, int src);
template <>
__device__ __forceinline__ double lane_val<double>(double v, int src) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(u), src);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(u >> 32), src);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <>
__device__ __forceinline__ float lane_val<float>(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
// 1/sqrt(d) for d > 0: the hardware estimate refined by Newton steps (two
// from the FP64 estimate's ~2^-23, one in FP32); tiny d takes sqrt and a
// division, where the estimate loses range.
template <class S>
__device__ __forceinline__ S rsqrt_p(S d);
template <>
__device__ __forceinline__ double rsqrt_p<double>(double d) {
  if (d < 0x1p-1000) return 1.0 / sqrt(d);
  double y = __builtin_amdgcn_rsq(d);
  const double h = 0.5 * d;
  y = y * (1.5 - h * y * y);
  y = y * (1.5 - h * y * y);
  return y;
}
template <>
__device__ __forceinline__ float rsqrt_p<float>(float d) {
  if (d < 0x1p-100f) return 1.0f / sqrtf(d);
  float y = __builtin_amdgcn_rsqf(d);
  y = y * (1.5f - 0.5f * d * y * y);
  return y;
}

template <class S, int T>
__device__ int factor_diag(S* __restrict__ A, int lda, int J, int w, S* V, S* cb) {
  constexpr int CB = 8;  // broadcast values read per batch
  const int i = threadIdx.x;  // wave 0 only
  S x[T];
#pragma unroll
  for (int c = 0; c < T; c++)
    x[c] = (i < w && c <= i) ? A[(J + i) + (size_t)(J + c) * lda] : S(0);
  int f = 0;  // uniform: the loop is fully unrolled, its branches uniform
#pragma unroll
  for (int j = 0; j < T; j++) {
    if (j < w && f == 0) {
      const S d = lane_val(x[j], j);
      if (!(d > S(0))) {
        f = j + 1;  // d is the same in every lane
      } else {
        const S y = rsqrt_p(d);
        const S L = (i == j) ? d * y : x[j] * y;
        if (i >= j) x[j] = L;
        const S Lm = (i > j && i < w) ? L : S(0);  // lanes <= j: no update
        cb[i] = L;
#pragma unroll
        for (int c0 = j + 1; c0 < T; c0 += CB) {
          S cv[CB];
#pragma unroll
          for (int u = 0; u < CB; u++)
            if (c0 + u < T) cv[u] = cb[c0 + u];
#pragma unroll
          for (int u = 0; u < CB; u++)
            if (c0 + u < T) x[c0 + u] -= Lm * cv[u];
        }
      }
    }
  }
  if (f) return f;
#pragma unroll
  for (int c = 0; c < T; c++)
    if (i < w && c <= i) A[(J + i) + (size_t)(J + c) * lda] = x[c];
  if (i < kTmax) {
#pragma unroll
    for (int c = 0; c < kTmax; c++) {
      S v = (i == c) ? S(1) : S(0);
      if (c < T && i < w && c <= i) v = x[c];
      V[i * kDP + c] = v;
    }
  }
  return 0;
}

// ---- T: A[J+w:n, J:J+w] <- A[J+w:n, J:J+w] * V^T ---------------------------
// With Y (a fused forward solve, y_J already in Y[J:J+w]): also
// Y[J+w:n] -= A[J+w:n, J:J+w] y_J, from the new rows still in registers; the
// four lanes of a row each hold a quarter of its columns, summed by two
// shuffles, and lane quarter q writes right-hand side q.
template <class S, int T>
__device__ void trsm_block(int n, S* __restrict__ A, int lda, int J, int w, const S* V,
                           S* __restrict__ Y) {
  typedef Mfma<S> M;
  constexpr int CT = T / 16, KQ = T / 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = J + w;
  const int RT = (n - r0 + 15) >> 4;
  for (int rt = wave; rt < RT; rt += kBW) {
    const int row = r0 + 16 * rt + li;
    const bool rv = row < n;
    S x2[KQ];
#pragma unroll
    for (int q = 0; q < KQ; q++) {
      const int k = 4 * q + lk;
      x2[q] = (rv && k < w) ? A[row + (size_t)(J + k) * lda] : S(0);
    }
    S part[4] = {S(0), S(0), S(0), S(0)};
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
      typename M::acc_t acc = {0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < KQ; q++) acc = M::op(V[(16 * ct + li) * kDP + 4 * q + lk], x2[q], acc);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = 16 * ct + M::row(lane, r);
        if (rv && col < w) A[row + (size_t)(J + col) * lda] = acc[r];
        if (Y && col < w) {
#pragma unroll
          for (int q = 0; q < 4; q++) part[q] += acc[r] * Y[(J + col) * 4 + q];
        }
      }
    }
    if (Y) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        part[q] += __shfl_xor(part[q], 16, 64);
        part[q] += __shfl_xor(part[q], 32, 64);
      }
      const S mine = lk == 0 ? part[0] : (lk == 1 ? part[1] : (lk == 2 ? part[2] : part[3]));
      if (rv) Y[row * 4 + lk] -= mine;
    }
  }
}

// Whether a factorization can carry the forward solve of nrhs right-hand
// sides in LDS.
__device__ __forceinline__ bool fuse_fits(int n, int nrhs) {
  return nrhs > 0 && nrhs <= 4 && 4 * n <= kYmax;
}

// Factors one system; returns 0 or the LAPACK info (first failing minor).
// With B (fuse_fits(n, nrhs)): also solves L Y = B, by the right-looking
// forward substitution in LDS (y_J = inv(L_JJ) b_J, then the block column's
// TRSM subtracts L[J+w:n, J] y_J), and on success writes Y over B; B is left
// as it was when the factorization fails.
template <class S, int T>
__device__ int chol_sys(int n, S* __restrict__ A, int lda, Smem& sm, S* __restrict__ B = nullptr,
                        int ldb = 0, int nrhs = 0) {
  S* D = reinterpret_cast<S*>(sm.D);
  S* V = reinterpret_cast<S*>(sm.V);
  S* Y = B ? reinterpret_cast<S*>(sm.Y) : nullptr;
  const int t = threadIdx.x;
  if (Y) {
    for (int e = t; e < 4 * n; e += kBT) {
      const int i = e >> 2, q = e & 3;
      Y[e] = q < nrhs ? B[i + (size_t)q * ldb] : S(0);
    }
  }
  for (int J = 0; J < n; J += T) {
    const int w = min(T, n - J);
    if (J > 0) update_block<S, T>(n, A, lda, J, w, D);  // stages in D and V's space
    __syncthreads();
    if (t < 64) {
      const int f = factor_diag<S, T>(A, lda, J, w, V, D);  // D: the broadcast line
      if (t == 0) sm.flag = f ? J + f : 0;
    }
    __syncthreads();
    const int f = sm.flag;
    if (f) return f;
    lower_inverse_inplace<S, kTmax>(V, w);  // V <- inv(L_JJ), all threads
    if (Y) {  // y_J = inv(L_JJ) b_J in place: every read, a barrier, the writes
      const int i = t >> 2, q = t & 3;
      S y = S(0);
      if (i < w)
        for (int c = 0; c <= i; c++) y += V[i * kDP + c] * Y[(J + c) * 4 + q];
      __syncthreads();
      if (i < w) Y[(J + i) * 4 + q] = y;
      __syncthreads();
    }
    trsm_block<S, T>(n, A, lda, J, w, V, Y);
    __syncthreads();
  }
  if (Y) {
    for (int e = t; e < 4 * n; e += kBT) {
      const int i = e >> 2, q = e & 3;
      if (q < nrhs) B[i + (size_t)q * ldb] = Y[e];
    }
    __syncthreads();
  }
  return 0;
}

// acc[q] += sum_{c < w} p[c*st] * b[c + (r0+q)*ldb] for the right-hand sides
// r0 + q < nrhs, in order of c, four loads of p in flight at a time (the
// mixed solver's FP64 residual).
template <class S>
__device__ __forceinline__ void row_dot4(const S* p, size_t st, int w, const S* b, int sc, int sq,
                                         int r0, int nrhs, S (&acc)[4]) {
  int c = 0;
#pragma unroll 1
  for (; c + 4 <= w; c += 4) {
    S lv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) lv[u] = p[(size_t)(c + u) * st];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (r0 + q < nrhs) acc[q] += lv[u] * b[(size_t)(c + u) * sc + (size_t)q * sq];
  }
  for (; c < w; c++) {
    const S lv = p[(size_t)c * st];
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (r0 + q < nrhs) acc[q] += lv * b[(size_t)c * sc + (size_t)q * sq];
  }
}

// Wave 0: solves the w x w diagonal block, L_JJ y = b (TR: L_JJ^T y = b),
// for every right-hand side in place in B; the first four go to BS too.
// Lane (i, g) works on row i and right-hand side r0 + g.  FP32: the block's
// row (TR: column) i in registers, step j scales y_j by its reciprocal pivot
// in lane j and shuffles it to the group.  FP64: the block in LDS with the
// reciprocal pivots, which measured faster there (the registers of a row of
// 32 doubles cost the kernel spills elsewhere).
template <class S, int T, bool TR>
__device__ void diag_solve(const S* __restrict__ L, int lda, int J, int w, S* __restrict__ B,
                           int ldb, int nrhs, S* BS, Smem& sm) {
  constexpr int RG = 64 / T;  // right-hand sides per pass
  const int l = threadIdx.x & 63, i = l % T, g = l / T;
  if constexpr (sizeof(S) == 4) {
    S x[T];
#pragma unroll
    for (int c = 0; c < T; c++) {
      const bool in = i < w && c < w && (TR ? c >= i : c <= i);
      x[c] = in ? (TR ? L[(J + c) + (size_t)(J + i) * lda] : L[(J + i) + (size_t)(J + c) * lda])
                : S(0);
    }
    const S rinv = i < w ? S(1) / L[(J + i) + (size_t)(J + i) * lda] : S(0);
    for (int r0 = 0; r0 < nrhs; r0 += RG) {
      const int r = r0 + g;
      const bool act = i < w && r < nrhs;
      S y = act ? B[(J + i) + (size_t)r * ldb] : S(0);
#pragma unroll
      for (int jj = 0; jj < T; jj++) {
        const int j = TR ? T - 1 - jj : jj;
        if (j < w) {
          if (i == j) y *= rinv;
          const S yj = __shfl(y, g * T + j, 64);
          if (TR ? i < j : i > j) y -= x[j] * yj;
        }
      }
      if (act) B[(J + i) + (size_t)r * ldb] = y;
      if (act && r < 4) BS[i * 4 + r] = y;
    }
  } else {
    S* D = reinterpret_cast<S*>(sm.D);
    S* RD = reinterpret_cast<S*>(sm.rd);
    load_diag<S, T>(L, lda, J, w, D);
    wave_sync();
    if (l < w) RD[l] = S(1) / D[l * kDP + l];  // one division per pivot, not per step and lane
    wave_sync();
    for (int r0 = 0; r0 < nrhs; r0 += RG) {
      const int r = r0 + g;
      const bool act = i < w && r < nrhs;
      S y = act ? B[(J + i) + (size_t)r * ldb] : S(0);
      for (int jj = 0; jj < w; jj++) {
        const int j = TR ? w - 1 - jj : jj;
        const S yj = __shfl(y, g * T + j, 64) * RD[j];
        if (i == j) y = yj;
        else if (TR ? i < j : i > j) y -= (TR ? D[j * kDP + i] : D[i * kDP + j]) * yj;
      }
      if (act) B[(J + i) + (size_t)r * ldb] = y;
      if (act && r < 4) BS[i * 4 + r] = y;
    }
  }
}

// The rows outside the diagonal block take its solved values on the matrix
// cores: forward, B[J+w:n] -= L[J+w:n, J:J+w] Y; backward (TR), B[0:J] -=
// L[J:J+w, 0:J]^T X.  Each wave takes 16-row tiles and computes, four
// right-hand sides at a time, the tile's update transposed (D = Y^T L^T), so
// that a lane's result is one row of one right-hand side and the loads of L
// cover 16 consecutive rows; all of a tile's loads are in flight together.
template <class S, int T, bool TR>
__device__ void offdiag_update(const S* __restrict__ L, int lda, int J, int w, int n,
                               S* __restrict__ B, int ldb, int nrhs, const S* BS) {
  typedef Mfma<S> M;
  constexpr int KQ = T / 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const int rbeg = TR ? 0 : J + w, rend = TR ? J : n;
  const int RT = (rend - rbeg + 15) >> 4;
  for (int r0 = 0; r0 < nrhs; r0 += 4) {
    for (int rt = wave; rt < RT; rt += kBW) {
      const int row = rbeg + 16 * rt + li;
      const bool rv = row < rend;
      S lv[KQ];
#pragma unroll
      for (int q = 0; q < KQ; q++) {
        const int c = 4 * q + lk;
        lv[q] = (rv && c < w) ? (TR ? L[(J + c) + (size_t)row * lda] : L[row + (size_t)(J + c) * lda])
                              : S(0);
      }
      typename M::acc_t acc = {0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < KQ; q++) {
        const int c = 4 * q + lk;
        S yv = S(0);
        if (li < 4 && c < w && r0 + li < nrhs)
          yv = r0 == 0 ? BS[c * 4 + li] : B[(J + c) + (size_t)(r0 + li) * ldb];
        acc = M::op(yv, lv[q], acc);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int q = M::row(lane, r);  // right-hand side within the group of four
        if (rv && q < 4 && r0 + q < nrhs) B[row + (size_t)(r0 + q) * ldb] -= acc[r];
      }
    }
  }
}

// Solves L L^T X = B in place (B: n x nrhs, column-major, ldb); with fwd
// false B already holds Y = L^-1 B and only L^T X = Y is solved.
template <class S, int T>
__device__ void solve_sys(int n, const S* __restrict__ L, int lda, S* __restrict__ B, int ldb,
                          int nrhs, Smem& sm, bool fwd = true) {
  if (n <= 0) return;
  S* BS = reinterpret_cast<S*>(sm.bs);  // solved block, [c][q] for q < 4
  const int t = threadIdx.x;
  // forward: L Y = B
  for (int J = 0; fwd && J < n; J += T) {
    const int w = min(T, n - J);
    if (t < 64) diag_solve<S, T, false>(L, lda, J, w, B, ldb, nrhs, BS, sm);
    __syncthreads();
    offdiag_update<S, T, false>(L, lda, J, w, n, B, ldb, nrhs, BS);
    __syncthreads();
  }
  // backward: L^T X = Y, last block row first
  for (int J = ((n - 1) / T) * T; J >= 0; J -= T) {
    const int w = min(T, n - J);
    if (t < 64) diag_solve<S, T, true>(L, lda, J, w, B, ldb, nrhs, BS, sm);
    __syncthreads();
    offdiag_update<S, T, true>(L, lda, J, w, n, B, ldb, nrhs, BS);
    __syncthreads();
  }
}

// width: the system's block-column width (16 or 32), chosen on the host
template <class S>
__device__ int chol_any(int n, S* A, int lda, int width, Smem& sm, S* B = nullptr, int ldb = 0,
                        int nrhs = 0) {
  return (width == 16) ? chol_sys<S, 16>(n, A, lda, sm, B, ldb, nrhs)
                       : chol_sys<S, 32>(n, A, lda, sm, B, ldb, nrhs);
}
template <class S>
__device__ void solve_any(int n, const S* L, int lda, S* B, int ldb, int nrhs, int width,
                          Smem& sm, bool fwd = true) {
  if (width == 16) solve_sys<S, 16>(n, L, lda, B, ldb, nrhs, sm, fwd);
  else solve_sys<S, 32>(n, L, lda, B, ldb, nrhs, sm, fwd);
}

// Takes the next queue entry (or -1 when drained); every thread gets it.
__device__ int next_system(unsigned* counter, int nsys, Smem& sm) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned q = atomicAdd(counter, 1u);
    sm.cur = (q < (unsigned)nsys) ? (int)q : -1;
  }
  __syncthreads();
  return sm.cur;
}

template <class S>
struct SysDesc {
  S* A;
  S* B;
  const S* Bsrc;  // copied into B first when not null
  int n, lda, ldb, idx, width;
};

// mode: 0 factor, 1 factor + solve, 2 solve with a given factor.
// Occupancy: the compiler's unconstrained allocation leaves two 256-thread workgroups per CU;
// capping registers trades spills for workgroups in flight, which hide the latency the phases
// wait on.  10 000 uneven systems, n 16..512, 4 right-hand sides: 4 waves per SIMD 16.3 ms,
// 3 waves 16.4, 5 waves 19.2 (before the register factor of the diagonal block, 5 waves was
// the best: 21.2 ms; 25.3 unconstrained).
template <class S>
__global__ void __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(4)))
posv_batched_kernel(int nsys, const SysDesc<S>* __restrict__ desc, int nrhs, int mode,
                    int* __restrict__ info, unsigned* __restrict__ counter) {
  __shared__ Smem sm;
  for (;;) {
    const int q = next_system(counter, nsys, sm);
    if (q < 0) return;
    const SysDesc<S> d = desc[q];
    if (d.Bsrc) {
      for (int r = 0; r < nrhs; r++)
        for (int i = threadIdx.x; i < d.n; i += kBT)
          d.B[i + (size_t)r * d.ldb] = d.Bsrc[i + (size_t)r * d.ldb];
      __syncthreads();
    }
    // factor + solve: the forward substitution runs inside the factorization
    // when the right-hand sides fit in LDS
    const bool fuse = mode == 1 && fuse_fits(d.n, nrhs);
    int f = 0;
    if (mode != 2)
      f = fuse ? chol_any<S>(d.n, d.A, d.lda, d.width, sm, d.B, d.ldb, nrhs)
               : chol_any<S>(d.n, d.A, d.lda, d.width, sm);
    if (f == 0 && mode != 0) solve_any<S>(d.n, d.A, d.lda, d.B, d.ldb, nrhs, d.width, sm, !fuse);
    if (threadIdx.x == 0 && mode != 2) info[d.idx] = f;
  }
}

// ---- mixed precision: FP32 factor, FP64 residual and update ---------------
struct MixDesc {
  const double* A;  // FP64 matrix (lower triangle read)
  const double* B;  // FP64 right-hand sides
  double* X;        // FP64 solution
  float* A32;       // workspace n x n (ld n)
  float* W32;       // workspace n x nrhs (ld n)
  int n, lda, ldb, ldx, idx, width;
};

__device__ double block_max(double v, Smem& sm) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sm.red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = sm.red[0];
  for (int k = 1; k < kBW; k++) m = fmax(m, sm.red[k]);
  return m;
}

// r = B - A X for the rows this thread owns; returns max |x| and max |r| of
// right-hand side r0..r0+3 across the workgroup, writing float(r) to W32.
__device__ void residual(const MixDesc& d, int nrhs, Smem& sm, bool* conv, double cte) {
  const int n = d.n;
  // X staged in LDS when it fits (n <= 528): the rows' dot products then
  // read it there instead of from global memory
  const bool stage = 4 * n <= 2 * kTmax * kDP;
  for (int r0 = 0; r0 < nrhs; r0 += 4) {
    double xm[4] = {0, 0, 0, 0}, rm[4] = {0, 0, 0, 0};
    if (stage) {
      for (int e = threadIdx.x; e < 4 * n; e += kBT) {
        const int j = e >> 2, q = e & 3;
        sm.XS[e] = (r0 + q < nrhs) ? d.X[j + (size_t)(r0 + q) * d.ldx] : 0.0;
      }
      __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += kBT) {
      double acc[4] = {0, 0, 0, 0};
      // row i of the lower triangle, then column i below the diagonal
      if (stage) {
        row_dot4<double>(d.A + i, (size_t)d.lda, i + 1, sm.XS, 4, 1, r0, nrhs, acc);
        row_dot4<double>(d.A + (size_t)i * d.lda + i + 1, 1, n - i - 1, sm.XS + 4 * (i + 1), 4,
                         1, r0, nrhs, acc);
      } else {
        const double* xr = d.X + (size_t)r0 * d.ldx;
        row_dot4<double>(d.A + i, (size_t)d.lda, i + 1, xr, 1, d.ldx, r0, nrhs, acc);
        row_dot4<double>(d.A + (size_t)i * d.lda + i + 1, 1, n - i - 1, xr + i + 1, 1, d.ldx, r0,
                         nrhs, acc);
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (r0 + q < nrhs) {
          const double rv = d.B[i + (size_t)(r0 + q) * d.ldb] - acc[q];
          xm[q] = fmax(xm[q], fabs(d.X[i + (size_t)(r0 + q) * d.ldx]));
          rm[q] = fmax(rm[q], fabs(rv));
          d.W32[i + (size_t)(r0 + q) * n] = (float)rv;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (r0 + q >= nrhs) break;
      const double xn = block_max(xm[q], sm), rn = block_max(rm[q], sm);
      if (!(rn < xn * cte)) *conv = false;
    }
  }
}

// Same trade as above, measured separately (dsposv on the same 10 000 systems: 26.7 ms at 4
// waves per SIMD, 31.2 at 3, 29.5 at 5; 48.6 unconstrained before the staged update).
__global__ void __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(4)))
mixed_posv_kernel(int nsys, const MixDesc* __restrict__ desc, int nrhs, int itermax,
                  int* __restrict__ info, int* __restrict__ iter, unsigned* __restrict__ counter) {
  __shared__ Smem sm;
  for (;;) {
    const int q = next_system(counter, nsys, sm);
    if (q < 0) return;
    const MixDesc d = desc[q];
    const int n = d.n;
    // ||A||_inf of the symmetric matrix, A32 = float(lower(A)), overflow check
    double rs = 0.0;
    bool over = false;
    for (int i = threadIdx.x; i < n; i += kBT) {
      // row i of the lower triangle (converted to A32), then column i below
      // the diagonal; four loads in flight, sums in order of j
      double s = 0.0;
      int j = 0;
#pragma unroll 1
      for (; j + 4 <= i + 1; j += 4) {
        double av[4];
#pragma unroll
        for (int u = 0; u < 4; u++) av[u] = d.A[i + (size_t)(j + u) * d.lda];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          s += fabs(av[u]);
          if (fabs(av[u]) > (double)FLT_MAX) over = true;
          d.A32[i + (size_t)(j + u) * n] = (float)av[u];
        }
      }
      for (; j <= i; j++) {
        const double a = d.A[i + (size_t)j * d.lda];
        s += fabs(a);
        if (fabs(a) > (double)FLT_MAX) over = true;
        d.A32[i + (size_t)j * n] = (float)a;
      }
      const double* col = d.A + (size_t)i * d.lda;
      j = i + 1;
#pragma unroll 1
      for (; j + 4 <= n; j += 4) {
        double av[4];
#pragma unroll
        for (int u = 0; u < 4; u++) av[u] = col[j + u];
#pragma unroll
        for (int u = 0; u < 4; u++) s += fabs(av[u]);
      }
      for (; j < n; j++) s += fabs(col[j]);
      rs = fmax(rs, s);
    }
    const double anrm = block_max(rs, sm);
    const double ov = block_max(over ? 1.0 : 0.0, sm);
    int it = 0;
    if (n == 0) {
      it = 0;
    } else if (ov != 0.0) {
      it = -2;
    } else {
      // W32 = float(B); the factorization runs its forward solve alongside
      // when the right-hand sides fit in LDS
      for (int r = 0; r < nrhs; r++)
        for (int i = threadIdx.x; i < n; i += kBT)
          d.W32[i + (size_t)r * n] = (float)d.B[i + (size_t)r * d.ldb];
      __syncthreads();
      const bool fuse = fuse_fits(n, nrhs);
      it = (fuse ? chol_any<float>(n, d.A32, n, d.width, sm, d.W32, n, nrhs)
                 : chol_any<float>(n, d.A32, n, d.width, sm)) != 0 ? -3 : 0;
    }
    if (it == 0 && n > 0) {
      // X = float solve of B
      solve_any<float>(n, d.A32, n, d.W32, n, nrhs, d.width, sm, !fuse_fits(n, nrhs));
      for (int r = 0; r < nrhs; r++)
        for (int i = threadIdx.x; i < n; i += kBT)
          d.X[i + (size_t)r * d.ldx] = (double)d.W32[i + (size_t)r * n];
      __syncthreads();
      const double cte = anrm * (DBL_EPSILON * 0.5) * sqrt((double)n);
      it = -(itermax + 1);
      for (int k = 0; k <= itermax; k++) {
        bool conv = true;
        residual(d, nrhs, sm, &conv, cte);  // ends with a barrier (block_max)
        if (conv) {
          it = k;
          break;
        }
        if (k == itermax) break;
        solve_any<float>(n, d.A32, n, d.W32, n, nrhs, d.width, sm);
        for (int r = 0; r < nrhs; r++)
          for (int i = threadIdx.x; i < n; i += kBT)
            d.X[i + (size_t)r * d.ldx] += (double)d.W32[i + (size_t)r * n];
        __syncthreads();
      }
    }
    if (threadIdx.x == 0) {
      iter[d.idx] = it;
      info[d.idx] = 0;
    }
  }
}

__global__ void zero_counter(unsigned* c) { *c = 0u; }

// ---- host side -------------------------------------------------------------

// Pinned staging for the descriptor upload, reused once the previous copy
// from it has completed.
struct Staging {
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;
};
std::mutex g_stage_mu;
Staging g_stage;

void* stage_upload(hipStream_t s, const void* src, size_t bytes, Slot slot) {
  void* dev = scratch(s, slot, bytes);
  std::lock_guard<std::mutex> lk(g_stage_mu);
  if (g_stage.done) ATLG_HIP(hipEventSynchronize(g_stage.done));
  if (bytes > g_stage.cap) {
    if (g_stage.host) ATLG_HIP(hipHostFree(g_stage.host));
    g_stage.host = nullptr;
    g_stage.cap = 0;
    ATLG_HIP(hipHostMalloc(&g_stage.host, bytes, hipHostMallocDefault));
    g_stage.cap = bytes;
  }
  if (!g_stage.done) ATLG_HIP(hipEventCreateWithFlags(&g_stage.done, hipEventDisableTiming));
  std::memcpy(g_stage.host, src, bytes);
  ATLG_HIP(hipMemcpyAsync(dev, g_stage.host, bytes, hipMemcpyHostToDevice, s));
  ATLG_HIP(hipEventRecord(g_stage.done, s));
  return dev;
}

// Longest-first order of the systems (n^3 decides; ties by index).
std::vector<int> lpt_order(int nsys, const int* n) {
  std::vector<int> ord(nsys);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [n](int a, int b) { return n[a] > n[b]; });
  return ord;
}

int grid_for(hipStream_t s, const BatchTune& tune, int nsys) {
  (void)s;
  int dev = 0, cus = 256;
  ATLG_HIP(hipGetDevice(&dev));
  ATLG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(1, std::min(nsys, cus * std::max(1, tune.wg_per_cu)));
}

unsigned* reset_counter(hipStream_t s) {
  unsigned* counter = static_cast<unsigned*>(scratch(s, kSlotBatchCtr, 256));
  hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, s, counter);
  ATLG_LAUNCHED();
  return counter;
}

BatchTune g_tune[2] = {{-1, ATLG_BATCH_WPC_D}, {-1, ATLG_BATCH_WPC_S}};
std::mutex g_tune_mu;

// Block-column width of one system: the override threshold when set, else
// the generated decision tree of the empirical search.
int width_for(const BatchTune& tune, bool single, int n) {
  if (tune.t16_max >= 0) return n <= tune.t16_max ? 16 : 32;
  return single ? atlg_batched_width_s(n) : atlg_batched_width_d(n);
}

}  // namespace

BatchTune batch_tune(bool single) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_tune[single ? 1 : 0];
}
void set_batch_tune(bool single, const BatchTune& t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[single ? 1 : 0] = t;
}

template <class S>
void posv_batched(hipStream_t s, int mode, int nsys, const int* n, S* const* A, const int* lda,
                  S* const* B, const int* ldb, const S* const* Bsrc, int nrhs, int* info,
                  const BatchTune& tune, const int* orig) {
  if (nsys <= 0) return;
  const std::vector<int> ord = lpt_order(nsys, n);
  std::vector<SysDesc<S>> desc(nsys);
  for (int q = 0; q < nsys; q++) {
    const int k = ord[q];
    desc[q] = SysDesc<S>{A[k], B ? B[k] : nullptr, Bsrc ? Bsrc[k] : nullptr, n[k], lda[k],
                         ldb ? ldb[k] : 1, orig ? orig[k] : k,
                         width_for(tune, std::is_same<S, float>::value, n[k])};
  }
  auto* dd = static_cast<const SysDesc<S>*>(
      stage_upload(s, desc.data(), sizeof(SysDesc<S>) * nsys, kSlotBatchDesc));
  unsigned* counter = reset_counter(s);
  hipLaunchKernelGGL(posv_batched_kernel<S>, dim3(grid_for(s, tune, nsys)), dim3(kBT), 0, s, nsys,
                     dd, nrhs, mode, info, counter);
  ATLG_LAUNCHED();
}

template void posv_batched<double>(hipStream_t, int, int, const int*, double* const*, const int*,
                                   double* const*, const int*, const double* const*, int, int*,
                                   const BatchTune&, const int*);
template void posv_batched<float>(hipStream_t, int, int, const int*, float* const*, const int*,
                                  float* const*, const int*, const float* const*, int, int*,
                                  const BatchTune&, const int*);

void dsposv_batched(hipStream_t s, int nsys, const int* n, double* const* A, const int* lda,
                    const double* const* B, const int* ldb, double* const* X, const int* ldx,
                    int nrhs, int* info, int* iter, int itermax, const BatchTune& tune) {
  if (nsys <= 0) return;
  const std::vector<int> ord = lpt_order(nsys, n);
  size_t ws = 0;
  std::vector<size_t> off(nsys);
  for (int q = 0; q < nsys; q++) {
    off[q] = ws;
    const size_t nn = (size_t)n[ord[q]];
    ws += (nn * nn + nn * (size_t)nrhs + 63) & ~(size_t)63;  // floats, 256-byte aligned
  }
  float* W = static_cast<float*>(scratch(s, kSlotBatchWork, std::max<size_t>(ws, 64) * 4));
  std::vector<MixDesc> desc(nsys);
  for (int q = 0; q < nsys; q++) {
    const int k = ord[q];
    const size_t nn = (size_t)n[k];
    desc[q] = MixDesc{A[k], B[k], X[k], W + off[q], W + off[q] + nn * nn, n[k], lda[k], ldb[k],
                      ldx[k], k, width_for(tune, true, n[k])};
  }
  auto* dd = static_cast<const MixDesc*>(
      stage_upload(s, desc.data(), sizeof(MixDesc) * nsys, kSlotBatchDesc));
  unsigned* counter = reset_counter(s);
  hipLaunchKernelGGL(mixed_posv_kernel, dim3(grid_for(s, tune, nsys)), dim3(kBT), 0, s, nsys, dd,
                     nrhs, itermax, info, iter, counter);
  ATLG_LAUNCHED();
  // Systems whose refinement did not converge (or whose FP32 factor failed)
  // get the FP64 factor and solve, as LAPACK's dsposv does; A then holds the
  // FP64 factor and X the FP64 solution.
  std::vector<int> it(nsys);
  ATLG_HIP(hipMemcpyAsync(it.data(), iter, sizeof(int) * nsys, hipMemcpyDeviceToHost, s));
  ATLG_HIP(hipStreamSynchronize(s));
  std::vector<int> fn, flda, fldx, forig;
  std::vector<double*> fA, fX;
  std::vector<const double*> fB;
  for (int k = 0; k < nsys; k++) {
    if (it[k] >= 0) continue;
    fn.push_back(n[k]);
    flda.push_back(lda[k]);
    fldx.push_back(ldx[k]);
    forig.push_back(k);
    fA.push_back(A[k]);
    fX.push_back(X[k]);
    if (ldb[k] == ldx[k]) {
      fB.push_back(B[k]);  // copied into X inside the launch
    } else {
      ATLG_HIP(hipMemcpy2DAsync(X[k], sizeof(double) * ldx[k], B[k], sizeof(double) * ldb[k],
                                sizeof(double) * n[k], nrhs, hipMemcpyDeviceToDevice, s));
      fB.push_back(nullptr);
    }
  }
  if (fn.empty()) return;
  posv_batched<double>(s, 1, (int)fn.size(), fn.data(), fA.data(), flda.data(), fX.data(),
                       fldx.data(), fB.data(), nrhs, info, batch_tune(false), forig.data());
  ATLG_HIP(hipStreamSynchronize(s));
}

}  // namespace atlg
