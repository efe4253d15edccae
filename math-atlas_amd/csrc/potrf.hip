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
constexpr int P_THREADS = kTriThreads;
static_assert(DB == kTriN, "leaf order must match the block inverter");

// Broadcast of lane `src`'s value (uniform src) through the scalar unit.
__device__ __forceinline__ double lane_value(double v, int src) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(u), src);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(u >> 32), src);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// 1/sqrt(d) for d > 0: the hardware estimate and two Newton steps
// (quadratic convergence from its ~2^-23 relative error to below an ulp);
// values below 2^-1000, where the estimate loses range, take sqrt and a
// division.
__device__ __forceinline__ double rsqrt_pos(double d) {
  if (d < 0x1p-1000) return 1.0 / sqrt(d);
  double y = __builtin_amdgcn_rsq(d);
  const double h = 0.5 * d;
  y = y * (1.5 - h * y * y);
  y = y * (1.5 - h * y * y);
  return y;
}

constexpr int SB = 32;  // sub-block order of the in-workgroup factorization
#ifndef ATLG_POTRF_RB
#define ATLG_POTRF_RB 384
#endif
constexpr long RB = ATLG_POTRF_RB;  // block width of the right-looking outer steps

// Factors the n x n (n <= DB) diagonal block at A and writes inv(L) of the
// block to Dinv (DB x DB, column-major, zeros above the diagonal).
// lower: A = L L^T from the lower triangle; else A = U^T U from the upper
// triangle, handled as L = U^T, so one algorithm serves both.  The other
// triangle of A is never read or written.
//
// Factorization (right-looking, 32-wide panels, the block held in LDS;
// order beyond n is padded with the identity), per panel:
//   1. wave 0 holds the panel's 32 x 32 diagonal block and the 32 rows
//      below it in registers, lane i one row, and factors them column by
//      column: the pivot from lane j by v_readlane, multipliers scaled by
//      the reciprocal square root (ATLAS's potrfL closed forms multiply by
//      the reciprocal too), the column's diagonal-block part broadcast to
//      the other lanes through LDS; no barrier inside the panel;
//   2. the remaining rows solve X L11^T = A21 right-looking (one row per
//      thread, the 31 - k updates of step k independent of each other);
//   3. the trailing lower triangle A22 -= X X^T on the matrix cores.
// The factor is then inverted in LDS (lower_inverse_inplace).
__global__ void __launch_bounds__(P_THREADS)
potf2_inv_kernel(int n, double* __restrict__ A, long lda, int lower, int* __restrict__ dinfo,
                 int info_off, double* __restrict__ Dinv) {
  if (*dinfo != 0) return;
  __shared__ double S[DB * LDS_LD];
  __shared__ double LT[SB * SB];   // the panel's diagonal block, transposed
  __shared__ double rinv_sh[SB];   // its reciprocal pivots
  __shared__ double cbuf[64];      // wave 0's column broadcast
  __shared__ int fail_sh;
  const int t = threadIdx.x;
  // fill in batches of 16 elements per thread, all loads of a batch in
  // flight before the LDS stores
  constexpr int PER = DB * DB / P_THREADS, BATCH = 16;
  static_assert(PER % BATCH == 0, "fill batches");
  for (int b0 = 0; b0 < PER; b0 += BATCH) {
    double v[BATCH];
#pragma unroll
    for (int u = 0; u < BATCH; u++) {
      const int idx = t + (b0 + u) * P_THREADS;
      const int i = idx % DB, k = idx / DB;
      double x = 0.0;
      if (i >= k) {
        if (i < n && k < n) x = lower ? A[i + (size_t)k * lda] : A[k + (size_t)i * lda];
        else if (i == k) x = 1.0;
      }
      v[u] = x;
    }
#pragma unroll
    for (int u = 0; u < BATCH; u++) {
      const int idx = t + (b0 + u) * P_THREADS;
      S[(idx % DB) * LDS_LD + idx / DB] = v[u];
    }
  }
  if (t == 0) fail_sh = -1;
  __syncthreads();

  int fail = -1;
  for (int kb = 0; kb < DB; kb += SB) {
    // 1. wave 0: lane i holds panel row kb + i (the 32 x 32 diagonal block
    //    for i < 32, the next 32 rows below it for i >= 32) and factors the
    //    panel right-looking: per column j the pivot comes from lane j
    //    (v_readlane), the multipliers L(i, j) are scaled by the reciprocal
    //    square root, and column j's part inside the diagonal block reaches
    //    the other lanes through a wave-private LDS line (ordered within the
    //    wave: no barrier)
    if (t < 64) {
      const int i = t, r = kb + i;
      const bool have = r < DB;
      double x[SB];
#pragma unroll
      for (int c = 0; c < SB; c++)
        x[c] = (have && (i >= SB || c <= i)) ? S[r * LDS_LD + kb + c] : 0.0;
      // The chain from one pivot to the next is kept short: the next
      // column's update (one FMA with the multiplier of row j+1, taken by
      // v_readlane) and the next pivot's reciprocal square root come first;
      // the other columns' updates read the column from LDS off that path.
      int f = -1;  // first non-positive pivot (uniform)
      double d = lane_value(x[0], 0);
      int bad = __builtin_amdgcn_readfirstlane((int)!(d > 0.0));
      double y = bad ? 0.0 : rsqrt_pos(d);
      double yi = 0.0;  // lane i < 32: its own reciprocal pivot
#pragma unroll
      for (int j = 0; j < SB; j++) {
        if (bad) {
          f = j;
          break;
        }
        const double L = (i == j) ? d * y : x[j] * y;
        if (i >= j) x[j] = L;
        if (i == j) yi = y;
        const double Lm = (i > j) ? L : 0.0;  // lanes <= j: no update
        // column j to LDS and back (all reads in flight together), the
        // next pivot computed while they travel
        cbuf[i] = L;
        double cv[SB];
#pragma unroll
        for (int c = j + 2; c < SB; c++) cv[c] = cbuf[c];
        __builtin_amdgcn_sched_barrier(0);
        if (j + 1 < SB) {
          x[j + 1] -= Lm * lane_value(L, j + 1);
          d = lane_value(x[j + 1], j + 1);
          bad = __builtin_amdgcn_readfirstlane((int)!(d > 0.0));
          y = bad ? 0.0 : rsqrt_pos(d);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = j + 2; c < SB; c++) x[c] -= Lm * cv[c];
      }
      if (have) {
#pragma unroll
        for (int c = 0; c < SB; c++)
          if (i >= SB || c <= i) S[r * LDS_LD + kb + c] = x[c];
      }
      if (i < SB) {  // the diagonal block transposed: LT[k][c] = L(kb+c, kb+k)
        rinv_sh[i] = yi;
#pragma unroll
        for (int k = 0; k < SB; k++) LT[k * SB + i] = x[k];
      }
      if (t == 0) fail_sh = f >= 0 ? kb + f : -1;
    }
    __syncthreads();
    fail = fail_sh;
    if (fail >= 0) break;
    const int r0 = kb + SB, nrow = DB - r0;
    if (nrow <= 0) break;
    // 2. the rows beyond wave 0's (kb + 64 ..): X L11^T = A21 right-looking,
    //    the column updates of one step independent of each other
    if (t >= 64 && kb + t < DB) {
      const int r = kb + t;
      double x[SB];
#pragma unroll
      for (int c = 0; c < SB; c++) x[c] = S[r * LDS_LD + kb + c];
#pragma unroll
      for (int k = 0; k < SB; k++) {
        double cv[SB];  // row k of LT, all reads in flight together
        const double rk = rinv_sh[k];
#pragma unroll
        for (int c = k + 1; c < SB; c++) cv[c] = LT[k * SB + c];
        __builtin_amdgcn_sched_barrier(0);
        x[k] *= rk;
#pragma unroll
        for (int c = k + 1; c < SB; c++) x[c] -= x[k] * cv[c];
      }
#pragma unroll
      for (int c = 0; c < SB; c++) S[r * LDS_LD + kb + c] = x[c];
    }
    __syncthreads();
    // 3. trailing lower triangle A22 -= X X^T on the matrix cores: 16 x 16
    //    tiles (ti >= tj) of the nrow x nrow block, K = 32 in eight steps
    {
      typedef TriMfma<double> TM;
      const int lane = t & 63, wv = t >> 6, lx = lane & 15, lk = lane >> 4;
      const int T = nrow / 16, ntile = T * (T + 1) / 2;
      for (int u = wv; u < ntile; u += P_THREADS / 64) {
        int ti = (int)((sqrtf(8.0f * u + 1.0f) - 1.0f) * 0.5f);
        while (ti * (ti + 1) / 2 > u) ti--;
        while ((ti + 1) * (ti + 2) / 2 <= u) ti++;
        const int tj = u - ti * (ti + 1) / 2;
        const int ra = r0 + 16 * ti + lx, rb = r0 + 16 * tj + lx;
        double av[SB / 4], bv[SB / 4];
#pragma unroll
        for (int q = 0; q < SB / 4; q++) {
          av[q] = S[ra * LDS_LD + kb + 4 * q + lk];
          bv[q] = S[rb * LDS_LD + kb + 4 * q + lk];
        }
        TM::acc_t acc = {0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < SB / 4; q++) acc = TM::op(av[q], bv[q], acc);
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int gi = r0 + 16 * ti + TM::row(lane, e), gk = r0 + 16 * tj + lx;
          if (gi >= gk) S[gi * LDS_LD + gk] -= acc[e];
        }
      }
    }
    __syncthreads();
  }
  if (fail >= 0 && t == 0) *dinfo = info_off + fail + 1;
  // write back (on failure: the partially factored block, as ATLAS leaves it)
  for (int idx = t; idx < DB * DB; idx += P_THREADS) {
    const int i = idx % DB, k = idx / DB;
    if (i >= k && i < n && k < n) {
      if (lower) A[i + (size_t)k * lda] = S[i * LDS_LD + k];
      else A[k + (size_t)i * lda] = S[i * LDS_LD + k];
    }
  }
  if (fail >= 0 || Dinv == nullptr) return;

  // ---- inverse of the factor (upper part of S is zero) ----
  __syncthreads();
  lower_inverse_inplace(S, n);
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
  // Right-looking by RB-column blocks (then ATLAS's halving inside a block):
  // each step's update is one large SYRK instead of the many small GEMMs of
  // a halving recursion over the whole order.
  long nl = (N > 2 * RB) ? RB : ((N / 2) / DB) * DB;
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

// Look-ahead (1) or plain right-looking steps (0).  At N = 16384 the
// look-ahead with the panels on 32 compute units and the SYRKs on the rest
// takes 55 ms against 65 ms for plain steps; on a shared device (no masked
// streams) both took 65 ms: the 128-block leaves, which need 132 KB of LDS,
// wait for a compute unit free of SYRK workgroups.
// Without masked streams: grid cap of the look-ahead's trailing SYRK
// (workgroups).
#ifndef ATLG_CHOL_LOOKAHEAD
#define ATLG_CHOL_LOOKAHEAD 1
#endif
#ifndef ATLG_CHOL_CAP
#define ATLG_CHOL_CAP 384
#endif
// Compute units of the panel stream when the look-ahead runs panels and
// updates on a CU-masked stream pair (0: both share the device).
#ifndef ATLG_CHOL_PANEL_CUS
#define ATLG_CHOL_PANEL_CUS 32
#endif
constexpr long kCholLaMin = 2048;

// Look-ahead step width (ATLG_CHOL_STEP when set, a multiple of 128): the
// chain of diagonal blocks (three 128-block leaves and a panel solve per 384
// columns) is the longer path at small orders, the trailing SYRK at large
// ones.  Measured on MI355X (ms, panels on 32 compute units, panel solves
// moved to the update units where chol_solve_wide says): N=8192 128: 10.9,
// 256: 11.3, 512: 12.5; N=12288 128: 24.2, 256: 21.3, 512: 22.6; N=16384
// 128: 47.7, 256: 39.4, 384: 37.3, 512: 38.1 (before the move, 384: 39.8);
// 48 or 64 panel units were slower at every order.
#ifndef ATLG_CHOL_STEP
#define ATLG_CHOL_STEP 0
#endif
// Widest block (columns) whose look-ahead update goes in two parts, the
// diagonal block's triangle and then the rest of the block column, both by
// 32-row strips (gemm_narrow): the leaf starts after the first part.  N=8192
// (128-column steps): 10.9 -> 10.3 ms; for the 256/384-column steps of
// N=12288/16384 the split, by the 128x128 tiling, was 1 ms slower.
#ifndef ATLG_CHOL_SPLIT_MAX
#define ATLG_CHOL_SPLIT_MAX 128
#endif
inline long chol_step(long N) {
  if (ATLG_CHOL_STEP > 0) return ATLG_CHOL_STEP;
  return N < 10240 ? 128 : (N < 14336 ? 256 : 384);
}

// Whether step k's panel solve (rest x nb1, after the diagonal block of nb1
// columns) goes to the update side.  Per step, the update side runs the
// block column's update and the trailing SYRK (U), the panel side the leaves
// (P) and, when narrow, the solve on its ATLG_CHOL_PANEL_CUS units (Sn); a
// wide solve (Sw) runs after U on the update side instead.  The step takes
// max(U, P + Sn) narrow and max(U, P) + Sw wide, so the solve goes wide when
// U < P + Sn - Sw: the test for blocks of at most ATLG_CHOL_MODEL2_MAX
// columns, with 32 TFLOP/s for their 128-deep updates (measured at
// N = 8192; 26 and 40 were slower).  Wider blocks keep the simpler test
// T < P + Sn (T the trailing SYRK alone) with the 50 TFLOP/s found by a
// sweep at N = 16384 (the full test measured the same there).
#ifndef ATLG_CHOL_WIDE_SOLVE
#define ATLG_CHOL_WIDE_SOLVE 1
#endif
#ifndef ATLG_CHOL_UPD_TFLOPS
#define ATLG_CHOL_UPD_TFLOPS 50.0
#endif
#ifndef ATLG_CHOL_UPD128_TFLOPS
#define ATLG_CHOL_UPD128_TFLOPS 32.0
#endif
#ifndef ATLG_CHOL_MODEL2_MAX
#define ATLG_CHOL_MODEL2_MAX 128
#endif
inline bool chol_solve_wide(long rest, long nb1) {
  if (!ATLG_CHOL_WIDE_SOLVE || rest <= 0) return false;
  const double upd_rate = (nb1 <= 128 ? ATLG_CHOL_UPD128_TFLOPS : ATLG_CHOL_UPD_TFLOPS) * 1e12,
               panel_rate = upd_rate * ATLG_CHOL_PANEL_CUS / (256.0 - ATLG_CHOL_PANEL_CUS);
  const double r = (double)rest, b = (double)nb1;
  const double t_solve_n = 2.0 * r * b * b / panel_rate, t_solve_w = 2.0 * r * b * b / upd_rate;
  if (nb1 > ATLG_CHOL_MODEL2_MAX) return r * r * b / upd_rate < b / 128.0 * 95e-6 + t_solve_n;
  const double t_upd = r * r * b / upd_rate + 2.0 * r * b * b / upd_rate;
  return t_upd < b / 128.0 * 95e-6 + t_solve_n - t_solve_w;
}

// Right-looking Cholesky by chol_step(N)-column blocks with one step of
// look-ahead:
// while the caller's stream applies step k's SYRK to the columns right of
// block k+1, the auxiliary stream factors block k+1 (updated first) and
// solves its panel.  lower: A = L L^T, panel L21 = A21 L11^-T, update
// A22 -= L21 L21^T; upper: A = U^T U, U12 = U11^-T A12, A22 -= U12^T U12.
void potrf_lookahead(hipStream_t s0, bool lower, long N, double* A, long lda, int* dinfo,
                     double* W) {
  const auto lock = aux_lock();
  hipStream_t s1 = aux_stream();
  hipEvent_t e_upd = aux_event(0), e_panel = aux_event(1);
  // with a CU-masked pair the 128-block leaves (132 KB of LDS) never wait
  // for a compute unit free of SYRK workgroups
  hipStream_t caller = s0, sp = nullptr, sg = nullptr;
  const bool split = ATLG_CHOL_PANEL_CUS > 0 && split_streams(ATLG_CHOL_PANEL_CUS, &sp, &sg);
  if (split) {
    ATLG_HIP(hipEventRecord(aux_event(2), caller));
    ATLG_HIP(hipStreamWaitEvent(sg, aux_event(2), 0));
    ATLG_HIP(hipStreamWaitEvent(sp, aux_event(2), 0));
    s0 = sg;
    s1 = sp;
  }
  auto at = [&](long i, long j) { return A + i + j * lda; };
  // the panel of the block at k (kb columns) below/right of it, rows/cols r
  auto panel_ptr = [&](long k, long kb) { return lower ? at(k + kb, k) : at(k, k + kb); };
  auto solve_panel = [&](hipStream_t s, long k, long kb, long r) {
    if (r <= 0) return;
    const double* Winv = W + (size_t)(k / DB) * DB * DB;
    if (lower)
      trsm(s, kRight, kLower, kTrans, kNonUnit, r, kb, 1.0, at(k, k), lda, panel_ptr(k, kb), lda,
           dinfo, Winv, 1);
    else
      trsm(s, kLeft, kUpper, kTrans, kNonUnit, kb, r, 1.0, at(k, k), lda, panel_ptr(k, kb), lda,
           dinfo, Winv, 0);
  };
  // A(c0.., c0..c0+nc) -= P_c P_c^T restricted to the stored triangle, P the
  // panel of block k (kb wide) from its row/column c0 on; m rows/cols in all
  auto update = [&](hipStream_t s, long k, long kb, long c0, long m, long nc, int cap) {
    if (m <= 0 || nc <= 0) return;
    if (lower) {
      const double* P = at(c0, k);
      gemm(s, kNoTrans, kTrans, m, nc, kb, -1.0, P, lda, P, lda, 1.0, at(c0, c0), lda, 1, dinfo,
           cap);
    } else {
      const double* P = at(k, c0);
      gemm(s, kTrans, kNoTrans, nc, m, kb, -1.0, P, lda, P, lda, 1.0, at(c0, c0), lda, 2, dinfo,
           cap);
    }
  };
  // the off-diagonal part of block column c0 (nc wide, r rows/columns past
  // it) minus the product of block k's panel rows: narrow blocks by strips
  auto update_col = [&](hipStream_t s, long k, long kb, long c0, long nc, long r) {
    if (r <= 0) return;
    if (lower) {
      if (!gemm_narrow(s, kNoTrans, kTrans, r, nc, kb, -1.0, at(c0 + nc, k), lda, at(c0, k), lda,
                       1.0, at(c0 + nc, c0), lda, 0, dinfo))
        gemm(s, kNoTrans, kTrans, r, nc, kb, -1.0, at(c0 + nc, k), lda, at(c0, k), lda, 1.0,
             at(c0 + nc, c0), lda, 0, dinfo);
    } else {
      if (!gemm_narrow(s, kTrans, kNoTrans, nc, r, kb, -1.0, at(k, c0), lda, at(k, c0 + nc), lda,
                       1.0, at(c0, c0 + nc), lda, 0, dinfo))
        gemm(s, kTrans, kNoTrans, nc, r, kb, -1.0, at(k, c0), lda, at(k, c0 + nc), lda, 1.0,
             at(c0, c0 + nc), lda, 0, dinfo);
    }
  };
  // the stored triangle of diagonal block c0 (nc wide) likewise
  auto update_diag = [&](hipStream_t s, long k, long kb, long c0, long nc) {
    const double* P = lower ? at(c0, k) : at(k, c0);
    if (lower ? gemm_narrow(s, kNoTrans, kTrans, nc, nc, kb, -1.0, P, lda, P, lda, 1.0,
                            at(c0, c0), lda, 1, dinfo)
              : gemm_narrow(s, kTrans, kNoTrans, nc, nc, kb, -1.0, P, lda, P, lda, 1.0,
                            at(c0, c0), lda, 2, dinfo))
      return;
    update(s, k, kb, c0, nc, nc, 0);
  };
  hipEvent_t e_col = aux_event(4);
  const long rb = chol_step(N);
  const long kb0 = std::min(rb, N);
  potrf_rec(s0, lower, kb0, A, lda, dinfo, 0, W);
  solve_panel(s0, 0, kb0, N - kb0);
  for (long k0 = 0; k0 + kb0 < N; k0 += rb) {
    const long kb = std::min(rb, N - k0), k1 = k0 + kb, nb1 = std::min(rb, N - k1);
    const long rest = N - k1 - nb1;
    // block k+1 first, then the rest of step k's update, beside which the
    // auxiliary stream factors block k+1 and solves its panel; for narrow
    // blocks its diagonal block alone goes first (all the leaf needs), then
    // the rest of its block column (what the panel solve needs)
    const bool diag_first = nb1 <= ATLG_CHOL_SPLIT_MAX;
    if (diag_first) {
      update_diag(s0, k0, kb, k1, nb1);
      ATLG_HIP(hipEventRecord(e_upd, s0));
      update_col(s0, k0, kb, k1, nb1, rest);
      ATLG_HIP(hipEventRecord(e_col, s0));
    } else {
      update(s0, k0, kb, k1, N - k1, nb1, 0);
      ATLG_HIP(hipEventRecord(e_upd, s0));
    }
    update(s0, k0, kb, k1 + nb1, rest, rest, split ? 0 : ATLG_CHOL_CAP);
    ATLG_HIP(hipStreamWaitEvent(s1, e_upd, 0));
    potrf_rec(s1, lower, nb1, at(k1, k1), lda, dinfo, (int)k1, W);
    // The panel solve is a GEMM-shaped TRSM of rest x nb1: on the panel
    // units while the trailing update keeps the others busy, on the update
    // side once that update is the shorter path (its units would wait).
    const bool solve_wide = split && chol_solve_wide(rest, nb1);
    if (!solve_wide) {
      if (diag_first) ATLG_HIP(hipStreamWaitEvent(s1, e_col, 0));
      solve_panel(s1, k1, nb1, rest);
    }
    ATLG_HIP(hipEventRecord(e_panel, s1));
    ATLG_HIP(hipStreamWaitEvent(s0, e_panel, 0));
    if (solve_wide) solve_panel(s0, k1, nb1, rest);
  }
  if (split) {  // every panel was joined into s0 above
    ATLG_HIP(hipEventRecord(aux_event(3), s0));
    ATLG_HIP(hipStreamWaitEvent(caller, aux_event(3), 0));
  }
}

}  // namespace

void potrf(hipStream_t s, int uplo, long N, double* A, long lda, int* dinfo) {
  set_int(s, dinfo, 0);
  if (N <= 0) return;
  const long nblk = (N + DB - 1) / DB;
  double* W = static_cast<double*>(scratch(s, kSlotFactInv, sizeof(double) * nblk * DB * DB));
  if (ATLG_CHOL_LOOKAHEAD && N >= kCholLaMin)
    potrf_lookahead(s, uplo == kLower, N, A, lda, dinfo, W);
  else
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
