// LU with partial pivoting (column-major) and the solves built on it.
//
// Replaces ATL_dgetrfC (AtlasBase/Clint/atlas-lp.base:1497-1618), its panel
// routines ATL_dgetf2 (:5204-5390), LU2 (:1301-1426) and the MN==1 case
// (:1581-1600), ATL_dlaswp (:1053-1166) and ATL_dgetrs (:1631-1720).
//
// The recursion is ATLAS's: split the columns in half, factor the left half,
// swap+TRSM+GEMM the right half, factor it, and back-apply its interchanges
// to the left half.  The recursion stops at panels of at most PW columns,
// which are factored by
//   * panel_regs: one workgroup, one panel row per thread held in registers,
//     pivot search as a wavefront argmax over cross-lane shuffles + LDS,
//     used when the panel has at most 512 rows;
//   * panel_step: one launch per column for taller panels; each launch
//     applies the previous column's scaling and rank-1 update to its rows,
//     finds its rows' candidate pivot, and the last workgroup to finish
//     reduces the candidates, swaps the two panel rows and publishes the
//     pivot.
// Pivot choice follows cblas_idamax (AtlasBase/Antoine/atlrefblas.base:753+):
// first index of the largest |a|, NaN never chosen.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>

#include "atlg_internal.hpp"

namespace atlg {
namespace {

constexpr int PW = 32;           // leaf panel width
constexpr int SMALL_ROWS = 512;  // panel_regs limit (one row per thread)
constexpr int STEP_THREADS = 256;
constexpr int KC = 256;          // interchanges composed per laswp chunk
constexpr int CW = 16;           // columns per laswp workgroup

// ---------------------------------------------------------------- helpers

struct Cand {
  double key;  // |a|, or -1 for NaN / out of range
  int idx;
};

__device__ inline double pkey(double v) { return (v != v) ? -1.0 : fabs(v); }

__device__ inline bool better(double k1, int i1, double k2, int i2) {
  return k1 > k2 || (k1 == k2 && i1 < i2);
}

// wavefront argmax: after the call every lane holds the wave's best pair
__device__ inline void wave_argmax(double& k, int& i) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double k2 = __shfl_xor(k, off, 64);
    const int i2 = __shfl_xor(i, off, 64);
    if (better(k2, i2, k, i)) {
      k = k2;
      i = i2;
    }
  }
}

// workgroup argmax (blockDim multiple of 64, <= 1024); result in every thread
__device__ inline void block_argmax(double& k, int& i, double* sk, int* si) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  wave_argmax(k, i);
  if (lane == 0) {
    sk[w] = k;
    si[w] = i;
  }
  __syncthreads();
  if (w == 0) {
    double kk = lane < nw ? sk[lane] : -2.0;
    int ii = lane < nw ? si[lane] : INT_MAX;
    wave_argmax(kk, ii);
    if (lane == 0) {
      sk[0] = kk;
      si[0] = ii;
    }
  }
  __syncthreads();
  k = sk[0];
  i = si[0];
  __syncthreads();
}

// ---------------------------------------------------------------- panels

// Factors an M x w panel (M <= SMALL_ROWS, w <= PW) held one row per thread.
__global__ void __launch_bounds__(SMALL_ROWS)
panel_regs(int M, int w, double* __restrict__ A, long lda, int* __restrict__ ipiv,
           int* __restrict__ dinfo, int info_off) {
  __shared__ double prow[PW], jrow[PW];
  __shared__ double sk[16];
  __shared__ int si[16];
  const int r = threadIdx.x;
  double v[PW];
#pragma unroll
  for (int c = 0; c < PW; c++) v[c] = (r < M && c < w) ? A[r + (size_t)c * lda] : 0.0;

  const int mn = min(M, w);
  for (int j = 0; j < mn; j++) {
    double key = -1.0;
#pragma unroll
    for (int c = 0; c < PW; c++)
      if (c == j) key = (r >= j && r < M) ? pkey(v[c]) : -2.0;
    int idx = (r >= j && r < M) ? r : INT_MAX;
    block_argmax(key, idx, sk, si);
    const int p = idx;
    if (r == p) {
#pragma unroll
      for (int c = 0; c < PW; c++) prow[c] = v[c];
    }
    if (r == j) {
#pragma unroll
      for (int c = 0; c < PW; c++) jrow[c] = v[c];
    }
    __syncthreads();
    double piv = 0.0;
#pragma unroll
    for (int c = 0; c < PW; c++)
      if (c == j) piv = prow[c];
    if (r == 0) {
      ipiv[j] = p;
      if (piv == 0.0 && *dinfo == 0) *dinfo = info_off + j + 1;
    }
    if (piv != 0.0) {
      if (r == j && p != j) {
#pragma unroll
        for (int c = 0; c < PW; c++) v[c] = prow[c];
      } else if (r == p && p != j) {
#pragma unroll
        for (int c = 0; c < PW; c++) v[c] = jrow[c];
      }
      if (r > j && r < M) {
        const bool recip = fabs(piv) >= DBL_MIN;
        const double rinv = 1.0 / piv;
        double l = 0.0;
#pragma unroll
        for (int c = 0; c < PW; c++)
          if (c == j) {
            v[c] = recip ? v[c] * rinv : v[c] / piv;
            l = v[c];
          }
#pragma unroll
        for (int c = 0; c < PW; c++)
          if (c > j && c < w) v[c] -= l * prow[c];
      }
    }
    __syncthreads();
  }
  if (r < M) {
#pragma unroll
    for (int c = 0; c < PW; c++)
      if (c < w) A[r + (size_t)c * lda] = v[c];
  }
}

// Pivot record published by the last workgroup of a column step.
struct PivRec {
  double piv;
  int p;
  int pad;
};

// One column step of a tall panel.  Launch j (0 <= j <= mn) first applies
// column j-1's scaling and rank-1 update to rows [j, M), then (j < mn) finds
// the pivot of column j.  `scr` holds per-workgroup candidates, the ticket
// counter and the pivot record.
__global__ void __launch_bounds__(STEP_THREADS)
panel_step(int M, int w, int j, int mn, double* __restrict__ A, long lda,
           int* __restrict__ ipiv, int* __restrict__ dinfo, int info_off,
           Cand* __restrict__ cand, unsigned* __restrict__ ticket, PivRec* __restrict__ rec) {
  __shared__ double urow[PW];
  __shared__ double sk[4];
  __shared__ int si[4];
  __shared__ bool last;
  const int t = threadIdx.x;
  const int r = j + blockIdx.x * STEP_THREADS + t;  // this thread's row
  const bool mine = r < M;

  if (j > 0) {
    const PivRec pr = rec[j - 1];
    if (t < w) urow[t] = A[(j - 1) + (size_t)t * lda];
    __syncthreads();
    if (mine && pr.piv != 0.0) {
      double* a = A + r;
      double l = a[(size_t)(j - 1) * lda];
      l = (fabs(pr.piv) >= DBL_MIN) ? l * (1.0 / pr.piv) : l / pr.piv;
      a[(size_t)(j - 1) * lda] = l;
      for (int c = j; c < w; c++) a[(size_t)c * lda] -= l * urow[c];
    }
  }
  if (j >= mn) return;

  double key = mine ? pkey(A[r + (size_t)j * lda]) : -2.0;
  int idx = mine ? r : INT_MAX;
  block_argmax(key, idx, sk, si);
  if (t == 0) {
    cand[blockIdx.x].key = key;
    cand[blockIdx.x].idx = idx;
    __threadfence();
    const unsigned tk = atomicAdd(ticket, 1u);
    last = (tk == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;

  // last workgroup: reduce the candidates, swap rows j and p, publish
  __threadfence();
  double k2 = -2.0;
  int i2 = INT_MAX;
  for (int b = t; b < (int)gridDim.x; b += STEP_THREADS) {
    const volatile Cand* cv = cand + b;
    const double kb = cv->key;
    const int ib = cv->idx;
    if (better(kb, ib, k2, i2)) {
      k2 = kb;
      i2 = ib;
    }
  }
  block_argmax(k2, i2, sk, si);
  const int p = i2;
  const double piv = A[p + (size_t)j * lda];
  if (piv != 0.0 && p != j && t < w) {
    double* a = A + (size_t)t * lda;
    const double x = a[j];
    a[j] = a[p];
    a[p] = x;
  }
  if (t == 0) {
    ipiv[j] = p;
    if (piv == 0.0 && *dinfo == 0) *dinfo = info_off + j + 1;
    rec[j].piv = piv;
    rec[j].p = p;
    *ticket = 0u;
  }
}

// ---------------------------------------------------------------- laswp

// Composes interchanges ipiv[i] for i in [k1, k2) (ascending if inci > 0,
// descending otherwise; at most KC of them) into the list of rows whose
// content changes: after the swaps, row dst[t] holds the original row
// src[t].  One workgroup; each thread owns one step.
//   A_s: content at a_s = i_s after step s;  B_s: content at b_s = ipiv[i_s].
//   A_s = content before step s at b_s,  B_s = content before s at a_s,
// where "content before step s at x" is the value left by the latest earlier
// step touching x, or x itself.  References are resolved by pointer jumping.
__global__ void __launch_bounds__(KC)
laswp_compose(const int* __restrict__ ipiv, int k1, int k2, int inci, int* __restrict__ out) {
  __shared__ int pa[KC], pb[KC];   // positions of step s
  __shared__ int val[2 * KC];      // >= 0: original row; < 0: -(ref+1)
  __shared__ int cnt;
  const int s = threadIdx.x;
  const int k = k2 - k1;
  if (s == 0) cnt = 0;
  if (s < k) {
    const int i = inci > 0 ? k1 + s : k2 - 1 - s;
    pa[s] = i;
    pb[s] = ipiv[i];
  }
  __syncthreads();
  if (s < k) {
    const int a = pa[s], b = pb[s];
    int ra = -1, rb = -1;  // latest earlier refs for b (-> A_s) and a (-> B_s)
    for (int q = s - 1; q >= 0 && (ra < 0 || rb < 0); q--) {
      if (ra < 0) {
        if (pa[q] == b) ra = 2 * q;
        else if (pb[q] == b) ra = 2 * q + 1;
      }
      if (rb < 0) {
        if (pa[q] == a) rb = 2 * q;
        else if (pb[q] == a) rb = 2 * q + 1;
      }
    }
    val[2 * s] = ra >= 0 ? -(ra + 1) : b;
    val[2 * s + 1] = rb >= 0 ? -(rb + 1) : a;
  }
  __syncthreads();
  for (int round = 0; round < 10; round++) {  // 2^10 > 2*KC
    int v0 = 0, v1 = 0;
    if (s < k) {
      v0 = val[2 * s];
      v1 = val[2 * s + 1];
      if (v0 < 0) v0 = val[-v0 - 1];
      if (v1 < 0) v1 = val[-v1 - 1];
    }
    __syncthreads();
    if (s < k) {
      val[2 * s] = v0;
      val[2 * s + 1] = v1;
    }
    __syncthreads();
  }
  if (s < k) {
    const int a = pa[s], b = pb[s];
    bool la = true, lb = (a != b);
    for (int q = s + 1; q < k; q++) {
      if (pa[q] == a || pb[q] == a) la = false;
      if (pb[q] == b || pa[q] == b) lb = false;
    }
    if (la && val[2 * s] != a) {
      const int t = atomicAdd(&cnt, 1);
      out[1 + 2 * t] = a;
      out[2 + 2 * t] = val[2 * s];
    }
    if (lb && val[2 * s + 1] != b) {
      const int t = atomicAdd(&cnt, 1);
      out[1 + 2 * t] = b;
      out[2 + 2 * t] = val[2 * s + 1];
    }
  }
  __syncthreads();
  if (s == 0) out[0] = cnt;
}

// Applies a composed list to CW columns per workgroup: gather the source rows
// into LDS, then scatter them to their destinations.
__global__ void __launch_bounds__(256)
laswp_apply(long N, double* __restrict__ A, long lda, const int* __restrict__ lst) {
  __shared__ double buf[2 * KC * CW];
  const int T = lst[0];
  const long c0 = (long)blockIdx.x * CW;
  const int nc = (int)min((long)CW, N - c0);
  for (int e = threadIdx.x; e < T * CW; e += blockDim.x) {
    const int t = e / CW, c = e % CW;
    if (c < nc) buf[e] = A[(size_t)lst[2 + 2 * t] + (size_t)(c0 + c) * lda];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < T * CW; e += blockDim.x) {
    const int t = e / CW, c = e % CW;
    if (c < nc) A[(size_t)lst[1 + 2 * t] + (size_t)(c0 + c) * lda] = buf[e];
  }
}

__global__ void ipiv_add_kernel(int* ipiv, int k1, int k2, int off) {
  const int i = k1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k2) ipiv[i] += off;
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

// ---------------------------------------------------------------- drivers

struct PanelScratch {
  Cand* cand;
  unsigned* ticket;
  PivRec* rec;
};

void panel(hipStream_t s, long M, int w, double* A, long lda, int* ipiv, int* dinfo,
           int info_off, const PanelScratch& ps) {
  const int mn = (int)std::min<long>(M, w);
  if (M <= SMALL_ROWS) {
    hipLaunchKernelGGL(panel_regs, dim3(1), dim3(SMALL_ROWS), 0, s, (int)M, w, A, lda, ipiv,
                       dinfo, info_off);
    ATLG_LAUNCHED();
    return;
  }
  for (int j = 0; j <= mn; j++) {
    const long rows = M - j;
    const unsigned nb = (unsigned)std::max<long>(1, (rows + STEP_THREADS - 1) / STEP_THREADS);
    hipLaunchKernelGGL(panel_step, dim3(nb), dim3(STEP_THREADS), 0, s, (int)M, w, j, mn, A, lda,
                       ipiv, dinfo, info_off, ps.cand, ps.ticket, ps.rec);
    ATLG_LAUNCHED();
  }
}

void getrf_rec(hipStream_t s, long M, long N, double* A, long lda, int* ipiv, int* dinfo,
               int info_off, const PanelScratch& ps) {
  const long MN = std::min(M, N);
  if (MN == 0) return;
  if (N <= PW) {
    panel(s, M, (int)N, A, lda, ipiv, dinfo, info_off, ps);
    return;
  }
  if (MN <= PW) {  // short and wide: factor the square part, fix up the rest
    panel(s, M, (int)MN, A, lda, ipiv, dinfo, info_off, ps);
    laswp(s, N - MN, A + MN * lda, lda, 0, (int)MN, ipiv, 1);
    trsm(s, kLeft, kLower, kNoTrans, kUnit, MN, N - MN, 1.0, A, lda, A + MN * lda, lda);
    return;
  }
  long nl = ((MN / 2) / PW) * PW;
  if (nl < PW) nl = PW;
  const long nr = N - nl;
  getrf_rec(s, M, nl, A, lda, ipiv, dinfo, info_off, ps);
  double* Ac = A + nl * lda;
  laswp(s, nr, Ac, lda, 0, (int)nl, ipiv, 1);
  trsm(s, kLeft, kLower, kNoTrans, kUnit, nl, nr, 1.0, A, lda, Ac, lda);
  gemm(s, kNoTrans, kNoTrans, M - nl, nr, nl, -1.0, A + nl, lda, Ac, lda, 1.0, Ac + nl, lda);
  getrf_rec(s, M - nl, nr, Ac + nl, lda, ipiv + nl, dinfo, info_off + (int)nl, ps);
  ipiv_add(s, ipiv, (int)nl, (int)MN, (int)nl);
  laswp(s, nl, A, lda, (int)nl, (int)MN, ipiv, 1);
}

}  // namespace

void set_int(hipStream_t s, int* p, int v) {
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, p, v);
  ATLG_LAUNCHED();
}

void ipiv_add(hipStream_t s, int* ipiv, int k1, int k2, int off) {
  if (k2 <= k1) return;
  const int n = k2 - k1;
  hipLaunchKernelGGL(ipiv_add_kernel, dim3((n + 255) / 256), dim3(256), 0, s, ipiv, k1, k2, off);
  ATLG_LAUNCHED();
}

void laswp(hipStream_t s, long N, double* A, long lda, int K1, int K2, const int* ipiv,
           int inci) {
  if (N <= 0 || K2 <= K1) return;
  int* lst = static_cast<int*>(scratch(s, kSlotLaswp, sizeof(int) * (1 + 4 * KC)));
  const unsigned nb = (unsigned)((N + CW - 1) / CW);
  const int nchunk = (K2 - K1 + KC - 1) / KC;
  for (int c = 0; c < nchunk; c++) {
    int a, b;
    if (inci > 0) {
      a = K1 + c * KC;
      b = std::min(K2, a + KC);
    } else {
      b = K2 - c * KC;
      a = std::max(K1, b - KC);
    }
    hipLaunchKernelGGL(laswp_compose, dim3(1), dim3(KC), 0, s, ipiv, a, b, inci, lst);
    ATLG_LAUNCHED();
    hipLaunchKernelGGL(laswp_apply, dim3(nb), dim3(256), 0, s, N, A, lda, lst);
    ATLG_LAUNCHED();
  }
}

void getrf(hipStream_t s, long M, long N, double* A, long lda, int* ipiv, int* dinfo) {
  set_int(s, dinfo, 0);
  if (M == 0 || N == 0) return;
  const long maxblk = (M + STEP_THREADS - 1) / STEP_THREADS + 1;
  char* base = static_cast<char*>(
      scratch(s, kSlotPanel, 256 + sizeof(Cand) * maxblk + sizeof(PivRec) * PW));
  PanelScratch ps;
  ps.ticket = reinterpret_cast<unsigned*>(base);
  ps.rec = reinterpret_cast<PivRec*>(base + 64);
  ps.cand = reinterpret_cast<Cand*>(base + 64 + sizeof(PivRec) * PW + 64);
  set_int(s, reinterpret_cast<int*>(ps.ticket), 0);
  getrf_rec(s, M, N, A, lda, ipiv, dinfo, 0, ps);
}

void getrs(hipStream_t s, int trans, long N, long NRHS, const double* A, long lda,
           const int* ipiv, double* B, long ldb) {
  if (N == 0 || NRHS == 0) return;
  if (!is_trans(trans)) {
    laswp(s, NRHS, B, ldb, 0, (int)N, ipiv, 1);
    trsm(s, kLeft, kLower, kNoTrans, kUnit, N, NRHS, 1.0, A, lda, B, ldb);
    trsm(s, kLeft, kUpper, kNoTrans, kNonUnit, N, NRHS, 1.0, A, lda, B, ldb);
  } else {
    trsm(s, kLeft, kUpper, kTrans, kNonUnit, N, NRHS, 1.0, A, lda, B, ldb);
    trsm(s, kLeft, kLower, kTrans, kUnit, N, NRHS, 1.0, A, lda, B, ldb);
    laswp(s, NRHS, B, ldb, 0, (int)N, ipiv, -1);
  }
}

}  // namespace atlg
