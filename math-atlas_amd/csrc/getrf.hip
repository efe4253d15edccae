// LU with partial pivoting (column-major) and the solves built on it.
//
// Replaces ATL_dgetrfC (AtlasBase/Clint/atlas-lp.base:1497-1618), its panel
// routines ATL_dgetf2 (:5204-5390), LU2 (:1301-1426) and the MN==1 case
// (:1581-1600), ATL_dlaswp (:1053-1166) and ATL_dgetrs (:1631-1720).
//
// The recursion is ATLAS's: split the columns, factor the left part, TRSM +
// GEMM the right part, factor it.  It stops at panels of at most PW columns
// ("leaves").  Two things differ from a line-by-line reading of getrfC:
//
//  * Row interchanges are applied by each leaf, at once, to every column of
//    the matrix outside the leaf, instead of by the laswp calls of every
//    recursion level.  This gives the same matrix: the columns to the right
//    have not been touched yet when a leaf runs, the columns to the left are
//    final L columns, and the swaps still happen in pivot order.  A leaf
//    composes its swaps into one "row i receives row j" list, so a leaf's
//    interchanges cost one gather/scatter pass over the other columns.
//
//  * A leaf panel is factored by one cooperative launch (panel_coop): the
//    panel rows live in the registers of G <= 64 workgroups; per column each
//    workgroup finds its candidate pivot (wavefront argmax over cross-lane
//    shuffles, then LDS) and publishes it with the candidate row into the
//    column's write-once mailbox; every workgroup polls the mailbox slots
//    until all are written (one round trip, no barrier), picks the same
//    global pivot, swaps, scales and applies the rank-1 update to its rows.
//    The waits are bounded; on overrun every workgroup leaves, the panel's
//    pivots are written as the identity, *dinfo = INT_MIN and the stream's
//    failure word gets ATLG_FAIL_PANEL (checked by the host entry points).
//
// Pivot choice follows cblas_idamax (AtlasBase/Antoine/atlrefblas.base:753+):
// first index of the largest |a|, NaN never chosen; the column is scaled by
// the reciprocal pivot when |pivot| >= SAFMIN, else divided.
//
// Everything is a template on the scalar type: double for ATL_dgetrf, float
// for the single-precision factor of the mixed-precision solver (the
// ATL_sgetrf instance of the same ATLAS source).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <limits>
#include <type_traits>
#include <vector>

#include "atlg_internal.hpp"
#include "tri_inv.hpp"

namespace atlg {
namespace {

// Leaf panel width (32 or 64): wider leaves mean fewer recursion nodes
// between leaves, at more registers and exchange volume per column.
#ifndef ATLG_PANEL_PW
#define ATLG_PANEL_PW 32
#endif
constexpr int PW = ATLG_PANEL_PW;  // leaf panel width
static_assert(PW == 32 || PW == 64, "leaf panel width");
constexpr int CP_THREADS = 256;   // workgroup of the cooperative panel
// Most workgroups of the cooperative panel (64 or 128): one panel row per
// thread up to CP_MAXG * 256 rows, more rows per thread beyond.
#ifndef ATLG_PANEL_MAXG
#define ATLG_PANEL_MAXG 64
#endif
constexpr int CP_MAXG = ATLG_PANEL_MAXG;
static_assert(CP_MAXG == 64 || CP_MAXG == 128, "panel workgroups");
constexpr int STEP_THREADS = 256;
constexpr int KC = 256;           // interchanges composed per laswp chunk
constexpr int CW = 16;            // columns per laswp workgroup
constexpr int kLeafListInts = 1 + 4 * PW;  // one leaf's composed list: count + pairs
constexpr int kJrowT = 128;  // first thread that gathers row j in the panel exchange
constexpr long kSpinLimit = 1L << 20;  // default mailbox polls (each one round trip) before giving up
// Fewest panel rows per thread of the cooperative panel (1, 2 or 4): more
// rows per thread means fewer workgroups exchanging candidates per column.
#ifndef ATLG_PANEL_RPT_MIN
#define ATLG_PANEL_RPT_MIN 1
#endif
// Look-ahead driver (1) or ATLAS's recursion over the whole matrix (0), and
// its block-column width and the smallest min(M, N) that uses it.
#ifndef ATLG_LU_LOOKAHEAD
#define ATLG_LU_LOOKAHEAD 1
#endif
#ifndef ATLG_LA_COOP
#define ATLG_LA_COOP 0
#endif
// Wave priority of the panel kernel (0-3): above the GEMM waves it shares
// SIMDs with while the look-ahead overlaps them.
#ifndef ATLG_PANEL_PRIO
#define ATLG_PANEL_PRIO 3
#endif
// Compute units reserved for the look-ahead panels (0: panels and updates
// share the device).  64 units give a 64-workgroup panel grid one unit per
// workgroup; a grid that cannot all be resident ends at the mailbox spin
// limit with an error rather than waiting.
#ifndef ATLG_LA_PANEL_CUS
#define ATLG_LA_PANEL_CUS 64
#endif
// Without reserved units: grid cap of the look-ahead trailing update
// (workgroups; 0 = no cap), applied while the update is smaller than
// ATLG_LA_CAP_AREA elements: then the panel beside it is the longer path and
// keeps compute units; a larger update is the longer path itself and gets
// the whole device.
#ifndef ATLG_LA_GEMM_CAP
#define ATLG_LA_GEMM_CAP 384
#endif
#ifndef ATLG_LA_CAP_AREA
#define ATLG_LA_CAP_AREA (16384L * 16384L)
#endif
// Look-ahead block width: ATLG_LA_NB when set (a multiple of 128), else
// about min(M,N)/32 rounded down to 128, within [512, 1024]: the panel
// chain's cost per column does not depend on the width, while wider blocks
// make deeper (more efficient) trailing updates.  Measured: N=16384 512
// (157 ms; 768: 159), N=24576 768 (339 ms; 512: 346), N=32768 1024 (665 ms;
// 512: 698, 768: 678).
#ifndef ATLG_LA_NB
#define ATLG_LA_NB 0
#endif
inline long la_block(long mn) {
  if (ATLG_LA_NB > 0) return ATLG_LA_NB;
  return std::min(1024L, std::max(512L, mn / 32 / 128 * 128));
}
constexpr long kLaMin = 2048;
// Split column of a look-ahead block of width nb (0: not split), as the
// recursion would choose it.
#ifndef ATLG_LA_SPLIT
#define ATLG_LA_SPLIT 1
#endif
// Share of step k's trailing update that the panel units take on after the
// next panel, while the update is the longer path (FP64; the low-precision
// updates are fast enough that the panel always is).  Balance of the two
// sides: the update side's rate ATLG_LA_UPD_TFLOPS on 192 units against a
// third of it on the 64 panel units, and ATLG_LA_PANEL_US per panel column.
// Measured (DGETRF, uniform [-1,1], ms at N=16384 / 32768): rate 40:
// 122.1 / 497; 55: 123.6 / 504; 70: 124.4 / 513; 30: 123.7 / 517; 20: 130 /
// 539; no balancing: 124 / 516 (the panel-time constant 5 instead of 7
// with rate 55: 122.7 / 501).
#ifndef ATLG_LA_BALANCE
#define ATLG_LA_BALANCE 1
#endif
#ifndef ATLG_LA_UPD_TFLOPS
#define ATLG_LA_UPD_TFLOPS 40.0
#endif
#ifndef ATLG_LA_PANEL_US
#define ATLG_LA_PANEL_US 7.0
#endif
template <class S>
long la_chunk(long rest, long rows, long kb, long nb1, long nb2) {
  if (!ATLG_LA_BALANCE || !std::is_same<S, double>::value || rows <= 0 || nb1 <= 0) return 0;
  const double panel_s = ATLG_LA_PANEL_US * 1e-6 * (double)nb1;
  const double per_col = 2.0 * (double)rows * (double)kb;  // flops per trailing column
  const double c = ((double)rest - panel_s * ATLG_LA_UPD_TFLOPS * 1e12 / per_col) / 4.0;
  long ci = c > 0 ? ((long)c / 128) * 128 : 0;
  ci = std::min(ci, rest - nb2);  // never the next block's columns
  return ci >= 256 ? ci : 0;
}
inline long la_split(long nb) {
  if (!ATLG_LA_SPLIT || nb < 2 * kTrsmNB) return 0;
  return ((nb / 2) / kTrsmNB) * kTrsmNB;
}

// ---------------------------------------------------------------- helpers

struct Cand {
  double key;  // |a|, -1 for NaN, -2 for "no row"
  int idx;
  int pad;
};

template <class S>
__device__ inline double pkey(S v) { return (v != v) ? -1.0 : fabs((double)v); }

__device__ inline bool better(double k1, int i1, double k2, int i2) {
  return k1 > k2 || (k1 == k2 && i1 < i2);
}

// Cross-lane moves of the wavefront argmax: DPP within each 16-lane row
// (quad_perm xor 1 and xor 2, then the half-row and row mirrors: after the
// four steps every lane holds its row's best pair), v_readlane across the
// four rows.  All 64 lanes must be active.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long u = __double_as_longlong(x);
  const int lo = dpp_i<CTRL>(static_cast<int>(u));
  const int hi = dpp_i<CTRL>(static_cast<int>(u >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long u = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(u), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(u >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <int CTRL>
__device__ __forceinline__ void argmax_step(double& k, int& i) {
  const double k2 = dpp_d<CTRL>(k);
  const int i2 = dpp_i<CTRL>(i);
  if (better(k2, i2, k, i)) {
    k = k2;
    i = i2;
  }
}

// wavefront argmax: after the call every lane holds the wave's best pair
__device__ inline void wave_argmax(double& k, int& i) {
  argmax_step<0xB1>(k, i);   // quad_perm [1,0,3,2]
  argmax_step<0x4E>(k, i);   // quad_perm [2,3,0,1]
  argmax_step<0x141>(k, i);  // row_half_mirror
  argmax_step<0x140>(k, i);  // row_mirror
  double kb = readlane_d(k, 0);
  int ib = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
  for (int r = 1; r < 4; r++) {
    const double k2 = readlane_d(k, 16 * r);
    const int i2 = __builtin_amdgcn_readlane(i, 16 * r);
    if (better(k2, i2, kb, ib)) {
      kb = k2;
      ib = i2;
    }
  }
  k = kb;
  i = ib;
}

// workgroup argmax (blockDim multiple of 64, <= 1024); result in every
// thread.  One barrier: every thread reduces the per-wave results itself.
// sk/si are overwritten by the next call, so calls must be separated by
// another barrier.
__device__ inline void block_argmax(double& k, int& i, double* sk, int* si) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  wave_argmax(k, i);
  if (lane == 0) {
    sk[w] = k;
    si[w] = i;
  }
  __syncthreads();
  k = sk[0];
  i = si[0];
  for (int q = 1; q < nw; q++) {
    const double k2 = sk[q];
    const int i2 = si[q];
    if (better(k2, i2, k, i)) {
      k = k2;
      i = i2;
    }
  }
}

// Element j of a register row (j runtime but uniform, indices
// compile-time): a switch on the uniform j is a few scalar branches and one
// move, where a select over all PW elements would cost PW vector selects.
#define ATLG_RG(c) \
  case c:          \
    if constexpr (c < PW) return r[c]; else return S(0);
#define ATLG_RP(c)                      \
  case c:                               \
    if constexpr (c < PW) r[c] = x;     \
    break;
template <class S>
__device__ __forceinline__ S rget(const S (&r)[PW], int j) {
  switch (j) {
    ATLG_RG(0)
    ATLG_RG(1)
    ATLG_RG(2)
    ATLG_RG(3)
    ATLG_RG(4)
    ATLG_RG(5)
    ATLG_RG(6)
    ATLG_RG(7)
    ATLG_RG(8)
    ATLG_RG(9)
    ATLG_RG(10)
    ATLG_RG(11)
    ATLG_RG(12)
    ATLG_RG(13)
    ATLG_RG(14)
    ATLG_RG(15)
    ATLG_RG(16)
    ATLG_RG(17)
    ATLG_RG(18)
    ATLG_RG(19)
    ATLG_RG(20)
    ATLG_RG(21)
    ATLG_RG(22)
    ATLG_RG(23)
    ATLG_RG(24)
    ATLG_RG(25)
    ATLG_RG(26)
    ATLG_RG(27)
    ATLG_RG(28)
    ATLG_RG(29)
    ATLG_RG(30)
    ATLG_RG(31)
    ATLG_RG(32)
    ATLG_RG(33)
    ATLG_RG(34)
    ATLG_RG(35)
    ATLG_RG(36)
    ATLG_RG(37)
    ATLG_RG(38)
    ATLG_RG(39)
    ATLG_RG(40)
    ATLG_RG(41)
    ATLG_RG(42)
    ATLG_RG(43)
    ATLG_RG(44)
    ATLG_RG(45)
    ATLG_RG(46)
    ATLG_RG(47)
    ATLG_RG(48)
    ATLG_RG(49)
    ATLG_RG(50)
    ATLG_RG(51)
    ATLG_RG(52)
    ATLG_RG(53)
    ATLG_RG(54)
    ATLG_RG(55)
    ATLG_RG(56)
    ATLG_RG(57)
    ATLG_RG(58)
    ATLG_RG(59)
    ATLG_RG(60)
    ATLG_RG(61)
    ATLG_RG(62)
    ATLG_RG(63)
  }
  return S(0);
}
template <class S>
__device__ __forceinline__ void rput(S (&r)[PW], int j, S x) {
  switch (j) {
    ATLG_RP(0)
    ATLG_RP(1)
    ATLG_RP(2)
    ATLG_RP(3)
    ATLG_RP(4)
    ATLG_RP(5)
    ATLG_RP(6)
    ATLG_RP(7)
    ATLG_RP(8)
    ATLG_RP(9)
    ATLG_RP(10)
    ATLG_RP(11)
    ATLG_RP(12)
    ATLG_RP(13)
    ATLG_RP(14)
    ATLG_RP(15)
    ATLG_RP(16)
    ATLG_RP(17)
    ATLG_RP(18)
    ATLG_RP(19)
    ATLG_RP(20)
    ATLG_RP(21)
    ATLG_RP(22)
    ATLG_RP(23)
    ATLG_RP(24)
    ATLG_RP(25)
    ATLG_RP(26)
    ATLG_RP(27)
    ATLG_RP(28)
    ATLG_RP(29)
    ATLG_RP(30)
    ATLG_RP(31)
    ATLG_RP(32)
    ATLG_RP(33)
    ATLG_RP(34)
    ATLG_RP(35)
    ATLG_RP(36)
    ATLG_RP(37)
    ATLG_RP(38)
    ATLG_RP(39)
    ATLG_RP(40)
    ATLG_RP(41)
    ATLG_RP(42)
    ATLG_RP(43)
    ATLG_RP(44)
    ATLG_RP(45)
    ATLG_RP(46)
    ATLG_RP(47)
    ATLG_RP(48)
    ATLG_RP(49)
    ATLG_RP(50)
    ATLG_RP(51)
    ATLG_RP(52)
    ATLG_RP(53)
    ATLG_RP(54)
    ATLG_RP(55)
    ATLG_RP(56)
    ATLG_RP(57)
    ATLG_RP(58)
    ATLG_RP(59)
    ATLG_RP(60)
    ATLG_RP(61)
    ATLG_RP(62)
    ATLG_RP(63)
  }
}
#undef ATLG_RG
#undef ATLG_RP

// coherent (agent-scope) loads/stores for data shared between workgroups
__device__ inline double ld_agent(const double* p) {
  const unsigned long long u = __hip_atomic_load(
      reinterpret_cast<unsigned long long*>(const_cast<double*>(p)), __ATOMIC_RELAXED,
      __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double(static_cast<long long>(u));
}
__device__ inline void st_agent(double* p, double x) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                     static_cast<unsigned long long>(__double_as_longlong(x)), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_agent(const int* p) {
  return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_agent(int* p, int x) {
  __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float ld_agent(const float* p) {
  const unsigned u = __hip_atomic_load(reinterpret_cast<unsigned*>(const_cast<float*>(p)),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __uint_as_float(u);
}
__device__ inline void st_agent(float* p, float x) {
  __hip_atomic_store(reinterpret_cast<unsigned*>(p), __float_as_uint(x), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------ swap composition

// Composes k <= blockDim.x interchanges, step s swapping positions pa[s] and
// pb[s] (in order s = 0, 1, ...), into the list of rows whose content
// changes: afterwards row out[1+2t] holds the original row out[2+2t];
// out[0] = count.  Every thread of the block must call it.
//   A_s: content at pa[s] after step s;  B_s: content at pb[s] after step s.
//   A_s = content before step s at pb[s], B_s = content before s at pa[s],
// where "content before step s at x" is what the latest earlier step
// touching x left there, or x itself.  References are resolved by pointer
// jumping; a position's final content is what its last step left.
__device__ void compose_swaps(const int* pa, const int* pb, int k, int* val, int* cnt,
                              int* __restrict__ out) {
  const int s = threadIdx.x;
  if (s == 0) *cnt = 0;
  if (s < k) {
    const int a = pa[s], b = pb[s];
    int ra = -1, rb = -1;
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
  for (int span = 1; span < 2 * k; span *= 2) {  // pointer jumping
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
      const int t = atomicAdd(cnt, 1);
      out[1 + 2 * t] = a;
      out[2 + 2 * t] = val[2 * s];
    }
    if (lb && val[2 * s + 1] != b) {
      const int t = atomicAdd(cnt, 1);
      out[1 + 2 * t] = b;
      out[2 + 2 * t] = val[2 * s + 1];
    }
  }
  __syncthreads();
  if (s == 0) out[0] = *cnt;
}

// The same composition for k <= 64 interchanges by one wave, without
// barriers: lane s holds step s; the scans over the other steps read their
// rows from registers (readlane), the pointer jumping exchanges through
// cross-lane permutes, and the list is written in (step, a before b) order.
// Called by all 64 lanes of the wave.
__device__ void compose_swaps_wave(const int* pa, const int* pb, int k, int* __restrict__ out) {
  const int s = threadIdx.x & 63;
  const bool act = s < k;
  const int a = act ? pa[s] : -1, b = act ? pb[s] : -1;
  int ra = -1, rb = -1;  // latest earlier step touching b / a, as 2q (its pa) or 2q+1 (its pb)
  bool la = act, lb = act && a != b;  // no later step touches a / b
  for (int q = 0; q < k; q++) {
    const int aq = __builtin_amdgcn_readlane(a, q), bq = __builtin_amdgcn_readlane(b, q);
    if (q < s) {
      if (aq == b) ra = 2 * q;
      else if (bq == b) ra = 2 * q + 1;
      if (aq == a) rb = 2 * q;
      else if (bq == a) rb = 2 * q + 1;
    } else if (q > s) {
      if (aq == a || bq == a) la = false;
      if (bq == b || aq == b) lb = false;
    }
  }
  int v0 = !act ? 0 : (ra >= 0 ? -(ra + 1) : b);
  int v1 = !act ? 0 : (rb >= 0 ? -(rb + 1) : a);
  for (int span = 1; span < 2 * k; span *= 2) {  // pointer jumping
    const int i0 = v0 < 0 ? -v0 - 1 : 0, i1 = v1 < 0 ? -v1 - 1 : 0;
    const int x00 = __shfl(v0, i0 >> 1), x01 = __shfl(v1, i0 >> 1);
    const int x10 = __shfl(v0, i1 >> 1), x11 = __shfl(v1, i1 >> 1);
    if (v0 < 0) v0 = (i0 & 1) ? x01 : x00;
    if (v1 < 0) v1 = (i1 & 1) ? x11 : x10;
  }
  const bool e0 = la && v0 != a, e1 = lb && v1 != b;
  const unsigned long long m0 = __ballot(e0), m1 = __ballot(e1);
  const unsigned long long lt = (1ull << s) - 1ull;
  const int pos = __popcll(m0 & lt) + __popcll(m1 & lt);
  if (e0) {
    out[1 + 2 * pos] = a;
    out[2 + 2 * pos] = v0;
  }
  if (e1) {
    out[1 + 2 * (pos + e0)] = b;
    out[2 + 2 * (pos + e0)] = v1;
  }
  if (s == 0) out[0] = __popcll(m0) + __popcll(m1);
}

// Interchanges ipiv[i], i in [k1, k2) (ascending if inci > 0, else
// descending; at most KC of them) composed into a row list.
__global__ void __launch_bounds__(KC)
laswp_compose(const int* __restrict__ ipiv, int k1, int k2, int inci, int* __restrict__ out) {
  __shared__ int pa[KC], pb[KC], val[2 * KC], cnt;
  const int s = threadIdx.x, k = k2 - k1;
  if (s < k) {
    const int i = inci > 0 ? k1 + s : k2 - 1 - s;
    pa[s] = i;
    pb[s] = ipiv[i];
  }
  __syncthreads();
  compose_swaps(pa, pb, k, val, &cnt, out);
}

// Applies a composed list (at most 2*KC rows) to CW columns per workgroup:
// gather the source rows into LDS, then scatter them to their destinations.
template <class S>
__global__ void __launch_bounds__(256)
laswp_apply(long N, S* __restrict__ A, long lda, const int* __restrict__ lst) {
  __shared__ S buf[2 * KC * CW];
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

// ---- ATL_dlaswp as one composed permutation ---------------------------
//
// The interchanges are cut into chunks of KC; every chunk is composed into
// its own "row i receives row j" list in parallel (one workgroup each), the
// lists are folded in order into one row map held in LDS, and every column
// is then permuted in one pass: its touched row range read once into LDS
// (coalesced), written back once in the new order (coalesced).  Used while
// the leading dimension fits the LDS staging (kSwapRows rows).
constexpr int kSwapRows = 20224;          // rows staged per column (158 KiB of doubles)
constexpr int kSwapRowsSmall = 4096;      // staging of short columns (more workgroups per CU)
constexpr int kChunkInts = 1 + 4 * KC;    // one chunk's composed list

// Chunk c of the interchanges [k1, k2) in application order.
__device__ inline void swap_chunk(int c, int k1, int k2, int inci, int& a, int& b) {
  if (inci > 0) {
    a = k1 + c * KC;
    b = min(k2, a + KC);
  } else {
    b = k2 - c * KC;
    a = max(k1, b - KC);
  }
}

__global__ void __launch_bounds__(KC)
laswp_compose_chunks(const int* __restrict__ ipiv, int k1, int k2, int inci,
                     int* __restrict__ lists) {
  __shared__ int pa[KC], pb[KC], val[2 * KC], cnt;
  int a, b;
  swap_chunk(blockIdx.x, k1, k2, inci, a, b);
  const int s = threadIdx.x, k = b - a;
  if (s < k) {
    const int i = inci > 0 ? a + s : b - 1 - s;
    pa[s] = i;
    pb[s] = ipiv[i];
  }
  __syncthreads();
  compose_swaps(pa, pb, k, val, &cnt, lists + (size_t)blockIdx.x * kChunkInts);
}

// Folds the chunk lists, in order, into the row map of rows [0, R): map[r]
// = the original row whose content ends in row r.  hdr[0..1] receives the
// range [lo, hi) of rows that change (lo >= hi: none).
__global__ void __launch_bounds__(1024)
laswp_build_map(int nchunk, int R, const int* __restrict__ lists, int* __restrict__ map,
                int* __restrict__ hdr) {
  __shared__ int smap[kSwapRows];
  __shared__ int slo[16], shi[16];
  const int t = threadIdx.x;
  for (int r = t; r < R; r += 1024) smap[r] = r;
  __syncthreads();
  for (int c = 0; c < nchunk; c++) {
    const int* lst = lists + (size_t)c * kChunkInts;
    const int T = lst[0];  // <= 2 * KC <= 1024
    const int v = t < T ? smap[lst[2 + 2 * t]] : 0;
    __syncthreads();
    if (t < T) smap[lst[1 + 2 * t]] = v;
    __syncthreads();
  }
  int lo = INT_MAX, hi = -1;
  for (int r = t; r < R; r += 1024) {
    const int m = smap[r];
    map[r] = m;
    if (m != r) {
      lo = min(lo, r);
      hi = max(hi, r);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  if ((t & 63) == 0) {
    slo[t >> 6] = lo;
    shi[t >> 6] = hi;
  }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < 16; w++) {
      lo = min(lo, slo[w]);
      hi = max(hi, shi[w]);
    }
    hdr[0] = hi < 0 ? 0 : lo;
    hdr[1] = hi + 1;
  }
}

// One workgroup per column: rows [lo, hi) into LDS (all of a thread's loads
// in flight together), then row r <- staged row map[r].
template <class S, int CAP>
__global__ void __launch_bounds__(1024)
laswp_cols(S* __restrict__ A, long lda, const int* __restrict__ map, const int* __restrict__ hdr) {
  __shared__ S col[CAP];
  constexpr int U = (CAP + 1023) / 1024;
  const int lo = hdr[0], hi = hdr[1], t = threadIdx.x;
  if (hi <= lo) return;
  S* a = A + (size_t)blockIdx.x * lda;
  S v[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int r = lo + t + 1024 * u;
    if (r < hi) v[u] = a[r];
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int r = lo + t + 1024 * u;
    if (r < hi) col[r - lo] = v[u];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int r = lo + t + 1024 * u;
    if (r < hi) v[u] = col[map[r] - lo];
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int r = lo + t + 1024 * u;
    if (r < hi) a[r] = v[u];
  }
}

// Applies the composed lists of consecutive leaves (lists[l], l < nl, each
// kLeafListInts ints, in order) to the columns [c0, c0 + n0) and
// [c1, c1 + n1): per list, every thread reads its (at most NE) source
// elements into registers, all loads in flight, then after a barrier
// writes them to their destinations.  The same workgroup reads what the
// previous list wrote, so the barrier between lists is enough.
template <class S>
__global__ void __launch_bounds__(256)
laswp_lists(S* __restrict__ A, long lda, long c0, long n0, long c1, long n1,
            const int* __restrict__ lists, int nl) {
  constexpr int NE = 2 * PW * CW / 256;  // a list moves at most 2*PW rows
  static_assert(NE * 256 == 2 * PW * CW, "list elements per thread");
  const long nb0 = (n0 + CW - 1) / CW, b = blockIdx.x;
  const long cb = b < nb0 ? c0 + b * CW : c1 + (b - nb0) * CW;
  const long ce = b < nb0 ? c0 + n0 : c1 + n1;
  const int nc = (int)min((long)CW, ce - cb);
  for (int l = 0; l < nl; l++) {
    const int* lst = lists + (size_t)l * kLeafListInts;
    const int T = lst[0];
    S v[NE];
#pragma unroll
    for (int u = 0; u < NE; u++) {
      const int e = threadIdx.x + 256 * u, t = e / CW, c = e % CW;
      if (e < T * CW && c < nc) v[u] = A[(size_t)lst[2 + 2 * t] + (size_t)(cb + c) * lda];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NE; u++) {
      const int e = threadIdx.x + 256 * u, t = e / CW, c = e % CW;
      if (e < T * CW && c < nc) A[(size_t)lst[1 + 2 * t] + (size_t)(cb + c) * lda] = v[u];
    }
    __syncthreads();
  }
}

// Deferred interchanges of the look-ahead LU.  Q_b (b = 1 .. nblk-1) is
// the row map of block b's interchanges applied to the identity: each of
// its leaf lists ("row dst receives row src") in order.  One workgroup per
// block; Q holds nblk-1 arrays of M rows (Q_b at (b-1)*M).
__global__ void __launch_bounds__(256)
la_block_maps(int M, int NB, int mn, const int* __restrict__ lists, int* __restrict__ Q) {
  __shared__ int tmp[2 * PW];
  const int b = blockIdx.x + 1, t = threadIdx.x;
  int* q = Q + (size_t)blockIdx.x * M;
  for (int i = t; i < M; i += blockDim.x) q[i] = i;
  __syncthreads();
  const int l0 = b * NB / PW, l1 = (min(mn, (b + 1) * NB) + PW - 1) / PW;
  for (int l = l0; l < l1; l++) {
    const int* lst = lists + (size_t)l * kLeafListInts;
    const int T = lst[0];
    if (t < T) tmp[t] = q[lst[2 + 2 * t]];
    __syncthreads();
    if (t < T) q[lst[1 + 2 * t]] = tmp[t];
    __syncthreads();
  }
}

// P_j = Q_{j+1} o P_{j+1}, P_{nblk-1} = identity: the source row of row i
// in column block j once the interchanges of every later block are applied.
// P holds nblk-1 arrays of M rows (P_j at j*M).
__global__ void la_suffix_maps(int M, int nblk, const int* __restrict__ Q, int* __restrict__ P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  int p = i;
  for (int j = nblk - 2; j >= 0; j--) {
    p = Q[(size_t)j * M + p];  // Q_{j+1}
    P[(size_t)j * M + i] = p;
  }
}

// T(i - r, c) = A(P[i], c0 + c) for rows i in [r, M): one workgroup per
// column (the column's reads stay in one XCD's L2), then copied back.
template <class S>
__global__ void __launch_bounds__(256)
la_gather_rows(int M, int r, S* __restrict__ A, long lda, long c0, const int* __restrict__ P,
               S* __restrict__ T) {
  const long c = blockIdx.x;
  const S* a = A + (c0 + c) * lda;
  S* tc = T + c * (long)(M - r);
  for (int i = r + threadIdx.x; i < M; i += blockDim.x) tc[i - r] = a[P[i]];
}
template <class S>
__global__ void __launch_bounds__(256)
la_copy_rows(int M, int r, S* __restrict__ A, long lda, long c0, const S* __restrict__ T) {
  const long c = blockIdx.x;
  S* a = A + (c0 + c) * lda;
  const S* tc = T + c * (long)(M - r);
  for (int i = r + threadIdx.x; i < M; i += blockDim.x) a[i] = tc[i - r];
}

// ---------------------------------------------------------------- panels

// Words of the column mailboxes, handled as raw bits: a slot holding the
// sentinel has not been written yet.  The sentinel is a signalling-NaN
// pattern; a value with exactly that pattern is published as the quiet NaN
// (only the payload changes), so no written slot ever reads as empty.
template <class S>
struct Word;
template <>
struct Word<double> {
  typedef unsigned long long U;
  static constexpr U kSent = 0x7FF4A5A5A5A5A5A5ull;
  static constexpr U kQNaN = 0x7FF8000000000000ull;
  __device__ static U bits(double x) { return static_cast<U>(__double_as_longlong(x)); }
  __device__ static double val(U u) { return __longlong_as_double(static_cast<long long>(u)); }
};
template <>
struct Word<float> {
  typedef unsigned U;
  static constexpr U kSent = 0x7FA5A5A5u;
  static constexpr U kQNaN = 0x7FC00000u;
  __device__ static U bits(float x) { return __float_as_uint(x); }
  __device__ static float val(U u) { return __uint_as_float(u); }
};
constexpr unsigned long long kKeySent = Word<double>::kSent;  // keys are never NaN
constexpr unsigned kIdxSent = 0xFFFFFFFFu;                    // rows are never that

template <class U>
__device__ inline U ld_bits(const U* p) {
  return __hip_atomic_load(const_cast<U*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class U>
__device__ inline void st_bits(U* p, U v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class S>
__device__ inline typename Word<S>::U pub_bits(S x) {
  const typename Word<S>::U u = Word<S>::bits(x);
  return u == Word<S>::kSent ? Word<S>::kQNaN : u;
}

// Panel element i of column c through a buffer resource on the panel: the
// column offset travels in a scalar register, the thread's row offset in one
// VGPR shared by all 32 columns, so loads and stores need no per-element
// address registers (a store's address VGPRs could only be reused after the
// store completes).  Usable while PW columns span less than 2 GiB.
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
constexpr int kBufDword3 = 0x00020000;  // raw buffer, 32-bit dwords (gfx9 family)
__device__ inline double buf_ld(double, __amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
__device__ inline float buf_ld(float, __amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
__device__ inline void buf_st(double x, __amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, x), r, vo, so, 0);
}
__device__ inline void buf_st(float x, __amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, vo, so, 0);
}

// One leaf's exchange area: per column j and workgroup g the candidate
// (key, row) and its row, plus row j.  Every slot is written at most once
// per leaf, so a reader simply polls the slots it needs until none holds
// the sentinel.  Two areas alternate between consecutive leaves: a leaf
// refills the other one with sentinels for the next leaf (the previous
// leaf, which used it, has finished: launches on a stream are ordered).
template <class S>
struct Mailbox {
  typedef typename Word<S>::U U;
  unsigned long long key[PW][CP_MAXG];
  unsigned idx[PW][CP_MAXG];
  U crow[PW][CP_MAXG][PW];
  U jrow[PW][PW];
  int err;  // 1 when a wait ran past its bound (in any workgroup)
};

// Global scratch of the cooperative panel: the two alternating mailboxes.
template <class S>
struct CoopBuf {
  Mailbox<S> mb[2];
};

// Fills mailbox `m` with sentinels for the workgroups g < gclear (the
// previous user's grid) and clears its error word; the calling grid's
// threads share the work (thread gt of nt).
template <class S>
__device__ void mb_clear(Mailbox<S>* m, int gclear, int gt, int nt) {
  for (int e = gt; e < PW * gclear; e += nt) {
    const int j = e / gclear, g = e - j * gclear;
    m->key[j][g] = kKeySent;
    m->idx[j][g] = kIdxSent;
  }
  const int per = gclear * PW;
  for (int e = gt; e < PW * per; e += nt) {
    const int j = e / per, r = e - j * per;
    m->crow[j][r / PW][r % PW] = Word<S>::kSent;
  }
  for (int e = gt; e < PW * PW; e += nt) m->jrow[e / PW][e % PW] = Word<S>::kSent;
  if (gt == 0) m->err = 0;
}

template <class S>
__global__ void mb_init(CoopBuf<S>* cb) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  mb_clear(&cb->mb[0], CP_MAXG, gt, nt);
  mb_clear(&cb->mb[1], CP_MAXG, gt, nt);
}

// Factors the M x w (w <= PW) panel P (whose row 0 is absolute row `off` of
// the matrix, on the diagonal).  Row i of the panel is held by global thread
// i mod (G*256), RPT rows per thread.  Writes ipiv[off + j] (absolute),
// *dinfo (first zero pivot), and the composed interchange list.  With
// G > 1 the workgroups exchange through mailbox `par` (and refill mailbox
// par^1, last used by a grid of gclear workgroups); G == 1 stays in LDS.
template <class S, int RPT>
__global__ void __launch_bounds__(CP_THREADS)
panel_coop(int M, int w, int off, S* __restrict__ P, long lda, int* __restrict__ ipiv,
           int* __restrict__ dinfo, CoopBuf<S>* __restrict__ cb, int par, int gclear,
           int* __restrict__ list, int* __restrict__ status, long limit) {
  typedef typename Word<S>::U U;
  if (ATLG_PANEL_PRIO > 0) __builtin_amdgcn_s_setprio(ATLG_PANEL_PRIO);
  __shared__ S jrow[PW];
  __shared__ S crows[CP_MAXG * PW];
  __shared__ S crowm[CP_MAXG * PW];  // the same with columns <= j and >= w zeroed
  __shared__ S stage_c[PW], stage_j[PW];  // this workgroup's candidate row / row j
  __shared__ int sbad;
  __shared__ double sk[CP_THREADS / 64];
  __shared__ int si[CP_THREADS / 64];
  __shared__ double gk[2];  // best candidate of workgroups 0..63, 64..127
  __shared__ int gi[2];
  __shared__ int pa[PW], pb[PW];
  __shared__ int zcol;  // first column with a zero pivot (workgroup 0), -1: none
  const int t = threadIdx.x, G = gridDim.x, GT = G * CP_THREADS;
  const int g = blockIdx.x * CP_THREADS + t;
  Mailbox<S>* mb = &cb->mb[par];
  if (G > 1) mb_clear(&cb->mb[par ^ 1], gclear, g, GT);
  if (t == 0) {
    sbad = 0;
    zcol = -1;
  }
  const bool bufio = (long)PW * lda * (long)sizeof(S) < (1L << 31);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(P, 0, 0x7fffffff, kBufDword3);
  const int cstride = static_cast<int>(lda * (long)sizeof(S));  // used only when bufio
  S v[RPT][PW];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int i = g + q * GT;
    const unsigned boff = static_cast<unsigned>(i) * sizeof(S);  // M <= 65536 rows
#pragma unroll
    for (int c = 0; c < PW; c++) {
      S x = S(0);
      if (i < M && c < w) x = bufio ? buf_ld(S(0), rs, boff, c * cstride) : P[i + (size_t)c * lda];
      v[q][c] = x;
    }
  }
  const int mn = min(M, w);
  bool ok = true;
  for (int j = 0; j < mn; j++) {
    double key = -2.0;
    int idx = INT_MAX;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int i = g + q * GT;
      if (i >= j && i < M) {
        const double kq = pkey(rget(v[q], j));
        if (better(kq, i, key, idx)) {
          key = kq;
          idx = i;
        }
      }
    }
    block_argmax(key, idx, sk, si);
    int win = idx;  // the global pivot row (G == 1: this workgroup's)
    if (G == 1) {
      // the whole panel is in this workgroup: exchange through LDS only
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const int i = g + q * GT;
        if (i == idx) {
#pragma unroll
          for (int c = 0; c < PW; c++) {
            crows[c] = (c < w) ? v[q][c] : S(0);
            crowm[c] = (c > j && c < w) ? v[q][c] : S(0);
          }
        }
        if (i == j) {
#pragma unroll
          for (int c = 0; c < PW; c++) jrow[c] = (c < w) ? v[q][c] : S(0);
        }
      }
      __syncthreads();
    } else {
      // Publish: the owners copy the candidate row and row j to LDS, then one
      // lane per element stores them (an agent-scope atomic store waits for
      // the stores before it, so one lane storing a whole row would pay a
      // round trip per element).  A workgroup without rows >= j publishes
      // key -2 and a zero row (never chosen: rows j.. always exist).
      const bool has_c = idx != INT_MAX;
      const bool has_j = (j % GT) / CP_THREADS == (int)blockIdx.x;
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const int i = g + q * GT;
        if (i == idx) {
#pragma unroll
          for (int c = 0; c < PW; c++) stage_c[c] = v[q][c];
        }
        if (i == j) {
#pragma unroll
          for (int c = 0; c < PW; c++) stage_j[c] = v[q][c];
        }
      }
      __syncthreads();
      if (t < PW) {
        if (t < w) st_bits(&mb->crow[j][blockIdx.x][t], pub_bits(has_c ? stage_c[t] : S(0)));
      } else if (t < 2 * PW) {
        if (has_j && t - PW < w) st_bits(&mb->jrow[j][t - PW], pub_bits(stage_j[t - PW]));
      } else if (t == 2 * PW) {
        st_bits(&mb->key[j][blockIdx.x], Word<double>::bits(key));
      } else if (t == 2 * PW + 1) {
        st_bits(&mb->idx[j][blockIdx.x], static_cast<unsigned>(idx));
      }
      // Gather: all G candidate rows (8 slots per thread, all loads in
      // flight), the G candidates (threads < G) and row j (threads 128..),
      // each slot polled until written; one round trip once the last
      // workgroup has published.
      constexpr int NU = CP_MAXG * PW / CP_THREADS;
      U cw[NU];
#pragma unroll
      for (int u = 0; u < NU; u++) cw[u] = Word<S>::kSent;
      unsigned long long kw = kKeySent;
      unsigned iw = kIdxSent;
      U jw = Word<S>::kSent;
      const bool want_k = t < G, want_j = t >= kJrowT && t - kJrowT < w;
      long spins = 0;
      for (;;) {
        // issue every pending load before looking at any result, so that
        // they are all in flight together (one round trip per poll)
        bool pend[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) {
          const int e = t + u * CP_THREADS, c = e % PW;
          pend[u] = e < G * PW && c < w && cw[u] == Word<S>::kSent;
        }
        const bool pk = want_k && kw == kKeySent, pi = want_k && iw == kIdxSent;
        const bool pj = want_j && jw == Word<S>::kSent;
#pragma unroll
        for (int u = 0; u < NU; u++) {
          const int e = t + u * CP_THREADS, bq = e / PW, c = e - bq * PW;
          if (pend[u]) cw[u] = ld_bits(&mb->crow[j][bq][c]);
        }
        if (pk) kw = ld_bits(&mb->key[j][t]);
        if (pi) iw = ld_bits(&mb->idx[j][t]);
        if (pj) jw = ld_bits(&mb->jrow[j][t - kJrowT]);
        bool waiting = (want_k && (kw == kKeySent || iw == kIdxSent)) ||
                       (want_j && jw == Word<S>::kSent);
#pragma unroll
        for (int u = 0; u < NU; u++) {
          const int e = t + u * CP_THREADS, c = e % PW;
          waiting |= e < G * PW && c < w && cw[u] == Word<S>::kSent;
        }
        if (!waiting) break;
        // past the bound, or another workgroup gave up (its error word is
        // read every 256 polls)
        if (++spins >= limit || ((spins & 255) == 0 && ld_agent(&mb->err))) {
          st_agent(&mb->err, 1);
          sbad = 1;
          break;
        }
      }
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int e = t + u * CP_THREADS, c = e % PW;
        if (e < G * PW) {
          const S x = Word<S>::val(cw[u]);
          crows[e] = c < w ? x : S(0);
          crowm[e] = (c > j && c < w) ? x : S(0);
        }
      }
      if (t >= kJrowT && t < kJrowT + PW) jrow[t - kJrowT] = want_j ? Word<S>::val(jw) : S(0);
      if (t < 128) {  // the candidates of workgroups 0..63 and 64..127
        double kk = want_k ? Word<double>::val(kw) : -3.0;
        int ii = want_k ? static_cast<int>(iw) : INT_MAX;
        wave_argmax(kk, ii);
        if ((t & 63) == 0) {
          gk[t >> 6] = kk;
          gi[t >> 6] = ii;
        }
      }
      __syncthreads();
      if (sbad) {
        ok = false;
        break;
      }
      win = better(gk[1], gi[1], gk[0], gi[0]) ? gi[1] : gi[0];
    }  // G > 1
    // rows j..M-1 always hold a candidate, so p is in [j, M); the clamp only
    // keeps a corrupted exchange from indexing outside the panel
    const int p = (win >= j && win < M) ? win : j;
    const int bw = (p % GT) / CP_THREADS;  // workgroup that published row p
    const S* prow = crows + bw * PW;        // the pivot row
    const S* pmsk = crowm + bw * PW;        // ... with columns <= j and >= w zeroed
    // the masked pivot row into registers first, every read in flight
    // together and ahead of the reciprocal (read one by one inside the
    // update, each FMA waited on its own LDS round trip)
    S pm[PW];
#pragma unroll
    for (int c = 0; c < PW; c++) pm[c] = pmsk[c];
    const S piv = prow[j];
    // the pivots and the first zero pivot stay in LDS until the end: a
    // global store here would stall this wave on its completion (stores
    // and loads share one counter), delaying every workgroup's next column
    if (blockIdx.x == 0 && t == 0) {
      pa[j] = off + j;
      pb[j] = off + p;
      if (piv == S(0) && zcol < 0) zcol = j;
    }
    if (piv != S(0)) {
      const bool recip = fabs(piv) >= std::numeric_limits<S>::min();
      const S rinv = S(1) / piv;
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const int i = g + q * GT;
        if (p != j) {
          if (i == j) {
#pragma unroll
            for (int c = 0; c < PW; c++) v[q][c] = prow[c];
          } else if (i == p) {
#pragma unroll
            for (int c = 0; c < PW; c++) v[q][c] = jrow[c];
          }
        }
        if (i > j && i < M) {
          // scale column j, then the rank-1 update.  With a finite
          // multiplier the masked pivot row makes the update one FMA per
          // column (columns <= j and >= w see a zero); otherwise only the
          // columns j < c < w are touched, as in the unmasked form.
          const S x = rget(v[q], j);
          S l;
          if (recip) l = x * rinv;
          else l = x / piv;
          if (isfinite(l)) {
#pragma unroll
            for (int c = 0; c < PW; c++) v[q][c] -= l * pm[c];
          } else {
#pragma unroll
            for (int c = 0; c < PW; c++)
              if (c > j && c < w) v[q][c] -= l * prow[c];
          }
          rput(v[q], j, l);
        }
      }
    }
  }
  if (bufio && w == PW) {
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int i = g + q * GT;
      const unsigned boff = static_cast<unsigned>(i) * sizeof(S);
      if (i < M) {
#pragma unroll
        for (int c = 0; c < PW; c++) buf_st(v[q][c], rs, boff, c * cstride);
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int i = g + q * GT;
      if (i < M) {
#pragma unroll
        for (int c = 0; c < PW; c++)
          if (c < w) P[i + (size_t)c * lda] = v[q][c];
      }
    }
  }
  if (!ok) {
    if (t == 0) {
      *dinfo = INT_MIN;
      atomicOr(status, ATLG_FAIL_PANEL);
    }
    if (blockIdx.x == 0) {
      if (t < mn) ipiv[off + t] = off + t;  // identity: later users stay in bounds
      if (t == 0) list[0] = 0;              // apply no interchanges
    }
    return;
  }
  if (blockIdx.x == 0) {
    __syncthreads();
    if (t < mn) ipiv[off + t] = pb[t];
    if (t == 0 && zcol >= 0 && *dinfo == 0) *dinfo = off + zcol + 1;
    if (t < 64) compose_swaps_wave(pa, pb, mn, list);
  }
}

// Pivot record published by the last workgroup of a column step.
template <class S>
struct PivRec {
  S piv;
  int p;
  int pad;
};

// Fallback for panels taller than the cooperative kernel holds: one launch
// per column.  Launch j (0 <= j <= mn) first applies column j-1's scaling
// and rank-1 update to rows [j, M), then (j < mn) finds the pivot of column
// j; the last workgroup to finish reduces the candidates, swaps rows j and
// p across the panel and publishes the pivot (ipiv absolute: off + p).
template <class S>
__global__ void __launch_bounds__(STEP_THREADS)
panel_step(int M, int w, int j, int mn, int off, S* __restrict__ A, long lda,
           int* __restrict__ ipiv, int* __restrict__ dinfo, Cand* __restrict__ cand,
           unsigned* __restrict__ ticket, PivRec<S>* __restrict__ rec) {
  __shared__ S urow[PW];
  __shared__ double sk[4];
  __shared__ int si[4];
  __shared__ bool last;
  const int t = threadIdx.x;
  const int r = j + blockIdx.x * STEP_THREADS + t;
  const bool mine = r < M;
  if (j > 0) {
    const PivRec<S> pr = rec[j - 1];
    if (t < w) urow[t] = A[(j - 1) + (size_t)t * lda];
    __syncthreads();
    if (mine && pr.piv != S(0)) {
      S* a = A + r;
      S l = a[(size_t)(j - 1) * lda];
      l = (fabs(pr.piv) >= std::numeric_limits<S>::min()) ? l * (S(1) / pr.piv) : l / pr.piv;
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
  const S piv = A[p + (size_t)j * lda];
  if (piv != S(0) && p != j && t < w) {
    S* a = A + (size_t)t * lda;
    const S x = a[j];
    a[j] = a[p];
    a[p] = x;
  }
  if (t == 0) {
    ipiv[off + j] = off + p;
    if (piv == S(0) && *dinfo == 0) *dinfo = off + j + 1;
    rec[j].piv = piv;
    rec[j].p = p;
    *ticket = 0u;
  }
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

// ---------------------------------------------------------------- small nodes

// A recursion node of width K (32 or 64) with nc <= 64 columns to its right,
// in one launch: U12 = inv(L11) A12 and A22 -= L21 U12, L11 being the node's
// K x K unit lower block.  Every workgroup inverts L11 in LDS and forms U12
// itself, so no workgroup waits for another.  The last workgroup to have
// read A12 (a ticket counted after the reads have returned, reset by that
// workgroup for the next launch) writes U12 over it.  Each workgroup then updates 64 rows of A22 on the matrix cores (one
// wave per 16 rows), with U12 from LDS as the first operand so that a lane's
// results are consecutive rows of one column.  Its rows of L21 and A22 are
// loaded first and arrive while the inverse is formed.  This is the TRSM and
// GEMM of getrfC's recursion step (AtlasBase/Clint/atlas-lp.base:1540-1575)
// for the narrowest nodes, where three launches cost more than the work.
template <class S, int K>
__global__ void __launch_bounds__(kTriThreads)
lu_node_small(int mrows, int nc, const S* __restrict__ L11, S* __restrict__ A12,
              const S* __restrict__ L21, S* __restrict__ A22, long lda,
              unsigned* __restrict__ ticket) {
  static_assert(K == 32 || K == 64, "node width");
  typedef TriMfma<S> Mf;
  typedef typename Mf::acc_t acc_t;
  constexpr int LD = K + 1;   // pitch of the inverse, [row][col]
  constexpr int UP = 64 + 1;  // pitch of A12 / U12, [k][col]
  constexpr int KT = K / 16;  // row tiles of U12
  __shared__ S Li[K * LD];
  __shared__ S U[K * UP];
  __shared__ int writer;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, lx = lane & 15, lk = lane >> 4;
  const int row = blockIdx.x * 64 + 16 * wv + lx;  // row of A22 / L21
  const bool rv = row < mrows;
  S l21[K / 4], c[4][4];
#pragma unroll
  for (int q = 0; q < K / 4; q++) l21[q] = rv ? L21[row + (size_t)(4 * q + lk) * lda] : S(0);
#pragma unroll
  for (int ct = 0; ct < 4; ct++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int col = 16 * ct + Mf::row(lane, r);
      c[ct][r] = (rv && col < nc) ? A22[row + (size_t)col * lda] : S(0);
    }
  {
    constexpr int PL = K * K / kTriThreads, PU = K * 64 / kTriThreads;
    S vl[PL], vu[PU];
#pragma unroll
    for (int u = 0; u < PL; u++) {
      const int e = t + u * kTriThreads, i = e % K, j = e / K;
      vl[u] = i > j ? L11[i + (size_t)j * lda] : (i == j ? S(1) : S(0));
    }
#pragma unroll
    for (int u = 0; u < PU; u++) {
      const int e = t + u * kTriThreads, i = e % K, j = e / K;
      vu[u] = j < nc ? A12[i + (size_t)j * lda] : S(0);
    }
#pragma unroll
    for (int u = 0; u < PL; u++) {
      const int e = t + u * kTriThreads;
      Li[(e % K) * LD + e / K] = vl[u];
    }
#pragma unroll
    for (int u = 0; u < PU; u++) {
      const int e = t + u * kTriThreads;
      U[(e % K) * UP + e / K] = vu[u];
    }
  }
  __syncthreads();  // every read of A12 has returned (its values are in LDS)
  if (t == 0) {
    writer = atomicAdd(ticket, 1u) == gridDim.x - 1;
    if (writer) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  lower_inverse_inplace<S, K>(Li, K);  // ends with a barrier
  // U12 = inv(L11) A12: wave wv forms column tile wv, every row tile
  const bool cw = 16 * wv < nc;
  acc_t u12[KT];
#pragma unroll
  for (int rt = 0; rt < KT; rt++) {
    u12[rt] = acc_t{0, 0, 0, 0};
    if (cw) {
      for (int k0 = 0; k0 < 16 * (rt + 1); k0 += 4)  // inv(L11) is lower
        u12[rt] = Mf::op(Li[(16 * rt + lx) * LD + k0 + lk], U[(k0 + lk) * UP + 16 * wv + lx],
                         u12[rt]);
    }
  }
  __syncthreads();
  if (cw) {
#pragma unroll
    for (int rt = 0; rt < KT; rt++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * rt + Mf::row(lane, r), col = 16 * wv + lx;
        U[i * UP + col] = u12[rt][r];
        if (writer && col < nc) A12[i + (size_t)col * lda] = u12[rt][r];
      }
  }
  __syncthreads();
  // A22 -= L21 U12 as (U12^T L21^T)^T
#pragma unroll
  for (int ct = 0; ct < 4; ct++) {
    if (16 * ct >= nc) continue;
    acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < K / 4; q++) acc = Mf::op(U[(4 * q + lk) * UP + 16 * ct + lx], l21[q], acc);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int col = 16 * ct + Mf::row(lane, r);
      if (rv && col < nc) A22[row + (size_t)col * lda] = c[ct][r] - acc[r];
    }
  }
}

// ---------------------------------------------------------------- drivers

template <class S>
struct LU {
  hipStream_t s;
  S* A;
  long lda, M, N;
  int* ipiv;
  int* dinfo;
  CoopBuf<S>* cb;
  int* list;
  Cand* cand;        // panel_step scratch
  unsigned* ticket;
  PivRec<S>* rec;
  S* Winv;           // inverses of the 128-blocks of L
  std::vector<char> inv_done;
  int gemm_prec;     // kGemmNative or kGemmBf16 (float only)
  long win_lo, win_hi;  // leaves apply their interchanges to these columns only
  bool coop = true;     // cooperative launch of the panel kernel
  int gemm_cap = 0;     // grid cap of the trailing updates (0: none)
  int* leaf_lists = nullptr;  // look-ahead: leaf at column c keeps its list at c/PW
  int* status = nullptr;      // the caller's stream's failure word
  long limit = kSpinLimit;    // mailbox polls before a panel gives up
  bool small_lds = false;     // runs beside a GEMM: keep kernels' LDS small
  int mb_par = 0;             // mailbox the next multi-workgroup panel uses
  int mb_g[2] = {0, 0};       // grid of each mailbox's last user (0: clean)
};

// Trailing update C -= A * B of a recursion node.
inline void node_update(LU<double>& f, long m, long n, long k, const double* A, double* B,
                        double* C) {
  gemm<double>(f.s, kNoTrans, kNoTrans, m, n, k, -1.0, A, f.lda, B, f.lda, 1.0, C, f.lda, 0,
               nullptr, f.gemm_cap);
}
// The BF16 forms (two operand conversions and the product) serve the wide
// updates, k >= 256; narrower ones inside the panels stay on the FP32 MFMA.
inline void node_update(LU<float>& f, long m, long n, long k, const float* A, float* B,
                        float* C) {
  if (f.gemm_prec != kGemmNative && k >= 256)
    gemm_bf16(f.s, m, n, k, -1.0f, A, f.lda, B, f.lda, 1.0f, C, f.lda,
              f.gemm_prec == kGemmBf16x6 ? 3 : (f.gemm_prec == kGemmBf16x3 ? 2 : 1));
  else
    gemm<float>(f.s, kNoTrans, kNoTrans, m, n, k, -1.0f, A, f.lda, B, f.lda, 1.0f, C, f.lda, 0,
                nullptr, f.gemm_cap);
}

template <class S>
S* at(const LU<S>& f, long i, long j) { return f.A + i + j * f.lda; }

// Applies a composed list to the window's columns except [c0, c0 + w).  The
// list is a leaf's: w <= PW interchanges, so at most 2*PW rows, the most
// laswp_lists holds per list.
template <class S>
void apply_outside(LU<S>& f, long c0, long w, const int* lst) {
  if (w > PW) throw HipFailure{"getrf: leaf wider than the interchange list capacity"};
  const long nl = std::max(0L, c0 - f.win_lo);
  const long r0 = c0 + w, nr = std::max(0L, f.win_hi - r0);
  const long nblk = (nl + CW - 1) / CW + (nr + CW - 1) / CW;
  if (nblk == 0) return;
  // both column ranges in one launch
  hipLaunchKernelGGL(laswp_lists<S>, dim3((unsigned)nblk), dim3(256), 0, f.s, f.A, f.lda,
                     f.win_lo, nl, r0, nr, lst, 1);
  ATLG_LAUNCHED();
}

template <class S, int RPT>
void launch_coop(LU<S>& f, int M, int w, int off, int G, int* lst) {
  S* P = at(f, off, off);
  long lda = f.lda;
  // a multi-workgroup panel takes the next mailbox and refills the other
  int par = 0, gclear = 0;
  if (G > 1) {
    par = f.mb_par;
    gclear = f.mb_g[par ^ 1];
    f.mb_g[par ^ 1] = 0;
    f.mb_g[par] = G;
    f.mb_par = par ^ 1;
  }
  if (!f.coop) {
    // Ordinary launch (look-ahead panel beside a running GEMM): the G <= 64
    // workgroups fit the device many times over, so any of them still
    // waiting for a slot gets one as the independent GEMM's workgroups
    // retire; the mailbox waits are bounded anyway.
    hipLaunchKernelGGL((panel_coop<S, RPT>), dim3(G), dim3(CP_THREADS), 0, f.s, M, w, off, P, lda,
                       f.ipiv, f.dinfo, f.cb, par, gclear, lst, f.status, f.limit);
    ATLG_LAUNCHED();
    return;
  }
  void* args[] = {&M,      &w,      &off, &P,   &lda,      &f.ipiv,  &f.dinfo,
                  &f.cb,   &par,    &gclear, &lst, &f.status, &f.limit};
  ATLG_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&panel_coop<S, RPT>), dim3(G),
                                      dim3(CP_THREADS), args, 0, f.s));
}

// Factors the leaf panel at (off, off): m rows, w columns, and applies its
// interchanges to every other column.
template <class S>
void leaf(LU<S>& f, long off, long m, int w) {
  int* lst = f.leaf_lists ? f.leaf_lists + (off / PW) * kLeafListInts : f.list;
  const long cap = (long)CP_MAXG * CP_THREADS;
  if (m <= 4 * cap) {
    const int rpt = std::max(ATLG_PANEL_RPT_MIN, m <= cap ? 1 : (m <= 2 * cap ? 2 : 4));
    const int G = (int)((m + (long)CP_THREADS * rpt - 1) / ((long)CP_THREADS * rpt));
    if (rpt == 1) launch_coop<S, 1>(f, (int)m, w, (int)off, G, lst);
    else if (rpt == 2) launch_coop<S, 2>(f, (int)m, w, (int)off, G, lst);
    else launch_coop<S, 4>(f, (int)m, w, (int)off, G, lst);
    apply_outside(f, off, w, lst);
    return;
  }
  const int mn = (int)std::min<long>(m, w);
  S* P = at(f, off, off);
  for (int j = 0; j <= mn; j++) {
    const unsigned nb = (unsigned)std::max<long>(1, (m - j + STEP_THREADS - 1) / STEP_THREADS);
    hipLaunchKernelGGL(panel_step<S>, dim3(nb), dim3(STEP_THREADS), 0, f.s, (int)m, w, j, mn,
                       (int)off, P, f.lda, f.ipiv, f.dinfo, f.cand, f.ticket, f.rec);
    ATLG_LAUNCHED();
  }
  hipLaunchKernelGGL(laswp_compose, dim3(1), dim3(KC), 0, f.s, f.ipiv, (int)off, (int)off + mn,
                     1, lst);
  ATLG_LAUNCHED();
  apply_outside(f, off, w, lst);
}

// U12 <- L11^-1 A12 for the node at `off` with nl pivot columns, on the nc
// columns starting at column c0 (c0 >= off + nl).
template <class S>
void node_trsm_cols(LU<S>& f, long off, long nl, long c0, long nc) {
  if (nc <= 0) return;
  if (f.small_lds && nl == kTrsmNB) {
    // beside a running GEMM: two 64-column halves, whose diagonal-block
    // inverses need a quarter of the LDS of a 128-block inverse (which
    // would wait for a compute unit free of GEMM workgroups)
    const long h = nl / 2;
    trsm<S>(f.s, kLeft, kLower, kNoTrans, kUnit, h, nc, S(1), at(f, off, off), f.lda,
            at(f, off, c0), f.lda);
    gemm<S>(f.s, kNoTrans, kNoTrans, h, nc, h, S(-1), at(f, off + h, off), f.lda, at(f, off, c0),
            f.lda, S(1), at(f, off + h, c0), f.lda);
    trsm<S>(f.s, kLeft, kLower, kNoTrans, kUnit, h, nc, S(1), at(f, off + h, off + h), f.lda,
            at(f, off + h, c0), f.lda);
    return;
  }
  if (nl >= kTrsmNB && off % kTrsmNB == 0 && nl % kTrsmNB == 0) {
    const long b0 = off / kTrsmNB, b1 = (off + nl) / kTrsmNB;
    long first = b0;
    while (first < b1 && f.inv_done[first]) first++;
    if (first < b1) {
      tri_inverse_blocks<S>(f.s, (b1 - first) * kTrsmNB, at(f, first * kTrsmNB, first * kTrsmNB),
                         f.lda, kNoTrans, kLower, kUnit,
                         f.Winv + (size_t)first * kTrsmNB * kTrsmNB);
      for (long b = first; b < b1; b++) f.inv_done[b] = 1;
    }
    trsm<S>(f.s, kLeft, kLower, kNoTrans, kUnit, nl, nc, S(1), at(f, off, off), f.lda,
            at(f, off, c0), f.lda, nullptr, f.Winv + (size_t)b0 * kTrsmNB * kTrsmNB, 0);
  } else {
    trsm<S>(f.s, kLeft, kLower, kNoTrans, kUnit, nl, nc, S(1), at(f, off, off), f.lda,
            at(f, off, c0), f.lda);
  }
}

template <class S>
void node_trsm(LU<S>& f, long off, long nl, long nr) {
  node_trsm_cols(f, off, nl, off + nl, nr);
}

// The node at `off` (nl pivot columns, nr columns to their right) in one
// lu_node_small launch when it is narrow enough; false otherwise.
#ifndef ATLG_LU_NODE_SMALL
#define ATLG_LU_NODE_SMALL 1
#endif
template <class S>
bool node_small(LU<S>& f, long off, long nl, long nr) {
  if (!ATLG_LU_NODE_SMALL || !(nl == 32 || nl == 64) || nr <= 0 || nr > 64) return false;
  const long mrows = f.M - off - nl;
  if (mrows >= INT_MAX - 64) return false;
  const unsigned grid = (unsigned)std::max(1L, (mrows + 63) / 64);
  const S* L11 = at(f, off, off);
  S* A12 = at(f, off, off + nl);
  const S* L21 = at(f, off + nl, off);
  S* A22 = at(f, off + nl, off + nl);
  if (nl == 32)
    hipLaunchKernelGGL((lu_node_small<S, 32>), dim3(grid), dim3(kTriThreads), 0, f.s, (int)mrows,
                       (int)nr, L11, A12, L21, A22, f.lda, f.ticket);
  else
    hipLaunchKernelGGL((lu_node_small<S, 64>), dim3(grid), dim3(kTriThreads), 0, f.s, (int)mrows,
                       (int)nr, L11, A12, L21, A22, f.lda, f.ticket);
  ATLG_LAUNCHED();
  return true;
}

// Factors rows [off, M) x columns [off, off + n).
template <class S>
void getrf_rec(LU<S>& f, long off, long n) {
  const long m = f.M - off, mn = std::min(m, n);
  if (mn <= 0) return;
  if (n <= PW) {
    leaf(f, off, m, (int)n);
    return;
  }
  if (mn <= PW) {  // short and wide: factor the square part, solve the rest
    leaf(f, off, m, (int)mn);
    node_trsm(f, off, mn, n - mn);
    return;
  }
  long nl;
  if (mn >= 2 * kTrsmNB) nl = ((mn / 2) / kTrsmNB) * kTrsmNB;
  else nl = std::max<long>(PW, ((mn / 2) / PW) * PW);
  const long nr = n - nl;
  getrf_rec(f, off, nl);
  if (!node_small(f, off, nl, nr)) {
    node_trsm(f, off, nl, nr);
    node_update(f, m - nl, nr, nl, at(f, off + nl, off), at(f, off, off + nl),
                at(f, off + nl, off + nl));
  }
  getrf_rec(f, off + nl, nr);
}

}  // namespace

void set_int(hipStream_t s, int* p, int v) {
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, p, v);
  ATLG_LAUNCHED();
}

template <class S>
void laswp(hipStream_t s, long N, S* A, long lda, int K1, int K2, const int* ipiv, int inci) {
  if (N <= 0 || K2 <= K1) return;
  // every row index (K1..K2 and the pivots) is below lda: one composed
  // permutation, one pass over the columns
  if (lda <= kSwapRows && N <= INT_MAX) {
    const int nchunk = (K2 - K1 + KC - 1) / KC, R = (int)lda;
    int* ws = static_cast<int*>(
        scratch(s, kSlotLaswp, sizeof(int) * ((size_t)nchunk * kChunkInts + R + 64)));
    int* lists = ws;
    int* map = ws + (size_t)nchunk * kChunkInts;
    int* hdr = map + R;
    hipLaunchKernelGGL(laswp_compose_chunks, dim3((unsigned)nchunk), dim3(KC), 0, s, ipiv, K1, K2,
                       inci, lists);
    ATLG_LAUNCHED();
    hipLaunchKernelGGL(laswp_build_map, dim3(1), dim3(1024), 0, s, nchunk, R, lists, map, hdr);
    ATLG_LAUNCHED();
    if (lda <= kSwapRowsSmall)
      hipLaunchKernelGGL((laswp_cols<S, kSwapRowsSmall>), dim3((unsigned)N), dim3(1024), 0, s, A,
                         lda, map, hdr);
    else
      hipLaunchKernelGGL((laswp_cols<S, kSwapRows>), dim3((unsigned)N), dim3(1024), 0, s, A, lda,
                         map, hdr);
    ATLG_LAUNCHED();
    return;
  }
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
    hipLaunchKernelGGL(laswp_apply<S>, dim3(nb), dim3(256), 0, s, N, A, lda, lst);
    ATLG_LAUNCHED();
  }
}

namespace {

// Scratch and bookkeeping of one factorization stream.
template <class S>
void init_lu(LU<S>& f, hipStream_t s, long M, long N, S* A, long lda, int* ipiv, int* dinfo,
             int gemm_prec) {
  f.s = s;
  f.A = A;
  f.lda = lda;
  f.M = M;
  f.N = N;
  f.ipiv = ipiv;
  f.dinfo = dinfo;
  f.gemm_prec = gemm_prec;
  f.win_lo = 0;
  f.win_hi = N;
  const long maxblk = (M + STEP_THREADS - 1) / STEP_THREADS + 1;
  const size_t b_cb = (sizeof(CoopBuf<S>) + 255) & ~size_t(255);
  const size_t b_list = 256 + sizeof(int) * (1 + 4 * KC);
  const size_t b_rec = 256 + sizeof(PivRec<S>) * PW;
  const size_t b_cand = sizeof(Cand) * maxblk;
  char* base = static_cast<char*>(scratch(s, kSlotPanel, b_cb + b_list + 256 + b_rec + b_cand));
  f.cb = reinterpret_cast<CoopBuf<S>*>(base);
  f.list = reinterpret_cast<int*>(base + b_cb);
  f.ticket = reinterpret_cast<unsigned*>(base + b_cb + b_list);
  f.rec = reinterpret_cast<PivRec<S>*>(base + b_cb + b_list + 256);
  f.cand = reinterpret_cast<Cand*>(base + b_cb + b_list + 256 + b_rec);
  set_int(s, reinterpret_cast<int*>(f.ticket), 0);
  hipLaunchKernelGGL(mb_init<S>, dim3(128), dim3(256), 0, s, f.cb);
  ATLG_LAUNCHED();
  f.mb_par = 0;
  f.mb_g[0] = f.mb_g[1] = 0;
  const long nblk = (std::min(M, N) + kTrsmNB - 1) / kTrsmNB;
  f.Winv = static_cast<S*>(
      scratch(s, kSlotFactInv, sizeof(S) * std::max<long>(nblk, 1) * kTrsmNB * kTrsmNB));
  f.inv_done.assign(nblk + 1, 0);
}

// Applies the interchanges of the leaves covering columns [k0, k0 + kb) to
// the columns [c0, c0 + n0) and [c1, c1 + n1), in pivot order.
template <class S>
void apply_block_lists(LU<S>& f, long k0, long kb, long c0, long n0, long c1, long n1) {
  n0 = std::max(0L, n0);
  n1 = std::max(0L, n1);
  const long nblk = (n0 + CW - 1) / CW + (n1 + CW - 1) / CW;
  if (nblk == 0) return;
  hipLaunchKernelGGL(laswp_lists<S>, dim3((unsigned)nblk), dim3(256), 0, f.s, f.A, f.lda, c0, n0,
                     c1, n1, f.leaf_lists + (k0 / PW) * kLeafListInts, (int)((kb + PW - 1) / PW));
  ATLG_LAUNCHED();
}

// Applies, to every column block j but the last, the interchanges of all
// later blocks (deferred there from the look-ahead steps): one composed
// row map per block, each column permuted once through a staging buffer.
template <class S>
void apply_deferred_left(LU<S>& f, long NB, long mn) {
  const long M = f.M, nblk = (mn + NB - 1) / NB;
  if (nblk < 2) return;
  int* Q = static_cast<int*>(scratch(f.s, kSlotLaPerm, sizeof(int) * 2 * (nblk - 1) * M));
  int* P = Q + (nblk - 1) * M;
  S* T = static_cast<S*>(scratch(f.s, kSlotLaTemp, sizeof(S) * M * NB));
  hipLaunchKernelGGL(la_block_maps, dim3((unsigned)(nblk - 1)), dim3(256), 0, f.s, (int)M,
                     (int)NB, (int)mn, f.leaf_lists, Q);
  ATLG_LAUNCHED();
  hipLaunchKernelGGL(la_suffix_maps, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, f.s,
                     (int)M, (int)nblk, Q, P);
  ATLG_LAUNCHED();
  for (long j = 0; j + 1 < nblk; j++) {
    const long r = (j + 1) * NB;
    if (r >= M) break;
    hipLaunchKernelGGL(la_gather_rows<S>, dim3((unsigned)NB), dim3(256), 0, f.s, (int)M, (int)r,
                       f.A, f.lda, j * NB, P + j * M, T);
    ATLG_LAUNCHED();
    hipLaunchKernelGGL(la_copy_rows<S>, dim3((unsigned)NB), dim3(256), 0, f.s, (int)M, (int)r,
                       f.A, f.lda, j * NB, T);
    ATLG_LAUNCHED();
  }
}

// Right-looking LU by la_block(min(M,N))-wide block columns with one step of
// look-ahead.
// Block k is factored by ATLAS's recursion (getrf_rec) with its leaves'
// interchanges confined to the block; every leaf keeps its composed list.
// Step k then, on the caller's stream:
//   1. brings block k+1 up to date: block k's interchanges, U = L_kk^-1 A,
//      and the GEMM update (the critical path);
//   2. applies block k's interchanges to the columns right of block k+1, and
//      their TRSM and GEMM update (the columns left of block k receive all
//      later interchanges at the end, composed into one row map per block);
//   3. hands block k+1 to the auxiliary stream, which factors it while 2
//      runs (it waits for 1 only);
//   4. waits for the panel before step k+1.
// A block of at least two 128-blocks is factored as the recursion's top node
// written out: left half, its TRSM and GEMM on the right half, right half,
// with each half's leaf interchanges confined to the half.  As soon as the
// left half is done, the caller's stream applies the left half's share of
// step 1 for the next block (interchanges, TRSM and a GEMM with K = half)
// while the right half is factored, so that only the right half's share
// waits for the whole panel; the left half receives the right half's
// interchanges afterwards, before step 2 reads the block.  While step 2 is
// the longer path, its GEMM leaves its last columns (la_chunk) to the panel
// stream, which updates them after the panel; step k+1's step 2 waits for
// them.
// The columns each stream touches are disjoint while both run.  Partial
// pivoting picks from the same updated columns as the plain recursion, so
// the pivots agree wherever they are well separated.
template <class S>
void getrf_lookahead(LU<S>& f0) {
  const auto lock = aux_lock();
  hipStream_t s0 = f0.s, s1 = aux_stream();
  hipEvent_t e_upd = aux_event(0), e_panel = aux_event(1);
  // With ATLG_LA_PANEL_CUS > 0 the panels run on a stream confined to that
  // many compute units and everything else on a stream confined to the
  // rest, so the latency-bound panel never shares a SIMD with the GEMM; the
  // caller's stream forks to the pair and joins at the end.
  hipStream_t caller = s0, sp = nullptr, sg = nullptr;
  const bool split = ATLG_LA_PANEL_CUS > 0 && split_streams(ATLG_LA_PANEL_CUS, &sp, &sg);
  if (split) {
    ATLG_HIP(hipEventRecord(aux_event(2), caller));
    ATLG_HIP(hipStreamWaitEvent(sg, aux_event(2), 0));
    ATLG_HIP(hipStreamWaitEvent(sp, aux_event(2), 0));
    s0 = sg;
    s1 = sp;
    f0.s = sg;
  }
  const long M = f0.M, N = f0.N, mn = std::min(M, N);
  const long nleaf = (mn + PW - 1) / PW;
  int* lists = static_cast<int*>(scratch(s0, kSlotLaLists, sizeof(int) * kLeafListInts * nleaf));
  LU<S> f1;
  init_lu(f1, s1, f0.M, f0.N, f0.A, f0.lda, f0.ipiv, f0.dinfo, f0.gemm_prec);
  f1.status = f0.status;  // failures are reported on the caller's stream
  f1.limit = f0.limit;
  f1.coop = ATLG_LA_COOP != 0;
  f1.small_lds = true;
  f0.leaf_lists = lists;
  f1.leaf_lists = lists;
  // block 0 on the caller's stream
  f0.win_lo = 0;
  const long kLaNB = la_block(mn);
  f0.win_hi = std::min(kLaNB, mn);
  getrf_rec(f0, 0, f0.win_hi);
  hipEvent_t e_left = aux_event(4), e_trsm = aux_event(5), e_chunk = aux_event(6);
  long hk = 0;  // split column of block k within it (0: factored whole)
  bool chunk_pending = false;  // the panel stream updates columns [N - chunk, N)
  for (long k0 = 0; k0 < mn; k0 += kLaNB) {
    const long kb = std::min(kLaNB, mn - k0), k1 = k0 + kb;
    const long nb1 = std::min(kLaNB, mn - k1);  // next block (none after the last)
    if (nb1 > 0) {
      if (hk > 0) {
        // the left half's share of this update ran beside the right half's
        // panel (below); the right half's remains
        const long kr = k0 + hk, wr = kb - hk;
        apply_block_lists(f0, kr, wr, k1, nb1, 0, 0);
        node_trsm_cols(f0, kr, wr, k1, nb1);
        node_update(f0, M - k1, nb1, wr, at(f0, k1, kr), at(f0, kr, k1), at(f0, k1, k1));
      } else {
        apply_block_lists(f0, k0, kb, k1, nb1, 0, 0);
        node_trsm_cols(f0, k0, kb, k1, nb1);
        node_update(f0, M - k1, nb1, kb, at(f0, k1, k0), at(f0, k0, k1), at(f0, k1, k1));
      }
      ATLG_HIP(hipEventRecord(e_upd, s0));
    }
    // a split block's left half has not seen its right half's interchanges
    if (hk > 0) apply_block_lists(f0, k0 + hk, kb - hk, k0, hk, 0, 0);
    // The rest of step k is queued before the panel's many small launches,
    // so s0 has work while the host issues them.  While the update is the
    // longer path, its last `chunk` columns go to the panel stream, after
    // the next panel.
    const long r0 = k1 + nb1, rest = N - r0;
    const long nb2n = std::min(kLaNB, std::max(0L, mn - r0));
    const long chunk = (split && nb1 > 0 && M > k1) ? la_chunk<S>(rest, M - k1, kb, nb1, nb2n) : 0;
    if (chunk_pending) {  // the previous step's share on the panel stream
      ATLG_HIP(hipStreamWaitEvent(s0, e_chunk, 0));
      chunk_pending = false;
    }
    apply_block_lists(f0, k0, kb, r0, rest, 0, 0);
    if (rest > 0) {
      node_trsm_cols(f0, k0, kb, r0, rest);
      if (chunk > 0) ATLG_HIP(hipEventRecord(e_trsm, s0));
      if (M > k1) {
        f0.gemm_cap = !split && (M - k1) * rest < ATLG_LA_CAP_AREA ? ATLG_LA_GEMM_CAP : 0;
        node_update(f0, M - k1, rest - chunk, kb, at(f0, k1, k0), at(f0, k0, r0),
                    at(f0, k1, r0));
        f0.gemm_cap = 0;
      }
    }
    hk = 0;
    if (nb1 > 0) {
      const long h1 = la_split(nb1), k2 = k1 + nb1, nb2 = std::min(kLaNB, mn - k2);
      ATLG_HIP(hipStreamWaitEvent(s1, e_upd, 0));
      if (h1 > 0) {
        // the recursion's top node by hand, each half's leaves confined to
        // the half, so that the left half's L columns stay as they are
        // while the right half is factored
        f1.win_lo = k1;
        f1.win_hi = k1 + h1;
        getrf_rec(f1, k1, h1);
        ATLG_HIP(hipEventRecord(e_left, s1));
        apply_block_lists(f1, k1, h1, k1 + h1, nb1 - h1, 0, 0);
        node_trsm_cols(f1, k1, h1, k1 + h1, nb1 - h1);
        node_update(f1, M - k1 - h1, nb1 - h1, h1, at(f1, k1 + h1, k1), at(f1, k1, k1 + h1),
                    at(f1, k1 + h1, k1 + h1));
        f1.win_lo = k1 + h1;
        f1.win_hi = k1 + nb1;
        getrf_rec(f1, k1 + h1, nb1 - h1);
      } else {
        f1.win_lo = k1;
        f1.win_hi = k1 + nb1;
        getrf_rec(f1, k1, nb1);
      }
      ATLG_HIP(hipEventRecord(e_panel, s1));
      if (chunk > 0) {
        const long c0 = N - chunk;
        ATLG_HIP(hipStreamWaitEvent(s1, e_trsm, 0));
        node_update(f1, M - k1, chunk, kb, at(f1, k1, k0), at(f1, k0, c0), at(f1, k1, c0));
        ATLG_HIP(hipEventRecord(e_chunk, s1));
        chunk_pending = true;
      }
      if (h1 > 0 && nb2 > 0) {
        // the left half's share of the next block's update, beside the
        // right half's panel (s0 has finished the rest of step k)
        ATLG_HIP(hipStreamWaitEvent(s0, e_left, 0));
        apply_block_lists(f0, k1, h1, k2, nb2, 0, 0);
        node_trsm_cols(f0, k1, h1, k2, nb2);
        node_update(f0, M - k1 - h1, nb2, h1, at(f0, k1 + h1, k1), at(f0, k1, k2),
                    at(f0, k1 + h1, k2));
      }
      ATLG_HIP(hipStreamWaitEvent(s0, e_panel, 0));
      hk = h1;
    }
  }
  if (chunk_pending) ATLG_HIP(hipStreamWaitEvent(s0, e_chunk, 0));
  apply_deferred_left(f0, kLaNB, mn);
  if (split) {  // every panel and panel-side update was joined into s0 above
    ATLG_HIP(hipEventRecord(aux_event(3), s0));
    ATLG_HIP(hipStreamWaitEvent(caller, aux_event(3), 0));
    f0.s = caller;
  }
}

}  // namespace

template <class S>
void getrf(hipStream_t s, long M, long N, S* A, long lda, int* ipiv, int* dinfo, int gemm_prec) {
  set_int(s, dinfo, 0);
  if (M == 0 || N == 0) return;
  LU<S> f;
  init_lu(f, s, M, N, A, lda, ipiv, dinfo, gemm_prec);
  f.status = status_word(s);
  f.limit = wait_limit(kSpinLimit);
  if (ATLG_LU_LOOKAHEAD && std::min(M, N) >= kLaMin)
    getrf_lookahead(f);
  else
    getrf_rec(f, 0, N);
}

template <class S>
void getrs(hipStream_t s, int trans, long N, long NRHS, const S* A, long lda, const int* ipiv,
           S* B, long ldb) {
  if (N == 0 || NRHS == 0) return;
  if (!is_trans(trans)) {
    laswp<S>(s, NRHS, B, ldb, 0, (int)N, ipiv, 1);
    trsm<S>(s, kLeft, kLower, kNoTrans, kUnit, N, NRHS, S(1), A, lda, B, ldb);
    trsm<S>(s, kLeft, kUpper, kNoTrans, kNonUnit, N, NRHS, S(1), A, lda, B, ldb);
  } else {
    trsm<S>(s, kLeft, kUpper, kTrans, kNonUnit, N, NRHS, S(1), A, lda, B, ldb);
    trsm<S>(s, kLeft, kLower, kTrans, kUnit, N, NRHS, S(1), A, lda, B, ldb);
    laswp<S>(s, NRHS, B, ldb, 0, (int)N, ipiv, -1);
  }
}

template void laswp<double>(hipStream_t, long, double*, long, int, int, const int*, int);
template void laswp<float>(hipStream_t, long, float*, long, int, int, const int*, int);
template void getrf<double>(hipStream_t, long, long, double*, long, int*, int*, int);
template void getrf<float>(hipStream_t, long, long, float*, long, int*, int*, int);
template void getrs<double>(hipStream_t, int, long, long, const double*, long, const int*,
                            double*, long);
template void getrs<float>(hipStream_t, int, long, long, const float*, long, const int*, float*,
                           long);

}  // namespace atlg
