// Triangular solve with many right-hand sides on the matrix cores (FP64, and
// FP32 for the single-precision factorization of the mixed solver).
//
// Replaces ATL_dtrsm (AtlasBase/Clint/trsm-micro.base:200-420: recursive
// halving trsm_recL/R, then iptrsm/trsmL_LNUT/utrsm with reciprocal
// diagonals, :550-1700, and the nano-kernels of
// AtlasBase/Students/trsm-nano.base:297-344).  ATLAS keeps reciprocal
// diagonals and runs everything off the diagonal as amm GEMM; this version
// goes one step further:
//   1. one launch inverts every NB x NB diagonal block of op(A) in LDS
//      (one workgroup per block: 8x8 register inverses, then recursive
//      doubling), writing full NB x NB inverses;
//   2. a recursive halving driver (the same split as trsm_recL/R) applies
//      each inverse to its block row/column as an in-place GEMM and updates
//      the rest with large GEMMs.
// All flops but the small inversions run on the 16x16x4 MFMA GEMM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "atlg_internal.hpp"
#include "tri_inv.hpp"

namespace atlg {
namespace {

constexpr int NB = kTrsmNB;    // diagonal block order (<= GEMM tile edge)
constexpr int TRI_THREADS = kTriThreads;
static_assert(NB == kTriN, "TRSM blocks are inverted by the 128-block inverter");

// Element (i, j) of op(A).
template <class S>
__device__ inline S opa(const S* A, long lda, bool trans, long i, long j) {
  return trans ? A[j + i * lda] : A[i + j * lda];
}

// Inverts the diagonal blocks of op(A) (order na).  Block b covers rows and
// columns [b*NB, b*NB + nbb).  An upper-triangular op(A) block U is handled
// as the lower-triangular J U J (J = order reversal), so one in-place lower
// inversion serves all cases.  Output: Dinv + b*NB*NB, column-major, ld NB,
// zeros outside the triangle; Unit diagonals are never read.  NT is the
// order held in LDS: NB for the general case; 32 or 64 when na <= NT (one
// block), which then writes only the leading NT x NT part of the block (the
// solves read the leading na x na part) and needs a fraction of the LDS, so
// the launch fits beside GEMM workgroups on a busy device.
template <class R, int NT>
__global__ void __launch_bounds__(TRI_THREADS)
trtri_diag_kernel(int na, const R* __restrict__ A, long lda, int trans, int lower, int unit,
                  R* __restrict__ Dinv, const int* __restrict__ skip) {
  if (skip && *skip) return;
  constexpr int LD = NT + 1;
  __shared__ R S[NT * LD];
  const int b = blockIdx.x, t = threadIdx.x;
  const long o = (long)b * NB;
  const int n = (int)min((long)NT, na - o);
  const bool tr = trans != 0;
  // the fill goes in batches of 16 elements per thread whose loads are all
  // issued before any of them is stored to LDS (one memory latency per
  // batch instead of one per element)
  constexpr int PER = NT * NT / TRI_THREADS, BATCH = PER < 16 ? PER : 16;
  for (int b0 = 0; b0 < PER; b0 += BATCH) {
    R v[BATCH];
#pragma unroll
    for (int u = 0; u < BATCH; u++) {
      const int idx = t + (b0 + u) * TRI_THREADS;
      const int i = idx % NT, j = idx / NT;
      R x = R(0);
      if (i >= j) {
        if (i < n && j < n) {
          if (i == j && unit) x = R(1);
          else x = lower ? opa(A, lda, tr, o + i, o + j)
                         : opa(A, lda, tr, o + n - 1 - i, o + n - 1 - j);
        } else if (i == j) {
          x = R(1);
        }
      }
      v[u] = x;
    }
#pragma unroll
    for (int u = 0; u < BATCH; u++) {
      const int idx = t + (b0 + u) * TRI_THREADS;
      S[(idx % NT) * LD + idx / NT] = v[u];
    }
  }
  __syncthreads();
  lower_inverse_inplace<R, NT>(S, n);
  R* D = Dinv + (size_t)b * NB * NB;
  for (int idx = t; idx < NT * NT; idx += TRI_THREADS) {
    const int i = idx % NT, j = idx / NT;
    R v = R(0);
    if (i < n && j < n) {
      if (lower) v = (i >= j) ? S[i * LD + j] : R(0);
      else v = (i <= j) ? S[(n - 1 - i) * LD + (n - 1 - j)] : R(0);
    }
    D[i + (size_t)j * NB] = v;
  }
}

template <class S>
struct Ctx {
  hipStream_t s;
  bool left, lower_eff, trans;
  const S* A;
  long lda;
  S* B;
  long ldb;
  long M, N;   // B is M x N
  const S* Dinv;
  int dtrans;  // kTrans when Dinv holds the transposes of inv(op(A)_bb)
  const int* skip;
};

// op(A)[i0:, j0:] as a GEMM operand: pointer and transpose flag.
template <class S>
inline void opa_block(const Ctx<S>& c, long i0, long j0, const S** p, int* t) {
  if (!c.trans) {
    *p = c.A + i0 + j0 * c.lda;
    *t = kNoTrans;
  } else {
    *p = c.A + j0 + i0 * c.lda;
    *t = kTrans;
  }
}

long split(long n) {
  long n1 = ((n / 2) / NB) * NB;
  return n1 > 0 ? n1 : NB;
}

// Solves on rows/cols [r0, r0+n) of op(A) with scale alpha.
template <class S>
void solve(const Ctx<S>& c, long r0, long n, S alpha) {
  const S one = S(1), mone = S(-1), zero = S(0);
  if (n <= NB) {
    const S* D = c.Dinv + (size_t)(r0 / NB) * NB * NB;
    if (c.left) {
      S* Bb = c.B + r0;
      gemm<S>(c.s, c.dtrans, kNoTrans, n, c.N, n, alpha, D, NB, Bb, c.ldb, zero, Bb, c.ldb, 0,
              c.skip);
    } else {
      S* Bb = c.B + r0 * c.ldb;
      gemm<S>(c.s, kNoTrans, c.dtrans, c.M, n, n, alpha, Bb, c.ldb, D, NB, zero, Bb, c.ldb, 0,
              c.skip);
    }
    return;
  }
  const long n1 = split(n), n2 = n - n1;
  const S* Ap;
  int ta;
  if (c.left) {
    S* B1 = c.B + r0;
    S* B2 = c.B + r0 + n1;
    if (c.lower_eff) {
      // [L11 0; L21 L22]: X1 first, then B2 <- alpha*B2 - L21*X1
      solve(c, r0, n1, alpha);
      opa_block(c, r0 + n1, r0, &Ap, &ta);
      gemm<S>(c.s, ta, kNoTrans, n2, c.N, n1, mone, Ap, c.lda, B1, c.ldb, alpha, B2, c.ldb, 0,
           c.skip);
      solve(c, r0 + n1, n2, one);
    } else {
      // [U11 U12; 0 U22]: X2 first, then B1 <- alpha*B1 - U12*X2
      solve(c, r0 + n1, n2, alpha);
      opa_block(c, r0, r0 + n1, &Ap, &ta);
      gemm<S>(c.s, ta, kNoTrans, n1, c.N, n2, mone, Ap, c.lda, B2, c.ldb, alpha, B1, c.ldb, 0,
           c.skip);
      solve(c, r0, n1, one);
    }
  } else {
    S* B1 = c.B + r0 * c.ldb;
    S* B2 = c.B + (r0 + n1) * c.ldb;
    if (c.lower_eff) {
      // X [L11 0; L21 L22] = B: X2 first, then B1 <- alpha*B1 - X2*L21
      solve(c, r0 + n1, n2, alpha);
      opa_block(c, r0 + n1, r0, &Ap, &ta);
      gemm<S>(c.s, kNoTrans, ta, c.M, n1, n2, mone, B2, c.ldb, Ap, c.lda, alpha, B1, c.ldb, 0,
           c.skip);
      solve(c, r0, n1, one);
    } else {
      // X [U11 U12; 0 U22] = B: X1 first, then B2 <- alpha*B2 - X1*U12
      solve(c, r0, n1, alpha);
      opa_block(c, r0, r0 + n1, &Ap, &ta);
      gemm<S>(c.s, kNoTrans, ta, c.M, n2, n1, mone, B1, c.ldb, Ap, c.lda, alpha, B2, c.ldb, 0,
           c.skip);
      solve(c, r0 + n1, n2, one);
    }
  }
}

// ---- few right-hand sides: one sync-free cooperative launch ----------------
//
// op(A) = A (Left, NoTrans) with at most 4 right-hand sides.  The rows are cut
// into 128-row blocks; workgroup w solves the blocks w, w+G, w+2G, ... in
// solve order (top down for lower, bottom up for upper).  For its block it
// accumulates A(block, q) x_q over the earlier blocks q as soon as each x_q
// is published (a ready flag per block), then applies the block's inverse
// from tri_inverse_blocks and publishes x.  The smallest unfinished block
// always has all its inputs, so the launch cannot deadlock while every
// workgroup is resident (cooperative launch); the waits are bounded anyway:
// an overrun sets the launch's error word, which makes every workgroup
// leave, and ORs ATLG_FAIL_TRSV into the stream's failure word.  Published x and the flags go through
// agent-scope (coherence point) atomic loads and stores, as in getrf's
// cooperative panel.
constexpr int SV_T = 256;                  // threads: 2 per block row
constexpr long kTrsvMin = 1024;            // shortest triangle that takes this path
constexpr long kSvSpin = 1L << 22;         // default polls (round trips) before giving up

template <class S>
__device__ inline S ld_co(const S* p);
template <>
__device__ inline double ld_co<double>(const double* p) {
  const unsigned long long u = __hip_atomic_load(
      reinterpret_cast<unsigned long long*>(const_cast<double*>(p)), __ATOMIC_RELAXED,
      __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double(static_cast<long long>(u));
}
template <>
__device__ inline float ld_co<float>(const float* p) {
  return __uint_as_float(__hip_atomic_load(reinterpret_cast<unsigned*>(const_cast<float*>(p)),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline void st_co(double* p, double x) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                     static_cast<unsigned long long>(__double_as_longlong(x)), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_co(float* p, float x) {
  __hip_atomic_store(reinterpret_cast<unsigned*>(p), __float_as_uint(x), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_flag(const int* p) {
  return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_flag(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Publication of a solved block.  Everything that crosses workgroups (the
// solved x values and the flags) is accessed only with agent-scope atomic
// loads and stores, which gfx950 performs coherently at the device level
// (the sc1 form: stores write through, loads do not take a stale per-XCD L2
// copy).  What the protocol needs is order: every thread waits until its
// own x stores are acknowledged (s_waitcnt 0) before the workgroup barrier,
// the flag is stored after the barrier, and a reader issues its x loads
// only after its flag load has returned 1 (and a barrier).  Agent-scope
// release/acquire fences would give the same order but also write back and
// invalidate the whole L2 of the XCD at every block (buffer_wbl2 /
// buffer_inv), which made each solve several times slower.

__global__ void sv_reset(int* flags, int n) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    flags[e] = 0;
}

template <class S, int R, bool TR>
__global__ void __launch_bounds__(SV_T)
trsv_sf_kernel(int n, int nb, int upper, S alpha, const S* __restrict__ A, long lda,
               const S* __restrict__ Dinv, S* __restrict__ B, long ldb, int* __restrict__ flags,
               int* __restrict__ err, int* __restrict__ status, long limit) {
  __shared__ S part[2][NB][R];
  __shared__ S xs[NB][R];
  __shared__ int ok;
  const int t = threadIdx.x, r = t & (NB - 1), h = t >> 7;
  for (int li = blockIdx.x; li < nb; li += gridDim.x) {
    const int p = upper ? nb - 1 - li : li;
    const long r0 = (long)p * NB;
    const int rows = (int)min((long)NB, n - r0);
    S acc[R];
#pragma unroll
    for (int k = 0; k < R; k++) acc[k] = S(0);
    for (int lj = 0; lj < li; lj++) {
      const int q = upper ? nb - 1 - lj : lj;
      const long c0 = (long)q * NB;
      const int cols = (int)min((long)NB, n - c0);
      // this thread's 64 entries of op(A)(block p, block q): issued before
      // the wait, they do not depend on x_q (TR: op(A) = A^T, the thread's
      // entries are contiguous in A)
      S va[64];
      {
        const int cend = (r < rows) ? min(64, cols - h * 64) : 0;
        if (TR) {
          const S* a = A + (c0 + h * 64) + (size_t)(r0 + r) * lda;
#pragma unroll
          for (int c = 0; c < 64; c++) va[c] = (c < cend) ? a[c] : S(0);
        } else {
          const S* a = A + (r0 + r) + (size_t)(c0 + h * 64) * lda;
#pragma unroll
          for (int c = 0; c < 64; c++) va[c] = (c < cend) ? a[(size_t)c * lda] : S(0);
        }
      }
      if (t == 0) {
        // one round trip per poll; the error word only every 256 polls
        bool good = true;
        long spins = 0;
        while (ld_flag(&flags[q]) == 0) {
          if (++spins >= limit || ((spins & 255) == 0 && ld_flag(err))) {
            st_flag(err, 1);
            atomicOr(status, ATLG_FAIL_TRSV);
            good = false;
            break;
          }
        }
        ok = good;
      }
      __syncthreads();  // x_q is read (ld_co) only after the flag was seen
      if (!ok) return;  // same value in every thread
      if (t < NB) {
#pragma unroll
        for (int k = 0; k < R; k++) xs[t][k] = (t < cols) ? ld_co(&B[c0 + t + (size_t)k * ldb]) : S(0);
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 64; c++)
#pragma unroll
        for (int k = 0; k < R; k++) acc[k] += va[c] * xs[h * 64 + c][k];
      __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < R; k++) part[h][r][k] = acc[k];
    __syncthreads();
    if (t < NB) {
#pragma unroll
      for (int k = 0; k < R; k++)
        xs[t][k] = (t < rows) ? alpha * B[r0 + t + (size_t)k * ldb] - part[0][t][k] - part[1][t][k]
                              : S(0);
    }
    __syncthreads();
    // x = inv(A_pp) (alpha b - acc): thread (r, h) sums 64 of the 128 columns
    const S* D = Dinv + (size_t)p * NB * NB + r + (size_t)h * 64 * NB;
    S y[R];
#pragma unroll
    for (int k = 0; k < R; k++) y[k] = S(0);
#pragma unroll
    for (int c0 = 0; c0 < 64; c0 += 16) {
      S d[16];
#pragma unroll
      for (int c = 0; c < 16; c++) d[c] = D[(size_t)(c0 + c) * NB];
#pragma unroll
      for (int c = 0; c < 16; c++)
#pragma unroll
        for (int k = 0; k < R; k++) y[k] += d[c] * xs[h * 64 + c0 + c][k];
    }
#pragma unroll
    for (int k = 0; k < R; k++) part[h][r][k] = y[k];
    __syncthreads();
    if (t < rows) {
#pragma unroll
      for (int k = 0; k < R; k++) st_co(&B[r0 + t + (size_t)k * ldb], part[0][t][k] + part[1][t][k]);
    }
    // every thread's x stores are acknowledged before the barrier, and the
    // flag is stored after it
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t == 0) st_flag(&flags[p], 1);
  }
}

// Few-RHS Left solve with op(A) = A (tr false) or A^T (tr true), `lower`
// saying whether op(A) is lower triangular, and Dinv holding the inverses
// of op(A)'s diagonal blocks; returns false when the shape is not handled.
template <class S>
bool trsv_few(hipStream_t s, bool lower, bool tr, long n, long nrhs, S alpha, const S* A,
              long lda, const S* Dinv, S* B, long ldb) {
  if (nrhs < 1 || nrhs > 4) return false;
  const int nb = (int)((n + NB - 1) / NB);
  int* flags = static_cast<int*>(scratch(s, kSlotTrsv, sizeof(int) * (nb + 64)));
  int* err = flags + nb;
  hipLaunchKernelGGL(sv_reset, dim3((unsigned)std::min(64, (nb + 64 + 255) / 256)), dim3(256), 0,
                     s, flags, nb + 1);
  ATLG_LAUNCHED();
  int dev = 0, cus = 0;
  ATLG_HIP(hipGetDevice(&dev));
  ATLG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int nn = (int)n, nbk = nb, up = lower ? 0 : 1;
  int G = std::max(1, std::min(nb, cus));
  int* status = status_word(s);
  long limit = wait_limit(kSvSpin);
  void* args[] = {&nn, &nbk, &up, &alpha, &A, &lda, &Dinv, &B, &ldb, &flags, &err, &status, &limit};
  const void* fn = nullptr;
  switch (nrhs + (tr ? 4 : 0)) {
    case 1: fn = reinterpret_cast<const void*>(&trsv_sf_kernel<S, 1, false>); break;
    case 2: fn = reinterpret_cast<const void*>(&trsv_sf_kernel<S, 2, false>); break;
    case 3: fn = reinterpret_cast<const void*>(&trsv_sf_kernel<S, 3, false>); break;
    case 4: fn = reinterpret_cast<const void*>(&trsv_sf_kernel<S, 4, false>); break;
    case 5: fn = reinterpret_cast<const void*>(&trsv_sf_kernel<S, 1, true>); break;
    case 6: fn = reinterpret_cast<const void*>(&trsv_sf_kernel<S, 2, true>); break;
    case 7: fn = reinterpret_cast<const void*>(&trsv_sf_kernel<S, 3, true>); break;
    default: fn = reinterpret_cast<const void*>(&trsv_sf_kernel<S, 4, true>); break;
  }
  ATLG_HIP(hipLaunchCooperativeKernel(fn, dim3(G), dim3(SV_T), args, 0, s));
  return true;
}

}  // namespace

template <class S>
void tri_inverse_blocks(hipStream_t s, long na, const S* A, long lda, int ta, int uplo, int diag,
                        S* Dinv, const int* skip) {
  if (na <= 0) return;
  const bool tr = is_trans(ta);
  const long nblk = (na + NB - 1) / NB;
  const int trn = tr ? 1 : 0, low = (uplo == kLower) != tr ? 1 : 0, un = diag == kUnit ? 1 : 0;
  if (na <= 32)
    hipLaunchKernelGGL((trtri_diag_kernel<S, 32>), dim3(1), dim3(TRI_THREADS), 0, s, (int)na, A,
                       lda, trn, low, un, Dinv, skip);
  else if (na <= 64)
    hipLaunchKernelGGL((trtri_diag_kernel<S, 64>), dim3(1), dim3(TRI_THREADS), 0, s, (int)na, A,
                       lda, trn, low, un, Dinv, skip);
  else
    hipLaunchKernelGGL((trtri_diag_kernel<S, NB>), dim3((unsigned)nblk), dim3(TRI_THREADS), 0, s,
                       (int)na, A, lda, trn, low, un, Dinv, skip);
  ATLG_LAUNCHED();
}

template <class S>
void trsm(hipStream_t s, int side, int uplo, int ta, int diag, long M, long N, S alpha,
          const S* A, long lda, S* B, long ldb, const int* skip, const S* dinv, int dinv_trans) {
  if (M <= 0 || N <= 0) return;
  if (alpha == S(0)) {
    scale<S>(s, M, N, S(0), B, ldb, 0, skip);
    return;
  }
  Ctx<S> c;
  c.s = s;
  c.left = (side == kLeft);
  c.trans = is_trans(ta);
  c.lower_eff = ((uplo == kLower) != c.trans);
  c.A = A;
  c.lda = lda;
  c.B = B;
  c.ldb = ldb;
  c.M = M;
  c.N = N;
  c.skip = skip;
  const long na = c.left ? M : N;
  c.dtrans = dinv ? (dinv_trans ? kTrans : kNoTrans) : kNoTrans;
  if (dinv) {
    c.Dinv = dinv;
    if (c.left && !dinv_trans && !skip && na >= kTrsvMin &&
        trsv_few<S>(s, c.lower_eff, c.trans, na, N, alpha, A, lda, dinv, B, ldb))
      return;
  } else {
    const long nblk = (na + NB - 1) / NB;
    S* D = static_cast<S*>(scratch(s, kSlotTrsmInv, sizeof(S) * nblk * NB * NB));
    tri_inverse_blocks<S>(s, na, A, lda, ta, uplo, diag, D, skip);
    c.Dinv = D;
    // a few right-hand sides on a long triangle: one sync-free launch
    if (c.left && !skip && na >= kTrsvMin &&
        trsv_few<S>(s, c.lower_eff, c.trans, na, N, alpha, A, lda, D, B, ldb))
      return;
  }
  solve(c, 0, na, alpha);
}

void* noalloc_scratch();

// The allocation-free form: the diagonal-block inverses go to the file-scope
// device array (at most kNoAllocMaxOrder / kTrsmNB blocks), and the solve is
// the blocked recursion of trsm() (no few-RHS launch, whose ready flags and
// failure word would need memory of their own).
template <class S>
void trsm_noalloc(hipStream_t s, int side, int uplo, int ta, int diag, long M, long N, S alpha,
                  const S* A, long lda, S* B, long ldb) {
  if (M <= 0 || N <= 0) return;
  if (alpha == S(0)) {
    scale<S>(s, M, N, S(0), B, ldb, 0, nullptr);
    return;
  }
  const long na = side == kLeft ? M : N;
  if (na > kNoAllocMaxOrder) throw HipFailure{"trsm_noalloc: order above the allocation-free maximum"};
  S* D = static_cast<S*>(noalloc_scratch());
  tri_inverse_blocks<S>(s, na, A, lda, ta, uplo, diag, D, nullptr);
  Ctx<S> c;
  c.s = s;
  c.left = (side == kLeft);
  c.trans = is_trans(ta);
  c.lower_eff = ((uplo == kLower) != c.trans);
  c.A = A;
  c.lda = lda;
  c.B = B;
  c.ldb = ldb;
  c.M = M;
  c.N = N;
  c.skip = nullptr;
  c.Dinv = D;
  c.dtrans = kNoTrans;
  solve(c, 0, na, alpha);
}

template void trsm_noalloc<double>(hipStream_t, int, int, int, int, long, long, double,
                                   const double*, long, double*, long);

template void tri_inverse_blocks<double>(hipStream_t, long, const double*, long, int, int, int,
                                         double*, const int*);
template void tri_inverse_blocks<float>(hipStream_t, long, const float*, long, int, int, int,
                                        float*, const int*);
template void trsm<double>(hipStream_t, int, int, int, int, long, long, double, const double*,
                           long, double*, long, const int*, const double*, int);
template void trsm<float>(hipStream_t, int, int, int, int, long, long, float, const float*, long,
                          float*, long, const int*, const float*, int);

}  // namespace atlg

// ---- the device array behind atlg_dtrsm_noalloc ---------------------------
// kNoAllocMaxOrder / kTrsmNB inverses of 128 x 128 doubles (8 MiB), one copy
// per device, held until the process ends.  Its name is only a host-side
// stand-in for the symbol: host code uses the address that
// hipGetSymbolAddress returns for the current device, never the name itself.
static __device__ __align__(256) unsigned char scratch[atlg::kNoAllocBytes];

namespace atlg {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

// Address of the array on the current device, resolved once per device
// (with that device current) and kept.
void* noalloc_scratch() {
  int dev = 0;
  ATLG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw HipFailure{"trsm_noalloc: device index out of range"};
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    ATLG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    g_noalloc_ptr[dev] = p;
  }
  return g_noalloc_ptr[dev];
}

}  // namespace atlg
