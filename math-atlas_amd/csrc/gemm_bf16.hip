// BF16 trailing update with FP32 accumulation for the low-precision LU of
// the mixed-precision solver: C(f32) += alpha * op16(A) * op16(B), where
// op16 rounds the FP32 operand to bfloat16 (round to nearest even).  No
// ATLAS counterpart: it stands in for the GEMM of ATL_sgetrf's update step
// (AtlasBase/Clint/atlas-lp.base:1497-1618) when the factor is computed with
// BF16 products.
//
// Split variant (NPROD = 6): each FP32 operand x is carried as three BF16
// terms, hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid), which
// hold all 24 bits of its significand, and the product is accumulated as
// hi*hi + hi*mid + mid*hi + mid*mid + hi*lo + lo*hi (the terms below FP32
// resolution are dropped: "bf16x6" of Henry, Tang and Heinecke, "Leveraging
// the bfloat16 artificial intelligence datatype for higher-precision
// computations", ARITH 2019).  Six BF16 MFMAs per tile step give FP32
// accuracy.  Two-term variant (NPROD = 3, "bf16x3" of the same paper):
// hi = bf16(x), lo = bf16(x - hi) (16 significand bits) and
// hi*hi + hi*lo + lo*hi, about 2^-15 relative error per product.
//
// The operands are first converted into K-contiguous BF16 copies (A
// transposed through LDS; B's columns are already K-contiguous), with K
// padded to the tile depth with zeros.  The GEMM then streams 16-byte
// vectors of 8 BF16 into double-buffered LDS tiles and runs
// v_mfma_f32_16x16x32_bf16: 128x128 C per 256-thread workgroup, 64x64 per
// wave, K consumed 64 at a time (32 for the split variants, whose two or
// three tiles per operand share the LDS).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "atlg_internal.hpp"

// No implicit contraction anywhere in this file: the split terms and the
// epilogue round exactly as written, whatever the compiler's defaults, so the
// factor (and whether refinement converges on it) does not depend on the
// build.
#pragma clang fp contract(off)

namespace atlg {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128, NT = 256;
constexpr int WM = 64, WN = 64, TM = WM / 16, TN = WN / 16;
constexpr int KPAD = 64;     // K of the converted copies is a multiple of this
// LDS row padding (BF16 elements, a multiple of 8: rows stay 16-byte aligned)
// of the single and the split kernels.
#ifndef ATLG_BF16_PAD1
#define ATLG_BF16_PAD1 8
#endif
#ifndef ATLG_BF16_PAD6
#define ATLG_BF16_PAD6 8
#endif
static_assert(ATLG_BF16_PAD1 % 8 == 0 && ATLG_BF16_PAD6 % 8 == 0, "16-byte LDS rows");

// Register layout of v_mfma_f32_16x16x32_bf16: lane l supplies
// X1[l&15][8(l>>4) .. 8(l>>4)+7] and X2[8(l>>4) .. +7][l&15]; D[i][j] is in
// lane j + 16*(i>>2), register i&3.  The kernel passes the A tile as X1 and
// the B tile as X2, so lane l, register r carries C[m = 4(l>>4)+r][n = l&15]:
// four consecutive rows of one column, one 16-byte access of column-major C.

// dst[x][k] (k contiguous, pitch ldd, x < X, k < Kp) = bf16(src(x, k)) where
// src(x, k) = S[x + k*lds] (TRANS: x contiguous in the source) or
// S[k + x*lds]; zero for k >= K.
// With nterm > 1 also writes the remainders to D1 (and D2): two terms
// hi, bf16(x - hi); three terms hi, mid = bf16(x - hi), lo = bf16(x - hi -
// mid) (each difference is exact in FP32).
__device__ inline void split_store(float v, int nterm, __bf16* D0, __bf16* D1, __bf16* D2,
                                   size_t at) {
  const __bf16 h = (__bf16)v;
  D0[at] = h;
  if (nterm > 1) {
    const float r1 = v - (float)h;
    const __bf16 m = (__bf16)r1;
    D1[at] = m;
    if (nterm > 2) D2[at] = (__bf16)(r1 - (float)m);
  }
}

template <bool TRANS>
__global__ void __launch_bounds__(256)
to_bf16_kernel(long X, long K, long Kp, const float* __restrict__ S, long lds,
               __bf16* __restrict__ D, long ldd, __bf16* __restrict__ D1,
               __bf16* __restrict__ D2, int nterm) {
  if (TRANS) {
    __shared__ float tile[64][65];
    const long x0 = (long)blockIdx.x * 64, k0 = (long)blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    for (int r = ty; r < 64; r += 4) {
      const long x = x0 + tx, k = k0 + r;
      tile[r][tx] = (x < X && k < K) ? S[x + k * lds] : 0.0f;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
      const long x = x0 + r, k = k0 + tx;
      if (x < X && k < Kp) split_store(tile[tx][r], nterm, D, D1, D2, x * ldd + k);
    }
  } else {
    const long x = blockIdx.x;
    for (long k = threadIdx.x; k < Kp; k += blockDim.x)
      split_store((k < K) ? S[k + x * lds] : 0.0f, nterm, D, D1, D2, x * ldd + k);
  }
}

// C[m][n] = alpha * sum_k At[m][k] * Bt[n][k] + beta * C[m][n]; with
// NPROD = 6 each operand comes as hi/mid/lo copies (A0/A1/A2, B0/B1/B2),
// with NPROD = 3 as hi/lo copies (A0/A1, B0/B1).
// VEC: C is 16-byte aligned and ldc a multiple of 4 (vector C accesses).
// NPROD = 1: double-buffered LDS, the C tile read before the K loop (its
// latency overlaps the MFMAs).  NPROD = 3, 6: one LDS buffer (40 / 60 KB)
// and C read at the end, which fit two workgroups per compute unit: one
// computes while the other waits at its barriers.
template <int NPROD, bool VEC>
__global__ void __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu(NPROD == 1 ? 1 : 2, 2)))
gemm_bf16_kernel(int M, int N, int Kp, float alpha, const __bf16* __restrict__ A0,
                 const __bf16* __restrict__ A1, const __bf16* __restrict__ A2, long lda,
                 const __bf16* __restrict__ B0, const __bf16* __restrict__ B1,
                 const __bf16* __restrict__ B2, long ldb, float beta, float* __restrict__ C,
                 long ldc) {
  static_assert(NPROD == 1 || NPROD == 3 || NPROD == 6, "product count");
  constexpr int NOP = NPROD == 1 ? 1 : (NPROD == 3 ? 2 : 3);  // operand copies per side
  constexpr int BK = (NPROD == 1) ? 64 : 32;
  constexpr int LDK = BK + (NPROD == 1 ? ATLG_BF16_PAD1 : ATLG_BF16_PAD6);  // LDS row pitch
  constexpr int VPR = BK / 8;                // 16-byte vectors per row
  constexpr int NV = BM * VPR / NT;          // vectors per thread per operand copy
  constexpr int NS = (NPROD == 1) ? 2 : 1;   // LDS stages
  constexpr bool PREC = NPROD == 1;          // C read before the K loop
  __shared__ __attribute__((aligned(16))) __bf16 As[NS][NOP][BM * LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[NS][NOP][BN * LDK];
  const int tilesM = (M + BM - 1) / BM;
  const int tm = blockIdx.x % tilesM, tn = blockIdx.x / tilesM;
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const __bf16* Ap[3] = {A0, A1, A2};
  const __bf16* Bp[3] = {B0, B1, B2};

  us8 ra[NOP][NV], rb[NOP][NV];
  auto gload = [&](int k0) {
#pragma unroll
    for (int o = 0; o < NOP; o++)
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int e = t + NT * q, row = e / VPR, kv = (e % VPR) * 8;
        const int gm = m0 + row, gn = n0 + row;
        ra[o][q] = (gm < M) ? *reinterpret_cast<const us8*>(Ap[o] + (size_t)gm * lda + k0 + kv)
                            : us8{0, 0, 0, 0, 0, 0, 0, 0};
        rb[o][q] = (gn < N) ? *reinterpret_cast<const us8*>(Bp[o] + (size_t)gn * ldb + k0 + kv)
                            : us8{0, 0, 0, 0, 0, 0, 0, 0};
      }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int o = 0; o < NOP; o++)
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int e = t + NT * q, row = e / VPR, kv = (e % VPR) * 8;
        *reinterpret_cast<us8*>(&As[buf][o][row * LDK + kv]) = ra[o][q];
        *reinterpret_cast<us8*>(&Bs[buf][o][row * LDK + kv]) = rb[o][q];
      }
  };

  f4 acc[TM][TN], cin[TM][TN];
  const int cm = 4 * (lane >> 4), cn = lane & 15;  // this lane's C rows cm..cm+3, column cn
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      acc[i][j] = f4{0, 0, 0, 0};
      cin[i][j] = f4{0, 0, 0, 0};
      if (PREC && beta != 0.0f) {
        const int gm = m0 + wm0 + i * 16 + cm, gn = n0 + wn0 + j * 16 + cn;
        const float* c = C + (size_t)gm + (size_t)gn * ldc;
        if (VEC && gm + 3 < M && gn < N) {
          cin[i][j] = *reinterpret_cast<const f4*>(c);
        } else if (gn < N) {
#pragma unroll
          for (int r = 0; r < 4; r++)
            if (gm + r < M) cin[i][j][r] = c[r];
        }
      }
    }

  const int nk = Kp / BK;
  gload(0);
  if (NS == 2) {
    lstore(0);
    __syncthreads();
  }
  const int fx = lane & 15, fk = (lane >> 4) * 8;
  for (int kt = 0; kt < nk; kt++) {
    const int cur = NS == 2 ? (kt & 1) : 0;
    const bool more = kt + 1 < nk;
    if (NS == 1) {  // the previous step's reads are done before the overwrite
      if (kt > 0) __syncthreads();
      lstore(0);
      __syncthreads();
    }
    if (more) gload((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      bf16x8 af[NOP][TM], bfr[NOP][TN];
#pragma unroll
      for (int o = 0; o < NOP; o++) {
#pragma unroll
        for (int i = 0; i < TM; i++)
          af[o][i] =
              *reinterpret_cast<const bf16x8*>(&As[cur][o][(wm0 + i * 16 + fx) * LDK + ks + fk]);
#pragma unroll
        for (int j = 0; j < TN; j++)
          bfr[o][j] =
              *reinterpret_cast<const bf16x8*>(&Bs[cur][o][(wn0 + j * 16 + fx) * LDK + ks + fk]);
      }
      // Products in the order small terms first: (six) hi*lo, lo*hi,
      // mid*mid, hi*mid, mid*hi, hi*hi; (three) hi*lo, lo*hi, hi*hi; (one)
      // hi*hi.  The product loop is outermost so consecutive MFMAs go to
      // different accumulators.
      constexpr int kPA6[6] = {0, 2, 1, 0, 1, 0}, kPB6[6] = {2, 0, 1, 1, 0, 0};
      constexpr int kPA3[3] = {0, 1, 0}, kPB3[3] = {1, 0, 0};
#pragma unroll
      for (int p = 0; p < NPROD; p++) {
        const int pa = NPROD == 6 ? kPA6[p] : (NPROD == 3 ? kPA3[p] : 0);
        const int pb = NPROD == 6 ? kPB6[p] : (NPROD == 3 ? kPB3[p] : 0);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[pa][i], bfr[pb][j], acc[i][j],
                                                                0, 0, 0);
      }
    }
    if (NS == 2) {
      if (more) lstore(cur ^ 1);
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int gm = m0 + wm0 + i * 16 + cm, gn = n0 + wn0 + j * 16 + cn;
      float* c = C + (size_t)gm + (size_t)gn * ldc;
      if (!PREC && beta != 0.0f) {
        if (VEC && gm + 3 < M && gn < N) {
          cin[i][j] = *reinterpret_cast<const f4*>(c);
        } else if (gn < N) {
#pragma unroll
          for (int r = 0; r < 4; r++)
            if (gm + r < M) cin[i][j][r] = c[r];
        }
      }
      f4 v;
      // rounding pinned: beta * c rounded, then one fused multiply-add
      // (the file is compiled with contraction off)
#pragma unroll
      for (int r = 0; r < 4; r++)
        v[r] = (beta == 0.0f) ? alpha * acc[i][j][r]
                              : __builtin_fmaf(alpha, acc[i][j][r], beta * cin[i][j][r]);
      if (VEC && gm + 3 < M && gn < N) {
        *reinterpret_cast<f4*>(c) = v;
      } else if (gn < N) {
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (gm + r < M) c[r] = v[r];
      }
    }
}

}  // namespace

void gemm_bf16(hipStream_t s, long M, long N, long K, float alpha, const float* A, long lda,
               const float* B, long ldb, float beta, float* C, long ldc, int terms) {
  if (M <= 0 || N <= 0) return;
  if (K <= 0 || alpha == 0.0f) {
    scale<float>(s, M, N, beta, C, ldc);
    return;
  }
  const long Kp = (K + KPAD - 1) / KPAD * KPAD;
  const size_t ba = ((size_t)M * Kp * sizeof(__bf16) + 255) & ~size_t(255);
  const size_t bb = ((size_t)N * Kp * sizeof(__bf16) + 255) & ~size_t(255);
  const int copies = terms;  // BF16 copies per operand: 1, 2 (hi, lo) or 3 (hi, mid, lo)
  char* w = static_cast<char*>(scratch(s, kSlotBf16, copies * (ba + bb)));
  __bf16* Ac[3];
  __bf16* Bc[3];
  for (int c = 0; c < 3; c++) {
    Ac[c] = reinterpret_cast<__bf16*>(w + (size_t)(c < copies ? c : 0) * (ba + bb));
    Bc[c] = reinterpret_cast<__bf16*>(w + (size_t)(c < copies ? c : 0) * (ba + bb) + ba);
  }
  hipLaunchKernelGGL(to_bf16_kernel<true>, dim3((unsigned)((M + 63) / 64), (unsigned)(Kp / 64)),
                     dim3(256), 0, s, M, K, Kp, A, lda, Ac[0], Kp, Ac[1], Ac[2], terms);
  ATLG_LAUNCHED();
  hipLaunchKernelGGL(to_bf16_kernel<false>, dim3((unsigned)N), dim3(256), 0, s, N, K, Kp, B, ldb,
                     Bc[0], Kp, Bc[1], Bc[2], terms);
  ATLG_LAUNCHED();
  const long tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  const bool vec = reinterpret_cast<uintptr_t>(C) % 16 == 0 && ldc % 4 == 0;
  auto kern = terms == 3   ? (vec ? gemm_bf16_kernel<6, true> : gemm_bf16_kernel<6, false>)
              : terms == 2 ? (vec ? gemm_bf16_kernel<3, true> : gemm_bf16_kernel<3, false>)
                           : (vec ? gemm_bf16_kernel<1, true> : gemm_bf16_kernel<1, false>);
  hipLaunchKernelGGL(kern, dim3((unsigned)tiles), dim3(NT), 0, s, (int)M, (int)N, (int)Kp, alpha,
                     Ac[0], Ac[1], Ac[2], Kp, Bc[0], Bc[1], Bc[2], Kp, beta, C, ldc);
  ATLG_LAUNCHED();
}

}  // namespace atlg
