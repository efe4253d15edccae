// BF16 trailing update with FP32 accumulation for the low-precision LU of
// the mixed-precision solver: C(f32) += alpha * op16(A) * op16(B), where
// op16 rounds the FP32 operand to bfloat16 (round to nearest even).  No
// ATLAS counterpart: it stands in for the GEMM of ATL_sgetrf's update step
// (AtlasBase/Clint/atlas-lp.base:1497-1618) when the factor is computed with
// BF16 products.
//
// The operands are first converted into K-contiguous BF16 copies (A
// transposed through LDS; B's columns are already K-contiguous), with K
// padded to the tile depth with zeros.  The GEMM then streams 16-byte
// vectors of 8 BF16 into double-buffered LDS tiles and runs
// v_mfma_f32_16x16x32_bf16: 128x128 C per 256-thread workgroup, 64x64 per
// wave, K consumed 64 at a time.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "atlg_internal.hpp"

namespace atlg {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128, BK = 64, NT = 256;
constexpr int WM = 64, WN = 64, TM = WM / 16, TN = WN / 16;
constexpr int LDK = BK + 8;  // LDS row pitch in BF16 (16-byte pad)
constexpr int KPAD = BK;     // K of the converted copies is a multiple of this

// Register layout of v_mfma_f32_16x16x32_bf16: lane l supplies
// X1[l&15][8(l>>4) .. 8(l>>4)+7] and X2[8(l>>4) .. +7][l&15]; D[i][j] is in
// lane j + 16*(i>>2), register i&3.  The kernel passes the B tile as X1 and
// the A tile as X2, so lane l, register r carries C[m = l&15][n = 4(l>>4)+r].

// dst[x][k] (k contiguous, pitch ldd, x < X, k < Kp) = bf16(src(x, k)) where
// src(x, k) = S[x + k*lds] (TRANS: x contiguous in the source) or
// S[k + x*lds]; zero for k >= K.
template <bool TRANS>
__global__ void __launch_bounds__(256)
to_bf16_kernel(long X, long K, long Kp, const float* __restrict__ S, long lds,
               __bf16* __restrict__ D, long ldd) {
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
      if (x < X && k < Kp) D[x * ldd + k] = (__bf16)tile[tx][r];
    }
  } else {
    const long x = blockIdx.x;
    for (long k = threadIdx.x; k < Kp; k += blockDim.x)
      D[x * ldd + k] = (k < K) ? (__bf16)S[k + x * lds] : (__bf16)0.0f;
  }
}

// C[m][n] = alpha * sum_k At[m][k] * Bt[n][k] + beta * C[m][n]
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_bf16_kernel(int M, int N, int Kp, float alpha, const __bf16* __restrict__ At, long lda,
                 const __bf16* __restrict__ Bt, long ldb, float beta, float* __restrict__ C,
                 long ldc) {
  __shared__ __attribute__((aligned(16))) __bf16 As[2][BM * LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][BN * LDK];
  const int tilesM = (M + BM - 1) / BM;
  const int tm = blockIdx.x % tilesM, tn = blockIdx.x / tilesM;
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;

  // each thread moves 4 vectors of 8 BF16 per operand tile: row = e >> 3,
  // k-vector = e & 7 for e = t + 256 q
  us8 ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = t + NT * q, row = e >> 3, kv = (e & 7) * 8;
      const int gm = m0 + row, gn = n0 + row;
      ra[q] = (gm < M) ? *reinterpret_cast<const us8*>(At + (size_t)gm * lda + k0 + kv)
                       : us8{0, 0, 0, 0, 0, 0, 0, 0};
      rb[q] = (gn < N) ? *reinterpret_cast<const us8*>(Bt + (size_t)gn * ldb + k0 + kv)
                       : us8{0, 0, 0, 0, 0, 0, 0, 0};
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = t + NT * q, row = e >> 3, kv = (e & 7) * 8;
      *reinterpret_cast<us8*>(&As[buf][row * LDK + kv]) = ra[q];
      *reinterpret_cast<us8*>(&Bs[buf][row * LDK + kv]) = rb[q];
    }
  };

  f4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = f4{0, 0, 0, 0};

  const int nk = Kp / BK;
  gload(0);
  lstore(0);
  __syncthreads();
  const int fx = lane & 15, fk = (lane >> 4) * 8;
  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) gload((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      bf16x8 af[TM], bfr[TN];
#pragma unroll
      for (int i = 0; i < TM; i++)
        af[i] = *reinterpret_cast<const bf16x8*>(&As[cur][(wm0 + i * 16 + fx) * LDK + ks + fk]);
#pragma unroll
      for (int j = 0; j < TN; j++)
        bfr[j] = *reinterpret_cast<const bf16x8*>(&Bs[cur][(wn0 + j * 16 + fx) * LDK + ks + fk]);
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j], af[i], acc[i][j], 0, 0, 0);
    }
    if (more) lstore(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gm = m0 + wm0 + i * 16 + (lane & 15);
        const int gn = n0 + wn0 + j * 16 + 4 * (lane >> 4) + r;
        if (gm < M && gn < N) {
          float* c = C + (size_t)gm + (size_t)gn * ldc;
          const float v = alpha * acc[i][j][r];
          *c = (beta == 0.0f) ? v : v + beta * (*c);
        }
      }
}

}  // namespace

void gemm_bf16(hipStream_t s, long M, long N, long K, float alpha, const float* A, long lda,
               const float* B, long ldb, float beta, float* C, long ldc) {
  if (M <= 0 || N <= 0) return;
  if (K <= 0 || alpha == 0.0f) {
    scale<float>(s, M, N, beta, C, ldc);
    return;
  }
  const long Kp = (K + KPAD - 1) / KPAD * KPAD;
  const size_t ba = ((size_t)M * Kp * sizeof(__bf16) + 255) & ~size_t(255);
  const size_t bb = (size_t)N * Kp * sizeof(__bf16);
  char* w = static_cast<char*>(scratch(s, kSlotBf16, ba + bb));
  __bf16* At = reinterpret_cast<__bf16*>(w);
  __bf16* Bt = reinterpret_cast<__bf16*>(w + ba);
  hipLaunchKernelGGL(to_bf16_kernel<true>, dim3((unsigned)((M + 63) / 64), (unsigned)(Kp / 64)),
                     dim3(256), 0, s, M, K, Kp, A, lda, At, Kp);
  ATLG_LAUNCHED();
  hipLaunchKernelGGL(to_bf16_kernel<false>, dim3((unsigned)N), dim3(256), 0, s, N, K, Kp, B, ldb,
                     Bt, Kp);
  ATLG_LAUNCHED();
  const long tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3((unsigned)tiles), dim3(NT), 0, s, (int)M, (int)N,
                     (int)Kp, alpha, At, Kp, Bt, Kp, beta, C, ldc);
  ATLG_LAUNCHED();
}

}  // namespace atlg
