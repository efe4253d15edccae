// DGEMM family for gfx950 on the FP64 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Replaces the work ATLAS does in ATL_dgemm/ATL_dammm (recursion + shape
// dispatch, AtlasBase/Clint/atlas-ammm.base:17266-17810), its copy-to-block
// and loop nests (ATL_ammmKMNK/KNMK, doAmmBlk, :29419-31100) and the amm
// microkernels (ammkern_t, :577): here one tiled kernel reads column-major
// operands straight from HBM into double-buffered LDS tiles and accumulates
// 128x128 C tiles in registers.  The same kernel, with a triangle mask,
// serves SYRK (ATL_dsyrk, AtlasBase/Clint/l3micro.base:2358-2720).
//
// Tile: 128x128 C per 256-thread workgroup (4 waves in 2x2), each wave a
// 64x64 sub-tile = 4x4 MFMA 16x16 tiles = 64 FP64 accumulators per lane.
// K is consumed 16 at a time (4 MFMA k-steps).  Global->LDS loads are 128-bit
// where the operand's base and leading dimension allow it.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "atlg_internal.hpp"

namespace atlg {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 16, NT = 256;
constexpr int WM = 64, WN = 64;           // wave tile
constexpr int TM = WM / 16, TN = WN / 16; // MFMA tiles per wave
constexpr int PADA = 4, PADB = 4;         // LDS row padding (doubles)
constexpr int LDA_S = BM + PADA, LDB_S = BN + PADB;

// Register layout of v_mfma_f64_16x16x4_f64 (D = A*B + C, A 16x4, B 4x16),
// measured on gfx950 with one-hot operands: lane l supplies A[l&15][l>>4]
// and B[l>>4][l&15]; D[i][j] is in lane j + 16*(i&3), register i>>2.
// The kernel passes the B tile as the first operand and the A tile as the
// second, so the accumulator holds C^T: lane l, register r carries
// C[m = l&15][n = (l>>4) + 4r], and a register's 16 lanes of one row group
// are 16 consecutive rows of a column (one 128-byte store).
struct MfmaF64 {
  __device__ static int frag_x(int l) { return l & 15; }   // m (or n) of a fragment
  __device__ static int frag_k(int l) { return l >> 4; }   // k of a fragment
  __device__ static int c_m(int l) { return l & 15; }
  __device__ static int c_n(int l, int r) { return (l >> 4) + 4 * r; }
};

// Loads one BK x 128 operand tile into registers.  Element (x, k) of the
// tile (x = m for A, n for B) lives at P[x + k*ld] when XCONTIG (A NoTrans,
// B Trans), else at P[k + x*ld] (A Trans, B NoTrans).
template <bool XCONTIG, bool VEC>
struct TileLoader {
  double v[8];
  __device__ void load(const double* __restrict__ P, long ld, int x0, int k0,
                       int Xtot, int Ktot, int t) {
    if (XCONTIG) {
      const int k = t >> 4, x = (t & 15) * 8;
      const int gk = k0 + k, gx = x0 + x;
      const double* p = P + (size_t)gx + (size_t)gk * ld;
      if (gk < Ktot && gx + 8 <= Xtot) {
        if (VEC) {
          const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
          for (int i = 0; i < 4; i++) {
            double2 w = q[i];
            v[2 * i] = w.x;
            v[2 * i + 1] = w.y;
          }
        } else {
#pragma unroll
          for (int i = 0; i < 8; i++) v[i] = p[i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = (gk < Ktot && gx + i < Xtot) ? p[i] : 0.0;
      }
    } else {
      const int x = t >> 1, k = (t & 1) * 8;
      const int gk = k0 + k, gx = x0 + x;
      const double* p = P + (size_t)gk + (size_t)gx * ld;
      if (gx < Xtot && gk + 8 <= Ktot) {
        if (VEC) {
          const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
          for (int i = 0; i < 4; i++) {
            double2 w = q[i];
            v[2 * i] = w.x;
            v[2 * i + 1] = w.y;
          }
        } else {
#pragma unroll
          for (int i = 0; i < 8; i++) v[i] = p[i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = (gx < Xtot && gk + i < Ktot) ? p[i] : 0.0;
      }
    }
  }
  // LDS tile is [BK][LD] with x contiguous.
  template <int LD>
  __device__ void store(double* S, int t) const {
    if (XCONTIG) {
      const int k = t >> 4, x = (t & 15) * 8;
      double2* q = reinterpret_cast<double2*>(S + k * LD + x);
#pragma unroll
      for (int i = 0; i < 4; i++) q[i] = make_double2(v[2 * i], v[2 * i + 1]);
    } else {
      const int x = t >> 1, k = (t & 1) * 8;
#pragma unroll
      for (int i = 0; i < 8; i++) S[(k + i) * LD + x] = v[i];
    }
  }
};

// Maps the linear workgroup id to a C tile so that the workgroups resident on
// one XCD (hardware deals ids round-robin over the 8 XCDs) cover a compact
// block of tiles, sharing A row panels and B column panels in that XCD's L2.
__device__ inline void tile_of(int bid, int nblk, int tilesM, int tilesN, int& tm,
                               int& tn) {
  const int per = nblk / 8, rem = nblk % 8;
  const int xcd = bid % 8, local = bid / 8;
  const int id = (xcd < rem) ? xcd * (per + 1) + local
                             : rem * (per + 1) + (xcd - rem) * per + local;
  constexpr int GM = 8;  // tile rows per group
  const int group = id / (GM * tilesN);
  const int first = group * GM;
  const int gsz = min(tilesM - first, GM);
  const int in = id - group * GM * tilesN;
  tm = first + in % gsz;
  tn = in / gsz;
}

template <bool AT, bool BT, bool VA, bool VB>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
dgemm_mfma(int M, int N, int K, double alpha, const double* A, long lda,
           const double* B, long ldb, double beta, double* C, long ldc, int mode,
           const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
  int tm, tn;
  tile_of(blockIdx.x, gridDim.x, tilesM, tilesN, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  if (mode == 1 && n0 >= m0 + BM) return;  // tile strictly above the diagonal
  if (mode == 2 && m0 >= n0 + BN) return;  // tile strictly below the diagonal

  __shared__ __attribute__((aligned(16))) double As[2][BK * LDA_S];
  __shared__ __attribute__((aligned(16))) double Bs[2][BK * LDB_S];

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;

  // A tile: x = m; contiguous in m unless transposed.  B tile: x = n;
  // contiguous in n only when transposed.
  TileLoader<!AT, VA> la;
  TileLoader<BT, VB> lb;

  d4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  const int nk = (K + BK - 1) / BK;
  la.load(A, lda, m0, 0, M, K, t);
  lb.load(B, ldb, n0, 0, N, K, t);
  la.template store<LDA_S>(As[0], t);
  lb.template store<LDB_S>(Bs[0], t);
  __syncthreads();

  const int fx = MfmaF64::frag_x(lane), fk = MfmaF64::frag_k(lane);

  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    const bool more = (kt + 1 < nk);
    if (more) {
      la.load(A, lda, m0, (kt + 1) * BK, M, K, t);
      lb.load(B, ldb, n0, (kt + 1) * BK, N, K, t);
    }
    const double* as = As[cur];
    const double* bs = Bs[cur];
#pragma unroll
    for (int ks = 0; ks < BK / 4; ks++) {
      double af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) af[i] = as[(ks * 4 + fk) * LDA_S + wm0 + i * 16 + fx];
#pragma unroll
      for (int j = 0; j < TN; j++) bf[j] = bs[(ks * 4 + fk) * LDB_S + wn0 + j * 16 + fx];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
    if (more) {
      la.template store<LDA_S>(As[cur ^ 1], t);
      lb.template store<LDB_S>(Bs[cur ^ 1], t);
    }
    __syncthreads();
  }

  // Epilogue: C = alpha*acc + beta*C, column-major, masked to the matrix and
  // (for SYRK) to the requested triangle.
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi = m0 + wm0 + i * 16 + MfmaF64::c_m(lane);
        const int gj = n0 + wn0 + j * 16 + MfmaF64::c_n(lane, r);
        if (gi >= M || gj >= N) continue;
        if (mode == 1 && gi < gj) continue;
        if (mode == 2 && gi > gj) continue;
        double* c = C + (size_t)gi + (size_t)gj * ldc;
        const double v = alpha * acc[i][j][r];
        *c = (beta == 0.0) ? v : v + beta * (*c);
      }
}

__global__ void scale_kernel(int M, int N, double beta, double* C, long ldc, int mode,
                             const int* __restrict__ skip) {
  if (skip && *skip) return;
  const long total = (long)M * N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % M), j = (int)(idx / M);
    if (mode == 1 && i < j) continue;
    if (mode == 2 && i > j) continue;
    double* c = C + (size_t)i + (size_t)j * ldc;
    *c = (beta == 0.0) ? 0.0 : beta * (*c);
  }
}

// C(i,j) <- C(j,i) for i > j: the lower triangle becomes the mirror of the
// upper one (ATL_dsyreflect as used by cblas_dgemm, cblas.base:224-231).
__global__ void syreflect_kernel(int N, double* C, long ldc) {
  const long total = (long)N * N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % N), j = (int)(idx / N);
    if (i > j) C[(size_t)i + (size_t)j * ldc] = C[(size_t)j + (size_t)i * ldc];
  }
}

template <bool AT, bool BT>
void launch_t(hipStream_t s, dim3 g, bool va, bool vb, int M, int N, int K, double alpha,
              const double* A, long lda, const double* B, long ldb, double beta, double* C,
              long ldc, int mode, const int* skip) {
#define ATLG_GO(VA_, VB_)                                                          \
  hipLaunchKernelGGL((dgemm_mfma<AT, BT, VA_, VB_>), g, dim3(NT), 0, s, M, N, K, alpha, \
                     A, lda, B, ldb, beta, C, ldc, mode, skip)
  if (va && vb) ATLG_GO(true, true);
  else if (va) ATLG_GO(true, false);
  else if (vb) ATLG_GO(false, true);
  else ATLG_GO(false, false);
#undef ATLG_GO
}

bool aligned16(const void* p, long ld) {
  return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && (ld % 2 == 0);
}

}  // namespace

void scale(hipStream_t s, long M, long N, double beta, double* C, long ldc, int mode,
           const int* skip) {
  if (M <= 0 || N <= 0 || beta == 1.0) return;
  const long total = M * N;
  const int nb = (int)std::min<long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(256), 0, s, (int)M, (int)N, beta, C, ldc,
                     mode, skip);
  ATLG_LAUNCHED();
}

void syreflect_upper(hipStream_t s, long N, double* C, long ldc) {
  if (N <= 1) return;
  const int nb = (int)std::min<long>((N * N + 255) / 256, 4096);
  hipLaunchKernelGGL(syreflect_kernel, dim3(nb), dim3(256), 0, s, (int)N, C, ldc);
  ATLG_LAUNCHED();
}

void gemm(hipStream_t s, int TA, int TB, long M, long N, long K, double alpha,
          const double* A, long lda, const double* B, long ldb, double beta, double* C,
          long ldc, int mode, const int* skip) {
  if (M <= 0 || N <= 0) return;
  if (K == 0 || alpha == 0.0) {
    scale(s, M, N, beta, C, ldc, mode, skip);
    return;
  }
  const long tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
  const dim3 g((unsigned)(tilesM * tilesN));
  const bool at = is_trans(TA), bt = is_trans(TB);
  const bool va = aligned16(A, lda), vb = aligned16(B, ldb);
  if (!at && !bt)
    launch_t<false, false>(s, g, va, vb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode, skip);
  else if (!at && bt)
    launch_t<false, true>(s, g, va, vb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode, skip);
  else if (at && !bt)
    launch_t<true, false>(s, g, va, vb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode, skip);
  else
    launch_t<true, true>(s, g, va, vb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode, skip);
  ATLG_LAUNCHED();
}

}  // namespace atlg
