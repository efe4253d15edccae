// GEMM family for gfx950 on the matrix cores: FP64 (v_mfma_f64_16x16x4_f64)
// and, with the same tiling, FP32 (v_mfma_f32_16x16x4_f32) for the
// single-precision factorization of the mixed-precision solver.
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
#include <type_traits>

#include "atlg_internal.hpp"

namespace atlg {
namespace {

constexpr int BM = 128, BN = 128, BK = 16, NT = 256;
static_assert(BM == BN, "triangular tile enumeration assumes square tiles");
// Tile rows per group of the XCD-aware tile order, and whether workgroup ids
// are regrouped per XCD at all (tuned constants).
#ifndef ATLG_GEMM_GM
#define ATLG_GEMM_GM 8
#endif
#ifndef ATLG_GEMM_XCD
#define ATLG_GEMM_XCD 1
#endif
constexpr int WM = 64, WN = 64;           // wave tile
constexpr int TM = WM / 16, TN = WN / 16; // MFMA tiles per wave
constexpr int PADA = 4, PADB = 4;         // LDS row padding (elements)
constexpr int LDA_S = BM + PADA, LDB_S = BN + PADB;

// The 16x16x4 MFMA of each precision (D = X1*X2 + C, X1 16x4, X2 4x16).
// Operand layout (both precisions): lane l supplies X1[l&15][l>>4] and
// X2[l>>4][l&15].  Result layout: FP64 keeps D[i][j] in lane j + 16*(i&3),
// register i>>2 (measured on gfx950 with one-hot operands); FP32 keeps it in
// lane j + 16*(i>>2), register i&3 (the CDNA layout of the single-block
// 16x16 FP32 MFMA, confirmed by the SGEMM parity tests).  The kernel passes the B tile as X1 and the A tile as X2, so the
// accumulator holds C^T: lane l, register r carries C[m = l&15][n = c_n(l, r)],
// and a register's 16 lanes of one row group are 16 consecutive rows of a
// column (one coalesced store).
template <class S>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t op(double x1, double x2, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x1, x2, c, 0, 0, 0);
  }
  __device__ static int c_n(int l, int r) { return (l >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t op(float x1, float x2, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x1, x2, c, 0, 0, 0);
  }
  __device__ static int c_n(int l, int r) { return 4 * (l >> 4) + r; }
};
__device__ inline int frag_x(int l) { return l & 15; }  // m (or n) of a fragment
__device__ inline int frag_k(int l) { return l >> 4; }  // k of a fragment
__device__ inline int c_m(int l) { return l & 15; }

// Loads one BK x 128 operand tile into registers (8 elements per thread).
// Element (x, k) of the tile (x = m for A, n for B) lives at P[x + k*ld] when
// XCONTIG (A NoTrans, B Trans), else at P[k + x*ld] (A Trans, B NoTrans).
// VEC: the 8 elements are read with 16-byte loads.
template <class S, bool XCONTIG, bool VEC>
struct TileLoader {
  static constexpr int VW = 16 / sizeof(S);  // elements per 16-byte vector
  typedef S vec_t __attribute__((ext_vector_type(VW)));
  S v[8];
  __device__ void load8(const S* p) {
    if (VEC) {
      const vec_t* q = reinterpret_cast<const vec_t*>(p);
#pragma unroll
      for (int i = 0; i < 8 / VW; i++) {
        const vec_t w = q[i];
#pragma unroll
        for (int e = 0; e < VW; e++) v[VW * i + e] = w[e];
      }
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = p[i];
    }
  }
  __device__ void load(const S* __restrict__ P, long ld, int x0, int k0, int Xtot, int Ktot,
                       int t) {
    if (XCONTIG) {
      const int k = t >> 4, x = (t & 15) * 8;
      const int gk = k0 + k, gx = x0 + x;
      const S* p = P + (size_t)gx + (size_t)gk * ld;
      if (gk < Ktot && gx + 8 <= Xtot) {
        load8(p);
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = (gk < Ktot && gx + i < Xtot) ? p[i] : S(0);
      }
    } else {
      const int x = t >> 1, k = (t & 1) * 8;
      const int gk = k0 + k, gx = x0 + x;
      const S* p = P + (size_t)gk + (size_t)gx * ld;
      if (gx < Xtot && gk + 8 <= Ktot) {
        load8(p);
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = (gx < Xtot && gk + i < Ktot) ? p[i] : S(0);
      }
    }
  }
  // LDS tile is [BK][LD] with x contiguous.
  template <int LD>
  __device__ void store(S* T, int t) const {
    if (XCONTIG) {
      const int k = t >> 4, x = (t & 15) * 8;
      vec_t* q = reinterpret_cast<vec_t*>(T + k * LD + x);
#pragma unroll
      for (int i = 0; i < 8 / VW; i++) {
        vec_t w;
#pragma unroll
        for (int e = 0; e < VW; e++) w[e] = v[VW * i + e];
        q[i] = w;
      }
    } else {
      const int x = t >> 1, k = (t & 1) * 8;
#pragma unroll
      for (int i = 0; i < 8; i++) T[(k + i) * LD + x] = v[i];
    }
  }
};

// Maps the linear workgroup id to a C tile so that the workgroups resident on
// one XCD (hardware deals ids round-robin over the 8 XCDs) cover a compact
// block of tiles, sharing A row panels and B column panels in that XCD's L2.
__device__ inline void tile_of(int bid, int nblk, int tilesM, int tilesN, int& tm,
                               int& tn, bool compact = true) {
  const int per = nblk / 8, rem = nblk % 8;
  const int xcd = bid % 8, local = bid / 8;
  const int id = (!ATLG_GEMM_XCD || !compact) ? bid
                 : (xcd < rem) ? xcd * (per + 1) + local
                               : rem * (per + 1) + (xcd - rem) * per + local;
  constexpr int GM = ATLG_GEMM_GM;  // tile rows per group
  const int group = id / (GM * tilesN);
  const int first = group * GM;
  const int gsz = min(tilesM - first, GM);
  const int in = id - group * GM * tilesN;
  tm = first + in % gsz;
  tn = in / gsz;
}

// Tiles of one triangle of a tilesM x tilesN tile grid (BM == BN): mode 1
// keeps the tiles with tn <= tm (the lower triangle and everything below
// it), mode 2 those with tm <= tn.  The kept tiles are enumerated row by row
// (mode 1) or column by column (mode 2), so a triangular update launches
// only the tiles it computes.
__host__ __device__ inline int tri_tiles(int mode, int tilesM, int tilesN) {
  const int R = mode == 1 ? tilesM : tilesN, Cn = mode == 1 ? tilesN : tilesM;
  const int P = R < Cn ? R : Cn;
  return P * (P + 1) / 2 + (R - P) * Cn;
}
__device__ inline void tri_tile_of(int id, int mode, int tilesM, int tilesN, int& tm, int& tn) {
  const int R = mode == 1 ? tilesM : tilesN, Cn = mode == 1 ? tilesN : tilesM;
  const int P = R < Cn ? R : Cn, T0 = P * (P + 1) / 2;
  int a, b;  // a: index along R, b <= a along Cn
  if (id < T0) {
    a = (int)((sqrtf(8.0f * (float)id + 1.0f) - 1.0f) * 0.5f);
    while ((a + 1) * (a + 2) / 2 <= id) a++;
    while (a * (a + 1) / 2 > id) a--;
    b = id - a * (a + 1) / 2;
  } else {
    const int r = id - T0;
    a = P + r / Cn;
    b = r - (r / Cn) * Cn;
  }
  if (mode == 1) {
    tm = a;
    tn = b;
  } else {
    tn = a;
    tm = b;
  }
}

// The same tiles in groups of ATLG_GEMM_GM tile rows (mode 1; tile columns
// for mode 2), each group column by column, like tile_of's groups: the
// workgroups resident on one XCD then cover a few tile rows by a few tile
// columns (a few A and B panels in its L2) instead of one or two whole rows
// of the triangle (as many B panels as the row is long).
#ifndef ATLG_GEMM_TRI_GROUPED
#define ATLG_GEMM_TRI_GROUPED 1
#endif
__device__ inline void tri_tile_grouped(int id, int mode, int tilesM, int tilesN, int& tm,
                                        int& tn) {
  constexpr int GM = ATLG_GEMM_GM;
  const int R = mode == 1 ? tilesM : tilesN, Cn = mode == 1 ? tilesN : tilesM;
  for (int r0 = 0;; r0 += GM) {
    const int r1 = min(r0 + GM, R), h = r1 - r0;
    const int full = min(r0, Cn);                  // columns with all h rows
    const int tc = max(0, min(r1, Cn) - r0);       // then column r0 + j has h - j rows
    const int cnt = full * h + tc * h - tc * (tc - 1) / 2;
    if (id < cnt || r1 >= R) {
      int a, b;
      if (id < full * h) {
        b = id / h;
        a = r0 + id % h;
      } else {
        int o = id - full * h, j = 0;
        while (j + 1 < tc && o >= h - j) {
          o -= h - j;
          j++;
        }
        b = r0 + j;
        a = r0 + j + o;
      }
      if (mode == 1) {
        tm = a;
        tn = b;
      } else {
        tn = a;
        tm = b;
      }
      return;
    }
    id -= cnt;
  }
}

template <class S, bool AT, bool BT, bool VA, bool VB>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_mfma(int M, int N, int K, S alpha, const S* A, long lda, const S* B, long ldb, S beta,
          S* C, long ldc, int mode, const int* __restrict__ skip) {
  typedef Mfma<S> MF;
  if (skip && *skip) return;
  const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
  const int ntiles = mode == 0 ? tilesM * tilesN : tri_tiles(mode, tilesM, tilesN);

  __shared__ __attribute__((aligned(16))) S As[2][BK * LDA_S];
  __shared__ __attribute__((aligned(16))) S Bs[2][BK * LDB_S];

  // One tile per workgroup, or (a capped grid, its size a multiple of 8 so
  // that a workgroup keeps its XCD) a strided loop over the tiles.
  for (int bid = blockIdx.x; bid < ntiles; bid += gridDim.x) {
  int tm, tn;
  if (mode == 0) {
    tile_of(bid, ntiles, tilesM, tilesN, tm, tn);
  } else {
    // the same XCD-compact renumbering, then the triangle's enumeration
    const int per = ntiles / 8, rem = ntiles % 8, xcd = bid % 8, local = bid / 8;
    const int id = !ATLG_GEMM_XCD ? bid
                   : (xcd < rem) ? xcd * (per + 1) + local
                                 : rem * (per + 1) + (xcd - rem) * per + local;
    if (ATLG_GEMM_TRI_GROUPED) tri_tile_grouped(id, mode, tilesM, tilesN, tm, tn);
    else tri_tile_of(id, mode, tilesM, tilesN, tm, tn);
  }
  const int m0 = tm * BM, n0 = tn * BN;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;

  // A tile: x = m; contiguous in m unless transposed.  B tile: x = n;
  // contiguous in n only when transposed.
  TileLoader<S, !AT, VA> la;
  TileLoader<S, BT, VB> lb;

  typename MF::acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = typename MF::acc_t{0, 0, 0, 0};

  const int nk = (K + BK - 1) / BK;
  la.load(A, lda, m0, 0, M, K, t);
  lb.load(B, ldb, n0, 0, N, K, t);
  la.template store<LDA_S>(As[0], t);
  lb.template store<LDB_S>(Bs[0], t);
  __syncthreads();

  const int fx = frag_x(lane), fk = frag_k(lane);

  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    const bool more = (kt + 1 < nk);
    if (more) {
      la.load(A, lda, m0, (kt + 1) * BK, M, K, t);
      lb.load(B, ldb, n0, (kt + 1) * BK, N, K, t);
    }
    const S* as = As[cur];
    const S* bs = Bs[cur];
#pragma unroll
    for (int ks = 0; ks < BK / 4; ks++) {
      S af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) af[i] = as[(ks * 4 + fk) * LDA_S + wm0 + i * 16 + fx];
#pragma unroll
      for (int j = 0; j < TN; j++) bf[j] = bs[(ks * 4 + fk) * LDB_S + wn0 + j * 16 + fx];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = MF::op(bf[j], af[i], acc[i][j]);
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
        const int gi = m0 + wm0 + i * 16 + c_m(lane);
        const int gj = n0 + wn0 + j * 16 + MF::c_n(lane, r);
        if (gi >= M || gj >= N) continue;
        if (mode == 1 && gi < gj) continue;
        if (mode == 2 && gi > gj) continue;
        S* c = C + (size_t)gi + (size_t)gj * ldc;
        const S v = alpha * acc[i][j][r];
        *c = (beta == S(0)) ? v : v + beta * (*c);
      }
  __syncthreads();  // the next tile's first loads overwrite As/Bs[0]
  }
}

// FP64 C = alpha*op(A)*op(B) + beta*C for A and B 16-byte aligned with
// even leading dimensions, K a multiple of 32: the A operand
// goes from memory straight into the MFMA operand registers, and only B is
// staged through (double-buffered) LDS.  Halving the LDS stores per K step
// is what lifts this form above gemm_mfma (measured on MI355X at
// M=N=K=16384: 72.9 against 66.5 TFLOP/s for the same tiling with both
// operands staged).
//
// Tile 128x128 per 256-thread workgroup, 2x2 waves of 64x64, K in steps of
// 16 (four 16x16x4 MFMA steps).  Lane (fx, fk) = (l & 15, l >> 4) of wave
// tile i = 2h + e (rows) supplies A row wm0 + 32h + 2fx + e at k = 4ks + fk:
// one 16-byte load holds the rows of tiles 2h and 2h+1, and the C tile keeps
// the same row order, so the epilogue also moves 16 bytes per lane.  A for
// K tile kt + 2 is loaded right after the MFMAs of K tile kt have consumed
// the registers; B for K tile kt + 2 is loaded one step after B for K tile
// kt + 1 is written to LDS.  The loop body is branch-free (loads past the
// end repeat the last K tile and are never used), so the memory counters'
// waits stay exact.  Rows past M and columns past N are clamped on load and
// masked on store.  One barrier per K tile, placed before the last MFMA
// step so that step covers the next tile's first LDS reads.
//
// AT: op(A) = A^T, each lane loading its rows' k runs instead of row pairs.
// BT selects op(B): false, B is K x N (a tile column of 16 k is 128
// contiguous bytes, loaded as 8 column chunks per wave and transposed into
// LDS by two 8-byte stores); true, op(B) = B^T with B N x K (a tile row of
// 128 n is contiguous, loaded and stored to LDS 16 bytes per lane).  mode
// 1 / 2 computes only the tiles and elements on or below / above the
// diagonal (SYRK and the Cholesky update), enumerated as gemm_mfma does.
constexpr int NN_DA = 2;  // K tiles of A in flight
constexpr int kBufDword3 = 0x00020000;  // raw buffer, 32-bit dwords (gfx9 family)
// XT: the extended form (a batch over blockIdx.y at element strides sa, sb,
// sc, and the per-tile K range of ktrim); without it the code is that of the
// plain product, whose register allocation it would otherwise disturb.
template <bool AT, bool BT, bool XT>
__device__ __forceinline__ void gemm_rega_f64(int M, int N, int K, double alpha,
                                              const double* __restrict__ A, long lda,
                                              const double* B, long ldb, double beta, double* C,
                                              long ldc, int vecc, int mode, long sa, long sb,
                                              long sc, int ktrim) {
  typedef Mfma<double> MF;
  typedef double d2 __attribute__((ext_vector_type(2)));
  constexpr int LDB_S = BN + PADB;
  __shared__ __attribute__((aligned(16))) double Bs[2][BK * LDB_S];
  const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
  const int ntiles = mode == 0 ? tilesM * tilesN : tri_tiles(mode, tilesM, tilesN);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const int fx = lane & 15, fk = lane >> 4;
  if (XT) {  // batch member blockIdx.y: operands at fixed element strides
    A += (size_t)blockIdx.y * sa;
    B += (size_t)blockIdx.y * sb;
    C += (size_t)blockIdx.y * sc;
  }
  // With K-trimmed tiles the work per tile follows its row or column, and
  // the XCD-compact order (each XCD a contiguous range of tiles) would give
  // one XCD all the long ones: those keep the dispatch order, which deals
  // consecutive tiles to the XCDs in turn.
  const bool compact = !(XT && ktrim != 0);
  for (int bid = blockIdx.x; bid < ntiles; bid += gridDim.x) {
    int tm, tn;
    if (mode == 0) {
      tile_of(bid, ntiles, tilesM, tilesN, tm, tn, compact);
    } else {
      const int per = ntiles / 8, rem = ntiles % 8, xcd = bid % 8, local = bid / 8;
      const int id = (!ATLG_GEMM_XCD || !compact) ? bid
                     : (xcd < rem) ? xcd * (per + 1) + local
                                   : rem * (per + 1) + (xcd - rem) * per + local;
      if (ATLG_GEMM_TRI_GROUPED && compact) tri_tile_grouped(id, mode, tilesM, tilesN, tm, tn);
      else tri_tile_of(id, mode, tilesM, tilesN, tm, tn);
    }
    const int m0 = tm * BM, n0 = tn * BN;
    // K range of the tile: all of K, or with a triangular operand whose
    // other triangle is zero only the part that can be nonzero (bit 1: B
    // lower, k >= n0; bit 2: A^T lower, k >= m0; bit 4: A lower, k < m0+BM).
    // The bounds are multiples of 2*BK (m0, n0 and K are); the callers'
    // shapes keep every range nonempty (m0, n0 < K), which the prologue's
    // loads assume.
    int kb = 0, ke = K;
    if (XT) {
      if (ktrim & 1) kb = max(kb, n0);
      if (ktrim & 2) kb = max(kb, m0);
      if (ktrim & 4) ke = min(ke, m0 + BM);
    }
    const int nk = max(0, ke - kb) / BK;
    // Addresses: a wave-uniform base per tile and K step plus a 32-bit
    // per-lane element offset inside the tile.  A: this lane's row pairs,
    // clamped so a pair starts at most at M-1 (lda is even, so the pair
    // stays inside the allocation).  A^T: the lane's four rows (clamped to
    // M-1), each holding k = 4fk .. 4fk+3 contiguously; the MFMA steps then
    // take k = 4fk + ks (any k order serves, as long as B's matches).
    const double* At = AT ? A + (size_t)m0 * lda + kb : A + m0 + (size_t)kb * lda;
    unsigned aoff[AT ? 4 : 2];  // bytes
    if (AT) {
#pragma unroll
      for (int i = 0; i < 4; i++)
        aoff[i] = 8u * (unsigned)(min(wm0 + 32 * (i >> 1) + 2 * fx + (i & 1), M - 1 - m0) *
                                      (int)lda + 4 * fk);
    } else {
#pragma unroll
      for (int h = 0; h < 2; h++)
        aoff[h] = 8u * (unsigned)(min(wm0 + 32 * h + 2 * fx, ((M - 1) & ~1) - m0) + fk * (int)lda);
    }
    // A registers per K tile: NoTrans, slot 2ks + h holds rows of tiles 2h,
    // 2h+1 at step ks; A^T, slot 4p + i holds tile i's k pair p (steps 2p,
    // 2p+1)
    d2 ar[NN_DA][8];
    auto aval = [&](int d, int ks, int i) -> double {
      return AT ? ar[d][(ks >> 1) * 4 + i][ks & 1] : ar[d][ks * 2 + (i >> 1)][i & 1];
    };
    // loads the registers step ks of K tile slot d has just consumed
    auto aload = [&](int d, int ks, int k0) {
      if (AT) {
        if (ks & 1) {
          const int p = ks >> 1;
          const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
              const_cast<double*>(At + k0 + 2 * p), 0, 0x7fffffff, kBufDword3);
#pragma unroll
          for (int i = 0; i < 4; i++)
            ar[d][4 * p + i] =
                __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, aoff[i], 0, 0));
        }
      } else {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<double*>(At + (size_t)(k0 + 4 * ks) * lda), 0, 0x7fffffff, kBufDword3);
#pragma unroll
        for (int h = 0; h < 2; h++)
          ar[d][2 * ks + h] =
              __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, aoff[h], 0, 0));
      }
    };
    // B (not transposed): 4 chunks of 8 columns x 16 k per wave, 16 bytes
    // per lane, columns past N clamped to N-1.  B^T: k rows 4i + wave, 2
    // columns per lane, the pair start clamped to at most N-1 (ldb is even,
    // so the pair stays inside the allocation).
    constexpr int B_PER = BN / 8 / (NT / 64);
    const double* Bt = BT ? B + n0 + (size_t)kb * ldb : B + (size_t)n0 * ldb + kb;
    unsigned boff[B_PER];  // bytes
#pragma unroll
    for (int i = 0; i < B_PER; i++)
      boff[i] = BT ? 8u * (unsigned)(min(2 * lane, ((N - 1) & ~1) - n0) + (4 * i + wave) * (int)ldb)
                   : 8u * (unsigned)(min((wave * B_PER + i) * 8 + (lane >> 3), N - 1 - n0) *
                                         (int)ldb + 2 * (lane & 7));
    d2 rb[B_PER];
    auto bload = [&](int k0) {
      const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<double*>(BT ? Bt + (size_t)k0 * ldb : Bt + k0), 0, 0x7fffffff, kBufDword3);
#pragma unroll
      for (int i = 0; i < B_PER; i++)
        rb[i] = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, boff[i], 0, 0));
    };
    auto bstore = [&](int buf) {
#pragma unroll
      for (int i = 0; i < B_PER; i++) {
        if (BT) {
          *reinterpret_cast<d2*>(&Bs[buf][(4 * i + wave) * LDB_S + 2 * lane]) = rb[i];
        } else {
          const int n = (wave * B_PER + i) * 8 + (lane >> 3), k = 2 * (lane & 7);
          Bs[buf][k * LDB_S + n] = rb[i][0];
          Bs[buf][(k + 1) * LDB_S + n] = rb[i][1];
        }
      }
    };
    double bf[2][TN];
    auto bread = [&](int buf, int ks, int f) {
#pragma unroll
      for (int j = 0; j < TN; j++)
        bf[f][j] = Bs[buf][(AT ? 4 * fk + ks : ks * 4 + fk) * LDB_S + wn0 + j * 16 + fx];
    };
    MF::acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = MF::acc_t{0, 0, 0, 0};
    auto mma = [&](int d, int ks, int f) {
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = MF::op(bf[f][j], aval(d, ks, i), acc[i][j]);
    };
    bload(0);
    bstore(0);
    bload(min(1, nk - 1) * BK);
#pragma unroll
    for (int d = 0; d < NN_DA; d++)
#pragma unroll
      for (int ks = 0; ks < 4; ks++) aload(d, ks, min(d, nk - 1) * BK);
    __syncthreads();
    bread(0, 0, 0);
#define ATLG_SB() __builtin_amdgcn_sched_barrier(0)
    for (int kt0 = 0; kt0 < nk; kt0 += NN_DA) {
#pragma unroll
      for (int d = 0; d < NN_DA; d++) {  // nk is a multiple of NN_DA
        const int kt = kt0 + d, cur = d & 1;
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
          if (ks < 3) bread(cur, ks + 1, (ks + 1) & 1);
          else bread(cur ^ 1, 0, 0);
          if (ks == 1) {
            bstore(cur ^ 1);
            bload(min(kt + 2, nk - 1) * BK);
          }
          ATLG_SB();
          mma(d, ks, ks & 1);
          ATLG_SB();
          aload(d, ks, min(kt + NN_DA, nk - 1) * BK);
          if (ks == 2) __syncthreads();
        }
      }
    }
#undef ATLG_SB
    // epilogue: the rows of tiles 2h and 2h+1 are adjacent (16-byte C
    // accesses when C allows them and both rows are kept); beta == 0 never
    // reads C
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int j = 0; j < TN; j++) {
        const int gi = m0 + wm0 + 32 * h + 2 * fx;
        double cv[4][2];
        bool keep[4][2];
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int gj = n0 + wn0 + j * 16 + MF::c_n(lane, r);
#pragma unroll
          for (int e = 0; e < 2; e++)
            keep[r][e] = gi + e < M && gj < N &&
                         (mode == 0 || (mode == 1 ? gi + e >= gj : gi + e <= gj));
          double* c = C + (size_t)gi + (size_t)gj * ldc;
          cv[r][0] = cv[r][1] = 0.0;
          if (beta != 0.0) {
            if (vecc && keep[r][0] && keep[r][1]) {
              const d2 w = *reinterpret_cast<const d2*>(c);
              cv[r][0] = w[0];
              cv[r][1] = w[1];
            } else {
              if (keep[r][0]) cv[r][0] = c[0];
              if (keep[r][1]) cv[r][1] = c[1];
            }
          }
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int gj = n0 + wn0 + j * 16 + MF::c_n(lane, r);
          double* c = C + (size_t)gi + (size_t)gj * ldc;
          d2 o;
          o[0] = alpha * acc[2 * h][j][r];
          o[1] = alpha * acc[2 * h + 1][j][r];
          if (beta != 0.0) {
            o[0] += beta * cv[r][0];
            o[1] += beta * cv[r][1];
          }
          if (vecc && keep[r][0] && keep[r][1]) {
            *reinterpret_cast<d2*>(c) = o;
          } else {
            if (keep[r][0]) c[0] = o[0];
            if (keep[r][1]) c[1] = o[1];
          }
        }
      }
    __syncthreads();  // the next tile's first B stores overwrite Bs[0]
  }
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_nn_f64(int M, int N, int K, double alpha, const double* __restrict__ A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc, int vecc, int mode,
            const int* __restrict__ skip) {
  if (skip && *skip) return;
  gemm_rega_f64<false, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, vecc, mode, 0, 0,
                               0, 0);
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_nn_f64_x(int M, int N, int K, double alpha, const double* __restrict__ A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc, int vecc, int mode,
            const int* __restrict__ skip, long sa, long sb, long sc, int ktrim) {
  if (skip && *skip) return;
  gemm_rega_f64<false, false, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, vecc, mode, sa, sb,
                              sc, ktrim);
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_nt_f64(int M, int N, int K, double alpha, const double* __restrict__ A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc, int vecc, int mode,
            const int* __restrict__ skip) {
  if (skip && *skip) return;
  gemm_rega_f64<false, true, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, vecc, mode, 0, 0,
                               0, 0);
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_nt_f64_x(int M, int N, int K, double alpha, const double* __restrict__ A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc, int vecc, int mode,
            const int* __restrict__ skip, long sa, long sb, long sc, int ktrim) {
  if (skip && *skip) return;
  gemm_rega_f64<false, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, vecc, mode, sa, sb,
                              sc, ktrim);
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_tn_f64(int M, int N, int K, double alpha, const double* __restrict__ A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc, int vecc, int mode,
            const int* __restrict__ skip) {
  if (skip && *skip) return;
  gemm_rega_f64<true, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, vecc, mode, 0, 0,
                               0, 0);
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_tn_f64_x(int M, int N, int K, double alpha, const double* __restrict__ A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc, int vecc, int mode,
            const int* __restrict__ skip, long sa, long sb, long sc, int ktrim) {
  if (skip && *skip) return;
  gemm_rega_f64<true, false, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, vecc, mode, sa, sb,
                              sc, ktrim);
}

template <class S>
__global__ void scale_kernel(int M, int N, S beta, S* C, long ldc, int mode,
                             const int* __restrict__ skip) {
  if (skip && *skip) return;
  const long total = (long)M * N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % M), j = (int)(idx / M);
    if (mode == 1 && i < j) continue;
    if (mode == 2 && i > j) continue;
    S* c = C + (size_t)i + (size_t)j * ldc;
    *c = (beta == S(0)) ? S(0) : beta * (*c);
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

// Products with one short side: C = alpha op(A) B + beta C with M, K <= 128
// (LEFT), or C = alpha A op(B) + beta C with N, K <= 128.  C may be B (LEFT)
// or A (else): the diagonal-block applications of TRSM and TRMM, in place.
// The 128x128 tiling would give them one workgroup per 128 columns (rows);
// here a workgroup owns a 32-wide strip of the whole short side instead (4x
// the workgroups), reads all of its strip (and op(A) / op(B), shared through
// L2) before a barrier, and only then writes the strip back.  16x16 output
// tiles, four per wave: LEFT, row tiles 2w, 2w+1 by both column tiles; else
// both row tiles by column tiles 2w, 2w+1.  T: the short operand transposed.
// tri: 0 all of C, 1 only C(i,j) with i >= j, 2 only i <= j (the stored
// triangle of a symmetric diagonal block).  SL: 16-deep K slabs whose
// operands are loaded per round trip; SL > 1 (for launches of fewer
// workgroups than compute units, where each round trip's latency is the
// cost) also loads the old C before the K loop.
constexpr int kStrip = 32;
template <class S, bool LEFT, bool T, int SL>
__global__ void __launch_bounds__(256)
gemm_strip(int M, int N, int K, S alpha, const S* __restrict__ A, long lda,
           const S* __restrict__ B, long ldb, S beta, S* C, long ldc,
           const int* __restrict__ skip, int tri) {
  typedef Mfma<S> MF;
  if (skip && *skip) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, fx = lane & 15, fk = lane >> 4;
  const int r0 = LEFT ? 0 : (int)blockIdx.x * kStrip;  // first row / column of the strip
  const int c0 = LEFT ? (int)blockIdx.x * kStrip : 0;
  int rt[2], ct[2];  // this wave's row / column tile origins
#pragma unroll
  for (int e = 0; e < 2; e++) {
    rt[e] = LEFT ? r0 + 16 * (2 * w + e) : r0 + 16 * e;
    ct[e] = LEFT ? c0 + 16 * e : c0 + 16 * (2 * w + e);
  }
  // element (m, k) of op(A) and (k, n) of op(B); zero outside the matrices
  auto aval = [&](int m, int k) -> S {
    if (m >= M || k >= K) return S(0);
    return (LEFT && T) ? A[k + (size_t)m * lda] : A[m + (size_t)k * lda];
  };
  auto bval = [&](int k, int n) -> S {
    if (k >= K || n >= N) return S(0);
    return (!LEFT && T) ? B[n + (size_t)k * ldb] : B[k + (size_t)n * ldb];
  };
  typename MF::acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = typename MF::acc_t{0, 0, 0, 0};
  // lane l, register r: C[rt + (l & 15)][ct + c_n(l, r)]
  S old[2][2][4];
  auto load_old = [&]() {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int gi = rt[i] + fx, gj = ct[j] + MF::c_n(lane, r);
          old[i][j][r] = (beta != S(0) && gi < M && gj < N) ? C[gi + (size_t)gj * ldc] : S(0);
        }
  };
  if (SL > 1) load_old();
  for (int k0 = 0; k0 < K; k0 += 16 * SL) {  // 4 SL MFMA steps' operands in flight
    S av[SL][4][2], bv[SL][4][2];
#pragma unroll
    for (int l = 0; l < SL; l++)
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int e = 0; e < 2; e++) {
          av[l][q][e] = aval(rt[e] + fx, k0 + 16 * l + 4 * q + fk);
          bv[l][q][e] = bval(k0 + 16 * l + 4 * q + fk, ct[e] + fx);
        }
#pragma unroll
    for (int l = 0; l < SL; l++)
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) acc[i][j] = MF::op(bv[l][q][j], av[l][q][i], acc[i][j]);
  }
  if (SL == 1) load_old();
  __syncthreads();  // every read of the strip precedes every write
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi = rt[i] + fx, gj = ct[j] + MF::c_n(lane, r);
        if (gi < M && gj < N && !(tri == 1 && gi < gj) && !(tri == 2 && gi > gj))
          C[gi + (size_t)gj * ldc] =
              (beta == S(0)) ? alpha * acc[i][j][r] : alpha * acc[i][j][r] + beta * old[i][j][r];
      }
}

// Strip launches of fewer workgroups than this (a fraction of the device's
// 256 compute units, latency-bound) load four K slabs per round trip.
#ifndef ATLG_STRIP_DEEP_MAX
#define ATLG_STRIP_DEEP_MAX 256
#endif
template <class S>
void launch_strip(hipStream_t s, bool left, bool t, long M, long N, long K, S alpha, const S* A,
                  long lda, const S* B, long ldb, S beta, S* C, long ldc, const int* skip,
                  int tri) {
  const long nstrips = ((left ? N : M) + kStrip - 1) / kStrip;
  const dim3 gs((unsigned)nstrips);
  const bool deep = nstrips < ATLG_STRIP_DEEP_MAX;
#define ATLG_STRIP(L_, T_)                                                                   \
  do {                                                                                       \
    if (deep)                                                                                \
      hipLaunchKernelGGL((gemm_strip<S, L_, T_, 4>), gs, dim3(256), 0, s, (int)M, (int)N,   \
                         (int)K, alpha, A, lda, B, ldb, beta, C, ldc, skip, tri);            \
    else                                                                                     \
      hipLaunchKernelGGL((gemm_strip<S, L_, T_, 1>), gs, dim3(256), 0, s, (int)M, (int)N,   \
                         (int)K, alpha, A, lda, B, ldb, beta, C, ldc, skip, tri);            \
  } while (0)
  if (left) {
    if (t) ATLG_STRIP(true, true);
    else ATLG_STRIP(true, false);
  } else {
    if (t) ATLG_STRIP(false, true);
    else ATLG_STRIP(false, false);
  }
#undef ATLG_STRIP
  ATLG_LAUNCHED();
}

template <class S, bool AT, bool BT>
void launch_t(hipStream_t s, dim3 g, bool va, bool vb, int M, int N, int K, S alpha, const S* A,
              long lda, const S* B, long ldb, S beta, S* C, long ldc, int mode, const int* skip) {
#define ATLG_GO(VA_, VB_)                                                                  \
  hipLaunchKernelGGL((gemm_mfma<S, AT, BT, VA_, VB_>), g, dim3(NT), 0, s, M, N, K, alpha, A, \
                     lda, B, ldb, beta, C, ldc, mode, skip)
  if (va && vb) ATLG_GO(true, true);
  else if (va) ATLG_GO(true, false);
  else if (vb) ATLG_GO(false, true);
  else ATLG_GO(false, false);
#undef ATLG_GO
}

template <class S>
bool aligned16(const void* p, long ld) {
  return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && ((ld * (long)sizeof(S)) % 16 == 0);
}

}  // namespace

template <class S>
void scale(hipStream_t s, long M, long N, S beta, S* C, long ldc, int mode, const int* skip) {
  if (M <= 0 || N <= 0 || beta == S(1)) return;
  const long total = M * N;
  const int nb = (int)std::min<long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(scale_kernel<S>, dim3(nb), dim3(256), 0, s, (int)M, (int)N, beta, C, ldc,
                     mode, skip);
  ATLG_LAUNCHED();
}

void syreflect_upper(hipStream_t s, long N, double* C, long ldc) {
  if (N <= 1) return;
  const int nb = (int)std::min<long>((N * N + 255) / 256, 4096);
  hipLaunchKernelGGL(syreflect_kernel, dim3(nb), dim3(256), 0, s, (int)N, C, ldc);
  ATLG_LAUNCHED();
}

template <class S>
void gemm(hipStream_t s, int TA, int TB, long M, long N, long K, S alpha, const S* A, long lda,
          const S* B, long ldb, S beta, S* C, long ldc, int mode, const int* skip, int max_wg) {
  if (M <= 0 || N <= 0) return;
  if (K == 0 || alpha == S(0)) {
    scale<S>(s, M, N, beta, C, ldc, mode, skip);
    return;
  }
  const long tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
  long nwg = mode == 0 ? tilesM * tilesN : tri_tiles(mode, (int)tilesM, (int)tilesN);
  if (max_wg > 0 && nwg > max_wg) nwg = std::max(8L, (long)(max_wg / 8) * 8);
  const dim3 g((unsigned)nwg);
  const bool at = is_trans(TA), bt = is_trans(TB);
  // in-place diagonal-block applications (C is B, or C is A) with a short
  // side of at most 128: strips of 32 across the long side
  if (mode == 0 && C == B && C != A && M <= BM && K <= BM && !bt) {
    launch_strip<S>(s, true, at, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, skip, 0);
    return;
  }
  if (mode == 0 && C == A && C != B && N <= BN && K <= BN && !at) {
    launch_strip<S>(s, false, bt, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, skip, 0);
    return;
  }
  const bool va = aligned16<S>(A, lda), vb = aligned16<S>(B, ldb);
  // FP64 with aligned operands, K a multiple of 32 and C not one of the
  // operands (TRSM's in-place diagonal apply): A straight to registers, B
  // through LDS (per-lane byte offsets inside a tile are 32-bit: 128
  // columns or 16 rows of B, 4 columns plus 128 rows or 128 columns of A).
  // Not for A^T x B^T, whose form needs more than 256 registers.
  if (std::is_same<S, double>::value && !(at && bt) && va && vb && K % (2 * BK) == 0 &&
      C != A && C != B &&
      8L * (at ? BM * lda + BK : 4 * lda + BM) < (1L << 31) &&
      8L * (bt ? BK * ldb + BN : BN * ldb + BK) < (1L << 31)) {
    const int vecc = aligned16<S>(C, ldc) ? 1 : 0;
    hipLaunchKernelGGL(at ? gemm_tn_f64 : (bt ? gemm_nt_f64 : gemm_nn_f64),
                       g, dim3(NT), 0, s, (int)M, (int)N, (int)K, (double)alpha,
                       reinterpret_cast<const double*>(A), lda,
                       reinterpret_cast<const double*>(B), ldb, (double)beta,
                       reinterpret_cast<double*>(C), ldc, vecc, mode, skip);
    ATLG_LAUNCHED();
    return;
  }
  if (!at && !bt)
    launch_t<S, false, false>(s, g, va, vb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode,
                              skip);
  else if (!at && bt)
    launch_t<S, false, true>(s, g, va, vb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode,
                             skip);
  else if (at && !bt)
    launch_t<S, true, false>(s, g, va, vb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode,
                             skip);
  else
    launch_t<S, true, true>(s, g, va, vb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode,
                            skip);
  ATLG_LAUNCHED();
}

template <class S>
bool gemm_narrow(hipStream_t s, int TA, int TB, long M, long N, long K, S alpha, const S* A,
                 long lda, const S* B, long ldb, S beta, S* C, long ldc, int tri,
                 const int* skip) {
  const bool at = is_trans(TA), bt = is_trans(TB);
  if (M <= 0 || N <= 0 || K <= 0 || K > BM || M > (1L << 30) || N > (1L << 30)) return false;
  if (N <= BN && !at)  // strips of 32 rows by all N columns
    launch_strip<S>(s, false, bt, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, skip, tri);
  else if (M <= BM && !bt)  // strips of 32 columns by all M rows
    launch_strip<S>(s, true, at, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, skip, tri);
  else
    return false;
  return true;
}

bool gemm_batched(hipStream_t s, int TA, int TB, long M, long N, long K, double alpha,
                  const double* A, long lda, long sa, const double* B, long ldb, long sb,
                  double beta, double* C, long ldc, long sc, int batch, int mode, int ktrim,
                  const int* skip) {
  const bool at = is_trans(TA), bt = is_trans(TB);
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0 || batch > 65535 || (at && bt) ||
      K % (2 * BK) != 0 || ((sa | sb | sc) & 1) || !aligned16<double>(A, lda) ||
      !aligned16<double>(B, ldb) || 8L * (at ? BM * lda + BK : 4 * lda + BM) >= (1L << 31) ||
      8L * (bt ? BK * ldb + BN : BN * ldb + BK) >= (1L << 31))
    return false;
  // every tile's K range must be nonempty (the kernel's prologue loads its
  // first K tiles unconditionally)
  // (bit 4 with bit 1 only for the lower triangle of the output, whose
  // tiles have n0 <= m0)
  if (((ktrim & 1) && N > K) || ((ktrim & 2) && M > K) || ((ktrim & 4) && (ktrim & 2)) ||
      ((ktrim & 4) && (ktrim & 1) && mode != 1))
    return false;
  const long tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
  const long nwg = mode == 0 ? tilesM * tilesN : tri_tiles(mode, (int)tilesM, (int)tilesN);
  const int vecc = aligned16<double>(C, ldc) ? 1 : 0;
  hipLaunchKernelGGL(at ? gemm_tn_f64_x : (bt ? gemm_nt_f64_x : gemm_nn_f64_x),
                     dim3((unsigned)nwg, (unsigned)batch), dim3(NT), 0, s, (int)M, (int)N, (int)K,
                     alpha, A, lda, B, ldb, beta, C, ldc, vecc, mode, skip, sa, sb, sc, ktrim);
  ATLG_LAUNCHED();
  return true;
}

template void scale<double>(hipStream_t, long, long, double, double*, long, int, const int*);
template void scale<float>(hipStream_t, long, long, float, float*, long, int, const int*);
template void gemm<double>(hipStream_t, int, int, long, long, long, double, const double*, long,
                           const double*, long, double, double*, long, int, const int*, int);
template void gemm<float>(hipStream_t, int, int, long, long, long, float, const float*, long,
                          const float*, long, float, float*, long, int, const int*, int);
template bool gemm_narrow<double>(hipStream_t, int, int, long, long, long, double, const double*,
                                  long, const double*, long, double, double*, long, int,
                                  const int*);
template bool gemm_narrow<float>(hipStream_t, int, int, long, long, long, float, const float*,
                                 long, const float*, long, float, float*, long, int, const int*);

}  // namespace atlg
