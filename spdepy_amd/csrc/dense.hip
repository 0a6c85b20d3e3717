// Dense FP64 kernels for the multifrontal factorisation, written for CDNA4
// (gfx950): batched tile-list GEMM on v_mfma_f64_16x16x4f64 (which also does
// the panel solve below each diagonal block, X = S W^T with W = L11^-1) and
// the Cholesky factor + inverse of 64x64 diagonal blocks.
#include <hip/hip_runtime.h>

#include "internal.hpp"

#include <vector>

namespace spde {

typedef double d4 __attribute__((ext_vector_type(4)));

// ----------------------------------------------------------------- GEMM
// C(M x N) = alpha op(A) op(B) + beta C, column-major, FP64 MFMA.
// Block tile BM x BN x 16 (64 or 128), four waves in a 2x2 arrangement, each
// wave (BM/2) x (BN/2) = (BM/32) x (BN/32) MFMA 16x16 tiles.  LDS is double
// buffered and the next K-slice is prefetched into registers while the
// current one feeds the MFMAs.  The MFMA computes the transposed tile
// (B-fragment as the matrix "A" input) so that the 16 lanes of a row group own
// 16 consecutive rows of one column of C: stores are 128-byte coalesced in
// column-major memory.  Accumulator layout of v_mfma_f64_16x16x4f64
// (measured): lane l, register v holds D[l/16 + 4v][l%16].
// SPLIT: the block computes the chunk t.pad of `split` chunks of the tile's
// K range (each a multiple of BK deep) and writes the raw 64x64 partial tile
// to ws + blockIdx.x * 4096 (column-major, ld 64); split_reduce_kernel then
// applies alpha, beta and the lower-triangle rule.
// List entry of block b: the dispatcher hands blocks to the eight XCDs in
// turn (b and b + 8 share an XCD and its L2).  Inside every full group of 64
// blocks, XCD x takes the eight consecutive entries 8x .. 8x + 7, which share
// operand panels (a tile row of a problem is consecutive in the list); the
// list order, deepest tiles first, holds at the granularity of a group.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int g = b >> 6;
  if ((g + 1) * 64 > nb) return b;
  const int r = b & 63;
  return (g << 6) + ((r & 7) << 3) + (r >> 3);
}

template <bool TA, bool TB, int BM, int BN, bool SPLIT = false>
__device__ __forceinline__ void gemm_body(const GemmProb* __restrict__ probs, const GemmTile* __restrict__ tiles,
                                          const Bases& bases, int split = 1, double* __restrict__ ws = nullptr) {
  constexpr int BK = 16;
  // LDS layout follows each operand's memory order so that both the global
  // loads and the LDS stores are contiguous: an operand stored with m (n)
  // contiguous goes to a k-major tile (pitch BM+16: the four lane groups of a
  // fragment read land 32 banks apart), a k-contiguous one to an m-major
  // (n-major) tile with pitch BK+1.
  constexpr bool AK = !TA, BKM = TB;  // k-major tiles
  constexpr int PA = AK ? BM + 16 : BK + 1, PB = BKM ? BN + 16 : BK + 1;
  constexpr int SA = AK ? BK * PA : BM * PA, SB = BKM ? BK * PB : BN * PB;
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
  __shared__ double As[2][SA];
  __shared__ double Bs[2][SB];
  const GemmTile t = tiles[SPLIT ? (int)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x)];
  const GemmProb P = probs[t.prob];
  // A and C may be the same memory (the in-place panel solve X = S W^T,
  // one tile column): every tile reads all of its rows of A before its
  // epilogue writes them, and no tile reads rows another tile writes.
  const double* A = bases.p[P.a_base] + P.a_off;
  const double* __restrict__ B = bases.p[P.b_base] + P.b_off;
  double* C = bases.p[P.c_base] + P.c_off;
  const int M = P.M, N = P.N;
  const int m0 = t.tm * BM, n0 = t.tn * BN;
  int kb = 0, K = P.K;  // this block's K range [kb, K): the zeros of a triangular operand are skipped
  if (P.kmode == 1)
    kb = min(K, n0);
  else if (P.kmode == 2)
    kb = min(K, m0);
  else if (P.kmode == 3)
    K = min(K, m0 + BM);
  if constexpr (SPLIT) {
    const int kc = ((K - kb + split - 1) / split + 15) / 16 * 16;
    const int k0 = min(K, kb + t.pad * kc);
    K = min(K, k0 + kc);
    kb = k0;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;
  const bool interior = (m0 + BM <= M) && (n0 + BN <= N);
  const int am = TA ? tid / BK : tid % BM, ak = TA ? tid % BK : tid / BM;
  const int bn = TB ? tid % BN : tid / BK, bk = TB ? tid / BN : tid % BK;
  constexpr int ASTEP_M = TA ? 256 / BK : 0, ASTEP_K = TA ? 0 : 256 / BM;
  constexpr int BSTEP_N = TB ? 0 : 256 / BK, BSTEP_K = TB ? 256 / BN : 0;
  auto aidx = [](int k, int m) { return AK ? k * PA + m : m * PA + k; };
  auto bidx = [](int k, int n) { return BKM ? k * PB + n : n * PB + k; };
  d4 acc[TN][TM];
#pragma unroll
  for (int x = 0; x < TN; ++x)
#pragma unroll
    for (int y = 0; y < TM; ++y) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
  double ra[LA], rb[LB];
  auto load = [&](int k0) {
    k0 += kb;
    const bool full = interior && (k0 + BK <= K);
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int mm = am + q * ASTEP_M, kk = ak + q * ASTEP_K;
      const int gm = m0 + mm, gk = k0 + kk;
      const int64_t off = TA ? gk + (int64_t)gm * P.lda : gm + (int64_t)gk * P.lda;
      ra[q] = (full || (gm < M && gk < K)) ? A[off] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int nn = bn + q * BSTEP_N, kk = bk + q * BSTEP_K;
      const int gn = n0 + nn, gk = k0 + kk;
      const int64_t off = TB ? gn + (int64_t)gk * P.ldb : gk + (int64_t)gn * P.ldb;
      rb[q] = (full || (gn < N && gk < K)) ? B[off] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < LA; ++q) As[buf][aidx(ak + q * ASTEP_K, am + q * ASTEP_M)] = ra[q];
#pragma unroll
    for (int q = 0; q < LB; ++q) Bs[buf][bidx(bk + q * BSTEP_K, bn + q * BSTEP_N)] = rb[q];
  };
  const int nk = (K - kb + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
    double fa[2][TM], fb[2][TN];
#pragma unroll
    for (int mi = 0; mi < TM; ++mi) fa[0][mi] = As[cur][aidx(lk, wm + mi * 16 + lr)];
#pragma unroll
    for (int ni = 0; ni < TN; ++ni) fb[0][ni] = Bs[cur][bidx(lk, wn + ni * 16 + lr)];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int f = s & 1;
      if (s < 3) {
#pragma unroll
        for (int mi = 0; mi < TM; ++mi) fa[f ^ 1][mi] = As[cur][aidx(4 * (s + 1) + lk, wm + mi * 16 + lr)];
#pragma unroll
        for (int ni = 0; ni < TN; ++ni) fb[f ^ 1][ni] = Bs[cur][bidx(4 * (s + 1) + lk, wn + ni * 16 + lr)];
      }
#pragma unroll
      for (int ni = 0; ni < TN; ++ni)
#pragma unroll
        for (int mi = 0; mi < TM; ++mi)
          acc[ni][mi] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[f][ni], fa[f][mi], acc[ni][mi], 0, 0, 0);
    }
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }
  if constexpr (SPLIT) {
    double* w = ws + (int64_t)blockIdx.x * (BM * BN);
#pragma unroll
    for (int ni = 0; ni < TN; ++ni)
#pragma unroll
      for (int mi = 0; mi < TM; ++mi)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          w[(wm + mi * 16 + lr) + (wn + ni * 16 + lk + 4 * v) * BM] = acc[ni][mi][v];
    return;
  }
  const double alpha = P.alpha, beta = P.beta;
  // lower problems: the part of a diagonal tile above the diagonal is left
  // untouched (it is never read, and stays as initialised)
  const bool diag_tile = P.lower && t.tm == t.tn;
#pragma unroll
  for (int ni = 0; ni < TN; ++ni)
#pragma unroll
    for (int mi = 0; mi < TM; ++mi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + mi * 16 + lr;
        const int n = n0 + wn + ni * 16 + lk + 4 * v;
        if ((interior || (m < M && n < N)) && !(diag_tile && m < n)) {
          double* c = C + m + (int64_t)n * P.ldc;
          const double r = alpha * acc[ni][mi][v];
          *c = (beta == 0.0) ? r : r + beta * (*c);
        }
      }
}

// 64x64 tiles: four blocks per CU (LDS 38 KB, ~120 registers).
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm64_kernel(const GemmProb* __restrict__ probs,
                                                     const GemmTile* __restrict__ tiles, Bases bases) {
  gemm_body<TA, TB, 64, 64>(probs, tiles, bases);
}

// Split-K partial tiles (64x64) and their ordered sum.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm64_split_kernel(const GemmProb* __restrict__ probs,
                                                           const GemmTile* __restrict__ tiles, Bases bases, int split,
                                                           double* __restrict__ ws) {
  gemm_body<TA, TB, 64, 64, true>(probs, tiles, bases, split, ws);
}

// Block b: C tile of tiles[b * split] = alpha * sum_q partial[b * split + q]
// (q ascending) + beta C, inside the problem and, for lower problems, on or
// below the diagonal.
__global__ __launch_bounds__(256) void split_reduce_kernel(const GemmProb* __restrict__ probs,
                                                           const GemmTile* __restrict__ tiles, Bases bases, int split,
                                                           const double* __restrict__ ws) {
  const GemmTile t = tiles[(int64_t)blockIdx.x * split];
  const GemmProb P = probs[t.prob];
  double* C = bases.p[P.c_base] + P.c_off;
  const int m0 = t.tm * 64, n0 = t.tn * 64;
  const bool diag_tile = P.lower && t.tm == t.tn;
  const double* w = ws + (int64_t)blockIdx.x * split * 4096;
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int m = m0 + (e & 63), n = n0 + (e >> 6);
    if (m >= P.M || n >= P.N || (diag_tile && m < n)) continue;
    double acc = 0.0;
    for (int q = 0; q < split; ++q) acc += w[(int64_t)q * 4096 + e];
    double* c = C + m + (int64_t)n * P.ldc;
    const double r = P.alpha * acc;
    *c = (P.beta == 0.0) ? r : r + P.beta * (*c);
  }
}

// 128x128 tiles (each wave 64x64: half the LDS reads per MFMA): two blocks
// per CU (LDS 74 KB); the register budget is pinned to two waves per SIMD,
// without it the compiler takes ~360 registers and one wave.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm128_kernel(
    const GemmProb* __restrict__ probs, const GemmTile* __restrict__ tiles, Bases bases) {
  gemm_body<TA, TB, 128, 128>(probs, tiles, bases);
}

// y = alpha op(A) x + beta y for N <= kGemvMaxN columns (solves with few
// right-hand sides).  Bandwidth-bound: A is read once, coalesced.
// A not transposed (M x K): block = 64 rows (lane = row), the four waves take
// quarters of K, partial sums combined in LDS in wave order.
// A transposed (K x M stored): one wave per output row m (a column of A),
// lanes stride over k, butterfly sum across the wave.
// loads each lane issues before their products in the GEMV's inner loop
constexpr int kGemvDepth = 16;
template <bool TA, int NW = 4>
__global__ __launch_bounds__(NW * 64) void gemv_kernel(const GemmProb* __restrict__ probs,
                                                       const GemmTile* __restrict__ tiles, Bases bases) {
  const GemmTile t = tiles[blockIdx.x];
  const GemmProb P = probs[t.prob];
  const double* __restrict__ A = bases.p[P.a_base] + P.a_off;
  const double* __restrict__ X = bases.p[P.b_base] + P.b_off;
  double* __restrict__ Y = bases.p[P.c_base] + P.c_off;
  const int N = P.N, K = P.K;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc[kGemvMaxN];
#pragma unroll
  for (int c = 0; c < kGemvMaxN; ++c) acc[c] = 0.0;
  if constexpr (!TA) {
    __shared__ double red[NW][kGemvMaxN][64];
    const int m = t.tm * 64 + lane;
    const int q = (K + NW - 1) / NW, k0 = wave * q, k1 = min(K, k0 + q);
    if (m < P.M) {
      // batches of kGemvDepth loads issued before their products, so that
      // many columns' worth of A are in flight per lane; the sums still run
      // in ascending k
      const double* __restrict__ a = A + m + (int64_t)k0 * P.lda;
      int k = k0;
      for (; k + kGemvDepth <= k1; k += kGemvDepth, a += (int64_t)kGemvDepth * P.lda) {
        double av[kGemvDepth];
#pragma unroll
        for (int u = 0; u < kGemvDepth; ++u) av[u] = a[(int64_t)u * P.lda];
#pragma unroll
        for (int u = 0; u < kGemvDepth; ++u)
#pragma unroll
          for (int c = 0; c < kGemvMaxN; ++c)
            if (c < N) acc[c] += av[u] * X[k + u + (int64_t)c * P.ldb];
      }
      for (; k < k1; ++k, a += P.lda) {
        const double av = *a;
#pragma unroll
        for (int c = 0; c < kGemvMaxN; ++c)
          if (c < N) acc[c] += av * X[k + (int64_t)c * P.ldb];
      }
    }
#pragma unroll
    for (int c = 0; c < kGemvMaxN; ++c) red[wave][c][lane] = acc[c];
    __syncthreads();
    if (wave == 0 && m < P.M)
      for (int c = 0; c < N; ++c) {
        double v = red[0][c][lane];
#pragma unroll
        for (int w = 1; w < NW; ++w) v += red[w][c][lane];
        double* y = Y + m + (int64_t)c * P.ldc;
        const double r = P.alpha * v;
        *y = (P.beta == 0.0) ? r : r + P.beta * (*y);
      }
  } else {
    const int m = t.tm * 4 + wave;
    if (m >= P.M) return;
    const double* a = A + (int64_t)m * P.lda;
#pragma unroll 8
    for (int k = lane; k < K; k += 64) {
      const double av = a[k];
#pragma unroll
      for (int c = 0; c < kGemvMaxN; ++c)
        if (c < N) acc[c] += av * X[k + (int64_t)c * P.ldb];
    }
#pragma unroll
    for (int c = 0; c < kGemvMaxN; ++c)
      for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, 64);
    if (lane == 0)
      for (int c = 0; c < N; ++c) {
        double* y = Y + m + (int64_t)c * P.ldc;
        const double r = P.alpha * acc[c];
        *y = (P.beta == 0.0) ? r : r + P.beta * (*y);
      }
  }
}

void launch_gemm(const GemmLaunch& g, const Bases& b, hipStream_t s, unsigned lds) {
  if (g.ntiles == 0) return;
  dim3 grid(g.ntiles), block(256);
  if (g.gemv) {
    if (g.ta)
      hipLaunchKernelGGL(gemv_kernel<true>, grid, block, 0, s, g.d_probs, g.d_tiles, b);
    else if (g.gemv_wide)
      hipLaunchKernelGGL((gemv_kernel<false, 16>), grid, dim3(1024), 0, s, g.d_probs, g.d_tiles, b);
    else
      hipLaunchKernelGGL(gemv_kernel<false>, grid, block, 0, s, g.d_probs, g.d_tiles, b);
    return;
  }
  if (g.split > 1) {
    double* ws = b.ws + g.ws_off;
    if (!g.ta && !g.tb)
      hipLaunchKernelGGL((gemm64_split_kernel<false, false>), grid, block, 0, s, g.d_probs, g.d_tiles, b, g.split, ws);
    else if (!g.ta && g.tb)
      hipLaunchKernelGGL((gemm64_split_kernel<false, true>), grid, block, 0, s, g.d_probs, g.d_tiles, b, g.split, ws);
    else if (g.ta && !g.tb)
      hipLaunchKernelGGL((gemm64_split_kernel<true, false>), grid, block, 0, s, g.d_probs, g.d_tiles, b, g.split, ws);
    else
      hipLaunchKernelGGL((gemm64_split_kernel<true, true>), grid, block, 0, s, g.d_probs, g.d_tiles, b, g.split, ws);
    hipLaunchKernelGGL(split_reduce_kernel, dim3(g.ntiles / g.split), block, 0, s, g.d_probs, g.d_tiles, b, g.split,
                       ws);
    return;
  }
  if (g.tile == 128) {
    if (!g.ta && !g.tb)
      hipLaunchKernelGGL((gemm128_kernel<false, false>), grid, block, lds, s, g.d_probs, g.d_tiles, b);
    else if (!g.ta && g.tb)
      hipLaunchKernelGGL((gemm128_kernel<false, true>), grid, block, lds, s, g.d_probs, g.d_tiles, b);
    else if (g.ta && !g.tb)
      hipLaunchKernelGGL((gemm128_kernel<true, false>), grid, block, lds, s, g.d_probs, g.d_tiles, b);
    else
      hipLaunchKernelGGL((gemm128_kernel<true, true>), grid, block, lds, s, g.d_probs, g.d_tiles, b);
  } else {
    if (!g.ta && !g.tb)
      hipLaunchKernelGGL((gemm64_kernel<false, false>), grid, block, lds, s, g.d_probs, g.d_tiles, b);
    else if (!g.ta && g.tb)
      hipLaunchKernelGGL((gemm64_kernel<false, true>), grid, block, lds, s, g.d_probs, g.d_tiles, b);
    else if (g.ta && !g.tb)
      hipLaunchKernelGGL((gemm64_kernel<true, false>), grid, block, lds, s, g.d_probs, g.d_tiles, b);
    else
      hipLaunchKernelGGL((gemm64_kernel<true, true>), grid, block, lds, s, g.d_probs, g.d_tiles, b);
  }
}

// ------------------------------------------------ diagonal block Cholesky
// One 256-thread block per 64x64 diagonal block.  Thread (r, g) (lane r of
// wave g) keeps row r, columns c = 4q + g (q = 0..15), in registers, so each
// step of four pivot columns j0..j0+3 has exactly one of them in every
// thread.  Step: the threads publish their pivot-column entry (one LDS write
// each, double buffered), every thread forms the 4x4 pivot block's factor
// itself, finishes its row's four entries, writes its own one to the L copy
// Y in LDS, and the rank-4 update of its columns right of the pivot block,
// with the finished rows read back from Y (broadcast reads, no
// recomputation), follows in the next step.  Entries right of the diagonal are
// never read, so the update runs unconditionally over them.  W = L^-1 is
// then built by recursive blocking in LDS (below).  Pivots use v_rsq_f64 with
// two Newton steps (relative error below 2^-52) instead of a square root and
// a division.  Blocks narrower than 64 are padded with the identity.  A
// non-positive pivot is recorded in *status as (global elimination column
// + 1) and replaced by 1 so the sweep completes.
struct Piv4 {
  double i0, i1, i2, i3, l10, l20, l30, l21, l31, l32, d0, d1, d2, d3;
};

// Keeps an updated register value where it is computed: without it the
// compiler defers the rank-4 updates of later columns to their first use and
// keeps every step's operands live (hundreds of registers, spilled).
__device__ inline void pin(double& v) { asm volatile("" : "+v"(v)); }

// 1/sqrt(x) for x > 0 (hardware estimate refined twice), x for the factor.
__device__ inline double pivot_rsq(double x, bool& bad) {
  if (!(x > 0.0)) {
    bad = true;
    x = 1.0;
  }
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = y * (1.5 - hx * y * y);
  y = y * (1.5 - hx * y * y);
  return y;
}

// 4x4 Cholesky factor of D(i, j) = D[i * P + j] (lower part read).
template <int P>
__device__ inline void chol4(const double* D, Piv4& c, int& badcol) {
  bool b0 = false, b1 = false, b2 = false, b3 = false;
  double x = D[0];
  c.i0 = pivot_rsq(x, b0);
  c.d0 = b0 ? 1.0 : x * c.i0;
  c.l10 = D[P] * c.i0;
  c.l20 = D[2 * P] * c.i0;
  c.l30 = D[3 * P] * c.i0;
  x = D[P + 1] - c.l10 * c.l10;
  c.i1 = pivot_rsq(x, b1);
  c.d1 = b1 ? 1.0 : x * c.i1;
  c.l21 = (D[2 * P + 1] - c.l20 * c.l10) * c.i1;
  c.l31 = (D[3 * P + 1] - c.l30 * c.l10) * c.i1;
  x = D[2 * P + 2] - c.l20 * c.l20 - c.l21 * c.l21;
  c.i2 = pivot_rsq(x, b2);
  c.d2 = b2 ? 1.0 : x * c.i2;
  c.l32 = (D[3 * P + 2] - c.l30 * c.l20 - c.l31 * c.l21) * c.i2;
  x = D[3 * P + 3] - c.l30 * c.l30 - c.l31 * c.l31 - c.l32 * c.l32;
  c.i3 = pivot_rsq(x, b3);
  c.d3 = b3 ? 1.0 : x * c.i3;
  badcol = b0 ? 0 : (b1 ? 1 : (b2 ? 2 : (b3 ? 3 : -1)));
}

__global__ __launch_bounds__(256) void potrf_diag_kernel(const PotrfDesc* __restrict__ descs,
                                                         double* __restrict__ F, double* __restrict__ Z,
                                                         int* __restrict__ status) {
  constexpr int YP = 66;                // pitch of Y (16-byte aligned rows)
  __shared__ double pub[2][64 * 4];     // pivot columns, row-major [r][0..3]
  __shared__ double Y[64 * YP];         // finished L (zeros right of the diagonal)
  __shared__ double idg[64];            // 1 / L(k, k)
  const PotrfDesc d = descs[blockIdx.x];
  const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
  const int kb = d.kb;
  double* Fb = F + d.foff + d.k0 + (int64_t)d.k0 * d.ld;
  double a[16];
  const int rr = r < kb ? r : kb - 1;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = 4 * q + g;
    const int cc = c < kb ? c : kb - 1;
    const double v = Fb[rr + (int64_t)cc * d.ld];  // clamped: always inside the block
    a[q] = (r < kb && c < kb) ? (c <= r ? v : 0.0) : (r == c ? 1.0 : 0.0);
  }
  // The rank-4 update of step s - 1 is applied to the next pivot column
  // first and to the later columns during step s, beside its serial 4x4
  // factor and row solve (same operations in the same order per entry).
  double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int j0 = 4 * s, buf = s & 1, jp = j0 - 4;
    if (s > 0) {
      const double* yc = &Y[(4 * s + g) * YP + jp];
      a[s] -= p0 * yc[0] + p1 * yc[1] + p2 * yc[2] + p3 * yc[3];
    }
    pub[buf][4 * r + g] = a[s];
    __syncthreads();
    Piv4 c;
    int bad;
    chol4<4>(&pub[buf][4 * j0], c, bad);
    if (bad >= 0 && tid == 0 && j0 + bad < kb) atomicCAS(status, 0, d.colbase + j0 + bad + 1);
    if (s > 0) {
#pragma unroll
      for (int q = s + 1; q < 16; ++q) {
        const double* yc = &Y[(4 * q + g) * YP + jp];
        a[q] -= p0 * yc[0] + p1 * yc[1] + p2 * yc[2] + p3 * yc[3];
      }
    }
    // row r's four entries; on the pivot rows this yields the factor's
    // entries left of and on the diagonal, the rest (and all of the rows
    // above) are cleared
    const double* pr = &pub[buf][4 * r];
    double y0 = pr[0] * c.i0;
    double y1 = (pr[1] - y0 * c.l10) * c.i1;
    double y2 = (pr[2] - y0 * c.l20 - y1 * c.l21) * c.i2;
    double y3 = (pr[3] - y0 * c.l30 - y1 * c.l31 - y2 * c.l32) * c.i3;
    y0 = (r >= j0) ? y0 : 0.0;
    y1 = (r >= j0 + 1) ? y1 : 0.0;
    y2 = (r >= j0 + 2) ? y2 : 0.0;
    y3 = (r >= j0 + 3) ? y3 : 0.0;
    const double mine = (g == 0) ? y0 : (g == 1 ? y1 : (g == 2 ? y2 : y3));
    a[s] = mine;
    Y[r * YP + j0 + g] = mine;
    if (tid < 4) idg[j0 + tid] = (tid == 0) ? c.i0 : (tid == 1 ? c.i1 : (tid == 2 ? c.i2 : c.i3));
    __syncthreads();
    p0 = y0;
    p1 = y1;
    p2 = y2;
    p3 = y3;
#pragma unroll
    for (int q = s + 1; q < 16; ++q) pin(a[q]);
  }
  if (r < kb) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = 4 * q + g;
      if (c < kb && c <= r) Fb[r + (int64_t)c * d.ld] = a[q];
    }
  }
  // W = L^-1 by recursive blocking: the four 16x16 diagonal blocks are
  // inverted by column-oriented forward substitution (wave g, lane j < 16:
  // column j of block g), then W21 = -W22 (L21 W11) for the two 32x32 and
  // the 64x64 block rows, the products spread over the block.  Wl is
  // row-major with zeros above the diagonal; blocks beyond kb hold the
  // identity padding of L, so W is the identity there too.
  constexpr int WP = 65;
  __shared__ double Wl[64 * WP];
  __shared__ double Xs[32 * 33];
  for (int e = tid; e < 64 * WP; e += 256) Wl[e] = 0.0;
  __syncthreads();
  if (r < 16) {
    const int o = 16 * g, j = r;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      x[kk] *= idg[o + kk];
#pragma unroll
      for (int i = kk + 1; i < 16; ++i) x[i] -= Y[(o + i) * YP + o + kk] * x[kk];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) Wl[(o + i) * WP + o + j] = x[i];
  }
  __syncthreads();
  {
    // 32 level, pair P = rows/cols 32P .. 32P + 31: X = L(o+16.., o..) W00,
    // W10 = -W11 X; thread: pair tid >> 7, row i, columns j0, j0 + 1
    const int P = tid >> 7, t = tid & 127, i = t >> 3, j0 = (t & 7) * 2, o = 32 * P;
    double x0 = 0.0, x1 = 0.0;
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      const double l = Y[(o + 16 + i) * YP + o + kk];
      x0 += l * Wl[(o + kk) * WP + o + j0];
      x1 += l * Wl[(o + kk) * WP + o + j0 + 1];
    }
    Xs[(P * 16 + i) * 33 + j0] = x0;
    Xs[(P * 16 + i) * 33 + j0 + 1] = x1;
    __syncthreads();
    double w0 = 0.0, w1 = 0.0;
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      const double v = Wl[(o + 16 + i) * WP + o + 16 + kk];
      w0 += v * Xs[(P * 16 + kk) * 33 + j0];
      w1 += v * Xs[(P * 16 + kk) * 33 + j0 + 1];
    }
    Wl[(o + 16 + i) * WP + o + j0] = -w0;
    Wl[(o + 16 + i) * WP + o + j0 + 1] = -w1;
  }
  __syncthreads();
  {
    // 64 level: X = L(32.., 0..31) W_aa, W_ba = -W_bb X; thread: column
    // j = tid & 31, rows i0 + 8u
    const int j = tid & 31, i0 = tid >> 5;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 32; ++kk) {
      const double wv = Wl[kk * WP + j];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] += Y[(32 + i0 + 8 * u) * YP + kk] * wv;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) Xs[(i0 + 8 * u) * 33 + j] = x[u];
    __syncthreads();
    double w[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 32; ++kk) {
      const double xv = Xs[kk * 33 + j];
#pragma unroll
      for (int u = 0; u < 4; ++u) w[u] += Wl[(32 + i0 + 8 * u) * WP + 32 + kk] * xv;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) Wl[(32 + i0 + 8 * u) * WP + j] = -w[u];
  }
  __syncthreads();
  double* Zb = Z + d.zoff + d.k0 + (int64_t)d.k0 * d.zld;
  if (r < kb) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = 4 * q + g;
      if (c < kb) Zb[r + (int64_t)c * d.zld] = Wl[r * WP + c];
    }
  }
}

__global__ void fill_kernel(double* __restrict__ p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Lane-layout probe of v_mfma_f64_16x16x4f64 (64 lanes, 4 outputs each).
// which=0: a = 1 on lanes 0..15, b = lane+1 on lanes 0..15;
// which=1: a = lane+1 on lanes 0..15, b = 1 on lanes 0..15.
__global__ __launch_bounds__(64) void mfma_probe_kernel(int which, double* __restrict__ out) {
  const int l = threadIdx.x;
  double a = 0.0, b = 0.0;
  if (l < 16) {
    a = which == 0 ? 1.0 : (double)(l + 1);
    b = which == 0 ? (double)(l + 1) : 1.0;
  }
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  for (int v = 0; v < 4; ++v) out[4 * l + v] = acc[v];
}

void launch_potrf(const PotrfDesc* d, int n, double* F, double* Z, int* status, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(potrf_diag_kernel, dim3(n), dim3(256), 0, s, d, F, Z, status);
}
void launch_fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

}  // namespace spde

using namespace spde;

extern "C" int spde_debug_mfma_probe(int which, double* out, void* stream) {
  if (!out || which < 0 || which > 1) {
    set_error("spde_debug_mfma_probe: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, which, out);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

extern "C" int spde_debug_gemm(int ta, int tb, int M, int N, int K, double alpha, const double* A, int lda,
                               const double* B, int ldb, double beta, double* C, int ldc, int tile, void* stream) {
  if (M <= 0 || N <= 0 || K < 0 || !A || !B || !C || lda < (ta ? K : M) || ldb < (tb ? N : K) || ldc < M ||
      (tile != 64 && tile != 128)) {
    set_error("spde_debug_gemm: invalid argument");
    return SPDE_ERR_ARG;
  }
  GemmProb P{};  // every field zero (kmode 0: no triangular operand)
  P.a_off = P.b_off = P.c_off = 0;
  P.lda = lda;
  P.ldb = ldb;
  P.ldc = ldc;
  P.M = M;
  P.N = N;
  P.K = K;
  P.a_base = 0;
  P.b_base = 1;
  P.c_base = 2;
  P.lower = 0;
  P.alpha = alpha;
  P.beta = beta;
  std::vector<GemmTile> tiles;
  for (int a = 0; a < (M + tile - 1) / tile; ++a)
    for (int b = 0; b < (N + tile - 1) / tile; ++b) tiles.push_back({0, a, b, 0});
  GemmLaunch g;
  g.tile = tile;
  g.ta = ta != 0;
  g.tb = tb != 0;
  g.ntiles = (int)tiles.size();
  SPDE_HIP(hipMalloc((void**)&g.d_probs, sizeof(GemmProb)));
  SPDE_HIP(hipMalloc((void**)&g.d_tiles, sizeof(GemmTile) * tiles.size()));
  SPDE_HIP(hipMemcpy(g.d_probs, &P, sizeof(GemmProb), hipMemcpyHostToDevice));
  SPDE_HIP(hipMemcpy(g.d_tiles, tiles.data(), sizeof(GemmTile) * tiles.size(), hipMemcpyHostToDevice));
  Bases bs;
  for (int q = 0; q < kNumBases; ++q) bs.p[q] = nullptr;
  bs.p[0] = const_cast<double*>(A);
  bs.p[1] = const_cast<double*>(B);
  bs.p[2] = C;
  launch_gemm(g, bs, (hipStream_t)stream, 0);
  SPDE_HIP(hipStreamSynchronize((hipStream_t)stream));
  hipFree(g.d_probs);
  hipFree(g.d_tiles);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}
