// Dense FP64 kernels for the multifrontal factorisation, written for CDNA4
// (gfx950): batched tile-list GEMM on v_mfma_f64_16x16x4f64, a single-wave
// Cholesky + inverse of 64x64 diagonal blocks, and the row-strip triangular
// solve of the panel below each diagonal block.
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
template <bool TA, bool TB, int BM, int BN>
__global__ __launch_bounds__(256) void gemm_kernel(const GemmProb* __restrict__ probs,
                                                   const GemmTile* __restrict__ tiles, Bases bases) {
  constexpr int BK = 16, PAD = 4;
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
  __shared__ double As[2][BK][BM + PAD];
  __shared__ double Bs[2][BK][BN + PAD];
  const GemmTile t = tiles[blockIdx.x];
  const GemmProb P = probs[t.prob];
  const double* __restrict__ A = bases.p[P.a_base] + P.a_off;
  const double* __restrict__ B = bases.p[P.b_base] + P.b_off;
  double* __restrict__ C = bases.p[P.c_base] + P.c_off;
  const int M = P.M, N = P.N, K = P.K;
  const int m0 = t.tm * BM, n0 = t.tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc[TN][TM];
#pragma unroll
  for (int a = 0; a < TN; ++a)
#pragma unroll
    for (int b = 0; b < TM; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  double ra[LA], rb[LB];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int idx = tid + q * 256;
      const int mm = TA ? idx / BK : idx % BM, kk = TA ? idx % BK : idx / BM;
      const int gm = m0 + mm, gk = k0 + kk;
      ra[q] = (gm < M && gk < K) ? (TA ? A[gk + (int64_t)gm * P.lda] : A[gm + (int64_t)gk * P.lda]) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int idx = tid + q * 256;
      const int nn = TB ? idx % BN : idx / BK, kk = TB ? idx / BN : idx % BK;
      const int gn = n0 + nn, gk = k0 + kk;
      rb[q] = (gn < N && gk < K) ? (TB ? B[gn + (int64_t)gk * P.ldb] : B[gk + (int64_t)gn * P.ldb]) : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int idx = tid + q * 256;
      const int mm = TA ? idx / BK : idx % BM, kk = TA ? idx % BK : idx / BM;
      As[buf][kk][mm] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int idx = tid + q * 256;
      const int nn = TB ? idx % BN : idx / BK, kk = TB ? idx / BN : idx % BK;
      Bs[buf][kk][nn] = rb[q];
    }
  };
  const int nk = (K + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double a[TM], b[TN];
#pragma unroll
      for (int mi = 0; mi < TM; ++mi) a[mi] = As[cur][kk + lk][wm + mi * 16 + lr];
#pragma unroll
      for (int ni = 0; ni < TN; ++ni) b[ni] = Bs[cur][kk + lk][wn + ni * 16 + lr];
#pragma unroll
      for (int ni = 0; ni < TN; ++ni)
#pragma unroll
        for (int mi = 0; mi < TM; ++mi)
          acc[ni][mi] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ni], a[mi], acc[ni][mi], 0, 0, 0);
    }
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }
  const double alpha = P.alpha, beta = P.beta;
#pragma unroll
  for (int ni = 0; ni < TN; ++ni)
#pragma unroll
    for (int mi = 0; mi < TM; ++mi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + mi * 16 + lr;
        const int n = n0 + wn + ni * 16 + lk + 4 * v;
        if (m < M && n < N) {
          double* c = C + m + (int64_t)n * P.ldc;
          const double r = alpha * acc[ni][mi][v];
          *c = (beta == 0.0) ? r : r + beta * (*c);
        }
      }
}

template <int T>
static void launch_gemm_t(const GemmLaunch& g, const Bases& b, hipStream_t s) {
  dim3 grid(g.ntiles), block(256);
  if (!g.ta && !g.tb)
    hipLaunchKernelGGL((gemm_kernel<false, false, T, T>), grid, block, 0, s, g.d_probs, g.d_tiles, b);
  else if (!g.ta && g.tb)
    hipLaunchKernelGGL((gemm_kernel<false, true, T, T>), grid, block, 0, s, g.d_probs, g.d_tiles, b);
  else if (g.ta && !g.tb)
    hipLaunchKernelGGL((gemm_kernel<true, false, T, T>), grid, block, 0, s, g.d_probs, g.d_tiles, b);
  else
    hipLaunchKernelGGL((gemm_kernel<true, true, T, T>), grid, block, 0, s, g.d_probs, g.d_tiles, b);
}

void launch_gemm(const GemmLaunch& g, const Bases& b, hipStream_t s) {
  if (g.ntiles == 0) return;
  if (g.tile == 128)
    launch_gemm_t<128>(g, b, s);
  else
    launch_gemm_t<64>(g, b, s);
}

// ------------------------------------------------ diagonal block Cholesky
// One 256-thread block per 64x64 diagonal block.  Thread (r, g) keeps row r,
// columns 16g..16g+15 in registers.  Right-looking: at step j the owners of
// column j publish it to LDS and every thread applies the rank-1 update to
// its own entries (two barriers per column).  Then W = L^-1 by right-looking
// forward substitution (row k published per step), written to the node's Z
// block.  Blocks narrower than 64 are padded with the identity.  A
// non-positive pivot is recorded in *status as (global elimination column +
// 1) and replaced by 1 so the sweep completes.
__global__ __launch_bounds__(256) void potrf_diag_kernel(const PotrfDesc* __restrict__ descs,
                                                         double* __restrict__ F, double* __restrict__ Z,
                                                         int* __restrict__ status) {
  __shared__ double col[64];
  __shared__ double dg[64];
  __shared__ double Ls[64][65];
  __shared__ double wrow[64];
  const PotrfDesc d = descs[blockIdx.x];
  const int tid = threadIdx.x, r = tid & 63, cg = tid >> 6;
  const int kb = d.kb;
  double* Fb = F + d.foff + d.k0 + (int64_t)d.k0 * d.ld;
  double a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = 16 * cg + q;
    double v;
    if (r < kb && c < kb)
      v = (c <= r) ? Fb[r + (int64_t)c * d.ld] : 0.0;
    else
      v = (r == c) ? 1.0 : 0.0;
    a[q] = v;
  }
#pragma unroll 1
  for (int g = 0; g < 4; ++g) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int j = 16 * g + q;
      if (cg == g) col[r] = a[q];
      __syncthreads();
      double piv = col[j];
      if (!(piv > 0.0)) {
        if (tid == 0 && j < kb) atomicCAS(status, 0, d.colbase + j + 1);
        piv = 1.0;
      }
      const double dj = sqrt(piv), idj = 1.0 / dj;
      const double lr = col[r] * idj;
#pragma unroll
      for (int q2 = 0; q2 < 16; ++q2) {
        const int c = 16 * cg + q2;
        if (c > j && c <= r) a[q2] -= lr * (col[c] * idj);
      }
      if (tid == 0) dg[j] = dj;
      __syncthreads();
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = 16 * cg + q;
    const double v = (r > c) ? a[q] / dg[c] : (r == c ? dg[c] : 0.0);
    Ls[r][c] = v;
    if (r < kb && c < kb && c <= r) Fb[r + (int64_t)c * d.ld] = v;
  }
  __syncthreads();
  double w[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) w[q] = (r == 16 * cg + q) ? 1.0 : 0.0;
#pragma unroll 1
  for (int k = 0; k < 64; ++k) {
    if (r == k) {
      const double inv = 1.0 / Ls[k][k];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        w[q] *= inv;
        wrow[16 * cg + q] = w[q];
      }
    }
    __syncthreads();
    if (r > k) {
      const double l = Ls[r][k];
#pragma unroll
      for (int q = 0; q < 16; ++q) w[q] -= l * wrow[16 * cg + q];
    }
    __syncthreads();
  }
  if (r < kb) {
    double* Zb = Z + d.zoff + d.k0 + (int64_t)d.k0 * d.zld;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = 16 * cg + q;
      if (c < kb) Zb[r + (int64_t)c * d.zld] = w[q];
    }
  }
}

// ------------------------------------------------------------ panel solve
// Rows r0..r0+nr of the panel below a diagonal block: X = S W^T, in place.
__global__ __launch_bounds__(256) void panel_kernel(const PanelDesc* __restrict__ descs, double* __restrict__ F,
                                                    const double* __restrict__ Z) {
  __shared__ double S[64][65];
  __shared__ double W[64][65];
  const PanelDesc d = descs[blockIdx.x];
  const int tid = threadIdx.x;
  double* Fp = F + d.foff + d.r0 + (int64_t)d.k0 * d.ld;
  const double* Zb = Z + d.zoff + d.k0 + (int64_t)d.k0 * d.zld;
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    const int rr = idx & 63, k = idx >> 6;
    S[rr][k] = (rr < d.nr && k < d.kb) ? Fp[rr + (int64_t)k * d.ld] : 0.0;
    W[rr][k] = (rr < d.kb && k < d.kb && k <= rr) ? Zb[rr + (int64_t)k * d.zld] : 0.0;
  }
  __syncthreads();
  const int rr = tid & 63, cb = (tid >> 6) * 16;
  double out[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) out[q] = 0.0;
  for (int k = 0; k < 64; ++k) {
    const double s = S[rr][k];
#pragma unroll
    for (int q = 0; q < 16; ++q) out[q] += s * W[cb + q][k];
  }
  if (rr < d.nr) {
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (cb + q < d.kb) Fp[rr + (int64_t)(cb + q) * d.ld] = out[q];
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
void launch_panel(const PanelDesc* d, int n, double* F, const double* Z, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(panel_kernel, dim3(n), dim3(256), 0, s, d, F, Z);
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
  GemmProb P;
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
  launch_gemm(g, bs, (hipStream_t)stream);
  SPDE_HIP(hipStreamSynchronize((hipStream_t)stream));
  hipFree(g.d_probs);
  hipFree(g.d_tiles);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}
