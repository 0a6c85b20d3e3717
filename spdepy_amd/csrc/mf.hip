// Sparse <-> frontal data movement of the multifrontal method: scatter of Q
// into the fronts, extend-add of update blocks, log-determinant reduction,
// gathers for the triangular solves and the selected inverse.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace spde {

// F[map[e]] = q[e] for every lower structural entry.
__global__ void scatter_kernel(const double* __restrict__ q, const int64_t* __restrict__ map, int64_t nnz,
                               double* __restrict__ F) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = map[e];
    if (m >= 0) F[m] = q[e];
  }
}

// F[map[diag[r]]] += add[r]
__global__ void diag_add_kernel(const double* __restrict__ add, const int64_t* __restrict__ diag,
                                const int64_t* __restrict__ map, int64_t n, double* __restrict__ F) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    F[map[diag[r]]] += add[r];
}

// Extend-add of the lower triangle of a child's update block (rows/cols p..n
// of its front) into the parent front at relpos.  Tile 64x64 of the child's
// b x b block; tiles above the diagonal are not listed.  Phase 0 adds the
// columns that land in the parent's own columns, phase 1 the rest (into the
// parent's update block, after the parent's SYRK has written it).
__global__ __launch_bounds__(256) void extend_add_kernel(const ExtAddDesc* __restrict__ descs,
                                                         const ExtAddTile* __restrict__ tiles,
                                                         const int32_t* __restrict__ relpos,
                                                         double* __restrict__ F) {
  __shared__ int rpj[64];
  const ExtAddTile t = tiles[blockIdx.x];
  const ExtAddDesc d = descs[t.desc];
  const int32_t* rp = relpos + d.relpos_off;
  const double* U = F + d.cfoff + d.cp + (int64_t)d.cp * d.cld;
  double* P = F + d.pfoff;
  const int tid = threadIdx.x;
  if (tid < 64) rpj[tid] = (t.tj * 64 + tid < d.b) ? rp[t.tj * 64 + tid] : -1;
  __syncthreads();
  const int i = t.ti * 64 + (tid & 63);
  if (i >= d.b) return;
  const int pi = rp[i];
  const int jlo = t.phase ? d.split : 0, jhi = t.phase ? d.b : d.split;
#pragma unroll 4
  for (int jj = (tid >> 6); jj < 64; jj += 4) {
    const int j = t.tj * 64 + jj;
    if (j > i || j < jlo || j >= jhi) continue;
    P[pi + (int64_t)rpj[jj] * d.pld] += U[i + (int64_t)j * d.cld];
  }
}

// Zero the lower triangle of a front, one 64x64 tile (node, ti, tj) per
// block (tiles above the diagonal are not listed; the parts of the fronts
// above the diagonal are never written after the arena's first clearing).
__global__ __launch_bounds__(256) void zero_lower_kernel(const NodeDev* __restrict__ nodes, const int4* __restrict__ tiles,
                                                         double* __restrict__ F) {
  const int4 t = tiles[blockIdx.x];
  const NodeDev nd = nodes[t.x];
  double* f = F + nd.foff;
  const int i = t.y * 64 + (threadIdx.x & 63);
  if (i >= nd.n) return;
  for (int jj = threadIdx.x >> 6; jj < 64; jj += 4) {
    const int j = t.z * 64 + jj;
    if (j <= i) f[i + (int64_t)j * nd.n] = 0.0;
  }
}

// partial[b] = sum of log L_ii over the variables i of block b's range of
// kLogdetPer (the diagonal of L sits where the diagonal of Q was scattered,
// F[map[diag[i]]]); each thread takes kLogdetPer / 256 entries, all loads in
// flight at once.
constexpr int kLogdetPer = 256;
__global__ __launch_bounds__(256) void logdet_partial_kernel(const int64_t* __restrict__ diag,
                                                             const int64_t* __restrict__ map, int64_t n,
                                                             const double* __restrict__ F,
                                                             double* __restrict__ partial) {
  __shared__ double red[256];
  const int64_t i0 = blockIdx.x * (int64_t)kLogdetPer + threadIdx.x;
  double v[kLogdetPer / 256];
#pragma unroll
  for (int u = 0; u < kLogdetPer / 256; ++u) {
    const int64_t i = i0 + 256 * u;
    v[u] = i < n ? F[map[diag[i]]] : 1.0;
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < kLogdetPer / 256; ++u) s += log(v[u]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[0] = 2 * sum partial (fixed order: deterministic)
__global__ __launch_bounds__(256) void sum_kernel(const double* __restrict__ partial, int n, double scale,
                                                  double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) s += partial[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = scale * red[0];
}

// ----------------------------------------------------------------- solves
// xy[q, c] = B[perm[q], c]   (lattice order -> elimination order)
__global__ void permute_in_kernel(const double* __restrict__ B, int64_t ldb, const int32_t* __restrict__ perm,
                                  int64_t n, int nrhs, double* __restrict__ xy) {
  const int64_t tot = n * nrhs;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = e % n, c = e / n;
    xy[e] = B[perm[q] + c * ldb];
  }
}
__global__ void permute_out_kernel(const double* __restrict__ xy, const int32_t* __restrict__ perm, int64_t n,
                                   int nrhs, double* __restrict__ B, int64_t ldb) {
  const int64_t tot = n * nrhs;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = e % n, c = e / n;
    B[perm[q] + c * ldb] = xy[e];
  }
}

// Forward-solve front vectors of the nodes of a level: v[0:p] = xy[own],
// v[p:n] = 0, then the children's update parts are added (child slot order).
// One block per node.
__global__ __launch_bounds__(256) void fwd_gather_kernel(const int32_t* __restrict__ lvl_nodes, const NodeDev* __restrict__ nodes,
                                                         const int32_t* __restrict__ children, const int32_t* __restrict__ relpos,
                                                         const double* __restrict__ xy, int64_t n, int nrhs,
                                                         double* __restrict__ vec) {
  // block (node, column c): right-hand sides are independent, so a level's
  // gather spreads over nodes x columns instead of one block per node
  const int node = lvl_nodes[blockIdx.x];
  const int c = blockIdx.y;
  const NodeDev nd = nodes[node];
  double* v = vec + nd.voff * nrhs + (int64_t)c * nd.n;
  for (int r = threadIdx.x; r < nd.n; r += 256) v[r] = (r < nd.p) ? xy[nd.start + r + (int64_t)c * n] : 0.0;
  __syncthreads();
  for (int s = 0; s < 2; ++s) {
    const int ch = children[2 * node + s];
    if (ch < 0) continue;
    const NodeDev cd = nodes[ch];
    const int32_t* rp = relpos + cd.relpos_off;
    const double* u = vec + cd.voff * nrhs + (int64_t)c * cd.n;
    for (int r = threadIdx.x; r < cd.nb; r += 256) v[rp[r]] += u[cd.p + r];
    __syncthreads();
  }
}

// Backward-solve: v[0:p] = xy[own] (forward result), v[p:n] = xy[B] (solved
// ancestors).
__global__ __launch_bounds__(256) void bwd_gather_kernel(const int32_t* __restrict__ lvl_nodes, const NodeDev* __restrict__ nodes,
                                                         const int32_t* __restrict__ bnd, const double* __restrict__ xy,
                                                         int64_t n, int nrhs, double* __restrict__ vec) {
  const int node = lvl_nodes[blockIdx.x];
  const int c = blockIdx.y;  // one column per block, as the forward gather
  const NodeDev nd = nodes[node];
  double* v = vec + nd.voff * nrhs + (int64_t)c * nd.n;
  const int32_t* B = bnd + nd.relpos_off;
  for (int r = threadIdx.x; r < nd.n; r += 256) {
    const int64_t g = (r < nd.p) ? (int64_t)nd.start + r : (int64_t)B[r - nd.p];
    v[r] = xy[g + (int64_t)c * n];
  }
}

// --------------------------------------------------------- selected inverse
// Sigma22 of a node = symmetric gather from the lower triangle of the
// parent's Sigma front.  One block per 64x64 tile (node, ti, tj), ti >= tj,
// of the b x b block: relpos increases, so such a tile lies wholly in the
// parent's lower triangle (diagonal tiles: their lower half).  The block
// reads it once, consecutive threads on consecutive rows, writes it in place
// and writes its mirror (tj, ti) through an LDS transpose, again with
// consecutive threads on consecutive rows.  All sixteen reads of a thread
// are issued before the first store.
__global__ __launch_bounds__(256) void sigma_gather_kernel(const int32_t* __restrict__ lvl_nodes, const NodeDev* __restrict__ nodes,
                                                           const int32_t* __restrict__ relpos, double* __restrict__ S,
                                                           const int4* __restrict__ tiles) {
  __shared__ double tile[64][65];
  __shared__ int rpi[64], rpj[64];
  const int4 tt = tiles[blockIdx.x];
  const int node = lvl_nodes[tt.x];
  const NodeDev nd = nodes[node];
  const NodeDev pa = nodes[nd.parent];
  const int32_t* rp = relpos + nd.relpos_off;
  const int b = nd.nb;
  const int i0 = tt.y * 64, j0 = tt.z * 64;
  const bool diag = tt.y == tt.z;
  const double* src = S + pa.foff;
  const int64_t pn = pa.n;
  const int tid = threadIdx.x, u = tid & 63, w0 = tid >> 6;
  // parent positions of the tile's rows and columns (-1 outside the block)
  if (tid < 64) rpi[tid] = (i0 + tid < b) ? rp[i0 + tid] : -1;
  else if (tid < 128) rpj[tid - 64] = (j0 + tid - 64 < b) ? rp[j0 + tid - 64] : -1;
  __syncthreads();
  const int pi = rpi[u];
  double v[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int jj = w0 + 4 * t, pj = rpj[jj];
    v[t] = (pi >= 0 && pj >= 0 && (!diag || u >= jj)) ? src[pi + (int64_t)pj * pn] : 0.0;
  }
  double* dst = S + nd.foff + nd.p + (int64_t)nd.p * nd.n;
  const int i = i0 + u;
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int jj = w0 + 4 * t, j = j0 + jj;
    if (i < b && j < b && (!diag || u >= jj)) dst[i + (int64_t)j * nd.n] = v[t];
    tile[u][jj] = v[t];
  }
  __syncthreads();
  // mirror: entry (j0 + u, i0 + c) is the lower entry (i0 + c, j0 + u)
  const int row = j0 + u;
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int c = w0 + 4 * t, col = i0 + c;
    if (row < b && col < b && (!diag || u < c)) dst[row + (int64_t)col * nd.n] = tile[c][u];
  }
}

// out[e] = S[mapsym[e]]
__global__ void gather_pattern_kernel(const double* __restrict__ S, const int64_t* __restrict__ mapsym, int64_t nnz,
                                      double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = S[mapsym[e]];
}

static unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

void launch_scatter(const double* q, const int64_t* map, int64_t nnz, double* F, hipStream_t s) {
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, q, map, nnz, F);
}
void launch_diag_add(const double* add, const int64_t* diag, const int64_t* map, int64_t n, double* F, hipStream_t s) {
  hipLaunchKernelGGL(diag_add_kernel, dim3(grid_for(n)), dim3(256), 0, s, add, diag, map, n, F);
}
void launch_extend_add(const ExtAddDesc* d, const ExtAddTile* t, int ntiles, const int32_t* relpos, double* F,
                       hipStream_t s) {
  if (ntiles > 0) hipLaunchKernelGGL(extend_add_kernel, dim3(ntiles), dim3(256), 0, s, d, t, relpos, F);
}
int logdet_parts(int64_t n) { return (int)((n + kLogdetPer - 1) / kLogdetPer); }
void launch_logdet(const int64_t* diag, const int64_t* map, int64_t n, const double* F, double* partial, double* out,
                   hipStream_t s) {
  const int parts = logdet_parts(n);
  if (parts > 0) hipLaunchKernelGGL(logdet_partial_kernel, dim3(parts), dim3(256), 0, s, diag, map, n, F, partial);
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, partial, parts, 2.0, out);
}
void launch_permute_in(const double* B, int64_t ldb, const int32_t* perm, int64_t n, int nrhs, double* xy,
                       hipStream_t s) {
  hipLaunchKernelGGL(permute_in_kernel, dim3(grid_for(n * nrhs)), dim3(256), 0, s, B, ldb, perm, n, nrhs, xy);
}
void launch_permute_out(const double* xy, const int32_t* perm, int64_t n, int nrhs, double* B, int64_t ldb,
                        hipStream_t s) {
  hipLaunchKernelGGL(permute_out_kernel, dim3(grid_for(n * nrhs)), dim3(256), 0, s, xy, perm, n, nrhs, B, ldb);
}
void launch_fwd_gather(const int32_t* lvl_nodes, int nn, const NodeDev* nodes, const int32_t* children,
                       const int32_t* relpos, const double* xy, int64_t n, int nrhs, double* vec, hipStream_t s) {
  if (nn > 0)
    hipLaunchKernelGGL(fwd_gather_kernel, dim3(nn, nrhs), dim3(256), 0, s, lvl_nodes, nodes, children, relpos, xy, n,
                       nrhs, vec);
}
void launch_bwd_gather(const int32_t* lvl_nodes, int nn, const NodeDev* nodes, const int32_t* bnd, const double* xy,
                       int64_t n, int nrhs, double* vec, hipStream_t s) {
  if (nn > 0)
    hipLaunchKernelGGL(bwd_gather_kernel, dim3(nn, nrhs), dim3(256), 0, s, lvl_nodes, nodes, bnd, xy, n, nrhs, vec);
}
void launch_sigma_gather(const int32_t* lvl_nodes, const NodeDev* nodes, const int32_t* relpos, double* S,
                         const int4* tiles, int ntiles, hipStream_t s) {
  if (ntiles > 0)
    hipLaunchKernelGGL(sigma_gather_kernel, dim3(ntiles), dim3(256), 0, s, lvl_nodes, nodes, relpos, S, tiles);
}
void launch_zero_lower(const NodeDev* nodes, const int4* tiles, int ntiles, double* F, hipStream_t s) {
  if (ntiles > 0) hipLaunchKernelGGL(zero_lower_kernel, dim3(ntiles), dim3(256), 0, s, nodes, tiles, F);
}
void launch_gather_pattern(const double* S, const int64_t* mapsym, int64_t nnz, double* out, hipStream_t s) {
  hipLaunchKernelGGL(gather_pattern_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, S, mapsym, nnz, out);
}

}  // namespace spde
