// Separable space-time models: Q = Qt (x) Qs with a tridiagonal Qt
// (seperable_spatial_temporal2D.py:64-124, _ha2D.py, _idiffusion2D.py:62-98).
//
// The plan is a wide-time lattice: row (t,i) holds the spatial neighbours j of
// cell i (the one-slice row i, in the same order) once in each of the slices
// t-1, t, t+1 (check_wide_layout), so entry ((t,i),(t+d,j)) sits at
//     indptr[(t,i)] + (t > 0 ? cnt_i : 0) + d cnt_i + q
// where q is the position of j in the one-slice row i and cnt_i its length.
//
// Qt is passed as qt[2T-1] = diagonal | off-diagonal (symmetric).
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace spde {

namespace {

__device__ inline double qt_coef(const double* qt, int T, int t, int d) {
  return d == 0 ? qt[t] : qt[T + (d < 0 ? t - 1 : t)];
}

// out[e] = Qt[t, t+d] Qs[i, j] for every structural entry of the wide plan.
// Thread per one-slice entry s = (i, q) and slice t.
__global__ __launch_bounds__(256) void kron_expand_kernel(int64_t Ns, int T, const int64_t* __restrict__ indptr,
                                                          const int64_t* __restrict__ sp_indptr,
                                                          const int32_t* __restrict__ sp_row,
                                                          int64_t sp_nnz, const double* __restrict__ qs,
                                                          const double* __restrict__ qt, double* __restrict__ out) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (s >= sp_nnz) return;
  const int64_t i = sp_row[s];
  const int64_t b0 = sp_indptr[i], cnt = sp_indptr[i + 1] - b0, q = s - b0;
  const double v = qs[s];
  const int64_t base = indptr[(int64_t)t * Ns + i] + (t > 0 ? cnt : 0) + q;
  for (int d = -1; d <= 1; ++d) {
    if (t + d < 0 || t + d >= T) continue;
    out[base + d * cnt] = qt_coef(qt, T, t, d) * v;
  }
}

// weights[s] = ws sigma_s[s] + sum_t sum_d Qt[t,t+d] (-1/2 r sigma_c[e] - 1/2 sum_c mu mu)
// (thread per one-slice entry; fixed summation order).
__global__ __launch_bounds__(256) void kron_weights_kernel(int64_t Ns, int64_t n, int T, int r,
                                                           const int64_t* __restrict__ indptr,
                                                           const int64_t* __restrict__ sp_indptr,
                                                           const int32_t* __restrict__ sp_row,
                                                           const int32_t* __restrict__ sp_col, int64_t sp_nnz,
                                                           const double* __restrict__ qt,
                                                           const double* __restrict__ sigma_s, double ws,
                                                           const double* __restrict__ sigma_c,
                                                           const double* __restrict__ mu,
                                                           double* __restrict__ weights) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= sp_nnz) return;
  const int64_t i = sp_row[s], j = sp_col[s];
  const int64_t b0 = sp_indptr[i], cnt = sp_indptr[i + 1] - b0, q = s - b0;
  double acc = 0.0;
  for (int t = 0; t < T; ++t) {
    const int64_t row = (int64_t)t * Ns + i;
    const int64_t base = indptr[row] + (t > 0 ? cnt : 0) + q;
    for (int d = -1; d <= 1; ++d) {
      if (t + d < 0 || t + d >= T) continue;
      const int64_t col = (int64_t)(t + d) * Ns + j;
      double mm = 0.0;
      for (int c = 0; c < r; ++c) mm += mu[row + (int64_t)c * n] * mu[col + (int64_t)c * n];
      acc += qt_coef(qt, T, t, d) * (-0.5 * r * sigma_c[base + d * cnt] - 0.5 * mm);
    }
  }
  weights[s] = ws * sigma_s[s] + acc;
}

// Block (t, d+1): the three sums of one time pair over all one-slice entries.
__global__ __launch_bounds__(256) void kron_blocks_kernel(int64_t Ns, int64_t n, int T, int r,
                                                          const int64_t* __restrict__ indptr,
                                                          const int64_t* __restrict__ sp_indptr,
                                                          const int32_t* __restrict__ sp_row,
                                                          const int32_t* __restrict__ sp_col, int64_t sp_nnz,
                                                          const double* __restrict__ qs,
                                                          const double* __restrict__ sigma_c,
                                                          const double* __restrict__ mu,
                                                          const double* __restrict__ diag_add,
                                                          double* __restrict__ blocks) {
  __shared__ double red[3][256];
  const int t = blockIdx.x, d = (int)blockIdx.y - 1;
  double sc = 0.0, sm = 0.0, st = 0.0;
  if (t + d >= 0 && t + d < T) {
    for (int64_t s = threadIdx.x; s < sp_nnz; s += blockDim.x) {
      const int64_t i = sp_row[s], j = sp_col[s];
      const int64_t b0 = sp_indptr[i], cnt = sp_indptr[i + 1] - b0, q = s - b0;
      const int64_t row = (int64_t)t * Ns + i, col = (int64_t)(t + d) * Ns + j;
      const int64_t e = indptr[row] + (t > 0 ? cnt : 0) + d * cnt + q;
      const double v = qs[s];
      sc += sigma_c[e] * v;
      double mm = 0.0;
      for (int c = 0; c < r; ++c) mm += mu[row + (int64_t)c * n] * mu[col + (int64_t)c * n];
      sm += mm * v;
      if (d == 0 && j == i && diag_add) st += diag_add[row] * sigma_c[e];
    }
  }
  red[0][threadIdx.x] = sc;
  red[1][threadIdx.x] = sm;
  red[2][threadIdx.x] = st;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* o = blocks + ((int64_t)t * 3 + d + 1) * 3;
    o[0] = red[0][0];
    o[1] = red[1][0];
    o[2] = red[2][0];
  }
}

// Checks the pair of plans; builds the row and column of every one-slice
// entry on the spatial plan at its first use.
int kron_side(const spde_plan* pl, spde_plan* sp, const char* who) {
  if (!pl || pl->host_only || !sp || sp->host_only || !pl->lat.wide || sp->lat.T != 1 || sp->lat.M != pl->lat.M ||
      sp->lat.N != pl->lat.N || sp->lat.bc != pl->lat.bc) {
    set_error(std::string(who) + ": needs a wide-time plan and the one-slice plan of the same M, N, bc");
    return SPDE_ERR_ARG;
  }
  const int64_t Ns = sp->lat.Ns();
  for (int64_t i = 0; i < Ns; ++i) {  // the wide rows (0, i) repeat the one-slice rows
    const int64_t cnt = sp->st.indptr[i + 1] - sp->st.indptr[i];
    const int64_t len = pl->st.indptr[i + 1] - pl->st.indptr[i];
    if (len != cnt * (pl->lat.T > 1 ? 2 : 1)) {
      set_error(std::string(who) + ": the one-slice plan does not match the wide-time plan's rows");
      return SPDE_ERR_ARG;
    }
  }
  if (sp->d_sprow) return SPDE_OK;
  const int64_t nnz = (int64_t)sp->st.indices.size();
  std::vector<int32_t> row(nnz), col(nnz);
  for (int64_t i = 0; i < Ns; ++i)
    for (int64_t e = sp->st.indptr[i]; e < sp->st.indptr[i + 1]; ++e) {
      row[e] = (int32_t)i;
      col[e] = sp->st.indices[e];
    }
  SPDE_HIP(hipMalloc((void**)&sp->d_sprow, sizeof(int32_t) * nnz));
  SPDE_HIP(hipMalloc((void**)&sp->d_spcol, sizeof(int32_t) * nnz));
  SPDE_HIP(hipMemcpy(sp->d_sprow, row.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  SPDE_HIP(hipMemcpy(sp->d_spcol, col.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  return SPDE_OK;
}

}  // namespace

}  // namespace spde

using namespace spde;

extern "C" {

int spde_assemble_kron(spde_plan* pl, spde_plan* sp, const double* qs, const double* qt, double* out,
                       void* stream) {
  if (!qs || !qt || !out) {
    set_error("spde_assemble_kron: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (int rc = kron_side(pl, sp, "spde_assemble_kron")) return rc;
  const int64_t nnz = (int64_t)sp->st.indices.size();
  const int T = pl->lat.T;
  hipLaunchKernelGGL(kron_expand_kernel, dim3((unsigned)((nnz + 255) / 256), (unsigned)T), dim3(256), 0,
                     (hipStream_t)stream, pl->lat.Ns(), T, pl->d_indptr, sp->d_indptr, sp->d_sprow, nnz, qs, qt, out);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_kron_reduce(spde_plan* pl, spde_plan* sp, const double* qs, const double* qt, const double* sigma_s,
                     double ws, const double* sigma_c, const double* mu, int r, const double* diag_add,
                     double* weights, double* blocks, void* stream) {
  if (!qs || !qt || !sigma_s || !sigma_c || !mu || r < 1 || !weights || !blocks) {
    set_error("spde_kron_reduce: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (int rc = kron_side(pl, sp, "spde_kron_reduce")) return rc;
  const int64_t nnz = (int64_t)sp->st.indices.size();
  const int T = pl->lat.T;
  const int64_t Ns = pl->lat.Ns(), n = pl->sym.n;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(kron_weights_kernel, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s, Ns, n, T, r,
                     pl->d_indptr, sp->d_indptr, sp->d_sprow, sp->d_spcol, nnz, qt, sigma_s, ws, sigma_c, mu, weights);
  hipLaunchKernelGGL(kron_blocks_kernel, dim3((unsigned)T, 3), dim3(256), 0, s, Ns, n, T, r, pl->d_indptr,
                     sp->d_indptr, sp->d_sprow, sp->d_spcol, nnz, qs, sigma_c, mu, diag_add, blocks);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

}  // extern "C"
