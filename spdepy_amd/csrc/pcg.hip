// Mixed-precision iterative path for Q_c = Q + diag(diag_add) on a space-time
// lattice: an incomplete Cholesky factor IC(0) on Q's structural pattern in a
// multi-colour ordering, FP32 preconditioned conjugate gradients with k
// right-hand sides, FP64 iterative refinement, Lanczos for stochastic
// log-determinants, Rademacher probes and the Hutchinson estimate of Q_c^-1
// on the pattern.
//
// The reference evaluates log|Q_c| and tr(Q_c^-1 dQ) with CHOLMOD and a
// Hutchinson estimator (advection_idiffusion2D.py:173-189,
// var_advection_var_diffusion2D.py:205-221); this path replaces the
// factorisation of Q_c by PCG solves.
//
// Colouring.  Row (i, j, t) of Q couples to the 5x5 spatial window of its own
// slice and the 3x3 windows of the neighbouring slices, so the colour
// (i mod 3) + 3 (j mod 3) + 9 (t mod 2) (18 colours) never repeats inside a
// row: rows of one colour are mutually independent, and IC(0) in colour order
// is computed, and its triangular sweeps applied, one colour per launch.
//
// Factor storage.  Per row i, the lower list holds the entries (i, j) with
// colour(j) < colour(i), sorted by (colour(j), j), with L_ij; the upper list
// the entries with colour(j) > colour(i), with L_ji (gathered from row j's
// lower list).  Vectors with k right-hand sides are row-major n x k (the k
// values of a row contiguous), k a power of two <= 64: the factor and matrix
// values of a row are read once for all right-hand sides.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

// Pattern, colouring and factor of one precision: built by spde_ic_create,
// written by spde_ic_factor / spde_ic_correct, read by every solve.  A view
// (spde_ic_view) takes these fields from its parent at each call.
struct IcShared {
  spde_plan* pl = nullptr;
  int64_t n = 0, nnz = 0, nlow = 0;
  int64_t max_low = 0;               // longest lower list of a row
  int M = 0, N = 0, T = 0;
  std::vector<int64_t> coff;        // 19 offsets into crow
  int32_t* d_crow = nullptr;
  int64_t *d_lptr = nullptr, *d_uptr = nullptr, *d_lsrc = nullptr, *d_ulow = nullptr, *d_lch = nullptr;
  int32_t *d_lcol = nullptr, *d_ucol = nullptr;
  int64_t* d_lkey = nullptr;         // colour * n + index of each lower entry's column
  double *d_l64 = nullptr, *d_dg64 = nullptr, *d_adiag = nullptr;  // factor (FP64), Q_c diagonal
  float *d_l32 = nullptr, *d_u32 = nullptr, *d_dinv = nullptr, *d_qc32 = nullptr;
  const double* qval = nullptr;      // Q values of the last factorisation (structural order)
  const double* diag_add = nullptr;
  bool factored = false;
};
// Host copies of the lists (spde_ic_export), kept by the handle that built them.
struct IcHost {
  std::vector<int32_t> h_colour;    // colour per lattice index
  std::vector<int64_t> h_lptr, h_uptr, h_lsrc, h_lch;
  std::vector<int32_t> h_lcol, h_ucol;
  std::vector<int64_t> h_ulow;
};
// Buffers of the solve in flight: every handle, parent or view, has its own,
// so a parent and its views can solve on different streams at once.
struct IcWork {
  int* d_status = nullptr;
  double* d_part = nullptr;          // reduction partials [kRedBlocks * 64 * 2]
  double* d_scal = nullptr;          // per right-hand side scalars [64 * 8]
  int* d_active = nullptr;           // [64]
  int* d_iters = nullptr;            // [64]
  int* d_state = nullptr;            // [64] per right-hand side after a block residual replacement
  float* d_work = nullptr;           // 4 vectors n x 64 (PCG: r, z, p, q; Lanczos: q_prev, q_cur, w, u)
  double* d_r64 = nullptr;           // refinement residual (n x 64)
  float *d_b32 = nullptr, *d_x32 = nullptr;  // inner right-hand side and solution (n x 64)
  double* d_rn = nullptr;            // refinement: ||r||^2 and threshold per right-hand side [2 * 64]
  double* d_lanczos = nullptr;       // alpha, beta [2 * steps * 64]
  int64_t lanczos_cap = 0;
  double* d_klw = nullptr;           // L'^T x of the KL loss (n x K)
  int64_t kl_cap = 0;
  double* d_hist = nullptr;          // PCG alpha, beta per iteration and right-hand side [2 cap 64]
  int hist_cap = 0;
  hipGraphExec_t it_graph[7] = {};   // PCG iterations on the refinement buffers, per log2 of the rhs count
};
struct spde_ic : IcShared, IcHost, IcWork {
  spde_ic* parent = nullptr;  // set on a view
  int nviews = 0;             // live views of this handle
  bool released = false;      // destroyed while views were live: freed with the last view
};

namespace spde {

constexpr int kColours = 18;
constexpr int kRedBlocks = 2048;     // blocks of the reduction kernels (8 waves per SIMD on 256 CUs)
constexpr int kPairs = 22;           // correction channels per cell: diagonal + 21 forward-half pairs

// Scalars per right-hand side in d_scal[s * 64 + r].
enum : int { S_RZ = 0, S_PQ = 1, S_ALPHA = 2, S_BETA = 3, S_RR = 4, S_BN = 5, S_TOL2 = 6, S_X = 7 };

__device__ inline int lat_colour(int64_t idx, int M, int N) {
  const int64_t Ns = (int64_t)M * N;
  const int64_t t = idx / Ns, rem = idx - t * Ns;
  const int j = (int)(rem / M), i = (int)(rem - (int64_t)j * M);
  return (i % 3) + 3 * (j % 3) + 9 * (int)(t & 1);
}

// ------------------------------------------------------------ IC(0)
// One thread per row of colour c.  L_ij = (A_ij - sum_k L_ik L_jk) / L_jj over
// the k adjacent to both with colour(k) < colour(j), in (colour, index) order
// of the row's lower list; L_ii = sqrt((1 + shift) A_ii - sum_j L_ij^2).  The
// common k are found by one merge of the two sorted lists (lkey = colour * n +
// index of each entry), in row j's order.
__global__ __launch_bounds__(256) void ic_factor_kernel(int64_t r0, int64_t r1, const int32_t* __restrict__ crow,
                                                        const int64_t* __restrict__ lptr,
                                                        const int32_t* __restrict__ lcol,
                                                        const int64_t* __restrict__ lkey,
                                                        const int64_t* __restrict__ lsrc,
                                                        const double* __restrict__ qval,
                                                        const double* __restrict__ adiag, double shift,
                                                        double* __restrict__ l64, double* __restrict__ dg64,
                                                        int* __restrict__ status) {
  const int64_t rr = r0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (rr >= r1) return;
  const int i = crow[rr];
  const int64_t lo = lptr[i], hi = lptr[i + 1];
  double ssq = 0.0;
  for (int64_t q = lo; q < hi; ++q) {
    const int j = lcol[q];
    double a = qval[lsrc[q]];
    int64_t sq = lo, p = lptr[j];
    const int64_t pe = lptr[j + 1];
    while (sq < q && p < pe) {
      const int64_t ks = lkey[sq], kp = lkey[p];
      if (ks == kp) {
        a -= l64[sq] * l64[p];
        ++sq;
        ++p;
      } else if (ks < kp) {
        ++sq;
      } else {
        ++p;
      }
    }
    const double v = a / dg64[j];
    l64[q] = v;
    ssq += v * v;
  }
  const double d = (1.0 + shift) * adiag[i] - ssq;
  if (!(d > 0.0)) {
    atomicMax(status, i + 1);
    dg64[i] = 1.0;
  } else {
    dg64[i] = sqrt(d);
  }
}

// The same factor with one wave per row (rows of at most kIcRowMax lower
// entries): for each entry q in turn, the lanes take the entries of row j's
// lower list, find each one's column among the row's finished entries lo..q-1
// by binary search on lkey, and the products are summed over the wave in a
// fixed order.  The row's finished values stay in LDS (written by lane 0, read
// by the other lanes of the same wave later in program order).
constexpr int kIcRowMax = 128;
__global__ __launch_bounds__(256) void ic_factorw_kernel(int64_t r0, int64_t r1, const int32_t* __restrict__ crow,
                                                         const int64_t* __restrict__ lptr,
                                                         const int32_t* __restrict__ lcol,
                                                         const int64_t* __restrict__ lkey,
                                                         const int64_t* __restrict__ lsrc,
                                                         const double* __restrict__ qval,
                                                         const double* __restrict__ adiag, double shift,
                                                         double* __restrict__ l64, double* __restrict__ dg64,
                                                         int* __restrict__ status) {
  __shared__ double vals[4][kIcRowMax];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t rr = r0 + blockIdx.x * (int64_t)4 + wave;
  if (rr >= r1) return;
  const int i = crow[rr];
  const int64_t lo = lptr[i], hi = lptr[i + 1];
  double* v_row = vals[wave];
  double ssq = 0.0;
  for (int64_t q = lo; q < hi; ++q) {
    const int j = lcol[q];
    const double aq = qval[lsrc[q]];
    const int64_t pe = lptr[j + 1];
    double part = 0.0;
    for (int64_t p = lptr[j] + lane; p < pe; p += 64) {
      const int64_t kp = lkey[p];
      int64_t a = lo, b = q;  // first key >= kp in [lo, q)
      while (a < b) {
        const int64_t mid = (a + b) >> 1;
        if (lkey[mid] < kp) a = mid + 1;
        else b = mid;
      }
      if (a < q && lkey[a] == kp) part += v_row[a - lo] * l64[p];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    const double v = (aq - part) / dg64[j];
    if (lane == 0) {
      v_row[q - lo] = v;
      l64[q] = v;
    }
    __builtin_amdgcn_wave_barrier();
    ssq += v * v;
  }
  if (lane == 0) {
    const double d = (1.0 + shift) * adiag[i] - ssq;
    if (!(d > 0.0)) {
      atomicMax(status, i + 1);
      dg64[i] = 1.0;
    } else {
      dg64[i] = sqrt(d);
    }
  }
}

// Q_c diagonal (FP64) per row, and Q_c values in FP32 on the structural CSR.
__global__ void qc_values_kernel(int64_t n, const int64_t* __restrict__ indptr, const int64_t* __restrict__ diagpos,
                                 const double* __restrict__ qval, const double* __restrict__ dadd,
                                 double* __restrict__ adiag, float* __restrict__ qc32) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t dp = diagpos[i];
  for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) qc32[e] = (float)(qval[e] + (e == dp && dadd ? dadd[i] : 0.0));
  adiag[i] = qval[dp] + (dadd ? dadd[i] : 0.0);
}

// FP32 copies of the factor: lower values, upper values (mirrors), 1 / L_ii.
__global__ void ic_cast_lower_kernel(int64_t nlow, const double* __restrict__ l64, float* __restrict__ l32) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < nlow) l32[q] = (float)l64[q];
}
__global__ void ic_cast_upper_kernel(int64_t nlow, const int64_t* __restrict__ ulow, const double* __restrict__ l64,
                                     float* __restrict__ u32) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < nlow) u32[q] = (float)l64[ulow[q]];
}
__global__ void ic_cast_diag_kernel(int64_t n, const double* __restrict__ dg64, float* __restrict__ dinv) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dinv[i] = (float)(1.0 / dg64[i]);
}

// Learned correction in the Jacobi-scaled space (A~ = D^-1/2 A D^-1/2,
// L = D^1/2 L~ row-scaled by sqrt(A_ii) of the row): L~_ij += corr, and
// L~_ii *= exp(corr) for the diagonal channel.  corr[cell * 22 + ch] with the
// channel of each lower entry in lch (-1: no channel).
__global__ void ic_correct_lower_kernel(int64_t n, const int64_t* __restrict__ lptr, const int64_t* __restrict__ lch,
                                        const double* __restrict__ adiag, const float* __restrict__ corr,
                                        double* __restrict__ l64) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = sqrt(adiag[i]);
  for (int64_t q = lptr[i]; q < lptr[i + 1]; ++q) {
    const int64_t c = lch[q];
    if (c >= 0) l64[q] += s * (double)corr[c];
  }
}
__global__ void ic_correct_diag_kernel(int64_t n, const float* __restrict__ corr, double* __restrict__ dg64) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dg64[i] *= exp((double)corr[i * kPairs]);
}

// ------------------------------------------------------------ reductions
// partial[b][r] over threads of block b, for r = thread & (k-1).
template <int NV>
__device__ inline void block_reduce_store(double (&v)[NV], int k, double* __restrict__ part) {
  __shared__ double red[NV][256];
#pragma unroll
  for (int a = 0; a < NV; ++a) red[a][threadIdx.x] = v[a];
  __syncthreads();
  // threads with the same (thread & (k-1)) share a right-hand side
  for (int stride = 128; stride >= k; stride >>= 1) {
    if ((int)threadIdx.x < stride)
#pragma unroll
      for (int a = 0; a < NV; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + stride];
    __syncthreads();
  }
  if ((int)threadIdx.x < k)
#pragma unroll
    for (int a = 0; a < NV; ++a) part[((int64_t)a * kRedBlocks + blockIdx.x) * 64 + threadIdx.x] = red[a][threadIdx.x];
}

// ------------------------------------------------------------ SpMM, sweeps
// y = A x for A on the structural CSR (FP32), k right-hand sides; when part is
// set, the block partials of x^T y per right-hand side.
__global__ __launch_bounds__(256) void spmm32_kernel(int64_t n, int logk, const int64_t* __restrict__ indptr,
                                                     const int32_t* __restrict__ indices, const float* __restrict__ a,
                                                     const float* __restrict__ x, float* __restrict__ y,
                                                     double* __restrict__ part) {
  const int k = 1 << logk;
  const int64_t units = n << logk;
  double acc[1] = {0.0};
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = u >> logk;
    const int r = (int)(u & (k - 1));
    float s = 0.f;
    for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) s += a[e] * x[((int64_t)indices[e] << logk) + r];
    y[u] = s;
    acc[0] += (double)s * (double)x[u];
  }
  if (part) block_reduce_store<1>(acc, k, part);
}

// FP64 residual r = b - (Q + diag(dadd)) x for one right-hand side.
__global__ __launch_bounds__(256) void resid64_kernel(int64_t n, const int64_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ indices,
                                                      const int64_t* __restrict__ diagpos,
                                                      const double* __restrict__ qval, const double* __restrict__ dadd,
                                                      const double* __restrict__ b, const double* __restrict__ x,
                                                      double* __restrict__ r, double* __restrict__ part) {
  double acc[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) s += qval[e] * x[indices[e]];
    if (dadd) s += dadd[i] * x[i];
    const double v = b[i] - s;
    r[i] = v;
    acc[0] += v * v;
  }
  block_reduce_store<1>(acc, 1, part);
}

// Forward sweep of one colour: z_i = (r_i - sum_lower L_ij z_j) / L_ii.
__global__ __launch_bounds__(256) void fwd_kernel(int64_t r0, int64_t r1, int logk, const int32_t* __restrict__ crow,
                                                  const int64_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
                                                  const float* __restrict__ l32, const float* __restrict__ dinv,
                                                  const float* r, float* z, int fold) {
  const int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t rr = r0 + (u >> logk);
  if (rr >= r1) return;
  const int rhs = (int)(u & ((1 << logk) - 1));
  const int i = crow[rr];
  float s = r[((int64_t)i << logk) + rhs];
  for (int64_t p = lptr[i]; p < lptr[i + 1]; ++p) s -= l32[p] * z[((int64_t)lcol[p] << logk) + rhs];
  float v = s * dinv[i];
  if (fold) v *= dinv[i];
  z[((int64_t)i << logk) + rhs] = v;
}

// Backward sweep of one colour: x_i = (y_i - sum_upper L_ji x_j) / L_ii
// (y and x may be the same array).
__global__ __launch_bounds__(256) void bwd_kernel(int64_t r0, int64_t r1, int logk, const int32_t* __restrict__ crow,
                                                  const int64_t* __restrict__ uptr, const int32_t* __restrict__ ucol,
                                                  const float* __restrict__ u32, const float* __restrict__ dinv,
                                                  const float* y, float* x) {
  const int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t rr = r0 + (u >> logk);
  if (rr >= r1) return;
  const int rhs = (int)(u & ((1 << logk) - 1));
  const int i = crow[rr];
  float s = y[((int64_t)i << logk) + rhs];
  for (int64_t p = uptr[i]; p < uptr[i + 1]; ++p) s -= u32[p] * x[((int64_t)ucol[p] << logk) + rhs];
  x[((int64_t)i << logk) + rhs] = s * dinv[i];
}

// One right-hand side: a row is shared by kRowLanes consecutive lanes (entries
// strided by kRowLanes, then a fixed butterfly over the group), so a colour's
// rows are not each a single thread's chain of dependent gathers.
constexpr int kRowLanes = 8;
// Sum over p = p0, p0 + kRowLanes, ... < p1 of w[p] v[col[p]] in that order,
// four entries' loads issued before their products, so a row's few entries
// per lane cost one round of dependent loads instead of one per entry.
__device__ inline float strided_dot(int64_t p0, int64_t p1, const int32_t* __restrict__ col,
                                    const float* __restrict__ w, const float* v) {
  float acc = 0.f;
  for (int64_t p = p0; p < p1; p += 4 * kRowLanes) {
    int c[4];
    float a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool in = p + q * kRowLanes < p1;
      c[q] = in ? col[p + q * kRowLanes] : 0;
      a[q] = in ? w[p + q * kRowLanes] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) b[q] = v[c[q]];  // index 0 past the end: a valid address, value unused
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (p + q * kRowLanes < p1) acc += a[q] * b[q];
  }
  return acc;
}
__global__ __launch_bounds__(256) void fwd1_kernel(int64_t r0, int64_t r1, const int32_t* __restrict__ crow,
                                                   const int64_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
                                                   const float* __restrict__ l32, const float* __restrict__ dinv,
                                                   const float* r, float* z, int fold) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t rr = r0 + g / kRowLanes;
  const int lane = (int)(g % kRowLanes);
  if (rr >= r1) return;  // uniform over the group
  const int i = crow[rr];
  float acc = strided_dot(lptr[i] + lane, lptr[i + 1], lcol, l32, z);
  for (int o = kRowLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kRowLanes);
  if (lane == 0) {
    float v = (r[i] - acc) * dinv[i];
    if (fold) v *= dinv[i];
    z[i] = v;
  }
}
// (y and x may be the same array: row i reads only rows of other colours)
__global__ __launch_bounds__(256) void bwd1_kernel(int64_t r0, int64_t r1, const int32_t* __restrict__ crow,
                                                   const int64_t* __restrict__ uptr, const int32_t* __restrict__ ucol,
                                                   const float* __restrict__ u32, const float* __restrict__ dinv,
                                                   const float* y, float* x) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t rr = r0 + g / kRowLanes;
  const int lane = (int)(g % kRowLanes);
  if (rr >= r1) return;
  const int i = crow[rr];
  float acc = strided_dot(uptr[i] + lane, uptr[i + 1], ucol, u32, x);
  for (int o = kRowLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kRowLanes);
  if (lane == 0) x[i] = (y[i] - acc) * dinv[i];
}
// y = A x for one right-hand side, a group of kRowLanes lanes per row (grid-
// stride over the rows); the block partials of x^T y as spmm32_kernel's.
__global__ __launch_bounds__(256) void spmm1_kernel(int64_t n, const int64_t* __restrict__ indptr,
                                                    const int32_t* __restrict__ indices, const float* __restrict__ a,
                                                    const float* __restrict__ x, float* __restrict__ y,
                                                    double* __restrict__ part) {
  const int lane = (int)(threadIdx.x % kRowLanes);
  const int64_t groups = (int64_t)gridDim.x * blockDim.x / kRowLanes;
  double acc[1] = {0.0};
  for (int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kRowLanes; i < n; i += groups) {
    float s = strided_dot(indptr[i] + lane, indptr[i + 1], indices, a, x);
    for (int o = kRowLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kRowLanes);
    if (lane == 0) {
      y[i] = s;
      acc[0] += (double)s * (double)x[i];
    }
  }
  if (part) block_reduce_store<1>(acc, 1, part);
}

// K = 16, 32 or 64 right-hand sides: a wave per row, lane r < K = right-hand
// side r (the other lanes idle).  The row index is made wave-uniform
// (readfirstlane), so the row's pattern and values come through the scalar
// unit and each neighbour costs one contiguous 4K-byte vector load; eight
// (SpMM) or four (sweeps) neighbours' loads are issued before their sums.
template <int K>
__global__ __launch_bounds__(256) void spmmw_kernel(int64_t n, const int64_t* __restrict__ indptr,
                                                    const int32_t* __restrict__ indices,
                                                    const float* __restrict__ a, const float* __restrict__ x,
                                                    float* __restrict__ y, double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const bool on = lane < K;
  const int wpb = blockDim.x >> 6;
  double acc[1] = {0.0};
  for (int64_t row = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6); row < n; row += (int64_t)gridDim.x * wpb) {
    const int i = __builtin_amdgcn_readfirstlane((int)row);
    const int64_t e1 = indptr[i + 1];
    int64_t e = indptr[i];
    float s = 0.f;
    for (; e + 8 <= e1; e += 8) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = on ? x[(int64_t)indices[e + q] * K + lane] : 0.f;
#pragma unroll
      for (int q = 0; q < 8; ++q) s += a[e + q] * v[q];
    }
    for (; e < e1; ++e) s += a[e] * (on ? x[(int64_t)indices[e] * K + lane] : 0.f);
    if (on) {
      const int64_t u = (int64_t)i * K + lane;
      y[u] = s;
      acc[0] += (double)s * (double)x[u];
    }
  }
  if (part) block_reduce_store<1>(acc, K, part);
}
// The colour sweeps with K right-hand sides, a wave per row as above.
template <int K>
__global__ __launch_bounds__(256) void fwdw_kernel(int64_t r0, int64_t r1, const int32_t* __restrict__ crow,
                                                   const int64_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
                                                   const float* __restrict__ l32, const float* __restrict__ dinv,
                                                   const float* r, float* z, int fold) {
  const int lane = threadIdx.x & 63;
  const int64_t rw = r0 + (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (rw >= r1 || lane >= K) return;  // the row is uniform over the wave
  const int i = crow[__builtin_amdgcn_readfirstlane((int)rw)];
  float s = r[(int64_t)i * K + lane];
  const int64_t p1 = lptr[i + 1];
  int64_t p = lptr[i];
  for (; p + 4 <= p1; p += 4) {
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = z[(int64_t)lcol[p + q] * K + lane];
#pragma unroll
    for (int q = 0; q < 4; ++q) s -= l32[p + q] * v[q];
  }
  for (; p < p1; ++p) s -= l32[p] * z[(int64_t)lcol[p] * K + lane];
  float v = s * dinv[i];
  if (fold) v *= dinv[i];
  z[(int64_t)i * K + lane] = v;
}
template <int K>
__global__ __launch_bounds__(256) void bwdw_kernel(int64_t r0, int64_t r1, const int32_t* __restrict__ crow,
                                                   const int64_t* __restrict__ uptr, const int32_t* __restrict__ ucol,
                                                   const float* __restrict__ u32, const float* __restrict__ dinv,
                                                   const float* y, float* x) {
  const int lane = threadIdx.x & 63;
  const int64_t rw = r0 + (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (rw >= r1 || lane >= K) return;
  const int i = crow[__builtin_amdgcn_readfirstlane((int)rw)];
  float s = y[(int64_t)i * K + lane];
  const int64_t p1 = uptr[i + 1];
  int64_t p = uptr[i];
  for (; p + 4 <= p1; p += 4) {
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = x[(int64_t)ucol[p + q] * K + lane];
#pragma unroll
    for (int q = 0; q < 4; ++q) s -= u32[p + q] * v[q];
  }
  for (; p < p1; ++p) s -= u32[p] * x[(int64_t)ucol[p] * K + lane];
  x[(int64_t)i * K + lane] = s * dinv[i];
}

// ------------------------------------------------------------ PCG steps
// Per right-hand side sums of the reduction partials, out[a][r] = sum_b
// part[a][b][r] for a < NA, r < k (k <= 64), by a whole block of kRedThreads:
// with kp the power of two >= k, thread t = q kp + r adds blocks q, q + G,
// ... (G = kRedThreads / kp lanes per right-hand side, all 1024 for one), then
// the G values of each r are added by a fixed tree over q (deterministic).
// `out` is shared memory; every thread returns after the block has it.
constexpr int kRedThreads = 1024;
template <int NA>
__device__ inline void reduce_parts(const double* __restrict__ part, int k, int nblk, double (*out)[64]) {
  __shared__ double red[NA][kRedThreads];
  int kp = 1;
  while (kp < k) kp <<= 1;
  const int t = threadIdx.x, r = t & (kp - 1), q = t / kp, G = kRedThreads / kp;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    double acc = 0.0;
    if (r < k)
      for (int b = q; b < nblk; b += G) acc += part[((int64_t)a * kRedBlocks + b) * 64 + r];
    red[a][t] = acc;
  }
  __syncthreads();
  for (int st = kRedThreads / 2; st >= kp; st >>= 1) {
    if (t < st)
#pragma unroll
      for (int a = 0; a < NA; ++a) red[a][t] += red[a][t + st];
    __syncthreads();
  }
  if (t < kp)
#pragma unroll
    for (int a = 0; a < NA; ++a) out[a][t] = red[a][t];
  __syncthreads();
}

// out[a] = sum_b part[a][b][r] for one right-hand side r by a whole block of
// kRedThreads (thread t adds blocks t, t + kRedThreads, ..., then a fixed
// tree: deterministic); valid in thread 0.  The per-column scalar kernels
// run one such block per right-hand side.
template <int NA>
__device__ inline void reduce_col(const double* __restrict__ part, int r, int nblk, double (&out)[NA]) {
  __shared__ double red[NA][kRedThreads];
  const int t = threadIdx.x;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    double acc = 0.0;
    for (int b = t; b < nblk; b += kRedThreads) acc += part[((int64_t)a * kRedBlocks + b) * 64 + r];
    red[a][t] = acc;
  }
  __syncthreads();
  for (int st = kRedThreads / 2; st > 0; st >>= 1) {
    if (t < st)
#pragma unroll
      for (int a = 0; a < NA; ++a) red[a][t] += red[a][t + st];
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < NA; ++a) out[a] = red[a][0];
}

// x = 0, p = z, rz = r^T z, bn = b^T b (r = b), active = bn > 0.
// One block per right-hand side r = blockIdx.x (as the kernels below).
__global__ __launch_bounds__(kRedThreads) void pcg_init_scalars_kernel(int nblk, const double* __restrict__ part,
                                                                      double tol2, double* __restrict__ sc,
                                                                      int* __restrict__ active,
                                                                      int* __restrict__ iters) {
  double v[2];
  reduce_col<2>(part, blockIdx.x, nblk, v);
  const int r = blockIdx.x;
  if (threadIdx.x) return;
  const double rz = v[0], bn = v[1];
  sc[S_RZ * 64 + r] = rz;
  sc[S_BN * 64 + r] = bn;
  sc[S_TOL2 * 64 + r] = tol2;
  sc[S_RR * 64 + r] = bn;
  active[r] = bn > 0.0 ? 1 : 0;
  iters[r] = 0;
}

__global__ __launch_bounds__(kRedThreads) void pcg_alpha_kernel(int nblk, const double* __restrict__ part,
                                                               double* __restrict__ sc, const int* __restrict__ active) {
  double v[1];
  reduce_col<1>(part, blockIdx.x, nblk, v);
  const int r = blockIdx.x;
  if (threadIdx.x) return;
  const double pq = v[0];
  sc[S_PQ * 64 + r] = pq;
  sc[S_ALPHA * 64 + r] = (active[r] && pq > 0.0) ? sc[S_RZ * 64 + r] / pq : 0.0;
}

__global__ __launch_bounds__(256) void pcg_update_kernel(int64_t units, int logk, const double* __restrict__ sc,
                                                         const float* __restrict__ p, const float* __restrict__ q,
                                                         float* __restrict__ x, float* __restrict__ r) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const float a = (float)sc[S_ALPHA * 64 + (int)(u & ((1 << logk) - 1))];
    x[u] += a * p[u];
    r[u] -= a * q[u];
  }
}

// Partials of r^T z and r^T r (two sums per right-hand side).
__global__ __launch_bounds__(256) void dot2_kernel(int64_t units, int logk, const float* __restrict__ r,
                                                   const float* __restrict__ z, double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const double rv = r[u];
    acc[0] += rv * (double)z[u];
    acc[1] += rv * rv;
  }
  block_reduce_store<2>(acc, 1 << logk, part);
}

// After a residual replacement (one right-hand side): r^T z and r^T r of the
// replaced residual; the search direction is kept.
__global__ __launch_bounds__(kRedThreads) void pcg_reseed_kernel(int nblk, const double* __restrict__ part,
                                                                double* __restrict__ sc, int* __restrict__ active) {
  __shared__ double sums[2][64];
  reduce_parts<2>(part, 1, nblk, sums);
  if (threadIdx.x != 0) return;
  sc[S_RZ * 64] = sums[0][0];
  sc[S_RR * 64] = sums[1][0];
  active[0] = sums[0][0] > 0.0 ? 1 : 0;
}

// After a block residual replacement, one block per right-hand side r: state
// 0 stops the column, 1 keeps its search direction with r^T z and r^T r of
// the replaced residual, 2 restarts it from them (p = z by copy_cols_kernel).
__global__ __launch_bounds__(kRedThreads) void pcg_replace_scalars_kernel(int nblk, const double* __restrict__ part,
                                                                         const int* __restrict__ state,
                                                                         double* __restrict__ sc,
                                                                         int* __restrict__ active) {
  double v[2];
  reduce_col<2>(part, blockIdx.x, nblk, v);
  const int r = blockIdx.x;
  if (threadIdx.x) return;
  if (state[r] == 0) {
    active[r] = 0;
    return;
  }
  sc[S_RZ * 64 + r] = v[0];
  sc[S_RR * 64 + r] = v[1];
  active[r] = v[0] > 0.0 ? 1 : 0;
}

// beta = rz_new / rz; convergence: ||r||^2 <= tol^2 ||b||^2 stops a right-hand side.
// hist (optional, cap iterations): the step's alpha and beta per right-hand
// side at hist[(2 j) 64 + r], hist[(2 j + 1) 64 + r] for its j-th iteration,
// the coefficients that give the Lanczos tridiagonal of L^-1 Q_c L^-T.
__global__ __launch_bounds__(kRedThreads) void pcg_beta_kernel(int nblk, const double* __restrict__ part,
                                                              double* __restrict__ sc, int* __restrict__ active,
                                                              int* __restrict__ iters, double* __restrict__ hist,
                                                              int cap) {
  double v[2];
  reduce_col<2>(part, blockIdx.x, nblk, v);
  const int r = blockIdx.x;
  if (threadIdx.x) return;
  const double rz = v[0], rr = v[1];
  if (!active[r]) {
    sc[S_BETA * 64 + r] = 0.0;
    return;
  }
  const int j = iters[r];
  iters[r] = j + 1;
  sc[S_RR * 64 + r] = rr;
  const double old = sc[S_RZ * 64 + r];
  const double beta = old != 0.0 ? rz / old : 0.0;
  sc[S_BETA * 64 + r] = beta;
  sc[S_RZ * 64 + r] = rz;
  if (hist && j < cap) {
    hist[(int64_t)(2 * j) * 64 + r] = sc[S_ALPHA * 64 + r];
    hist[(int64_t)(2 * j + 1) * 64 + r] = beta;
  }
  if (rr <= sc[S_TOL2 * 64 + r] * sc[S_BN * 64 + r] || !(rz > 0.0)) active[r] = 0;
}

__global__ __launch_bounds__(256) void pcg_p_kernel(int64_t units, int logk, const double* __restrict__ sc,
                                                    const int* __restrict__ active, const float* __restrict__ z,
                                                    float* __restrict__ p) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(u & ((1 << logk) - 1));
    if (active[r]) p[u] = z[u] + (float)sc[S_BETA * 64 + r] * p[u];
  }
}

__global__ __launch_bounds__(256) void fill_kernel(int64_t units, float v, float* __restrict__ x) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x)
    x[u] = v;
}
__global__ __launch_bounds__(256) void copy_kernel(int64_t units, const float* __restrict__ a, float* __restrict__ b) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x)
    b[u] = a[u];
}

// ------------------------------------------------------------ refinement
// b32 = r64 / s (s = ||r64||, its square at *s2 on the device), for the FP32 inner solve.
__global__ __launch_bounds__(256) void scale_down_kernel(int64_t n, const double* __restrict__ s2,
                                                         const double* __restrict__ r64, float* __restrict__ b32) {
  const double s = sqrt(*s2);
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b32[i] = (float)(r64[i] * inv);
}
// x64 += s * d32 (s^2 at *s2).
__global__ __launch_bounds__(256) void correct_kernel(int64_t n, const double* __restrict__ s2,
                                                      const float* __restrict__ d32, double* __restrict__ x64) {
  const double s = sqrt(*s2);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x64[i] += s * (double)d32[i];
}
// r32 = r64 in FP32 (residual replacement: the FP64 residual becomes the recurrence's).
__global__ __launch_bounds__(256) void cast_down_kernel(int64_t n, const double* __restrict__ r64,
                                                        float* __restrict__ r32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r32[i] = (float)r64[i];
}
// x64 += a x32 (a = 1: a segment's FP32 iterate into the FP64 solution; -1: taken back).
__global__ __launch_bounds__(256) void axpy32_kernel(int64_t n, double a, const float* __restrict__ x32,
                                                     double* __restrict__ x64) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x64[i] += a * (double)x32[i];
}
// x64 += coef[r] x32 per right-hand side (block residual replacement).
__global__ __launch_bounds__(256) void axpy32k_kernel(int64_t units, int logk, const double* __restrict__ coef,
                                                      const float* __restrict__ x32, double* __restrict__ x64) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const double c = coef[(int)(u & ((1 << logk) - 1))];
    if (c != 0.0) x64[u] += c * (double)x32[u];
  }
}
// Y = L W (row-major n x k, FP32): the diagonal of the factor and its lower entries.
__global__ __launch_bounds__(256) void ic_lmul_kernel(int64_t units, int logk, const int64_t* __restrict__ lptr,
                                                      const int32_t* __restrict__ lcol, const float* __restrict__ l32,
                                                      const double* __restrict__ dg64, const float* __restrict__ w,
                                                      float* __restrict__ y) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = u >> logk;
    const int r = (int)(u & ((1 << logk) - 1));
    float s = (float)dg64[i] * w[u];
    for (int64_t p = lptr[i]; p < lptr[i + 1]; ++p) s += l32[p] * w[((int64_t)lcol[p] << logk) + r];
    y[u] = s;
  }
}
// p = z for the right-hand sides whose recurrence restarts (state 2).
__global__ __launch_bounds__(256) void copy_cols_kernel(int64_t units, int logk, const int* __restrict__ state,
                                                        const float* __restrict__ z, float* __restrict__ p) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x)
    if (state[(int)(u & ((1 << logk) - 1))] == 2) p[u] = z[u];
}
__global__ __launch_bounds__(kRedThreads) void sum_one_kernel(int nblk, const double* __restrict__ part,
                                                             double* __restrict__ out) {
  __shared__ double sums[1][64];
  reduce_parts<1>(part, 1, nblk, sums);
  if (threadIdx.x == 0) out[0] = sums[0][0];
}
// out[0] = sum over the k right-hand sides of the partials (in order).
__global__ __launch_bounds__(kRedThreads) void sum_rhs_kernel(int nblk, int k, const double* __restrict__ part,
                                                             double* __restrict__ out) {
  __shared__ double sums[1][64];
  reduce_parts<1>(part, k, nblk, sums);
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int r = 0; r < k; ++r) s += sums[0][r];
  out[0] = s;
}

// FP64 residual R = B - (Q + diag(dadd)) X for k right-hand sides (row-major
// n x k) with the block partials of ||R_r||^2 per right-hand side.
__global__ __launch_bounds__(256) void resid64k_kernel(int64_t n, int logk, const int64_t* __restrict__ indptr,
                                                       const int32_t* __restrict__ indices,
                                                       const double* __restrict__ qval,
                                                       const double* __restrict__ dadd, const double* __restrict__ B,
                                                       const double* __restrict__ X, double* __restrict__ R,
                                                       double* __restrict__ part) {
  const int k = 1 << logk;
  const int64_t units = n << logk;
  double acc[1] = {0.0};
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = u >> logk;
    const int r = (int)(u & (k - 1));
    double s = 0.0;
    for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) s += qval[e] * X[((int64_t)indices[e] << logk) + r];
    if (dadd) s += dadd[i] * X[u];
    const double v = B[u] - s;
    R[u] = v;
    acc[0] += v * v;
  }
  block_reduce_store<1>(acc, k, part);
}
// out[r] = sum of the partials of right-hand side r (in block order).
__global__ __launch_bounds__(kRedThreads) void norms_kernel(int nblk, const double* __restrict__ part,
                                                           double* __restrict__ out) {
  double v[1];
  reduce_col<1>(part, blockIdx.x, nblk, v);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}
// b32 = r64 / ||r64|| per right-hand side; zero for those already within
// their threshold (||r||^2 <= thr), whose inner solve is then empty.
__global__ __launch_bounds__(256) void scale_down_k_kernel(int64_t units, int logk, const double* __restrict__ rn,
                                                           const double* __restrict__ thr,
                                                           const double* __restrict__ r64, float* __restrict__ b32) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(u & ((1 << logk) - 1));
    const double s2 = rn[r];
    b32[u] = (s2 > thr[r] && s2 > 0.0) ? (float)(r64[u] / sqrt(s2)) : 0.f;
  }
}
// x64 += ||r|| d32 per right-hand side.
__global__ __launch_bounds__(256) void correct_k_kernel(int64_t units, int logk, const double* __restrict__ rn,
                                                        const double* __restrict__ thr, const float* __restrict__ d32,
                                                        double* __restrict__ x64) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(u & ((1 << logk) - 1));
    const double s2 = rn[r];
    if (s2 > thr[r] && s2 > 0.0) x64[u] += sqrt(s2) * (double)d32[u];
  }
}
__global__ __launch_bounds__(256) void fill64_kernel(int64_t units, double* __restrict__ x) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x)
    x[u] = 0.0;
}

// ------------------------------------------------------------ Lanczos
// alpha_j = q^T w per right-hand side (partials), then w -= alpha q + beta_prev q_prev
// with the partials of ||w||^2; then q_prev = q, q = w / beta_j.
__global__ __launch_bounds__(kRedThreads) void lanczos_alpha_kernel(int nblk, const double* __restrict__ part,
                                                                   double* __restrict__ sc,
                                                                   double* __restrict__ alpha_out) {
  double v[1];
  reduce_col<1>(part, blockIdx.x, nblk, v);
  const int r = blockIdx.x;
  if (threadIdx.x) return;
  const double a = v[0];
  sc[S_ALPHA * 64 + r] = a;
  alpha_out[r] = a;
}
__global__ __launch_bounds__(256) void lanczos_orth_kernel(int64_t units, int logk, const double* __restrict__ sc,
                                                           const float* __restrict__ q, const float* __restrict__ qp,
                                                           float* __restrict__ w, double* __restrict__ part) {
  double acc[1] = {0.0};
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(u & ((1 << logk) - 1));
    const float v = w[u] - (float)sc[S_ALPHA * 64 + r] * q[u] - (float)sc[S_BETA * 64 + r] * qp[u];
    w[u] = v;
    acc[0] += (double)v * (double)v;
  }
  block_reduce_store<1>(acc, 1 << logk, part);
}
__global__ __launch_bounds__(kRedThreads) void lanczos_beta_kernel(int nblk, const double* __restrict__ part,
                                                                  double* __restrict__ sc,
                                                                  double* __restrict__ beta_out) {
  double v[1];
  reduce_col<1>(part, blockIdx.x, nblk, v);
  const int r = blockIdx.x;
  if (threadIdx.x) return;
  const double b = sqrt(v[0]);
  sc[S_BETA * 64 + r] = b;
  beta_out[r] = b;
}
// q_prev = q; q = w / beta (beta 0: q = 0).
__global__ __launch_bounds__(256) void lanczos_next_kernel(int64_t units, int logk, const double* __restrict__ sc,
                                                           float* __restrict__ q, float* __restrict__ qp,
                                                           const float* __restrict__ w) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const double b = sc[S_BETA * 64 + (int)(u & ((1 << logk) - 1))];
    qp[u] = q[u];
    q[u] = b > 0.0 ? (float)(w[u] / b) : 0.f;
  }
}

// ------------------------------------------------------------ probes
__device__ __host__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// V[i][r] = +-scale from the top bit of splitmix64(seed * 2^32 + i * k + r)... see
// spde_rademacher for the exact counter.
__global__ __launch_bounds__(256) void rademacher_kernel(int64_t units, uint64_t seed, float scale,
                                                         float* __restrict__ v) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h = splitmix64(seed ^ splitmix64((uint64_t)u));
    v[u] = (h >> 63) ? -scale : scale;
  }
}

// z[u] ~ N(0, 1) by Box-Muller from two 53-bit uniforms of the counter:
// h1 = splitmix64(seed ^ splitmix64(2u)), h2 = splitmix64(seed ^ splitmix64(2u + 1)),
// U1 = ((h1 >> 11) + 0.5) 2^-53 in (0, 1), U2 = (h2 >> 11) 2^-53,
// z = sqrt(-2 log U1) cos(2 pi U2).
__global__ __launch_bounds__(256) void gaussian_kernel(int64_t units, uint64_t seed, double* __restrict__ z) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h1 = splitmix64(seed ^ splitmix64(2 * (uint64_t)u));
    const uint64_t h2 = splitmix64(seed ^ splitmix64(2 * (uint64_t)u + 1));
    const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    z[u] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
}

// sigma[e] = (1/k) sum_r (U[i][r] V[j][r] + U[j][r] V[i][r]) / 2 for every
// structural entry e = (i, j); one wave per row, lanes = right-hand sides.
__global__ __launch_bounds__(256) void hutch_pattern_kernel(int64_t n, int logk, const int64_t* __restrict__ indptr,
                                                            const int32_t* __restrict__ indices,
                                                            const float* __restrict__ U, const float* __restrict__ V,
                                                            double alpha, double beta, double* __restrict__ sigma) {
  const int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int k = 1 << logk;
  const bool on = lane < k;
  const double ui = on ? (double)U[(i << logk) + lane] : 0.0, vi = on ? (double)V[(i << logk) + lane] : 0.0;
  for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) {
    const int64_t j = indices[e];
    double s = on ? 0.5 * (ui * (double)V[(j << logk) + lane] + (double)U[(j << logk) + lane] * vi) : 0.0;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) sigma[e] = (beta != 0.0 ? beta * sigma[e] : 0.0) + alpha * (s / k);
  }
}

__global__ void logdet_diag_kernel(int64_t n, const double* __restrict__ dg64, double* __restrict__ part) {
  double acc[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc[0] += log(dg64[i]);
  block_reduce_store<1>(acc, 1, part);
}


// ------------------------------------------------------------ learned correction: inputs and loss
// Seven Jacobi-scaled stencil coefficients of Q_c per lattice entry (one time
// slice of the network's input): [tau count / A_ii, A~ to E, W, N, S in the
// slice, A~ to the same cell in the next and the previous slice, 0], with
// A~_ij = A_ij / sqrt(A_ii A_jj) (0 where the neighbour is outside).
__global__ __launch_bounds__(256) void features_kernel(int64_t n, const int64_t* __restrict__ indptr,
                                                       const int32_t* __restrict__ indices,
                                                       const uint8_t* __restrict__ code,
                                                       const double* __restrict__ qval,
                                                       const double* __restrict__ dadd,
                                                       const double* __restrict__ adiag, float* __restrict__ feat) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double ai = adiag[i];
  f[0] = (float)((dadd ? dadd[i] : 0.0) / ai);
  for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) {
    int slot = -1;
    switch (code[e]) {
      case 13: slot = 1; break;  // (dy, dx) = (0, +1): E
      case 11: slot = 2; break;  // (0, -1): W
      case 17: slot = 3; break;  // (+1, 0): N
      case 7: slot = 4; break;   // (-1, 0): S
      case 38: slot = 5; break;  // next slice, same cell
      case 29: slot = 6; break;  // previous slice, same cell
      default: break;
    }
    if (slot > 0) f[slot] = (float)(qval[e] / sqrt(ai * adiag[indices[e]]));
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) feat[i * 8 + c] = f[c];
}

// w = L'^T x for K samples (row-major n x K doubles), FP64 factor, and the
// block partials of sum ||w||^2.
__global__ __launch_bounds__(256) void kl_w_kernel(int64_t n, int logk, const int64_t* __restrict__ uptr,
                                                   const int32_t* __restrict__ ucol, const int64_t* __restrict__ ulow,
                                                   const double* __restrict__ l64, const double* __restrict__ dg64,
                                                   const double* __restrict__ x, double* __restrict__ w,
                                                   double* __restrict__ part) {
  double acc[1] = {0.0};
  const int64_t units = n << logk;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = u >> logk;
    const int s = (int)(u & ((1 << logk) - 1));
    double v = dg64[j] * x[u];
    for (int64_t q = uptr[j]; q < uptr[j + 1]; ++q) v += l64[ulow[q]] * x[((int64_t)ucol[q] << logk) + s];
    w[u] = v;
    acc[0] += v * v;
  }
  block_reduce_store<1>(acc, 1, part);
}

// Gradient of KL(N(0, A^-1) || N(0, (L' L'^T)^-1)) = 1/2 E||L'^T x||^2 - sum log L'_ii + const
// with respect to the correction channels (scaled by `scale`): for the lower
// entry (i, j) of channel c, d/dcorr = sqrt(A_ii) E[x_i w_j]; for the diagonal
// channel, L'_ii E[x_i w_i] - 1.  One wave per row, lanes = samples.
__global__ __launch_bounds__(256) void kl_grad_kernel(int64_t n, int logk, const int64_t* __restrict__ lptr,
                                                      const int32_t* __restrict__ lcol,
                                                      const int64_t* __restrict__ lch,
                                                      const double* __restrict__ dg64,
                                                      const double* __restrict__ adiag,
                                                      const double* __restrict__ x, const double* __restrict__ w,
                                                      double scale, float* __restrict__ dout) {
  const int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int k = 1 << logk;
  const bool on = lane < k;
  const double xi = on ? x[(i << logk) + lane] : 0.0;
  const double inv = 1.0 / k;
  const double si = sqrt(adiag[i]);
  for (int64_t q = lptr[i]; q < lptr[i + 1]; ++q) {
    double g = on ? xi * w[((int64_t)lcol[q] << logk) + lane] : 0.0;
    for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o, 64);
    if (lane == 0 && lch[q] >= 0) dout[lch[q]] = (float)(scale * si * g * inv);
  }
  double g = on ? xi * w[(i << logk) + lane] : 0.0;
  for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o, 64);
  if (lane == 0) dout[i * kPairs] = (float)(scale * (dg64[i] * g * inv - 1.0));
}


// Gradient of the symmetrised divergence J = 1/2 E||L'^T x||^2 + 1/2 E[y^T A y] - n
// (y = L'^-T z, z Rademacher; E over samples) = KL(p||q) + KL(q||p) of
// p = N(0, A^-1), q = N(0, (L' L'^T)^-1): for the pair (i, j) of a lower
// entry, dJ/dL'_ij = E[x_i w_j] - E[y_i u_j] with w = L'^T x, u = L'^-1 A y;
// the correction channels follow as in kl_grad_kernel (sqrt(A_ii) for pairs,
// L'_ii for the diagonal's log-scale).  One wave per row, lanes = samples.
__global__ __launch_bounds__(256) void jeffreys_grad_kernel(int64_t n, int logk, const int64_t* __restrict__ lptr,
                                                            const int32_t* __restrict__ lcol,
                                                            const int64_t* __restrict__ lch,
                                                            const double* __restrict__ dg64,
                                                            const double* __restrict__ adiag,
                                                            const double* __restrict__ x, const double* __restrict__ w,
                                                            const float* __restrict__ y, const float* __restrict__ u,
                                                            double scale, float* __restrict__ dout) {
  const int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int k = 1 << logk;
  const bool on = lane < k;
  const double xi = on ? x[(i << logk) + lane] : 0.0;
  const double yi = on ? (double)y[(i << logk) + lane] : 0.0;
  const double inv = 1.0 / k;
  const double si = sqrt(adiag[i]);
  for (int64_t q = lptr[i]; q < lptr[i + 1]; ++q) {
    const int64_t j = lcol[q];
    double g = on ? xi * w[(j << logk) + lane] - yi * (double)u[(j << logk) + lane] : 0.0;
    for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o, 64);
    if (lane == 0 && lch[q] >= 0) dout[lch[q]] = (float)(scale * si * g * inv);
  }
  double g = on ? xi * w[(i << logk) + lane] - yi * (double)u[(i << logk) + lane] : 0.0;
  for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o, 64);
  if (lane == 0) dout[i * kPairs] = (float)(scale * dg64[i] * g * inv);
}

inline unsigned grid_for(int64_t units) {
  const int64_t b = (units + 255) / 256;
  return (unsigned)(b < kRedBlocks ? (b < 1 ? 1 : b) : kRedBlocks);
}
inline unsigned blocks_of(int64_t units) { return (unsigned)((units + 255) / 256); }

template <typename T>
int upload_vec(T** dst, const std::vector<T>& src) {
  const size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
  SPDE_HIP(hipMalloc((void**)dst, bytes));
  if (!src.empty()) SPDE_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return SPDE_OK;
}

int log2_of(int k) {
  int l = 0;
  while ((1 << l) < k) ++l;
  return (1 << l) == k && k <= 64 ? l : -1;
}

// The last colour with rows: its rows have no neighbour of a later colour.
int last_colour(const spde_ic* ic) {
  int c = kColours - 1;
  while (c > 0 && ic->coff[c + 1] <= ic->coff[c]) --c;
  return c;
}
// Forward sweep z = L^-1 r, colour by colour.  fold: the rows of the last
// colour also take their backward step (x_i = z_i / L_ii: no later colour
// couples to them), which the backward sweep of ic_apply then skips.
// lds: bytes of LDS each block of the 64-column sweeps reserves (0: none), which
// caps how many of their blocks a CU holds (kShareLds).
int ic_fwd(spde_ic* ic, const float* r, float* z, int logk, hipStream_t s, bool fold = false, unsigned lds = 0) {
  const int last = last_colour(ic);
  for (int c = 0; c < kColours; ++c) {
    const int64_t r0 = ic->coff[c], r1 = ic->coff[c + 1];
    if (r1 <= r0) continue;
    const int f = fold && c == last ? 1 : 0;
    if (logk == 0)
      hipLaunchKernelGGL(fwd1_kernel, dim3(blocks_of((r1 - r0) * kRowLanes)), dim3(256), 0, s, r0, r1, ic->d_crow,
                         ic->d_lptr, ic->d_lcol, ic->d_l32, ic->d_dinv, r, z, f);
    else if (logk >= 4) {
      auto kern = logk == 6 ? fwdw_kernel<64> : logk == 5 ? fwdw_kernel<32> : fwdw_kernel<16>;
      hipLaunchKernelGGL(kern, dim3(blocks_of((r1 - r0) << 6)), dim3(256), logk == 6 ? lds : 0, s, r0, r1, ic->d_crow,
                         ic->d_lptr,
                         ic->d_lcol, ic->d_l32, ic->d_dinv, r, z, f);
    }
    else
      hipLaunchKernelGGL(fwd_kernel, dim3(blocks_of((r1 - r0) << logk)), dim3(256), 0, s, r0, r1, logk, ic->d_crow,
                         ic->d_lptr, ic->d_lcol, ic->d_l32, ic->d_dinv, r, z, f);
  }
  return SPDE_OK;
}
// Backward sweep x = L^-T y (x may be y), colour by colour from the last;
// skip_last: the last colour was folded into the forward sweep.
int ic_bwd(spde_ic* ic, const float* y, float* x, int logk, hipStream_t s, bool skip_last = false,
           unsigned lds = 0) {
  const int last = last_colour(ic);
  for (int c = kColours - 1; c >= 0; --c) {
    const int64_t r0 = ic->coff[c], r1 = ic->coff[c + 1];
    if (r1 <= r0 || (skip_last && c == last)) continue;
    if (logk == 0)
      hipLaunchKernelGGL(bwd1_kernel, dim3(blocks_of((r1 - r0) * kRowLanes)), dim3(256), 0, s, r0, r1, ic->d_crow,
                         ic->d_uptr, ic->d_ucol, ic->d_u32, ic->d_dinv, y, x);
    else if (logk >= 4) {
      auto kern = logk == 6 ? bwdw_kernel<64> : logk == 5 ? bwdw_kernel<32> : bwdw_kernel<16>;
      hipLaunchKernelGGL(kern, dim3(blocks_of((r1 - r0) << 6)), dim3(256), logk == 6 ? lds : 0, s, r0, r1, ic->d_crow,
                         ic->d_uptr,
                         ic->d_ucol, ic->d_u32, ic->d_dinv, y, x);
    }
    else
      hipLaunchKernelGGL(bwd_kernel, dim3(blocks_of((r1 - r0) << logk)), dim3(256), 0, s, r0, r1, logk, ic->d_crow,
                         ic->d_uptr, ic->d_ucol, ic->d_u32, ic->d_dinv, y, x);
  }
  return SPDE_OK;
}
// z = M^-1 r = (L L^T)^-1 r: forward from r into z, then backward in place,
// the last colour's two steps in one launch.
int ic_apply(spde_ic* ic, const float* r, float* z, int logk, hipStream_t s, unsigned lds = 0) {
  ic_fwd(ic, r, z, logk, s, true, lds);
  return ic_bwd(ic, z, z, logk, s, true, lds);
}
// y = Q_c x in FP32 (k = 2^logk right-hand sides) with the block partials of
// x^T y when part is set; `g` blocks (the caller's reduction grid).
void launch_spmm(spde_ic* ic, int logk, unsigned g, const float* x, float* y, double* part, hipStream_t s,
                 unsigned lds = 0) {
  if (logk == 0)
    hipLaunchKernelGGL(spmm1_kernel, dim3(g), dim3(256), 0, s, ic->n, ic->pl->d_indptr, ic->pl->d_indices,
                       ic->d_qc32, x, y, part);
  else if (logk >= 4)
    hipLaunchKernelGGL(logk == 6 ? spmmw_kernel<64> : logk == 5 ? spmmw_kernel<32> : spmmw_kernel<16>, dim3(g),
                       dim3(256), logk == 6 ? lds : 0, s, ic->n, ic->pl->d_indptr, ic->pl->d_indices, ic->d_qc32, x,
                       y, part);
  else
    hipLaunchKernelGGL(spmm32_kernel, dim3(g), dim3(256), 0, s, ic->n, logk, ic->pl->d_indptr, ic->pl->d_indices,
                       ic->d_qc32, x, y, part);
}

// Scalars and partials of one handle (the n-sized buffers come with its first solve).
int alloc_small_work(spde_ic* ic) {
  if (hipMalloc((void**)&ic->d_status, sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&ic->d_part, sizeof(double) * kRedBlocks * 64 * 2) != hipSuccess ||
      hipMalloc((void**)&ic->d_scal, sizeof(double) * 64 * 8) != hipSuccess ||
      hipMalloc((void**)&ic->d_active, sizeof(int) * 64) != hipSuccess ||
      hipMalloc((void**)&ic->d_iters, sizeof(int) * 64) != hipSuccess ||
      hipMalloc((void**)&ic->d_state, sizeof(int) * 64) != hipSuccess) {
    set_error("spde_ic: device allocation failed");
    return SPDE_ERR_ALLOC;
  }
  return SPDE_OK;
}
void free_work(spde_ic* ic) {
  for (void* p : {(void*)ic->d_status, (void*)ic->d_part, (void*)ic->d_scal, (void*)ic->d_active,
                  (void*)ic->d_iters, (void*)ic->d_state, (void*)ic->d_work, (void*)ic->d_r64, (void*)ic->d_b32, (void*)ic->d_x32,
                  (void*)ic->d_rn, (void*)ic->d_lanczos, (void*)ic->d_klw, (void*)ic->d_hist})
    if (p) hipFree(p);
  for (hipGraphExec_t g : ic->it_graph)
    if (g) hipGraphExecDestroy(g);
  static_cast<IcWork&>(*ic) = IcWork();
}
void free_shared(spde_ic* ic) {
  for (void* p : {(void*)ic->d_crow, (void*)ic->d_lptr, (void*)ic->d_uptr, (void*)ic->d_lsrc, (void*)ic->d_ulow,
                  (void*)ic->d_lch, (void*)ic->d_lcol, (void*)ic->d_lkey, (void*)ic->d_ucol, (void*)ic->d_l64,
                  (void*)ic->d_dg64,
                  (void*)ic->d_adiag, (void*)ic->d_l32, (void*)ic->d_u32, (void*)ic->d_dinv, (void*)ic->d_qc32})
    if (p) hipFree(p);
}
// A view takes its parent's pattern and factor at the start of every call.
inline void sync_view(spde_ic* ic) {
  if (ic && ic->parent) static_cast<IcShared&>(*ic) = static_cast<const IcShared&>(*ic->parent);
}

int ensure_work(spde_ic* ic) {
  // each buffer once (a failed allocation leaves the others for the next call)
  if (!ic->d_work) SPDE_HIP(hipMalloc((void**)&ic->d_work, sizeof(float) * ic->n * 64 * 4));
  if (!ic->d_r64) SPDE_HIP(hipMalloc((void**)&ic->d_r64, sizeof(double) * ic->n * 64));
  if (!ic->d_b32) SPDE_HIP(hipMalloc((void**)&ic->d_b32, sizeof(float) * ic->n * 64));
  if (!ic->d_x32) SPDE_HIP(hipMalloc((void**)&ic->d_x32, sizeof(float) * ic->n * 64));
  if (!ic->d_rn) SPDE_HIP(hipMalloc((void**)&ic->d_rn, sizeof(double) * 64 * 2));
  return SPDE_OK;
}

constexpr int kPoll = 16;     // iterations between the host's looks at the active flags
// LDS each block of the 64-column launches (SpMM, sweeps, vector steps)
// reserves in spde_pcg_solve_coeffs, whose probe solves run beside another
// stream's solve (the conditional mean): 32 KB of the CU's 160 KB holds five
// instead of eight of their blocks per CU and leaves wave slots to the other
// stream's short launches (iterative log-likelihood at 100x100x20: 2.95 ->
// 3.35 evals/s; 24 KB 3.11, 40 KB 3.07, 48 KB 2.83).
constexpr unsigned kShareLds = 32 * 1024;
constexpr int kSegMax = 512;  // residual replacement: at most this many iterations between replacements

// One PCG iteration on the handle's work vectors (R, Z, P, Qv) and X.
void enqueue_iteration(spde_ic* ic, int logk, unsigned g, float* X, hipStream_t s, double* hist = nullptr,
                       int cap = 0, unsigned lds = 0) {
  const int k = 1 << logk;
  const int64_t n = ic->n, units = n << logk;
  float* R = ic->d_work;
  float* Z = R + n * 64;
  float* P = Z + n * 64;
  float* Qv = P + n * 64;
  launch_spmm(ic, logk, g, P, Qv, ic->d_part, s, lds);
  hipLaunchKernelGGL(pcg_alpha_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal, ic->d_active);
  hipLaunchKernelGGL(pcg_update_kernel, dim3(g), dim3(256), lds, s, units, logk, ic->d_scal, P, Qv, X, R);
  ic_apply(ic, R, Z, logk, s, lds);
  hipLaunchKernelGGL(dot2_kernel, dim3(g), dim3(256), lds, s, units, logk, R, Z, ic->d_part);
  hipLaunchKernelGGL(pcg_beta_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal, ic->d_active,
                     ic->d_iters, hist, cap);
  hipLaunchKernelGGL(pcg_p_kernel, dim3(g), dim3(256), lds, s, units, logk, ic->d_scal, ic->d_active, Z, P);
}

// kPoll iterations with X = the handle's d_x32 as one HIP graph, captured the
// first time and replayed afterwards (the buffers it names never move): one
// launch instead of about forty per iteration, which is what bounds a solve
// with one right-hand side.  Null when the stream cannot be captured (the
// legacy default stream): the caller then launches kernel by kernel.
hipGraphExec_t iteration_graph(spde_ic* ic, int logk, unsigned g, hipStream_t s) {
  if (ic->it_graph[logk]) return ic->it_graph[logk];
  if (s == nullptr) return nullptr;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  for (int j = 0; j < kPoll; ++j) enqueue_iteration(ic, logk, g, ic->d_x32, s);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s, &graph);
  hipGraphExec_t exec = nullptr;
  if (e == hipSuccess && graph) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) hipGraphDestroy(graph);
  if (e != hipSuccess || !exec) {
    (void)hipGetLastError();
    return nullptr;
  }
  ic->it_graph[logk] = exec;
  return exec;
}

// FP32 PCG with k = 2^logk right-hand sides; X (initial guess 0) and B in
// row-major n x k.  Host polls the active flags every kPoll iterations.
// graph: iterate through iteration_graph (X must be the handle's d_x32).
int pcg_run(spde_ic* ic, const float* B, float* X, int logk, double tol, int maxit, int32_t* iters, hipStream_t s,
            bool graph = false, double* hist = nullptr, int cap = 0, unsigned lds = 0) {
  const int k = 1 << logk;
  const int64_t n = ic->n, units = n << logk;
  int rc = ensure_work(ic);
  if (rc) return rc;
  float* R = ic->d_work;
  float* Z = R + n * 64;
  float* P = Z + n * 64;
  const unsigned g = grid_for(units);
  hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, units, 0.f, X);
  hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, s, units, B, R);
  ic_apply(ic, R, Z, logk, s);
  hipLaunchKernelGGL(dot2_kernel, dim3(g), dim3(256), 0, s, units, logk, R, Z, ic->d_part);
  hipLaunchKernelGGL(pcg_init_scalars_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part, tol * tol, ic->d_scal,
                     ic->d_active, ic->d_iters);
  hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, s, units, Z, P);
  SPDE_HIP(hipGetLastError());
  hipGraphExec_t ge = graph && !hist && X == ic->d_x32 ? iteration_graph(ic, logk, g, s) : nullptr;
  std::vector<int> act(k);
  for (int it = 0; it < maxit;) {
    if (ge && maxit - it >= kPoll) {
      SPDE_HIP(hipGraphLaunch(ge, s));
      it += kPoll;
    } else {
      enqueue_iteration(ic, logk, g, X, s, hist, cap, lds);
      ++it;
    }
    if (it % kPoll == 0 || it == maxit) {
      SPDE_HIP(hipMemcpyAsync(act.data(), ic->d_active, sizeof(int) * k, hipMemcpyDeviceToHost, s));
      SPDE_HIP(hipStreamSynchronize(s));
      bool any = false;
      for (int r = 0; r < k; ++r) any = any || act[r];
      if (!any) break;
    }
  }
  SPDE_HIP(hipGetLastError());
  if (iters) {
    SPDE_HIP(hipMemcpyAsync(iters, ic->d_iters, sizeof(int) * k, hipMemcpyDeviceToHost, s));
    SPDE_HIP(hipStreamSynchronize(s));
  }
  return SPDE_OK;
}

}  // namespace spde

using namespace spde;

extern "C" {

int spde_ic_create(spde_plan* pl, spde_ic** out) {
  if (!pl || !out || pl->host_only || pl->lat.wide) {
    set_error("spde_ic_create: invalid argument (a device space-time plan of the advection-diffusion pattern)");
    return SPDE_ERR_ARG;
  }
  const Lattice& L = pl->lat;
  if (L.bc == 2 && (L.M % 3 || L.N % 3 || (L.T > 1 && L.T % 2))) {
    set_error("spde_ic_create: periodic lattices need M, N divisible by 3 and T even for the 18-colour ordering");
    return SPDE_ERR_ARG;
  }
  *out = nullptr;
  spde_ic* ic = new spde_ic();
  ic->pl = pl;
  ic->M = L.M;
  ic->N = L.N;
  ic->T = L.T;
  const int64_t n = L.n(), Ns = L.Ns();
  ic->n = n;
  ic->nnz = (int64_t)pl->st.indices.size();
  const auto& ip = pl->st.indptr;
  const auto& ix = pl->st.indices;
  auto colour = [&](int64_t v) {
    const int64_t t = v / Ns, rem = v % Ns;
    const int j = (int)(rem / L.M), i = (int)(rem % L.M);
    return (i % 3) + 3 * (j % 3) + 9 * (int)(t & 1);
  };
  ic->h_colour.resize(n);
  for (int64_t v = 0; v < n; ++v) ic->h_colour[v] = colour(v);
  // rows grouped by colour
  std::vector<int32_t> crow;
  crow.reserve(n);
  ic->coff.assign(kColours + 1, 0);
  for (int c = 0; c < kColours; ++c) {
    ic->coff[c] = (int64_t)crow.size();
    for (int64_t v = 0; v < n; ++v)
      if (ic->h_colour[v] == c) crow.push_back((int32_t)v);
  }
  ic->coff[kColours] = (int64_t)crow.size();
  // lower lists sorted by (colour, column); channel of each pair
  ic->h_lptr.assign(n + 1, 0);
  ic->h_uptr.assign(n + 1, 0);
  for (int64_t v = 0; v < n; ++v) {
    for (int64_t e = ip[v]; e < ip[v + 1]; ++e) {
      const int64_t w = ix[e];
      if (w == v) continue;
      if (ic->h_colour[w] == ic->h_colour[v]) {
        set_error("spde_ic_create: colouring is not an independent-set partition of the pattern");
        delete ic;
        return SPDE_ERR_STATE;
      }
      if (ic->h_colour[w] < ic->h_colour[v]) ++ic->h_lptr[v + 1];
      else ++ic->h_uptr[v + 1];
    }
  }
  for (int64_t v = 0; v < n; ++v) {
    ic->max_low = std::max(ic->max_low, ic->h_lptr[v + 1]);
    ic->h_lptr[v + 1] += ic->h_lptr[v];
    ic->h_uptr[v + 1] += ic->h_uptr[v];
  }
  ic->nlow = ic->h_lptr[n];
  if (ic->h_uptr[n] != ic->nlow) {
    set_error("spde_ic_create: pattern is not symmetric");
    delete ic;
    return SPDE_ERR_STATE;
  }
  ic->h_lcol.resize(ic->nlow);
  ic->h_lsrc.resize(ic->nlow);
  ic->h_lch.assign(ic->nlow, -1);
  ic->h_ucol.resize(ic->nlow);
  ic->h_ulow.resize(ic->nlow);
  // forward-half pair channels: same slice (dy > 0, or dy == 0 and dx > 0),
  // then the next slice's 3x3 window; channel 0 is the diagonal
  auto pair_channel = [&](int dx, int dy, int dt) -> int {
    if (dt == 0) {
      if (!(dy > 0 || (dy == 0 && dx > 0))) return -1;
      const int lin = (dy + 2) * 5 + (dx + 2);  // 13..24 for the forward half
      return 1 + (lin - 13);                    // 1..12
    }
    if (dt == 1) return 13 + (dy + 1) * 3 + (dx + 1);  // 13..21
    return -1;
  };
  for (int64_t v = 0; v < n; ++v) {
    std::vector<std::pair<int, int64_t>> lo;  // (colour, structural position)
    for (int64_t e = ip[v]; e < ip[v + 1]; ++e) {
      const int64_t w = ix[e];
      if (w != v && ic->h_colour[w] < ic->h_colour[v]) lo.push_back({ic->h_colour[w] * (int)n + 0, e});
    }
    std::sort(lo.begin(), lo.end(), [&](const std::pair<int, int64_t>& a, const std::pair<int, int64_t>& b) {
      const int ca = ic->h_colour[ix[a.second]], cb = ic->h_colour[ix[b.second]];
      return ca != cb ? ca < cb : ix[a.second] < ix[b.second];
    });
    int64_t q = ic->h_lptr[v];
    for (auto& pr : lo) {
      const int64_t e = pr.second, w = ix[e];
      ic->h_lcol[q] = (int32_t)w;
      ic->h_lsrc[q] = e;
      // lattice offset of the pair (v, w), wrapped for periodic lattices
      const int64_t tv = v / Ns, tw = w / Ns;
      int dx = (int)(w % Ns % L.M) - (int)(v % Ns % L.M);
      int dy = (int)(w % Ns / L.M) - (int)(v % Ns / L.M);
      int dt = (int)(tw - tv);
      if (L.bc == 2) {
        if (dx > 2) dx -= L.M;
        if (dx < -2) dx += L.M;
        if (dy > 2) dy -= L.N;
        if (dy < -2) dy += L.N;
      }
      int ch = pair_channel(dx, dy, dt);
      int64_t cell = v;
      if (ch < 0) {
        ch = pair_channel(-dx, -dy, -dt);
        cell = w;
      }
      if (ch >= 1) ic->h_lch[q] = cell * kPairs + ch;
      ++q;
    }
  }
  // upper lists: entries (v, w) with colour(w) > colour(v); value L_wv in row w's lower list
  for (int64_t v = 0; v < n; ++v) {
    int64_t q = ic->h_uptr[v];
    for (int64_t e = ip[v]; e < ip[v + 1]; ++e) {
      const int64_t w = ix[e];
      if (w == v || ic->h_colour[w] < ic->h_colour[v]) continue;
      int64_t pos = -1;
      for (int64_t p = ic->h_lptr[w]; p < ic->h_lptr[w + 1]; ++p)
        if (ic->h_lcol[p] == (int32_t)v) pos = p;
      if (pos < 0) {
        set_error("spde_ic_create: mirror entry missing");
        delete ic;
        return SPDE_ERR_STATE;
      }
      ic->h_ucol[q] = (int32_t)w;
      ic->h_ulow[q] = pos;
      ++q;
    }
  }
  int rc = SPDE_OK;
  auto fail = [&](int code) {
    spde_ic_destroy(ic);
    return code;
  };
  if ((rc = upload_vec(&ic->d_crow, crow))) return fail(rc);
  if ((rc = upload_vec(&ic->d_lptr, ic->h_lptr))) return fail(rc);
  if ((rc = upload_vec(&ic->d_uptr, ic->h_uptr))) return fail(rc);
  if ((rc = upload_vec(&ic->d_lcol, ic->h_lcol))) return fail(rc);
  {
    std::vector<int64_t> lkey(ic->nlow);
    for (int64_t q = 0; q < ic->nlow; ++q) lkey[q] = (int64_t)ic->h_colour[ic->h_lcol[q]] * n + ic->h_lcol[q];
    if ((rc = upload_vec(&ic->d_lkey, lkey))) return fail(rc);
  }
  if ((rc = upload_vec(&ic->d_lsrc, ic->h_lsrc))) return fail(rc);
  if ((rc = upload_vec(&ic->d_lch, ic->h_lch))) return fail(rc);
  if ((rc = upload_vec(&ic->d_ucol, ic->h_ucol))) return fail(rc);
  if ((rc = upload_vec(&ic->d_ulow, ic->h_ulow))) return fail(rc);
  const int64_t nl = std::max<int64_t>(ic->nlow, 1);
  if (hipMalloc((void**)&ic->d_l64, sizeof(double) * nl) != hipSuccess ||
      hipMalloc((void**)&ic->d_dg64, sizeof(double) * n) != hipSuccess ||
      hipMalloc((void**)&ic->d_adiag, sizeof(double) * n) != hipSuccess ||
      hipMalloc((void**)&ic->d_l32, sizeof(float) * nl) != hipSuccess ||
      hipMalloc((void**)&ic->d_u32, sizeof(float) * nl) != hipSuccess ||
      hipMalloc((void**)&ic->d_dinv, sizeof(float) * n) != hipSuccess ||
      hipMalloc((void**)&ic->d_qc32, sizeof(float) * ic->nnz) != hipSuccess) {
    set_error("spde_ic_create: device allocation failed");
    return fail(SPDE_ERR_ALLOC);
  }
  if ((rc = alloc_small_work(ic))) return fail(rc);
  *out = ic;
  return SPDE_OK;
}

int spde_ic_view(spde_ic* parent, spde_ic** out) {
  if (!parent || !out || parent->parent || parent->released) {
    set_error("spde_ic_view: invalid argument (a solver made by spde_ic_create)");
    return SPDE_ERR_ARG;
  }
  *out = nullptr;
  spde_ic* v = new spde_ic();
  static_cast<IcShared&>(*v) = static_cast<const IcShared&>(*parent);
  v->parent = parent;
  const int rc = alloc_small_work(v);
  if (rc) {
    free_work(v);
    delete v;
    return rc;
  }
  parent->nviews++;
  *out = v;
  return SPDE_OK;
}

void spde_ic_destroy(spde_ic* ic) {
  if (!ic) return;
  hipDeviceSynchronize();
  free_work(ic);
  if (spde_ic* p = ic->parent) {
    delete ic;
    if (--p->nviews == 0 && p->released) {
      free_shared(p);
      delete p;
    }
    return;
  }
  if (ic->nviews > 0) {  // the views still read the factor
    ic->released = true;
    return;
  }
  free_shared(ic);
  delete ic;
}

int spde_ic_info(const spde_ic* ic, int64_t* nlow, int32_t* ncolours) {
  if (!ic) {
    set_error("spde_ic_info: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (nlow) *nlow = ic->nlow;
  if (ncolours) *ncolours = kColours;
  return SPDE_OK;
}

int spde_ic_export(const spde_ic* ic, int32_t* colour, int64_t* lptr, int32_t* lcol, int64_t* lch, int64_t* uptr,
                   int32_t* ucol) {
  if (!ic) {
    set_error("spde_ic_export: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (ic->parent) ic = ic->parent;
  if (colour) std::memcpy(colour, ic->h_colour.data(), sizeof(int32_t) * ic->n);
  if (lptr) std::memcpy(lptr, ic->h_lptr.data(), sizeof(int64_t) * (ic->n + 1));
  if (lcol) std::memcpy(lcol, ic->h_lcol.data(), sizeof(int32_t) * ic->nlow);
  if (lch) std::memcpy(lch, ic->h_lch.data(), sizeof(int64_t) * ic->nlow);
  if (uptr) std::memcpy(uptr, ic->h_uptr.data(), sizeof(int64_t) * (ic->n + 1));
  if (ucol) std::memcpy(ucol, ic->h_ucol.data(), sizeof(int32_t) * ic->nlow);
  return SPDE_OK;
}

int spde_ic_factor(spde_ic* ic, const double* qval, const double* diag_add, double shift, double* logdet_m,
                   void* stream) {
  if (ic && ic->parent) {
    set_error("spde_ic_factor: a view cannot change its parent's factor");
    return SPDE_ERR_ARG;
  }
  if (!ic || !qval || shift < 0.0) {
    set_error("spde_ic_factor: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  spde_plan* pl = ic->pl;
  const int64_t n = ic->n;
  ic->qval = qval;
  ic->diag_add = diag_add;
  ic->factored = false;
  hipLaunchKernelGGL(qc_values_kernel, dim3(blocks_of(n)), dim3(256), 0, s, n, pl->d_indptr, pl->d_diag, qval,
                     diag_add, ic->d_adiag, ic->d_qc32);
  SPDE_HIP(hipMemsetAsync(ic->d_status, 0, sizeof(int), s));
  for (int c = 0; c < kColours; ++c) {
    const int64_t r0 = ic->coff[c], r1 = ic->coff[c + 1];
    if (r1 <= r0) continue;
    if (ic->max_low <= kIcRowMax)
      hipLaunchKernelGGL(ic_factorw_kernel, dim3((unsigned)((r1 - r0 + 3) / 4)), dim3(256), 0, s, r0, r1, ic->d_crow,
                         ic->d_lptr, ic->d_lcol, ic->d_lkey, ic->d_lsrc, qval, ic->d_adiag, shift, ic->d_l64,
                         ic->d_dg64, ic->d_status);
    else
      hipLaunchKernelGGL(ic_factor_kernel, dim3(blocks_of(r1 - r0)), dim3(256), 0, s, r0, r1, ic->d_crow, ic->d_lptr,
                         ic->d_lcol, ic->d_lkey, ic->d_lsrc, qval, ic->d_adiag, shift, ic->d_l64, ic->d_dg64,
                         ic->d_status);
  }
  SPDE_HIP(hipGetLastError());
  int st = 0;
  SPDE_HIP(hipMemcpyAsync(&st, ic->d_status, sizeof(int), hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipStreamSynchronize(s));
  if (st) {
    set_error("spde_ic_factor: non-positive pivot at lattice index " + std::to_string(st - 1) +
              " (raise the diagonal shift)");
    return SPDE_ERR_NOT_SPD;
  }
  return spde_ic_correct(ic, nullptr, logdet_m, stream);
}

int spde_ic_correct(spde_ic* ic, const float* corr, double* logdet_m, void* stream) {
  if (ic && ic->parent) {
    set_error("spde_ic_correct: a view cannot change its parent's factor");
    return SPDE_ERR_ARG;
  }
  if (!ic || !ic->qval) {
    set_error("spde_ic_correct: no factorisation");
    return SPDE_ERR_STATE;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = ic->n;
  if (corr) {
    hipLaunchKernelGGL(ic_correct_lower_kernel, dim3(blocks_of(n)), dim3(256), 0, s, n, ic->d_lptr, ic->d_lch,
                       ic->d_adiag, corr, ic->d_l64);
    hipLaunchKernelGGL(ic_correct_diag_kernel, dim3(blocks_of(n)), dim3(256), 0, s, n, corr, ic->d_dg64);
  }
  hipLaunchKernelGGL(ic_cast_lower_kernel, dim3(blocks_of(ic->nlow)), dim3(256), 0, s, ic->nlow, ic->d_l64, ic->d_l32);
  hipLaunchKernelGGL(ic_cast_upper_kernel, dim3(blocks_of(ic->nlow)), dim3(256), 0, s, ic->nlow, ic->d_ulow,
                     ic->d_l64, ic->d_u32);
  hipLaunchKernelGGL(ic_cast_diag_kernel, dim3(blocks_of(n)), dim3(256), 0, s, n, ic->d_dg64, ic->d_dinv);
  const unsigned g = grid_for(n);
  hipLaunchKernelGGL(logdet_diag_kernel, dim3(g), dim3(256), 0, s, n, ic->d_dg64, ic->d_part);
  hipLaunchKernelGGL(sum_one_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal + S_X * 64);
  SPDE_HIP(hipGetLastError());
  double ld = 0.0;
  SPDE_HIP(hipMemcpyAsync(&ld, ic->d_scal + S_X * 64, sizeof(double), hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipStreamSynchronize(s));
  if (logdet_m) *logdet_m = 2.0 * ld;
  ic->factored = true;
  return SPDE_OK;
}

int spde_ic_values(spde_ic* ic, double* lower, double* diag, double* adiag, void* stream) {
  sync_view(ic);
  if (!ic || !ic->factored) {
    set_error("spde_ic_values: no factorisation");
    return SPDE_ERR_STATE;
  }
  hipStream_t s = (hipStream_t)stream;
  if (lower) SPDE_HIP(hipMemcpyAsync(lower, ic->d_l64, sizeof(double) * ic->nlow, hipMemcpyDeviceToDevice, s));
  if (diag) SPDE_HIP(hipMemcpyAsync(diag, ic->d_dg64, sizeof(double) * ic->n, hipMemcpyDeviceToDevice, s));
  if (adiag) SPDE_HIP(hipMemcpyAsync(adiag, ic->d_adiag, sizeof(double) * ic->n, hipMemcpyDeviceToDevice, s));
  return SPDE_OK;
}

int spde_ic_apply(spde_ic* ic, const float* r, float* z, int k, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !r || !z || logk < 0) {
    set_error("spde_ic_apply: invalid argument (k a power of two <= 64, after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  ic_apply(ic, r, z, logk, (hipStream_t)stream);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_ic_sweep(spde_ic* ic, int which, const float* in, float* out, int k, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !in || !out || logk < 0 || (which != 1 && which != 2)) {
    set_error("spde_ic_sweep: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (which == 1) ic_fwd(ic, in, out, logk, (hipStream_t)stream);
  else ic_bwd(ic, in, out, logk, (hipStream_t)stream);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_ic_spmm(spde_ic* ic, const float* x, float* y, int k, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !x || !y || logk < 0) {
    set_error("spde_ic_spmm: invalid argument");
    return SPDE_ERR_ARG;
  }
  launch_spmm(ic, logk, grid_for(ic->n << logk), x, y, nullptr, (hipStream_t)stream);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_pcg_solve(spde_ic* ic, const float* B, float* X, int k, double tol, int maxit, int32_t* iters,
                   void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !B || !X || logk < 0 || !(tol > 0.0) || maxit < 1) {
    set_error("spde_pcg_solve: invalid argument (k a power of two <= 64, after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  return pcg_run(ic, B, X, logk, tol, maxit, iters, (hipStream_t)stream);
}

int spde_pcg_solve_coeffs(spde_ic* ic, const float* B, float* X, int k, double tol, int maxit, int32_t* iters,
                          double* alpha, double* beta, int cap, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !B || !X || logk < 0 || !(tol > 0.0) || maxit < 1 || cap < 1 || !alpha || !beta ||
      !iters) {
    set_error("spde_pcg_solve_coeffs: invalid argument (k a power of two <= 64, cap >= 1, after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  if (ic->hist_cap < cap) {
    if (ic->d_hist) SPDE_HIP(hipFree(ic->d_hist));
    ic->d_hist = nullptr;
    ic->hist_cap = 0;
    SPDE_HIP(hipMalloc((void**)&ic->d_hist, sizeof(double) * 2 * 64 * cap));
    ic->hist_cap = cap;
  }
  int rc = pcg_run(ic, B, X, logk, tol, maxit, iters, s, false, ic->d_hist, cap, kShareLds);
  if (rc) return rc;
  std::vector<double> h((size_t)2 * 64 * cap);
  SPDE_HIP(hipMemcpyAsync(h.data(), ic->d_hist, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipStreamSynchronize(s));
  for (int r = 0; r < k; ++r)
    for (int j = 0; j < cap; ++j) {
      const bool ok = j < iters[r];
      alpha[(int64_t)r * cap + j] = ok ? h[(size_t)(2 * j) * 64 + r] : 0.0;
      beta[(int64_t)r * cap + j] = ok ? h[(size_t)(2 * j + 1) * 64 + r] : 0.0;
    }
  return SPDE_OK;
}

int spde_ic_lmul(spde_ic* ic, const float* W, float* Y, int k, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !W || !Y || logk < 0) {
    set_error("spde_ic_lmul: invalid argument (k a power of two <= 64, after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  const int64_t units = ic->n << logk;
  hipLaunchKernelGGL(ic_lmul_kernel, dim3(grid_for(units)), dim3(256), 0, (hipStream_t)stream, units, logk,
                     ic->d_lptr, ic->d_lcol, ic->d_l32, ic->d_dg64, W, Y);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_pcg_refine(spde_ic* ic, const double* b, double* x, double tol, double inner_tol, int max_rounds, int maxit,
                    int32_t* iters, int32_t* rounds, double* relres, void* stream) {
  sync_view(ic);
  if (!ic || !ic->factored || !b || !x || !(tol > 0.0) || !(inner_tol > 0.0) || max_rounds < 1 || maxit < 1) {
    set_error("spde_pcg_refine: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_work(ic);
  if (rc) return rc;
  spde_plan* pl = ic->pl;
  const int64_t n = ic->n;
  const unsigned g = grid_for(n);
  // ||b||^2 through the residual of x = 0
  SPDE_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
  hipLaunchKernelGGL(resid64_kernel, dim3(g), dim3(256), 0, s, n, pl->d_indptr, pl->d_indices, pl->d_diag, ic->qval,
                     ic->diag_add, b, x, ic->d_r64, ic->d_part);
  double bn = 0.0, rn = 0.0;
  hipLaunchKernelGGL(sum_one_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal + S_X * 64);
  SPDE_HIP(hipMemcpyAsync(&bn, ic->d_scal + S_X * 64, sizeof(double), hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipStreamSynchronize(s));
  rn = bn;
  int total = 0, r = 0;
  // ||r||^2 of the current round stays at d_scal[S_X * 64] (pcg_run does not
  // touch that slot) for the scaling before and the correction after the solve
  while (r < max_rounds && rn > tol * tol * bn && bn > 0.0) {
    hipLaunchKernelGGL(scale_down_kernel, dim3(g), dim3(256), 0, s, n, ic->d_scal + S_X * 64, ic->d_r64, ic->d_b32);
    int32_t it = 0;
    rc = pcg_run(ic, ic->d_b32, ic->d_x32, 0, inner_tol, maxit, &it, s, true);
    if (rc) return rc;
    total += it;
    hipLaunchKernelGGL(correct_kernel, dim3(g), dim3(256), 0, s, n, ic->d_scal + S_X * 64, ic->d_x32, x);
    hipLaunchKernelGGL(resid64_kernel, dim3(g), dim3(256), 0, s, n, pl->d_indptr, pl->d_indices, pl->d_diag,
                       ic->qval, ic->diag_add, b, x, ic->d_r64, ic->d_part);
    hipLaunchKernelGGL(sum_one_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal + S_X * 64);
    SPDE_HIP(hipMemcpyAsync(&rn, ic->d_scal + S_X * 64, sizeof(double), hipMemcpyDeviceToHost, s));
    SPDE_HIP(hipStreamSynchronize(s));
    ++r;
  }
  SPDE_HIP(hipGetLastError());
  if (iters) *iters = total;
  if (rounds) *rounds = r;
  if (relres) *relres = bn > 0.0 ? std::sqrt(rn / bn) : 0.0;
  return SPDE_OK;
}

int spde_pcg_replace(spde_ic* ic, const double* b, double* x, double tol, double delta, int maxit, int32_t* iters,
                     int32_t* updates, double* relres, void* stream) {
  sync_view(ic);
  if (!ic || !ic->factored || !b || !x || !(tol > 0.0) || !(delta > 0.0 && delta < 1.0) || maxit < 1) {
    set_error("spde_pcg_replace: invalid argument (tol > 0, 0 < delta < 1, maxit >= 1, after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_work(ic);
  if (rc) return rc;
  spde_plan* pl = ic->pl;
  const int64_t n = ic->n;
  const unsigned g = grid_for(n);
  float* R = ic->d_work;
  float* Z = R + n * 64;
  float* P = Z + n * 64;
  double* rn_dev = ic->d_scal + S_X * 64;
  auto residual = [&](double* out) -> int {  // r64 = b - Q_c x, *out = ||r64||^2
    hipLaunchKernelGGL(resid64_kernel, dim3(g), dim3(256), 0, s, n, pl->d_indptr, pl->d_indices, pl->d_diag,
                       ic->qval, ic->diag_add, b, x, ic->d_r64, ic->d_part);
    hipLaunchKernelGGL(sum_one_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g, ic->d_part, rn_dev);
    SPDE_HIP(hipMemcpyAsync(out, rn_dev, sizeof(double), hipMemcpyDeviceToHost, s));
    SPDE_HIP(hipStreamSynchronize(s));
    return SPDE_OK;
  };
  // r32 = r64, z = M^-1 r32 and the partials of r^T z, r^T r
  auto precondition = [&]() {
    hipLaunchKernelGGL(cast_down_kernel, dim3(g), dim3(256), 0, s, n, ic->d_r64, R);
    ic_apply(ic, R, Z, 0, s);
    hipLaunchKernelGGL(dot2_kernel, dim3(g), dim3(256), 0, s, n, 0, R, Z, ic->d_part);
  };
  // a fresh recurrence from the current residual (tolerance 0: the host ends the solve)
  auto restart = [&]() {
    precondition();
    hipLaunchKernelGGL(pcg_init_scalars_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g, ic->d_part, 0.0,
                       ic->d_scal, ic->d_active, ic->d_iters);
    hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, s, n, Z, P);
  };
  SPDE_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
  double bn = 0.0;
  if ((rc = residual(&bn))) return rc;
  double rn = bn;
  int it = 0, upd = 0, stall = 0;
  if (bn > 0.0) {
    hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, n, 0.f, ic->d_x32);
    restart();
    SPDE_HIP(hipGetLastError());
    hipGraphExec_t ge = iteration_graph(ic, 0, g, s);
    double seg = bn;   // ||r||^2 at the start of the segment
    int seg_it = 0;    // iterations of the segment
    while (it < maxit) {
      const int before = it;
      if (ge && maxit - it >= kPoll) {
        SPDE_HIP(hipGraphLaunch(ge, s));
        it += kPoll;
      } else {
        for (int j = 0; j < kPoll && it < maxit; ++j, ++it) enqueue_iteration(ic, 0, g, ic->d_x32, s);
      }
      seg_it += it - before;
      int act = 0;
      double rr = 0.0;
      SPDE_HIP(hipMemcpyAsync(&act, ic->d_active, sizeof(int), hipMemcpyDeviceToHost, s));
      SPDE_HIP(hipMemcpyAsync(&rr, ic->d_scal + S_RR * 64, sizeof(double), hipMemcpyDeviceToHost, s));
      SPDE_HIP(hipStreamSynchronize(s));
      if (act && rr > delta * delta * seg && seg_it < kSegMax && it < maxit) continue;
      // residual replacement: the segment's FP32 iterate into x, the true FP64 residual into the recurrence
      hipLaunchKernelGGL(axpy32_kernel, dim3(g), dim3(256), 0, s, n, 1.0, ic->d_x32, x);
      double rn_new = 0.0;
      if ((rc = residual(&rn_new))) return rc;
      ++upd;
      seg_it = 0;
      if (!(rn_new <= rn)) {
        // the segment made the true residual worse (the FP32 recurrence lost its
        // way): take it back and restart the recurrence from the previous residual
        hipLaunchKernelGGL(axpy32_kernel, dim3(g), dim3(256), 0, s, n, -1.0, ic->d_x32, x);
        hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, n, 0.f, ic->d_x32);
        if ((rc = residual(&rn))) return rc;
        if (++stall >= 3 || it >= maxit) break;
        restart();
        seg = rn;
        continue;
      }
      hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, n, 0.f, ic->d_x32);
      stall = rn_new > 0.81 * rn ? stall + 1 : 0;  // the true residual fell by less than 10%
      rn = rn_new;
      if (rn <= tol * tol * bn || it >= maxit || stall >= 3) break;
      if (act && stall == 0) {
        precondition();
        hipLaunchKernelGGL(pcg_reseed_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal,
                           ic->d_active);
      } else {
        restart();  // breakdown of the recurrence, or little progress: new search directions
      }
      seg = rn;
    }
  }
  SPDE_HIP(hipGetLastError());
  if (iters) *iters = it;
  if (updates) *updates = upd;
  if (relres) *relres = bn > 0.0 ? std::sqrt(rn / bn) : 0.0;
  return SPDE_OK;
}

int spde_pcg_replace_block(spde_ic* ic, const double* B, double* X, int k, double tol, double delta, int maxit,
                           int32_t* iters, int32_t* updates, double* relres, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !B || !X || logk < 0 || !(tol > 0.0) || !(delta > 0.0 && delta < 1.0) || maxit < 1) {
    set_error("spde_pcg_replace_block: invalid argument (k a power of two <= 64, tol > 0, 0 < delta < 1)");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_work(ic);
  if (rc) return rc;
  spde_plan* pl = ic->pl;
  const int64_t n = ic->n, units = n << logk;
  const unsigned g = grid_for(units);
  float* R = ic->d_work;
  float* Z = R + n * 64;
  float* P = Z + n * 64;
  double* rn_dev = ic->d_rn;
  double* coef_dev = ic->d_rn + 64;
  // ||B - Q_c X||^2 per column into out (host)
  auto residual = [&](double* out) -> int {
    hipLaunchKernelGGL(resid64k_kernel, dim3(g), dim3(256), 0, s, n, logk, pl->d_indptr, pl->d_indices, ic->qval,
                       ic->diag_add, B, X, ic->d_r64, ic->d_part);
    hipLaunchKernelGGL(norms_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part, rn_dev);
    SPDE_HIP(hipMemcpyAsync(out, rn_dev, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    SPDE_HIP(hipStreamSynchronize(s));
    return SPDE_OK;
  };
  auto axpy = [&](const double* coef) -> int {
    SPDE_HIP(hipMemcpyAsync(coef_dev, coef, sizeof(double) * k, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(axpy32k_kernel, dim3(g), dim3(256), 0, s, units, logk, coef_dev, ic->d_x32, X);
    SPDE_HIP(hipStreamSynchronize(s));  // coef lives on the host stack
    return SPDE_OK;
  };
  // r32 = r64, z = M^-1 r32, partials of r^T z and r^T r
  auto precondition = [&]() {
    hipLaunchKernelGGL(cast_down_kernel, dim3(g), dim3(256), 0, s, units, ic->d_r64, R);
    ic_apply(ic, R, Z, logk, s);
    hipLaunchKernelGGL(dot2_kernel, dim3(g), dim3(256), 0, s, units, logk, R, Z, ic->d_part);
  };
  hipLaunchKernelGGL(fill64_kernel, dim3(g), dim3(256), 0, s, units, X);
  double bn[64], cur[64], fresh[64], coef[64];
  int state[64], stall[64], act[64];
  bool done[64];
  if ((rc = residual(bn))) return rc;
  for (int r = 0; r < k; ++r) {
    cur[r] = bn[r];
    stall[r] = 0;
    done[r] = !(bn[r] > 0.0);
  }
  int it = 0, upd = 0;
  auto all_done = [&]() {
    for (int r = 0; r < k; ++r)
      if (!done[r]) return false;
    return true;
  };
  if (!all_done()) {
    hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, units, 0.f, ic->d_x32);
    precondition();
    hipLaunchKernelGGL(pcg_init_scalars_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part, 0.0,
                       ic->d_scal, ic->d_active, ic->d_iters);
    hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, s, units, Z, P);
    SPDE_HIP(hipGetLastError());
    hipGraphExec_t ge = iteration_graph(ic, logk, g, s);
    double seg[64], rr[64];
    for (int r = 0; r < k; ++r) seg[r] = bn[r];
    int seg_it = 0;
    while (it < maxit) {
      const int before = it;
      if (ge && maxit - it >= kPoll) {
        SPDE_HIP(hipGraphLaunch(ge, s));
        it += kPoll;
      } else {
        for (int j = 0; j < kPoll && it < maxit; ++j, ++it) enqueue_iteration(ic, logk, g, ic->d_x32, s);
      }
      seg_it += it - before;
      SPDE_HIP(hipMemcpyAsync(act, ic->d_active, sizeof(int) * k, hipMemcpyDeviceToHost, s));
      SPDE_HIP(hipMemcpyAsync(rr, ic->d_scal + S_RR * 64, sizeof(double) * k, hipMemcpyDeviceToHost, s));
      SPDE_HIP(hipStreamSynchronize(s));
      bool ready = seg_it >= kSegMax || it >= maxit;
      if (!ready) {
        ready = true;  // every live column's recursive residual has fallen by delta (or it broke down)
        for (int r = 0; r < k; ++r)
          if (!done[r] && act[r] && rr[r] > delta * delta * seg[r]) ready = false;
      }
      if (!ready) continue;
      // replacement for every column: the segment's FP32 iterate into X, the FP64 residual into the recurrence
      for (int r = 0; r < k; ++r) coef[r] = 1.0;
      if ((rc = axpy(coef))) return rc;
      if ((rc = residual(fresh))) return rc;
      ++upd;
      seg_it = 0;
      bool back = false;
      for (int r = 0; r < k; ++r) {
        coef[r] = 0.0;
        if (done[r]) {
          state[r] = 0;
        } else if (!(fresh[r] <= cur[r])) {
          coef[r] = -1.0;  // the segment made this column's residual worse: take it back, restart the column
          back = true;
          state[r] = 2;
          if (++stall[r] >= 3) done[r] = true, state[r] = 0;
        } else {
          stall[r] = fresh[r] > 0.81 * cur[r] ? stall[r] + 1 : 0;
          cur[r] = fresh[r];
          if (cur[r] <= tol * tol * bn[r] || stall[r] >= 3) {
            done[r] = true;
            state[r] = 0;
          } else {
            state[r] = act[r] && stall[r] == 0 ? 1 : 2;
          }
        }
      }
      if (back) {
        if ((rc = axpy(coef))) return rc;
        if ((rc = residual(fresh))) return rc;
        for (int r = 0; r < k; ++r)
          if (coef[r] != 0.0) cur[r] = fresh[r];
      }
      hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, units, 0.f, ic->d_x32);
      if (all_done() || it >= maxit) break;
      SPDE_HIP(hipMemcpyAsync(ic->d_state, state, sizeof(int) * k, hipMemcpyHostToDevice, s));
      precondition();
      hipLaunchKernelGGL(pcg_replace_scalars_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part,
                         ic->d_state, ic->d_scal, ic->d_active);
      hipLaunchKernelGGL(copy_cols_kernel, dim3(g), dim3(256), 0, s, units, logk, ic->d_state, Z, P);
      SPDE_HIP(hipStreamSynchronize(s));  // state lives on the host stack
      for (int r = 0; r < k; ++r) seg[r] = cur[r];
    }
    if (iters) {
      SPDE_HIP(hipMemcpyAsync(iters, ic->d_iters, sizeof(int) * k, hipMemcpyDeviceToHost, s));
      SPDE_HIP(hipStreamSynchronize(s));
    }
  } else if (iters) {
    for (int r = 0; r < k; ++r) iters[r] = 0;
  }
  SPDE_HIP(hipGetLastError());
  if (updates) *updates = upd;
  if (relres)
    for (int r = 0; r < k; ++r) relres[r] = bn[r] > 0.0 ? std::sqrt(cur[r] / bn[r]) : 0.0;
  return SPDE_OK;
}

int spde_pcg_refine_block(spde_ic* ic, const double* B, double* X, int k, double tol, double inner_tol,
                          int max_rounds, int maxit, int32_t* iters, int32_t* rounds, double* relres, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !B || !X || logk < 0 || !(tol > 0.0) || !(inner_tol > 0.0) || max_rounds < 1 ||
      maxit < 1) {
    set_error("spde_pcg_refine_block: invalid argument (k a power of two <= 64, after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_work(ic);
  if (rc) return rc;
  spde_plan* pl = ic->pl;
  const int64_t n = ic->n, units = n << logk;
  const unsigned g = grid_for(units);
  double* rn = ic->d_rn;
  double* thr = ic->d_rn + 64;
  hipLaunchKernelGGL(fill64_kernel, dim3(g), dim3(256), 0, s, units, X);
  hipLaunchKernelGGL(resid64k_kernel, dim3(g), dim3(256), 0, s, n, logk, pl->d_indptr, pl->d_indices, ic->qval,
                     ic->diag_add, B, X, ic->d_r64, ic->d_part);
  hipLaunchKernelGGL(norms_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part, rn);
  double bn[64], cur[64], th[64];
  SPDE_HIP(hipMemcpyAsync(bn, rn, sizeof(double) * k, hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipStreamSynchronize(s));
  for (int r = 0; r < k; ++r) {
    th[r] = tol * tol * bn[r];
    cur[r] = bn[r];
    if (iters) iters[r] = 0;
  }
  SPDE_HIP(hipMemcpyAsync(thr, th, sizeof(double) * k, hipMemcpyHostToDevice, s));
  int32_t its[64];
  int rd = 0;
  auto pending = [&]() {
    for (int r = 0; r < k; ++r)
      if (cur[r] > th[r] && bn[r] > 0.0) return true;
    return false;
  };
  while (rd < max_rounds && pending()) {
    hipLaunchKernelGGL(scale_down_k_kernel, dim3(g), dim3(256), 0, s, units, logk, rn, thr, ic->d_r64, ic->d_b32);
    rc = pcg_run(ic, ic->d_b32, ic->d_x32, logk, inner_tol, maxit, its, s, true);
    if (rc) return rc;
    if (iters)
      for (int r = 0; r < k; ++r) iters[r] += its[r];
    hipLaunchKernelGGL(correct_k_kernel, dim3(g), dim3(256), 0, s, units, logk, rn, thr, ic->d_x32, X);
    hipLaunchKernelGGL(resid64k_kernel, dim3(g), dim3(256), 0, s, n, logk, pl->d_indptr, pl->d_indices, ic->qval,
                       ic->diag_add, B, X, ic->d_r64, ic->d_part);
    hipLaunchKernelGGL(norms_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part, rn);
    SPDE_HIP(hipMemcpyAsync(cur, rn, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    SPDE_HIP(hipStreamSynchronize(s));
    ++rd;
  }
  SPDE_HIP(hipGetLastError());
  if (rounds) *rounds = rd;
  if (relres)
    for (int r = 0; r < k; ++r) relres[r] = bn[r] > 0.0 ? std::sqrt(cur[r] / bn[r]) : 0.0;
  return SPDE_OK;
}

int spde_gaussian(int64_t n, int k, uint64_t seed, double* Z, void* stream) {
  if (n < 1 || k < 1 || !Z) {
    set_error("spde_gaussian: invalid argument");
    return SPDE_ERR_ARG;
  }
  const int64_t units = n * k;
  hipLaunchKernelGGL(gaussian_kernel, dim3(grid_for(units)), dim3(256), 0, (hipStream_t)stream, units, seed, Z);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_rademacher(int64_t n, int k, uint64_t seed, float scale, float* V, void* stream) {
  if (n < 1 || k < 1 || !V) {
    set_error("spde_rademacher: invalid argument");
    return SPDE_ERR_ARG;
  }
  const int64_t units = n * k;
  hipLaunchKernelGGL(rademacher_kernel, dim3(grid_for(units)), dim3(256), 0, (hipStream_t)stream, units, seed, scale,
                     V);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_slq(spde_ic* ic, int k, int steps, uint64_t seed, double* alpha, double* beta, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || logk < 0 || steps < 1 || !alpha || !beta) {
    set_error("spde_slq: invalid argument (k a power of two <= 64, after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_work(ic);
  if (rc) return rc;
  if (ic->lanczos_cap < (int64_t)steps * 64 * 2) {
    if (ic->d_lanczos) SPDE_HIP(hipFree(ic->d_lanczos));
    ic->d_lanczos = nullptr;
    SPDE_HIP(hipMalloc((void**)&ic->d_lanczos, sizeof(double) * steps * 64 * 2));
    ic->lanczos_cap = (int64_t)steps * 64 * 2;
  }
  spde_plan* pl = ic->pl;
  const int64_t n = ic->n, units = n << logk;
  float* Qp = ic->d_work;
  float* Qc = Qp + n * 64;
  float* W = Qc + n * 64;
  float* U = W + n * 64;
  const unsigned g = grid_for(units);
  // q_0 = v / sqrt(n) for Rademacher v; q_-1 = 0; beta_-1 = 0
  hipLaunchKernelGGL(rademacher_kernel, dim3(g), dim3(256), 0, s, units, seed, (float)(1.0 / std::sqrt((double)n)),
                     Qc);
  hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, units, 0.f, Qp);
  SPDE_HIP(hipMemsetAsync(ic->d_scal + S_BETA * 64, 0, sizeof(double) * 64, s));
  double* da = ic->d_lanczos;
  double* db = ic->d_lanczos + (int64_t)steps * 64;
  for (int j = 0; j < steps; ++j) {
    ic_bwd(ic, Qc, W, logk, s);                                                  // W = L^-T q
    launch_spmm(ic, logk, g, W, U, nullptr, s);                                  // U = Q_c W
    ic_fwd(ic, U, W, logk, s);                                                   // W = L^-1 U
    hipLaunchKernelGGL(dot2_kernel, dim3(g), dim3(256), 0, s, units, logk, Qc, W, ic->d_part);
    hipLaunchKernelGGL(lanczos_alpha_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal,
                       da + (int64_t)j * 64);
    hipLaunchKernelGGL(lanczos_orth_kernel, dim3(g), dim3(256), 0, s, units, logk, ic->d_scal, Qc, Qp, W, ic->d_part);
    hipLaunchKernelGGL(lanczos_beta_kernel, dim3(k), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal,
                       db + (int64_t)j * 64);
    hipLaunchKernelGGL(lanczos_next_kernel, dim3(g), dim3(256), 0, s, units, logk, ic->d_scal, Qc, Qp, W);
  }
  SPDE_HIP(hipGetLastError());
  std::vector<double> ha((size_t)steps * 64), hb((size_t)steps * 64);
  SPDE_HIP(hipMemcpyAsync(ha.data(), da, sizeof(double) * steps * 64, hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipMemcpyAsync(hb.data(), db, sizeof(double) * steps * 64, hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < steps; ++j)
    for (int r = 0; r < k; ++r) {
      alpha[(int64_t)r * steps + j] = ha[(size_t)j * 64 + r];
      beta[(int64_t)r * steps + j] = hb[(size_t)j * 64 + r];
    }
  return SPDE_OK;
}


int spde_ic_features(spde_ic* ic, float* feat, void* stream) {
  sync_view(ic);
  if (!ic || !ic->qval || !feat) {
    set_error("spde_ic_features: invalid argument (after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  spde_plan* pl = ic->pl;
  hipLaunchKernelGGL(features_kernel, dim3(blocks_of(ic->n)), dim3(256), 0, (hipStream_t)stream, ic->n, pl->d_indptr,
                     pl->d_indices, pl->d_code, ic->qval, ic->diag_add, ic->d_adiag, feat);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_ic_kl_grad(spde_ic* ic, const double* x, int k, double scale, float* dout, double* loss, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !x || !dout || logk < 0) {
    set_error("spde_ic_kl_grad: invalid argument (k a power of two <= 64, after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = ic->n;
  if (!ic->kl_cap || ic->kl_cap < (n << logk)) {
    if (ic->d_klw) SPDE_HIP(hipFree(ic->d_klw));
    ic->d_klw = nullptr;
    SPDE_HIP(hipMalloc((void**)&ic->d_klw, sizeof(double) * (n << logk)));
    ic->kl_cap = n << logk;
  }
  SPDE_HIP(hipMemsetAsync(dout, 0, sizeof(float) * n * kPairs, s));
  const unsigned g = grid_for(n << logk);
  hipLaunchKernelGGL(kl_w_kernel, dim3(g), dim3(256), 0, s, n, logk, ic->d_uptr, ic->d_ucol, ic->d_ulow, ic->d_l64,
                     ic->d_dg64, x, ic->d_klw, ic->d_part);
  hipLaunchKernelGGL(sum_one_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal + S_X * 64 + 1);
  const unsigned g2 = grid_for(n);
  hipLaunchKernelGGL(kl_grad_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, n, logk, ic->d_lptr, ic->d_lcol,
                     ic->d_lch, ic->d_dg64, ic->d_adiag, x, ic->d_klw, scale, dout);
  hipLaunchKernelGGL(logdet_diag_kernel, dim3(g2), dim3(256), 0, s, n, ic->d_dg64, ic->d_part);
  hipLaunchKernelGGL(sum_one_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g2, ic->d_part, ic->d_scal + S_X * 64 + 2);
  SPDE_HIP(hipGetLastError());
  if (loss) {
    double h[2];
    SPDE_HIP(hipMemcpyAsync(h, ic->d_scal + S_X * 64 + 1, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    SPDE_HIP(hipStreamSynchronize(s));
    *loss = 0.5 * h[0] / k - h[1];
  }
  return SPDE_OK;
}


int spde_ic_jeffreys_grad(spde_ic* ic, const double* x, const float* z, int k, double scale, float* dout,
                          double* loss, void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !ic->factored || !x || !z || !dout || logk < 0) {
    set_error("spde_ic_jeffreys_grad: invalid argument (k a power of two <= 64, after spde_ic_factor)");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = ic->n;
  int rc = ensure_work(ic);
  if (rc) return rc;
  if (!ic->kl_cap || ic->kl_cap < (n << logk)) {
    if (ic->d_klw) SPDE_HIP(hipFree(ic->d_klw));
    ic->d_klw = nullptr;
    SPDE_HIP(hipMalloc((void**)&ic->d_klw, sizeof(double) * (n << logk)));
    ic->kl_cap = n << logk;
  }
  float* Y = ic->d_work;
  float* V = Y + n * 64;
  float* U = V + n * 64;
  SPDE_HIP(hipMemsetAsync(dout, 0, sizeof(float) * n * kPairs, s));
  const unsigned g = grid_for(n << logk);
  hipLaunchKernelGGL(kl_w_kernel, dim3(g), dim3(256), 0, s, n, logk, ic->d_uptr, ic->d_ucol, ic->d_ulow, ic->d_l64,
                     ic->d_dg64, x, ic->d_klw, ic->d_part);
  hipLaunchKernelGGL(sum_one_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g, ic->d_part, ic->d_scal + S_X * 64 + 1);
  ic_bwd(ic, z, Y, logk, s);                                                     // y = L^-T z
  launch_spmm(ic, logk, g, Y, V, ic->d_part, s);                                 // v = A y, partials of y^T A y
  hipLaunchKernelGGL(sum_rhs_kernel, dim3(1), dim3(kRedThreads), 0, s, (int)g, k, ic->d_part, ic->d_scal + S_X * 64 + 2);
  ic_fwd(ic, V, U, logk, s);                                                     // u = L^-1 A y
  hipLaunchKernelGGL(jeffreys_grad_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, n, logk, ic->d_lptr,
                     ic->d_lcol, ic->d_lch, ic->d_dg64, ic->d_adiag, x, ic->d_klw, Y, U, scale, dout);
  SPDE_HIP(hipGetLastError());
  if (loss) {
    double h[2];
    SPDE_HIP(hipMemcpyAsync(h, ic->d_scal + S_X * 64 + 1, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    SPDE_HIP(hipStreamSynchronize(s));
    *loss = 0.5 * h[0] / k + 0.5 * h[1] / k - (double)n;
  }
  return SPDE_OK;
}

int spde_hutch_pattern(spde_ic* ic, const float* U, const float* V, int k, double alpha, double beta, double* sigma,
                       void* stream) {
  sync_view(ic);
  const int logk = log2_of(k);
  if (!ic || !U || !V || !sigma || logk < 0) {
    set_error("spde_hutch_pattern: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipLaunchKernelGGL(hutch_pattern_kernel, dim3((unsigned)((ic->n + 3) / 4)), dim3(256), 0, (hipStream_t)stream, ic->n,
                     logk, ic->pl->d_indptr, ic->pl->d_indices, U, V, alpha, beta, sigma);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

}  // extern "C"
