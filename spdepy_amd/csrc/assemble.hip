// Finite-volume stencils and the assembly of the space-time precision Q.
//
// The per-cell stencil functions reproduce the reference's native stencils
// (berild/spdepy src/spdepy/spdes/ccode/{AcH,AH,Acw,Aw}_2D_b{1,2,3}.cpp)
// operation for operation; floating-point contraction is disabled in this
// file so that the slot values match the reference bit for bit.
//
// Assembly of Q (makeQ, advection_idiffusion2D.py:85-116 and
// var_advection_var_diffusion2D.py:83-116) is restructured for the GPU:
//   1. canonical 9-point A (and A0) per cell, duplicates merged;
//   2. per-cell value templates of the five distinct block rows
//      (Q00 = Q0 + c Qs, Qtt = c(B^T Qs B + Qs), Q_{T-1} = c B^T Qs B,
//       Q_{t,t+1} = -c Qs B, Q_{t+1,t} = -c B^T Qs; B = A/V, c = 1/(dt sigma));
//   3. a bandwidth-bound expansion of the templates onto the structural CSR.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "stencil.hpp"

namespace spde {

// ------------------------------------------------- reference-layout kernels
__global__ void stencil_diffusion_kernel(int M, int N, const double* __restrict__ H, int variable, double hx,
                                         double hy, int bc, int32_t* __restrict__ row, int32_t* __restrict__ col,
                                         double* __restrict__ val) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M * N) return;
  const int i = k % M, j = k / M;
  Cell9 c;
  if (variable) {
    diffusion_var(i, j, M, N, bc, hx, hy, H + 16 * (int64_t)k, c);
  } else {
    double h[4] = {H[0], H[1], H[2], H[3]};
    diffusion_const(i, j, M, N, bc, hx, hy, h, c);
  }
  for (int s = 0; s < 9; ++s) {
    row[9 * k + s] = c.keep[s] ? k : M * N;
    col[9 * k + s] = c.col[s];
    val[9 * k + s] = c.val[s];
  }
}

__global__ void stencil_advection_kernel(int M, int N, const double* __restrict__ G, const double* __restrict__ dG,
                                         int variable, double hx, double hy, int diff, int bc,
                                         int32_t* __restrict__ row, int32_t* __restrict__ col,
                                         double* __restrict__ val) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M * N) return;
  const int i = k % M, j = k / M;
  Cell5 c;
  if (variable)
    advection_var(i, j, M, N, bc, hx, hy, G + 4 * (int64_t)k, dG ? dG + 4 * (int64_t)k : nullptr, diff, c);
  else
    advection_const(i, j, M, N, bc, hx, hy, G[0], G[1], diff, c);
  for (int s = 0; s < 5; ++s) {
    row[5 * k + s] = c.keep[s] ? k : M * N;
    col[5 * k + s] = c.col[s];
    val[5 * k + s] = c.val[s];
  }
}

// ------------------------------------------------------ canonical operator
struct CoeffArgs {
  int variable, stationary_form;
  double hx, hy, dt, V;
  double kappa_c, H_c[4], G_c[2];
  const double* kappa;
  const double* H;
  const double* G;
  double kappa0, H0[4];
  int a0_variable;
  int mix;  // with variable: 0 all fields, 1 constant advection G_c, 2 constant kappa_c and H_c
};

// A = Dv + (Dv Dk - A_H + A_w) dt (B-spline models) or
// A = Dv + Dv Dk dt - A_H dt + A_w dt (stationary models); A0 = Dv kappa0 - A_H(H0).
__global__ void canonical_kernel(int M, int N, int bc, CoeffArgs a, double* __restrict__ acan,
                                 double* __restrict__ a0can, double* __restrict__ q) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M * N) return;
  const int i = k % M, j = k / M;
  Cell9 d;
  Cell5 w;
  double kap;
  const bool kvar = a.variable && a.mix != 2, dvar = a.variable && a.mix != 2, gvar = a.variable && a.mix != 1;
  if (dvar)
    diffusion_var(i, j, M, N, bc, a.hx, a.hy, a.H + 16 * (int64_t)k, d);
  else
    diffusion_const(i, j, M, N, bc, a.hx, a.hy, a.H_c, d);
  if (gvar)
    advection_var(i, j, M, N, bc, a.hx, a.hy, a.G + 4 * (int64_t)k, nullptr, 3, w);
  else
    advection_const(i, j, M, N, bc, a.hx, a.hy, a.G_c[0], a.G_c[1], 3, w);
  kap = kvar ? a.kappa[k] : a.kappa_c;
  double ah[9], aw[9];
  for (int o = 0; o < 9; ++o) ah[o] = aw[o] = 0.0;
  for (int s = 0; s < 9; ++s)
    if (d.keep[s]) {
      const int o = canon(d.col[s], i, j, M, N);
      ah[o] = ah[o] + d.val[s];
    }
  for (int s = 0; s < 5; ++s)
    if (w.keep[s]) {
      const int o = canon(w.col[s], i, j, M, N);
      aw[o] = aw[o] + w.val[s];
    }
  const double V = a.V, dt = a.dt;
  for (int o = 0; o < 9; ++o) {
    double v;
    if (a.stationary_form)
      v = (o == 4) ? ((V + V * kap * dt) - ah[o] * dt) + aw[o] * dt : -(ah[o] * dt) + aw[o] * dt;
    else
      v = (o == 4) ? V + ((V * kap - ah[o]) + aw[o]) * dt : (-ah[o] + aw[o]) * dt;
    acan[9 * (int64_t)k + o] = v;
  }
  q[k] = ((V * kap) * (1.0 / V)) * (V * kap);
  if (a.a0_variable) {  // A0 = Dv Dk - A_H(H) with the variable fields
    for (int o = 0; o < 9; ++o) a0can[9 * (int64_t)k + o] = (o == 4) ? V * kap - ah[o] : -ah[o];
    return;
  }
  diffusion_const(i, j, M, N, bc, a.hx, a.hy, a.H0, d);
  double a0[9];
  for (int o = 0; o < 9; ++o) a0[o] = 0.0;
  for (int s = 0; s < 9; ++s)
    if (d.keep[s]) {
      const int o = canon(d.col[s], i, j, M, N);
      a0[o] = a0[o] + d.val[s];
    }
  for (int o = 0; o < 9; ++o) a0can[9 * (int64_t)k + o] = (o == 4) ? V * a.kappa0 - a0[o] : -a0[o];
}

// Template layout per cell (kTmpl = 143 doubles): T0[25] | Tmid[25] | Tlast[25]
// | Tup[9] | Tlo[9] | Gram[25] | Q0[25].  Gram = B^T B with B = A / V and
// Q0 = A0^T A0 / V are the spatial matrices of the closed-form log|Q|
// (Q = F^T D F with F block bidiagonal: log|Q| = log|Q0| + (T-1) log|B^T B|
// + (T-1) sum_k log(q_k c)).
__global__ void template_kernel(int M, int N, int bc, double V, double c, const double* __restrict__ acan,
                                const double* __restrict__ a0can, const double* __restrict__ q,
                                double* __restrict__ tmpl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M * N) return;
  const int i = k % M, j = k / M;
  double* T = tmpl + kTmpl * (int64_t)k;
  const double iV = 1.0 / V;
  for (int wy = -2; wy <= 2; ++wy)
    for (int wx = -2; wx <= 2; ++wx) {
      double x = 0.0, x0 = 0.0, xg = 0.0;
      for (int d1y = -1; d1y <= 1; ++d1y)
        for (int d1x = -1; d1x <= 1; ++d1x) {
          const int d2x = wx + d1x, d2y = wy + d1y;
          if (d2x < -1 || d2x > 1 || d2y < -1 || d2y > 1) continue;
          const int m = nbr_cell(i, j, -d1x, -d1y, M, N, bc);
          if (m < 0) continue;
          const int o1 = (d1y + 1) * 3 + d1x + 1, o2 = (d2y + 1) * 3 + d2x + 1;
          const double* am = acan + 9 * (int64_t)m;
          const double* bm = a0can + 9 * (int64_t)m;
          x += ((am[o1] * iV) * q[m]) * iV * am[o2];
          x0 += (bm[o1] * iV) * bm[o2];
          xg += (am[o1] * iV) * (am[o2] * iV);
        }
      const int w = (wy + 2) * 5 + wx + 2;
      const double diag = (wx == 0 && wy == 0) ? q[k] : 0.0;
      T[w] = x0 + c * diag;
      T[25 + w] = c * (x + diag);
      T[50 + w] = c * x;
      T[93 + w] = xg;
      T[118 + w] = x0;
    }
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      const int o = (dy + 1) * 3 + dx + 1;
      T[75 + o] = -c * (q[k] * iV * acan[9 * (int64_t)k + o]);
      const int jn = nbr_cell(i, j, dx, dy, M, N, bc);
      double lo = 0.0;
      if (jn >= 0) lo = -c * (acan[9 * (int64_t)jn + (8 - o)] * iV * q[jn]);
      T[84 + o] = lo;
    }
}


// ------------------------------------------------------ tangent (dQ_i)
// Direction of one parameter of the reference's dQ list, in coefficient
// terms (see spde_direction in the C header).
struct DirArgs {
  int kind;                 // 0 coefficient fields, 1 log sigma, 2 initial-state operator
  const double* dkappa;     // per cell, or NULL (then dkappa_c)
  double dkappa_c;
  const double* dH;         // per cell faces W,E,S,N (16), or NULL (then dH_c)
  double dH_c[4];
  int has_dH;
  const double* dG;         // variable advection direction (4 faces per cell), or NULL
  int adv_diff;             // 0 none; 1 d/dwx, 2 d/dwy derivative stencil
  double dkappa0, dH0[4];
};

// Canonical 9-point dA, dA0 and dq of the direction at the plan's base point
// (the fields of the last spde_assemble): the reference forms each dQ_i from
// dA = dt (V dkappa I - A_H(dH) + Aw'(G; dG)) and dQs = 2 V kappa dkappa
// (var_advection_var_diffusion2D.py:131-183, advection_idiffusion2D.py:
// 120-155), with NaN slots of the advection derivative set to 0 (:255).
__global__ void tangent_canonical_kernel(int M, int N, int bc, CoeffArgs a, DirArgs d, double* __restrict__ dacan,
                                         double* __restrict__ da0can, double* __restrict__ dq) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M * N) return;
  const int i = k % M, j = k / M;
  double dah[9], daw[9], da0[9];
  for (int o = 0; o < 9; ++o) dah[o] = daw[o] = da0[o] = 0.0;
  double dkap = 0.0;
  if (d.kind == 0) {
    dkap = d.dkappa ? d.dkappa[k] : d.dkappa_c;
    if (d.has_dH) {
      Cell9 c9;
      if (d.dH)
        diffusion_var(i, j, M, N, bc, a.hx, a.hy, d.dH + 16 * (int64_t)k, c9);
      else
        diffusion_const(i, j, M, N, bc, a.hx, a.hy, d.dH_c, c9);
      for (int s = 0; s < 9; ++s)
        if (c9.keep[s]) {
          const int o = canon(c9.col[s], i, j, M, N);
          dah[o] = dah[o] + c9.val[s];
        }
    }
    if (d.adv_diff) {
      Cell5 c5;
      if (d.dG)
        advection_var(i, j, M, N, bc, a.hx, a.hy, a.G + 4 * (int64_t)k, d.dG + 4 * (int64_t)k, d.adv_diff, c5);
      else
        advection_const(i, j, M, N, bc, a.hx, a.hy, a.G_c[0], a.G_c[1], d.adv_diff, c5);
      for (int s = 0; s < 5; ++s)
        if (c5.keep[s]) {
          const double v = c5.val[s];
          const int o = canon(c5.col[s], i, j, M, N);
          daw[o] = daw[o] + (v == v ? v : 0.0);
        }
    }
  } else if (d.kind == 2) {
    Cell9 c9;
    diffusion_const(i, j, M, N, bc, a.hx, a.hy, d.dH0, c9);
    for (int s = 0; s < 9; ++s)
      if (c9.keep[s]) {
        const int o = canon(c9.col[s], i, j, M, N);
        da0[o] = da0[o] + c9.val[s];
      }
  }
  const double V = a.V, dt = a.dt;
  const double kap = (a.variable && a.mix != 2) ? a.kappa[k] : a.kappa_c;
  for (int o = 0; o < 9; ++o) {
    dacan[9 * (int64_t)k + o] = ((o == 4 ? V * dkap : 0.0) - dah[o] + daw[o]) * dt;
    if (d.kind == 2)
      da0can[9 * (int64_t)k + o] = (o == 4 ? V * d.dkappa0 : 0.0) - da0[o];
    else if (a.a0_variable)
      da0can[9 * (int64_t)k + o] = (o == 4 ? V * dkap : 0.0) - dah[o];
    else
      da0can[9 * (int64_t)k + o] = 0.0;
  }
  dq[k] = 2.0 * V * kap * dkap;
}

// Directional derivative of the templates: the products A^T Qs A / V^2,
// A0^T A0 / V and the Qs terms of template_kernel differentiated by the
// product rule, and dlc = d log c / d theta (-1 for log sigma, else 0) for
// the factor c = 1 / (dt sigma).
__global__ void template_tangent_kernel(int M, int N, int bc, double V, double c, double dlc,
                                        const double* __restrict__ acan, const double* __restrict__ a0can,
                                        const double* __restrict__ q, const double* __restrict__ dacan,
                                        const double* __restrict__ da0can, const double* __restrict__ dq,
                                        double* __restrict__ tmpl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M * N) return;
  const int i = k % M, j = k / M;
  double* T = tmpl + kTmpl * (int64_t)k;
  const double iV = 1.0 / V;
  for (int wy = -2; wy <= 2; ++wy)
    for (int wx = -2; wx <= 2; ++wx) {
      double x = 0.0, xt = 0.0, x0t = 0.0;
      for (int d1y = -1; d1y <= 1; ++d1y)
        for (int d1x = -1; d1x <= 1; ++d1x) {
          const int d2x = wx + d1x, d2y = wy + d1y;
          if (d2x < -1 || d2x > 1 || d2y < -1 || d2y > 1) continue;
          const int m = nbr_cell(i, j, -d1x, -d1y, M, N, bc);
          if (m < 0) continue;
          const int o1 = (d1y + 1) * 3 + d1x + 1, o2 = (d2y + 1) * 3 + d2x + 1;
          const double* am = acan + 9 * (int64_t)m;
          const double* dm = dacan + 9 * (int64_t)m;
          const double* bm = a0can + 9 * (int64_t)m;
          const double* em = da0can + 9 * (int64_t)m;
          x += ((am[o1] * iV) * q[m]) * iV * am[o2];
          xt += ((dm[o1] * iV) * q[m]) * iV * am[o2] + ((am[o1] * iV) * q[m]) * iV * dm[o2] +
                ((am[o1] * iV) * dq[m]) * iV * am[o2];
          x0t += (em[o1] * iV) * bm[o2] + (bm[o1] * iV) * em[o2];
        }
      const int w = (wy + 2) * 5 + wx + 2;
      const bool ctr = wx == 0 && wy == 0;
      const double diag = ctr ? q[k] : 0.0, ddiag = ctr ? dq[k] : 0.0;
      T[w] = x0t + c * (dlc * diag + ddiag);
      T[25 + w] = c * (dlc * (x + diag) + xt + ddiag);
      T[50 + w] = c * (dlc * x + xt);
      T[118 + w] = x0t;
    }
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      const int o = (dy + 1) * 3 + dx + 1;
      const double* ak = acan + 9 * (int64_t)k;
      const double* dk = dacan + 9 * (int64_t)k;
      T[75 + o] = -c * (dlc * (q[k] * iV * ak[o]) + dq[k] * iV * ak[o] + q[k] * iV * dk[o]);
      const int jn = nbr_cell(i, j, dx, dy, M, N, bc);
      double lo = 0.0;
      if (jn >= 0) {
        const double* aj = acan + 9 * (int64_t)jn;
        const double* dj = dacan + 9 * (int64_t)jn;
        lo = -c * (dlc * (aj[8 - o] * iV * q[jn]) + dj[8 - o] * iV * q[jn] + aj[8 - o] * iV * dq[jn]);
      }
      T[84 + o] = lo;
    }
}

// Expansion of the templates onto the structural CSR: one block per 128 rows,
// row of an entry found by binary search over the block's row pointers in LDS.
__global__ __launch_bounds__(256) void expand_kernel(int64_t n, int64_t Ns, int T, const int64_t* __restrict__ indptr,
                                                     const uint8_t* __restrict__ code, const double* __restrict__ tmpl,
                                                     double* __restrict__ qval) {
  __shared__ int64_t rp[129];
  const int64_t r0 = (int64_t)blockIdx.x * 128;
  const int nr = (int)((n - r0) < 128 ? (n - r0) : 128);
  for (int q = threadIdx.x; q <= nr; q += 256) rp[q] = indptr[r0 + q];
  __syncthreads();
  const int64_t e0 = rp[0], e1 = rp[nr];
  for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rp[mid] <= e) lo = mid; else hi = mid - 1;
    }
    const int64_t r = r0 + lo;
    const int64_t cell = r % Ns;
    const int t = (int)(r / Ns);
    const int cd = code[e];
    const double* Tc = tmpl + kTmpl * cell;
    double v;
    if (cd < 25)
      v = (t == 0) ? Tc[cd] : ((t == T - 1) ? Tc[50 + cd] : Tc[25 + cd]);
    else if (cd < 34)
      v = Tc[84 + (cd - 25)];
    else
      v = Tc[75 + (cd - 34)];
    qval[e] = v;
  }
}

// Spatial matrix (Gram or Q0 block of the templates) onto the structural
// CSR of a one-slice plan (row = cell, code = 5x5 offset).
__global__ void spatial_expand_kernel(int64_t Ns, const int64_t* __restrict__ indptr, const uint8_t* __restrict__ code,
                                      const double* __restrict__ tmpl, int off, double* __restrict__ out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= Ns) return;
  const double* Tc = tmpl + kTmpl * r + off;
  for (int64_t e = indptr[r]; e < indptr[r + 1]; ++e) out[e] = Tc[code[e]];
}

// out[0] = sum_k log(q_k c), one block, fixed order (deterministic).
__global__ __launch_bounds__(1024) void logq_sum_kernel(int64_t Ns, const double* __restrict__ q, double c,
                                                        double* __restrict__ out) {
  __shared__ double red[1024];
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < Ns; k += 1024) acc += log(q[k] * c);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// --------------------------------------------------------- prune and SpMV
__global__ void count_nz_kernel(int64_t n, const int64_t* __restrict__ indptr, const double* __restrict__ q,
                                int64_t* __restrict__ cnt) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t c = 0;
  for (int64_t e = indptr[r]; e < indptr[r + 1]; ++e) c += (q[e] != 0.0);
  cnt[r] = c;
}

// Exclusive scan of cnt[0..n) into out[0..n], single block, in chunks.
__global__ __launch_bounds__(1024) void scan_kernel(int64_t n, const int64_t* __restrict__ cnt, int64_t* __restrict__ out) {
  __shared__ int64_t buf[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += 1024) {
    const int64_t r = base + threadIdx.x;
    const int64_t v = (r < n) ? cnt[r] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int64_t add = (threadIdx.x >= (unsigned)off) ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (r < n) out[r] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ void prune_write_kernel(int64_t n, const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                   const double* __restrict__ q, const int64_t* __restrict__ outptr,
                                   int32_t* __restrict__ oind, double* __restrict__ oval) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t o = outptr[r];
  for (int64_t e = indptr[r]; e < indptr[r + 1]; ++e)
    if (q[e] != 0.0) {
      oind[o] = indices[e];
      oval[o] = q[e];
      ++o;
    }
}

// Sixteen lanes per row: lane l takes entries l, l + 16, ... of the row
// (consecutive lanes read consecutive entries), then a fixed butterfly over
// the sixteen lanes (deterministic).
__global__ __launch_bounds__(256) void spmv_kernel(int64_t n, int nrhs, const int64_t* __restrict__ indptr,
                                                   const int32_t* __restrict__ indices, const double* __restrict__ q,
                                                   const double* __restrict__ x, double* __restrict__ y) {
  const int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 4;
  const int l = threadIdx.x & 15;
  if (r >= n) return;  // uniform over the sixteen lanes of a row
  const int64_t e0 = indptr[r], e1 = indptr[r + 1];
  for (int c = 0; c < nrhs; ++c) {
    double s = 0.0;
    for (int64_t e = e0 + l; e < e1; e += 16) s += q[e] * x[indices[e] + (int64_t)c * n];
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (l == 0) y[r + (int64_t)c * n] = s;
  }
}

static unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// *out (device) = sum_k log(q_k c), enqueued without waiting.
void launch_logq_sum(int64_t Ns, const double* q, double c, double* out, hipStream_t s) {
  hipLaunchKernelGGL(logq_sum_kernel, dim3(1), dim3(1024), 0, s, Ns, q, c, out);
}

}  // namespace spde

using namespace spde;


// ------------------------------------------------------ F-form square root
// out = F^T D^(1/2) z for Q = F^T D F: F has V^(-1/2) A0 on block 0, A / V on
// blocks t >= 1 and -I below the diagonal; D = blockdiag(I, c diag(q), ...).
// With w_0 = z_0, w_t = sqrt(c q) z_t:  (F^T w)_t = F_tt^T w_t - w_(t+1).
// A^T w at cell k sums A[m, k] w_m over the rows m whose canonical slot
// (d1x, d1y) points at k (m = k - (d1x, d1y)), as the block formula does.
// One thread per (lattice index, right-hand side); z, out row-major n x K.
__global__ __launch_bounds__(256) void fform_t_kernel(int M, int N, int T, int bc, int K, double V, double c,
                                                      const double* __restrict__ acan,
                                                      const double* __restrict__ a0can,
                                                      const double* __restrict__ q, const double* __restrict__ z,
                                                      double* __restrict__ out) {
  const int64_t Ns = (int64_t)M * N;
  const int64_t units = Ns * T * K;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(u % K);
    const int64_t idx = u / K;
    const int t = (int)(idx / Ns);
    const int k = (int)(idx - (int64_t)t * Ns);
    const int i = k % M, j = k / M;
    const double* op = t == 0 ? a0can : acan;
    const double* zt = z + (int64_t)t * Ns * K;
    double acc = 0.0;
    for (int d1y = -1; d1y <= 1; ++d1y)
      for (int d1x = -1; d1x <= 1; ++d1x) {
        const int m = nbr_cell(i, j, -d1x, -d1y, M, N, bc);
        if (m < 0) continue;
        const int o1 = (d1y + 1) * 3 + d1x + 1;
        const double wm = t == 0 ? zt[(int64_t)m * K + r] : sqrt(c * q[m]) * zt[(int64_t)m * K + r];
        acc += op[9 * (int64_t)m + o1] * wm;
      }
    double v = t == 0 ? acc / sqrt(V) : acc / V;
    if (t + 1 < T) v -= sqrt(c * q[k]) * z[((int64_t)(t + 1) * Ns + k) * K + r];
    out[u] = v;
  }
}

// out[rows[a] * K + r] += w[a * K + r] for distinct rows (one thread per entry).
__global__ __launch_bounds__(256) void add_rows_kernel(int64_t nrows, int K, const int64_t* __restrict__ rows,
                                                       const double* __restrict__ w, double* __restrict__ out) {
  const int64_t units = nrows * K;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < units; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = u / K;
    const int r = (int)(u - a * K);
    out[rows[a] * K + r] += w[u];
  }
}

extern "C" {

int spde_stencil_diffusion(int M, int N, const double* H, int variable, double hx, double hy, int bc, int32_t* row,
                           int32_t* col, double* val, void* stream) {
  if (M < 1 || N < 1 || !H || !row || !col || !val || bc < 1 || bc > 3) {
    set_error("spde_stencil_diffusion: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipLaunchKernelGGL(stencil_diffusion_kernel, dim3(blocks_for((int64_t)M * N, 128)), dim3(128), 0,
                     (hipStream_t)stream, M, N, H, variable, hx, hy, bc, row, col, val);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_stencil_advection(int M, int N, const double* G, const double* dG, int variable, double hx, double hy,
                           int diff, int bc, int32_t* row, int32_t* col, double* val, void* stream) {
  if (M < 1 || N < 1 || !G || !row || !col || !val || bc < 1 || bc > 3 || diff < 1 || diff > 3) {
    set_error("spde_stencil_advection: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipLaunchKernelGGL(stencil_advection_kernel, dim3(blocks_for((int64_t)M * N, 128)), dim3(128), 0,
                     (hipStream_t)stream, M, N, G, dG, variable, hx, hy, diff, bc, row, col, val);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_assemble(spde_plan* pl, const spde_coeffs* c, double* qval, void* stream) {
  if (!pl || pl->host_only || !c || !qval) {
    set_error("spde_assemble: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (c->variable && (c->mix < 0 || c->mix > 2 || (c->mix != 2 && (!c->kappa || !c->H)) || (c->mix != 1 && !c->G))) {
    set_error("spde_assemble: variable coefficients need their fields (kappa, H unless mix = 2; G unless mix = 1)");
    return SPDE_ERR_ARG;
  }
  if (pl->lat.wide) {
    set_error("spde_assemble: a wide-time plan takes its values from spde_assemble_kron");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int M = pl->lat.M, N = pl->lat.N, bc = pl->lat.bc;
  const int64_t Ns = pl->lat.Ns();
  CoeffArgs a;
  a.variable = c->variable;
  a.stationary_form = !c->variable;
  a.hx = c->hx;
  a.hy = c->hy;
  a.dt = c->dt;
  a.V = c->V;
  a.kappa_c = c->kappa_c;
  for (int q = 0; q < 4; ++q) a.H_c[q] = c->H_c[q], a.H0[q] = c->H0[q];
  a.G_c[0] = c->G_c[0];
  a.G_c[1] = c->G_c[1];
  a.kappa = c->kappa;
  a.H = c->H;
  a.G = c->G;
  a.kappa0 = c->kappa0;
  a.a0_variable = c->a0_variable;
  a.mix = c->variable ? c->mix : 0;
  if (a.a0_variable && !c->variable) {
    set_error("spde_assemble: a0_variable needs the variable fields");
    return SPDE_ERR_ARG;
  }
  double* qdiag = pl->d_qdiag;
  hipLaunchKernelGGL(canonical_kernel, dim3(blocks_for(Ns, 128)), dim3(128), 0, s, M, N, bc, a, pl->d_acan,
                     pl->d_a0can, qdiag);
  const double cc = 1.0 / (c->dt * c->sigma);
  hipLaunchKernelGGL(template_kernel, dim3(blocks_for(Ns, 128)), dim3(128), 0, s, M, N, bc, c->V, cc, pl->d_acan,
                     pl->d_a0can, qdiag, pl->d_tmpl);
  const int64_t n = pl->sym.n;
  hipLaunchKernelGGL(expand_kernel, dim3(blocks_for(n, 128)), dim3(256), 0, s, n, Ns, pl->lat.T, pl->d_indptr,
                     pl->d_code, pl->d_tmpl, qval);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_assemble_spatial(spde_plan* st, spde_plan* sp, int which, double* out, void* stream) {
  if (!st || st->host_only || !sp || sp->host_only || !out || which < 0 || which > 1) {
    set_error("spde_assemble_spatial: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (sp->lat.T != 1 || sp->lat.M != st->lat.M || sp->lat.N != st->lat.N || sp->lat.bc != st->lat.bc ||
      st->lat.wide) {
    set_error("spde_assemble_spatial: the spatial plan must be the one-slice lattice of the space-time plan");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t Ns = sp->lat.Ns();
  hipLaunchKernelGGL(spatial_expand_kernel, dim3(blocks_for(Ns, 128)), dim3(128), 0, s, Ns, sp->d_indptr, sp->d_code,
                     st->d_tmpl, which == 0 ? 118 : 93, out);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_logq_sum(spde_plan* pl, double c, double* out, void* stream) {
  if (!pl || pl->host_only || !out || !(c > 0.0) || pl->lat.wide) {
    set_error("spde_logq_sum: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(logq_sum_kernel, dim3(1), dim3(1024), 0, s, pl->lat.Ns(), pl->d_qdiag, c, pl->d_red);
  SPDE_HIP(hipGetLastError());
  SPDE_HIP(hipMemcpyAsync(out, pl->d_red, sizeof(double), hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipStreamSynchronize(s));
  return SPDE_OK;
}


int spde_assemble_tangent(spde_plan* pl, const spde_coeffs* c, const spde_direction* dir, double* dqval, void* stream) {
  if (!pl || pl->host_only || !c || !dir || !dqval || pl->lat.wide || dir->kind < 0 || dir->kind > 2 ||
      dir->adv_diff < 0 || dir->adv_diff > 2 || (dir->dG && !c->variable)) {
    set_error("spde_assemble_tangent: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int M = pl->lat.M, N = pl->lat.N, bc = pl->lat.bc;
  const int64_t Ns = pl->lat.Ns();
  if (!pl->d_tan) {
    SPDE_HIP(hipStreamSynchronize(s));
    SPDE_HIP(hipMalloc((void**)&pl->d_tan, sizeof(double) * Ns * (kTmpl + 19)));
  }
  double* ttan = pl->d_tan;
  double* dacan = ttan + Ns * kTmpl;
  double* da0can = dacan + Ns * 9;
  double* dq = da0can + Ns * 9;
  CoeffArgs a{};
  a.variable = c->variable;
  a.stationary_form = !c->variable;
  a.hx = c->hx;
  a.hy = c->hy;
  a.dt = c->dt;
  a.V = c->V;
  a.kappa_c = c->kappa_c;
  for (int q = 0; q < 4; ++q) a.H_c[q] = c->H_c[q], a.H0[q] = c->H0[q];
  a.G_c[0] = c->G_c[0];
  a.G_c[1] = c->G_c[1];
  a.kappa = c->kappa;
  a.H = c->H;
  a.G = c->G;
  a.kappa0 = c->kappa0;
  a.a0_variable = c->a0_variable;
  a.mix = c->variable ? c->mix : 0;
  DirArgs d{};
  d.kind = dir->kind;
  d.dkappa = dir->dkappa;
  d.dkappa_c = dir->dkappa_c;
  d.dH = dir->dH;
  for (int q = 0; q < 4; ++q) d.dH_c[q] = dir->dH_c[q], d.dH0[q] = dir->dH0[q];
  d.has_dH = dir->has_dH;
  d.dG = dir->dG;
  d.adv_diff = dir->adv_diff;
  d.dkappa0 = dir->dkappa0;
  hipLaunchKernelGGL(tangent_canonical_kernel, dim3(blocks_for(Ns, 128)), dim3(128), 0, s, M, N, bc, a, d, dacan,
                     da0can, dq);
  const double cc = 1.0 / (c->dt * c->sigma);
  hipLaunchKernelGGL(template_tangent_kernel, dim3(blocks_for(Ns, 128)), dim3(128), 0, s, M, N, bc, c->V, cc,
                     dir->kind == 1 ? -1.0 : 0.0, pl->d_acan, pl->d_a0can, pl->d_qdiag, dacan, da0can, dq, ttan);
  const int64_t n = pl->sym.n;
  if (pl->lat.T == 1)
    hipLaunchKernelGGL(spatial_expand_kernel, dim3(blocks_for(Ns, 128)), dim3(128), 0, s, Ns, pl->d_indptr, pl->d_code,
                       ttan, 118, dqval);
  else
    hipLaunchKernelGGL(expand_kernel, dim3(blocks_for(n, 128)), dim3(256), 0, s, n, Ns, pl->lat.T, pl->d_indptr,
                       pl->d_code, ttan, dqval);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_prune(spde_plan* pl, const double* qval, int64_t* indptr, int32_t* indices, double* data, int64_t* nnz,
               void* stream) {
  if (!pl || pl->host_only || !qval || !indptr || !indices || !data) {
    set_error("spde_prune: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = pl->sym.n;
  int64_t* cnt = nullptr;
  SPDE_HIP(hipMallocAsync((void**)&cnt, sizeof(int64_t) * n, s));
  hipLaunchKernelGGL(count_nz_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, pl->d_indptr, qval, cnt);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, n, cnt, indptr);
  hipLaunchKernelGGL(prune_write_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, pl->d_indptr, pl->d_indices,
                     qval, indptr, indices, data);
  SPDE_HIP(hipFreeAsync(cnt, s));
  if (nnz) {
    SPDE_HIP(hipMemcpyAsync(nnz, indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SPDE_HIP(hipStreamSynchronize(s));
  }
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_spmv(spde_plan* pl, const double* qval, const double* x, double* y, int nrhs, void* stream) {
  if (!pl || pl->host_only || !qval || !x || !y || nrhs < 1) {
    set_error("spde_spmv: invalid argument");
    return SPDE_ERR_ARG;
  }
  const int64_t n = pl->sym.n;
  hipLaunchKernelGGL(spmv_kernel, dim3(blocks_for(n * 16, 256)), dim3(256), 0, (hipStream_t)stream, n, nrhs,
                     pl->d_indptr, pl->d_indices, qval, x, y);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_fform_apply_t(spde_plan* pl, const spde_coeffs* c, const double* z, double* out, int K, void* stream) {
  if (!pl || pl->host_only || !c || !z || !out || K < 1 || pl->lat.wide || !(c->V > 0.0) ||
      !(c->dt * c->sigma > 0.0)) {
    set_error("spde_fform_apply_t: invalid argument (a device plan of an advection-diffusion lattice)");
    return SPDE_ERR_ARG;
  }
  const int64_t units = pl->sym.n * (int64_t)K;
  hipLaunchKernelGGL(fform_t_kernel, dim3(blocks_for(units, 256)), dim3(256), 0, (hipStream_t)stream, pl->lat.M,
                     pl->lat.N, pl->lat.T, pl->lat.bc, K, c->V, 1.0 / (c->dt * c->sigma), pl->d_acan, pl->d_a0can,
                     pl->d_qdiag, z, out);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_add_rows(int64_t n, int K, int64_t nrows, const int64_t* rows, const double* w, double* out, void* stream) {
  if (n < 1 || K < 1 || nrows < 0 || (nrows > 0 && (!rows || !w)) || !out) {
    set_error("spde_add_rows: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (nrows == 0) return SPDE_OK;
  hipLaunchKernelGGL(add_rows_kernel, dim3(blocks_for(nrows * K, 256)), dim3(256), 0, (hipStream_t)stream, nrows, K,
                     rows, w, out);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

}  // extern "C"
