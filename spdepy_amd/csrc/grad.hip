// Gradient of the log-likelihood by the adjoint of the assembly.
//
// The reference forms one matrix dQ_i per parameter and contracts it with the
// trace and data terms (advection_idiffusion2D.py:186-189,
// var_advection_var_diffusion2D.py:218-221):
//     g_i = <G, dQ_i>,  G = 1/2 r (Q^-1 - Q_c^-1) - 1/2 mu_c mu_c^T  (on pattern(Q)).
// Every dQ_i is the directional derivative of the same block formula, so the
// contraction is done once, backwards: G is reduced over time into per-cell
// windows, differentiated through the block formula into adjoints of the
// spatial operator A (per cell and canonical offset), of diag(Qs) and of A0,
// and then through the per-cell stencils (unit perturbations of the very
// functions the assembly uses, so the reference's boundary folds, the bc=1
// stale-index zeroing and the NaN->0 rule of the advection derivative carry
// over exactly).  Cost O(nnz(Q)) instead of O(#params * nnz(Q)).
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "stencil.hpp"

namespace spde {

struct GradKArgs {
  int mode, r, nh1, T, variable, bc, M, N, a0_variable, mix;
  int64_t Ns, n;
  double c, V, dt, hx, hy;
  double H_c[4], G_c[2], kappa_c;
  const double* kappa;
  const double* H;
  const double* G;
  const int64_t* indptr;
  const int32_t* indices;
  const uint8_t* code;
  const double* qval;
  const double* sq;
  const double* sqc;
  const double* trq;
  const double* trqc;
  const double* probes;
  const double* mu;
  const double* diag_add;
  const double* acan;
  const double* a0can;
  const double* qdiag;
  // closed-form log|Q| (mode 0 without Sigma_Q): selected inverses of the
  // spatial matrices B^T B and Q0 on the one-slice pattern
  int closed;
  const double* sgram;
  const double* sq0;
  const int32_t* pos25;
  const double* wts;  // one-slice weights in place of G (separable models)
  const double* adv_dir;  // direction of a scaled advection field (cov-advection models)
};

constexpr int kWin = 60;  // Gd[25] | Gu[9] | G00[25] | Gq
constexpr int kOut = 40;

constexpr int kPart = kWin + 3;  // window | dotGQ, dot00, tautr

// Pass 1a: per cell and time slice, the slice's share of the windows (thread
// per (cell, slice)); pass 1b sums the slices in order.
// The windows live in LDS (indexed by the entry's code; a register array
// indexed at run time would go to scratch memory): thread x of the block owns
// wl[q * 128 + x].  part is [slice][q][cell] so that pass 1b reads it
// coalesced.
__global__ __launch_bounds__(128) void grad_reduce_kernel(GradKArgs a, double* __restrict__ part) {
  __shared__ double wl[kWin * 128];
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double* w = wl + threadIdx.x;  // w[q * 128]
  for (int q = 0; q < kWin; ++q) w[q * 128] = 0.0;
  if (idx >= a.Ns * a.T) return;
  const int64_t i = idx % a.Ns;
  const int t = (int)(idx / a.Ns);
  double dotGQ = 0.0, dot00 = 0.0, tautr = 0.0;
  const double inh = a.nh1 > 0 ? 1.0 / a.nh1 : 0.0;
  {
    const int64_t r = i + t * a.Ns;
    for (int64_t e = a.indptr[r]; e < a.indptr[r + 1]; ++e) {
      const int64_t col = a.indices[e];
      const int cd = a.code[e];
      double Ge;
      double tr_c = 0.0;
      if (a.wts) {
        Ge = a.wts[e];
      } else if (a.mode == 0) {
        Ge = 0.5 * a.r * ((a.closed ? 0.0 : a.sq[e]) - a.sqc[e]);
        if (cd == 12) tr_c = a.sqc[e];
      } else {
        double s = 0.0;
        for (int p = 0; p < a.nh1; ++p) {
          const int64_t o = (int64_t)p * a.n;
          s += (a.trq[r + o] - a.trqc[r + o]) * a.probes[col + o] + (a.trq[col + o] - a.trqc[col + o]) * a.probes[r + o];
          if (cd == 12) tr_c += a.trqc[r + o] * a.probes[r + o];
        }
        Ge = 0.5 * a.r * inh * 0.5 * s;
        tr_c *= inh;
      }
      if (!a.wts) {
        double mm = 0.0;
        for (int c = 0; c < a.r; ++c) mm += a.mu[r + (int64_t)c * a.n] * a.mu[col + (int64_t)c * a.n];
        Ge -= 0.5 * mm;
      }
      const double qv = a.qval[e];
      dotGQ += Ge * qv;
      if (cd < 25) {
        if (t >= 1) w[cd * 128] += Ge;
        if (t == 0) {
          w[(34 + cd) * 128] += Ge;
          dot00 += Ge * ((cd == 12) ? qv - a.c * a.qdiag[i] : qv);
        }
        if (cd == 12) {
          if (t <= a.T - 2) w[59 * 128] += Ge;
          if (a.diag_add) tautr += a.diag_add[r] * tr_c;
        }
      } else if (cd >= 34 && t <= a.T - 2) {
        w[(25 + cd - 34) * 128] += Ge;
      }
    }
  }
  // part[(t * kPart + q) * Ns + i]
  double* p = part + (int64_t)t * kPart * a.Ns + i;
  for (int q = 0; q < kWin; ++q) p[(int64_t)q * a.Ns] = w[q * 128];
  p[(int64_t)kWin * a.Ns] = dotGQ;
  p[(int64_t)(kWin + 1) * a.Ns] = dot00;
  p[(int64_t)(kWin + 2) * a.Ns] = tautr;
}

// Thread per (cell i, entry q): the sum over time slices in slice order
// (deterministic); consecutive threads take consecutive cells of one entry.
__global__ void grad_sum_kernel(GradKArgs a, const double* __restrict__ part, double* __restrict__ win,
                                double* __restrict__ out) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= a.Ns * kPart) return;
  const int64_t i = g % a.Ns;
  const int q = (int)(g / a.Ns);
  double acc = 0.0;
  for (int t = 0; t < a.T; ++t) acc += part[((int64_t)t * kPart + q) * a.Ns + i];
  if (q < kWin)
    win[i * kWin + q] = acc;
  else
    out[i * kOut + 36 + (q - kWin)] = acc;
}

__device__ inline int widx(int dx, int dy) { return (dy + 2) * 5 + dx + 2; }

// Pass 2: per cell m, adjoints of A(m,:), q_m and A0(m,:), then through the
// cell's own stencils.
__global__ void grad_cell_kernel(GradKArgs a, const double* __restrict__ win, double* __restrict__ out) {
  const int64_t mlin = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (mlin >= a.Ns) return;
  const int m = (int)mlin;
  const int M = a.M, N = a.N, bc = a.bc;
  const int i = m % M, j = m / M;
  const double* Am = a.acan + 9 * (int64_t)m;
  const double* A0m = a.a0can + 9 * (int64_t)m;
  const double qm = a.qdiag[m];
  const double iV = 1.0 / a.V;
  double abar[9], a0bar[9];
  double qbar_quad = 0.0, qbar_u = 0.0;
  for (int o = 0; o < 9; ++o) {
    const int dox = o % 3 - 1, doy = o / 3 - 1;
    double bg = 0.0, ag0 = 0.0;
    for (int o2 = 0; o2 < 9; ++o2) {
      const int d2x = o2 % 3 - 1, d2y = o2 / 3 - 1;
      const int jc = nbr_cell(i, j, d2x, d2y, M, N, bc);
      if (jc < 0) continue;
      const int wv = widx(dox - d2x, doy - d2y);
      bg += Am[o2] * win[(int64_t)jc * kWin + wv];
      ag0 += A0m[o2] * win[(int64_t)jc * kWin + 34 + wv];
    }
    const double gu = win[mlin * kWin + 25 + o];
    abar[o] = 2.0 * a.c * iV * qm * (iV * bg - gu);
    a0bar[o] = 2.0 * iV * ag0;
    qbar_quad += iV * iV * bg * Am[o];
    qbar_u += iV * gu * Am[o];
  }
  double qbar = a.c * (qbar_quad + win[mlin * kWin + 59] - 2.0 * qbar_u);
  if (a.closed) {
    // 1/2 r d log|Q| with log|Q| = log|Q0| + (T-1) log|B^T B| + (T-1) sum log(q c):
    // d log|B^T B| / dA(m,k) = 2 (A^-T)(m,k) = 2 (A Sigma_B)(m,k) / V^2,
    // d log|Q0| / dA0(m,k) = 2 (A0 Sigma_0)(m,k) / V, d/dq_m = (T-1) / q_m.
    const double st = (double)a.r * (a.T - 1);
    for (int o = 0; o < 9; ++o) {
      const int dox = o % 3 - 1, doy = o / 3 - 1;
      if (nbr_cell(i, j, dox, doy, M, N, bc) < 0) continue;
      double rt = 0.0, r0 = 0.0;
      for (int o1 = 0; o1 < 9; ++o1) {
        const int d1x = o1 % 3 - 1, d1y = o1 / 3 - 1;
        const int j1 = nbr_cell(i, j, d1x, d1y, M, N, bc);
        if (j1 < 0) continue;
        const int e = a.pos25[(int64_t)j1 * 25 + (doy - d1y + 2) * 5 + (dox - d1x + 2)];
        if (e < 0) continue;
        rt += Am[o1] * a.sgram[e];
        r0 += A0m[o1] * a.sq0[e];
      }
      abar[o] += st * rt * iV * iV;
      a0bar[o] += a.r * r0 * iV;
    }
    qbar += 0.5 * st / qm;
  }
  double* o = out + mlin * kOut;
  for (int q = 0; q < 36; ++q) o[q] = 0.0;
  double canv[9];
  // initial-state operator A0 = V kappa0 I - A_H(H0)
  o[25] = a0bar[4] * a.V;
  for (int e = 0; e < 4; ++e) {
    double h[4] = {0.0, 0.0, 0.0, 0.0};
    h[e] = 1.0;
    Cell9 c9;
    diffusion_const(i, j, M, N, bc, a.hx, a.hy, h, c9);
    canon9(c9, i, j, M, N, canv);
    double s = 0.0;
    for (int q = 0; q < 9; ++q) s += a0bar[q] * canv[q];
    o[26 + e] = -s;
  }
  // kappa, diffusion and advection each constant or a per-cell field
  const bool kvar = a.variable && a.mix != 2, dvar = a.variable && a.mix != 2, gvar = a.variable && a.mix != 1;
  const double kap = kvar ? a.kappa[m] : a.kappa_c;
  o[0] = abar[4] * a.V * a.dt + qbar * 2.0 * a.V * kap;
  if (dvar) {
    double Hk[16];
    for (int e = 0; e < 16; ++e) {
      for (int q = 0; q < 16; ++q) Hk[q] = 0.0;
      Hk[e] = 1.0;
      Cell9 c9;
      diffusion_var(i, j, M, N, bc, a.hx, a.hy, Hk, c9);
      canon9(c9, i, j, M, N, canv);
      double s = 0.0, s0 = 0.0;
      for (int q = 0; q < 9; ++q) {
        s += abar[q] * canv[q];
        s0 += a0bar[q] * canv[q];
      }
      o[1 + e] = -a.dt * s;
      if (a.a0_variable) o[1 + e] -= s0;  // A0 = Dv Dk - A_H(H) from the same fields
    }
    if (a.a0_variable) o[0] += a0bar[4] * a.V;
  } else {
    for (int e = 0; e < 4; ++e) {
      double h[4] = {0.0, 0.0, 0.0, 0.0};
      h[e] = 1.0;
      Cell9 c9;
      diffusion_const(i, j, M, N, bc, a.hx, a.hy, h, c9);
      canon9(c9, i, j, M, N, canv);
      double s = 0.0;
      for (int q = 0; q < 9; ++q) s += abar[q] * canv[q];
      o[30 + e] = -a.dt * s;
    }
  }
  if (gvar) {
    const double* Gk = a.G + 4 * (int64_t)m;
    for (int diff = 1; diff <= 2; ++diff)
      for (int f = 0; f < 4; ++f) {
        double dG[4] = {0.0, 0.0, 0.0, 0.0};
        dG[f] = 1.0;
        Cell5 c5;
        advection_var(i, j, M, N, bc, a.hx, a.hy, Gk, dG, diff, c5);
        canon5(c5, i, j, M, N, canv, true);
        double s = 0.0;
        for (int q = 0; q < 9; ++q) s += abar[q] * canv[q];
        o[17 + 4 * (diff - 1) + f] = a.dt * s;
      }
  } else {
    for (int diff = 1; diff <= 2; ++diff) {
      Cell5 c5;
      advection_const(i, j, M, N, bc, a.hx, a.hy, a.G_c[0], a.G_c[1], diff, c5);
      canon5(c5, i, j, M, N, canv, false);
      double s = 0.0;
      for (int q = 0; q < 9; ++q) s += abar[q] * canv[q];
      o[33 + diff] = a.dt * s;
    }
  }
  double sdir = 0.0;
  if (a.adv_dir) {
    Cell5 c5;
    advection_var(i, j, M, N, bc, a.hx, a.hy, a.adv_dir + 4 * (int64_t)m, nullptr, 3, c5);
    canon5(c5, i, j, M, N, canv, true);
    for (int q = 0; q < 9; ++q) sdir += abar[q] * canv[q];
  }
  o[39] = a.dt * sdir;
}

}  // namespace spde

using namespace spde;

extern "C" int spde_grad_cells(spde_plan* pl, const spde_grad_args* g, double* out, void* stream) {
  if (!pl || pl->host_only || !g || !out || !g->qval || (!g->mu && !g->weights) || pl->lat.wide) {
    set_error("spde_grad_cells: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (g->weights && (pl->lat.T != 1 || g->mode != 0)) {
    set_error("spde_grad_cells: weights are taken on one-slice plans in mode 0");
    return SPDE_ERR_ARG;
  }
  const bool closed = g->mode == 0 && !g->sigma_q && !g->weights;
  if (g->mode == 0 && !g->sigma_qc && !g->weights) {
    set_error("spde_grad_cells: exact mode needs the selected inverse of Q_c");
    return SPDE_ERR_ARG;
  }
  if (closed) {
    const spde_plan* sp = g->spatial;
    if (!g->sigma_gram || !g->sigma_q0 || !sp || sp->host_only || sp->lat.T != 1 || sp->lat.M != pl->lat.M ||
        sp->lat.N != pl->lat.N || sp->lat.bc != pl->lat.bc || !sp->d_pos25) {
      set_error("spde_grad_cells: without sigma_q, the spatial plan and the selected inverses of B^T B and Q0 "
                "are needed");
      return SPDE_ERR_ARG;
    }
  }
  if (g->mode == 1 && (!g->trq || !g->trqc || !g->probes || g->nh1 < 1)) {
    set_error("spde_grad_cells: Hutchinson mode needs trq, trqc and probes");
    return SPDE_ERR_ARG;
  }
  const spde_coeffs& c = g->coeffs;
  if (c.variable && (c.mix < 0 || c.mix > 2 || (c.mix != 2 && (!c.kappa || !c.H)) || (c.mix != 1 && !c.G))) {
    set_error("spde_grad_cells: variable coefficients need their fields (kappa, H unless mix = 2; G unless mix = 1)");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  GradKArgs a;
  a.mode = g->mode;
  a.r = g->r;
  a.nh1 = g->nh1;
  a.T = pl->lat.T;
  a.variable = c.variable;
  a.a0_variable = c.a0_variable;
  a.mix = c.variable ? c.mix : 0;
  a.bc = pl->lat.bc;
  a.M = pl->lat.M;
  a.N = pl->lat.N;
  a.Ns = pl->lat.Ns();
  a.n = pl->sym.n;
  a.c = 1.0 / (c.dt * c.sigma);
  a.V = c.V;
  a.dt = c.dt;
  a.hx = c.hx;
  a.hy = c.hy;
  for (int q = 0; q < 4; ++q) a.H_c[q] = c.H_c[q];
  a.G_c[0] = c.G_c[0];
  a.G_c[1] = c.G_c[1];
  a.kappa_c = c.kappa_c;
  a.kappa = c.kappa;
  a.H = c.H;
  a.G = c.G;
  a.indptr = pl->d_indptr;
  a.indices = pl->d_indices;
  a.code = pl->d_code;
  a.qval = g->qval;
  a.sq = g->sigma_q;
  a.sqc = g->sigma_qc;
  a.trq = g->trq;
  a.trqc = g->trqc;
  a.probes = g->probes;
  a.mu = g->mu;
  a.diag_add = g->diag_add;
  a.acan = pl->d_acan;
  a.a0can = pl->d_a0can;
  a.qdiag = pl->d_qdiag;
  a.closed = closed ? 1 : 0;
  a.sgram = closed ? g->sigma_gram : nullptr;
  a.sq0 = closed ? g->sigma_q0 : nullptr;
  a.pos25 = closed ? g->spatial->d_pos25 : nullptr;
  a.wts = g->weights;
  a.adv_dir = g->adv_dir;
  double* win = pl->d_win;
  if (!pl->d_gpart) {
    SPDE_HIP(hipStreamSynchronize(s));
    SPDE_HIP(hipMalloc((void**)&pl->d_gpart, sizeof(double) * a.Ns * a.T * kPart));
  }
  const unsigned blocks = (unsigned)((a.Ns + 127) / 128);
  hipLaunchKernelGGL(grad_reduce_kernel, dim3((unsigned)((a.Ns * a.T + 127) / 128)), dim3(128), 0, s, a,
                     pl->d_gpart);
  hipLaunchKernelGGL(grad_sum_kernel, dim3((unsigned)((a.Ns * kPart + 255) / 256)), dim3(256), 0, s, a,
                     pl->d_gpart, win, out);
  hipLaunchKernelGGL(grad_cell_kernel, dim3(blocks), dim3(128), 0, s, a, win, out);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}
