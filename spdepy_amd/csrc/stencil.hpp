// Per-cell finite-volume stencil functions shared by assembly and gradient.
// They reproduce berild/spdepy src/spdepy/spdes/ccode/{AcH,AH,Acw,Aw}_2D_b{1,2,3}.cpp
// operation for operation (floating-point contraction is disabled by the
// including translation units).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace spde {

struct Cell9 {
  int32_t col[9];
  double val[9];
  bool keep[9];
};
struct Cell5 {
  int32_t col[5];
  double val[5];
  bool keep[5];
};

__device__ inline void nbr_idx(int i, int j, int M, int N, int bc, int& iN, int& iP, int& jN, int& jP) {
  if (bc == 1) {
    iN = (i == 0) ? i : i - 1;
    iP = (i == M - 1 && i != 0) ? i : i + 1;
    jN = (j == 0) ? j : j - 1;
    jP = (j == N - 1 && j != 0) ? j : j + 1;
  } else if (bc == 2) {
    iN = (i == 0) ? M - 1 : i - 1;
    iP = (i == M - 1 && i != 0) ? 0 : i + 1;
    jN = (j == 0) ? N - 1 : j - 1;
    jP = (j == N - 1 && j != 0) ? 0 : j + 1;
  } else {
    iN = i - 1;
    iP = i + 1;
    jN = j - 1;
    jP = j + 1;
  }
}

// 9-point diffusion slots from the eight face coefficients of one cell.
__device__ void diffusion_cell(int i, int j, int M, int N, int bc, double hx, double hy, double W00, double E00,
                               double W10, double E10, double S01, double N01, double S11, double N11, Cell9& c) {
  int iN, iP, jN, jP;
  nbr_idx(i, j, M, N, bc, iN, iP, jN, jP);
  const int k = i + j * M;
  c.col[0] = k;
  c.col[1] = j * M + iP;
  c.col[2] = j * M + iN;
  c.col[3] = jP * M + i;
  c.col[4] = jN * M + i;
  c.col[5] = jP * M + iP;
  c.col[6] = jN * M + iN;
  c.col[7] = jP * M + iN;
  c.col[8] = jN * M + iP;
  const double rx = hy / hx, ry = hx / hy;
  const double aE = rx * E00, aW = rx * W00, aN = ry * N11, aS = ry * S11;
  const double bx = 0.25 * (N01 - S01), by = 0.25 * (E10 - W10);
  const double cNE = 0.25 * (N01 + E10), cSW = 0.25 * (S01 + W10);
  const double cNW = 0.25 * (N01 + W10), cSE = 0.25 * (S01 + E10);
  c.val[1] = aE + bx;
  c.val[2] = aW - bx;
  c.val[3] = aN + by;
  c.val[4] = aS - by;
  c.val[5] = cNE;
  c.val[6] = cSW;
  c.val[7] = -0.25 * (N01 + W10);
  c.val[8] = -0.25 * (S01 + E10);
  double cc = (-hy / hx) * (E00 + W00) - ry * (N11 + S11);
  for (int s = 0; s < 9; ++s) c.keep[s] = true;
  if (bc == 1) {
    double rem = 0.0;
    if (c.col[1] == k) { rem = rem + aE + bx; c.keep[1] = false; }
    if (c.col[2] == k) { rem = rem + aW - bx; c.keep[2] = false; }
    if (c.col[3] == k) { rem = rem + aN + by; c.keep[3] = false; }
    if (c.col[4] == k) { rem = rem + aS - by; c.keep[4] = false; }
    if (c.col[5] == k) { rem = rem + cNE; c.keep[5] = false; }
    if (c.col[6] == k) { rem = rem + cSW; c.keep[6] = false; }
    if (c.col[7] == k) { rem = rem - cNW; c.keep[7] = false; }
    if (c.col[8] == k) { rem = rem - cSE; c.keep[8] = false; }
    cc = cc + rem;
  } else {
    cc = cc + 0.0;
    if (bc == 3) {
      if (i == 0) c.keep[2] = c.keep[6] = c.keep[7] = false;
      else if (i == M - 1) c.keep[1] = c.keep[5] = c.keep[8] = false;
      if (j == 0) c.keep[4] = c.keep[6] = c.keep[8] = false;
      else if (j == N - 1) c.keep[3] = c.keep[5] = c.keep[7] = false;
    }
  }
  c.val[0] = cc;
}

// Variable H (faces W,E,S,N of cell k, 2x2 row-major each), AH_2D_b*.cpp.
// bc=1: only cell 0 has its W and S face coefficients zeroed (the stale
// index of AH_2D_b1.cpp:25-52).
__device__ void diffusion_var(int i, int j, int M, int N, int bc, double hx, double hy, const double* __restrict__ Hk,
                              Cell9& c) {
  double W00 = Hk[0], W10 = Hk[2], E00 = Hk[4], E10 = Hk[6];
  double S01 = Hk[9], S11 = Hk[11], N01 = Hk[13], N11 = Hk[15];
  if (bc == 1 && i == 0 && j == 0) {
    W00 = 0.0;
    W10 = 0.0;
    S11 = 0.0;
    S01 = 0.0;
  }
  diffusion_cell(i, j, M, N, bc, hx, hy, W00, E00, W10, E10, S01, N01, S11, N11, c);
}

// Constant H (AcH_2D_b{1,3}.cpp; bc=2 as the periodic variable stencil with
// every face equal to H).
__device__ void diffusion_const(int i, int j, int M, int N, int bc, double hx, double hy, const double H[4], Cell9& c) {
  if (bc == 3) {
    const int k = i + j * M;
    c.col[0] = k;
    c.col[1] = j * M + i + 1;
    c.col[2] = j * M + i - 1;
    c.col[3] = (j + 1) * M + i;
    c.col[4] = (j - 1) * M + i;
    c.col[5] = (j + 1) * M + i + 1;
    c.col[6] = (j - 1) * M + i - 1;
    c.col[7] = (j + 1) * M + i - 1;
    c.col[8] = (j - 1) * M + i + 1;
    const double e = hy / hx * H[0], nn = hx / hy * H[3], cr = 0.25 * (H[1] + H[2]);
    c.val[0] = (-2.0 * hy / hx) * H[0] - (2.0 * hx / hy) * H[3] + 0.0;
    c.val[1] = e;
    c.val[2] = e;
    c.val[3] = nn;
    c.val[4] = nn;
    c.val[5] = cr;
    c.val[6] = cr;
    c.val[7] = -0.25 * (H[1] + H[2]);
    c.val[8] = -0.25 * (H[1] + H[2]);
    for (int s = 0; s < 9; ++s) c.keep[s] = true;
    if (i == 0) c.keep[2] = c.keep[6] = c.keep[7] = false;
    else if (i == M - 1) c.keep[1] = c.keep[5] = c.keep[8] = false;
    if (j == 0) c.keep[4] = c.keep[6] = c.keep[8] = false;
    else if (j == N - 1) c.keep[3] = c.keep[5] = c.keep[7] = false;
    return;
  }
  if (bc == 2) {
    diffusion_cell(i, j, M, N, 2, hx, hy, H[0], H[0], H[2], H[2], H[1], H[1], H[3], H[3], c);
    return;
  }
  double W00 = H[0], E00 = H[0], W10 = H[2], E10 = H[2], S01 = H[2], N01 = H[2], S11 = H[3], N11 = H[3];
  if (i == 0) {
    W00 = 0.0;
    W10 = 0.0;
  } else if (i == M - 1) {
    E00 = 0.0;
    E10 = 0.0;
  }
  if (j == 0) {
    S11 = 0.0;
    S01 = 0.0;
  } else if (j == N - 1) {
    N11 = 0.0;
    N01 = 0.0;
  }
  diffusion_cell(i, j, M, N, 1, hx, hy, W00, E00, W10, E10, S01, N01, S11, N11, c);
}

__device__ inline void adv_cols(int i, int j, int M, int iN, int iP, int jN, int jP, Cell5& c) {
  const int k = i + j * M;
  c.col[0] = k;
  c.col[1] = j * M + iP;
  c.col[2] = j * M + iN;
  c.col[3] = jP * M + i;
  c.col[4] = jN * M + i;
}

// Constant velocity (Acw_2D_b*.cpp).
__device__ void advection_const(int i, int j, int M, int N, int bc, double hx, double hy, double G0, double G1,
                                int diff, Cell5& c) {
  int iN, iP, jN, jP;
  nbr_idx(i, j, M, N, bc, iN, iP, jN, jP);
  adv_cols(i, j, M, iN, iP, jN, jP, c);
  const double s0 = G0 / fabs(G0), s1 = G1 / fabs(G1);
  if (diff == 1) {
    c.val[0] = s0 * hy;
    c.val[1] = -(s0 - 1.0) * hy / 2;
    c.val[2] = -(s0 + 1.0) * hy / 2;
    c.val[3] = 0.0;
    c.val[4] = 0.0;
  } else if (diff == 2) {
    c.val[0] = s1 * hx;
    c.val[1] = 0.0;
    c.val[2] = 0.0;
    c.val[3] = -(s1 - 1.0) * hx / 2;
    c.val[4] = -(s1 + 1.0) * hx / 2;
  } else {
    c.val[0] = fabs(G0) * hy + fabs(G1) * hx;
    c.val[1] = -(fabs(G0) - G0) * hy / 2;
    c.val[2] = -(fabs(G0) + G0) * hy / 2;
    c.val[3] = -(fabs(G1) - G1) * hx / 2;
    c.val[4] = -(fabs(G1) + G1) * hx / 2;
  }
  for (int s = 0; s < 5; ++s) c.keep[s] = true;
  if (bc == 1) {
    if (i == 0) {
      c.val[0] = c.val[0] - fabs(G0) * hy / 2;
      c.keep[2] = false;
    } else if (i == M - 1) {
      c.val[0] = c.val[0] - fabs(G0) * hy / 2;
      c.keep[1] = false;
    }
    if (j == 0) {
      c.val[0] = c.val[0] - fabs(G1) * hx / 2;
      c.keep[4] = false;
    } else if (j == N - 1) {
      c.val[0] = c.val[0] - fabs(G1) * hx / 2;
      c.keep[3] = false;
    }
  } else {
    c.val[0] = c.val[0] + 0.0;
    if (bc == 3) {
      if (i == M - 1) c.keep[1] = false;
      if (i == 0) c.keep[2] = false;
      if (j == N - 1) c.keep[3] = false;
      if (j == 0) c.keep[4] = false;
    }
  }
}

// Variable velocity (Aw_2D_b*.cpp), G/dG faces E,N,W,S of cell k.
__device__ void advection_var(int i, int j, int M, int N, int bc, double hx, double hy, const double* __restrict__ Gk,
                              const double* __restrict__ dGk, int diff, Cell5& c) {
  int iN, iP, jN, jP;
  nbr_idx(i, j, M, N, bc, iN, iP, jN, jP);
  adv_cols(i, j, M, iN, iP, jN, jP, c);
  double g0 = Gk[0], g1 = Gk[1], g2 = Gk[2], g3 = Gk[3];
  double d0 = dGk ? dGk[0] : 0.0, d1 = dGk ? dGk[1] : 0.0, d2 = dGk ? dGk[2] : 0.0, d3 = dGk ? dGk[3] : 0.0;
  for (int s = 0; s < 5; ++s) c.keep[s] = true;
  if (bc == 1) {
    if (iP == i) { g0 = 0.0; d0 = 0.0; c.keep[1] = false; }
    if (iN == i) { g2 = 0.0; d2 = 0.0; c.keep[2] = false; }
    if (jP == j) { g1 = 0.0; d1 = 0.0; c.keep[3] = false; }
    if (jN == j) { g3 = 0.0; d3 = 0.0; c.keep[4] = false; }
  } else if (bc == 3) {
    if (i == M - 1) c.keep[1] = false;
    if (i == 0) c.keep[2] = false;
    if (j == N - 1) c.keep[3] = false;
    if (j == 0) c.keep[4] = false;
  }
  if (diff == 1) {
    const double u0 = g0 / fabs(g0) * d0, u2 = g2 / fabs(g2) * d2;
    c.val[0] = (u0 + d0 + u2 - d2) * hy / 2;
    c.val[1] = -(u0 - d0) * hy / 2;
    c.val[2] = -(u2 + d2) * hy / 2;
    c.val[3] = 0.0;
    c.val[4] = 0.0;
  } else if (diff == 2) {
    const double u1 = g1 / fabs(g1) * d1, u3 = g3 / fabs(g3) * d3;
    c.val[0] = (u1 + d1 + u3 - d3) * hx / 2;
    c.val[1] = 0.0;
    c.val[2] = 0.0;
    c.val[3] = -(u1 - d1) * hx / 2;
    c.val[4] = -(u3 + d3) * hx / 2;
  } else {
    c.val[0] = (fabs(g0) + g0 + fabs(g2) - g2) * hy / 2 + (fabs(g1) + g1 + fabs(g3) - g3) * hx / 2;
    c.val[1] = -(fabs(g0) - g0) * hy / 2;
    c.val[2] = -(fabs(g2) + g2) * hy / 2;
    c.val[3] = -(fabs(g1) - g1) * hx / 2;
    c.val[4] = -(fabs(g3) + g3) * hx / 2;
  }
}


// Offset index o = (dy+1)*3 + (dx+1) of column `c` relative to cell (i,j).
__device__ inline int canon(int c, int i, int j, int M, int N) {
  int dx = c % M - i, dy = c / M - j;
  if (dx > 1) dx -= M;
  if (dx < -1) dx += M;
  if (dy > 1) dy -= N;
  if (dy < -1) dy += N;
  return (dy + 1) * 3 + dx + 1;
}

__device__ inline int nbr_cell(int i, int j, int dx, int dy, int M, int N, int bc) {
  int a = i + dx, b = j + dy;
  if (bc == 2) {
    a = (a + M) % M;
    b = (b + N) % N;
  } else if (a < 0 || a >= M || b < 0 || b >= N) {
    return -1;
  }
  return a + b * M;
}

// Kept slots of a 9-point stencil merged onto the canonical 3x3 offsets.
__device__ inline void canon9(const Cell9& c, int i, int j, int M, int N, double out[9]) {
  for (int o = 0; o < 9; ++o) out[o] = 0.0;
  for (int s = 0; s < 9; ++s)
    if (c.keep[s]) {
      const int o = canon(c.col[s], i, j, M, N);
      out[o] = out[o] + c.val[s];
    }
}
__device__ inline void canon5(const Cell5& c, int i, int j, int M, int N, double out[9], bool nan_to_zero) {
  for (int o = 0; o < 9; ++o) out[o] = 0.0;
  for (int s = 0; s < 5; ++s)
    if (c.keep[s]) {
      const int o = canon(c.col[s], i, j, M, N);
      double v = c.val[s];
      if (nan_to_zero && v != v) v = 0.0;
      out[o] = out[o] + v;
    }
}

}  // namespace spde
