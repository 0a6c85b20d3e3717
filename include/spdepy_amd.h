/*
 * spdepy_amd -- MI355X-native GMRF hot path of berild/spdepy.
 *
 * C ABI of libspdepy_amd.so.  Plain C types only; every pointer documented as
 * "device" is HBM memory owned by the caller (e.g. a torch tensor's
 * data_ptr()); "host" pointers are ordinary process memory.  Every call
 * returns SPDE_OK or an error code and sets a message readable with
 * spde_last_error().  Calls are ordered on the HIP stream passed as `stream`
 * (NULL = default stream).  A plan is not thread-safe: one stream per plan.
 *
 * Reference interfaces each entry point replaces (berild/spdepy, path:line):
 *   stencils  -> AH_new/AH_Row/AH_Col/AH_Val/AH_delete
 *                (src/spdepy/spdes/ccode/AcH_2D_b1.cpp:170-185, AH_2D_b1.cpp:154-169)
 *                Aw_new/Aw_Row/Aw_Col/Aw_Val/Aw_delete
 *                (Acw_2D_b1.cpp:117-132, Aw_2D_b1.cpp:136-151)
 *   assembly  -> the block formula of makeQ
 *                (spdes/advection_idiffusion2D.py:85-116,
 *                 spdes/var_advection_var_diffusion2D.py:83-116)
 *   factor    -> sksparse.cholmod.cholesky(Q) and Factor.logdet()
 *                (advection_idiffusion2D.py:116,173,180)
 *   solve     -> Factor.solve_A / solve_Lt / apply_Pt (model.py:79-80,125-126;
 *                advection_idiffusion2D.py:176,183-184)
 *   selinv    -> Model.qinv via R-INLA (model.py:89-118, rqinv.R:18-70) and the
 *                trace terms of logLike (advection_idiffusion2D.py:182-189)
 *   gradient  -> the dQ loop of logLike (advection_idiffusion2D.py:186-189,
 *                var_advection_var_diffusion2D.py:218-221)
 */
#ifndef SPDEPY_AMD_H
#define SPDEPY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPDE_OK 0
#define SPDE_ERR_ARG 1      /* invalid argument */
#define SPDE_ERR_HIP 2      /* HIP runtime failure */
#define SPDE_ERR_NOT_SPD 3  /* factorisation hit a non-positive pivot */
#define SPDE_ERR_ALLOC 4    /* device allocation failed */
#define SPDE_ERR_STATE 5    /* call out of order (e.g. solve before factor) */

const char* spde_last_error(void);

/* ---------------------------------------------------------------- stencils
 * Finite-volume diffusion stencil in the reference's slot layout: for each
 * cell k = i + j*M nine slots [C,E,W,N,S,NE,SW,NW,SE]; row[9k+s] == M*N marks
 * a dropped slot.  variable=0: H is the 2x2 tensor (4 doubles, row-major,
 * AcH_2D_b{1,3}.cpp semantics); variable=1: H is (M*N,4 faces W,E,S,N,2,2)
 * (AH_2D_b{1,2,3}.cpp semantics, including the bc=1 stale-index zeroing of
 * AH_2D_b1.cpp:25-52).  H, row, col, val: device. */
int spde_stencil_diffusion(int M, int N, const double* H, int variable, double hx, double hy,
                           int bc, int32_t* row, int32_t* col, double* val, void* stream);

/* Upwind advection stencil, five slots [C,E,W,N,S] per cell.  variable=0: G
 * holds (wx, wy) (Acw_2D_b{1,2,3}.cpp); variable=1: G and dG are (M*N, 4 faces
 * E,N,W,S) (Aw_2D_b{1,2,3}.cpp).  diff: 3 operator, 1 d/dwx, 2 d/dwy.  dG may
 * be NULL (zeros).  All pointers device. */
int spde_stencil_advection(int M, int N, const double* G, const double* dG, int variable,
                           double hx, double hy, int diff, int bc, int32_t* row, int32_t* col,
                           double* val, void* stream);

/* ---------------------------------------------------------------- plan
 * A plan holds the structural pattern of Q for an (M x N x T, bc) lattice, the
 * nested-dissection ordering, the multifrontal schedule and the device
 * arenas for `nslots` factorisations.  flags: SPDE_PLAN_HOST_ONLY builds the
 * symbolic part only (no device memory; inspection and tests). */
typedef struct spde_plan spde_plan;
#define SPDE_PLAN_HOST_ONLY 1
/* Couple neighbouring time slices through the 5x5 spatial window (the pattern
 * of Qt (x) Qs for the separable models) instead of the 3x3 window of the
 * advection-diffusion models.  Such a plan takes its values from
 * spde_assemble_kron only. */
#define SPDE_PLAN_WIDE_TIME 2

typedef struct spde_plan_info {
  int64_t n, nnz_struct, nnz_L, front_doubles, z_doubles;
  int64_t nnodes, nlevels, max_front, max_p;
  double flops_factor, flops_selinv;
  int64_t split_doubles;  /* per-slot split-K workspace (factor, selected inverse, last solve) */
} spde_plan_info;

int spde_plan_create(int M, int N, int T, int bc, int leaf_max, int nslots, int flags,
                     spde_plan** plan);
void spde_plan_destroy(spde_plan* plan);
int spde_plan_get_info(const spde_plan* plan, spde_plan_info* info);

/* Host copies of the structure and the symbolic analysis (for inspection):
 * indptr[n+1], indices[nnz], code[nnz]; perm[n]; nodes[nnodes*8] = (start,
 * p, n, parent, child0, child1, level, nB); foff[nnodes]; bnd[sum nB];
 * relpos[sum nB]; map[nnz] (front-arena offset of each structural entry, -1
 * for the upper triangle).  Any pointer may be NULL. */
int spde_plan_export(const spde_plan* plan, int64_t* indptr, int32_t* indices, uint8_t* code,
                     int32_t* perm, int32_t* nodes, int64_t* foff, int32_t* bnd,
                     int32_t* relpos, int64_t* map);

/* ---------------------------------------------------------------- assembly
 * Coefficients of the space-time operator.  variable=0 (stationary models):
 * kappa_c, H_c (2x2 row-major) and G_c (wx, wy) with the constant-stencil
 * semantics; variable=1 (B-spline fields): device arrays kappa[Ns],
 * H[Ns*16] (faces W,E,S,N) and G[Ns*4] (faces E,N,W,S).  The initial-state
 * block is the Whittle-Matern precision of A0 = V kappa0 I - A_H(H0)
 * (constant stencil).  sigma = exp(log sigma). */
typedef struct spde_coeffs {
  int variable;
  double hx, hy, dt, V, sigma;
  double kappa_c, H_c[4], G_c[2];
  const double* kappa;
  const double* H;
  const double* G;
  double kappa0, H0[4];
  /* 1: A0 = Dv kappa - A_H(H) from the variable fields kappa[Ns], H[Ns*16]
   * (variable Whittle-Matern models, var_whittle_matern2D.py:61-65) instead
   * of the constant kappa0, H0. */
  int a0_variable;
  /* With variable = 1: 0 every field per cell; 1 constant advection G_c with
   * the constant stencil (advection_var_*diffusion2D.py); 2 constant kappa_c
   * and H_c with the constant stencil (var_advection_*diffusion2D.py). */
  int mix;
} spde_coeffs;

/* Values of Q on the structural pattern (qval[nnz_struct], device). */
int spde_assemble(spde_plan* plan, const spde_coeffs* c, double* qval, void* stream);

/* out = F^T D^(1/2) z with Q = F^T D F the precision of the last
 * spde_assemble on the plan (same c): F has V^(-1/2) A0 on the first time
 * block, A / V on the others and -I below the diagonal, D = blockdiag(I,
 * diag(q) / (dt sigma), ...).  For z ~ N(0, I) the result has covariance Q:
 * with tau S^T (y + eps / sqrt(tau)) added (spde_add_rows) it is the
 * right-hand side whose solve with Q_c is a conditional sample of x | y
 * (in place of the reference's P^T L^-T z + mu, model.py:73-87, for
 * solvers without a factor).  z, out: device FP64, row-major n x K. */
int spde_fform_apply_t(spde_plan* plan, const spde_coeffs* c, const double* z, double* out, int K, void* stream);
/* out[rows[a] * K + r] += w[a * K + r] (device; rows distinct, in [0, n)). */
int spde_add_rows(int64_t n, int K, int64_t nrows, const int64_t* rows, const double* w, double* out, void* stream);

/* Direction of one entry of the reference's dQ list (makeQ(par, grad=True),
 * var_advection_var_diffusion2D.py:118-197, advection_idiffusion2D.py:118-163):
 * kind 0 perturbs the coefficient fields -- dkappa (per cell, device, or the
 * constant dkappa_c when NULL), dH (has_dH: per-cell faces W,E,S,N x 2 x 2,
 * device, or the constant dH_c), the advection derivative stencil adv_diff
 * (1 d/dwx, 2 d/dwy; with dG per-cell faces E,N,W,S for variable advection);
 * kind 1 is log sigma; kind 2 the initial-state operator (dkappa0, dH0). */
typedef struct spde_direction {
  int kind;
  const double* dkappa;
  double dkappa_c;
  const double* dH;
  double dH_c[4];
  int has_dH;
  const double* dG;
  int adv_diff;
  double dkappa0, dH0[4];
} spde_direction;
/* dqval[nnz_struct] (device) = dQ along the direction at the coefficients of
 * the last spde_assemble on the plan (the same `c`): the product rule through
 * the block formula, exact to rounding (one-slice plans: dQ0). */
int spde_assemble_tangent(spde_plan* plan, const spde_coeffs* c, const spde_direction* dir, double* dqval,
                          void* stream);

/* Spatial matrices of the closed-form log-determinant, on a one-slice plan
 * (spde_plan_create with T = 1, same M, N, bc): which = 0 writes
 * Q0 = A0^T A0 / V, which = 1 writes B^T B with B = A / V, both from the last
 * spde_assemble on the space-time plan `st` (out[nnz_struct of sp], device).
 * log|Q| = log|Q0| + (T-1) log|B^T B| + (T-1) sum_k log(q_k c), c = 1/(dt
 * sigma), q the diagonal of Qs: the reference factorises Q itself
 * (advection_idiffusion2D.py:173, var_advection_var_diffusion2D.py:205); the
 * two give the same value (Q = F^T D F, F block bidiagonal). */
int spde_assemble_spatial(spde_plan* st, spde_plan* sp, int which, double* out, void* stream);
/* *out (host) = sum_k log(q_k c) for the last spde_assemble on the plan. */
int spde_logq_sum(spde_plan* plan, double c, double* out, void* stream);

/* Drop exact zeros: the reference's scipy pattern.  indptr[n+1] (int64),
 * indices[nnz_struct], data[nnz_struct] device; *nnz (host) receives the kept
 * count. */
int spde_prune(spde_plan* plan, const double* qval, int64_t* indptr, int32_t* indices,
               double* data, int64_t* nnz, void* stream);

/* y = Q x for Q on the structural pattern (x, y device, n x nrhs, column-major). */
int spde_spmv(spde_plan* plan, const double* qval, const double* x, double* y, int nrhs,
              void* stream);

/* ---------------------------------------------------------------- factor
 * Cholesky factorisation of qval (+ diag_add on the diagonal when non-NULL,
 * both device) into factor slot `slot`; *logdet (host) = log|Q|.  Returns
 * SPDE_ERR_NOT_SPD (message names the failing lattice index) when a pivot is
 * not positive. */
int spde_factor(spde_plan* plan, int slot, const double* qval, const double* diag_add,
                double* logdet, void* stream);

/* The same factorisation enqueued without waiting (factorisations of
 * different slots may run concurrently on different streams);
 * spde_factor_finish waits for it and returns *logdet or SPDE_ERR_NOT_SPD.
 * Solves and selected inverses may be enqueued on the same stream before
 * finishing. */
int spde_factor_async(spde_plan* plan, int slot, const double* qval, const double* diag_add,
                      void* stream);
int spde_factor_finish(spde_plan* plan, int slot, double* logdet);

/* Solves with the factor in `slot` (P Q P^T = L L^T, P this library's
 * nested-dissection ordering).  mode: 0 Q^-1 B, 1 L^-1 P B, 2 P^T L^-T B
 * (the reference's apply_Pt(solve_Lt(.)) as one call, model.py:79-80),
 * 3 P B and 4 P^T B (CHOLMOD's apply_P / apply_Pt; these need no factor),
 * 5 L^-1 B and 6 L^-T B (CHOLMOD's solve_L / solve_Lt with
 * use_LDLt_decomposition=False: B and the result in elimination order).
 * B device, n x nrhs column-major with leading dimension ldb, overwritten. */
#define SPDE_SOLVE_A 0
#define SPDE_SOLVE_L 1
#define SPDE_SOLVE_LT 2
#define SPDE_SOLVE_P 3
#define SPDE_SOLVE_PT 4
#define SPDE_SOLVE_L_NOPERM 5
#define SPDE_SOLVE_LT_NOPERM 6
int spde_solve(spde_plan* plan, int slot, int mode, double* B, int nrhs, int64_t ldb,
               void* stream);

/* Takahashi selected inverse of the factor in `slot`: Sigma = Q^-1 on the
 * structural pattern (sigma[nnz_struct], device).  Each slot has its own
 * arenas, so selected inverses of different slots may run concurrently.  The
 * products that need only the factor (X = L21 L11^-1, L11^-T L11^-1) run on a
 * second stream of the plan, ordered after `stream` by an event; the level
 * chain on `stream` waits for them, so the call's results are ordered on
 * `stream` alone. */
int spde_selinv(spde_plan* plan, int slot, double* sigma, void* stream);

/* ---------------------------------------------------------------- gradient
 * Per-cell adjoints of the log-likelihood gradient (the contraction
 * g_i = <G, dQ_i> of the reference's dQ loop, done once through the adjoint of
 * the assembly).  G = 1/2 r (Sigma_Q - Sigma_Qc) - 1/2 mu mu^T on the pattern
 * (mode 0, sigma_q/sigma_qc from spde_selinv) or its Hutchinson estimate from
 * solves trq = Q^-1 V, trqc = Q_c^-1 V with the probe matrix V (mode 1, all
 * n x nh1 column-major, lattice order).  The last spde_assemble on this plan
 * must have used the same coefficients.  out[Ns*40] (device) receives per
 * cell: [0] d/dkappa, [1..16] d/dH (faces W,E,S,N, 2x2), [17..20] d/d(dG)
 * of the x-derivative stencil, [21..24] of the y-derivative stencil, [25]
 * d/dkappa0 / kappa0 ... see spdepy_amd/engine.py for the full layout. */
typedef struct spde_grad_args {
  int mode;
  int r;
  int nh1;
  const double* qval;
  const double* sigma_q;
  const double* sigma_qc;
  const double* trq;
  const double* trqc;
  const double* probes;
  const double* mu;
  const double* diag_add;
  spde_coeffs coeffs;
  /* Closed-form log|Q| (mode 0 with sigma_q == NULL): the Sigma_Q term of G
   * is replaced by the derivative of log|Q| = log|Q0| + (T-1) log|B^T B| +
   * (T-1) sum_k log(q_k c), from the selected inverses of the spatial
   * matrices (spde_assemble_spatial on the one-slice plan `spatial`). */
  const double* sigma_gram;
  const double* sigma_q0;
  spde_plan* spatial;
  /* One-slice plans only: when set, G is these per-entry weights (on the
   * plan's structural pattern, both triangles) and sigma_q, sigma_qc, mu,
   * trq, trqc, probes are not read; the outputs are the adjoints of
   * <weights, dQ> (the separable models' spatial gradient, spde_kron_reduce). */
  const double* weights;
  /* Optional advection direction adv_dir[Ns*4] (faces E,N,W,S): column 39 of
   * out receives dt <Abar, Aw(adv_dir)>, the derivative along a scaling of the
   * advection field taken with the stencil of adv_dir itself, as the
   * cov-advection models form it (cov_advection_idiffusion2D.py:59,153-159:
   * dA = Aw(ww) dt for advection lambda ww). */
  const double* adv_dir;
} spde_grad_args;

/* out[Ns * 40] (device, row-major per spatial cell m) holds d like_raw with
 * respect to each cell's coefficients, summed over the time slices:
 *   [0]      kappa_m (the cell's kappa, through A, q and, with a0_variable, A0)
 *   [1..16]  variable diffusion H of the cell, faces W,E,S,N x (2x2 row-major)
 *   [17..20] d/dwx of the variable advection per face E,N,W,S
 *   [21..24] d/dwy of the variable advection per face
 *   [25]     kappa0 of the initial-state operator A0
 *   [26..29] H0 (2x2 row-major) of A0
 *   [30..33] constant diffusion H_c (2x2 row-major; constant models)
 *   [34], [35] constant advection derivatives d/dwx, d/dwy
 *   [36]     the cell's share of <G, Q>  [37] of <G00, Q0>
 *   [38]     of tau sum_k cnt_k Sigma_c,kk (the trace of Q_c^-1 tau S^T S)
 *   [39]     dt <Abar, Aw(adv_dir)> (with adv_dir; else 0)
 * The model's gradient is the contraction of these columns with its bases
 * (spde_fields_contract for the B-spline fields), plus the sigma and tau
 * entries from the column sums [36..38]. */
int spde_grad_cells(spde_plan* plan, const spde_grad_args* args, double* out, void* stream);

/* ---------------------------------------------------------------- fused evaluation
 * One exact-trace log-likelihood + gradient evaluation in one call (the body of
 * logLike, advection_idiffusion2D.py:167-207, var_advection_var_diffusion2D.py:
 * 201-237), with log|Q| in closed form from the one-slice plan:
 *   like_raw = 1/2 r log|Q| + 1/2 m r log tau - 1/2 r log|Q_c| - 1/2 mu^T Q mu
 *              - tau/2 ||y - S mu||^2,   Q_c = Q + tau S^T S,  mu = Q_c^-1 tau S^T y.
 * plan: space-time plan with >= 2 slots (slot 1 is used); spatial: its one-slice
 * plan (spde_plan_create with T = 1, same M, N, bc) with >= 2 slots; coeffs as
 * for spde_assemble.  obs_cols[m] (host): lattice index of each observation
 * (repeats allowed); y[m * r] (host, column-major m x r).  cells[Ns * 40]
 * (device) receives the per-cell adjoints of spde_grad_cells (closed form), from
 * which the model's gradient is contracted with its bases; out (host) the
 * scalars.  Blocks until the evaluation is done; uses a side stream of the plan
 * for the selected inverse of Q_c. */
typedef struct spde_loglik_out {
  double like;      /* -like_raw / (m r), the value logLike returns */
  double like_raw;
  double logdet_q, logdet_qc;
  double g_sigma;   /* d like_raw / d log sigma */
  double g_tau;     /* d like_raw / d log tau */
} spde_loglik_out;
int spde_loglik_grad(spde_plan* plan, spde_plan* spatial, const spde_coeffs* coeffs, const int64_t* obs_cols,
                     const double* y, int64_t m, int r, double log_tau, double* cells, spde_loglik_out* out,
                     void* stream);

/* The same evaluation in two halves: spde_loglik_enqueue returns once the
 * work is enqueued (every scalar it needs is reduced on the device and copied
 * to pinned host memory); spde_loglik_finish waits for it and fills `out`
 * (SPDE_ERR_NOT_SPD names the lattice index of a failed pivot).  One
 * evaluation per plan may be in flight. */
int spde_loglik_enqueue(spde_plan* plan, spde_plan* spatial, const spde_coeffs* coeffs, const int64_t* obs_cols,
                        const double* y, int64_t m, int r, double log_tau, double* cells, void* stream);
int spde_loglik_finish(spde_plan* plan, spde_loglik_out* out);
/* mu_c of the last evaluation on the plan (device, n x r column-major). */
int spde_loglik_mu(spde_plan* plan, double* mu, int r, void* stream);
/* Many models in one call (the loop of the reference's kriging drivers,
 * examples/server/get_results.py:6-51, as one batch): model i is
 * (plans[i], spatials[i], coeffs[i], obs_cols[i], y[i], m[i], log_taus[i]),
 * its per-cell adjoints go to cells[i] and its scalars to outs[i].  Models are
 * grouped into waves whose plans are distinct; a wave's models are enqueued
 * on streams[i mod nstreams] before any is finished, so small models share
 * the chip.  status[i] (host) = each model's return code; the call returns
 * the first non-zero one. */
int spde_loglik_batch(int B, spde_plan* const* plans, spde_plan* const* spatials, const spde_coeffs* coeffs,
                      const int64_t* const* obs_cols, const double* const* y, const int64_t* m, int r,
                      const double* log_taus, double* const* cells, spde_loglik_out* outs,
                      void* const* streams, int nstreams, int32_t* status);

/* ---------------------------------------------------------------- separable models
 * Q = Qt (x) Qs (seperable_spatial_temporal2D.py:80-81, _ha2D.py, _idiffusion2D.py:
 * sparse.kron(Qt, Qs)) on a plan created with SPDE_PLAN_WIDE_TIME; `spatial` is
 * the one-slice plan of the same M, N, bc holding Qs (its structural values,
 * qs[nnz(spatial)], e.g. from spde_assemble + spde_assemble_spatial of the
 * variable Whittle-Matern coefficients).  qt[2T-1] (device) = the diagonal of
 * the tridiagonal Qt, then its off-diagonal. */
int spde_assemble_kron(spde_plan* plan, spde_plan* spatial, const double* qs, const double* qt,
                       double* out, void* stream);
/* The contractions that replace the reference's dQ loop for dQ = Qt (x) dQs and
 * dQ = dQt (x) Qs (seperable_spatial_temporal2D.py:84-118,143-146), with
 * G = 1/2 r (Q^-1 - Q_c^-1) - 1/2 mu_c mu_c^T and Q^-1 = Qt^-1 (x) Qs^-1:
 *   weights[s] = ws sigma_s[s] + sum_{t,t'} Qt[t,t'] (-1/2 r sigma_c - 1/2 sum_c mu mu^T)[(t,i),(t',j)]
 * for every structural entry s = (i,j) of the one-slice plan (ws = r T / 2 and
 * sigma_s the selected inverse of Qs give the log|Q| term; pass the weights to
 * spde_grad_cells on the one-slice plan), and per time pair (t, t+d), d = -1, 0, 1:
 *   blocks[(3t + d + 1) * 3 + 0] = sum_{i,j} sigma_c[(t,i),(t+d,j)] Qs[i,j]
 *   blocks[(3t + d + 1) * 3 + 1] = sum_{c,i,j} mu[(t,i),c] mu[(t+d,j),c] Qs[i,j]
 *   blocks[(3t + 1) * 3 + 2]     = sum_i diag_add[(t,i)] sigma_c[(t,i),(t,i)]
 * (zeros where t+d is outside [0,T)).  sigma_c is the selected inverse of Q_c
 * on this plan, mu[n x r] column-major, diag_add[n]; blocks[9T] (device). */
int spde_kron_reduce(spde_plan* plan, spde_plan* spatial, const double* qs, const double* qt,
                     const double* sigma_s, double ws, const double* sigma_c, const double* mu, int r,
                     const double* diag_add, double* weights, double* blocks, void* stream);

/* ---------------------------------------------------------------- B-spline fields
 * Coefficient fields of the variable model from its basis coefficients
 * (var_advection_var_diffusion2D.py:91-99, spat2Dtemp_regular_mesh.py:251-275),
 * and the chain rule back onto the coefficients (the contraction the
 * reference's per-parameter dQ loop performs, :131-183).  Bases dense
 * row-major on the device: bs[Ns][nb], bsH[Ns][4][nb] (faces W,E,S,N),
 * bsA[Ns][4][nb] (faces E,N,W,S).  coef[6 nb] = kappa | gamma | vx | vy | wx |
 * wy.  Outputs (device): kappa[Ns], H[Ns][4][2][2], G[Ns][4] (the layouts of
 * spde_coeffs), fgv[Ns][4][3] = (gamma, vx, vy) per face for the contraction. */
int spde_fields_eval(int64_t Ns, int nb, const double* bs, const double* bsH, const double* bsA,
                     const double* coef, double* kappa, double* H, double* G, double* fgv,
                     void* stream);
/* out[6 nb + 40] (device) = d/dcoef of <cells, d(kappa, H, G)/dcoef> in the
 * coef layout, then the column sums of cells[Ns][40] (spde_grad_cells).
 * work: device doubles, count from spde_fields_contract_workspace. */
int spde_fields_contract_workspace(int64_t Ns, int nb, int64_t* doubles);
int spde_fields_contract(int64_t Ns, int nb, const double* bs, const double* bsH, const double* bsA,
                         const double* cells, const double* kappa, const double* fgv, double* work,
                         double* out, void* stream);
/* No counterpart in the reference: spde_fields_contract for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops), with
 * no `work` argument.  Its workspace is a static device array of
 * (240 + 1) * 424 doubles (0.8 MB) that each device holds until the process
 * ends -- the largest the contraction can use, so every (Ns, nb) fits; a larger
 * need would return SPDE_ERR_ARG.  Calls allocate nothing.  All calls on one
 * device share the array: calls that overlap on one device race. */
int spde_fields_contract_noalloc(int64_t Ns, int nb, const double* bs, const double* bsH,
                                 const double* bsA, const double* cells, const double* kappa,
                                 const double* fgv, double* out, void* stream);

/* ---------------------------------------------------------------- iterative path
 * Mixed-precision solves with Q_c = Q + diag(diag_add) on a space-time plan
 * (the pattern of the advection-diffusion models), replacing the factorisation
 * of Q_c that the reference hands to CHOLMOD (advection_idiffusion2D.py:173-189,
 * var_advection_var_diffusion2D.py:205-221): IC(0) in an 18-colour ordering
 * ((i mod 3) + 3 (j mod 3) + 9 (t mod 2)), FP32 preconditioned conjugate
 * gradients, FP64 iterative refinement, Lanczos coefficients for stochastic
 * log-determinants, Rademacher probes and the Hutchinson estimate of Q_c^-1
 * on the pattern.  Vectors with k right-hand sides are row-major n x k floats
 * (k a power of two <= 64). */
typedef struct spde_ic spde_ic;
int spde_ic_create(spde_plan* plan, spde_ic** ic);
/* A second handle on the same pattern and factor with its own work buffers
 * (no counterpart in the reference: it lets two solves with one factor run
 * at once on different streams, e.g. the conditional mean beside the
 * Hutchinson probes).  It reads the parent's current factor at every call;
 * spde_ic_factor / spde_ic_correct on a view return SPDE_ERR_ARG, and the
 * parent must not be refactored while a view's solve is in flight. */
int spde_ic_view(spde_ic* parent, spde_ic** view);
/* Frees a handle; a parent destroyed before its views is freed with the last
 * of them (the views keep working). */
void spde_ic_destroy(spde_ic* ic);
/* nlow: entries of the strictly lower factor (one per off-diagonal pair). */
int spde_ic_info(const spde_ic* ic, int64_t* nlow, int32_t* ncolours);
/* Host copies (any may be NULL): colour[n]; lower lists lptr[n+1], lcol[nlow]
 * (row i: columns of lower colour, by (colour, index)); lch[nlow] = cell * 22
 * + channel of each lower entry in the correction layout (-1 none); upper
 * lists uptr[n+1], ucol[nlow]. */
int spde_ic_export(const spde_ic* ic, int32_t* colour, int64_t* lptr, int32_t* lcol, int64_t* lch,
                   int64_t* uptr, int32_t* ucol);
/* IC(0) of Q_c with its diagonal scaled by (1 + shift); *logdet_m (host) =
 * log|L L^T|.  qval, diag_add (may be NULL) device, kept by reference until the
 * next factorisation.  SPDE_ERR_NOT_SPD names the lattice index of a
 * non-positive pivot. */
int spde_ic_factor(spde_ic* ic, const double* qval, const double* diag_add, double shift, double* logdet_m,
                   void* stream);
/* Correction of the factor in the Jacobi-scaled space (D = diag(Q_c)):
 * L~ = D^-1/2 L; L~_ij += corr[cell * 22 + ch] for the pair channel of each
 * lower entry, L~_ii *= exp(corr[i * 22]).  corr (device float, n * 22) is
 * the learned preconditioner's output; call once after spde_ic_factor
 * (NULL: refresh the FP32 copies only).  *logdet_m = log|L L^T|. */
int spde_ic_correct(spde_ic* ic, const float* corr, double* logdet_m, void* stream);
/* Device copies of the FP64 factor: lower[nlow], diag[n] (L_ii), adiag[n]
 * (the diagonal of Q_c); any may be NULL. */
int spde_ic_values(spde_ic* ic, double* lower, double* diag, double* adiag, void* stream);
/* z = (L L^T)^-1 r (FP32); which = 1: out = L^-1 in, which = 2: out = L^-T in. */
int spde_ic_apply(spde_ic* ic, const float* r, float* z, int k, void* stream);
int spde_ic_sweep(spde_ic* ic, int which, const float* in, float* out, int k, void* stream);
/* y = Q_c x in FP32. */
int spde_ic_spmm(spde_ic* ic, const float* x, float* y, int k, void* stream);
/* FP32 PCG for Q_c X = B from X = 0, each right-hand side until
 * ||r|| <= tol ||b|| or maxit iterations; iters[k] (host) per right-hand side. */
int spde_pcg_solve(spde_ic* ic, const float* B, float* X, int k, double tol, int maxit, int32_t* iters,
                   void* stream);
/* spde_pcg_solve that also returns each right-hand side's first cap step
 * coefficients (host, alpha[r * cap + j], beta[r * cap + j], zero past its
 * iterations): with M = L L^T they give the Lanczos tridiagonal of
 * L^-1 Q_c L^-T started at L^-1 b / ||L^-1 b|| (diagonal 1/alpha_1,
 * 1/alpha_j + beta_(j-1)/alpha_(j-1); off-diagonal sqrt(beta_j)/alpha_j), so
 * the solves of b = L w for Rademacher w give stochastic Lanczos quadrature
 * of log|Q_c| - log|M| on the probes w at no extra cost. */
int spde_pcg_solve_coeffs(spde_ic* ic, const float* B, float* X, int k, double tol, int maxit, int32_t* iters,
                          double* alpha, double* beta, int cap, void* stream);
/* Y = L W for the current FP32 factor (k right-hand sides, row-major n x k). */
int spde_ic_lmul(spde_ic* ic, const float* W, float* Y, int k, void* stream);
/* FP64 x = Q_c^-1 b (one right-hand side, device) by iterative refinement:
 * r = b - Q_c x in FP64, an FP32 PCG solve of Q_c d = r / ||r|| to inner_tol,
 * x += ||r|| d, until ||r|| <= tol ||b|| or max_rounds; *iters total inner
 * iterations, *rounds, *relres = ||b - Q_c x|| / ||b|| (host). */
int spde_pcg_refine(spde_ic* ic, const double* b, double* x, double tol, double inner_tol, int max_rounds,
                    int maxit, int32_t* iters, int32_t* rounds, double* relres, void* stream);
/* FP64 x = Q_c^-1 b (one right-hand side, device) by FP32 PCG with FP64
 * residual replacement (van der Vorst and Ye, SIAM J. Sci. Comput. 22, 2000):
 * iterative refinement without restarting the recurrence.  Whenever the
 * FP32 recursive residual has fallen by delta (checked every 16 iterations)
 * or 512 iterations have passed, the FP32 iterate is added into x,
 * r = b - Q_c x is formed in FP64 and replaces the recurrence's residual
 * (search direction kept), until ||r|| <= tol ||b||, maxit iterations, or
 * three replacements without a 10% fall.  A segment that leaves the true
 * residual larger is taken back; after it, a breakdown or a stall the
 * recurrence restarts from r.
 * *iters, *updates (replacements), *relres = ||b - Q_c x|| / ||b|| (host).
 * Replaces the refinement rounds' restarts of spde_pcg_refine, whose inner
 * solves each start a new Krylov space. */
int spde_pcg_replace(spde_ic* ic, const double* b, double* x, double tol, double delta, int maxit,
                     int32_t* iters, int32_t* updates, double* relres, void* stream);
/* X = Q_c^-1 B for k right-hand sides (k a power of two <= 64; B, X device
 * FP64, row-major n x k) by the FP32 block PCG with FP64 residual
 * replacement (as spde_pcg_replace): every column is replaced together once
 * every live column's recursive residual has fallen by delta, or after 512
 * iterations; per column a worse segment is taken back and the column
 * restarted, a column within tol (or stalled three times) stops.  iters[k],
 * relres[k] (host) per column; *updates (host). */
int spde_pcg_replace_block(spde_ic* ic, const double* B, double* X, int k, double tol, double delta, int maxit,
                           int32_t* iters, int32_t* updates, double* relres, void* stream);
/* X = Q_c^-1 B for k right-hand sides (k a power of two <= 64; B, X device
 * FP64, row-major n x k, entry i * k + r) by iterative refinement around the
 * FP32 block PCG: per column r = b - Q_c x in FP64, an inner solve of
 * Q_c d = r / ||r|| to inner_tol, x += ||r|| d, until ||r|| <= tol ||b|| for
 * every column or max_rounds; a column within tol takes no further inner
 * iterations.  iters[k], relres[k] (host): inner iterations and
 * ||b - Q_c x|| / ||b|| per column; *rounds (host). */
int spde_pcg_refine_block(spde_ic* ic, const double* B, double* X, int k, double tol, double inner_tol,
                          int max_rounds, int maxit, int32_t* iters, int32_t* rounds, double* relres, void* stream);
/* Z[u] ~ N(0, 1) for u < n k (device FP64): Box-Muller on two 53-bit uniforms,
 * U1 = ((h1 >> 11) + 0.5) 2^-53, U2 = (h2 >> 11) 2^-53 with
 * h1 = splitmix64(seed ^ splitmix64(2u)), h2 = splitmix64(seed ^ splitmix64(2u + 1)),
 * Z[u] = sqrt(-2 log U1) cos(2 pi U2). */
int spde_gaussian(int64_t n, int k, uint64_t seed, double* Z, void* stream);
/* V[u] = +-scale for u = i * k + r: the sign is the top bit of
 * splitmix64(seed ^ splitmix64(u)) (1: negative). */
int spde_rademacher(int64_t n, int k, uint64_t seed, float scale, float* V, void* stream);
/* Lanczos on L^-1 Q_c L^-T from k Rademacher probes (spde_rademacher with
 * seed, normalised): alpha[r * steps + j], beta[r * steps + j] (host) for
 * probe r; FP32 vectors, FP64 coefficients. */
int spde_slq(spde_ic* ic, int k, int steps, uint64_t seed, double* alpha, double* beta, void* stream);
/* sigma[e] = (1/k) sum_r (U[i][r] V[j][r] + U[j][r] V[i][r]) / 2 for every
 * structural entry e = (i, j) (device; U = Q_c^-1 V gives the Hutchinson
 * estimate of Q_c^-1 on the pattern), written as sigma = beta sigma + alpha
 * (estimate) so that batches of probes accumulate. */
int spde_hutch_pattern(spde_ic* ic, const float* U, const float* V, int k, double alpha, double beta,
                       double* sigma, void* stream);

/* ---------------------------------------------------------------- learned preconditioner
 * No reference counterpart: a correction of the IC(0) factor predicted by a
 * 4-layer 3x3 convolutional network over the time slices (8 input channels,
 * 32 hidden, 32 output of which 22 are used), trained on the GPU by the
 * gradient of KL(N(0, Q_c^-1) || N(0, (L' L'^T)^-1)) over samples of the
 * exact Cholesky path.  Activations pixel-major (pixel = lattice index, the
 * channels contiguous, FP32); layer weights W[k][co] FP32 with k = tap * cin +
 * ci, tap = 3 (dy + 1) + (dx + 1), packed as BF16 matrix-core fragments.
 *
 * feat[n * 8] (device float): per lattice entry [tau count / A_ii, A~ to the
 * E, W, N, S neighbours, A~ to the same cell in the next and previous slice,
 * 0] with A = Q_c, A~_ij = A_ij / sqrt(A_ii A_jj), from the last
 * spde_ic_factor. */
int spde_ic_features(spde_ic* ic, float* feat, void* stream);
/* dout[n * 22] (device float) = scale * d KL / d corr at the current
 * (corrected) factor from k exact samples x (n x k row-major FP64, lattice
 * order, x ~ N(0, Q_c^-1)); *loss (host, may be NULL) = 1/2 mean ||L'^T x||^2
 * - sum log L'_ii. */
int spde_ic_kl_grad(spde_ic* ic, const double* x, int k, double scale, float* dout, double* loss, void* stream);
/* The training loss: dout[n * 22] = scale * dJ / d corr of the symmetrised
 * divergence J = KL(p||q) + KL(q||p) = 1/2 E||L'^T x||^2 + 1/2 E[y^T Q_c y] - n,
 * p = N(0, Q_c^-1), q = N(0, (L' L'^T)^-1), over k exact samples x (n x k
 * row-major FP64) and y = L'^-T z for k probes z (n x k row-major float);
 * *loss (host, may be NULL) = J.  Both ends of the spectrum of L'^-1 Q_c L'^-T
 * are penalised (sum of lambda + 1/lambda - 2). */
int spde_ic_jeffreys_grad(spde_ic* ic, const double* x, const float* z, int k, double scale, float* dout,
                          double* loss, void* stream);
/* Packed BF16 fragment size (in halves) of a layer's operand; transposed = 1
 * for the backward-data operand (flipped taps, channels swapped). */
int spde_conv_packed_size(int cin, int transposed, int64_t* halves);
int spde_conv_pack(int cin, int transposed, const float* W, unsigned short* frag, void* stream);
/* out[P x 32] = act(conv3x3(in * (mask > 0)) + bias) over the T slices of an
 * M x N lattice (zero padding); cin 8 or 32 input channels; mask, bias may
 * be NULL; relu 0/1.  With transposed fragments and mask = the layer's
 * output this is the backward-data pass through the layer. */
int spde_conv_forward(int M, int N, int T, int cin, const float* in, const float* mask,
                      const unsigned short* frag, const float* bias, int relu, float* out, void* stream);
/* dW[9 cin x 32] and db[32] of a layer from its input X (P x cin) and the
 * gradient G (P x 32) of its output, masked by Y > 0 (the ReLU) when Y is not
 * NULL; operands rounded to BF16 as in the forward pass.  work: device
 * floats, count from spde_conv_wgrad_workspace. */
int spde_conv_wgrad_workspace(int64_t P, int cin, int64_t* floats);
int spde_conv_backward(int M, int N, int T, int cin, const float* X, const float* G, const float* Y, float* dW,
                       float* db, float* work, void* stream);
/* Adam step (step >= 1) on n FP32 parameters; mask (device bytes, may be NULL):
 * entries with 0 stay fixed. */
int spde_adam(int64_t n, const float* g, const unsigned char* mask, float lr, float b1, float b2, float eps,
              int step, float* theta, float* m, float* v, void* stream);

/* ---------------------------------------------------------------- measurement
 * Between spde_prof_begin and spde_prof_end every FP64 GEMM launch of the
 * plan (factorisation, inverse of L11, solves, selected inverse) is bracketed
 * by HIP events on its stream; spde_prof_end synchronises and returns the
 * summed per-launch time (ms), the summed algorithmic flops, the launch count
 * the busy time (ms during which at least one GEMM launch was in flight;
 * launches on different streams overlap, so busy <= summed time) and the
 * algorithmic bytes (A, B and C read once, C written once). */
int spde_prof_begin(spde_plan* plan);
/* The same for the GEMM launches of the given kinds only (bit k of kind_mask
 * for kind k, the code spde_prof_dump reports), so that timing one kernel
 * adds no events around the others (no reference counterpart). */
int spde_prof_begin_kinds(spde_plan* plan, int kind_mask);
int spde_prof_end(spde_plan* plan, double* gemm_ms, double* gemm_flops, int64_t* launches,
                  double* busy_ms, double* gemm_bytes);
/* Per-launch view of the timed launches so far (host arrays of length cap,
 * each may be NULL): kernel ms, algorithmic flops, algorithmic bytes, kernel
 * kind (4 for 128-wide tiles, + 2 when A is transposed, + 1 when B is:
 * gemm128_kernel<false, false> is kind 4) and tile count. */
int spde_prof_dump(spde_plan* plan, double* ms, double* flops, double* bytes, int32_t* kind,
                   int32_t* tiles, int64_t cap, int64_t* count);

/* ---------------------------------------------------------------- diagnostics
 * No counterpart in the reference: self-checks of the dense kernels.
 * spde_debug_mfma_probe writes the 64x4 accumulator lanes of one FP64 MFMA
 * for a fixed input pattern (out: 256 doubles, device).  spde_debug_gemm runs
 * the library's batched GEMM kernel on one problem (column-major, device). */
int spde_debug_mfma_probe(int which, double* out, void* stream);
int spde_debug_gemm(int ta, int tb, int M, int N, int K, double alpha, const double* A, int lda,
                    const double* B, int ldb, double beta, double* C, int ldc, int tile, void* stream);
/* The 64 x 4 accumulator lanes of one BF16 16x16x16 matrix instruction with
 * A[i][k] = 16 i + k and B = I (out: 256 floats, device). */
int spde_debug_bf16_probe(float* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
