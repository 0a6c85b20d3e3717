// One exact log-likelihood + gradient evaluation behind a single C call:
// assembly, factorisation of Q_c (main stream), its selected inverse (side
// stream), the solve for mu_c and the one-slice log|Q| work (main stream,
// concurrent with the selected inverse), then the adjoint pass.  The call is
// split into an enqueue half, which leaves every scalar it needs (log-
// determinants, pivot statuses, mu^T Q mu, the residual, the trace sums) in a
// device block copied to pinned host memory, and a finish half that waits
// for that copy: spde_loglik_batch enqueues many models on several streams
// before finishing any (the reference loops over models one by one,
// examples/server/get_results.py:6-51).
#include <hip/hip_runtime.h>

#include <thread>

#include <cmath>
#include <string>
#include <vector>

#include "internal.hpp"

namespace spde {

void launch_logq_sum(int64_t Ns, const double* q, double c, double* out, hipStream_t s);  // assemble.hip

namespace {

// device result block: doubles, then the three pivot statuses as doubles
enum : int { R_QUAD = 0, R_RSS, R_DOTGQ, R_DOT00, R_TAUTR, R_LOGQ, R_LDQC, R_LDQ0, R_LDG, R_ST0, R_ST1, R_ST2,
             kRes = 12 };
constexpr int kRedBlocks = 256;

struct LoglikWs {
  int64_t n = 0, nnz = 0, sp_nnz = 0, r = 0, m_cap = 0;
  double *qval = nullptr, *diag = nullptr, *b = nullptr, *qmu = nullptr, *sqc = nullptr;
  double *q0 = nullptr, *gram = nullptr, *sq0 = nullptr, *sgram = nullptr;
  int64_t* cols = nullptr;   // observation lattice indices (device)
  double* y = nullptr;       // observations, m x r column-major (device)
  double* part = nullptr;    // reduction partials
  double* res = nullptr;     // kRes result doubles (device)
  double* h_res = nullptr;   // pinned copy of res
  double *h_diag = nullptr, *h_b = nullptr, *h_y = nullptr;  // pinned staging
  int64_t* h_cols = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t factored = nullptr, done = nullptr, finished = nullptr;
  // the pending evaluation
  bool pending = false;
  int64_t m = 0;
  int rr = 0;
  double tau = 0.0, log_tau = 0.0;
  spde_plan* sp = nullptr;
};

int ensure_ws(spde_plan* pl, spde_plan* sp, int r, int64_t m, LoglikWs** out) {
  LoglikWs* w = static_cast<LoglikWs*>(pl->llws);
  if (!w) {
    w = new LoglikWs();
    pl->llws = w;
    w->n = pl->sym.n;
    w->nnz = (int64_t)pl->st.indices.size();
    SPDE_HIP(hipStreamCreateWithFlags(&w->side, hipStreamNonBlocking));
    SPDE_HIP(hipEventCreateWithFlags(&w->factored, hipEventDisableTiming));
    SPDE_HIP(hipEventCreateWithFlags(&w->done, hipEventDisableTiming));
    SPDE_HIP(hipEventCreateWithFlags(&w->finished, hipEventDisableTiming));
    SPDE_HIP(hipMalloc((void**)&w->qval, sizeof(double) * w->nnz));
    SPDE_HIP(hipMalloc((void**)&w->sqc, sizeof(double) * w->nnz));
    SPDE_HIP(hipMalloc((void**)&w->diag, sizeof(double) * w->n));
    SPDE_HIP(hipMalloc((void**)&w->part, sizeof(double) * kRedBlocks * 8));
    SPDE_HIP(hipMalloc((void**)&w->res, sizeof(double) * kRes));
    SPDE_HIP(hipHostMalloc((void**)&w->h_res, sizeof(double) * kRes, hipHostMallocDefault));
    SPDE_HIP(hipHostMalloc((void**)&w->h_diag, sizeof(double) * w->n, hipHostMallocDefault));
  }
  const int64_t snnz = (int64_t)sp->st.indices.size();
  if (w->sp_nnz != snnz) {
    SPDE_HIP(hipDeviceSynchronize());
    for (double** p : {&w->q0, &w->gram, &w->sq0, &w->sgram}) {
      hipFree(*p);
      *p = nullptr;
    }
    w->sp_nnz = 0;
    for (double** p : {&w->q0, &w->gram, &w->sq0, &w->sgram}) SPDE_HIP(hipMalloc((void**)p, sizeof(double) * snnz));
    w->sp_nnz = snnz;
  }
  if (w->r < r) {
    SPDE_HIP(hipDeviceSynchronize());
    hipFree(w->b);
    hipFree(w->qmu);
    hipHostFree(w->h_b);
    w->b = w->qmu = w->h_b = nullptr;
    w->r = 0;
    SPDE_HIP(hipMalloc((void**)&w->b, sizeof(double) * w->n * r));
    SPDE_HIP(hipMalloc((void**)&w->qmu, sizeof(double) * w->n * r));
    SPDE_HIP(hipHostMalloc((void**)&w->h_b, sizeof(double) * w->n * r, hipHostMallocDefault));
    w->r = r;
    w->m_cap = 0;  // y depends on r
  }
  if (w->m_cap < m) {
    SPDE_HIP(hipDeviceSynchronize());
    hipFree(w->cols);
    hipFree(w->y);
    hipHostFree(w->h_cols);
    hipHostFree(w->h_y);
    w->cols = nullptr;
    w->y = w->h_y = nullptr;
    w->h_cols = nullptr;
    SPDE_HIP(hipMalloc((void**)&w->cols, sizeof(int64_t) * m));
    SPDE_HIP(hipMalloc((void**)&w->y, sizeof(double) * m * w->r));
    SPDE_HIP(hipHostMalloc((void**)&w->h_cols, sizeof(int64_t) * m, hipHostMallocDefault));
    SPDE_HIP(hipHostMalloc((void**)&w->h_y, sizeof(double) * m * w->r, hipHostMallocDefault));
    w->m_cap = m;
  }
  *out = w;
  return SPDE_OK;
}

// Partial sums per block, then one thread sums them in block order
// (deterministic): quad = sum mu o (Q mu) over n x r; rss = sum over the
// observations of (y - mu[col])^2.
__global__ __launch_bounds__(256) void quad_rss_kernel(int64_t nr, const double* __restrict__ mu,
                                                       const double* __restrict__ qmu, int64_t m, int r, int64_t n,
                                                       const int64_t* __restrict__ cols, const double* __restrict__ y,
                                                       double* __restrict__ part) {
  __shared__ double red[2][256];
  double a = 0.0, b = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nr; i += (int64_t)gridDim.x * blockDim.x)
    a += mu[i] * qmu[i];
  const int64_t mr = m * r;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < mr; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = i / m, k = i - q * m;
    const double e = y[i] - mu[cols[k] + q * n];
    b += e * e;
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2] = red[0][0];
    part[blockIdx.x * 2 + 1] = red[1][0];
  }
}

// Column sums 36..38 of the per-cell adjoints (<G, Q>, <G00, Q0>, the tau trace).
__global__ __launch_bounds__(256) void colsum3_kernel(int64_t Ns, const double* __restrict__ cells,
                                                      double* __restrict__ part) {
  __shared__ double red[3][256];
  double a[3] = {0.0, 0.0, 0.0};
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < Ns; k += (int64_t)gridDim.x * blockDim.x)
    for (int c = 0; c < 3; ++c) a[c] += cells[k * 40 + 36 + c];
  for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = a[c];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) part[2 * kRedBlocks + blockIdx.x * 3 + c] = red[c][0];
}

// res[0..4] from the partials (block order); pivot statuses as doubles.
__global__ void finalize_kernel(int nb1, int nb2, const double* __restrict__ part, const int* __restrict__ st0,
                                const int* __restrict__ st1, const int* __restrict__ st2, double* __restrict__ res) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0.0, b = 0.0, c[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nb1; ++k) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  for (int k = 0; k < nb2; ++k)
    for (int q = 0; q < 3; ++q) c[q] += part[2 * kRedBlocks + 3 * k + q];
  res[R_QUAD] = a;
  res[R_RSS] = b;
  res[R_DOTGQ] = c[0];
  res[R_DOT00] = c[1];
  res[R_TAUTR] = c[2];
  res[R_ST0] = (double)st0[0];
  res[R_ST1] = (double)st1[0];
  res[R_ST2] = (double)st2[0];
}

inline unsigned grid_of(int64_t units) {
  const int64_t b = (units + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b < kRedBlocks ? b : kRedBlocks));
}

int check_args(spde_plan* pl, spde_plan* sp, const spde_coeffs* c, const int64_t* obs_cols, const double* y, int64_t m,
               int r, double* cells) {
  if (!pl || pl->host_only || !sp || sp->host_only || !c || !obs_cols || !y || m < 1 || r < 1 || !cells ||
      pl->lat.T < 2 || pl->lat.wide || pl->nslots < 2 || sp->lat.T != 1 || sp->nslots < 2 || sp->lat.M != pl->lat.M ||
      sp->lat.N != pl->lat.N || sp->lat.bc != pl->lat.bc || !(c->dt > 0.0) || !(c->sigma > 0.0)) {
    set_error("spde_loglik: invalid argument (space-time plan with 2 slots, its one-slice plan with 2 slots, "
              "observations, values and outputs)");
    return SPDE_ERR_ARG;
  }
  const int64_t n = pl->sym.n;
  for (int64_t i = 0; i < m; ++i)
    if (obs_cols[i] < 0 || obs_cols[i] >= n) {
      set_error("spde_loglik: observation index outside the lattice");
      return SPDE_ERR_ARG;
    }
  return SPDE_OK;
}

}  // namespace

void loglik_release(spde_plan* pl) {
  LoglikWs* w = static_cast<LoglikWs*>(pl->llws);
  if (!w) return;
  for (double* p : {w->qval, w->diag, w->b, w->qmu, w->sqc, w->q0, w->gram, w->sq0, w->sgram, w->y, w->part, w->res})
    hipFree(p);
  hipFree(w->cols);
  for (void* p : {(void*)w->h_res, (void*)w->h_diag, (void*)w->h_b, (void*)w->h_y, (void*)w->h_cols})
    if (p) hipHostFree(p);
  if (w->factored) hipEventDestroy(w->factored);
  if (w->done) hipEventDestroy(w->done);
  if (w->finished) hipEventDestroy(w->finished);
  if (w->side) hipStreamDestroy(w->side);
  delete w;
  pl->llws = nullptr;
}

}  // namespace spde

using namespace spde;

extern "C" {

int spde_loglik_enqueue(spde_plan* pl, spde_plan* sp, const spde_coeffs* c, const int64_t* obs_cols, const double* y,
                        int64_t m, int r, double log_tau, double* cells, void* stream) {
  int rc = check_args(pl, sp, c, obs_cols, y, m, r, cells);
  if (rc) return rc;
  LoglikWs* w = nullptr;
  if ((rc = ensure_ws(pl, sp, r, m, &w))) return rc;
  if (w->pending) {
    set_error("spde_loglik_enqueue: the plan has an evaluation in flight (finish it first)");
    return SPDE_ERR_STATE;
  }
  const int64_t n = pl->sym.n, Ns = pl->lat.Ns();
  hipStream_t s = (hipStream_t)stream;
  const double tau = std::exp(log_tau);
  // the pinned staging may still feed the previous evaluation's copies
  SPDE_HIP(hipEventSynchronize(w->finished));
  // tau S^T S (diagonal) and tau S^T y, in observation order
  for (int64_t i = 0; i < n; ++i) w->h_diag[i] = 0.0;
  for (int64_t i = 0; i < n * r; ++i) w->h_b[i] = 0.0;
  for (int64_t i = 0; i < m; ++i) w->h_diag[obs_cols[i]] += 1.0;
  for (int64_t i = 0; i < n; ++i) w->h_diag[i] *= tau;
  for (int q = 0; q < r; ++q)
    for (int64_t i = 0; i < m; ++i) w->h_b[obs_cols[i] + (int64_t)q * n] += y[i + (int64_t)q * m];
  for (int64_t i = 0; i < n * r; ++i) w->h_b[i] *= tau;
  for (int64_t i = 0; i < m; ++i) w->h_cols[i] = obs_cols[i];
  for (int64_t i = 0; i < m * r; ++i) w->h_y[i] = y[i];
  SPDE_HIP(hipMemcpyAsync(w->diag, w->h_diag, sizeof(double) * n, hipMemcpyHostToDevice, s));
  SPDE_HIP(hipMemcpyAsync(w->b, w->h_b, sizeof(double) * n * r, hipMemcpyHostToDevice, s));
  SPDE_HIP(hipMemcpyAsync(w->cols, w->h_cols, sizeof(int64_t) * m, hipMemcpyHostToDevice, s));
  SPDE_HIP(hipMemcpyAsync(w->y, w->h_y, sizeof(double) * m * r, hipMemcpyHostToDevice, s));
  if ((rc = spde_assemble(pl, c, w->qval, s))) return rc;
  // critical path: Q_c factorised here, its selected inverse on the side
  // stream (the wait is enqueued right after the event it waits on)
  if ((rc = spde_factor_async(pl, 1, w->qval, w->diag, s))) return rc;
  SPDE_HIP(hipEventRecord(w->factored, s));
  SPDE_HIP(hipStreamWaitEvent(w->side, w->factored, 0));
  if ((rc = spde_selinv(pl, 1, w->sqc, w->side))) return rc;
  // mu_c, Q mu_c and the one-slice log|Q| pieces, concurrent with it
  if ((rc = spde_solve(pl, 1, SPDE_SOLVE_A, w->b, r, n, s))) return rc;
  if ((rc = spde_spmv(pl, w->qval, w->b, w->qmu, r, s))) return rc;
  if ((rc = spde_assemble_spatial(pl, sp, 0, w->q0, s))) return rc;
  if ((rc = spde_assemble_spatial(pl, sp, 1, w->gram, s))) return rc;
  if ((rc = spde_factor_async(sp, 0, w->q0, nullptr, s))) return rc;
  if ((rc = spde_factor_async(sp, 1, w->gram, nullptr, s))) return rc;
  if ((rc = spde_selinv(sp, 0, w->sq0, s))) return rc;
  if ((rc = spde_selinv(sp, 1, w->sgram, s))) return rc;
  const unsigned g1 = grid_of(std::max<int64_t>(n * r, m * r));
  hipLaunchKernelGGL(quad_rss_kernel, dim3(g1), dim3(256), 0, s, n * r, w->b, w->qmu, m, r, n, w->cols, w->y, w->part);
  launch_logq_sum(Ns, pl->d_qdiag, 1.0 / (c->dt * c->sigma), w->res + R_LOGQ, s);
  SPDE_HIP(hipMemcpyAsync(w->res + R_LDQ0, sp->slots[0].ld, sizeof(double), hipMemcpyDeviceToDevice, s));
  SPDE_HIP(hipMemcpyAsync(w->res + R_LDG, sp->slots[1].ld, sizeof(double), hipMemcpyDeviceToDevice, s));
  SPDE_HIP(hipEventRecord(w->done, w->side));
  SPDE_HIP(hipStreamWaitEvent(s, w->done, 0));
  SPDE_HIP(hipMemcpyAsync(w->res + R_LDQC, pl->slots[1].ld, sizeof(double), hipMemcpyDeviceToDevice, s));
  // adjoint pass (closed-form log|Q| terms from the one-slice selected
  // inverses), then its three trace sums
  spde_grad_args g{};
  g.mode = 0;
  g.r = r;
  g.nh1 = 0;
  g.qval = w->qval;
  g.sigma_q = nullptr;
  g.sigma_qc = w->sqc;
  g.mu = w->b;
  g.diag_add = w->diag;
  g.coeffs = *c;
  g.sigma_gram = w->sgram;
  g.sigma_q0 = w->sq0;
  g.spatial = sp;
  if ((rc = spde_grad_cells(pl, &g, cells, s))) return rc;
  const unsigned g2 = grid_of(Ns);
  hipLaunchKernelGGL(colsum3_kernel, dim3(g2), dim3(256), 0, s, Ns, cells, w->part);
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(64), 0, s, (int)g1, (int)g2, w->part, pl->slots[1].status,
                     sp->slots[0].status, sp->slots[1].status, w->res);
  SPDE_HIP(hipMemcpyAsync(w->h_res, w->res, sizeof(double) * kRes, hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipEventRecord(w->finished, s));
  SPDE_HIP(hipGetLastError());
  w->pending = true;
  w->m = m;
  w->rr = r;
  w->tau = tau;
  w->log_tau = log_tau;
  w->sp = sp;
  return SPDE_OK;
}

int spde_loglik_finish(spde_plan* pl, spde_loglik_out* out) {
  if (!pl || pl->host_only || !out) {
    set_error("spde_loglik_finish: invalid argument");
    return SPDE_ERR_ARG;
  }
  LoglikWs* w = static_cast<LoglikWs*>(pl->llws);
  if (!w || !w->pending) {
    set_error("spde_loglik_finish: no evaluation in flight on this plan");
    return SPDE_ERR_STATE;
  }
  SPDE_HIP(hipEventSynchronize(w->finished));
  w->pending = false;
  const double* h = w->h_res;
  spde_plan* sp = w->sp;
  const int st[3] = {(int)h[R_ST0], (int)h[R_ST1], (int)h[R_ST2]};
  spde_plan* owner[3] = {pl, sp, sp};
  const int slot[3] = {1, 0, 1};
  for (int q = 0; q < 3; ++q)
    if (st[q] != 0) {
      owner[q]->slots[slot[q]].factored = false;
      set_error(std::string("spde_loglik: ") + (q == 0 ? "Q_c" : (q == 1 ? "Q0" : "B^T B")) +
                " is not positive definite (pivot at lattice index " +
                std::to_string(owner[q]->sym.perm[st[q] - 1]) + ")");
      return SPDE_ERR_NOT_SPD;
    }
  const int64_t n = pl->sym.n, Ns = pl->lat.Ns();
  const int T = pl->lat.T;
  const int r = w->rr;
  const int64_t m = w->m;
  const double tau = w->tau;
  const double ld_q = h[R_LDQ0] + (T - 1) * (h[R_LDG] + h[R_LOGQ]);
  const double like = 0.5 * ld_q * r + 0.5 * (double)m * r * w->log_tau - 0.5 * h[R_LDQC] * r - 0.5 * h[R_QUAD] -
                      0.5 * tau * h[R_RSS];
  out->like_raw = like;
  out->like = -like / ((double)m * r);
  out->logdet_q = ld_q;
  out->logdet_qc = h[R_LDQC];
  // <G, Q> without its cancellation: tr(Q_c^-1 Q) = n - tr(Q_c^-1 tau S^T S)
  const double dot_gq = -0.5 * r * ((double)n - h[R_TAUTR]) - 0.5 * h[R_QUAD];
  out->g_sigma = -(dot_gq - h[R_DOT00]) - 0.5 * r * (T - 1) * (double)Ns;
  out->g_tau = 0.5 * (double)m * r - 0.5 * r * h[R_TAUTR] - 0.5 * tau * h[R_RSS];
  return SPDE_OK;
}

int spde_loglik_grad(spde_plan* pl, spde_plan* sp, const spde_coeffs* c, const int64_t* obs_cols, const double* y,
                     int64_t m, int r, double log_tau, double* cells, spde_loglik_out* out, void* stream) {
  if (!out) {
    set_error("spde_loglik_grad: invalid argument");
    return SPDE_ERR_ARG;
  }
  int rc = spde_loglik_enqueue(pl, sp, c, obs_cols, y, m, r, log_tau, cells, stream);
  if (rc) return rc;
  return spde_loglik_finish(pl, out);
}

int spde_loglik_batch(int B, spde_plan* const* pls, spde_plan* const* sps, const spde_coeffs* cs,
                      const int64_t* const* obs_cols, const double* const* ys, const int64_t* ms, int r,
                      const double* log_taus, double* const* cells, spde_loglik_out* outs, void* const* streams,
                      int nstreams, int32_t* status) {
  if (B < 1 || !pls || !sps || !cs || !obs_cols || !ys || !ms || !log_taus || !cells || !outs || !streams ||
      nstreams < 1 || !status) {
    set_error("spde_loglik_batch: invalid argument");
    return SPDE_ERR_ARG;
  }
  // waves of models whose plans and one-slice plans are all distinct; within a
  // wave model i runs on stream i mod nstreams, and every model of a wave is
  // enqueued before any is finished
  int first_rc = SPDE_OK;
  std::string first_msg;
  int i0 = 0;
  while (i0 < B) {
    std::vector<spde_plan*> busy;
    int i1 = i0;
    while (i1 < B) {
      bool clash = false;
      for (spde_plan* p : busy) clash = clash || p == pls[i1] || p == sps[i1];
      if (clash) break;
      busy.push_back(pls[i1]);
      busy.push_back(sps[i1]);
      ++i1;
    }
    // enqueued from one host thread per stream (each model's plans are its
    // own within a wave, so the threads share no library state; the error
    // text is per thread)
    std::vector<int> enq(i1 - i0, 0);
    std::vector<std::string> msg(i1 - i0);
    const int nt = std::min(nstreams, i1 - i0);
    auto enqueue_part = [&](int t) {
      for (int i = i0 + t; i < i1; i += nt) {
        const int rc = spde_loglik_enqueue(pls[i], sps[i], &cs[i], obs_cols[i], ys[i], ms[i], r, log_taus[i],
                                           cells[i], streams[(i - i0) % nstreams]);
        status[i] = rc;
        enq[i - i0] = rc == SPDE_OK;
        if (rc) msg[i - i0] = spde_last_error();
      }
    };
    std::vector<std::thread> workers;
    for (int t = 1; t < nt; ++t) workers.emplace_back(enqueue_part, t);
    enqueue_part(0);
    for (auto& w : workers) w.join();
    for (int i = i0; i < i1; ++i)
      if (status[i] && !first_rc) {
        first_rc = status[i];
        first_msg = msg[i - i0];
      }
    for (int i = i0; i < i1; ++i) {
      if (!enq[i - i0]) continue;
      const int rc = spde_loglik_finish(pls[i], &outs[i]);
      status[i] = rc;
      if (rc && !first_rc) {
        first_rc = rc;
        first_msg = spde_last_error();
      }
    }
    i0 = i1;
  }
  if (first_rc) set_error("spde_loglik_batch: " + first_msg);
  return first_rc;
}

int spde_loglik_mu(spde_plan* pl, double* mu, int r, void* stream) {
  LoglikWs* w = pl ? static_cast<LoglikWs*>(pl->llws) : nullptr;
  if (!w || !mu || r < 1 || r > w->r) {
    set_error("spde_loglik_mu: no evaluation on this plan (or r too large)");
    return SPDE_ERR_STATE;
  }
  SPDE_HIP(hipMemcpyAsync(mu, w->b, sizeof(double) * w->n * r, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return SPDE_OK;
}

}  // extern "C"
