// One exact log-likelihood + gradient evaluation behind a single C call:
// assembly, factorisation of Q_c (main stream), its selected inverse (side
// stream), the solve for mu_c and the one-slice log|Q| work (main stream,
// concurrent with the selected inverse), then the adjoint pass.  The
// observation bookkeeping (tau S^T S, tau S^T y, the residual) is done on the
// host in observation order, as the data arrive from the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "internal.hpp"

namespace spde {

namespace {

struct LoglikWs {
  int64_t n = 0, nnz = 0, sp_nnz = 0, r = 0;
  double *qval = nullptr, *diag = nullptr, *b = nullptr, *qmu = nullptr, *sqc = nullptr;
  double *q0 = nullptr, *gram = nullptr, *sq0 = nullptr, *sgram = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t factored = nullptr, done = nullptr;
};

int ensure_ws(spde_plan* pl, spde_plan* sp, int r, LoglikWs** out) {
  LoglikWs* w = static_cast<LoglikWs*>(pl->llws);
  if (!w) {
    w = new LoglikWs();
    pl->llws = w;
    w->n = pl->sym.n;
    w->nnz = (int64_t)pl->st.indices.size();
    SPDE_HIP(hipStreamCreateWithFlags(&w->side, hipStreamNonBlocking));
    SPDE_HIP(hipEventCreateWithFlags(&w->factored, hipEventDisableTiming));
    SPDE_HIP(hipEventCreateWithFlags(&w->done, hipEventDisableTiming));
    SPDE_HIP(hipMalloc((void**)&w->qval, sizeof(double) * w->nnz));
    SPDE_HIP(hipMalloc((void**)&w->sqc, sizeof(double) * w->nnz));
    SPDE_HIP(hipMalloc((void**)&w->diag, sizeof(double) * w->n));
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
    w->b = w->qmu = nullptr;
    w->r = 0;
    SPDE_HIP(hipMalloc((void**)&w->b, sizeof(double) * w->n * r));
    SPDE_HIP(hipMalloc((void**)&w->qmu, sizeof(double) * w->n * r));
    w->r = r;
  }
  *out = w;
  return SPDE_OK;
}

}  // namespace

void loglik_release(spde_plan* pl) {
  LoglikWs* w = static_cast<LoglikWs*>(pl->llws);
  if (!w) return;
  for (double* p : {w->qval, w->diag, w->b, w->qmu, w->sqc, w->q0, w->gram, w->sq0, w->sgram}) hipFree(p);
  if (w->factored) hipEventDestroy(w->factored);
  if (w->done) hipEventDestroy(w->done);
  if (w->side) hipStreamDestroy(w->side);
  delete w;
  pl->llws = nullptr;
}

}  // namespace spde

using namespace spde;

extern "C" int spde_loglik_grad(spde_plan* pl, spde_plan* sp, const spde_coeffs* c, const int64_t* obs_cols,
                                const double* y, int64_t m, int r, double log_tau, double* cells,
                                spde_loglik_out* out, void* stream) {
  if (!pl || pl->host_only || !sp || sp->host_only || !c || !obs_cols || !y || m < 1 || r < 1 || !cells || !out ||
      pl->lat.T < 2 || pl->lat.wide || pl->nslots < 2 || sp->lat.T != 1 || sp->nslots < 2 || sp->lat.M != pl->lat.M ||
      sp->lat.N != pl->lat.N || sp->lat.bc != pl->lat.bc || !(c->dt > 0.0) || !(c->sigma > 0.0)) {
    set_error("spde_loglik_grad: invalid argument (space-time plan with 2 slots, its one-slice plan with 2 slots, "
              "observations, values and outputs)");
    return SPDE_ERR_ARG;
  }
  const int64_t n = pl->sym.n, Ns = pl->lat.Ns();
  for (int64_t i = 0; i < m; ++i)
    if (obs_cols[i] < 0 || obs_cols[i] >= n) {
      set_error("spde_loglik_grad: observation index outside the lattice");
      return SPDE_ERR_ARG;
    }
  LoglikWs* w = nullptr;
  if (int rc = ensure_ws(pl, sp, r, &w)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const double tau = std::exp(log_tau);
  // tau S^T S (diagonal) and tau S^T y, in observation order
  std::vector<double> diag(n, 0.0), b((size_t)n * r, 0.0);
  for (int64_t i = 0; i < m; ++i) diag[obs_cols[i]] += 1.0;
  for (int64_t i = 0; i < n; ++i) diag[i] *= tau;
  for (int q = 0; q < r; ++q)
    for (int64_t i = 0; i < m; ++i) b[obs_cols[i] + (int64_t)q * n] += y[i + (int64_t)q * m];
  for (double& v : b) v *= tau;
  SPDE_HIP(hipMemcpyAsync(w->diag, diag.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  SPDE_HIP(hipMemcpyAsync(w->b, b.data(), sizeof(double) * n * r, hipMemcpyHostToDevice, s));
  int rc = spde_assemble(pl, c, w->qval, s);
  if (rc) return rc;
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
  std::vector<double> mu((size_t)n * r), qmu((size_t)n * r);
  SPDE_HIP(hipMemcpyAsync(mu.data(), w->b, sizeof(double) * n * r, hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipMemcpyAsync(qmu.data(), w->qmu, sizeof(double) * n * r, hipMemcpyDeviceToHost, s));
  SPDE_HIP(hipEventRecord(w->done, w->side));
  SPDE_HIP(hipStreamWaitEvent(s, w->done, 0));
  // adjoint pass (closed-form log|Q| terms from the one-slice selected
  // inverses): it needs only device data, so it is enqueued, with the copy of
  // its result, before the host waits for the scalar pieces
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
  std::vector<double> hc((size_t)Ns * 40);
  SPDE_HIP(hipMemcpyAsync(hc.data(), cells, sizeof(double) * Ns * 40, hipMemcpyDeviceToHost, s));
  double ld_qc = 0.0, ld_q0 = 0.0, ld_g = 0.0, logq = 0.0;
  if ((rc = spde_factor_finish(pl, 1, &ld_qc))) return rc;
  if ((rc = spde_factor_finish(sp, 0, &ld_q0))) return rc;
  if ((rc = spde_factor_finish(sp, 1, &ld_g))) return rc;
  if ((rc = spde_logq_sum(pl, 1.0 / (c->dt * c->sigma), &logq, s))) return rc;
  const int T = pl->lat.T;
  const double ld_q = ld_q0 + (T - 1) * (ld_g + logq);
  double quad = 0.0, rss = 0.0;
  for (size_t i = 0; i < mu.size(); ++i) quad += mu[i] * qmu[i];
  for (int q = 0; q < r; ++q)
    for (int64_t i = 0; i < m; ++i) {
      const double e = y[i + (int64_t)q * m] - mu[obs_cols[i] + (int64_t)q * n];
      rss += e * e;
    }
  const double like = 0.5 * ld_q * r + 0.5 * (double)m * r * log_tau - 0.5 * ld_qc * r - 0.5 * quad - 0.5 * tau * rss;
  SPDE_HIP(hipStreamSynchronize(s));
  double dotgq = 0.0, dot00 = 0.0, tautr = 0.0;
  for (int64_t k = 0; k < Ns; ++k) {
    dotgq += hc[k * 40 + 36];
    dot00 += hc[k * 40 + 37];
    tautr += hc[k * 40 + 38];
  }
  out->like_raw = like;
  out->like = -like / ((double)m * r);
  out->logdet_q = ld_q;
  out->logdet_qc = ld_qc;
  // <G, Q> without its cancellation: tr(Q_c^-1 Q) = n - tr(Q_c^-1 tau S^T S)
  (void)dotgq;
  const double dot_gq = -0.5 * r * ((double)n - tautr) - 0.5 * quad;
  out->g_sigma = -(dot_gq - dot00) - 0.5 * r * (T - 1) * (double)Ns;
  out->g_tau = 0.5 * (double)m * r - 0.5 * r * tautr - 0.5 * tau * rss;
  return SPDE_OK;
}
