// Plan construction (device arenas, launch schedules) and the numeric
// multifrontal Cholesky, log-determinant, triangular solves and Takahashi
// selected inverse.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>

#include "internal.hpp"

namespace spde {

void launch_gemm(const GemmLaunch& g, const Bases& b, hipStream_t s, unsigned lds = 0);
void launch_potrf(const PotrfDesc* d, int n, double* F, double* Z, int* status, hipStream_t s);
void launch_fill(double* p, int64_t n, double v, hipStream_t s);
void launch_scatter(const double* q, const int64_t* map, int64_t nnz, double* F, hipStream_t s);
void launch_diag_add(const double* add, const int64_t* diag, const int64_t* map, int64_t n, double* F, hipStream_t s);
void launch_extend_add(const ExtAddDesc* d, const ExtAddTile* t, int ntiles, const int32_t* relpos, double* F,
                       hipStream_t s);
int logdet_parts(int64_t n);
void launch_logdet(const int64_t* diag, const int64_t* map, int64_t n, const double* F, double* partial, double* out,
                   hipStream_t s);
void launch_permute_in(const double* B, int64_t ldb, const int32_t* perm, int64_t n, int nrhs, double* xy,
                       hipStream_t s);
void launch_permute_out(const double* xy, const int32_t* perm, int64_t n, int nrhs, double* B, int64_t ldb,
                        hipStream_t s);
void launch_fwd_gather(const int32_t* lvl_nodes, int nn, const NodeDev* nodes, const int32_t* children,
                       const int32_t* relpos, const double* xy, int64_t n, int nrhs, double* vec, hipStream_t s);
void launch_bwd_gather(const int32_t* lvl_nodes, int nn, const NodeDev* nodes, const int32_t* bnd, const double* xy,
                       int64_t n, int nrhs, double* vec, hipStream_t s);
void launch_sigma_gather(const int32_t* lvl_nodes, const NodeDev* nodes, const int32_t* relpos, double* S,
                         const int4* tiles, int ntiles, hipStream_t s);
void launch_gather_pattern(const double* S, const int64_t* mapsym, int64_t nnz, double* out, hipStream_t s);
void launch_zero_lower(const NodeDev* nodes, const int4* tiles, int ntiles, double* F, hipStream_t s);
void loglik_release(spde_plan* pl);  // loglik.hip

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

// Clears a factorisation's status word (a kernel rather than a memset node,
// so that a captured factorisation replays correctly).
__global__ void clear_status_kernel(int* status) {
  if (threadIdx.x == 0) status[threadIdx.x] = 0;
}

// Device-side per-level index tables that are not part of LevelPlan.
struct LevelDev {
  int32_t* d_nodes = nullptr;
  int nn = 0;
  int4* d_gtiles = nullptr;
  int ngtiles = 0;
};

}  // namespace spde

using namespace spde;

namespace {

struct PlanExtra {            // kept beside the plan (not in the public struct)
  std::vector<LevelDev> lvl;
  int32_t* d_children = nullptr;
  int4* d_ztiles = nullptr;    // lower 64x64 tiles of every front (node, ti, tj)
  int nztiles = 0;
  // selected inverse: second stream for the products that need only the
  // factor, and its events (created at the first selected inverse)
  hipStream_t pre = nullptr;
  hipEvent_t ev_start = nullptr;
  std::vector<hipEvent_t> ev_pre;
};

PlanExtra* extra(spde_plan* p) { return static_cast<PlanExtra*>(p->extra); }

thread_local bool g_host_only_build = false;  // plan descriptors built for inspection only

template <class T>
int upload(T** dst, const std::vector<T>& src) {
  *dst = nullptr;
  if (src.empty() || g_host_only_build) return SPDE_OK;
  SPDE_HIP(hipMalloc((void**)dst, src.size() * sizeof(T)));
  SPDE_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return SPDE_OK;
}

// Arena sizes (doubles) per operand base, for the host-side bounds check.
struct ArenaSizes {
  int64_t sz[kNumBases];
};

bool operand_ok(int64_t off, int64_t ld, int64_t rows, int64_t cols, int64_t size) {
  if (rows <= 0 || cols <= 0) return true;
  if (off < 0 || ld < rows) return false;
  return off + (rows - 1) + (cols - 1) * ld < size;
}

// Every problem of a launch must touch only its operands' arenas.
int validate(const GemmLaunch& g, const ArenaSizes& a, const char* what) {
  for (const GemmProb& P : g.h_probs) {
    const int64_t ar = g.ta ? P.K : P.M, ac = g.ta ? P.M : P.K;
    const int64_t br = g.tb ? P.N : P.K, bc = g.tb ? P.K : P.N;
    const bool ok = operand_ok(P.a_off, P.lda, ar, ac, a.sz[P.a_base]) &&
                    operand_ok(P.b_off, P.ldb, br, bc, a.sz[P.b_base]) &&
                    operand_ok(P.c_off, P.ldc, P.M, P.N, a.sz[P.c_base]);
    if (!ok) {
      set_error(std::string("launch plan out of bounds in ") + what);
      return SPDE_ERR_STATE;
    }
  }
  for (const GemmTile& t : g.h_tiles)
    if (t.prob < 0 || t.prob >= (int)g.h_probs.size()) {
      set_error(std::string("tile list out of range in ") + what);
      return SPDE_ERR_STATE;
    }
  return SPDE_OK;
}

ArenaSizes arena_sizes(const spde_plan* pl, int nrhs) {
  ArenaSizes a;
  a.sz[kFront] = pl->sym.front_doubles;
  a.sz[kZ] = pl->sym.z_doubles;
  a.sz[kSigma] = pl->sym.front_doubles;
  a.sz[kWork] = pl->work_doubles;
  a.sz[kVec] = pl->sym.vec_doubles * nrhs;
  a.sz[kX] = pl->sym.n * nrhs;
  return a;
}

void make_tiles(GemmLaunch& g, int force_tile);

// Split-K for launches of 64-tiles that leave the chip's 1,024 block slots
// (four per CU) under-filled with deep problems: under 256 tiles and at least
// 512 deep, S = min(512 / tiles, K_max / 256, 64) chunks; under 1,024 tiles
// and at least 1,024 deep, S = min(round(2048 / tiles), K_max / 256) chunks
// (about two rounds of the chip).  Chunks are at least 256 deep; the partial
// tiles take S * tiles * 64 * 64 doubles of the workspace counted in *ws.
void plan_split(GemmLaunch& g, int64_t* ws) {
  if (!ws || g.gemv || g.tile != 64 || g.h_tiles.empty()) return;
  int kmax = 0;
  for (const GemmProb& P : g.h_probs) kmax = std::max(kmax, P.K);
  const int64_t nt = (int64_t)g.h_tiles.size();
  int S = 0;
  if (nt < 256 && kmax >= 512)
    S = (int)std::min<int64_t>(std::min<int64_t>(512 / nt, kmax / 256), 64);
  else if (nt < 1024 && kmax >= 1024)
    S = (int)std::min<int64_t>((2048 + nt / 2) / nt, kmax / 256);
  if (S < 2) return;
  std::vector<GemmTile> t;
  t.reserve(nt * S);
  for (const GemmTile& b : g.h_tiles)
    for (int q = 0; q < S; ++q) t.push_back({b.prob, b.tm, b.tn, q});
  g.h_tiles.swap(t);
  g.split = S;
  g.ws_off = *ws;
  *ws += (int64_t)g.h_tiles.size() * 64 * 64;
}

int finalize(GemmLaunch& g, int64_t* ws) {
  make_tiles(g, 0);
  plan_split(g, ws);
  g.nprobs = (int)g.h_probs.size();
  g.ntiles = (int)g.h_tiles.size();
  int rc = upload(&g.d_probs, g.h_probs);
  if (rc) return rc;
  return upload(&g.d_tiles, g.h_tiles);
}

// flop_scale: algorithmic share of 2MNK (0.5 for a triangular operand)
void add_prob(GemmLaunch& g, const GemmProb& P, bool lower, double flop_scale = 1.0) {
  if (P.M <= 0 || P.N <= 0) return;
  if (P.K <= 0 && P.beta == 1.0) return;
  GemmProb Q = P;
  Q.lower = lower ? 1 : 0;
  g.h_probs.push_back(Q);
  g.flops += 2.0 * P.M * (double)P.N * P.K * (lower ? 0.5 : 1.0) * flop_scale;
  const double cmn = P.M * (double)P.N * (lower ? 0.5 : 1.0);
  g.bytes += 8.0 * (P.M * (double)P.K + P.N * (double)P.K + cmn * (P.beta != 0.0 ? 2.0 : 1.0));
}

// Tile edge of a launch: 128 when the launch has no lower-triangular
// problem, a mean depth K of at least 256 and >= 512 tiles of 128 (two
// rounds of the chip at two blocks per CU); else 64 (short or triangular
// problems keep more blocks in flight and waste less on the diagonal).
// Then the tile list (tiles strictly above the diagonal of a lower problem
// are skipped).
void make_tiles(GemmLaunch& g, int force_tile) {
  // few right-hand sides: row blocks of 64 (A not transposed) or groups of 4
  // outputs (A transposed) for the bandwidth-bound kernel
  bool narrow = !g.h_probs.empty() && !g.tb && !force_tile;
  for (const GemmProb& P : g.h_probs) narrow = narrow && P.N <= kGemvMaxN && !P.lower;
  g.gemv = narrow;
  if (narrow) {
    g.tile = 0;
    g.h_tiles.clear();
    int kmax = 0;
    for (const GemmProb& P : g.h_probs) kmax = std::max(kmax, P.K);
    g.gemv_wide = !g.ta && kmax >= 1024;
    const int rows = g.ta ? 4 : 64;
    for (size_t i = 0; i < g.h_probs.size(); ++i)
      for (int a = 0; a < (g.h_probs[i].M + rows - 1) / rows; ++a) g.h_tiles.push_back({(int32_t)i, a, 0, 0});
    return;
  }
  int64_t t128 = 0;
  double ksum = 0.0;
  bool any_lower = false;
  int64_t t128_lower = 0;  // tiles of 128 actually launched (on or below the diagonal of lower problems)
  for (const GemmProb& P : g.h_probs) {
    const int64_t tm = (P.M + 127) / 128, tn = (P.N + 127) / 128;
    t128 += tm * tn;
    t128_lower += P.lower ? (tm <= tn ? tm * (tm + 1) / 2 : tm * tn - tn * (tn - 1) / 2) : tm * tn;
    ksum += P.K;
    any_lower = any_lower || P.lower;
  }
  // lower problems: 128 only when the launch still has two full rounds of
  // 128-tiles (the diagonal tiles do redundant work above the diagonal)
  const bool deep = !g.h_probs.empty() && ksum / g.h_probs.size() >= 256.0;
  const bool big = deep && (any_lower ? t128_lower >= 1024 : t128 >= 512);
  g.tile = force_tile ? force_tile : (big ? 128 : 64);
  g.h_tiles.clear();
  const int T = g.tile;
  for (size_t i = 0; i < g.h_probs.size(); ++i) {
    const GemmProb& P = g.h_probs[i];
    const int tm = (P.M + T - 1) / T, tn = (P.N + T - 1) / T;
    for (int a = 0; a < tm; ++a)
      for (int b = 0; b < tn; ++b)
        if (!P.lower || a >= b) g.h_tiles.push_back({(int32_t)i, a, b, 0});
  }
  // deepest tiles first: the blocks dispatched last are the shortest, which
  // shortens the launch's final, partly filled round
  auto depth = [&](const GemmTile& t) {
    const GemmProb& P = g.h_probs[t.prob];
    int kb = 0, ke = P.K;
    if (P.kmode == 1) kb = std::min(P.K, t.tn * T);
    if (P.kmode == 2) kb = std::min(P.K, t.tm * T);
    if (P.kmode == 3) ke = std::min(P.K, (t.tm + 1) * T);
    return ke - kb;
  };
  std::stable_sort(g.h_tiles.begin(), g.h_tiles.end(),
                   [&](const GemmTile& x, const GemmTile& y) { return depth(x) > depth(y); });
}

// A problem with a triangular operand whose zeros the kernel skips (kmode).
GemmProb tri(GemmProb P, int kmode) {
  P.kmode = (int8_t)kmode;
  return P;
}

GemmProb prob(int8_t ab, int64_t ao, int lda, int8_t bb, int64_t bo, int ldb, int8_t cb, int64_t co, int ldc, int M,
              int N, int K, double alpha, double beta) {
  GemmProb P{};
  P.a_off = ao;
  P.b_off = bo;
  P.c_off = co;
  P.lda = lda;
  P.ldb = ldb;
  P.ldc = ldc;
  P.M = M;
  P.N = N;
  P.K = K;
  P.a_base = ab;
  P.b_base = bb;
  P.c_base = cb;
  P.lower = 0;
  P.kmode = 0;
  P.alpha = alpha;
  P.beta = beta;
  return P;
}

int build_factor_plan(spde_plan* pl) {
  const Symbolic& sym = pl->sym;
  pl->split_fs = 0;
  pl->levels.assign(sym.levels.size(), LevelPlan());
  for (size_t l = 0; l < sym.levels.size(); ++l) {
    LevelPlan& L = pl->levels[l];
    L.nodes = sym.levels[l];
    for (int s = 0; s < 2; ++s) {
      for (int id : L.nodes) {
        const Node& nd = sym.nodes[id];
        if (s >= nd.nchild) continue;
        const Node& c = sym.nodes[nd.child[s]];
        const int b = (int)c.B.size();
        if (b == 0) continue;
        const int di = (int)L.h_ext[s].size();
        int split = 0;  // relpos increases: the first `split` columns map into the parent's own columns
        while (split < b && c.relpos[split] < nd.p) ++split;
        L.h_ext[s].push_back({c.foff, nd.foff, c.relpos_off, c.n(), c.p, nd.n(), b, split, 0});
        const int nt = (b + 63) / 64;
        for (int a = 0; a < nt; ++a)
          for (int bb = 0; bb <= a; ++bb) {
            if (bb * 64 < split) L.h_exttile[0][s].push_back({di, a, bb, 0});
            if ((bb + 1) * 64 > split) L.h_exttile[1][s].push_back({di, a, bb, 1});
          }
      }
      int rc = upload(&L.d_ext[s], L.h_ext[s]);
      if (rc) return rc;
      for (int ph = 0; ph < 2; ++ph) {
        L.next[ph][s] = (int)L.h_exttile[ph][s].size();
        rc = upload(&L.d_exttile[ph][s], L.h_exttile[ph][s]);
        if (rc) return rc;
      }
    }
    int pmax = 0;
    for (int id : L.nodes) pmax = std::max(pmax, sym.nodes[id].p);
    // Recursive column blocking of the nodes' own columns: rec(a, b) factors
    // columns [a, b) as rec(a, m), one update of columns [m, b) (all rows
    // below m) by columns [a, m) with depth m - a, rec(m, b); leaves are
    // 64-column steps (diagonal-block Cholesky + panel solve).  Same flops
    // as the right-looking sweep, but the trailing updates are few and deep
    // instead of one depth-64 pass over the trailing matrix per step.
    std::function<int(int, int)> rec = [&](int a, int b) -> int {
      Step st;
      st.panel.ta = false;
      st.panel.tb = true;
      st.update.ta = false;
      st.update.tb = true;
      if (b - a <= NB) {
        const int k0 = a;
        std::vector<PotrfDesc> pd;
        for (int id : L.nodes) {
          const Node& nd = sym.nodes[id];
          if (nd.p <= k0) continue;
          const int n = nd.n(), kb = std::min(NB, nd.p - k0);
          pd.push_back({nd.foff, nd.zoff, n, nd.p, k0, kb, id, nd.start + k0});
          const int M = n - k0 - kb;
          const int64_t off = nd.foff + (k0 + kb) + (int64_t)k0 * n;
          // in place (kb <= 64: one tile column); W is lower triangular
          add_prob(st.panel,
                   prob(kFront, off, n, kZ, nd.zoff + k0 + (int64_t)k0 * nd.p, nd.p, kFront, off, n, M, kb, kb, 1.0, 0.0),
                   false, 0.5);
        }
        st.npotrf = (int)pd.size();
        int rc = upload(&st.d_potrf, pd);
        if (rc) return rc;
        rc = finalize(st.panel, &pl->split_fs);
        if (rc) return rc;
        for (const GemmProb& P : st.panel.h_probs)
          if (P.N > st.panel.tile) {  // the in-place solve needs one tile column
            set_error("panel solve wider than a tile");
            return SPDE_ERR_STATE;
          }
        L.steps.push_back(std::move(st));
        return SPDE_OK;
      }
      int half = (b - a + 1) / 2;
      half = ((half + NB - 1) / NB) * NB;
      const int m = a + half;
      int rc = rec(a, m);
      if (rc) return rc;
      for (int id : L.nodes) {
        const Node& nd = sym.nodes[id];
        if (nd.p <= m) continue;
        const int n = nd.n(), M = n - m, N = std::min(b, nd.p) - m, K = m - a;
        const int64_t ab = nd.foff + m + (int64_t)a * n;
        add_prob(st.update, prob(kFront, ab, n, kFront, ab, n, kFront, nd.foff + m + (int64_t)m * n, n, M, N, K, -1.0, 1.0),
                 true);
      }
      rc = finalize(st.update, &pl->split_fs);
      if (rc) return rc;
      L.steps.push_back(std::move(st));
      return rec(m, b);
    };
    if (pmax > 0) {
      int rc = rec(0, pmax);
      if (rc) return rc;
    }
    L.syrk.ta = false;
    L.syrk.tb = true;
    for (int id : L.nodes) {
      const Node& nd = sym.nodes[id];
      const int n = nd.n(), b = (int)nd.B.size();
      // beta = 0: the update block is written here, the children's
      // contributions to it are added afterwards (extend-add phase 1)
      add_prob(L.syrk,
               prob(kFront, nd.foff + nd.p, n, kFront, nd.foff + nd.p, n, kFront, nd.foff + nd.p + (int64_t)nd.p * n, n,
                    b, b, nd.p, -1.0, 0.0),
               true);
    }
    int rc = finalize(L.syrk, &pl->split_fs);
    if (rc) return rc;
  }
  // inverse of L11 for every node: recursive halving, batched by height
  struct Op {
    int node, a, m, b;
  };
  std::vector<std::vector<Op>> ops;
  std::function<int(int, int, int)> rec = [&](int node, int a, int b) -> int {
    if (b - a <= NB) return 0;
    int half = (b - a + 1) / 2;
    half = ((half + NB - 1) / NB) * NB;
    const int m = a + half;
    if (m >= b) return 0;
    const int h = 1 + std::max(rec(node, a, m), rec(node, m, b));
    if ((int)ops.size() <= h) ops.resize(h + 1);
    ops[h].push_back({node, a, m, b});
    return h;
  };
  for (size_t q = 0; q < sym.nodes.size(); ++q) rec((int)q, 0, sym.nodes[q].p);
  pl->trtri.clear();
  int64_t work = 0;
  for (size_t h = 1; h < ops.size(); ++h) {
    GemmLaunch gx, gz;
    int64_t woff = 0;
    for (const Op& o : ops[h]) {
      const Node& nd = sym.nodes[o.node];
      const int n = nd.n(), p = nd.p;
      const int r = o.b - o.m, c = o.m - o.a;
      // X = L_ma Z_aa (Z_aa lower), Z_ma = -Z_mm X (Z_mm lower)
      add_prob(gx, tri(prob(kFront, nd.foff + o.m + (int64_t)o.a * n, n, kZ, nd.zoff + o.a + (int64_t)o.a * p, p,
                            kWork, woff, r, r, c, c, 1.0, 0.0),
                       1),
               false, 0.5);
      add_prob(gz, tri(prob(kZ, nd.zoff + o.m + (int64_t)o.m * p, p, kWork, woff, r, kZ,
                            nd.zoff + o.m + (int64_t)o.a * p, p, r, c, r, -1.0, 0.0),
                       3),
               false, 0.5);
      woff += (int64_t)r * c;
    }
    work = std::max(work, woff);
    int rc = finalize(gx, &pl->split_fs);
    if (rc) return rc;
    rc = finalize(gz, &pl->split_fs);
    if (rc) return rc;
    pl->trtri.push_back(std::move(gx));
    pl->trtri.push_back(std::move(gz));
  }
  pl->trtri_heights = (int)(pl->trtri.size() / 2);
  // selected-inverse workspace: per level sum of b*p
  for (auto& lv : sym.levels) {
    int64_t w = 0;
    for (int id : lv) w += (int64_t)sym.nodes[id].B.size() * sym.nodes[id].p;
    work = std::max(work, w);
  }
  work = std::max<int64_t>(work, pl->lat.Ns() * 60);
  pl->work_doubles = std::max<int64_t>(work, 1);
  const ArenaSizes as = arena_sizes(pl, 1);
  for (auto& L : pl->levels) {
    for (auto& st : L.steps) {
      int rc = validate(st.panel, as, "factor panel");
      if (rc) return rc;
      rc = validate(st.update, as, "factor update");
      if (rc) return rc;
    }
    int rc = validate(L.syrk, as, "factor syrk");
    if (rc) return rc;
    for (int c = 0; c < 2; ++c)
      for (const ExtAddDesc& d : L.h_ext[c])
        if (d.cfoff + (int64_t)d.cld * d.cld > pl->sym.front_doubles ||
            d.pfoff + (int64_t)d.pld * d.pld > pl->sym.front_doubles || d.cp + d.b != d.cld) {
          set_error("extend-add descriptor out of bounds");
          return SPDE_ERR_STATE;
        }
  }
  for (auto& g : pl->trtri) {
    int rc = validate(g, as, "inverse of L11");
    if (rc) return rc;
  }
  for (const Node& nd : sym.nodes)
    for (int rp : nd.relpos)
      if (nd.parent < 0 || rp < 0 || rp >= sym.nodes[nd.parent].n()) {
        set_error("relative position out of the parent front");
        return SPDE_ERR_STATE;
      }
  return SPDE_OK;
}

void free_launch(GemmLaunch& g);

int build_solve_plan(spde_plan* pl, int k) {
  auto hit = pl->solve_cache.find(k);
  if (hit != pl->solve_cache.end()) {
    const auto& v = hit->second.launches;
    for (size_t l = 0; l < pl->levels.size(); ++l) {
      LevelPlan& L = pl->levels[l];
      L.fwd_y = v[4 * l], L.fwd_u = v[4 * l + 1], L.bwd_t = v[4 * l + 2], L.bwd_x = v[4 * l + 3];
    }
    pl->split_solve = hit->second.split;
    pl->solve_nrhs = k;
    return SPDE_OK;
  }
  const Symbolic& sym = pl->sym;
  const int n = (int)sym.n;
  // solve launches take the split workspace after the factor and
  // selected-inverse regions (fixed when the plan is made)
  int64_t acc = pl->split_fs;
  spde_plan::SolveSet set;
  for (auto& L : pl->levels) {
    GemmLaunch fy, fu, bt, bx;
    fy.ta = false, fy.tb = false;
    fu.ta = false, fu.tb = false;
    bt.ta = true, bt.tb = false;
    bx.ta = true, bx.tb = false;
    for (int id : L.nodes) {
      const Node& nd = sym.nodes[id];
      const int nf = nd.n(), p = nd.p, b = (int)nd.B.size();
      const int64_t v = nd.voff * k;
      add_prob(fy, prob(kZ, nd.zoff, p, kVec, v, nf, kX, nd.start, n, p, k, p, 1.0, 0.0), false);
      add_prob(fu, prob(kFront, nd.foff + p, nf, kX, nd.start, n, kVec, v + p, nf, b, k, p, -1.0, 1.0), false);
      add_prob(bt, prob(kFront, nd.foff + p, nf, kVec, v + p, nf, kVec, v, nf, p, k, b, -1.0, 1.0), false);
      add_prob(bx, prob(kZ, nd.zoff, p, kVec, v, nf, kX, nd.start, n, p, k, p, 1.0, 0.0), false);
    }
    const ArenaSizes as = arena_sizes(pl, k);
    for (GemmLaunch* g : {&fy, &fu, &bt, &bx}) {
      int rc = validate(*g, as, "solve");
      if (!rc) rc = finalize(*g, &acc);
      if (rc) {
        for (auto& q : set.launches) free_launch(q);
        for (GemmLaunch* h : {&fy, &fu, &bt, &bx}) free_launch(*h);
        return rc;
      }
    }
    set.launches.push_back(fy);
    set.launches.push_back(fu);
    set.launches.push_back(bt);
    set.launches.push_back(bx);
  }
  set.split = acc - pl->split_fs;
  pl->solve_cache[k] = set;
  return build_solve_plan(pl, k);
}

int build_selinv_plan(spde_plan* pl) {
  const Symbolic& sym = pl->sym;
  PlanExtra* ex = extra(pl);
  // X = L21 Z of every level keeps its own region of the workspace: the X
  // products run ahead of the level chain on a second stream
  int64_t woff = 0;
  for (size_t l = 0; l < pl->levels.size(); ++l) {
    LevelPlan& L = pl->levels[l];
    L.si_x.ta = false, L.si_x.tb = false;
    L.si_s21.ta = false, L.si_s21.tb = false;
    L.si_s11a.ta = true, L.si_s11a.tb = false;
    L.si_s11b.ta = true, L.si_s11b.tb = false;
    std::vector<int4> gt;
    for (size_t q = 0; q < L.nodes.size(); ++q) {
      const Node& nd = sym.nodes[L.nodes[q]];
      const int nf = nd.n(), p = nd.p, b = (int)nd.B.size();
      if (b > 0 && nd.parent >= 0)
        for (int ti = 0; ti < (b + 63) / 64; ++ti)
          for (int tj = 0; tj <= ti; ++tj) gt.push_back(make_int4((int)q, ti, tj, 0));
      add_prob(L.si_x, tri(prob(kFront, nd.foff + p, nf, kZ, nd.zoff, p, kWork, woff, b, b, p, p, 1.0, 0.0), 1), false,
               0.5);
      add_prob(L.si_s21, prob(kSigma, nd.foff + p + (int64_t)p * nf, nf, kWork, woff, b, kSigma, nd.foff + p, nf, b, p,
                              b, -1.0, 0.0),
               false);
      add_prob(L.si_s11a, tri(prob(kZ, nd.zoff, p, kZ, nd.zoff, p, kSigma, nd.foff, nf, p, p, p, 1.0, 0.0), 2), true,
               1.0 / 3.0);
      add_prob(L.si_s11b, prob(kWork, woff, b, kSigma, nd.foff + p, nf, kSigma, nd.foff, nf, p, p, b, -1.0, 1.0), true);
      woff += (int64_t)b * p;
    }
    L.si_work = woff;
    if (woff > pl->work_doubles) {
      if (!pl->slots.empty()) {  // the slots' workspaces are allocated after this plan is built
        set_error("selected-inverse workspace too small");
        return SPDE_ERR_STATE;
      }
      pl->work_doubles = woff;
    }
    const ArenaSizes as = arena_sizes(pl, 1);
    for (GemmLaunch* g : {&L.si_x, &L.si_s21, &L.si_s11a, &L.si_s11b}) {
      int rc = validate(*g, as, "selected inverse");
      if (rc) return rc;
      rc = finalize(*g, &pl->split_fs);
      if (rc) return rc;
    }
    LevelDev& D = ex->lvl[l];
    D.ngtiles = (int)gt.size();
    int rc = upload(&D.d_gtiles, gt);
    if (rc) return rc;
  }
  pl->selinv_built = true;
  return SPDE_OK;
}

void free_launch(GemmLaunch& g) {
  hipFree(g.d_probs);
  hipFree(g.d_tiles);
  g.d_probs = nullptr;
  g.d_tiles = nullptr;
}

// LDS each 64-tile block of the selected inverse's second-stream products
// (X = L21 Z, Z^T Z, one level ahead of the chain) reserves: 48 KB holds one
// such block per CU, so the chain's own products on the first stream keep
// most of the chip.  128-tile launches run at full width: capping them too
// gained 0.5% of the evaluation rate but slowed each of their launches by a
// fifth (exact evaluation at 100x100x20, three alternating runs each: no cap
// 14.72 evals/s, 64-tiles capped 14.78, all capped 14.86).
constexpr unsigned kPreLds = 48 * 1024;
// GEMM launch, bracketed by HIP events when the plan's timing is enabled.
void run_gemm(spde_plan* pl, const GemmLaunch& g, const Bases& b, hipStream_t s, unsigned lds = 0) {
  if (g.ntiles == 0) return;
  const int kind = (g.tile == 128 ? 4 : 0) + (g.ta ? 2 : 0) + (g.tb ? 1 : 0);
  if (pl->prof_on && ((pl->prof_mask >> kind) & 1)) {
    while (pl->prof_ev.size() < 2 * (pl->prof_n + 1)) {
      hipEvent_t e;
      hipEventCreate(&e);
      pl->prof_ev.push_back(e);
    }
    hipEventRecord(pl->prof_ev[2 * pl->prof_n], s);
    launch_gemm(g, b, s, lds);
    hipEventRecord(pl->prof_ev[2 * pl->prof_n + 1], s);
    if (pl->prof_flops.size() <= pl->prof_n) {
      pl->prof_flops.resize(pl->prof_n + 1);
      pl->prof_bytes.resize(pl->prof_n + 1);
      pl->prof_tiles.resize(pl->prof_n + 1);
      pl->prof_kind.resize(pl->prof_n + 1);
    }
    pl->prof_flops[pl->prof_n] = g.flops;
    pl->prof_bytes[pl->prof_n] = g.bytes;
    pl->prof_tiles[pl->prof_n] = g.ntiles;
    pl->prof_kind[pl->prof_n] = kind;
    pl->prof_n++;
    pl->prof_stream = s;
    return;
  }
  launch_gemm(g, b, s, lds);
}

// The slot's split-K workspace covers the factor, selected-inverse and
// current solve regions (grows only when the solve plan grows, after the
// device has drained).
int ensure_splitws(spde_plan* pl, int slot) {
  Slot& sl = pl->slots[slot];
  const int64_t need = pl->split_fs + pl->split_solve;
  if (need <= sl.splitws_cap) return SPDE_OK;
  SPDE_HIP(hipDeviceSynchronize());
  hipFree(sl.splitws);
  sl.splitws = nullptr;
  sl.splitws_cap = 0;
  SPDE_HIP(hipMalloc((void**)&sl.splitws, sizeof(double) * need));
  sl.splitws_cap = need;
  return SPDE_OK;
}

Bases make_bases(spde_plan* pl, int slot, double* x) {
  Bases b;
  b.ws = pl->slots[slot].splitws;
  b.p[kFront] = pl->slots[slot].front;
  b.p[kZ] = pl->slots[slot].z;
  b.p[kSigma] = pl->slots[slot].sigma;
  b.p[kWork] = pl->slots[slot].work;
  b.p[kVec] = pl->slots[slot].vec;
  b.p[kX] = x;
  return b;
}

// Device arenas of a factor slot, allocated on its first factorisation.  The
// fronts and Z are cleared once (on the factorisation's stream): later
// factorisations clear only the lower triangles, and nothing writes above a
// front's diagonal.
int ensure_slot(spde_plan* pl, int slot, hipStream_t stream) {
  Slot& sl = pl->slots[slot];
  if (sl.front) return SPDE_OK;
  const Symbolic& sym = pl->sym;
  const size_t fb = sizeof(double) * std::max<int64_t>(sym.front_doubles, 1);
  const size_t zb = sizeof(double) * std::max<int64_t>(sym.z_doubles, 1);
  if (hipMalloc((void**)&sl.front, fb) != hipSuccess || hipMalloc((void**)&sl.z, zb) != hipSuccess ||
      hipMalloc((void**)&sl.work, sizeof(double) * pl->work_doubles) != hipSuccess ||
      hipMalloc((void**)&sl.partial, sizeof(double) * std::max(logdet_parts(sym.n), 1)) != hipSuccess ||
      hipMalloc((void**)&sl.ld, sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&sl.status, sizeof(int)) != hipSuccess) {
    for (void* p : {(void*)sl.front, (void*)sl.z, (void*)sl.work, (void*)sl.partial, (void*)sl.ld, (void*)sl.status})
      if (p) hipFree(p);
    sl.front = sl.z = sl.work = sl.partial = sl.ld = nullptr;
    sl.status = nullptr;
    set_error("factor slot " + std::to_string(slot) + ": device allocation of " +
              std::to_string((fb + zb) >> 20) + " MiB failed");
    return SPDE_ERR_ALLOC;
  }
  // cleared on the stream of the factorisation that follows, so the clear is
  // ordered before it whatever stream the caller uses
  SPDE_HIP(hipMemsetAsync(sl.front, 0, fb, stream));
  SPDE_HIP(hipMemsetAsync(sl.z, 0, zb, stream));
  SPDE_HIP(hipMemsetAsync(sl.status, 0, sizeof(int), stream));
  return SPDE_OK;
}

int check_slot(spde_plan* pl, int slot) {
  if (!pl || pl->host_only) {
    set_error("plan has no device state");
    return SPDE_ERR_STATE;
  }
  if (slot < 0 || slot >= pl->nslots) {
    set_error("factor slot out of range");
    return SPDE_ERR_ARG;
  }
  return SPDE_OK;
}

}  // namespace

extern "C" {

const char* spde_last_error(void) { return g_err.c_str(); }

int spde_plan_create(int M, int N, int T, int bc, int leaf_max, int nslots, int flags, spde_plan** out) {
  if (!out || bc < 1 || bc > 3 || M < 2 || N < 2 || T < 1 || nslots < 0) {
    set_error("spde_plan_create: invalid argument");
    return SPDE_ERR_ARG;
  }
  *out = nullptr;
  spde_plan* pl = new spde_plan();
  pl->lat = Lattice{M, N, T, bc};
  pl->lat.wide = (flags & SPDE_PLAN_WIDE_TIME) ? 1 : 0;
  try {
    build_structure(pl->lat, pl->st);
    if (pl->lat.wide) check_wide_layout(pl->lat, pl->st);
    build_symbolic(pl->lat, pl->st, leaf_max > 0 ? leaf_max : 128, pl->sym);
    build_scatter_map(pl->st, pl->sym, pl->map);
  } catch (const std::exception& e) {
    set_error(std::string("spde_plan_create: ") + e.what());
    delete pl;
    return SPDE_ERR_ARG;
  }
  const int64_t nnz = (int64_t)pl->st.indices.size();
  pl->mapsym.resize(nnz);
  for (int64_t e = 0; e < nnz; ++e) pl->mapsym[e] = pl->map[e] >= 0 ? pl->map[e] : pl->map[pl->st.mirror[e]];
  pl->host_only = (flags & SPDE_PLAN_HOST_ONLY) != 0;
  pl->nslots = nslots;
  PlanExtra* ex = new PlanExtra();
  pl->extra = ex;
  if (pl->host_only) {
    // build and bounds-check every launch schedule without touching the device
    g_host_only_build = true;
    ex->lvl.resize(pl->sym.levels.size());
    int rc = build_factor_plan(pl);
    if (!rc) rc = build_selinv_plan(pl);
    if (!rc) rc = build_solve_plan(pl, 3);
    g_host_only_build = false;
    if (rc) {
      spde_plan_destroy(pl);
      return rc;
    }
    *out = pl;
    return SPDE_OK;
  }
  const Symbolic& sym = pl->sym;
  int rc = SPDE_OK;
#define TRY(x)                   \
  do {                           \
    rc = (x);                    \
    if (rc) {                    \
      spde_plan_destroy(pl);     \
      return rc;                 \
    }                            \
  } while (0)
  TRY(upload(&pl->d_indptr, pl->st.indptr));
  TRY(upload(&pl->d_indices, pl->st.indices));
  TRY(upload(&pl->d_code, pl->st.code));
  // the diagonal add and the log-determinant read F[map[diag[i]]] for every variable
  if ((int64_t)pl->st.diag.size() != sym.n) TRY((set_error("diagonal positions do not match the order"), SPDE_ERR_STATE));
  for (int64_t i = 0; i < sym.n; ++i) {
    const int64_t e = pl->st.diag[i];
    if (e < 0 || e >= (int64_t)pl->map.size() || pl->map[e] < 0 || pl->map[e] >= sym.front_doubles)
      TRY((set_error("diagonal entry outside the fronts"), SPDE_ERR_STATE));
  }
  TRY(upload(&pl->d_diag, pl->st.diag));
  TRY(upload(&pl->d_map, pl->map));
  TRY(upload(&pl->d_mapsym, pl->mapsym));
  TRY(upload(&pl->d_perm, sym.perm));
  std::vector<int32_t> relpos(std::max<int64_t>(sym.relpos_total, 1)), bnd(std::max<int64_t>(sym.relpos_total, 1));
  std::vector<NodeDev> nodes(sym.nodes.size());
  std::vector<int32_t> children(2 * sym.nodes.size(), -1);
  for (size_t q = 0; q < sym.nodes.size(); ++q) {
    const Node& nd = sym.nodes[q];
    for (size_t k = 0; k < nd.B.size(); ++k) {
      bnd[nd.relpos_off + k] = nd.B[k];
      relpos[nd.relpos_off + k] = nd.relpos.empty() ? 0 : nd.relpos[k];
    }
    nodes[q] = NodeDev{nd.foff, nd.zoff, nd.voff, nd.relpos_off, nd.start, nd.p, nd.n(), nd.parent,
                       (int32_t)nd.B.size(), nd.level, 0, 0};
    for (int c = 0; c < nd.nchild; ++c) children[2 * q + c] = nd.child[c];
  }
  TRY(upload(&pl->d_relpos, relpos));
  TRY(upload(&pl->d_bnd, bnd));
  TRY(upload(&pl->d_nodes, nodes));
  TRY(upload(&ex->d_children, children));
  {
    std::vector<int4> zt;
    for (size_t q = 0; q < sym.nodes.size(); ++q) {
      // lower triangle of the own columns only (the update block is
      // written by the SYRK before anything adds to it)
      const int nt = (sym.nodes[q].n() + 63) / 64, pt = (sym.nodes[q].p + 63) / 64;
      for (int a = 0; a < nt; ++a)
        for (int b = 0; b <= a && b < pt; ++b) zt.push_back(int4{(int)q, a, b, 0});
    }
    if (zt.size() > (size_t)INT32_MAX) TRY((set_error("front tile list too long"), SPDE_ERR_STATE));
    ex->nztiles = (int)zt.size();
    TRY(upload(&ex->d_ztiles, zt));
  }
  ex->lvl.resize(sym.levels.size());
  for (size_t l = 0; l < sym.levels.size(); ++l) {
    ex->lvl[l].nn = (int)sym.levels[l].size();
    TRY(upload(&ex->lvl[l].d_nodes, sym.levels[l]));
  }
  TRY(build_factor_plan(pl));
  TRY(build_selinv_plan(pl));  // fixes the split workspace of factor and selected inverse
  const int64_t Ns = pl->lat.Ns();
  SPDE_HIP(hipMalloc((void**)&pl->d_tmpl, sizeof(double) * Ns * kTmpl));
  SPDE_HIP(hipMalloc((void**)&pl->d_red, sizeof(double)));
  if (pl->lat.T == 1) {
    std::vector<int32_t> pos((size_t)Ns * 25, -1);
    for (int64_t r = 0; r < Ns; ++r)
      for (int64_t e = pl->st.indptr[r]; e < pl->st.indptr[r + 1]; ++e)
        if (pl->st.code[e] < 25) pos[(size_t)r * 25 + pl->st.code[e]] = (int32_t)e;
    TRY(upload(&pl->d_pos25, pos));
  }
  SPDE_HIP(hipMalloc((void**)&pl->d_acan, sizeof(double) * Ns * 9));
  SPDE_HIP(hipMalloc((void**)&pl->d_a0can, sizeof(double) * Ns * 9));
  SPDE_HIP(hipMalloc((void**)&pl->d_qdiag, sizeof(double) * Ns));
  // slot arenas are allocated by the slot's first factorisation (ensure_slot):
  // a plan with two slots of which one is used holds one slot's memory
  pl->slots.resize(nslots);
  SPDE_HIP(hipMalloc((void**)&pl->d_win, sizeof(double) * Ns * 60));
  *out = pl;
  return SPDE_OK;
#undef TRY
}

void spde_plan_destroy(spde_plan* pl) {
  if (!pl) return;
  if (!pl->host_only) spde::loglik_release(pl);
  if (pl->host_only) {
    delete extra(pl);
  } else {
    hipDeviceSynchronize();
    for (auto* p : {(void*)pl->d_indptr, (void*)pl->d_indices, (void*)pl->d_code, (void*)pl->d_diag,
                    (void*)pl->d_map, (void*)pl->d_mapsym, (void*)pl->d_perm, (void*)pl->d_relpos,
                    (void*)pl->d_bnd, (void*)pl->d_nodes, (void*)pl->d_tmpl, (void*)pl->d_acan,
                    (void*)pl->d_a0can, (void*)pl->d_qdiag, (void*)pl->d_win, (void*)pl->d_red, (void*)pl->d_pos25,
                    (void*)pl->d_sprow, (void*)pl->d_spcol, (void*)pl->d_gpart, (void*)pl->d_tan})
      hipFree(p);
    for (auto& s : pl->slots) {
      hipFree(s.front);
      hipFree(s.z);
      hipFree(s.sigma);
      hipFree(s.work);
      hipFree(s.partial);
      hipFree(s.ld);
      hipFree(s.status);
      hipFree(s.vec);
      hipFree(s.xy);
      hipFree(s.splitws);
    }
    for (auto& L : pl->levels) {
      for (int s = 0; s < 2; ++s) {
        hipFree(L.d_ext[s]);
        hipFree(L.d_exttile[0][s]);
        hipFree(L.d_exttile[1][s]);
      }
      for (auto& st : L.steps) {
        hipFree(st.d_potrf);
        free_launch(st.panel);
        free_launch(st.update);
      }
      for (GemmLaunch* g : {&L.syrk, &L.si_x, &L.si_s21, &L.si_s11a, &L.si_s11b}) free_launch(*g);
    }
    for (auto& g : pl->trtri) free_launch(g);
    for (auto& kv : pl->solve_cache)
      for (auto& g : kv.second.launches) free_launch(g);
    for (auto e : pl->prof_ev) hipEventDestroy(e);
    if (pl->prof_base) hipEventDestroy(pl->prof_base);
    if (PlanExtra* ex = extra(pl)) {
      for (auto& d : ex->lvl) {
        hipFree(d.d_nodes);
        hipFree(d.d_gtiles);
      }
      hipFree(ex->d_children);
      hipFree(ex->d_ztiles);
      for (auto e : ex->ev_pre) hipEventDestroy(e);
      if (ex->ev_start) hipEventDestroy(ex->ev_start);
      if (ex->pre) hipStreamDestroy(ex->pre);
      delete ex;
    }
  }
  delete pl;
}

int spde_plan_get_info(const spde_plan* pl, spde_plan_info* info) {
  if (!pl || !info) {
    set_error("spde_plan_get_info: null argument");
    return SPDE_ERR_ARG;
  }
  info->n = pl->sym.n;
  info->nnz_struct = (int64_t)pl->st.indices.size();
  info->nnz_L = pl->sym.nnz_L;
  info->front_doubles = pl->sym.front_doubles;
  info->z_doubles = pl->sym.z_doubles;
  info->nnodes = (int64_t)pl->sym.nodes.size();
  info->nlevels = (int64_t)pl->sym.levels.size();
  info->max_front = pl->sym.max_front;
  info->max_p = pl->sym.max_p;
  info->flops_factor = pl->sym.flops_factor;
  info->flops_selinv = pl->sym.flops_selinv;
  info->split_doubles = pl->split_fs + pl->split_solve;
  return SPDE_OK;
}

int spde_plan_export(const spde_plan* pl, int64_t* indptr, int32_t* indices, uint8_t* code, int32_t* perm,
                     int32_t* nodes, int64_t* foff, int32_t* bnd, int32_t* relpos, int64_t* map) {
  if (!pl) {
    set_error("spde_plan_export: null plan");
    return SPDE_ERR_ARG;
  }
  const auto& st = pl->st;
  const auto& sym = pl->sym;
  if (indptr) std::memcpy(indptr, st.indptr.data(), st.indptr.size() * sizeof(int64_t));
  if (indices) std::memcpy(indices, st.indices.data(), st.indices.size() * sizeof(int32_t));
  if (code) std::memcpy(code, st.code.data(), st.code.size());
  if (perm) std::memcpy(perm, sym.perm.data(), sym.perm.size() * sizeof(int32_t));
  if (map) std::memcpy(map, pl->map.data(), pl->map.size() * sizeof(int64_t));
  for (size_t q = 0; q < sym.nodes.size(); ++q) {
    const Node& nd = sym.nodes[q];
    if (nodes) {
      int32_t* r = nodes + 8 * q;
      r[0] = nd.start;
      r[1] = nd.p;
      r[2] = nd.n();
      r[3] = nd.parent;
      r[4] = nd.child[0];
      r[5] = nd.child[1];
      r[6] = nd.level;
      r[7] = (int32_t)nd.B.size();
    }
    if (foff) foff[q] = nd.foff;
    for (size_t k = 0; k < nd.B.size(); ++k) {
      if (bnd) bnd[nd.relpos_off + k] = nd.B[k];
      if (relpos) relpos[nd.relpos_off + k] = nd.relpos.empty() ? -1 : nd.relpos[k];
    }
  }
  return SPDE_OK;
}

int spde_factor_async(spde_plan* pl, int slot, const double* qval, const double* diag_add, void* stream) {
  int rc = check_slot(pl, slot);
  if (rc) return rc;
  if (!qval) {
    set_error("spde_factor_async: null values");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  rc = ensure_slot(pl, slot, s);
  if (rc) return rc;
  Slot& sl = pl->slots[slot];
  const Symbolic& sym = pl->sym;
  PlanExtra* ex = extra(pl);
  rc = ensure_splitws(pl, slot);
  if (rc) return rc;
  hipLaunchKernelGGL(clear_status_kernel, dim3(1), dim3(64), 0, s, sl.status);
  launch_zero_lower(pl->d_nodes, ex->d_ztiles, ex->nztiles, sl.front, s);
  launch_scatter(qval, pl->d_map, (int64_t)pl->st.indices.size(), sl.front, s);
  if (diag_add) launch_diag_add(diag_add, pl->d_diag, pl->d_map, sym.n, sl.front, s);
  Bases b = make_bases(pl, slot, nullptr);
  for (size_t l = 0; l < pl->levels.size(); ++l) {
    LevelPlan& L = pl->levels[l];
    for (int c = 0; c < 2; ++c) launch_extend_add(L.d_ext[c], L.d_exttile[0][c], L.next[0][c], pl->d_relpos, sl.front, s);
    for (Step& st : L.steps) {  // a leaf (potrf + panel) or an update
      launch_potrf(st.d_potrf, st.npotrf, sl.front, sl.z, sl.status, s);
      run_gemm(pl, st.panel, b, s);
      run_gemm(pl, st.update, b, s);
    }
    run_gemm(pl, L.syrk, b, s);
    for (int c = 0; c < 2; ++c) launch_extend_add(L.d_ext[c], L.d_exttile[1][c], L.next[1][c], pl->d_relpos, sl.front, s);
  }
  for (auto& g : pl->trtri) run_gemm(pl, g, b, s);
  launch_logdet(pl->d_diag, pl->d_map, sym.n, sl.front, sl.partial, sl.ld, s);
  SPDE_HIP(hipGetLastError());
  sl.stream = s;
  sl.factored = true;  // a non-positive pivot is reported by spde_factor_finish
  return SPDE_OK;
}

int spde_factor_finish(spde_plan* pl, int slot, double* logdet) {
  int rc = check_slot(pl, slot);
  if (rc) return rc;
  Slot& sl = pl->slots[slot];
  if (!sl.front) {  // (the stream may be the null stream)
    set_error("spde_factor_finish: no factorisation was enqueued in this slot");
    return SPDE_ERR_STATE;
  }
  int status = 0;
  double ld = 0.0;
  SPDE_HIP(hipMemcpyAsync(&status, sl.status, sizeof(int), hipMemcpyDeviceToHost, sl.stream));
  SPDE_HIP(hipMemcpyAsync(&ld, sl.ld, sizeof(double), hipMemcpyDeviceToHost, sl.stream));
  SPDE_HIP(hipStreamSynchronize(sl.stream));
  if (status != 0) {
    sl.factored = false;
    const int col = status - 1;
    set_error("spde_factor: matrix is not positive definite (pivot at lattice index " +
              std::to_string(pl->sym.perm[col]) + ")");
    return SPDE_ERR_NOT_SPD;
  }
  if (logdet) *logdet = ld;
  return SPDE_OK;
}

int spde_factor(spde_plan* pl, int slot, const double* qval, const double* diag_add, double* logdet, void* stream) {
  int rc = spde_factor_async(pl, slot, qval, diag_add, stream);
  if (rc) return rc;
  return spde_factor_finish(pl, slot, logdet);
}

int spde_solve(spde_plan* pl, int slot, int mode, double* B, int nrhs, int64_t ldb, void* stream) {
  int rc = check_slot(pl, slot);
  if (rc) return rc;
  if (!B || nrhs <= 0 || ldb < pl->sym.n || mode < 0 || mode > SPDE_SOLVE_LT_NOPERM) {
    set_error("spde_solve: invalid argument");
    return SPDE_ERR_ARG;
  }
  if (mode == SPDE_SOLVE_P || mode == SPDE_SOLVE_PT) {
    // the permutation alone: no factor needed, the slot's staging buffer only
    Slot& sl = pl->slots[slot];
    const int64_t n = pl->sym.n;
    if (sl.xy_cap < n * nrhs) {
      SPDE_HIP(hipDeviceSynchronize());
      hipFree(sl.xy);
      sl.xy = nullptr;
      sl.xy_cap = 0;
      SPDE_HIP(hipMalloc((void**)&sl.xy, sizeof(double) * n * nrhs));
      sl.xy_cap = n * nrhs;
    }
    hipStream_t s = (hipStream_t)stream;
    if (mode == SPDE_SOLVE_P) {
      launch_permute_in(B, ldb, pl->d_perm, n, nrhs, sl.xy, s);
      SPDE_HIP(hipMemcpy2DAsync(B, ldb * sizeof(double), sl.xy, n * sizeof(double), n * sizeof(double), nrhs,
                                hipMemcpyDeviceToDevice, s));
    } else {
      SPDE_HIP(hipMemcpy2DAsync(sl.xy, n * sizeof(double), B, ldb * sizeof(double), n * sizeof(double), nrhs,
                                hipMemcpyDeviceToDevice, s));
      launch_permute_out(sl.xy, pl->d_perm, n, nrhs, B, ldb, s);
    }
    return SPDE_OK;
  }
  if (!pl->slots[slot].factored) {
    set_error("spde_solve: slot not factored");
    return SPDE_ERR_STATE;
  }
  hipStream_t s = (hipStream_t)stream;
  PlanExtra* ex = extra(pl);
  const Symbolic& sym = pl->sym;
  const int64_t n = sym.n;
  Slot& sl = pl->slots[slot];
  if (pl->solve_nrhs != nrhs && pl->solve_cache.count(nrhs)) {
    // launch lists of this nrhs exist (their descriptors never change): swap them in
    rc = build_solve_plan(pl, nrhs);
    if (rc) return rc;
  }
  if (sl.vec_cap < sym.vec_doubles * nrhs || sl.xy_cap < n * nrhs || pl->solve_nrhs != nrhs) {
    // buffers change or new launch lists are built: let every stream drain first
    SPDE_HIP(hipDeviceSynchronize());
    if (sl.vec_cap < sym.vec_doubles * nrhs) {
      hipFree(sl.vec);
      sl.vec = nullptr;
      sl.vec_cap = 0;
      SPDE_HIP(hipMalloc((void**)&sl.vec, sizeof(double) * sym.vec_doubles * nrhs));
      sl.vec_cap = sym.vec_doubles * nrhs;
    }
    if (sl.xy_cap < n * nrhs) {
      hipFree(sl.xy);
      sl.xy = nullptr;
      sl.xy_cap = 0;
      SPDE_HIP(hipMalloc((void**)&sl.xy, sizeof(double) * n * nrhs));
      sl.xy_cap = n * nrhs;
    }
    if (pl->solve_nrhs != nrhs) {
      rc = build_solve_plan(pl, nrhs);
      if (rc) return rc;
    }
  }
  rc = ensure_splitws(pl, slot);
  if (rc) return rc;
  Bases b = make_bases(pl, slot, sl.xy);
  const int nl = (int)pl->levels.size();
  // in: P B for modes A and L; B as it is (already in elimination order) for the others
  const bool perm_in = mode == SPDE_SOLVE_A || mode == SPDE_SOLVE_L;
  // out: P^T X for modes A and LT; X in elimination order for L and the unpermuted modes
  const bool perm_out = mode == SPDE_SOLVE_A || mode == SPDE_SOLVE_LT;
  const bool fwd = mode == SPDE_SOLVE_A || mode == SPDE_SOLVE_L || mode == SPDE_SOLVE_L_NOPERM;
  const bool bwd = mode == SPDE_SOLVE_A || mode == SPDE_SOLVE_LT || mode == SPDE_SOLVE_LT_NOPERM;
  if (perm_in) {
    launch_permute_in(B, ldb, pl->d_perm, n, nrhs, sl.xy, s);
  } else {
    SPDE_HIP(hipMemcpy2DAsync(sl.xy, n * sizeof(double), B, ldb * sizeof(double), n * sizeof(double), nrhs,
                              hipMemcpyDeviceToDevice, s));
  }
  if (fwd) {
    for (int l = 0; l < nl; ++l) {
      LevelPlan& L = pl->levels[l];
      launch_fwd_gather(ex->lvl[l].d_nodes, ex->lvl[l].nn, pl->d_nodes, ex->d_children, pl->d_relpos, sl.xy, n,
                        nrhs, sl.vec, s);
      run_gemm(pl, L.fwd_y, b, s);
      run_gemm(pl, L.fwd_u, b, s);
    }
  }
  if (bwd) {
    for (int l = nl - 1; l >= 0; --l) {
      LevelPlan& L = pl->levels[l];
      launch_bwd_gather(ex->lvl[l].d_nodes, ex->lvl[l].nn, pl->d_nodes, pl->d_bnd, sl.xy, n, nrhs, sl.vec, s);
      run_gemm(pl, L.bwd_t, b, s);
      run_gemm(pl, L.bwd_x, b, s);
    }
  }
  if (perm_out) {
    launch_permute_out(sl.xy, pl->d_perm, n, nrhs, B, ldb, s);
  } else {
    SPDE_HIP(hipMemcpy2DAsync(B, ldb * sizeof(double), sl.xy, n * sizeof(double), n * sizeof(double), nrhs,
                              hipMemcpyDeviceToDevice, s));
  }
  return SPDE_OK;
}

int spde_prof_begin(spde_plan* pl) { return spde_prof_begin_kinds(pl, 0xff); }

int spde_prof_begin_kinds(spde_plan* pl, int kind_mask) {
  if (!pl || pl->host_only || kind_mask <= 0 || kind_mask > 0xff) {
    set_error("spde_prof_begin: no device plan or empty kind mask");
    return SPDE_ERR_ARG;
  }
  pl->prof_mask = kind_mask;
  if (!pl->prof_base) SPDE_HIP(hipEventCreate(&pl->prof_base));
  SPDE_HIP(hipDeviceSynchronize());
  SPDE_HIP(hipEventRecord(pl->prof_base, nullptr));
  SPDE_HIP(hipEventSynchronize(pl->prof_base));
  pl->prof_on = true;
  pl->prof_n = 0;
  return SPDE_OK;
}

int spde_prof_dump(spde_plan* pl, double* ms, double* flops, double* bytes, int32_t* kind, int32_t* tiles, int64_t cap,
                   int64_t* count) {
  if (!pl || pl->host_only) {
    set_error("spde_prof_dump: no device plan");
    return SPDE_ERR_ARG;
  }
  if (pl->prof_n) SPDE_HIP(hipDeviceSynchronize());  // launches may sit on several streams
  const int64_t n = std::min<int64_t>(cap, (int64_t)pl->prof_n);
  for (int64_t i = 0; i < n; ++i) {
    float t = 0.0f;
    SPDE_HIP(hipEventElapsedTime(&t, pl->prof_ev[2 * i], pl->prof_ev[2 * i + 1]));
    if (ms) ms[i] = t;
    if (flops) flops[i] = pl->prof_flops[i];
    if (bytes) bytes[i] = pl->prof_bytes[i];
    if (kind) kind[i] = pl->prof_kind[i];
    if (tiles) tiles[i] = pl->prof_tiles[i];
  }
  if (count) *count = (int64_t)pl->prof_n;
  return SPDE_OK;
}

int spde_prof_end(spde_plan* pl, double* gemm_ms, double* gemm_flops, int64_t* launches, double* busy_ms,
                  double* gemm_bytes) {
  if (!pl || pl->host_only) {
    set_error("spde_prof_end: no device plan");
    return SPDE_ERR_ARG;
  }
  pl->prof_on = false;
  double ms = 0.0, fl = 0.0, by = 0.0;
  if (pl->prof_n) SPDE_HIP(hipDeviceSynchronize());  // launches may sit on several streams
  std::vector<std::pair<float, float>> iv(pl->prof_n);
  for (size_t i = 0; i < pl->prof_n; ++i) {
    float t = 0.0f, a = 0.0f, b = 0.0f;
    SPDE_HIP(hipEventElapsedTime(&t, pl->prof_ev[2 * i], pl->prof_ev[2 * i + 1]));
    SPDE_HIP(hipEventElapsedTime(&a, pl->prof_base, pl->prof_ev[2 * i]));
    SPDE_HIP(hipEventElapsedTime(&b, pl->prof_base, pl->prof_ev[2 * i + 1]));
    iv[i] = {a, b};
    ms += t;
    fl += pl->prof_flops[i];
    by += pl->prof_bytes[i];
  }
  // time during which at least one GEMM launch was in flight (launches on
  // different streams overlap)
  std::sort(iv.begin(), iv.end());
  double busy = 0.0, cur_a = 0.0, cur_b = -1.0;
  for (auto& p : iv) {
    if (p.first > cur_b) {
      if (cur_b > cur_a) busy += cur_b - cur_a;
      cur_a = p.first;
      cur_b = p.second;
    } else if (p.second > cur_b) {
      cur_b = p.second;
    }
  }
  if (cur_b > cur_a) busy += cur_b - cur_a;
  if (gemm_ms) *gemm_ms = ms;
  if (gemm_flops) *gemm_flops = fl;
  if (launches) *launches = (int64_t)pl->prof_n;
  if (busy_ms) *busy_ms = busy;
  if (gemm_bytes) *gemm_bytes = by;
  pl->prof_n = 0;
  return SPDE_OK;
}

int spde_selinv(spde_plan* pl, int slot, double* sigma, void* stream) {
  int rc = check_slot(pl, slot);
  if (rc) return rc;
  if (!pl->slots[slot].factored) {
    set_error("spde_selinv: slot not factored");
    return SPDE_ERR_STATE;
  }
  hipStream_t s = (hipStream_t)stream;
  PlanExtra* ex = extra(pl);
  const Symbolic& sym = pl->sym;
  Slot& sl = pl->slots[slot];
  if (!sl.sigma) {
    SPDE_HIP(hipStreamSynchronize(s));
    SPDE_HIP(hipMalloc((void**)&sl.sigma, sizeof(double) * std::max<int64_t>(sym.front_doubles, 1)));
  }
  if (!pl->selinv_built) {
    rc = build_selinv_plan(pl);
    if (rc) return rc;
  }
  rc = ensure_splitws(pl, slot);
  if (rc) return rc;
  if (!ex->pre) {
    SPDE_HIP(hipStreamCreateWithFlags(&ex->pre, hipStreamNonBlocking));  // default priority
    SPDE_HIP(hipEventCreateWithFlags(&ex->ev_start, hipEventDisableTiming));
    ex->ev_pre.resize(pl->levels.size());
    for (auto& e : ex->ev_pre) SPDE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  Bases b = make_bases(pl, slot, nullptr);
  // Top-down over the levels.  X = L21 Z and Sigma11 = Z^T Z need only the
  // factor: they run on the second stream, one level ahead of the chain
  // gather -> Sigma21 = -Sigma22 X -> Sigma11 -= X^T Sigma21 on this one,
  // which waits for its level's event.  Every wait is enqueued right after
  // the record it waits on.
  SPDE_HIP(hipEventRecord(ex->ev_start, s));
  SPDE_HIP(hipStreamWaitEvent(ex->pre, ex->ev_start, 0));
  for (int l = (int)pl->levels.size() - 1; l >= 0; --l) {
    LevelPlan& L = pl->levels[l];
    LevelDev& D = ex->lvl[l];
    for (const GemmLaunch* g : {&L.si_x, &L.si_s11a}) run_gemm(pl, *g, b, ex->pre, g->tile == 128 ? 0 : kPreLds);
    SPDE_HIP(hipEventRecord(ex->ev_pre[l], ex->pre));
    launch_sigma_gather(D.d_nodes, pl->d_nodes, pl->d_relpos, sl.sigma, D.d_gtiles, D.ngtiles, s);
    SPDE_HIP(hipStreamWaitEvent(s, ex->ev_pre[l], 0));
    run_gemm(pl, L.si_s21, b, s);
    run_gemm(pl, L.si_s11b, b, s);
  }
  if (sigma) launch_gather_pattern(sl.sigma, pl->d_mapsym, (int64_t)pl->st.indices.size(), sigma, s);
  return SPDE_OK;
}

}  // extern "C"
