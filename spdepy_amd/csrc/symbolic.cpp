// Host-side symbolic analysis: structural pattern of Q, geometric nested
// dissection on the lattice, separator tree, multifrontal index sets.
//
// This replaces the fill-reducing analysis CHOLMOD runs inside
// ``cholesky(Q)`` (berild/spdepy ``advection_idiffusion2D.py:116,173``); it is
// computed once per (lattice, bc) and reused for every numeric factorisation.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>

namespace spde {

void build_structure(const Lattice& lat, Structure& s) {
  if (lat.M < 2 || lat.N < 2 || lat.T < 1) throw std::runtime_error("lattice must be at least 2x2x1");
  if (lat.bc == 2 && (lat.M < 5 || lat.N < 5))
    throw std::runtime_error("periodic lattice needs M, N >= 5");
  const int64_t Ns = lat.Ns(), n = lat.n();
  s.indptr.assign(n + 1, 0);
  s.indices.clear();
  s.code.clear();
  s.indices.reserve(n * (lat.wide ? 75 : 43));
  s.code.reserve(n * (lat.wide ? 75 : 43));
  std::vector<std::pair<int32_t, uint8_t>> row;
  for (int t = 0; t < lat.T; ++t)
    for (int j = 0; j < lat.N; ++j)
      for (int i = 0; i < lat.M; ++i) {
        row.clear();
        const int64_t r = i + (int64_t)j * lat.M + (int64_t)t * Ns;
        const int rt = lat.wide ? 2 : 1;  // reach of the coupling to the neighbouring slices
        auto tcode = [&](int base_narrow, int base_wide, int dx, int dy) {
          return (uint8_t)(lat.wide ? base_wide + (dy + 2) * 5 + dx + 2 : base_narrow + (dy + 1) * 3 + dx + 1);
        };
        if (t > 0)
          for (int dy = -rt; dy <= rt; ++dy)
            for (int dx = -rt; dx <= rt; ++dx) {
              int64_t nb = lat.nbr(i, j, dx, dy);
              if (nb >= 0) row.push_back({(int32_t)(nb + (t - 1) * Ns), tcode(kPrevBase, kPrevWide, dx, dy)});
            }
        for (int dy = -2; dy <= 2; ++dy)
          for (int dx = -2; dx <= 2; ++dx) {
            int64_t nb = lat.nbr(i, j, dx, dy);
            if (nb >= 0) row.push_back({(int32_t)(nb + t * Ns), (uint8_t)((dy + 2) * 5 + dx + 2)});
          }
        if (t + 1 < lat.T)
          for (int dy = -rt; dy <= rt; ++dy)
            for (int dx = -rt; dx <= rt; ++dx) {
              int64_t nb = lat.nbr(i, j, dx, dy);
              if (nb >= 0) row.push_back({(int32_t)(nb + (t + 1) * Ns), tcode(kNextBase, kNextWide, dx, dy)});
            }
        std::sort(row.begin(), row.end());
        for (size_t q = 1; q < row.size(); ++q)
          if (row[q].first == row[q - 1].first) throw std::runtime_error("duplicate structural column");
        for (auto& e : row) {
          s.indices.push_back(e.first);
          s.code.push_back(e.second);
        }
        s.indptr[r + 1] = (int64_t)s.indices.size();
      }
  const int64_t nnz = (int64_t)s.indices.size();
  s.diag.assign(n, -1);
  s.mirror.assign(nnz, -1);
  for (int64_t r = 0; r < n; ++r)
    for (int64_t e = s.indptr[r]; e < s.indptr[r + 1]; ++e) {
      const int64_t c = s.indices[e];
      if (c == r) s.diag[r] = e;
      auto b = s.indices.begin() + s.indptr[c], en = s.indices.begin() + s.indptr[c + 1];
      auto it = std::lower_bound(b, en, (int32_t)r);
      if (it == en || *it != r) throw std::runtime_error("structural pattern not symmetric");
      s.mirror[e] = it - s.indices.begin();
    }
}

void check_wide_layout(const Lattice& lat, const Structure& s) {
  const int64_t Ns = lat.Ns();
  for (int64_t r = 0; r < lat.n(); ++r) {
    const int t = (int)(r / Ns);
    const int64_t b = s.indptr[r], len = s.indptr[r + 1] - b;
    int64_t cnt = 0;
    for (int64_t e = b; e < b + len; ++e) cnt += s.code[e] < kDiagSlots;
    const int nblk = 1 + (t > 0) + (t + 1 < lat.T);
    if (cnt * nblk != len) throw std::runtime_error("wide-time layout: unequal slice blocks");
    const int64_t same = b + (t > 0 ? cnt : 0);
    for (int64_t q = 0; q < cnt; ++q) {
      const int64_t col = s.indices[same + q] - (int64_t)t * Ns;
      const int cd = s.code[same + q];
      if (cd >= kDiagSlots) throw std::runtime_error("wide-time layout: slice block order");
      if (t > 0 && (s.indices[b + q] - (int64_t)(t - 1) * Ns != col || s.code[b + q] != kPrevWide + cd))
        throw std::runtime_error("wide-time layout: previous slice differs");
      if (t + 1 < lat.T &&
          (s.indices[same + cnt + q] - (int64_t)(t + 1) * Ns != col || s.code[same + cnt + q] != kNextWide + cd))
        throw std::runtime_error("wide-time layout: next slice differs");
    }
  }
}

namespace {

struct TNode {
  std::vector<int> own;  // lattice indices
  int child[2] = {-1, -1};
  int nchild = 0;
};

struct Dissector {
  const Lattice& lat;
  const Structure& s;
  int leaf_max;
  std::vector<int> mark;
  int epoch = 0;
  std::vector<TNode> tn;

  Dissector(const Lattice& l, const Structure& st, int lm)
      : lat(l), s(st), leaf_max(lm), mark(l.n(), 0) {}

  int coord(int v, int axis) const {
    const int64_t Ns = lat.Ns();
    if (axis == 0) return v % lat.M;
    if (axis == 1) return (int)((v / lat.M) % lat.N);
    return (int)(v / Ns);
  }

  // vertices of `from` with a neighbour carrying mark `tag`
  void boundary(const std::vector<int>& from, int tag, std::vector<int>& out) {
    out.clear();
    for (int v : from) {
      for (int64_t e = s.indptr[v]; e < s.indptr[v + 1]; ++e)
        if (mark[s.indices[e]] == tag) {
          out.push_back(v);
          break;
        }
    }
  }

  int build(std::vector<int>& V) {
    const int id = (int)tn.size();
    tn.emplace_back();
    if ((int)V.size() <= leaf_max) {
      tn[id].own = V;
      std::sort(tn[id].own.begin(), tn[id].own.end());
      return id;
    }
    int best_axis = -1;
    size_t best_size = SIZE_MAX;
    bool best_left = true;
    std::vector<int> L, R, sL, sR, bestL, bestR, bestS;
    for (int axis = 0; axis < 3; ++axis) {
      int lo = INT32_MAX, hi = INT32_MIN;
      for (int v : V) {
        int c = coord(v, axis);
        lo = std::min(lo, c);
        hi = std::max(hi, c);
      }
      if (hi == lo) continue;
      const int mid = (lo + hi + 1) / 2;
      L.clear();
      R.clear();
      for (int v : V) (coord(v, axis) < mid ? L : R).push_back(v);
      const int tagR = ++epoch;
      for (int v : R) mark[v] = tagR;
      boundary(L, tagR, sL);
      const int tagL = ++epoch;
      for (int v : L) mark[v] = tagL;
      boundary(R, tagL, sR);
      const bool left = sL.size() <= sR.size();
      const size_t sz = left ? sL.size() : sR.size();
      if (sz < best_size) {
        best_size = sz;
        best_axis = axis;
        best_left = left;
        bestL = L;
        bestR = R;
        bestS = left ? sL : sR;
      }
    }
    if (best_axis < 0) {  // a single lattice point repeated: cannot happen
      tn[id].own = V;
      return id;
    }
    const int tagS = ++epoch;
    for (int v : bestS) mark[v] = tagS;
    std::vector<int> a, b;
    for (int v : bestL)
      if (mark[v] != tagS) a.push_back(v);
    for (int v : bestR)
      if (mark[v] != tagS) b.push_back(v);
    (void)best_left;
    std::vector<int> own = bestS;
    std::sort(own.begin(), own.end());
    std::vector<int>().swap(V);
    int ca = a.empty() ? -1 : build(a);
    int cb = b.empty() ? -1 : build(b);
    tn[id].own = std::move(own);
    if (ca >= 0) tn[id].child[tn[id].nchild++] = ca;
    if (cb >= 0) tn[id].child[tn[id].nchild++] = cb;
    return id;
  }
};

}  // namespace

void build_symbolic(const Lattice& lat, const Structure& s, int leaf_max, Symbolic& sym) {
  const int64_t n = lat.n();
  if (n > INT32_MAX) throw std::runtime_error("lattice too large for 32-bit indices");
  sym = Symbolic();
  sym.n = n;
  Dissector d(lat, s, std::max(leaf_max, 1));
  std::vector<int> all(n);
  std::iota(all.begin(), all.end(), 0);
  const int root = d.build(all);

  // postorder: children first, then own vertices
  std::vector<int> order;  // temp ids in postorder
  {
    std::vector<std::pair<int, int>> st{{root, 0}};
    while (!st.empty()) {
      auto& [u, k] = st.back();
      if (k < d.tn[u].nchild) {
        int c = d.tn[u].child[k++];
        st.push_back({c, 0});
      } else {
        order.push_back(u);
        st.pop_back();
      }
    }
  }
  std::vector<int> newid(d.tn.size(), -1);
  for (size_t q = 0; q < order.size(); ++q) newid[order[q]] = (int)q;
  sym.nodes.resize(order.size());
  sym.perm.resize(n);
  sym.iperm.resize(n);
  sym.col_node.resize(n);
  int next = 0;
  for (size_t q = 0; q < order.size(); ++q) {
    TNode& t = d.tn[order[q]];
    Node& nd = sym.nodes[q];
    nd.nchild = t.nchild;
    for (int c = 0; c < t.nchild; ++c) {
      nd.child[c] = newid[t.child[c]];
      sym.nodes[nd.child[c]].parent = (int)q;
    }
    nd.start = next;
    nd.p = (int)t.own.size();
    for (int v : t.own) {
      sym.perm[next] = v;
      sym.iperm[v] = next;
      sym.col_node[next] = (int)q;
      ++next;
    }
  }
  if (next != n) throw std::runtime_error("nested dissection lost vertices");

  // boundary sets, bottom-up
  std::vector<int> mk(n, -1);
  for (size_t q = 0; q < sym.nodes.size(); ++q) {
    Node& nd = sym.nodes[q];
    const int end = nd.start + nd.p;
    std::vector<int>& B = nd.B;
    for (int x = nd.start; x < end; ++x) {
      const int v = sym.perm[x];
      for (int64_t e = s.indptr[v]; e < s.indptr[v + 1]; ++e) {
        const int y = sym.iperm[s.indices[e]];
        if (y >= end && mk[y] != (int)q) {
          mk[y] = (int)q;
          B.push_back(y);
        }
      }
    }
    for (int c = 0; c < nd.nchild; ++c)
      for (int y : sym.nodes[nd.child[c]].B) {
        if (y < nd.start) throw std::runtime_error("separator property violated (sibling coupling)");
        if (y >= end && mk[y] != (int)q) {
          mk[y] = (int)q;
          B.push_back(y);
        }
      }
    std::sort(B.begin(), B.end());
  }
  // relative positions in the parent front, levels, offsets
  int64_t foff = 0, zoff = 0, voff = 0, roff = 0;
  for (size_t q = 0; q < sym.nodes.size(); ++q) {
    Node& nd = sym.nodes[q];
    int lvl = 0;
    for (int c = 0; c < nd.nchild; ++c) lvl = std::max(lvl, sym.nodes[nd.child[c]].level + 1);
    nd.level = lvl;
    if (nd.parent >= 0) {
      // the parent is later in postorder; its B is not known yet -> second pass
    }
    const int64_t nn = nd.n();
    nd.foff = foff;
    foff += nn * nn;
    nd.zoff = zoff;
    zoff += (int64_t)nd.p * nd.p;
    nd.voff = voff;
    voff += nn;
    nd.relpos_off = roff;
    roff += (int64_t)nd.B.size();
    sym.max_front = std::max(sym.max_front, nd.n());
    sym.max_p = std::max(sym.max_p, nd.p);
    const double p = nd.p, b = (double)nd.B.size();
    sym.flops_factor += p * p * p / 3.0 + b * p * p + b * b * p;
    sym.flops_selinv += p * p * p / 3.0 + b * p * p + 2.0 * b * b * p + p * p * p / 3.0 + b * p * p;
    sym.nnz_L += (int64_t)(p * (p + 1) / 2 + b * p);
  }
  sym.front_doubles = foff;
  sym.z_doubles = zoff;
  sym.vec_doubles = voff;
  sym.relpos_total = roff;
  for (size_t q = 0; q < sym.nodes.size(); ++q) {
    Node& nd = sym.nodes[q];
    if (nd.parent < 0) continue;
    const Node& pa = sym.nodes[nd.parent];
    const int pend = pa.start + pa.p;
    nd.relpos.resize(nd.B.size());
    for (size_t k = 0; k < nd.B.size(); ++k) {
      const int y = nd.B[k];
      if (y >= pa.start && y < pend) {
        nd.relpos[k] = y - pa.start;
      } else {
        auto it = std::lower_bound(pa.B.begin(), pa.B.end(), y);
        if (it == pa.B.end() || *it != y) throw std::runtime_error("child boundary not in parent front");
        nd.relpos[k] = pa.p + (int)(it - pa.B.begin());
      }
    }
  }
  int maxlvl = 0;
  for (auto& nd : sym.nodes) maxlvl = std::max(maxlvl, nd.level);
  sym.levels.assign(maxlvl + 1, {});
  for (size_t q = 0; q < sym.nodes.size(); ++q) sym.levels[sym.nodes[q].level].push_back((int)q);
}

void build_scatter_map(const Structure& s, const Symbolic& sym, std::vector<int64_t>& map) {
  const int64_t nnz = (int64_t)s.indices.size();
  map.assign(nnz, -1);
  for (int64_t r = 0; r + 1 < (int64_t)s.indptr.size(); ++r) {
    const int pr = sym.iperm[r];
    for (int64_t e = s.indptr[r]; e < s.indptr[r + 1]; ++e) {
      const int pc = sym.iperm[s.indices[e]];
      if (pr < pc) continue;
      const Node& nd = sym.nodes[sym.col_node[pc]];
      const int col = pc - nd.start;
      int row;
      if (pr < nd.start + nd.p) {
        row = pr - nd.start;
      } else {
        auto it = std::lower_bound(nd.B.begin(), nd.B.end(), pr);
        if (it == nd.B.end() || *it != pr) throw std::runtime_error("entry outside its front");
        row = nd.p + (int)(it - nd.B.begin());
      }
      map[e] = nd.foff + row + (int64_t)col * nd.n();
    }
  }
}

}  // namespace spde
