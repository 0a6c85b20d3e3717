// Host driver for an AddressSanitizer / UndefinedBehaviorSanitizer build of
// the symbolic analysis (spdepy_amd/csrc/symbolic.cpp): structural pattern,
// nested dissection, fronts and scatter map for a set of lattices, with the
// invariants the device launches rely on checked on the host.
#include <cstdio>
#include <cstdlib>

#include "../../spdepy_amd/csrc/plan.hpp"

using namespace spde;

static int check(int M, int N, int T, int bc, int wide, int leaf) {
  Lattice lat{M, N, T, bc};
  lat.wide = wide;
  Structure st;
  Symbolic sym;
  std::vector<int64_t> map;
  build_structure(lat, st);
  if (wide) check_wide_layout(lat, st);
  build_symbolic(lat, st, leaf, sym);
  build_scatter_map(st, sym, map);
  const int64_t n = lat.n();
  if ((int64_t)st.indptr.size() != n + 1 || (int64_t)sym.perm.size() != n) return 1;
  std::vector<char> seen(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int p = sym.perm[i];
    if (p < 0 || p >= n || seen[p]) return 2;
    seen[p] = 1;
  }
  for (size_t e = 0; e < map.size(); ++e)
    if (map[e] >= sym.front_doubles) return 3;
  for (const Node& nd : sym.nodes)
    for (size_t k = 0; k < nd.relpos.size(); ++k)
      if (nd.parent >= 0 && (nd.relpos[k] < 0 || nd.relpos[k] >= sym.nodes[nd.parent].n())) return 4;
  std::printf("%dx%dx%d bc=%d wide=%d leaf=%d: n=%lld nodes=%zu nnzL=%lld ok\n", M, N, T, bc, wide, leaf,
              (long long)n, sym.nodes.size(), (long long)sym.nnz_L);
  return 0;
}

int main() {
  const int cases[][6] = {{8, 7, 4, 1, 0, 128}, {9, 9, 4, 2, 0, 64},  {12, 10, 6, 3, 0, 32},
                          {30, 30, 10, 1, 0, 128}, {9, 8, 6, 3, 1, 64}, {16, 16, 1, 1, 0, 128},
                          {5, 5, 2, 1, 0, 16}};
  for (auto& c : cases) {
    const int rc = check(c[0], c[1], c[2], c[3], c[4], c[5]);
    if (rc) {
      std::printf("invariant %d violated for %dx%dx%d\n", rc, c[0], c[1], c[2]);
      return rc;
    }
  }
  return 0;
}
