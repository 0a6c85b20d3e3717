// Structural pattern, nested-dissection ordering and multifrontal plan for the
// space-time precision Q on a regular M x N x T lattice.
//
// Q couples (i,j,t) with (i',j',t) inside a 5x5 spatial window (A^T Qs A with a
// 9-point A) and with (i',j',t+-1) inside a 3x3 window (Qs A), so a spatial cut
// needs a separator two cells wide and a temporal cut one slice.
#pragma once
#include <cstdint>
#include <vector>

namespace spde {

// Code of a structural entry of Q: 0..24 same time slice (5x5 window,
// (dy,dx) lexicographic), 25..33 previous slice (3x3), 34..42 next slice.
// Wide-time lattices (separable models, Qt (x) Qs) couple the neighbouring
// slices through the 5x5 window too: 43..67 previous, 68..92 next slice.
enum : int { kDiagSlots = 25, kPrevBase = 25, kNextBase = 34, kSlots = 43, kPrevWide = 43, kNextWide = 68,
             kSlotsWide = 93 };

struct Lattice {
  int M, N, T, bc;            // bc: 1 zero-flux, 2 periodic, 3 Dirichlet
  int wide = 0;               // 1: neighbouring slices coupled through the 5x5 window
  int64_t Ns() const { return (int64_t)M * N; }
  int64_t n() const { return Ns() * T; }
  // neighbour of cell (i,j) at offset (dx,dy) under the boundary rule, -1 if none
  int64_t nbr(int i, int j, int dx, int dy) const {
    int a = i + dx, b = j + dy;
    if (bc == 2) {
      a = ((a % M) + M) % M;
      b = ((b % N) + N) % N;
    } else if (a < 0 || a >= M || b < 0 || b >= N) {
      return -1;
    }
    return (int64_t)a + (int64_t)b * M;
  }
};

struct Structure {           // structural CSR of Q (both triangles)
  std::vector<int64_t> indptr;   // n+1
  std::vector<int32_t> indices;  // nnz
  std::vector<uint8_t> code;     // nnz, see enum above
  std::vector<int64_t> diag;     // position of the diagonal entry of each row
  std::vector<int64_t> mirror;   // position of (c,r) for entry (r,c)
};

struct Node {
  int start = 0, p = 0;          // own columns [start, start+p) in elimination order
  int parent = -1;
  int child[2] = {-1, -1};
  int nchild = 0;
  int level = 0;                 // 0 for leaves, 1 + max(children) otherwise
  std::vector<int> B;            // boundary rows (elimination indices), ascending
  std::vector<int> relpos;       // position of B[k] in the parent's front
  int n() const { return p + (int)B.size(); }
  int64_t foff = 0;              // front (n x n, column-major) offset in doubles
  int64_t zoff = 0;              // inverse of the diagonal block (p x p)
  int64_t relpos_off = 0;        // offset of relpos in the concatenated array
  int64_t voff = 0;              // front vector area (n x 1) for solves
};

struct Symbolic {
  int64_t n = 0;
  std::vector<int> perm;         // elimination position -> lattice index
  std::vector<int> iperm;        // lattice index -> elimination position
  std::vector<Node> nodes;       // in postorder (children before parents)
  std::vector<int> col_node;     // elimination position -> owning node
  std::vector<std::vector<int>> levels;  // node ids per level
  int64_t front_doubles = 0, z_doubles = 0, vec_doubles = 0, relpos_total = 0;
  int max_front = 0, max_p = 0;
  double flops_factor = 0, flops_selinv = 0;
  int64_t nnz_L = 0;
};

void build_structure(const Lattice& lat, Structure& s);
// Wide-time lattices: every row (t,i) holds the same spatial neighbours in
// each of its slices t-1, t, t+1, in the same order (throws otherwise).  The
// Kronecker kernels index Q's entries by that layout.
void check_wide_layout(const Lattice& lat, const Structure& s);
// Geometric nested dissection + multifrontal symbolic analysis.
// Throws std::runtime_error if the separator property is violated.
void build_symbolic(const Lattice& lat, const Structure& s, int leaf_max, Symbolic& sym);
// For each structural entry (r,c) with elimination(r) >= elimination(c): the
// offset of that entry in the front arena; -1 for the upper triangle.
void build_scatter_map(const Structure& s, const Symbolic& sym, std::vector<int64_t>& map);

}  // namespace spde
