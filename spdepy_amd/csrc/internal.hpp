// Internal plan object shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <map>
#include <vector>

#include "../../include/spdepy_amd.h"
#include "plan.hpp"

namespace spde {

void set_error(const std::string& msg);

#define SPDE_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::spde::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
      return SPDE_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

constexpr int NB = 64;  // diagonal block width of the dense partial factorisation
constexpr int kTmpl = 143;  // per-cell assembly template (doubles)

// Operand base pointers of a batched launch (selected per operand by id).
enum Base : int8_t { kFront = 0, kZ = 1, kSigma = 2, kWork = 3, kVec = 4, kX = 5, kNumBases = 6 };
struct Bases {
  double* p[kNumBases];
  double* ws;  // split-K partial tiles of the slot
};

// One GEMM problem: C = alpha op(A) op(B) + beta C, column-major.
struct GemmProb {
  int64_t a_off, b_off, c_off;
  int32_t lda, ldb, ldc;
  int32_t M, N, K;
  int8_t a_base, b_base, c_base, lower;
  // triangular operand, zeros skipped: 1 op(B) lower (k >= first column of
  // the tile), 2 op(A) upper (k >= first row), 3 op(A) lower (k < last row + 1)
  int8_t kmode;
  double alpha, beta;
};
struct GemmTile {
  int32_t prob, tm, tn, pad;
};

// A batched GEMM launch: problems + tile list (device), transposition flags.
struct GemmLaunch {
  bool ta = false, tb = false;
  int tile = 64;                 // block tile edge (64 or 128), chosen when finalised
  int ntiles = 0, nprobs = 0;
  GemmProb* d_probs = nullptr;
  GemmTile* d_tiles = nullptr;
  std::vector<GemmProb> h_probs;
  std::vector<GemmTile> h_tiles;
  double flops = 0;
  double bytes = 0;   // algorithmic HBM bytes: A, B, C read once (C only if beta != 0), C written once
  // split-K (few tiles, deep K): each tile's K range is cut into `split`
  // chunks computed by separate blocks into partial tiles at ws + ws_off (in
  // the slot's split workspace), summed in chunk order by a second kernel
  int split = 1;
  int64_t ws_off = 0;
  // few right-hand sides (N <= kGemvMaxN, op(B) = B, general C): a
  // bandwidth-bound kernel reading A once instead of 64-wide MFMA tiles
  bool gemv = false;
  bool gemv_wide = false;        // narrow launch with deep K: sixteen waves share each row block's K range
};
// One column only: at 100x100x20 a full solve takes 2.5 ms with the
// bandwidth-bound kernel for one right-hand side, but from two right-hand
// sides on the MFMA tiles are faster (3.2 ms for two, 4.9 ms for eight,
// against 3.8 and 10.6 ms with the per-column kernel).
constexpr int kGemvMaxN = 1;

struct PotrfDesc {      // one diagonal block of one front
  int64_t foff;         // front offset
  int64_t zoff;         // Z offset of the block's node
  int32_t ld, zld;      // front dimension, p
  int32_t k0, kb;       // block start (within own columns) and width
  int32_t node, colbase;
};
struct ExtAddDesc {     // extend-add of a child's update block into its parent
  int64_t cfoff, pfoff, relpos_off;
  int32_t cld, cp, pld, b;
  int32_t split, pad;   // child columns [0, split) land in the parent's own columns
};
struct ExtAddTile {
  int32_t desc, ti, tj, phase;  // phase 0: columns < split, 1: columns >= split
};

struct Step {           // one stage of the dense partial factorisation of a level
  int npotrf = 0;
  PotrfDesc* d_potrf = nullptr;
  GemmLaunch panel;     // rows below the diagonal blocks: X = S W^T, in place
  GemmLaunch update;    // trailing update of the nodes' own columns
};

struct LevelPlan {
  std::vector<int> nodes;
  // extend-add per child slot and phase: phase 0 (into the own columns) before
  // the level's factorisation, phase 1 (into the update blocks) after its
  // SYRK, which writes the update blocks without reading them
  int next[2][2] = {{0, 0}, {0, 0}};    // [phase][slot] tile counts
  ExtAddDesc* d_ext[2] = {nullptr, nullptr};
  ExtAddTile* d_exttile[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  std::vector<ExtAddDesc> h_ext[2];
  std::vector<ExtAddTile> h_exttile[2][2];
  std::vector<Step> steps;
  GemmLaunch syrk;
  // solves / selinv
  GemmLaunch fwd_y, fwd_u, bwd_t, bwd_x;   // the current nrhs' launches (owned by solve_cache)
  GemmLaunch si_x, si_s21, si_s11a, si_s11b;
  int64_t si_work = 0;
};

struct Slot {
  double* front = nullptr;  // factor fronts
  double* z = nullptr;      // inverses of the diagonal blocks of L
  double* sigma = nullptr;  // selected-inverse fronts (allocated on first use)
  double* work = nullptr;   // inverse-of-L11 and selected-inverse temporaries
  double* partial = nullptr;  // log-det partial sums per 256 variables
  double* ld = nullptr;     // device log-det
  int* status = nullptr;    // first non-positive pivot (elimination column + 1)
  double* vec = nullptr;    // solve front vectors
  int64_t vec_cap = 0;      // in doubles
  double* xy = nullptr;     // elimination-ordered solve buffer
  int64_t xy_cap = 0;
  hipStream_t stream = nullptr;  // stream of the last factorisation
  bool factored = false;
  double* splitws = nullptr;  // split-K partial tiles: [factor + selected inverse | solve] regions
  int64_t splitws_cap = 0;
};

struct NodeDev {            // per-node table on the device
  int64_t foff, zoff, voff, relpos_off;
  int32_t start, p, n, parent, nb, level, pad0, pad1;
};

}  // namespace spde

struct spde_plan {
  spde::Lattice lat;
  spde::Structure st;
  spde::Symbolic sym;
  std::vector<int64_t> map;      // scatter map (front offsets, -1 upper)
  std::vector<int64_t> mapsym;   // map of the lower mirror for every entry
  bool host_only = true;
  int nslots = 0;
  // device structure
  int64_t* d_indptr = nullptr;
  int32_t* d_indices = nullptr;
  uint8_t* d_code = nullptr;
  int64_t* d_diag = nullptr;
  int64_t* d_map = nullptr;
  int64_t* d_mapsym = nullptr;
  int32_t* d_perm = nullptr;
  int32_t* d_relpos = nullptr;
  int32_t* d_bnd = nullptr;       // concatenated boundary lists (elimination indices)
  spde::NodeDev* d_nodes = nullptr;
  double* d_tmpl = nullptr;       // per-cell value templates of the blocks of Q
  double* d_acan = nullptr;       // canonical 9-point spatial operator A
  double* d_a0can = nullptr;      // canonical 9-point initial-state operator A0
  double* d_qdiag = nullptr;      // diagonal of Qs per cell
  double* d_win = nullptr;        // per-cell gradient windows
  double* d_red = nullptr;        // one device scalar (reductions)
  int32_t* d_pos25 = nullptr;     // one-slice plans: entry index of (cell, 5x5 code), -1 if absent
  int64_t split_fs = 0;           // split-K workspace (doubles) of the factor and selected-inverse launches
  int64_t split_solve = 0;        // ... of the solve launches (current nrhs)
  double* d_gpart = nullptr;      // per (cell, slice) partial windows of the gradient reduction
  void* llws = nullptr;           // buffers and side stream of spde_loglik_grad (loglik.hip)
  void* extra = nullptr;          // device launch lists of the factor and selected inverse (numeric.hip)
  double* d_tan = nullptr;        // tangent assembly: templates + dA, dA0, dq of one direction
  int32_t* d_sprow = nullptr;     // one-slice plans: row and column of each entry (Kronecker kernels)
  int32_t* d_spcol = nullptr;
  int64_t work_doubles = 0;       // per-slot workspace size
  std::vector<spde::Slot> slots;
  std::vector<spde::LevelPlan> levels;
  std::vector<spde::GemmLaunch> trtri;  // per recursion height: pairs (X, Z2)
  int trtri_heights = 0;
  int solve_nrhs = -1;            // nrhs the solve launches are built for
  // solve launches per nrhs (four per level), built once each and swapped in
  // by value; their device descriptor arrays are freed with the plan
  struct SolveSet {
    std::vector<spde::GemmLaunch> launches;
    int64_t split = 0;
  };
  std::map<int, SolveSet> solve_cache;
  bool selinv_built = false;
  // optional event timing of the GEMM launches (measurement only)
  bool prof_on = false;
  int prof_mask = 0xff;           // kinds of GEMM launch timed (bit = kind code)
  std::vector<hipEvent_t> prof_ev;
  hipEvent_t prof_base = nullptr;
  std::vector<double> prof_flops;
  std::vector<double> prof_bytes;
  std::vector<int> prof_tiles;
  std::vector<int> prof_kind;
  size_t prof_n = 0;
  hipStream_t prof_stream = nullptr;
};
