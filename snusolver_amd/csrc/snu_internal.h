// Internal host-side types of the MI355X SnuSolver path.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/snusolver.h"

namespace snu {

int set_error(int code, const std::string &msg);
uint64_t splitmix64(uint64_t &s);
double u01(uint64_t &s);

// Per-leaf symbolic structure (one leaf = one rank of the reference).
// "core" = leaf rows followed by the ancestor separators bottom-up
// (kernel_sparse.cu:526-531).
struct LeafPlan {
  int rank = 0, leaf = 0, leaf_size = 0, core_n = 0, npath = 0;
  std::vector<int> path;    // path[k] = leaf >> k
  std::vector<int> cstart;  // core-local start of path[k], npath + 1 entries
  int core_nnz = 0, coo_off = 0;
  std::vector<int> core_rowptr, core_colidx, core_map;
  int lu_nnz = 0;
  std::vector<int> lu_rowptr, lu_colidx, lu_diag, lu_bias;
  // GPU-side views
  std::vector<int> wcol;    // workspace column of each LU entry
  int nbr = 0, nbc = 0;     // boundary rows (non-empty L_SL) / cols (U_LS)
  std::vector<int> brow;    // core rows of the boundary rows
  std::vector<int> bcol;    // core cols of the boundary cols
  std::vector<int> lev_ptr, lev_rows;    // factor / forward level sets
  std::vector<int> ulev_ptr, ulev_rows;  // backward level sets
  double flops = 0;         // structural leaf flops (LU + Schur)
  // multifrontal form of the leaf factorization (build_fronts), fronts
  // ordered by their largest pivot (children before parents):
  std::vector<int> fkp, fk;              // pivots K of each front (ascending)
  std::vector<int> fnk, ff;              // pivots, front order
  std::vector<int> fpar, fheight;        // parent front (-1: root), height
  std::vector<int> fidx_ptr, fidx;       // B of each front: core indices beyond K
  std::vector<int> frel;                 // per B entry: local index in the parent
  std::vector<int> fslot_ptr, fslot, fpos;  // owned LU slots, front position (col-major)
  std::vector<int> fchild_ptr, fchild;   // children, ascending
  std::vector<int> fsrow, fscol;         // root fronts: per B entry, boundary row / col (-1)
  long long fentries = 0;                // sum of front order squared
  int fheight_max = 0;
};

struct Plan {
  int n = 0, nnz = 0, np = 0, num_block = 0, max_level = 0, nk = 0;
  std::vector<int> order, sizes;
  std::vector<int> block_start, block_size, level;  // index 0..num_block
  std::vector<int> pa_rowptr, pa_colidx, perm_map;
  std::vector<int> a_rowptr, a_colidx;  // A's own pattern (FP32 refinement only)
  std::vector<int> coo_r, coo_c;
  std::vector<int> gnnz, goff;    // merged COO layout per grid id
  std::vector<int> mstart, msize; // merge region of each block
  std::vector<LeafPlan> leaves;   // by rank
  // dense single-copy separator storage:
  //   R_p = [D_p | Up_p], s_p x (s_p + nU_p), ld s_p (ancestors bottom-up)
  //   Lp_p, nL_p x s_p, ld nL_p
  // with the panels tile-compressed: of each ancestor q, the 64-row (Lp) /
  // 64-column (Up) tiles aligned to q's first index that hold a structural
  // nonzero; lmap_p / umap_p give the stored row / column of each ancestor-
  // space index (ancestors bottom-up, A_p entries; -1: not stored)
  std::vector<int> anc_sum;          // A_p
  std::vector<int> nL, nU;           // stored rows of Lp_p, stored columns of Up_p
  std::vector<std::vector<int>> lmap, umap;
  std::vector<int64_t> offR, offL;   // per separator, in the dense arena
  int64_t dense_total = 0;
  std::vector<int64_t> leaf_off;     // per rank, in the leaf value arena
  int64_t leaf_total = 0;
  std::vector<int64_t> a_dest;       // per original nnz: value arena index
                                     // (leaf arena, then dense arena)
  int grid_id(int i, int j) const {
    int x = i > j ? i : j, y = i < j ? i : j, k = 0;
    for (int jj = x; jj > y; jj /= 2) k++;
    return (x * nk + k) * 2 + (i >= j ? 0 : 1);
  }
  // offset of ancestor q inside p's ancestor list (bottom-up)
  int anc_off(int p, int q) const {
    int s = 0;
    for (int a = p / 2; a > q; a /= 2) s += block_size[a];
    return s;
  }
  // value-arena index of separator entry (bi, lr) x (bj, lc), -1 when it
  // lies in a panel tile that is not stored (structurally zero)
  int64_t sep_index(int bi, int lr, int bj, int lc) const;
  // Structural nonzeros of the dense separator panels: for separator p,
  // nzL[p][r] / nzU[p][c] flag the rows of Lp_p / columns of Up_p (ancestors
  // stacked bottom-up, A_p each) that can be nonzero; cntL / cntU are their
  // prefix counts (A_p + 1 entries).
  std::vector<std::vector<int>> cntL, cntU;
  bool anyL(int p, int a, int b) const { return b > a && cntL[p][b] > cntL[p][a]; }
  bool anyU(int p, int a, int b) const { return b > a && cntU[p][b] > cntU[p][a]; }
  double f_dense = 0, f_leaf = 0, b_leaf = 0, b_solve = 0;
};

int analyze(const snu_csr_t *A, const snu_opts_t *opts, Plan &P);
// Multifrontal structure of a leaf from its core pattern and LU pattern
// (boundary rows / cols set).
int build_fronts(LeafPlan &L);

int order_grid(int nx, int ny, int nz, int np, int *order, int *sizes);
// method 0: the smaller of the multilevel and the level-set separator at
// every bisection, 1: level-set separators, 2: multilevel separators
int order_nd(int n, const int *rowptr, const int *colidx, int np,
             uint64_t seed, int *order, int *sizes, int method = 0);
// Multilevel vertex separator of a symmetric graph without self loops (CSR):
// part[v] = 0 or 1 (the two sides, no edge between them) or 2 (separator).
void ml_separator(int n, const std::vector<int> &ptr, const std::vector<int> &adj, uint64_t seed,
                  std::vector<int> &part);

}  // namespace snu
