// Structure analysis of the construct phase (host, integer work).
//
// Produces, bit-exactly, the reference's integer structures for one process
// that owns every leaf:
//   block layout        calculate_block           SnuMat_construct.cu:21-38
//   PAP^T + perm_map    permutate                 SnuMat_construct.cu:40-80
//   merged COO layout   construct/setGrid/addElement  SnuMat_construct.cu:81-159
//   per-leaf core CSR   core_preprocess           kernel_sparse.cu:526-573
//   symbolic LU         core_symbfact             kernel_sparse.cu:368-524
//   core_map, LU_bias   core_preprocess           kernel_sparse.cu:576-597
// and the MI355X layout: one copy of every separator block (no per-rank
// replication), 64-bit offsets, the leaf value arena, the map from each
// entry of A to its slot, and per-leaf level sets for the GPU kernels.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <numeric>
#include <thread>

#include "snu_internal.h"

namespace snu {

int64_t Plan::sep_index(int bi, int lr, int bj, int lc) const {
  if (bi == bj) return offR[bi] + lr + (int64_t)lc * block_size[bi];
  if (bi > bj) {  // row block deeper: a stored column of the U panel of bi
    const int c = umap[bi][anc_off(bi, bj) + lc];
    return c < 0 ? -1 : offR[bi] + lr + (int64_t)(block_size[bi] + c) * block_size[bi];
  }
  const int r = lmap[bj][anc_off(bj, bi) + lr];  // a stored row of the L panel of bj
  return r < 0 ? -1 : offL[bj] + r + (int64_t)lc * nL[bj];
}

namespace {

int ilog2(int v) {
  int l = 0;
  while (v > 1) v >>= 1, l++;
  return l;
}

void dfs_block_order(std::vector<int> &v, int cur, int parts) {
  if (parts > 1) {
    dfs_block_order(v, cur * 2 + 1, parts / 2);
    dfs_block_order(v, cur * 2, parts / 2);
  }
  v.push_back(cur);
}

// core_symbfact (kernel_sparse.cu:368-524): filled pattern of every core row;
// separator rows close over leaf columns only.
void symbfact(LeafPlan &L) {
  int core_n = L.core_n, ls = L.leaf_size;
  L.lu_rowptr.assign(core_n + 1, 0);
  L.lu_diag.assign(core_n, 0);
  L.lu_colidx.clear();
  L.lu_colidx.reserve((size_t)L.core_nnz * 4 + 16);
  std::vector<char> vst(core_n + 1, 0);
  std::vector<int> Lj, Uj;
  Lj.reserve(core_n);
  Uj.reserve(core_n);
  for (int j = 0; j < core_n; j++) {
    Lj.clear();
    Uj.clear();
    for (int i = L.core_rowptr[j]; i < L.core_rowptr[j + 1]; i++) {
      int h = L.core_colidx[i];
      if (h < j) Lj.push_back(h);
      else Uj.push_back(h);
      vst[h] = 1;
    }
    for (size_t idx = 0; idx < Lj.size(); idx++) {
      int i = Lj[idx];
      for (int u = L.lu_diag[i]; u < L.lu_rowptr[i + 1]; u++) {
        int h = L.lu_colidx[u];
        if (j >= ls && h >= ls) break;
        if (vst[h]) continue;
        vst[h] = 1;
        if (h < j) Lj.push_back(h);
        else Uj.push_back(h);
      }
    }
    std::sort(Lj.begin(), Lj.end());
    std::sort(Uj.begin(), Uj.end());
    for (int h : Lj) L.lu_colidx.push_back(h), vst[h] = 0;
    L.lu_diag[j] = (int)L.lu_colidx.size();
    for (int h : Uj) L.lu_colidx.push_back(h), vst[h] = 0;
    L.lu_rowptr[j + 1] = (int)L.lu_colidx.size();
  }
  L.lu_nnz = (int)L.lu_colidx.size();
}

int build_leaf(const Plan &P, LeafPlan &L) {
  const int leaf = L.leaf;
  L.npath = 0;
  for (int i = leaf; i >= 1; i /= 2) L.npath++;
  L.path.resize(L.npath);
  L.cstart.resize(L.npath + 1);
  int sum = 0;
  for (int k = 0; k < L.npath; k++) {
    L.path[k] = leaf >> k;
    L.cstart[k] = sum;
    sum += P.block_size[leaf >> k];
  }
  L.cstart[L.npath] = sum;
  L.core_n = sum;
  L.leaf_size = P.block_size[leaf];
  L.coo_off = P.mstart[leaf];
  L.core_nnz = P.msize[leaf];
  const int core_n = L.core_n, ls = L.leaf_size;
  // core blocks in merge order: (leaf,leaf), (leaf,p1), (p1,leaf), ...
  struct Cb { int bi, bj, ki, kj; };
  std::vector<Cb> cb;
  cb.push_back({leaf, leaf, 0, 0});
  for (int k = 1; k < L.npath; k++) {
    cb.push_back({leaf, leaf >> k, 0, k});
    cb.push_back({leaf >> k, leaf, k, 0});
  }
  // core CSR: rows keep the merge order, which is column-ascending
  L.core_rowptr.assign(core_n + 1, 0);
  L.core_colidx.assign(L.core_nnz, 0);
  for (const Cb &e : cb) {
    int g = P.grid_id(e.bi, e.bj);
    for (int p = 0; p < P.gnnz[g]; p++) {
      int q = P.goff[g] + p;
      L.core_rowptr[P.coo_r[q] - P.block_start[e.bi] + L.cstart[e.ki] + 1]++;
    }
  }
  for (int i = 0; i < core_n; i++) L.core_rowptr[i + 1] += L.core_rowptr[i];
  {
    std::vector<int> fillp(L.core_rowptr.begin(), L.core_rowptr.end() - 1);
    for (const Cb &e : cb) {
      int g = P.grid_id(e.bi, e.bj);
      for (int p = 0; p < P.gnnz[g]; p++) {
        int q = P.goff[g] + p;
        int r = P.coo_r[q] - P.block_start[e.bi] + L.cstart[e.ki];
        int c = P.coo_c[q] - P.block_start[e.bj] + L.cstart[e.kj];
        L.core_colidx[fillp[r]++] = c;
      }
    }
  }
  symbfact(L);
  for (int i = 0; i < ls; i++)
    if (L.lu_diag[i] >= L.lu_rowptr[i + 1] || L.lu_colidx[L.lu_diag[i]] != i)
      return set_error(SNU_ERR_ZERO_DIAG, "structurally zero diagonal in leaf " + std::to_string(leaf));
  // core_map: merged-COO entry of the leaf -> LU slot
  L.core_map.assign(L.core_nnz, 0);
  {
    std::vector<int> cur(L.lu_rowptr.begin(), L.lu_rowptr.end() - 1);
    for (const Cb &e : cb) {
      int g = P.grid_id(e.bi, e.bj);
      for (int p = 0; p < P.gnnz[g]; p++) {
        int q = P.goff[g] + p;
        int r = P.coo_r[q] - P.block_start[e.bi] + L.cstart[e.ki];
        int c = P.coo_c[q] - P.block_start[e.bj] + L.cstart[e.kj];
        while (L.lu_colidx[cur[r]] != c) cur[r]++;
        L.core_map[q - L.coo_off] = cur[r];
      }
    }
  }
  L.lu_bias.assign(ls, 0);
  for (int i = 0; i < ls; i++) {
    int ptr = L.lu_rowptr[i];
    while (ptr < L.lu_rowptr[i + 1] && L.lu_colidx[ptr] < ls) ptr++;
    L.lu_bias[i] = ptr;
  }
  // boundary rows / cols and the workspace column of every LU entry
  int nsep = core_n - ls;
  std::vector<int> binv(nsep, -1);
  for (int r = 0; r < ls; r++)
    for (int p = L.lu_bias[r]; p < L.lu_rowptr[r + 1]; p++) binv[L.lu_colidx[p] - ls] = 0;
  L.bcol.clear();
  for (int c = 0; c < nsep; c++)
    if (binv[c] == 0) binv[c] = (int)L.bcol.size(), L.bcol.push_back(c + ls);
  L.nbc = (int)L.bcol.size();
  L.brow.clear();
  for (int r = ls; r < core_n; r++)
    if (L.lu_rowptr[r + 1] > L.lu_rowptr[r]) L.brow.push_back(r);
  L.nbr = (int)L.brow.size();
  L.wcol.resize(L.lu_nnz);
  for (int p = 0; p < L.lu_nnz; p++) {
    int c = L.lu_colidx[p];
    L.wcol[p] = c < ls ? c : ls + binv[c - ls];
  }
  // level sets of the leaf rows: factor/forward (L deps), backward (U deps)
  std::vector<int> lev(ls, 0);
  int nlev = 0;
  for (int j = 0; j < ls; j++) {
    int m = 0;
    for (int p = L.lu_rowptr[j]; p < L.lu_diag[j]; p++) m = std::max(m, lev[L.lu_colidx[p]] + 1);
    lev[j] = m;
    nlev = std::max(nlev, m + 1);
  }
  L.lev_ptr.assign(nlev + 1, 0);
  for (int j = 0; j < ls; j++) L.lev_ptr[lev[j] + 1]++;
  for (int l = 0; l < nlev; l++) L.lev_ptr[l + 1] += L.lev_ptr[l];
  L.lev_rows.resize(ls);
  {
    std::vector<int> f(L.lev_ptr.begin(), L.lev_ptr.end() - 1);
    for (int j = 0; j < ls; j++) L.lev_rows[f[lev[j]]++] = j;
  }
  std::vector<int> ulev(ls, 0);
  int nulev = 0;
  for (int j = ls - 1; j >= 0; j--) {
    int m = 0;
    for (int p = L.lu_diag[j] + 1; p < L.lu_bias[j]; p++) m = std::max(m, ulev[L.lu_colidx[p]] + 1);
    ulev[j] = m;
    nulev = std::max(nulev, m + 1);
  }
  L.ulev_ptr.assign(nulev + 1, 0);
  for (int j = 0; j < ls; j++) L.ulev_ptr[ulev[j] + 1]++;
  for (int l = 0; l < nulev; l++) L.ulev_ptr[l + 1] += L.ulev_ptr[l];
  L.ulev_rows.resize(ls);
  {
    std::vector<int> f(L.ulev_ptr.begin(), L.ulev_ptr.end() - 1);
    for (int j = ls - 1; j >= 0; j--) L.ulev_rows[f[ulev[j]]++] = j;
  }
  int frc = build_fronts(L);
  if (frc) return frc;
  // structural flops: every pivot of a row costs 1 division and 2 flops per
  // U entry of the pivot row it applies
  double fl = 0;
  for (int i = 0; i < core_n; i++) {
    int lend = i < ls ? L.lu_diag[i] : L.lu_rowptr[i + 1];
    for (int p = L.lu_rowptr[i]; p < lend; p++) {
      int k = L.lu_colidx[p];
      fl += 1.0 + 2.0 * (L.lu_rowptr[k + 1] - L.lu_diag[k] - 1);
    }
  }
  L.flops = fl;
  return SNU_OK;
}

template <class F>
void parallel_for(int n, F f) {
  int nt = (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 16) nt = 16;
  if (nt > n) nt = n;
  if (nt <= 1) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&]() {
      for (;;) {
        int i = next++;
        if (i >= n) break;
        f(i);
      }
    });
  for (auto &t : th) t.join();
}

}  // namespace

int analyze(const snu_csr_t *A, const snu_opts_t *opts, Plan &P) {
  if (!A || !opts) return set_error(SNU_ERR_BAD_INPUT, "null matrix or options");
  const int n = A->n, nnz = A->nnz, np = opts->np;
  if (n < 1 || A->m != n) return set_error(SNU_ERR_BAD_INPUT, "matrix must be square with n >= 1");
  if (nnz < 0 || !A->rowptr || (nnz && (!A->colidx || !A->data)))
    return set_error(SNU_ERR_BAD_INPUT, "null CSR arrays");
  if (np < 1 || (np & (np - 1)) || np > (1 << 20)) return set_error(SNU_ERR_BAD_INPUT, "np must be a power of two");
  if (A->rowptr[0] != 0 || A->rowptr[n] != nnz) return set_error(SNU_ERR_BAD_INPUT, "rowptr inconsistent with nnz");
  // Rows may come in any column order (the reference's reader keeps the
  // file order, src/read_matrix.cpp:28-35, and its permutate sorts,
  // src/SnuMat_construct.cu:40-80): nothing below needs sorted input rows.
  // Duplicate (row, col) entries are rejected.
  {
    std::vector<int> tmp;
    for (int i = 0; i < n; i++) {
      if (A->rowptr[i + 1] < A->rowptr[i]) return set_error(SNU_ERR_BAD_INPUT, "rowptr not monotone");
      bool sorted = true;
      for (int q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) {
        if (A->colidx[q] < 0 || A->colidx[q] >= n) return set_error(SNU_ERR_BAD_INPUT, "column index out of range");
        if (q > A->rowptr[i] && A->colidx[q] <= A->colidx[q - 1]) {
          if (A->colidx[q] == A->colidx[q - 1]) return set_error(SNU_ERR_DUPLICATE, "duplicate entry in a row");
          sorted = false;
        }
      }
      if (!sorted) {
        tmp.assign(A->colidx + A->rowptr[i], A->colidx + A->rowptr[i + 1]);
        std::sort(tmp.begin(), tmp.end());
        if (std::adjacent_find(tmp.begin(), tmp.end()) != tmp.end())
          return set_error(SNU_ERR_DUPLICATE, "duplicate entry in a row");
      }
    }
  }
  P.n = n; P.nnz = nnz; P.np = np;
  if (opts->precision != SNU_FP64 && opts->precision != SNU_FP32)
    return set_error(SNU_ERR_BAD_INPUT, "unknown precision kind");
  if (opts->precision == SNU_FP32) {  // pattern of A for the FP64 residuals
    P.a_rowptr.assign(A->rowptr, A->rowptr + n + 1);
    P.a_colidx.assign(A->colidx, A->colidx + nnz);
  }
  P.num_block = 2 * np - 1;
  P.max_level = ilog2(np);
  P.nk = P.max_level + 1;
  const int nb = P.num_block;
  // ---- ordering (the ParMETIS contract) ----
  P.order.assign(n, 0);
  P.sizes.assign(nb, 0);
  int rc = SNU_OK;
  if (opts->ordering == SNU_ORDER_GIVEN) {
    if (!opts->order || !opts->sizes) return set_error(SNU_ERR_BAD_INPUT, "given ordering without arrays");
    std::copy(opts->order, opts->order + n, P.order.begin());
    std::copy(opts->sizes, opts->sizes + nb, P.sizes.begin());
  } else if (opts->ordering == SNU_ORDER_GRID) {
    long long cells = (long long)opts->grid[0] * opts->grid[1] * opts->grid[2];
    if (cells != n) return set_error(SNU_ERR_BAD_INPUT, "grid dims do not match n");
    rc = order_grid(opts->grid[0], opts->grid[1], opts->grid[2], np, P.order.data(), P.sizes.data());
  } else if (opts->ordering == SNU_ORDER_ND || opts->ordering == SNU_ORDER_ND_LEVELSET) {
    rc = order_nd(n, A->rowptr, A->colidx, np, opts->seed, P.order.data(), P.sizes.data(),
                  opts->ordering == SNU_ORDER_ND ? 0 : 1);  // level sets only for ND_LEVELSET
  } else {
    return set_error(SNU_ERR_BAD_INPUT, "unknown ordering kind");
  }
  if (rc) return rc;
  {
    std::vector<char> seen(n, 0);
    long long s = 0;
    for (int i = 0; i < n; i++) {
      int o = P.order[i];
      if (o < 0 || o >= n || seen[o]) return set_error(SNU_ERR_BAD_INPUT, "order is not a permutation");
      seen[o] = 1;
    }
    for (int e = 0; e < nb; e++) {
      if (P.sizes[e] < 0) return set_error(SNU_ERR_BAD_INPUT, "negative block size");
      s += P.sizes[e];
    }
    if (s != n) return set_error(SNU_ERR_BAD_INPUT, "block sizes do not sum to n");
  }
  // ---- calculate_block ----
  P.block_start.assign(nb + 1, 0);
  P.block_size.assign(nb + 1, 0);
  P.level.assign(nb + 1, 0);
  {
    std::vector<int> bo;
    dfs_block_order(bo, 1, np);
    int cur = 0;
    for (int e : bo) {
      P.block_start[e] = cur;
      P.block_size[e] = P.sizes[np + np - 1 - e];
      cur += P.block_size[e];
    }
    for (int e = 1; e <= nb; e++) P.level[e] = ilog2(e);
  }
  // ---- permutate: PAP^T rows sorted by new column; perm_map ----
  P.pa_rowptr.assign(n + 1, 0);
  P.pa_colidx.assign(nnz, 0);
  P.perm_map.assign(nnz, 0);
  for (int i = 0; i < n; i++) P.pa_rowptr[P.order[i] + 1] = A->rowptr[i + 1] - A->rowptr[i];
  for (int i = 0; i < n; i++) P.pa_rowptr[i + 1] += P.pa_rowptr[i];
  {
    std::vector<std::pair<int, int>> v;
    for (int i = 0; i < n; i++) {
      v.clear();
      for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; j++) v.push_back({P.order[A->colidx[j]], j});
      std::sort(v.begin(), v.end());
      int base = P.pa_rowptr[P.order[i]];
      for (size_t jj = 0; jj < v.size(); jj++) {
        P.pa_colidx[base + jj] = v[jj].first;
        P.perm_map[v[jj].second] = base + (int)jj;
      }
    }
  }
  // ---- construct: merged COO layout ----
  std::vector<int> block(n);
  for (int b = 1; b <= nb; b++)
    for (int i = P.block_start[b]; i < P.block_start[b] + P.block_size[b]; i++) block[i] = b;
  for (int r = 0; r < n; r++)
    for (int q = P.pa_rowptr[r]; q < P.pa_rowptr[r + 1]; q++) {
      int x = block[r], y = block[P.pa_colidx[q]];
      if (x < y) std::swap(x, y);
      while (x > y) x /= 2;
      if (x != y) return set_error(SNU_ERR_STRUCTURE, "ordering violates the separator property");
    }
  const int ng = (nb + 1) * P.nk * 2;
  P.gnnz.assign(ng, 0);
  P.goff.assign(ng, 0);
  P.mstart.assign(nb + 1, 0);
  P.msize.assign(nb + 1, 0);
  for (int r = 0; r < n; r++)
    for (int q = P.pa_rowptr[r]; q < P.pa_rowptr[r + 1]; q++)
      P.gnnz[P.grid_id(block[r], block[P.pa_colidx[q]])]++;
  {
    std::vector<int> gend(ng, 0);
    int bias = 0;
    auto set = [&](int i, int j) {
      int g = P.grid_id(i, j);
      P.goff[g] = bias;
      bias += P.gnnz[g];
      gend[g] = bias;
    };
    for (int i = nb; i >= 1; i--) {
      P.mstart[i] = bias;
      set(i, i);
      for (int ii = i / 2; ii; ii /= 2) set(i, ii), set(ii, i);
      P.msize[i] = bias - P.mstart[i];
    }
    P.coo_r.assign(nnz, 0);
    P.coo_c.assign(nnz, 0);
    std::vector<int> cmap(nnz);
    for (int r = n - 1; r >= 0; r--)
      for (int idx = P.pa_rowptr[r + 1] - 1; idx >= P.pa_rowptr[r]; idx--) {
        int c = P.pa_colidx[idx];
        int pos = --gend[P.grid_id(block[r], block[c])];
        P.coo_r[pos] = r;
        P.coo_c[pos] = c;
        cmap[idx] = pos;
      }
    for (int i = 0; i < nnz; i++) P.perm_map[i] = cmap[P.perm_map[i]];
  }
  // ---- per-leaf structures (leaf-parallel) ----
  P.leaves.assign(np, LeafPlan());
  std::vector<int> rcs(np, SNU_OK);
  std::vector<std::string> msgs(np);
  parallel_for(np, [&](int r) {
    LeafPlan &L = P.leaves[r];
    L.rank = r;
    L.leaf = np + np - 1 - r;  // SnuMat_distribute_structure.cu:142
    rcs[r] = build_leaf(P, L);
  });
  for (int r = 0; r < np; r++)
    if (rcs[r]) return set_error(rcs[r], "structurally zero diagonal in leaf of rank " + std::to_string(r));
  P.anc_sum.assign(nb + 1, 0);
  for (int p = 1; p < np; p++) {
    int a = 0;
    for (int q = p / 2; q; q /= 2) a += P.block_size[q];
    P.anc_sum[p] = a;
  }
  P.leaf_off.assign(np + 1, 0);
  for (int r = 0; r < np; r++) P.leaf_off[r + 1] = P.leaf_off[r] + P.leaves[r].lu_nnz;
  P.leaf_total = P.leaf_off[np];
  // ---- structural nonzeros of the separator panels ----
  // Sources: A's separator-separator entries, every leaf's boundary Schur
  // block (its boundary rows x boundary cols), and the Schur update of each
  // separator into its ancestors (rows nonzero in its L panel x columns
  // nonzero in its U panel), propagated bottom-up.  The triangular solves of
  // the panels keep zero rows (L panel) and zero columns (U panel) zero.
  {
    std::vector<std::vector<char>> nzL(nb + 1), nzU(nb + 1);
    for (int p = 1; p < np; p++) nzL[p].assign(P.anc_sum[p], 0), nzU[p].assign(P.anc_sum[p], 0);
    for (int i = 1; i < np; i++)
      for (int ii = i / 2; ii; ii /= 2) {
        int g = P.grid_id(i, ii), off = P.anc_off(i, ii);
        for (int q = P.goff[g]; q < P.goff[g] + P.gnnz[g]; q++)
          nzU[i][off + P.coo_c[q] - P.block_start[ii]] = 1;
        g = P.grid_id(ii, i);
        for (int q = P.goff[g]; q < P.goff[g] + P.gnnz[g]; q++)
          nzL[i][off + P.coo_r[q] - P.block_start[ii]] = 1;
      }
    for (const LeafPlan &L : P.leaves) {
      std::vector<std::vector<int>> rows(L.npath), cols(L.npath);
      auto kof = [&](int c) {
        int k = 1;
        while (c >= L.cstart[k + 1]) k++;
        return k;
      };
      for (int r : L.brow) { int k = kof(r); rows[k].push_back(r - L.cstart[k]); }
      for (int c : L.bcol) { int k = kof(c); cols[k].push_back(c - L.cstart[k]); }
      for (int k1 = 1; k1 < L.npath; k1++)
        for (int k2 = 1; k2 < L.npath; k2++) {
          if (k1 == k2 || rows[k1].empty() || cols[k2].empty()) continue;
          int j1 = L.path[k1], j2 = L.path[k2];
          if (j1 > j2) {
            int off = P.anc_off(j1, j2);
            for (int c : cols[k2]) nzU[j1][off + c] = 1;
          } else {
            int off = P.anc_off(j2, j1);
            for (int r : rows[k1]) nzL[j2][off + r] = 1;
          }
        }
    }
    for (int l = P.max_level - 1; l >= 0; l--)
      for (int d = 1 << l; d < std::min(2 << l, np); d++)
        for (int j1 = d / 2; j1; j1 /= 2)
          for (int j2 = d / 2; j2; j2 /= 2) {
            if (j1 == j2) continue;
            int o1 = P.anc_off(d, j1), o2 = P.anc_off(d, j2);
            bool rowsnz = false, colsnz = false;
            for (int r = 0; r < P.block_size[j1] && !rowsnz; r++) rowsnz = nzL[d][o1 + r];
            for (int c = 0; c < P.block_size[j2] && !colsnz; c++) colsnz = nzU[d][o2 + c];
            if (!rowsnz || !colsnz) continue;
            if (j1 > j2) {
              int off = P.anc_off(j1, j2);
              for (int c = 0; c < P.block_size[j2]; c++) nzU[j1][off + c] |= nzU[d][o2 + c];
            } else {
              int off = P.anc_off(j2, j1);
              for (int r = 0; r < P.block_size[j1]; r++) nzL[j2][off + r] |= nzL[d][o1 + r];
            }
          }
    P.cntL.assign(nb + 1, std::vector<int>(1, 0));
    P.cntU.assign(nb + 1, std::vector<int>(1, 0));
    for (int p = 1; p < np; p++) {
      int a = P.anc_sum[p];
      P.cntL[p].assign(a + 1, 0);
      P.cntU[p].assign(a + 1, 0);
      for (int x = 0; x < a; x++) {
        P.cntL[p][x + 1] = P.cntL[p][x] + nzL[p][x];
        P.cntU[p][x + 1] = P.cntU[p][x] + nzU[p][x];
      }
    }
  }
  // ---- dense single-copy layout, tile-compressed panels ----
  // Per separator p: R_p = [D_p | Up_p] (s x (s + nU_p), ld s) and Lp_p
  // (nL_p x s, ld nL_p), where Up_p keeps, of each ancestor's columns, the
  // 64-column tiles (aligned to that ancestor's first column) holding a
  // structural nonzero, and Lp_p likewise its ancestors' 64-row tiles; every
  // structurally nonzero entry lies in a kept tile, and the tiles are those
  // of the dense factor, update and solve launches, so a kept tile is one
  // contiguous run of the compressed panel.
  P.nL.assign(nb + 1, 0);
  P.nU.assign(nb + 1, 0);
  P.lmap.assign(nb + 1, std::vector<int>());
  P.umap.assign(nb + 1, std::vector<int>());
  P.offR.assign(nb + 1, 0);
  P.offL.assign(nb + 1, 0);
  int64_t off = 0;
  for (int p = 1; p < np; p++) {
    const int a = P.anc_sum[p];
    std::vector<int> &lm = P.lmap[p], &um = P.umap[p];
    lm.assign(a, -1);
    um.assign(a, -1);
    int nl = 0, nu = 0;
    for (int q = p / 2; q; q /= 2) {
      const int o = P.anc_off(p, q), sq = P.block_size[q];
      for (int t = 0; t < sq; t += 64) {
        const int x0 = o + t, x1 = o + std::min(sq, t + 64);
        if (P.anyL(p, x0, x1))
          for (int x = x0; x < x1; x++) lm[x] = nl++;
        if (P.anyU(p, x0, x1))
          for (int x = x0; x < x1; x++) um[x] = nu++;
      }
    }
    P.nL[p] = nl;
    P.nU[p] = nu;
    int64_t s = P.block_size[p];
    P.offR[p] = off;
    off += s * (s + nu);
    P.offL[p] = off;
    off += (int64_t)nl * s;
  }
  P.dense_total = off;
  // ---- destination of every entry of A ----
  {
    std::vector<int64_t> dest_q(nnz, -1);
    for (int i = nb; i >= 1; i--) {
      for (int q = P.mstart[i]; q < P.mstart[i] + P.msize[i]; q++) {
        if (i >= np) {
          const LeafPlan &L = P.leaves[np + np - 1 - i];
          dest_q[q] = P.leaf_off[L.rank] + L.core_map[q - L.coo_off];
        } else {
          int bi = block[P.coo_r[q]], bj = block[P.coo_c[q]];
          const int64_t d = P.sep_index(bi, P.coo_r[q] - P.block_start[bi], bj, P.coo_c[q] - P.block_start[bj]);
          if (d < 0) return set_error(SNU_ERR_STRUCTURE, "separator entry outside the kept panel tiles (internal error)");
          dest_q[q] = P.leaf_total + d;
        }
      }
    }
    P.a_dest.assign(nnz, 0);
    for (int k = 0; k < nnz; k++) P.a_dest[k] = dest_q[P.perm_map[k]];
  }
  // ---- algorithmic counts ----
  double fd = 0, bsolve = 0, fl = 0, bleaf = 0;
  for (int p = 1; p < np; p++) {
    double s = P.block_size[p], a = P.anc_sum[p];
    fd += 2.0 / 3.0 * s * s * s + 2.0 * s * s * a + 2.0 * s * a * a;
    bsolve += 8.0 * (s * s + s * ((double)P.nL[p] + P.nU[p]));  // inverses' triangles + stored panels
  }
  for (const LeafPlan &L : P.leaves) {
    fl += L.flops;
    bleaf += 8.0 * L.core_nnz + 16.0 * L.lu_nnz + 4.0 * (L.lu_nnz + L.core_n) +
             16.0 * (double)L.nbr * L.nbc;
    bsolve += 12.0 * L.lu_nnz;
  }
  bsolve += 16.0 * n;
  P.f_dense = fd;
  P.f_leaf = fl;
  P.b_leaf = bleaf;
  P.b_solve = bsolve;
  return SNU_OK;
}

}  // namespace snu
