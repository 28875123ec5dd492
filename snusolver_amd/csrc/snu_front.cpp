// Multifrontal structure of a leaf (host, integer work).
//
// The leaf's no-pivot LU (the reference's core_numfact_v2 and core_GEMM,
// src/kernel_sparse.cu:740-781,953-971) is computed here as a sequence of
// dense partial factorizations of frontal matrices instead of row by row:
// a front holds a contiguous range K of leaf pivots and every later core
// index B that those pivots touch (their rows and columns beyond K, in the
// symmetrized pattern of the core); factoring the |K| pivots of a front
// gives the final L and U entries of those pivots and leaves the update
// -L(B,K) U(K,B) on B x B, which is added into the parent front (extend-add)
// and, at the leaf's root fronts, forms the leaf's boundary Schur block.
//
// Fronts follow the elimination tree of the symmetrized core pattern
// (symbolic Cholesky of A + A^T restricted to the leaf): every subtree of at
// most kSmallSubtree pivots whose indices are contiguous becomes one front,
// and above those, chains (pivot k whose only child is k - 1) share one.
// The union of the structures of a front's pivots beyond K equals the
// structure of its last pivot (column containment of the symbolic Cholesky
// factor), and a child's B minus its parent pivot lies inside the parent
// front, so the extend-add never meets an index the parent lacks.
//
// Every LU entry (i, j) of the reference's pattern belongs to the front whose
// K holds min(i, j): that front reads A's value from the entry's slot and
// writes the factor value back.  Positions of a front outside the LU pattern
// stay exactly zero (each of their updates multiplies a structural zero), so
// the factors are those of the reference's elimination, to rounding.
#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "snu_internal.h"

namespace snu {

namespace {
constexpr int kSmallSubtree = 32;
}

int build_fronts(LeafPlan &L) {
  const int ls = L.leaf_size, cn = L.core_n;
  L.fk0.clear();
  L.fnk.clear();
  L.ff.clear();
  L.fpar.clear();
  L.fheight.clear();
  L.fidx_ptr.assign(1, 0);
  L.fidx.clear();
  L.frel.clear();
  L.fslot_ptr.assign(1, 0);
  L.fslot.clear();
  L.fpos.clear();
  L.fchild_ptr.assign(1, 0);
  L.fchild.clear();
  L.fsrow.clear();
  L.fscol.clear();
  L.fentries = 0;
  L.fheight_max = 0;
  if (ls <= 0) return SNU_OK;
  // symmetrized adjacency of the leaf pivots (upper part: j > k)
  std::vector<std::vector<int>> adj(ls);
  for (int i = 0; i < cn; i++)
    for (int p = L.core_rowptr[i]; p < L.core_rowptr[i + 1]; p++) {
      const int j = L.core_colidx[p];
      if (i == j) continue;
      const int a = std::min(i, j), b = std::max(i, j);
      if (a < ls) adj[a].push_back(b);
    }
  // symbolic Cholesky: struct(k) = adj(k) U (struct(c) \ {k}) over children c
  std::vector<std::vector<int>> st(ls);
  std::vector<int> parent(ls, -1), mark(cn, -1);
  std::vector<std::vector<int>> kids(ls);
  for (int k = 0; k < ls; k++) {
    std::vector<int> &s = st[k];
    for (int j : adj[k])
      if (mark[j] != k) mark[j] = k, s.push_back(j);
    for (int c : kids[k])
      for (int j : st[c])
        if (j != k && mark[j] != k) mark[j] = k, s.push_back(j);
    std::sort(s.begin(), s.end());
    std::vector<int>().swap(adj[k]);
    if (!s.empty() && s[0] < ls) {
      parent[k] = s[0];
      kids[s[0]].push_back(k);
    }
  }
  // subtree sizes and first indices
  std::vector<int> sub(ls, 1), first(ls);
  std::iota(first.begin(), first.end(), 0);
  for (int k = 0; k < ls; k++)
    if (parent[k] >= 0) {
      sub[parent[k]] += sub[k];
      first[parent[k]] = std::min(first[parent[k]], first[k]);
    }
  auto croot = [&](int r) {
    return sub[r] <= kSmallSubtree && (parent[r] < 0 || sub[parent[r]] > kSmallSubtree) &&
           first[r] + sub[r] - 1 == r;
  };
  // fronts as contiguous pivot ranges, ascending (postorder of the front tree)
  std::vector<int> front_of(ls, -1);
  for (int k = 0; k < ls;) {
    int r = k;
    while (!croot(r) && sub[r] <= kSmallSubtree && parent[r] >= 0) r = parent[r];
    int end;
    if (croot(r) && first[r] == k) {
      end = r;
    } else {
      end = k;
      while (parent[end] == end + 1 && kids[end + 1].size() == 1 && sub[end + 1] > kSmallSubtree) end++;
    }
    for (int x = k; x <= end; x++) front_of[x] = (int)L.fk0.size();
    L.fk0.push_back(k);
    L.fnk.push_back(end - k + 1);
    k = end + 1;
  }
  const int nf = (int)L.fk0.size();
  L.ff.resize(nf);
  L.fpar.assign(nf, -1);
  L.fheight.assign(nf, 0);
  for (int f = 0; f < nf; f++) {
    const int k0 = L.fk0[f], k1 = k0 + L.fnk[f] - 1;
    const std::vector<int> &B = st[k1];
    // containment: every pivot's structure lies in K or in B
    for (int k = k0; k < k1; k++)
      for (int j : st[k])
        if (j > k1 && !std::binary_search(B.begin(), B.end(), j))
          return set_error(SNU_ERR_STRUCTURE, "leaf front structure is not closed (internal error)");
    L.fidx.insert(L.fidx.end(), B.begin(), B.end());
    L.fidx_ptr.push_back((int)L.fidx.size());
    L.ff[f] = L.fnk[f] + (int)B.size();
    L.fentries += (long long)L.ff[f] * L.ff[f];
    if (!B.empty() && B[0] < ls) L.fpar[f] = front_of[B[0]];
  }
  // children (ascending), heights
  std::vector<std::vector<int>> fk(nf);
  for (int f = 0; f < nf; f++)
    if (L.fpar[f] >= 0) fk[L.fpar[f]].push_back(f);
  for (int f = 0; f < nf; f++) {
    for (int c : fk[f]) L.fheight[f] = std::max(L.fheight[f], L.fheight[c] + 1);
    L.fchild.insert(L.fchild.end(), fk[f].begin(), fk[f].end());
    L.fchild_ptr.push_back((int)L.fchild.size());
    L.fheight_max = std::max(L.fheight_max, L.fheight[f]);
  }
  // local index of core index x inside front f (-1: not in it)
  auto loc = [&](int f, int x) -> int {
    const int k0 = L.fk0[f], nk = L.fnk[f];
    if (x >= k0 && x < k0 + nk) return x - k0;
    auto b = L.fidx.begin() + L.fidx_ptr[f], e = L.fidx.begin() + L.fidx_ptr[f + 1];
    auto it = std::lower_bound(b, e, x);
    return it != e && *it == x ? nk + (int)(it - b) : -1;
  };
  // extend-add maps: each B entry's local index in the parent front
  L.frel.assign(L.fidx.size(), -1);
  for (int f = 0; f < nf; f++) {
    const int p = L.fpar[f];
    if (p < 0) continue;
    for (int q = L.fidx_ptr[f]; q < L.fidx_ptr[f + 1]; q++) {
      L.frel[q] = loc(p, L.fidx[q]);
      if (L.frel[q] < 0) return set_error(SNU_ERR_STRUCTURE, "leaf front extend-add index missing (internal error)");
    }
  }
  // owned LU slots with their front positions (column-major), by position
  std::vector<std::vector<std::pair<int, int>>> own(nf);
  for (int i = 0; i < cn; i++)
    for (int q = L.lu_rowptr[i]; q < L.lu_rowptr[i + 1]; q++) {
      const int j = L.lu_colidx[q], m = std::min(i, j);
      if (m >= ls) return set_error(SNU_ERR_STRUCTURE, "LU entry outside the leaf columns (internal error)");
      const int f = front_of[m];
      const int r = loc(f, i), c = loc(f, j);
      if (r < 0 || c < 0) return set_error(SNU_ERR_STRUCTURE, "LU entry outside its front (internal error)");
      own[f].push_back({c * L.ff[f] + r, q});
    }
  for (int f = 0; f < nf; f++) {
    std::sort(own[f].begin(), own[f].end());
    for (auto &e : own[f]) {
      L.fpos.push_back(e.first);
      L.fslot.push_back(e.second);
    }
    L.fslot_ptr.push_back((int)L.fslot.size());
  }
  // root fronts: each B entry's row of the boundary Schur block (its index in
  // brow) and column (its index in bcol), -1 when it is neither
  L.fsrow.assign(L.fidx.size(), -1);
  L.fscol.assign(L.fidx.size(), -1);
  for (int f = 0; f < nf; f++) {
    if (L.fpar[f] >= 0) continue;
    for (int q = L.fidx_ptr[f]; q < L.fidx_ptr[f + 1]; q++) {
      const int x = L.fidx[q];
      auto r = std::lower_bound(L.brow.begin(), L.brow.end(), x);
      if (r != L.brow.end() && *r == x) L.fsrow[q] = (int)(r - L.brow.begin());
      auto c = std::lower_bound(L.bcol.begin(), L.bcol.end(), x);
      if (c != L.bcol.end() && *c == x) L.fscol[q] = (int)(c - L.bcol.begin());
    }
  }
  return SNU_OK;
}

}  // namespace snu
