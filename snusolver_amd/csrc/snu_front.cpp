// Multifrontal structure of a leaf (host, integer work).
//
// The leaf's no-pivot LU (the reference's core_numfact_v2 and core_GEMM,
// src/kernel_sparse.cu:740-781,953-971) is computed here as a sequence of
// dense partial factorizations of frontal matrices instead of row by row:
// a front holds a set K of leaf pivots and every later core index B that
// those pivots touch (their rows and columns outside K, in the symmetrized
// pattern of the core); factoring the |K| pivots of a front in ascending
// order gives the final L and U entries of those pivots and leaves the
// update -L(B,K) U(K,B) on B x B, which is added into the parent front
// (extend-add) and, at the leaf's root fronts, forms the leaf's boundary
// Schur block.
//
// Fronts follow the elimination tree of the symmetrized core pattern
// (symbolic Cholesky of A + A^T restricted to the leaf).  Every maximal
// subtree whose front (its pivots and the structure of its root) has order
// at most kSubtreeFront becomes one front, however its pivots are numbered;
// above those, a pivot whose only child is a pivot of the same kind joins
// that child's front (chains); then fronts of order above kSubtreeFront with
// at most kRelaxPivots pivots join their parent fronts.  A front's pivots
// are eliminated in
// ascending order; pivots outside its subtree are neither descendants nor
// ancestors of its pivots, so they update none of its entries, and the
// factors are those of the reference's elimination order.  The union of the
// structures of a front's pivots outside K equals the structure of its top
// pivot (column containment of the symbolic Cholesky factor), and a child's
// B minus the parent's pivots lies inside the parent front, so the
// extend-add never meets an index the parent lacks.
//
// Every LU entry (i, j) of the reference's pattern belongs to the front
// whose K holds min(i, j): that front reads A's value from the entry's slot
// and writes the factor value back.  Positions of a front outside the LU
// pattern stay exactly zero (each of their updates multiplies a structural
// zero), so the factors are those of the reference's elimination, to
// rounding.
#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "snu_internal.h"

namespace snu {

namespace {
constexpr int kSubtreeFront = 48;   // subtrees collapse into one front up to this order
constexpr int kRelaxPivots = 32;    // larger thin fronts with at most this many pivots merge into their parent
constexpr int kThin = 4;            // thin: pivots x kThin <= order
}

int build_fronts(LeafPlan &L) {
  const int ls = L.leaf_size, cn = L.core_n;
  L.fkp.assign(1, 0);
  L.fk.clear();
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
  // subtree pivot counts; a subtree collapses into one front when its order
  // (pivots + the root's structure) is within kSubtreeFront.  The order grows
  // towards the root (a child's structure lies in its parent's pivot and
  // structure), so the collapsed fronts are the maximal such subtrees.
  std::vector<int> sub(ls, 1);
  for (int k = 0; k < ls; k++)
    if (parent[k] >= 0) sub[parent[k]] += sub[k];
  auto coll = [&](int r) { return sub[r] + (long long)st[r].size() <= kSubtreeFront; };
  // representative pivot of every pivot's front: the root of its collapsed
  // subtree, or the bottom of its chain
  std::vector<int> rep(ls);
  for (int k = ls - 1; k >= 0; k--)  // parents first
    if (coll(k)) rep[k] = parent[k] >= 0 && coll(parent[k]) ? rep[parent[k]] : k;
  for (int k = 0; k < ls; k++) {  // children first
    if (coll(k)) continue;
    rep[k] = kids[k].size() == 1 && !coll(kids[k][0]) ? rep[kids[k][0]] : k;
  }
  // fronts ordered by their top (largest) pivot: a parent front's top is an
  // etree ancestor of its children's tops, so children come first
  std::vector<int> topof(ls, -1);  // per representative: the largest pivot of its front
  for (int k = 0; k < ls; k++) topof[rep[k]] = std::max(topof[rep[k]], k);
  std::vector<int> front_of(ls, -1), tops;
  for (int k = 0; k < ls; k++)
    if (topof[rep[k]] == k) front_of[k] = (int)tops.size(), tops.push_back(k);
  for (int k = 0; k < ls; k++) front_of[k] = front_of[topof[rep[k]]];
  // relaxed amalgamation: a front of order above kSubtreeFront with at most
  // kRelaxPivots pivots joins its parent front (its pivots become rows and
  // columns of the parent's front, whose structure B already holds the
  // child's), children first; this shortens the front tree, which is
  // processed height by height (each height of the larger fronts costs a
  // sequence of launches)
  {
    const int n0 = (int)tops.size();
    std::vector<int> cnt(n0, 0), into(n0, -1);
    for (int k = 0; k < ls; k++) cnt[front_of[k]]++;
    for (int f = 0; f < n0; f++) {
      const std::vector<int> &B = st[tops[f]];
      const long long order = cnt[f] + (long long)B.size();
      if (B.empty() || B[0] >= ls || cnt[f] > kRelaxPivots || order <= kSubtreeFront ||
          (long long)cnt[f] * kThin > order)
        continue;
      const int p = front_of[B[0]];  // not merged yet: p > f
      into[f] = p;
      cnt[p] += cnt[f];
    }
    std::vector<int> alive_id(n0, -1), ntops;
    for (int f = 0; f < n0; f++)
      if (into[f] < 0) alive_id[f] = (int)ntops.size(), ntops.push_back(tops[f]);
    for (int k = 0; k < ls; k++) {
      int f = front_of[k];
      while (into[f] >= 0) f = into[f];
      front_of[k] = alive_id[f];
    }
    tops.swap(ntops);
  }
  const int nf = (int)tops.size();
  L.fnk.assign(nf, 0);
  for (int k = 0; k < ls; k++) L.fnk[front_of[k]]++;
  for (int f = 0; f < nf; f++) L.fkp.push_back(L.fkp.back() + L.fnk[f]);
  L.fk.resize(ls);
  {
    std::vector<int> at(L.fkp.begin(), L.fkp.end() - 1);
    for (int k = 0; k < ls; k++) L.fk[at[front_of[k]]++] = k;  // ascending within a front
  }
  auto inK = [&](int f, int x) {
    return std::binary_search(L.fk.begin() + L.fkp[f], L.fk.begin() + L.fkp[f + 1], x);
  };
  L.ff.resize(nf);
  L.fpar.assign(nf, -1);
  L.fheight.assign(nf, 0);
  for (int f = 0; f < nf; f++) {
    const std::vector<int> &B = st[tops[f]];
    // containment: every pivot's structure lies in K or in B
    for (int q = L.fkp[f]; q < L.fkp[f + 1]; q++)
      for (int j : st[L.fk[q]])
        if (!inK(f, j) && !std::binary_search(B.begin(), B.end(), j))
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
    auto kb = L.fk.begin() + L.fkp[f], ke = L.fk.begin() + L.fkp[f + 1];
    auto kt = std::lower_bound(kb, ke, x);
    if (kt != ke && *kt == x) return (int)(kt - kb);
    auto b = L.fidx.begin() + L.fidx_ptr[f], e = L.fidx.begin() + L.fidx_ptr[f + 1];
    auto it = std::lower_bound(b, e, x);
    return it != e && *it == x ? L.fnk[f] + (int)(it - b) : -1;
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
