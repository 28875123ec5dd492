// Nested-dissection orderings in the ParMETIS_V3_NodeND output contract that
// the reference consumes (src/parmetis_driver.cpp:74-164,
// src/SnuMat_construct.cu:21-38):
//   order[i]  = new (permuted) index of vertex i;
//   sizes[2np-1]: sizes[2np-1-e] is the size of heap block e (root e = 1,
//   children 2e and 2e+1, leaves np..2np-1); the permuted numbering lists
//   each subtree in post-order with child 2e+1 first.
//
// Two orderers:
//  - order_grid: geometric ND of an nx*ny*nz grid (split the longest axis at
//    its middle with a one-vertex-thick plane; recurse inside leaves too).
//    A one-thick plane separates 5/7/9/27-point stencils alike.
//  - order_nd: graph ND of A + A^T without the diagonal
//    (addTransposeEliminateDiag, src/parmetis_driver.cpp:15-71).  Method 0
//    (the default): each bisection takes the smaller of the multilevel
//    vertex separator of snu_nd_ml.cpp and the middle BFS level set from a
//    pseudo-peripheral vertex (thinned).  Method 1: level sets only (the
//    simpler orderer); method 2: multilevel separators only.
#include <algorithm>
#include <cstring>
#include <queue>
#include <vector>

#include "snu_internal.h"

namespace snu {

namespace {

struct Box {
  int lo[3], hi[3];
  long long vol() const {
    long long v = 1;
    for (int a = 0; a < 3; a++) v *= (hi[a] > lo[a] ? hi[a] - lo[a] : 0);
    return v;
  }
};

struct GridCtx {
  int nx, ny, nz;
  int *order;
  int next = 0;
  void lex(const Box &b) {
    for (int z = b.lo[2]; z < b.hi[2]; z++)
      for (int y = b.lo[1]; y < b.hi[1]; y++)
        for (int x = b.lo[0]; x < b.hi[0]; x++) order[x + nx * (y + ny * z)] = next++;
  }
  // split b along its longest axis (ties: x, then y, then z)
  static void split(const Box &b, Box &l, Box &r, Box &s) {
    int ax = 0;
    for (int a = 1; a < 3; a++)
      if (b.hi[a] - b.lo[a] > b.hi[ax] - b.lo[ax]) ax = a;
    int mid = b.lo[ax] + (b.hi[ax] - b.lo[ax]) / 2;
    l = r = s = b;
    l.hi[ax] = mid;
    s.lo[ax] = mid;
    s.hi[ax] = mid + 1;
    r.lo[ax] = mid + 1;
  }
  void nd(const Box &b) {
    if (b.vol() <= 0) return;
    if (b.vol() <= 2) {
      lex(b);
      return;
    }
    Box l, r, s;
    split(b, l, r, s);
    nd(l);
    nd(r);
    lex(s);
  }
  void tree(const Box &b, int parts, int node, int np, int *sizes) {
    if (parts == 1) {
      int before = next;
      nd(b);
      sizes[2 * np - 1 - node] = next - before;
      return;
    }
    Box l, r, s;
    split(b, l, r, s);
    tree(l, parts / 2, 2 * node + 1, np, sizes);
    tree(r, parts / 2, 2 * node, np, sizes);
    int before = next;
    lex(s);
    sizes[2 * np - 1 - node] = next - before;
  }
};

// ---- graph ND ----
struct GraphCtx {
  int n;
  std::vector<int> ptr, adj;  // A + A^T, no diagonal
  std::vector<int> tag;       // current subset membership tag
  std::vector<int> dist;
  std::vector<int> where;     // 0 = A, 1 = B, 2 = S during a bisection
  int *order;
  int next = 0;
  int tagc = 0;
  uint64_t seed;

  // BFS inside the tagged subset from s; returns visit order; levels in lev
  void bfs(int s, int t, std::vector<int> &visit, std::vector<int> &lstart) {
    visit.clear();
    lstart.clear();
    visit.push_back(s);
    dist[s] = 0;
    size_t head = 0;
    int curlev = -1;
    while (head < visit.size()) {
      int v = visit[head];
      if (dist[v] != curlev) {
        curlev = dist[v];
        lstart.push_back((int)head);
      }
      head++;
      for (int q = ptr[v]; q < ptr[v + 1]; q++) {
        int u = adj[q];
        if (tag[u] == t && dist[u] < 0) {
          dist[u] = dist[v] + 1;
          visit.push_back(u);
        }
      }
    }
    lstart.push_back((int)visit.size());
  }
  void reset_dist(const std::vector<int> &vs) {
    for (int v : vs) dist[v] = -1;
  }

  int method = 0;
  std::vector<int> lid;
  // Splits vs (all carrying tag t) into A, B, S by the multilevel separator
  // of the induced subgraph (the seed varies per call, deterministically).
  void bisect_ml(const std::vector<int> &vs, int t, std::vector<int> &A,
                 std::vector<int> &B, std::vector<int> &S) {
    std::vector<int> sorted(vs);
    std::sort(sorted.begin(), sorted.end());
    const int m = (int)sorted.size();
    for (int i = 0; i < m; i++) lid[sorted[i]] = i;
    std::vector<int> lptr(m + 1, 0), ladj;
    for (int i = 0; i < m; i++) {
      const int v = sorted[i];
      for (int q = ptr[v]; q < ptr[v + 1]; q++)
        if (tag[adj[q]] == t) ladj.push_back(lid[adj[q]]);
      lptr[i + 1] = (int)ladj.size();
    }
    std::vector<int> part;
    ml_separator(m, lptr, ladj, seed + (uint64_t)tagc, part);
    A.clear(); B.clear(); S.clear();
    for (int i = 0; i < m; i++) (part[i] == 0 ? A : part[i] == 1 ? B : S).push_back(sorted[i]);
  }
  // Splits vs (all carrying tag t) into A, B, S.  Method 0 computes both
  // the multilevel and the level-set separator and keeps the smaller one
  // among those that leave both sides at least 30% of the rest (the more
  // balanced one if neither does): level sets win on graphs whose
  // breadth-first layers are thin (7-point grids, where they run
  // diagonally), the multilevel separator everywhere else.
  void bisect(const std::vector<int> &vs, int t, std::vector<int> &A,
              std::vector<int> &B, std::vector<int> &S) {
    if (method == 0) {
      std::vector<int> A2, B2, S2;
      bisect_ml(vs, t, A, B, S);
      bisect_ls(vs, t, A2, B2, S2);
      auto ok = [](const std::vector<int> &a, const std::vector<int> &b) {
        return 10 * std::min(a.size(), b.size()) >= 3 * (a.size() + b.size());
      };
      auto bal = [](const std::vector<int> &a, const std::vector<int> &b) {
        return (double)std::min(a.size(), b.size()) / std::max<size_t>(1, a.size() + b.size());
      };
      const bool ok1 = ok(A, B), ok2 = ok(A2, B2);
      const bool take2 = (ok1 && ok2) ? S2.size() < S.size() : ok2 ? true : ok1 ? false : bal(A2, B2) > bal(A, B);
      if (take2) {
        A.swap(A2);
        B.swap(B2);
        S.swap(S2);
      }
      return;
    }
    if (method == 2) bisect_ml(vs, t, A, B, S);
    else bisect_ls(vs, t, A, B, S);
  }
  // Level-set separator: the middle BFS level of each component's level
  // structure from a pseudo-peripheral vertex, thinned.
  void bisect_ls(const std::vector<int> &vs, int t, std::vector<int> &A,
                 std::vector<int> &B, std::vector<int> &S) {
    // level structures of every component, concatenated
    std::vector<std::vector<int>> levels;
    std::vector<int> visit, lstart;
    std::vector<int> sorted(vs);
    std::sort(sorted.begin(), sorted.end());
    for (int v : sorted) dist[v] = -1;
    for (size_t ci = 0; ci < sorted.size(); ci++) {
      int s = sorted[ci];
      if (dist[s] >= 0) continue;
      bfs(s, t, visit, lstart);
      std::vector<int> comp(visit);
      if (comp.size() > 1) {
        // pseudo-peripheral start: seed picks the first root of the search
        int start = comp[(size_t)(seed % comp.size())];
        int ecc = -1;
        for (int it = 0; it < 4; it++) {
          reset_dist(comp);
          bfs(start, t, visit, lstart);
          int e = (int)lstart.size() - 2;
          if (e <= ecc) break;
          ecc = e;
          int best = visit[lstart[lstart.size() - 2]], bd = 1 << 30;
          for (int q = lstart[lstart.size() - 2]; q < lstart.back(); q++) {
            int v = visit[q], d = ptr[v + 1] - ptr[v];
            if (d < bd || (d == bd && v < best)) bd = d, best = v;
          }
          start = best;
        }
        reset_dist(comp);
        bfs(start, t, visit, lstart);
      }
      for (size_t L = 0; L + 1 < lstart.size(); L++) {
        std::vector<int> lv(visit.begin() + lstart[L], visit.begin() + lstart[L + 1]);
        levels.push_back(lv);
      }
    }
    size_t total = vs.size(), cum = 0, m = 0;
    for (m = 0; m < levels.size(); m++) {
      if (cum + levels[m].size() >= (total + 1) / 2) break;
      cum += levels[m].size();
    }
    if (m >= levels.size()) m = levels.size() - 1;
    for (size_t L = 0; L < levels.size(); L++)
      for (int v : levels[L]) where[v] = L < m ? 0 : (L == m ? 2 : 1);
    // thin the separator: a separator vertex with no neighbour on one side
    // moves to the other side
    std::vector<int> lv(levels[m]);
    std::sort(lv.begin(), lv.end());
    for (int v : lv) {
      bool hasA = false, hasB = false;
      for (int q = ptr[v]; q < ptr[v + 1]; q++) {
        int u = adj[q];
        if (tag[u] != t) continue;
        if (where[u] == 0) hasA = true;
        if (where[u] == 1) hasB = true;
      }
      if (!hasB) where[v] = 0;
      else if (!hasA) where[v] = 1;
    }
    A.clear(); B.clear(); S.clear();
    for (int v : sorted) (where[v] == 0 ? A : where[v] == 1 ? B : S).push_back(v);
  }
  void number(std::vector<int> vs) {
    std::sort(vs.begin(), vs.end());
    for (int v : vs) order[v] = next++;
  }
  void nd_order(const std::vector<int> &vs) {
    if (vs.size() <= 16) {
      number(vs);
      return;
    }
    int t = ++tagc;
    for (int v : vs) tag[v] = t;
    std::vector<int> A, B, S;
    bisect(vs, t, A, B, S);
    if (A.empty() || B.empty()) {  // no useful split: number as is
      number(vs);
      return;
    }
    nd_order(A);
    nd_order(B);
    number(S);
  }
  void tree(const std::vector<int> &vs, int parts, int node, int np,
            int *sizes) {
    if (parts == 1) {
      int before = next;
      nd_order(vs);
      sizes[2 * np - 1 - node] = next - before;
      return;
    }
    std::vector<int> A, B, S;
    if (!vs.empty()) {
      int t = ++tagc;
      for (int v : vs) tag[v] = t;
      bisect(vs, t, A, B, S);
    }
    tree(A, parts / 2, 2 * node + 1, np, sizes);
    tree(B, parts / 2, 2 * node, np, sizes);
    int before = next;
    number(S);
    sizes[2 * np - 1 - node] = next - before;
  }
};

bool pow2(int v) { return v >= 1 && (v & (v - 1)) == 0; }

}  // namespace

int order_grid(int nx, int ny, int nz, int np, int *order, int *sizes) {
  if (nx < 1 || ny < 1 || nz < 1 || !pow2(np))
    return set_error(SNU_ERR_BAD_INPUT, "order_grid: bad grid or np");
  GridCtx g{nx, ny, nz, order};
  Box b{{0, 0, 0}, {nx, ny, nz}};
  g.tree(b, np, 1, np, sizes);
  return SNU_OK;
}

int order_nd(int n, const int *rowptr, const int *colidx, int np,
             uint64_t seed, int *order, int *sizes, int method) {
  if (n < 0 || !pow2(np)) return set_error(SNU_ERR_BAD_INPUT, "order_nd: bad n or np");
  if (method < 0 || method > 2) return set_error(SNU_ERR_BAD_INPUT, "order_nd: unknown method");
  GraphCtx g;
  g.n = n;
  g.order = order;
  g.seed = seed;
  g.method = method;
  // A + A^T without the diagonal, rows sorted and deduplicated
  std::vector<int> deg(n + 1, 0);
  for (int i = 0; i < n; i++)
    for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
      int j = colidx[q];
      if (j < 0 || j >= n) return set_error(SNU_ERR_BAD_INPUT, "order_nd: column out of range");
      if (j != i) deg[i]++, deg[j]++;
    }
  g.ptr.assign(n + 1, 0);
  for (int i = 0; i < n; i++) g.ptr[i + 1] = g.ptr[i] + deg[i];
  g.adj.resize(g.ptr[n]);
  std::vector<int> fill(g.ptr.begin(), g.ptr.end() - 1);
  for (int i = 0; i < n; i++)
    for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
      int j = colidx[q];
      if (j != i) g.adj[fill[i]++] = j, g.adj[fill[j]++] = i;
    }
  std::vector<int> nptr(n + 1, 0), nadj;
  nadj.reserve(g.adj.size());
  for (int i = 0; i < n; i++) {
    std::sort(g.adj.begin() + g.ptr[i], g.adj.begin() + g.ptr[i + 1]);
    int last = -1;
    for (int q = g.ptr[i]; q < g.ptr[i + 1]; q++)
      if (g.adj[q] != last) nadj.push_back(last = g.adj[q]);
    nptr[i + 1] = (int)nadj.size();
  }
  g.ptr.swap(nptr);
  g.adj.swap(nadj);
  g.tag.assign(n, 0);
  g.dist.assign(n, -1);
  g.where.assign(n, 0);
  g.lid.assign(n, 0);
  std::vector<int> all(n);
  for (int i = 0; i < n; i++) all[i] = i;
  g.tree(all, np, 1, np, sizes);
  return SNU_OK;
}

}  // namespace snu

extern "C" int snu_order_grid(int nx, int ny, int nz, int np, int *order,
                              int *sizes) {
  return snu::order_grid(nx, ny, nz, np, order, sizes);
}
extern "C" int snu_order_nd(int n, const int *rowptr, const int *colidx,
                            int np, uint64_t seed, int *order, int *sizes) {
  return snu::order_nd(n, rowptr, colidx, np, seed, order, sizes, 0);
}
extern "C" int snu_order_nd_ex(int n, const int *rowptr, const int *colidx, int np,
                               uint64_t seed, int method, int *order, int *sizes) {
  return snu::order_nd(n, rowptr, colidx, np, seed, order, sizes, method);
}
