// Multilevel vertex separator for the graph nested dissection (the role of
// ParMETIS_V3_NodeND's bisection, src/parmetis_driver.cpp:74-164; ParMETIS
// itself is third-party and absent, so this is the project's own method):
//
//   1. coarsening by heavy-edge matching in a seeded random visit order
//      (vertex and edge weights summed), until the graph stops shrinking or
//      is small;
//   2. initial edge bisection of the coarsest graph: greedy region growing
//      from several seeded start vertices, each refined by FM, best cut kept;
//   3. uncoarsening with Fiduccia-Mattheyses refinement of the edge cut at
//      every level (balance within a few percent);
//   4. the vertex separator of the final edge cut: a minimum vertex cover of
//      the bipartite graph of cut edges (maximum matching by Hopcroft-Karp,
//      then Koenig's construction), thinned (a separator vertex with no
//      neighbour on one side joins that side's opposite).
//
// Deterministic for a given graph and seed.
#include <algorithm>
#include <cstdint>
#include <queue>
#include <vector>

#include "snu_internal.h"

namespace snu {
namespace {

struct WGraph {
  int n = 0;
  std::vector<int> ptr, adj, ew, vw;
  long long total = 0;
};

struct Rng {
  uint64_t s;
  uint64_t next() { return splitmix64(s); }
  int below(int m) { return m <= 1 ? 0 : (int)(next() % (uint64_t)m); }
};

// heavy-edge matching; returns the coarse graph and cmap (fine -> coarse)
bool coarsen(const WGraph &g, Rng &rng, WGraph &c, std::vector<int> &cmap) {
  const int n = g.n;
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[rng.below(i + 1)]);
  const long long maxvw = std::max<long long>(1, (long long)(1.5 * g.total / 64));
  std::vector<int> match(n, -1);
  for (int v : perm) {
    if (match[v] >= 0) continue;
    int best = -1, bw = -1;
    for (int q = g.ptr[v]; q < g.ptr[v + 1]; q++) {
      int u = g.adj[q];
      if (match[u] >= 0 || u == v || g.vw[u] + g.vw[v] > maxvw) continue;
      if (g.ew[q] > bw || (g.ew[q] == bw && g.vw[u] < g.vw[best])) bw = g.ew[q], best = u;
    }
    if (best >= 0) match[v] = best, match[best] = v;
    else match[v] = v;
  }
  cmap.assign(n, -1);
  int cn = 0;
  for (int v = 0; v < n; v++)
    if (cmap[v] < 0) {
      cmap[v] = cn;
      if (match[v] != v) cmap[match[v]] = cn;
      cn++;
    }
  if (cn > 0.9 * n) return false;
  c.n = cn;
  c.vw.assign(cn, 0);
  c.total = g.total;
  c.ptr.assign(cn + 1, 0);
  c.adj.clear();
  c.ew.clear();
  std::vector<int> mark(cn, -1), pos(cn, 0);
  std::vector<int> members;
  int cv = 0;
  for (int v = 0; v < n; v++) {
    if (cmap[v] != cv) continue;  // first fine vertex of each coarse vertex
    members.clear();
    members.push_back(v);
    if (match[v] != v) members.push_back(match[v]);
    for (int x : members) {
      c.vw[cv] += g.vw[x];
      for (int q = g.ptr[x]; q < g.ptr[x + 1]; q++) {
        int cu = cmap[g.adj[q]];
        if (cu == cv) continue;
        if (mark[cu] != cv) {
          mark[cu] = cv;
          pos[cu] = (int)c.adj.size();
          c.adj.push_back(cu);
          c.ew.push_back(g.ew[q]);
        } else {
          c.ew[pos[cu]] += g.ew[q];
        }
      }
    }
    c.ptr[cv + 1] = (int)c.adj.size();
    cv++;
  }
  return true;
}

long long cut_of(const WGraph &g, const std::vector<int> &where) {
  long long cut = 0;
  for (int v = 0; v < g.n; v++)
    for (int q = g.ptr[v]; q < g.ptr[v + 1]; q++)
      if (where[g.adj[q]] != where[v]) cut += g.ew[q];
  return cut / 2;
}

// Fiduccia-Mattheyses refinement of a 2-way edge cut under a balance bound
void fm_refine(const WGraph &g, std::vector<int> &where, int passes) {
  const int n = g.n;
  if (n < 2) return;
  int maxv = 1;
  for (int v = 0; v < n; v++) maxv = std::max(maxv, g.vw[v]);
  const long long half = g.total / 2;
  const long long slack = std::max<long long>(maxv, (long long)(0.03 * g.total));
  const long long maxw = half + slack + (g.total & 1);
  std::vector<long long> gain(n);
  std::vector<int> ver(n, 0);
  std::vector<char> moved(n, 0);
  typedef std::pair<long long, std::pair<int, int>> Ent;  // gain, (vertex, version)
  for (int pass = 0; pass < passes; pass++) {
    long long tw[2] = {0, 0};
    for (int v = 0; v < n; v++) tw[where[v]] += g.vw[v];
    std::priority_queue<Ent> heap[2];
    for (int v = 0; v < n; v++) {
      long long ext = 0, in = 0;
      for (int q = g.ptr[v]; q < g.ptr[v + 1]; q++) (where[g.adj[q]] != where[v] ? ext : in) += g.ew[q];
      gain[v] = ext - in;
      moved[v] = 0;
      if (ext > 0) heap[where[v]].push({gain[v], {-v, ++ver[v]}});
    }
    long long cut = cut_of(g, where), best = cut;
    long long bestimb = std::llabs(tw[0] - tw[1]);
    std::vector<int> moves;
    int bestlen = 0, nonimp = 0;
    const int limit = std::max(64, n / 50);
    for (;;) {
      // candidates: the best valid move of each side that keeps the other
      // side within the bound; a side over the bound must give, otherwise
      // the larger gain moves (ties: from the heavier side)
      int top[2] = {-1, -1};
      for (int side = 0; side < 2; side++) {
        auto &H = heap[side];
        while (!H.empty()) {
          const int v = -H.top().second.first;
          if (moved[v] || H.top().second.second != ver[v] || where[v] != side) {
            H.pop();
            continue;
          }
          break;
        }
        if (!H.empty()) {
          const int v = -H.top().second.first;
          if (tw[1 - side] + g.vw[v] <= maxw) top[side] = v;
        }
      }
      int pick = -1;
      if (tw[0] > maxw) pick = top[0];
      else if (tw[1] > maxw) pick = top[1];
      else if (top[0] >= 0 && top[1] >= 0)
        pick = gain[top[0]] > gain[top[1]] || (gain[top[0]] == gain[top[1]] && tw[0] >= tw[1]) ? top[0] : top[1];
      else pick = top[0] >= 0 ? top[0] : top[1];
      if (pick < 0) break;
      const int v = pick, from = where[v], to = 1 - from;
      heap[from].pop();
      where[v] = to;
      moved[v] = 1;
      tw[from] -= g.vw[v];
      tw[to] += g.vw[v];
      cut -= gain[v];
      moves.push_back(v);
      for (int q = g.ptr[v]; q < g.ptr[v + 1]; q++) {
        const int u = g.adj[q];
        if (moved[u]) continue;
        gain[u] += where[u] == to ? -2LL * g.ew[q] : 2LL * g.ew[q];
        heap[where[u]].push({gain[u], {-u, ++ver[u]}});
      }
      const long long imb = std::llabs(tw[0] - tw[1]);
      if (cut < best || (cut == best && imb < bestimb)) {
        best = cut;
        bestimb = imb;
        bestlen = (int)moves.size();
        nonimp = 0;
      } else if (++nonimp > limit) {
        break;
      }
    }
    for (int k = (int)moves.size() - 1; k >= bestlen; k--) where[moves[k]] ^= 1;
    if (bestlen == 0) break;
  }
}

// greedy region growing from start vertices, each refined; best cut kept
void initial_bisect(const WGraph &g, Rng &rng, std::vector<int> &where) {
  const int n = g.n;
  std::vector<int> best;
  long long bestcut = -1, bestimb = 0;
  const int trials = n < 8 ? 1 : 8;
  for (int t = 0; t < trials; t++) {
    std::vector<int> w(n, 1);
    long long grown = 0;
    std::vector<char> seen(n, 0);
    std::vector<int> queue;
    size_t head = 0;
    int start = rng.below(n);
    while (grown < g.total / 2) {
      if (head == queue.size()) {  // next component
        int s = -1;
        for (int k = 0; k < n; k++) {
          int v = (start + k) % n;
          if (!seen[v]) { s = v; break; }
        }
        if (s < 0) break;
        seen[s] = 1;
        queue.push_back(s);
      }
      int v = queue[head++];
      if (grown + g.vw[v] > g.total / 2 + g.vw[v] / 2 && grown > 0) break;
      w[v] = 0;
      grown += g.vw[v];
      for (int q = g.ptr[v]; q < g.ptr[v + 1]; q++)
        if (!seen[g.adj[q]]) seen[g.adj[q]] = 1, queue.push_back(g.adj[q]);
    }
    fm_refine(g, w, 4);
    long long tw0 = 0;
    for (int v = 0; v < n; v++) tw0 += w[v] == 0 ? g.vw[v] : 0;
    const long long cut = cut_of(g, w), imb = std::llabs(2 * tw0 - g.total);
    if (bestcut < 0 || cut < bestcut || (cut == bestcut && imb < bestimb)) {
      bestcut = cut;
      bestimb = imb;
      best = w;
    }
  }
  where = best;
}

// minimum vertex cover of the cut edges (Hopcroft-Karp, then Koenig)
void cover_cut(const WGraph &g, std::vector<int> &part) {
  const int n = g.n;
  std::vector<int> lid(n, -1), rid(n, -1), L, R;
  for (int v = 0; v < n; v++) {
    bool b = false;
    for (int q = g.ptr[v]; q < g.ptr[v + 1] && !b; q++) b = part[g.adj[q]] != part[v];
    if (!b) continue;
    if (part[v] == 0) lid[v] = (int)L.size(), L.push_back(v);
    else rid[v] = (int)R.size(), R.push_back(v);
  }
  const int nl = (int)L.size(), nr = (int)R.size();
  if (!nl) return;
  std::vector<int> lptr(nl + 1, 0), ladj;
  for (int i = 0; i < nl; i++) {
    const int v = L[i];
    for (int q = g.ptr[v]; q < g.ptr[v + 1]; q++)
      if (part[g.adj[q]] == 1) ladj.push_back(rid[g.adj[q]]);
    lptr[i + 1] = (int)ladj.size();
  }
  std::vector<int> ml(nl, -1), mr(nr, -1), dist(nl);
  const int INF = 1 << 30;
  auto bfs = [&]() {
    std::vector<int> q;
    bool found = false;
    for (int i = 0; i < nl; i++) {
      if (ml[i] < 0) dist[i] = 0, q.push_back(i);
      else dist[i] = INF;
    }
    for (size_t h = 0; h < q.size(); h++) {
      const int i = q[h];
      for (int e = lptr[i]; e < lptr[i + 1]; e++) {
        const int j = mr[ladj[e]];
        if (j < 0) found = true;
        else if (dist[j] == INF) dist[j] = dist[i] + 1, q.push_back(j);
      }
    }
    return found;
  };
  std::vector<int> it(nl);
  // iterative augmenting DFS along the BFS layers
  auto dfs = [&](int root) {
    std::vector<int> stack{root};
    while (!stack.empty()) {
      const int i = stack.back();
      if (it[i] == lptr[i + 1]) {
        dist[i] = INF;
        stack.pop_back();
        continue;
      }
      const int r = ladj[it[i]];
      const int j = mr[r];
      if (j < 0) {  // free right vertex: flip the path
        for (int k = (int)stack.size() - 1; k >= 0; k--) {
          const int a = stack[k];
          const int rr = ladj[it[a]];
          const int prev = ml[a];
          ml[a] = rr;
          mr[rr] = a;
          (void)prev;
        }
        return true;
      }
      if (dist[j] == dist[i] + 1) {
        stack.push_back(j);
        continue;
      }
      it[i]++;
    }
    return false;
  };
  while (bfs()) {
    for (int i = 0; i < nl; i++) it[i] = lptr[i];
    for (int i = 0; i < nl; i++)
      if (ml[i] < 0) dfs(i);
  }
  // Koenig: Z = vertices reachable from free left vertices by alternating
  // paths; cover = (L \ Z) + (R n Z)
  std::vector<char> zl(nl, 0), zr(nr, 0);
  std::vector<int> q;
  for (int i = 0; i < nl; i++)
    if (ml[i] < 0) zl[i] = 1, q.push_back(i);
  for (size_t h = 0; h < q.size(); h++) {
    const int i = q[h];
    for (int e = lptr[i]; e < lptr[i + 1]; e++) {
      const int r = ladj[e];
      if (zr[r] || ml[i] == r) continue;
      zr[r] = 1;
      const int j = mr[r];
      if (j >= 0 && !zl[j]) zl[j] = 1, q.push_back(j);
    }
  }
  for (int i = 0; i < nl; i++)
    if (!zl[i]) part[L[i]] = 2;
  for (int r = 0; r < nr; r++)
    if (zr[r]) part[R[r]] = 2;
}

}  // namespace

void ml_separator(int n, const std::vector<int> &ptr, const std::vector<int> &adj, uint64_t seed,
                  std::vector<int> &part) {
  part.assign(n, 0);
  if (n < 3) {
    if (n == 2) part[1] = 1;
    return;
  }
  Rng rng{seed * 0x9e3779b97f4a7c15ULL + (uint64_t)n};
  std::vector<WGraph> levels(1);
  levels[0].n = n;
  levels[0].ptr = ptr;
  levels[0].adj = adj;
  levels[0].ew.assign(adj.size(), 1);
  levels[0].vw.assign(n, 1);
  levels[0].total = n;
  std::vector<std::vector<int>> cmaps;
  while (levels.back().n > 120) {
    WGraph c;
    std::vector<int> cmap;
    if (!coarsen(levels.back(), rng, c, cmap)) break;
    levels.push_back(std::move(c));
    cmaps.push_back(std::move(cmap));
  }
  std::vector<int> where;
  initial_bisect(levels.back(), rng, where);
  for (int l = (int)levels.size() - 2; l >= 0; l--) {
    const std::vector<int> &cm = cmaps[l];
    std::vector<int> fine(levels[l].n);
    for (int v = 0; v < levels[l].n; v++) fine[v] = where[cm[v]];
    where.swap(fine);
    fm_refine(levels[l], where, 6);
  }
  part = where;
  cover_cut(levels[0], part);
  // thinning: a separator vertex without neighbours on one side joins the
  // other side (ascending vertex order: deterministic)
  for (int v = 0; v < n; v++) {
    if (part[v] != 2) continue;
    bool has0 = false, has1 = false;
    for (int q = ptr[v]; q < ptr[v + 1]; q++) {
      if (part[adj[q]] == 0) has0 = true;
      if (part[adj[q]] == 1) has1 = true;
    }
    if (!has1) part[v] = 0;
    else if (!has0) part[v] = 1;
  }
}

}  // namespace snu
