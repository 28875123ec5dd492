// Host-side engine: device layout of one matrix, launch lists of its dense
// factorization and dense solves, merging of the lists of several matrices,
// and the executors of the factor and solve sweeps.
//
// Phase map (reference -> here):
//   core_run (kernel_sparse.cu:1011-1037)        -> prepare (zero + scatter),
//        k_leaf_rows, k_leaf_sep, k_assemble_schur
//   dense level loop (SnuMat_factorize.cu:258-284) -> per level: blocked
//        getrf/trsm (k_diag_lu + k_gemm by the block inverses), then one
//        grouped k_gemm for all Schur updates of the level (K-concatenated
//        over siblings: replaces reduction_gpu, SnuMat_factorize.cu:78-124)
//   forward solve (core_trsm/core_update_b, trsm_Lxb/gemm on b; fused into
//        the factorization in the reference) and backward solve
//        (SnuMat_factorize.cu:286-345)     -> run_sweep
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "snu_engine.h"

namespace snu {

int device_alloc(snumat *h, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 1;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(SNU_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed (" +
                                      hipGetErrorString(e) + ")");
  }
  h->allocs.push_back(*p);
  h->dev_bytes += bytes;
  return SNU_OK;
}

namespace {
template <class T>
int dalloc(snumat *h, T **p, size_t count) {
  return device_alloc(h, (void **)p, std::max<size_t>(1, count) * sizeof(T));
}
template <class T>
int dupload(snumat *h, T **p, const std::vector<T> &v) {
  int rc = dalloc(h, p, v.size());
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return SNU_OK;
}
}  // namespace

template <typename T>
int upload_list(T **p, const std::vector<T> &v) {
  *p = nullptr;
  HIPCHK(hipMalloc((void **)p, std::max<size_t>(1, v.size()) * sizeof(T)));
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return SNU_OK;
}
template int upload_list<int>(int **, const std::vector<int> &);
template int upload_list<LeafDev>(LeafDev **, const std::vector<LeafDev> &);
template int upload_list<VecDev>(VecDev **, const std::vector<VecDev> &);
template int upload_list<PermDev>(PermDev **, const std::vector<PermDev> &);
template int upload_list<PrepDev>(PrepDev **, const std::vector<PrepDev> &);
template int upload_list<IrDev>(IrDev **, const std::vector<IrDev> &);

float ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

// ---------------------------------------------------------------------------
// dense plan (structure only; pointers into the matrix's value arrays)
// ---------------------------------------------------------------------------
template <typename T>
int add_gemm(std::vector<GemmProbT<T>> &gp, std::vector<GemmSegT<T>> &gs, int &tiles,
             T *C, long long ldc, int M, int N, const std::vector<GemmSegT<T>> &segs,
             double &flops, int tb) {
  if (M <= 0 || N <= 0) return 0;
  GemmProbT<T> p;
  p.C = C;
  p.ldc = ldc;
  p.M = M;
  p.N = N;
  p.seg0 = (int)gs.size();
  p.nseg = 0;
  for (const GemmSegT<T> &s : segs)
    if (s.K > 0) {
      gs.push_back(s);
      p.nseg++;
      flops += 2.0 * M * N * s.K;
    }
  if (p.nseg == 0) return 0;
  p.tiles_n = (N + tb - 1) / tb;
  p.tile0 = tiles;
  tiles += ((M + tb - 1) / tb) * p.tiles_n;
  gp.push_back(p);
  return 1;
}
template int add_gemm<double>(std::vector<GemmProbT<double>> &, std::vector<GemmSegT<double>> &,
                              int &, double *, long long, int, int,
                              const std::vector<GemmSegT<double>> &, double &, int);
template int add_gemm<float>(std::vector<GemmProbT<float>> &, std::vector<GemmSegT<float>> &,
                             int &, float *, long long, int, int,
                             const std::vector<GemmSegT<float>> &, double &, int);

// ---------------------------------------------------------------------------
// tile-shape autotuner: per precision and size bucket of min(M, N), the
// faster of the 64 x 64 and 128 x 128 C -= AB kernels on a launch of many
// problems of the bucket's probe size (enough tiles to fill the device, as
// in the Schur-update launches), measured once per process on
// device-resident operands (both tiles compute every element with the same
// instruction sequence, so the choice never changes results)
// ---------------------------------------------------------------------------
namespace {
std::mutex g_tune_mu;
bool g_tuned[2] = {false, false};
int g_tile[2][3] = {{128, 128, 128}, {128, 128, 128}};
const int kBucketEdge[3] = {128, 512, 2048};
const int kBucketProbe[3] = {128, 1024, 4096};

template <typename T>
double time_gemm(int n, int count, int K, int tile, int reps) {
  T *dA = nullptr, *dB = nullptr, *dC = nullptr;
  GemmProbT<T> *dgp = nullptr;
  GemmSegT<T> *dgs = nullptr;
  double ms_out = -1;
  if (hipMalloc((void **)&dA, sizeof(T) * (size_t)n * K) == hipSuccess &&
      hipMalloc((void **)&dB, sizeof(T) * (size_t)K * n) == hipSuccess &&
      hipMalloc((void **)&dC, sizeof(T) * (size_t)n * n * count) == hipSuccess) {
    launch_fill(dA, (long long)n * K, nullptr);
    launch_fill(dB, (long long)K * n, nullptr);
    launch_fill(dC, (long long)n * n * count, nullptr);
    std::vector<GemmProbT<T>> gp;
    std::vector<GemmSegT<T>> gs;
    int tiles = 0;
    double fl = 0;
    for (int c = 0; c < count; c++)
      add_gemm<T>(gp, gs, tiles, dC + (size_t)n * n * c, n, n, n, {GemmSegT<T>{dA, dB, n, K, K}}, fl, tile);
    if (upload_list(&dgp, gp) == SNU_OK && upload_list(&dgs, gs) == SNU_OK) {
      auto run = [&]() {
        if (tile == 128) launch_gemm128<T>(dgp, count, dgs, tiles, 0, nullptr);
        else launch_gemm<T>(dgp, count, dgs, tiles, 0, nullptr);
      };
      hipEvent_t e0, e1;
      if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        run();
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; r++) run();
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) == hipSuccess) ms_out = ev_ms(e0, e1) / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
      }
    }
  }
  (void)hipDeviceSynchronize();
  if (dA) (void)hipFree(dA);
  if (dB) (void)hipFree(dB);
  if (dC) (void)hipFree(dC);
  if (dgp) (void)hipFree(dgp);
  if (dgs) (void)hipFree(dgs);
  return ms_out;
}
}  // namespace

template <typename T>
void autotune_tiles() {
  const int pi = std::is_same<T, float>::value ? 1 : 0;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_tuned[pi]) return;
  for (int b = 0; b < 3; b++) {
    const int n = kBucketProbe[b], count = (4096 / n) * (4096 / n);
    const double t64 = time_gemm<T>(n, count, 1024, 64, 3), t128 = time_gemm<T>(n, count, 1024, 128, 3);
    if (t64 > 0 && t128 > 0) g_tile[pi][b] = t128 < t64 ? 128 : 64;
  }
  g_tuned[pi] = true;
}
template void autotune_tiles<double>();
template void autotune_tiles<float>();

int tile_for(int prec, int M, int N) {
  const int m = std::min(M, N);
  if (m < kBucketEdge[0]) return 64;
  const int b = m < kBucketEdge[1] ? 0 : m < kBucketEdge[2] ? 1 : 2;
  return g_tile[prec == SNU_FP32 ? 1 : 0][b];
}
int tile_table(int prec, int *out3) {
  for (int b = 0; b < 3; b++) out3[b] = g_tile[prec == SNU_FP32 ? 1 : 0][b];
  return g_tuned[prec == SNU_FP32 ? 1 : 0] ? 1 : 0;
}

namespace {

// A grouped C -= sum A B launch, split by tile size: problems whose size
// bucket the autotuner gave to the 128 x 128 kernel take it when the launch
// holds at least kMin128Tiles of those tiles (16 per CU: on smaller launches
// the 64 x 64 tiles, four times as many and more workgroups per CU, balance
// better: 64^3 factor 37.2 -> 35.1 ms, 96^3 FP64 440 -> 433 ms against a
// threshold of 512, measured); the rest take 64 x 64 tiles (launch kind 5
// and kind 2).
constexpr int kMin128Tiles = 4096;
// FP32 factorizations whose 128 x 128 tiles run as bf16 splits
// (SNU_GEMM_SPLIT: about twice the rate of the FP32 64 x 64 tile) take the
// large tile for every problem with min(M, N) >= 128, from kMin128Split
// tiles per launch.
constexpr int kMin128Split = 512;
struct TileRule {
  int prec = SNU_FP64;
  bool split = false;
  int pick(int M, int N) const {
    if (split) return std::min(M, N) >= 128 ? 128 : 64;
    return tile_for(prec, M, N);
  }
  long long min128() const { return split ? kMin128Split : kMin128Tiles; }
};
template <typename T>
TileRule tile_rule(const snumat *h) {
  TileRule r;
  r.prec = std::is_same<T, float>::value ? SNU_FP32 : SNU_FP64;
  r.split = r.prec == SNU_FP32 && h->opts.fp32_gemm == SNU_GEMM_SPLIT;
  return r;
}
template <typename T>
struct GemmBatch {
  struct Pend {
    T *C;
    long long ldc;
    int M, N;
    std::vector<GemmSegT<T>> segs;
  };
  std::vector<Pend> pend;
  TileRule rule;
  void add(T *C, long long ldc, int M, int N, const std::vector<GemmSegT<T>> &segs) {
    if (M > 0 && N > 0) pend.push_back(Pend{C, ldc, M, N, segs});
  }
  void flush(std::vector<GemmProbT<T>> &gp, std::vector<GemmSegT<T>> &gs, std::vector<Launch> &out,
             long long key) {
    long long big = 0;
    for (const Pend &q : pend)
      if (rule.pick(q.M, q.N) == 128) big += (long long)((q.M + 127) / 128) * ((q.N + 127) / 128);
    const bool use128 = big >= rule.min128();
    std::vector<GemmProbT<T>> p[2];
    int tiles[2] = {0, 0};
    double flops[2] = {0, 0};
    for (const Pend &q : pend) {
      const int t = use128 && rule.pick(q.M, q.N) == 128 ? 1 : 0;
      add_gemm(p[t], gs, tiles[t], q.C, q.ldc, q.M, q.N, q.segs, flops[t], t ? 128 : 64);
    }
    pend.clear();
    for (int t = 0; t < 2; t++) {
      if (p[t].empty()) continue;
      Launch L{t ? 5 : 2, (int)gp.size(), (int)p[t].size(), tiles[t], flops[t], key};
      gp.insert(gp.end(), p[t].begin(), p[t].end());
      out.push_back(L);
    }
  }
};

// A grouped GEMV launch (the dense solves).
template <typename T>
struct GemvBatch {
  std::vector<GemvProbT<T>> p;
  int tiles = 0;
  double bytes = 0;
  void add(std::vector<GemvSegT<T>> &vs, T *y, long long ldy, int M, int tri, int nrhs,
           const std::vector<GemvSegT<T>> &segs) {
    if (M <= 0) return;
    GemvProbT<T> q;
    q.y = y;
    q.ldy = ldy;
    q.M = M;
    q.tri = tri;
    q.nrhs = nrhs;
    q.seg0 = (int)vs.size();
    q.nseg = 0;
    for (const GemvSegT<T> &g : segs)
      if (g.K > 0) {
        vs.push_back(g);
        q.nseg++;
        bytes += (double)sizeof(T) * M * g.K * (tri ? 0.5 : 1.0);
      }
    if (!q.nseg) return;
    q.tile0 = tiles;
    tiles += (M + 63) / 64;
    p.push_back(q);
  }
  void flush(int kind, std::vector<GemvProbT<T>> &vp, std::vector<Launch> &out, long long key) {
    if (p.empty()) return;
    out.push_back(Launch{kind, (int)vp.size(), (int)p.size(), tiles, bytes, key});
    vp.insert(vp.end(), p.begin(), p.end());
    p.clear();
    tiles = 0;
    bytes = 0;
  }
};

// Maximal runs [r0, r1) of 64-aligned tiles of [0, len) that hold a
// structural nonzero according to any(x0, x1) (half-open, relative).
template <class F>
std::vector<std::pair<int, int>> nz_runs(int len, F any) {
  std::vector<std::pair<int, int>> out;
  for (int t0 = 0; t0 < len; t0 += 64) {
    int t1 = std::min(len, t0 + 64);
    if (!any(t0, t1)) continue;
    if (!out.empty() && out.back().second == t0) out.back().second = t1;
    else out.push_back({t0, t1});
  }
  return out;
}

// Outer panel width of the blocked LU of the separators' diagonal blocks.
constexpr int kPanel = 512;

// Small fronts up to this order run in their own launch per height.
constexpr int kFrontLo = 64;

// Launch keys: tree height (1 = parents of the leaves) major, then the step
// inside the level.  Matrices of different depths align at equal height.
inline long long fkey(int height, long long step) { return (long long)height * 100000000LL + step; }

// Host-built launch descriptors of the dense factorization of one matrix.
// Per tree level, bottom-up:
//   for each 64-column step k of the diagonal blocks D: k_diag_lu (LU of the
//   diagonal block and the inverses of its triangles, written into the
//   separator's full L^-1 / U^-1), L21 <- L21 U11^-1 and U12 <- L11^-1 U12
//   inside D (GEMM by the block inverses), trailing update of D (GEMM);
//   then the full inverses L^-1, U^-1 of every separator of the level by
//   recursive doubling (two grouped GEMMs per doubling step);
//   then the panels in one pass each, Lp <- Lp U^-1 and Up <- L^-1 Up (GEMM
//   into scratch over the structurally nonzero rows / columns, K limited by
//   the triangle, then copied back): the panels are read and written once
//   instead of once per 64-column step;
//   then one grouped GEMM for all Schur updates of the ancestors,
//   K-concatenated over the level's separators below each of them.
template <typename T>
void build_dense_plan(snumat *h) {
  using GemmSeg = GemmSegT<T>;
  using DiagProb = DiagProbT<T>;
  const Plan &P = h->P;
  const TileRule rule = tile_rule<T>(h);
  Values<T> &VV = vals<T>(h);
  T *dense = VV.arena + P.leaf_total;
  auto R = [&](int p) { return dense + P.offR[p]; };
  auto Lp = [&](int p) { return dense + P.offL[p]; };
  auto Li = [&](int p) { return VV.inv + h->inv_off[p]; };
  auto Ui = [&](int p) { return VV.inv + h->inv_off[p] + (long long)P.block_size[p] * P.block_size[p]; };
  DensePlan<T> &F = VV.fact;
  auto &h_dp = F.dp;
  auto &h_gp = F.gp;
  auto &h_gs = F.gs;
  auto &out = F.launches;
  h_dp.clear();
  h_gp.clear();
  h_gs.clear();
  out.clear();
  for (int l = P.max_level - 1; l >= 0; l--) {
    const int height = P.max_level - l;
    int lo = 1 << l, hi = std::min((2 << l) - 1, P.np - 1);
    int smax = 0;
    for (int i = lo; i <= hi; i++) smax = std::max(smax, P.block_size[i]);
    for (int k = 0; k < smax; k += 64) {
      const long long ks = (long long)(k / 64) * 8;
      Launch d{0, (int)h_dp.size(), 0, 0, 0, fkey(height, ks + 0)};
      for (int i = lo; i <= hi; i++) {
        int s = P.block_size[i];
        if (s <= k) continue;
        int nb = std::min(64, s - k);
        DiagProb dp{};  // eps = 0: separator pivots are not clamped
        dp.D = R(i) + k + (long long)k * s;
        dp.ld = s;
        dp.nb = nb;
        dp.Linv = Li(i) + k + (long long)k * s;
        dp.Uinv = Ui(i) + k + (long long)k * s;
        dp.ldi = s;
        h_dp.push_back(dp);
        d.np++;
      }
      if (d.np) out.push_back(d);
      // L21 <- L21 U11^-1 (rows of D below the block) and U12 <- L11^-1 U12
      // (row block of D right of the diagonal block): independent products,
      // one grouped launch
      Launch tr{1, (int)h_gp.size(), 0, 0, 0, fkey(height, ks + 1)};
      for (int i = lo; i <= hi; i++) {
        int s = P.block_size[i];
        if (s <= k) continue;
        int nb = std::min(64, s - k);
        const T *U11i = Ui(i) + k + (long long)k * s;
        T *X1 = R(i) + (k + nb) + (long long)k * s;
        tr.np += add_gemm(h_gp, h_gs, tr.tiles, X1, s, s - k - nb, nb,
                          {GemmSeg{X1, U11i, s, s, nb}}, tr.flops);
        const T *L11i = Li(i) + k + (long long)k * s;
        T *X = R(i) + k + (long long)(k + nb) * s;
        tr.np += add_gemm(h_gp, h_gs, tr.tiles, X, s, nb, s - k - nb,
                          {GemmSeg{L11i, X, s, s, nb}}, tr.flops);
      }
      if (tr.np) out.push_back(tr);
      // trailing update of D, outer-blocked by kPanel columns: inside a
      // panel [K0, K1) a step updates only the L-shaped rest of the panel
      // (its columns below the step, its rows right of the panel); the last
      // step of the panel applies the panel's K1 - K0 rank-64 updates to the
      // trailing block D[K1:, K1:] as one product with K = K1 - K0 (a K = 64
      // update reads and writes the whole trailing block per step)
      GemmBatch<T> tu;
      tu.rule = rule;
      const int K0 = k - k % kPanel;
      for (int i = lo; i <= hi; i++) {
        const int s = P.block_size[i];
        if (s <= k) continue;
        const int nb = std::min(64, s - k), k2 = k + nb, K1 = std::min(s, K0 + kPanel);
        auto D = [&](int r, int c) { return R(i) + r + (long long)c * s; };
        if (k2 < K1) {
          tu.add(D(k2, k2), s, s - k2, K1 - k2, {GemmSeg{D(k2, k), D(k, k2), s, s, nb}});
          tu.add(D(k2, K1), s, K1 - k2, s - K1, {GemmSeg{D(k2, k), D(k, K1), s, s, nb}});
        } else {
          tu.add(D(K1, K1), s, s - K1, s - K1, {GemmSeg{D(K1, K0), D(K0, K1), s, s, K1 - K0}});
        }
      }
      tu.flush(h_gp, h_gs, out, fkey(height, ks + 3));
    }
    // full triangular inverses by recursive doubling:
    //   U: X[r0:, c0:] = -X[r0:, r0:] (U[r0:, c0:] X[c0:, c0:])
    //   L: X[c0:, r0:] = -X[c0:, c0:] (L[c0:, r0:] X[r0:, r0:])
    int step = 0;
    for (int hh = 64; hh < smax; hh *= 2, step++) {
      const long long base = 50000000LL + step * 4;
      Launch t1{1, (int)h_gp.size(), 0, 0, 0, fkey(height, base)};
      Launch t2{6, 0, 0, 0, 0, fkey(height, base + 1)};
      std::vector<GemmProbT<T>> fin;
      long long toff = 0;
      // per W-wide strip of each product, K limited to the triangle of the
      // triangular factor it meets (the rest of that factor is zero)
      long long big = 0;
      for (int i = lo; i <= hi; i++) {
        const int s = P.block_size[i];
        for (int r0 = 0; r0 + hh < s; r0 += 2 * hh) {
          const int hb = std::min(hh, s - r0 - hh);
          if (rule.pick(hh, hb) == 128) big += 2LL * ((hh + 127) / 128) * ((hb + 127) / 128);
        }
      }
      const int W = big >= rule.min128() ? 128 : 64;
      t1.kind = W == 128 ? 10 : 1;
      t2.kind = W == 128 ? 11 : 6;
      for (int i = lo; i <= hi; i++) {
        int s = P.block_size[i];
        for (int r0 = 0; r0 + hh < s; r0 += 2 * hh) {
          int c0 = r0 + hh, hb = std::min(hh, s - c0);
          T *TU = VV.tmp + toff;
          toff += (long long)hh * hb;
          T *TL = VV.tmp + toff;
          toff += (long long)hb * hh;
          // TU = U12 X22, X22 upper triangular: columns [n0, n0 + W) need
          // its rows [0, n0 + W)
          for (int n0 = 0; n0 < hb; n0 += W) {
            const int nt = std::min(W, hb - n0);
            t1.np += add_gemm(h_gp, h_gs, t1.tiles, TU + (long long)n0 * hh, hh, hh, nt,
                              {GemmSeg{R(i) + r0 + (long long)c0 * s, Ui(i) + c0 + (long long)(c0 + n0) * s,
                                       s, s, n0 + nt}},
                              t1.flops, W);
          }
          // TL = L21 X11, X11 lower triangular: columns [n0, n0 + W) need
          // its rows [n0, hh)
          for (int n0 = 0; n0 < hh; n0 += W) {
            const int nt = std::min(W, hh - n0);
            t1.np += add_gemm(h_gp, h_gs, t1.tiles, TL + (long long)n0 * hb, hb, hb, nt,
                              {GemmSeg{R(i) + c0 + (long long)(r0 + n0) * s,
                                       Li(i) + (r0 + n0) + (long long)(r0 + n0) * s, s, s, hh - n0}},
                              t1.flops, W);
          }
          // Ui[r0:, c0:] = -X11 TU, X11 upper triangular: rows [m0, m0 + W)
          // need its columns [m0, hh)
          for (int m0 = 0; m0 < hh; m0 += W) {
            const int mt = std::min(W, hh - m0);
            t2.np += add_gemm(fin, h_gs, t2.tiles, Ui(i) + r0 + m0 + (long long)c0 * s, s, mt, hb,
                              {GemmSeg{Ui(i) + (r0 + m0) + (long long)(r0 + m0) * s, TU + m0, s, hh, hh - m0}},
                              t2.flops, W);
          }
          // Li[c0:, r0:] = -X22 TL, X22 lower triangular: rows [m0, m0 + W)
          // need its columns [0, m0 + W)
          for (int m0 = 0; m0 < hb; m0 += W) {
            const int mt = std::min(W, hb - m0);
            t2.np += add_gemm(fin, h_gs, t2.tiles, Li(i) + c0 + m0 + (long long)r0 * s, s, mt, hh,
                              {GemmSeg{Li(i) + c0 + m0 + (long long)c0 * s, TL, s, hb, m0 + mt}},
                              t2.flops, W);
          }
        }
      }
      if (t1.np) out.push_back(t1);
      if (t2.np) {
        t2.p0 = (int)h_gp.size();
        h_gp.insert(h_gp.end(), fin.begin(), fin.end());
        out.push_back(t2);
      }
    }
    // panels: Lp <- Lp U^-1 per W-column strip of the result (U^-1 is upper
    // triangular: strip [n0, n0 + W) needs rows [0, n0 + W) of it), Up <-
    // L^-1 Up per W-row strip (rows [0, m0 + W) of Up), over the stored
    // (tile-compressed) rows / columns, into scratch; then copied back.
    // W = 128 (128 x 128 tiles) when the autotuner prefers that tile for the
    // level's panels and they hold enough such tiles to fill the device
    {
      long long big = 0;
      for (int i = lo; i <= hi; i++) {
        const int s = P.block_size[i], nl = P.nL[i], nu = P.nU[i];
        if (nl && rule.pick(nl, s) == 128) big += (long long)((nl + 127) / 128) * ((s + 127) / 128);
        if (nu && rule.pick(s, nu) == 128) big += (long long)((s + 127) / 128) * ((nu + 127) / 128);
      }
      const int W = big >= rule.min128() ? 128 : 64;
      Launch pm{W == 128 ? 10 : 1, (int)h_gp.size(), 0, 0, 0, fkey(height, 60000000LL)};
      std::vector<GemmProbT<T>> cp;
      Launch cl{3, 0, 0, 0, 0, fkey(height, 60000001LL)};
      double cfl = 0;
      long long toff = 0;
      for (int i = lo; i <= hi; i++) {
        const int s = P.block_size[i], m = P.nL[i], nc = P.nU[i];
        if (m && s) {
          T *dst = VV.tmp + toff;
          toff += (long long)m * s;
          T *src = Lp(i);
          for (int n0 = 0; n0 < s; n0 += W) {
            const int nt = std::min(W, s - n0), kk = n0 + nt;
            pm.np += add_gemm(h_gp, h_gs, pm.tiles, dst + (long long)n0 * m, m, m, nt,
                              {GemmSeg{src, Ui(i) + (long long)n0 * s, m, s, kk}}, pm.flops, W);
          }
          cl.np += add_gemm(cp, h_gs, cl.tiles, src, m, m, s, {GemmSeg{dst, dst, m, m, 1}}, cfl);
        }
        if (nc && s) {
          T *dst = VV.tmp + toff;
          toff += (long long)s * nc;
          T *src = R(i) + (long long)s * s;
          for (int m0 = 0; m0 < s; m0 += W) {
            const int mt = std::min(W, s - m0), kk = m0 + mt;
            pm.np += add_gemm(h_gp, h_gs, pm.tiles, dst + m0, s, mt, nc,
                              {GemmSeg{Li(i) + m0, src, s, s, kk}}, pm.flops, W);
          }
          cl.np += add_gemm(cp, h_gs, cl.tiles, src, s, s, nc, {GemmSeg{dst, dst, s, s, 1}}, cfl);
        }
      }
      if (pm.np) out.push_back(pm);
      if (cl.np) {
        cl.p0 = (int)h_gp.size();
        h_gp.insert(h_gp.end(), cp.begin(), cp.end());
        out.push_back(cl);
      }
    }
    // Schur updates of every ancestor, K-concatenated over the level's
    // separators below it (one grouped launch per level).  Emitted per
    // target tile (64 rows / columns aligned to the first index of the block
    // they belong to, the tiles of the compressed storage), with only the
    // contributors that store both the L-panel rows and the U-panel columns
    // of that tile (the other products are exactly zero).
    GemmBatch<T> sc;
    sc.rule = rule;
    for (int p = 1; p < lo; p++) {
      int lp = 31 - __builtin_clz(p);
      int first = p << (l - lp), last = ((p + 1) << (l - lp)) - 1;
      const int sp = P.block_size[p];
      // the tiles of p itself, then the stored tiles of p's ancestors (in
      // ancestor-space order; x: index in p's ancestor space), as
      // {position in the target, width, x}; x = -1 - t for p's own tile t
      std::vector<std::array<int, 3>> own, anL, anU;
      for (int t = 0; t < sp; t += 64) own.push_back({t, std::min(64, sp - t), -1 - t});
      for (int q = p / 2; q; q /= 2) {
        const int o = P.anc_off(p, q), sq = P.block_size[q];
        for (int t = 0; t < sq; t += 64) {
          const int w = std::min(64, sq - t);
          if (P.lmap[p][o + t] >= 0) anL.push_back({P.lmap[p][o + t], w, o + t});
          if (P.umap[p][o + t] >= 0) anU.push_back({sp + P.umap[p][o + t], w, o + t});
        }
      }
      for (int pass = 0; pass < 2; pass++) {  // 0: R_p = [D_p | Up_p], 1: Lp_p
        std::vector<std::array<int, 3>> rows = pass == 0 ? own : anL, cols = own;
        if (pass == 0) cols.insert(cols.end(), anU.begin(), anU.end());
        const int tm = (int)rows.size(), tn = (int)cols.size();
        if (!tm || !tn) continue;
        T *Cb = pass == 0 ? R(p) : Lp(p);
        const int ldc = pass == 0 ? sp : P.nL[p];
        // index of a tile in contributor i's ancestor space
        auto ix = [&](int i, const std::array<int, 3> &t) {
          return P.anc_off(i, p) + (t[2] < 0 ? -1 - t[2] : sp + t[2]);
        };
        // contributor set of every target tile
        std::vector<std::vector<int>> who((size_t)tm * tn);
        for (int a = 0; a < tm; a++)
          for (int b = 0; b < tn; b++)
            for (int i = first; i <= last; i++)
              if (P.lmap[i][ix(i, rows[a])] >= 0 && P.umap[i][ix(i, cols[b])] >= 0)
                who[(size_t)a * tn + b].push_back(i);
        auto segs_of = [&](const std::vector<int> &w, int a, int b) {
          std::vector<GemmSeg> segs;
          for (int i : w) {
            const int si = P.block_size[i];
            segs.push_back(GemmSeg{Lp(i) + P.lmap[i][ix(i, rows[a])],
                                   R(i) + (long long)(si + P.umap[i][ix(i, cols[b])]) * si, P.nL[i], si, si});
          }
          return segs;
        };
        auto emit = [&](int a, int b, int ta, int tb) {
          const std::vector<int> &w = who[(size_t)a * tn + b];
          int mt = 0, nt = 0;
          for (int x = 0; x < ta; x++) mt += rows[a + x][1];
          for (int y = 0; y < tb; y++) nt += cols[b + y][1];
          sc.add(Cb + rows[a][0] + (long long)cols[b][0] * ldc, ldc, mt, nt, segs_of(w, a, b));
        };
        auto nonempty = [&](int a, int b) { return a < tm && b < tn && !who[(size_t)a * tn + b].empty(); };
        // tiles a, a + 1 (rows) are consecutive in every contributor's L
        // panel, b, b + 1 (columns) in its U panel
        auto merge_ok = [&](int a, int b) {
          for (int i : who[(size_t)a * tn + b])
            if (P.lmap[i][ix(i, rows[a + 1])] != P.lmap[i][ix(i, rows[a])] + rows[a][1] ||
                P.umap[i][ix(i, cols[b + 1])] != P.umap[i][ix(i, cols[b])] + cols[b][1])
              return false;
          return true;
        };
        // one problem per target tile with its contributor set, or per 2 x 2
        // tiles when the four share one set and are contiguous in every
        // contributor's panels (the tile kernel is then chosen per size
        // bucket), emitted in strips of 8 tile rows, two tile columns at a
        // time: consecutive tiles (which one XCD runs together) share A rows
        // (same tile row and contributor) and B columns (same tile column),
        // so their operands meet in that XCD's L2
        constexpr int kStrip = 8;
        for (int a0 = 0; a0 < tm; a0 += kStrip)
          for (int b = 0; b < tn; b += 2)
            for (int a = a0; a < std::min(tm, a0 + kStrip); a += 2) {
              const std::vector<int> &w = who[(size_t)a * tn + b];
              if (nonempty(a, b) && nonempty(a + 1, b) && nonempty(a, b + 1) && nonempty(a + 1, b + 1) &&
                  who[(size_t)(a + 1) * tn + b] == w && who[(size_t)a * tn + b + 1] == w &&
                  who[(size_t)(a + 1) * tn + b + 1] == w && merge_ok(a, b)) {
                emit(a, b, 2, 2);
                continue;
              }
              for (int q = 0; q < 4; q++) {
                const int x = a + (q & 1), y = b + (q >> 1);
                if (nonempty(x, y)) emit(x, y, 1, 1);
              }
            }
      }
    }
    sc.flush(h_gp, h_gs, out, fkey(height, 90000000LL));
  }
}

// Scratch needed by the products of the triangular inversion (max over the
// doubling steps of every level) and by the panel products of a level.
long long inversion_tmp_size(const Plan &P) {
  long long best = 0;
  for (int l = P.max_level - 1; l >= 0; l--) {
    int lo = 1 << l, hi = std::min((2 << l) - 1, P.np - 1);
    int smax = 0;
    for (int i = lo; i <= hi; i++) smax = std::max(smax, P.block_size[i]);
    long long panels = 0;
    for (int i = lo; i <= hi; i++) panels += (long long)(P.nL[i] + P.nU[i]) * P.block_size[i];
    best = std::max(best, panels);
    for (int hh = 64; hh < smax; hh *= 2) {
      long long t = 0;
      for (int i = lo; i <= hi; i++) {
        int s = P.block_size[i];
        for (int r0 = 0; r0 + hh < s; r0 += 2 * hh) t += 2LL * hh * std::min(hh, s - r0 - hh);
      }
      best = std::max(best, t);
    }
  }
  return best;
}

}  // namespace

// Dense solves of one matrix, per level (y in ys, x overwriting b in bp):
//   forward (bottom-up):  y_i = L_i^-1 b_i;   b_p -= sum_i Lp_i[rows p] y_i
//   backward (top-down):  y_i -= Up_i x_anc;  x_i = U_i^-1 y_i
// Keys: forward launches ascend in height, backward ones descend, all
// backward after all forward.
template <typename T>
void build_solve_plan(const snumat *h, T *bp, T *ys, int nrhs, GemvPlan<T> &S) {
  using GemvSeg = GemvSegT<T>;
  const Values<T> &VV = vals<T>(h);
  const Plan &P = h->P;
  T *dense = VV.arena + P.leaf_total;
  const long long n = P.n;
  auto R = [&](int p) { return dense + P.offR[p]; };
  auto Lp = [&](int p) { return dense + P.offL[p]; };
  auto Li = [&](int p) { return VV.inv + h->inv_off[p]; };
  auto Ui = [&](int p) { return VV.inv + h->inv_off[p] + (long long)P.block_size[p] * P.block_size[p]; };
  auto bseg = [&](int p) { return bp + P.block_start[p]; };
  auto yseg = [&](int p) { return ys + P.block_start[p]; };
  S.vp.clear();
  S.vs.clear();
  S.fwd.clear();
  S.bwd.clear();
  GemvBatch<T> g;
  for (int l = P.max_level - 1; l >= 0; l--) {
    const int height = P.max_level - l;
    int lo = 1 << l, hi = std::min((2 << l) - 1, P.np - 1);
    for (int i = lo; i <= hi; i++) {
      int s = P.block_size[i];
      g.add(S.vs, yseg(i), n, s, 1, nrhs, {GemvSeg{Li(i), bseg(i), s, n, s}});
    }
    g.flush(7, S.vp, S.fwd, height * 4LL + 0);
    for (int p = 1; p < lo; p++) {
      int lp = 31 - __builtin_clz(p);
      int first = p << (l - lp), last = ((p + 1) << (l - lp)) - 1;
      int sp = P.block_size[p];
      for (int m0 = 0; m0 < sp; m0 += 64) {
        int mt = std::min(64, sp - m0);
        std::vector<GemvSeg> segs;
        for (int i = first; i <= last; i++) {
          const int r = P.lmap[i][P.anc_off(i, p) + m0];  // the tile's stored rows (or -1)
          if (r < 0) continue;
          segs.push_back(GemvSeg{Lp(i) + r, yseg(i), P.nL[i], n, P.block_size[i]});
        }
        g.add(S.vs, bseg(p) + m0, n, mt, 0, nrhs, segs);
      }
    }
    g.flush(8, S.vp, S.fwd, height * 4LL + 1);
  }
  for (int l = 0; l < P.max_level; l++) {
    const int height = P.max_level - l;
    const long long bkey = 1000000000LL - height * 4LL;
    int lo = 1 << l, hi = std::min((2 << l) - 1, P.np - 1);
    for (int i = lo; i <= hi; i++) {
      int s = P.block_size[i];
      std::vector<GemvSeg> segs;
      for (int a = i / 2; a >= 1; a /= 2) {
        int off = P.anc_off(i, a);
        // stored column tiles of a, merged into runs (consecutive stored
        // tiles of one ancestor are consecutive columns of the panel)
        for (auto cr : nz_runs(P.block_size[a], [&](int x0, int x1) { return P.umap[i][off + x0] >= 0; }))
          segs.push_back(GemvSeg{R(i) + (long long)(s + P.umap[i][off + cr.first]) * s, bseg(a) + cr.first, s, n,
                                 cr.second - cr.first});
      }
      g.add(S.vs, yseg(i), n, s, 0, nrhs, segs);
    }
    g.flush(8, S.vp, S.bwd, bkey + 0);
    for (int i = lo; i <= hi; i++) {
      int s = P.block_size[i];
      g.add(S.vs, bseg(i), n, s, 2, nrhs, {GemvSeg{Ui(i), yseg(i), s, n, s}});
    }
    g.flush(7, S.vp, S.bwd, bkey + 1);
  }
}

// Split-K layout: pieces per problem from its longest segment (the problem
// alone decides), workgroup and partial-sum offsets per launch; a launch
// whose problems are all single-piece runs unsplit.  Returns the partial-sum
// values the largest launch needs (launches run in turn and share them).
template <typename T>
size_t layout_gemv_plan(GemvPlan<T> &g) {
  size_t need_part = 0;
  for (std::vector<Launch> *ls : {&g.fwd, &g.bwd})
    for (Launch &L : *ls) {
      int wgs = 0, split = 0;
      long long off = 0;
      for (int q = 0; q < L.np; q++) {
        GemvProbT<T> &p = g.vp[L.p0 + q];
        int kmax = 1;
        for (int j = 0; j < p.nseg; j++) kmax = std::max(kmax, g.vs[p.seg0 + j].K);
        p.parts = (kmax + kGemvPiece - 1) / kGemvPiece;
        p.wg0 = wgs;
        p.poff = off;
        const int tiles = (p.M + 63) / 64;
        wgs += tiles * p.parts;
        off += (long long)tiles * p.parts * p.nrhs * 64;
        split |= p.parts > 1;
      }
      L.wgs = split ? wgs : 0;
      L.part = split ? off : 0;
      need_part = std::max(need_part, (size_t)L.part);
    }
  return need_part;
}

template <typename T>
int upload_gemv_plan(GemvPlan<T> &g, hipStream_t st) {
  auto grow = [](auto **p, size_t &cap, size_t need) -> int {
    if (need <= cap && *p) return SNU_OK;
    if (*p) HIPCHK(hipFree(*p));
    *p = nullptr;
    const size_t c = std::max<size_t>(need + need / 2, 64);
    HIPCHK(hipMalloc((void **)p, c * sizeof(**p)));
    cap = c;
    return SNU_OK;
  };
  const size_t need_part = layout_gemv_plan(g);
  int rc;
  if ((rc = grow(&g.d_vp, g.cap_vp, g.vp.size())) || (rc = grow(&g.d_vs, g.cap_vs, g.vs.size())) ||
      (need_part && (rc = grow(&g.d_part, g.cap_part, need_part))))
    return rc;
  if (!g.vp.empty())
    HIPCHK(hipMemcpyAsync(g.d_vp, g.vp.data(), g.vp.size() * sizeof(g.vp[0]), hipMemcpyHostToDevice, st));
  if (!g.vs.empty())
    HIPCHK(hipMemcpyAsync(g.d_vs, g.vs.data(), g.vs.size() * sizeof(g.vs[0]), hipMemcpyHostToDevice, st));
  return SNU_OK;
}
template <typename T>
void free_gemv_plan(GemvPlan<T> &g) {
  if (g.d_vp) (void)hipFree(g.d_vp);
  if (g.d_vs) (void)hipFree(g.d_vs);
  if (g.d_part) (void)hipFree(g.d_part);
  g = GemvPlan<T>();
}
template <typename T>
int upload_dense_plan(DensePlan<T> &d) {
  int rc;
  if ((rc = upload_list(&d.d_dp, d.dp)) || (rc = upload_list(&d.d_gp, d.gp)) ||
      (rc = upload_list(&d.d_gs, d.gs)))
    return rc;
  return SNU_OK;
}
template <typename T>
void free_dense_plan(DensePlan<T> &d) {
  if (d.d_dp) (void)hipFree(d.d_dp);
  if (d.d_gp) (void)hipFree(d.d_gp);
  if (d.d_gs) (void)hipFree(d.d_gs);
  d = DensePlan<T>();
}
// ---------------------------------------------------------------------------
// merging the launch lists of several matrices
// ---------------------------------------------------------------------------
namespace {
struct Ref {
  long long key;
  int kind, src, idx;
};
template <class GetLaunches>
std::vector<std::vector<Ref>> group_launches(int nsrc, GetLaunches get, bool descending = false) {
  std::vector<Ref> all;
  for (int s = 0; s < nsrc; s++) {
    const std::vector<Launch> &L = get(s);
    for (int i = 0; i < (int)L.size(); i++) all.push_back(Ref{L[i].key, L[i].kind, s, i});
  }
  std::stable_sort(all.begin(), all.end(), [&](const Ref &a, const Ref &b) {
    if (a.key != b.key) return descending ? a.key > b.key : a.key < b.key;
    return a.kind < b.kind;
  });
  std::vector<std::vector<Ref>> groups;
  for (size_t i = 0; i < all.size(); i++) {
    if (i == 0 || all[i].key != all[i - 1].key || all[i].kind != all[i - 1].kind) groups.emplace_back();
    groups.back().push_back(all[i]);
  }
  return groups;
}
}  // namespace

template <typename T>
void merge_dense(const std::vector<const DensePlan<T> *> &in, DensePlan<T> &out) {
  out.dp.clear();
  out.gp.clear();
  out.gs.clear();
  out.launches.clear();
  auto groups = group_launches((int)in.size(), [&](int s) -> const std::vector<Launch> & {
    return in[s]->launches;
  });
  for (auto &g : groups) {
    Launch M{g[0].kind, 0, 0, 0, 0, g[0].key};
    if (M.kind == 0) {
      M.p0 = (int)out.dp.size();
      for (const Ref &r : g) {
        const DensePlan<T> &D = *in[r.src];
        const Launch &L = D.launches[r.idx];
        out.dp.insert(out.dp.end(), D.dp.begin() + L.p0, D.dp.begin() + L.p0 + L.np);
        M.np += L.np;
      }
    } else {
      M.p0 = (int)out.gp.size();
      for (const Ref &r : g) {
        const DensePlan<T> &D = *in[r.src];
        const Launch &L = D.launches[r.idx];
        for (int q = 0; q < L.np; q++) {
          GemmProbT<T> p = D.gp[L.p0 + q];
          int s0 = (int)out.gs.size();
          out.gs.insert(out.gs.end(), D.gs.begin() + p.seg0, D.gs.begin() + p.seg0 + p.nseg);
          p.seg0 = s0;
          p.tile0 += M.tiles;
          out.gp.push_back(p);
        }
        M.np += L.np;
        M.tiles += L.tiles;
        M.flops += L.flops;
      }
    }
    out.launches.push_back(M);
  }
}

template <typename T>
void merge_gemv(const std::vector<const GemvPlan<T> *> &in, GemvPlan<T> &out) {
  out.vp.clear();
  out.vs.clear();
  out.fwd.clear();
  out.bwd.clear();
  for (int dir = 0; dir < 2; dir++) {
    auto groups = group_launches((int)in.size(), [&](int s) -> const std::vector<Launch> & {
      return dir == 0 ? in[s]->fwd : in[s]->bwd;
    });
    for (auto &g : groups) {
      Launch M{g[0].kind, (int)out.vp.size(), 0, 0, 0, g[0].key};
      for (const Ref &r : g) {
        const GemvPlan<T> &D = *in[r.src];
        const Launch &L = (dir == 0 ? D.fwd : D.bwd)[r.idx];
        for (int q = 0; q < L.np; q++) {
          GemvProbT<T> p = D.vp[L.p0 + q];
          int s0 = (int)out.vs.size();
          out.vs.insert(out.vs.end(), D.vs.begin() + p.seg0, D.vs.begin() + p.seg0 + p.nseg);
          p.seg0 = s0;
          p.tile0 += M.tiles;
          out.vp.push_back(p);
        }
        M.np += L.np;
        M.tiles += L.tiles;
        M.flops += L.flops;
      }
      (dir == 0 ? out.fwd : out.bwd).push_back(M);
    }
  }
}

// ---------------------------------------------------------------------------
// device layout of one matrix
// ---------------------------------------------------------------------------
namespace {
template <typename T>
int alloc_values(snumat *h, long long S_total) {
  Values<T> &VV = vals<T>(h);
  int rc;
  if ((rc = dalloc(h, &VV.arena, (size_t)h->arena_len)) || (rc = dalloc(h, &VV.S, (size_t)S_total)) ||
      (rc = dalloc(h, &VV.inv, (size_t)h->inv_total)) ||
      (rc = dalloc(h, &VV.tmp, (size_t)std::max(1LL, h->tmp_total))))
    return rc;
  build_dense_plan<T>(h);
  if ((rc = upload_dense_plan(VV.fact))) return rc;
  // leaf descriptors point at this value type's arrays
  for (size_t r = 0; r < h->h_leaves.size(); r++) {
    LeafDev &D = h->h_leaves[r];
    D.vals = VV.arena + h->P.leaf_off[r];
    D.dense = VV.arena + h->P.leaf_total;
  }
  return SNU_OK;
}
}  // namespace

// Multifrontal plan of one matrix's leaves (value arrays allocated): the
// front value arena, the integer tables (owned slots and positions,
// extend-add indices, the root fronts' boundary rows / columns), child and
// root descriptors, per-height front tables, the larger fronts' blocked
// partial LU (per 64-column step: diagonal block LU with the pivot clamp,
// both triangular solves by the block inverses, the trailing update of the
// whole remaining front) and the boundary Schur tiles.
template <typename T>
int build_front_plan(snumat *h) {
  using GemmSeg = GemmSegT<T>;
  Values<T> &VV = vals<T>(h);
  const Plan &P = h->P;
  const int np = P.np, fmax = front_small_max<T>();
  const T eps = (T)h->opts.eps;
  FrontPlan<T> &FP = VV.fronts;
  free_dense_plan(FP.dense);
  FP = FrontPlan<T>();
  int rc;
  // integer tables and front value offsets
  std::vector<int> ints;
  std::vector<size_t> o_slot(np), o_pos(np), o_rel(np);
  std::vector<std::vector<long long>> foff(np);
  std::vector<std::vector<std::pair<size_t, size_t>>> o_root(np);  // per root: arow, acol offsets
  long long ftotal = 0;
  int levels = 0;
  for (int r = 0; r < np; r++) {
    const LeafPlan &L = P.leaves[r];
    o_slot[r] = ints.size();
    ints.insert(ints.end(), L.fslot.begin(), L.fslot.end());
    o_pos[r] = ints.size();
    ints.insert(ints.end(), L.fpos.begin(), L.fpos.end());
    o_rel[r] = ints.size();
    ints.insert(ints.end(), L.frel.begin(), L.frel.end());
    const int nf = (int)L.ff.size();
    foff[r].resize(nf);
    for (int f = 0; f < nf; f++) {
      foff[r][f] = ftotal;
      ftotal += (long long)L.ff[f] * L.ff[f];
      if (L.ff[f] > fmax) ftotal += 2 * 64 * 64;  // inverses of the diagonal blocks
      if (L.fpar[f] >= 0) continue;
      std::vector<int> arow(L.nbr, -1), acol(L.nbc, -1);
      for (int q = L.fidx_ptr[f]; q < L.fidx_ptr[f + 1]; q++) {
        if (L.fsrow[q] >= 0) arow[L.fsrow[q]] = q - L.fidx_ptr[f];
        if (L.fscol[q] >= 0) acol[L.fscol[q]] = q - L.fidx_ptr[f];
      }
      const size_t oa = ints.size();
      ints.insert(ints.end(), arow.begin(), arow.end());
      const size_t oc = ints.size();
      ints.insert(ints.end(), acol.begin(), acol.end());
      o_root[r].push_back({oa, oc});
    }
    if (nf) levels = std::max(levels, L.fheight_max + 1);
  }
  int *d_ints = nullptr;
  if ((rc = dupload(h, &d_ints, ints)) || (rc = dalloc(h, &VV.farena, (size_t)std::max(1LL, ftotal)))) return rc;
  T *fa = VV.farena;
  // children
  std::vector<FrontChild> ch;
  std::vector<std::vector<int>> ch0(np);
  for (int r = 0; r < np; r++) {
    const LeafPlan &L = P.leaves[r];
    const int nf = (int)L.ff.size();
    ch0[r].resize(nf);
    for (int f = 0; f < nf; f++) {
      ch0[r][f] = (int)ch.size();
      for (int q = L.fchild_ptr[f]; q < L.fchild_ptr[f + 1]; q++) {
        const int c = L.fchild[q], fc = L.ff[c], nkc = L.fnk[c];
        ch.push_back(FrontChild{fa + foff[r][c] + nkc + (long long)nkc * fc,
                                d_ints + o_rel[r] + L.fidx_ptr[c], fc, fc - nkc});
      }
    }
  }
  FrontChild *d_ch = nullptr;
  if ((rc = dupload(h, &d_ch, ch))) return rc;
  // fronts by height
  std::vector<std::vector<FrontDev>> sm(levels), lg(levels);
  std::vector<std::vector<std::pair<int, int>>> lgid(levels);  // (leaf, front) of each larger front
  for (int r = 0; r < np; r++) {
    const LeafPlan &L = P.leaves[r];
    for (int f = 0; f < (int)L.ff.size(); f++) {
      FrontDev D;
      D.F = fa + foff[r][f];
      D.vals = VV.arena + P.leaf_off[r];
      D.slot = d_ints + o_slot[r] + L.fslot_ptr[f];
      D.pos = d_ints + o_pos[r] + L.fslot_ptr[f];
      D.child = d_ch + ch0[r][f];
      D.f = L.ff[f];
      D.nk = L.fnk[f];
      D.nslot = L.fslot_ptr[f + 1] - L.fslot_ptr[f];
      D.nchild = L.fchild_ptr[f + 1] - L.fchild_ptr[f];
      (D.f <= fmax ? sm : lg)[L.fheight[f]].push_back(D);
      if (D.f > fmax) lgid[L.fheight[f]].push_back({r, f});
    }
  }
  FP.levels.assign(levels, FrontLevel());
  DensePlan<T> &DP = FP.dense;
  for (int hh = 0; hh < levels; hh++) {
    FrontLevel &lv = FP.levels[hh];
    lv.small0 = (int)FP.small.size();
    lv.nsmall = (int)sm[hh].size();
    // fronts of order <= kFrontLo first: their launch needs a quarter of the
    // LDS or less, so more of them share a CU
    std::stable_partition(sm[hh].begin(), sm[hh].end(), [](const FrontDev &D) { return D.f <= kFrontLo; });
    for (const FrontDev &D : sm[hh]) {
      if (D.f <= kFrontLo) lv.nsmall_lo++, lv.max_f_lo = std::max(lv.max_f_lo, D.f);
      lv.max_f = std::max(lv.max_f, D.f);
    }
    FP.small.insert(FP.small.end(), sm[hh].begin(), sm[hh].end());
    lv.work0 = (int)FP.work.size();
    const int lbase = (int)FP.large.size();
    int kmax = 0;
    for (size_t q = 0; q < lg[hh].size(); q++) {
      const FrontDev &D = lg[hh][q];
      const LeafPlan &L = P.leaves[lgid[hh][q].first];
      const int fr = lgid[hh][q].second;
      const int *pos = L.fpos.data() + L.fslot_ptr[fr];
      for (int c0 = 0; c0 < D.f; c0 += 64) {
        const int c1 = std::min(D.f, c0 + 64);
        FP.work.push_back(make_int4(lbase + (int)q, c0, (int)FP.wrange.size(), 0));
        FP.wrange.push_back((int)(std::lower_bound(pos, pos + D.nslot, c0 * D.f) - pos));
        FP.wrange.push_back((int)(std::lower_bound(pos, pos + D.nslot, c1 * D.f) - pos));
        for (int k = L.fchild_ptr[fr]; k < L.fchild_ptr[fr + 1]; k++) {
          const int c = L.fchild[k];
          const int *rel = L.frel.data() + L.fidx_ptr[c], nb = L.fidx_ptr[c + 1] - L.fidx_ptr[c];
          FP.wrange.push_back((int)(std::lower_bound(rel, rel + nb, c0) - rel));
          FP.wrange.push_back((int)(std::lower_bound(rel, rel + nb, c1) - rel));
        }
      }
      kmax = std::max(kmax, D.nk);
    }
    lv.nwork = (int)FP.work.size() - lv.work0;
    FP.large.insert(FP.large.end(), lg[hh].begin(), lg[hh].end());
    // blocked partial LU of this height's larger fronts, grouped per step
    for (int k = 0; k < kmax; k += 64) {
      const int step = k / 64;
      Launch d{0, (int)DP.dp.size(), 0, 0, 0, front_key(hh, step, 0)};
      Launch tr{1, (int)DP.gp.size(), 0, 0, 0, front_key(hh, step, 1)};
      GemmBatch<T> tu;
      tu.rule = tile_rule<T>(h);
      for (const FrontDev &D : lg[hh]) {
        if (D.nk <= k) continue;
        const int f = D.f, nb = std::min(64, D.nk - k), m = f - k - nb;
        T *F = (T *)D.F;
        T *Li = F + (long long)f * f, *Ui = Li + 64 * 64;
        DiagProbT<T> dp{};
        dp.D = F + k + (long long)k * f;
        dp.ld = f;
        dp.nb = nb;
        dp.Linv = Li;
        dp.Uinv = Ui;
        dp.ldi = 64;
        dp.eps = eps;
        DP.dp.push_back(dp);
        d.np++;
        // L21 <- L21 U11^-1 (rows below), U12 <- L11^-1 U12 (columns right)
        T *X1 = F + (k + nb) + (long long)k * f;
        tr.np += add_gemm(DP.gp, DP.gs, tr.tiles, X1, f, m, nb, {GemmSeg{X1, Ui, f, 64, nb}}, tr.flops);
        T *X = F + k + (long long)(k + nb) * f;
        tr.np += add_gemm(DP.gp, DP.gs, tr.tiles, X, f, nb, m, {GemmSeg{Li, X, 64, f, nb}}, tr.flops);
        // trailing update of the remaining front, outer-blocked by kPanel
        // pivots as for the separators (the last step of a panel applies
        // its rank-64 updates to F[K1:, K1:] at once)
        const int k2 = k + nb, K0 = k - k % kPanel, K1 = std::min(D.nk, K0 + kPanel);
        auto at = [&](int r, int c) { return F + r + (long long)c * f; };
        if (k2 < K1) {
          tu.add(at(k2, k2), f, f - k2, K1 - k2, {GemmSeg{X1, X, f, f, nb}});
          tu.add(at(k2, K1), f, K1 - k2, f - K1, {GemmSeg{X1, at(k, K1), f, f, nb}});
        } else {
          tu.add(at(K1, K1), f, f - K1, f - K1, {GemmSeg{at(K1, K0), at(K0, K1), f, f, K1 - K0}});
        }
      }
      if (d.np) DP.launches.push_back(d);
      if (tr.np) DP.launches.push_back(tr);
      tu.flush(DP.gp, DP.gs, DP.launches, front_key(hh, step, 3));
    }
  }
  if ((rc = upload_dense_plan(DP))) return rc;
  // boundary Schur blocks from the root fronts
  std::vector<FrontRoot> roots;
  for (int r = 0; r < np; r++) {
    const LeafPlan &L = P.leaves[r];
    const int root0 = (int)roots.size();
    int k = 0;
    for (int f = 0; f < (int)L.ff.size(); f++) {
      if (L.fpar[f] >= 0) continue;
      const int ff = L.ff[f], nk = L.fnk[f];
      roots.push_back(FrontRoot{fa + foff[r][f] + nk + (long long)nk * ff, d_ints + o_root[r][k].first,
                                d_ints + o_root[r][k].second, ff});
      k++;
    }
    if (L.nbr > 0 && L.nbc > 0) {
      const int li = (int)FP.ls.size();
      FP.ls.push_back(LeafSchurDev{h->h_leaves[r].S, nullptr, L.nbr, L.nbc, (int)roots.size() - root0});
      FP.ls.back().root = (const FrontRoot *)(uintptr_t)root0;  // index until the table is uploaded
      for (int r0 = 0; r0 < L.nbr; r0 += 64)
        for (int c0 = 0; c0 < L.nbc; c0 += 64) FP.tiles.push_back(make_int4(li, r0, c0, 0));
    }
  }
  FrontRoot *d_roots = nullptr;
  if ((rc = dupload(h, &d_roots, roots))) return rc;
  for (LeafSchurDev &l : FP.ls) l.root = d_roots + (uintptr_t)l.root;
  if ((rc = dupload(h, &FP.d_small, FP.small)) || (rc = dupload(h, &FP.d_large, FP.large)) ||
      (rc = dupload(h, &FP.d_work, FP.work)) || (rc = dupload(h, &FP.d_wrange, FP.wrange)) ||
      (rc = dupload(h, &FP.d_ls, FP.ls)) ||
      (rc = dupload(h, &FP.d_tiles, FP.tiles)))
    return rc;
  return SNU_OK;
}

template <typename T>
void merge_fronts(const std::vector<const FrontPlan<T> *> &in, FrontPlan<T> &out) {
  out.small.clear();
  out.large.clear();
  out.work.clear();
  out.wrange.clear();
  std::vector<int> rbase;
  for (const FrontPlan<T> *p : in) {
    rbase.push_back((int)out.wrange.size());
    out.wrange.insert(out.wrange.end(), p->wrange.begin(), p->wrange.end());
  }
  out.ls.clear();
  out.tiles.clear();
  size_t levels = 0;
  for (const FrontPlan<T> *p : in) levels = std::max(levels, p->levels.size());
  out.levels.assign(levels, FrontLevel());
  for (size_t hh = 0; hh < levels; hh++) {
    FrontLevel &lv = out.levels[hh];
    lv.small0 = (int)out.small.size();
    lv.work0 = (int)out.work.size();
    for (size_t pi = 0; pi < in.size(); pi++) {
      const FrontPlan<T> *p = in[pi];
      if (hh >= p->levels.size()) continue;
      const FrontLevel &pl = p->levels[hh];
      out.small.insert(out.small.end(), p->small.begin() + pl.small0, p->small.begin() + pl.small0 + pl.nsmall_lo);
      lv.max_f = std::max(lv.max_f, pl.max_f);
      lv.max_f_lo = std::max(lv.max_f_lo, pl.max_f_lo);
      // this matrix's larger fronts of the height, in the order its work items name them
      int first = -1, last = -1;
      for (int q = pl.work0; q < pl.work0 + pl.nwork; q++) {
        if (first < 0) first = p->work[q].x;
        last = p->work[q].x;
      }
      if (first < 0) continue;
      const int base = (int)out.large.size();
      out.large.insert(out.large.end(), p->large.begin() + first, p->large.begin() + last + 1);
      for (int q = pl.work0; q < pl.work0 + pl.nwork; q++)
        out.work.push_back(make_int4(base + p->work[q].x - first, p->work[q].y, rbase[pi] + p->work[q].z, 0));
    }
    lv.nsmall_lo = (int)out.small.size() - lv.small0;
    for (const FrontPlan<T> *p : in)  // then every matrix's larger small fronts
      if (hh < p->levels.size()) {
        const FrontLevel &pl = p->levels[hh];
        out.small.insert(out.small.end(), p->small.begin() + pl.small0 + pl.nsmall_lo,
                         p->small.begin() + pl.small0 + pl.nsmall);
      }
    lv.nsmall = (int)out.small.size() - lv.small0;
    lv.nwork = (int)out.work.size() - lv.work0;
  }
  std::vector<const DensePlan<T> *> dp;
  for (const FrontPlan<T> *p : in) dp.push_back(&p->dense);
  merge_dense(dp, out.dense);
  for (const FrontPlan<T> *p : in) {
    const int base = (int)out.ls.size();
    out.ls.insert(out.ls.end(), p->ls.begin(), p->ls.end());
    for (int4 t : p->tiles) out.tiles.push_back(make_int4(base + t.x, t.y, t.z, t.w));
  }
}
template void merge_fronts<double>(const std::vector<const FrontPlan<double> *> &, FrontPlan<double> &);
template void merge_fronts<float>(const std::vector<const FrontPlan<float> *> &, FrontPlan<float> &);
template int build_front_plan<double>(snumat *);
template int build_front_plan<float>(snumat *);

int upload(snumat *h) {
  Plan &P = h->P;
  int rc;
  HIPCHK(hipSetDevice(h->device));
  if (h->prec == SNU_FP32) autotune_tiles<float>();
  else autotune_tiles<double>();
  const int np = P.np;
  // ---- leaf arrays (concatenated; descriptors hold pointers into them) ----
  std::vector<int> rowptr, diag, levptr, levrows, ulevptr, ulevrows;
  std::vector<int> brow, bcol_glob, bcol_blk, bcol_loc, wcol, ftask, btask, ubnd;
  struct Off { size_t rp, idx, lev, levr, ulev, ulevr, brow, bcol, task, row; long long S; };
  std::vector<Off> off(np);
  long long S_total = 0;
  h->h_leaves.assign(np, LeafDev());
  int maxw = 1, cb = 0;
  for (int r = 0; r < np; r++) {
    const LeafPlan &L = P.leaves[r];
    LeafDev &D = h->h_leaves[r];
    Off &o = off[r];
    o.rp = rowptr.size();
    o.idx = wcol.size();
    o.lev = levptr.size();
    o.levr = levrows.size();
    o.ulev = ulevptr.size();
    o.ulevr = ulevrows.size();
    o.brow = brow.size();
    o.bcol = bcol_glob.size();
    o.task = ftask.size();
    o.row = ubnd.size();
    o.S = S_total;
    S_total += (long long)L.nbr * L.nbc;
    rowptr.insert(rowptr.end(), L.lu_rowptr.begin(), L.lu_rowptr.end());
    diag.insert(diag.end(), L.lu_diag.begin(), L.lu_diag.end());
    diag.push_back(0);  // diag shares rowptr's offsets (core_n + 1 entries)
    // level-ordered solve tasks {row, first entry, end, diagonal}: forward
    // over the L part, backward over the U entries on leaf columns (they
    // precede the boundary ones: columns ascend), then the boundary start
    {
      const int ls = L.leaf_size;
      std::vector<int> pint(ls);
      for (int i = 0; i < ls; i++) {
        int q = L.lu_diag[i] + 1;
        while (q < L.lu_rowptr[i + 1] && L.lu_colidx[q] < ls) q++;
        pint[i] = q;
        ubnd.push_back(q);
      }
      for (int k = 0; k < ls; k++) {
        const int i = L.lev_rows[k];
        ftask.insert(ftask.end(), {i, L.lu_rowptr[i], L.lu_diag[i], L.lu_diag[i]});
        const int j = L.ulev_rows[k];
        btask.insert(btask.end(), {j, L.lu_diag[j] + 1, pint[j], L.lu_diag[j]});
      }
    }
    wcol.insert(wcol.end(), L.wcol.begin(), L.wcol.end());
    levptr.insert(levptr.end(), L.lev_ptr.begin(), L.lev_ptr.end());
    levrows.insert(levrows.end(), L.lev_rows.begin(), L.lev_rows.end());
    ulevptr.insert(ulevptr.end(), L.ulev_ptr.begin(), L.ulev_ptr.end());
    ulevrows.insert(ulevrows.end(), L.ulev_rows.begin(), L.ulev_rows.end());
    brow.insert(brow.end(), L.brow.begin(), L.brow.end());
    for (int c : L.bcol) {
      int k = 1;
      while (c >= L.cstart[k + 1]) k++;
      int b = L.path[k];
      bcol_blk.push_back(b);
      bcol_loc.push_back(c - L.cstart[k]);
      bcol_glob.push_back(P.block_start[b] + c - L.cstart[k]);
    }
    D.core_n = L.core_n;
    D.leaf_size = L.leaf_size;
    D.nbr = L.nbr;
    D.nbc = L.nbc;
    D.nlev = (int)L.lev_ptr.size() - 1;
    D.nulev = (int)L.ulev_ptr.size() - 1;
    D.gstart = P.block_start[L.leaf];
    D.cb_off = cb;
    D.vec = 0;
    cb += L.nbr;
    maxw = std::max(maxw, L.leaf_size + L.nbc);
  }
  h->max_leaf_w = maxw;
  h->wd.cb_total = cb;
  h->wd.max_leaf = 0;
  for (int r = 0; r < np; r++) h->wd.max_leaf = std::max(h->wd.max_leaf, P.leaves[r].leaf_size);
  h->wd.max_lev = 0;
  for (int r = 0; r < np; r++)
    h->wd.max_lev = std::max(h->wd.max_lev, (int)std::max(P.leaves[r].lev_ptr.size(), P.leaves[r].ulev_ptr.size()) - 1);
  // work lists: forward boundary workgroups and boundary targets (separator
  // rows receiving leaf contributions, each with its sources in ascending
  // rank order)
  WorkLists &W = h->wl;
  W = WorkLists();
  for (int r = 0; r < np; r++)
    for (int f = 0; f < P.leaves[r].nbr; f += 4) W.fwg_leaf.push_back(r), W.fwg_first.push_back(f);
  struct Src { int R, leaf, ri; };
  std::vector<Src> src;
  for (int r = 0; r < np; r++) {
    const LeafPlan &L = P.leaves[r];
    for (int ri = 0; ri < L.nbr; ri++) {
      int c = L.brow[ri], k = 1;
      while (c >= L.cstart[k + 1]) k++;
      src.push_back({P.block_start[L.path[k]] + c - L.cstart[k], r, ri});
    }
  }
  std::sort(src.begin(), src.end(), [](const Src &a, const Src &b) {
    return a.R != b.R ? a.R < b.R : a.leaf < b.leaf;
  });
  std::vector<int> blk_of(P.n);
  for (int b = 1; b <= P.num_block; b++)
    for (int i = P.block_start[b]; i < P.block_start[b] + P.block_size[b]; i++) blk_of[i] = b;
  W.tptr.assign(1, 0);
  for (size_t s = 0; s < src.size(); s++) {
    if (s == 0 || src[s].R != src[s - 1].R) {
      if (s) W.tptr.push_back((int)s);
      W.trow_glob.push_back(src[s].R);
      W.trow_blk.push_back(blk_of[src[s].R]);
      W.trow_loc.push_back(src[s].R - P.block_start[blk_of[src[s].R]]);
    }
    W.tsrc_leaf.push_back(src[s].leaf);
    W.tsrc_ri.push_back(src[s].ri);
  }
  if (!src.empty()) W.tptr.push_back((int)src.size());
  // assembly tiles: 64 consecutive rows of one separator block; per tile the
  // contributing leaves in ascending rank, each with the boundary row (ri)
  // feeding every row of the tile (-1: none)
  W.tile_blk.clear();
  W.tile_row0.clear();
  W.tile_ptr.assign(1, 0);
  W.seg_leaf.clear();
  W.seg_ri.clear();
  {
    size_t s0 = 0;
    while (s0 < src.size()) {
      const int b = blk_of[src[s0].R];
      const int row0 = ((src[s0].R - P.block_start[b]) / 64) * 64;
      const int R0 = P.block_start[b] + row0, R1 = std::min(R0 + 64, P.block_start[b] + P.block_size[b]);
      size_t s1 = s0;
      while (s1 < src.size() && src[s1].R < R1) s1++;
      std::vector<Src> part(src.begin() + s0, src.begin() + s1);
      std::stable_sort(part.begin(), part.end(), [](const Src &a, const Src &c) { return a.leaf < c.leaf; });
      for (size_t q = 0; q < part.size(); q++) {
        if (q == 0 || part[q].leaf != part[q - 1].leaf) {
          W.seg_leaf.push_back(part[q].leaf);
          W.seg_ri.insert(W.seg_ri.end(), 64, -1);
        }
        W.seg_ri[W.seg_ri.size() - 64 + (part[q].R - R0)] = part[q].ri;
      }
      W.tile_blk.push_back(b);
      W.tile_row0.push_back(row0);
      W.tile_ptr.push_back((int)W.seg_leaf.size());
      s0 = s1;
    }
  }
  // ---- dense tables ----
  std::vector<int> aoff((P.num_block + 1) * P.nk, 0);
  for (int p = 1; p <= P.num_block; p++)
    for (int d = 1; (p >> d) >= 1; d++) aoff[p * P.nk + d] = P.anc_off(p, p >> d);
  h->inv_off.assign(P.num_block + 1, 0);
  long long inv_total = 0;
  for (int p = 1; p < np; p++) {
    h->inv_off[p] = inv_total;
    inv_total += 2LL * P.block_size[p] * P.block_size[p];
  }
  h->inv_total = inv_total;
  h->tmp_total = inversion_tmp_size(P);
  h->arena_len = P.leaf_total + P.dense_total;
  // ---- structure uploads ----
  int *d_rowptr, *d_diag, *d_wcol, *d_levptr, *d_levrows, *d_ulevptr, *d_ulevrows;
  int *d_brow, *d_bcol_glob, *d_bcol_blk, *d_bcol_loc, *d_ftask, *d_btask, *d_ubnd;
  if ((rc = dalloc(h, &h->d_aval, (size_t)P.nnz))) return rc;
  if ((rc = dupload(h, &h->d_adest, std::vector<long long>(P.a_dest.begin(), P.a_dest.end())))) return rc;
  if ((rc = dupload(h, &d_rowptr, rowptr)) || (rc = dupload(h, &d_diag, diag)) ||
      (rc = dupload(h, &d_wcol, wcol)) || (rc = dupload(h, &d_levptr, levptr)) ||
      (rc = dupload(h, &d_levrows, levrows)) || (rc = dupload(h, &d_ulevptr, ulevptr)) ||
      (rc = dupload(h, &d_ulevrows, ulevrows)) || (rc = dupload(h, &d_brow, brow)) ||
      (rc = dupload(h, &d_bcol_glob, bcol_glob)) || (rc = dupload(h, &d_bcol_blk, bcol_blk)) ||
      (rc = dupload(h, &d_bcol_loc, bcol_loc)) || (rc = dupload(h, &h->d_order, P.order)) ||
      (rc = dupload(h, &d_ftask, ftask)) || (rc = dupload(h, &d_btask, btask)) ||
      (rc = dupload(h, &d_ubnd, ubnd)))
    return rc;
  WorkDev &wd = h->wd;
  wd.nleaves = np;
  wd.nfwg = (int)W.fwg_leaf.size();
  wd.ntarget = (int)W.trow_glob.size();
  wd.ntile = (int)W.tile_blk.size();
  if ((rc = dupload(h, &wd.fwg_leaf, W.fwg_leaf)) || (rc = dupload(h, &wd.fwg_first, W.fwg_first)) ||
      (rc = dupload(h, &wd.trow_blk, W.trow_blk)) || (rc = dupload(h, &wd.trow_loc, W.trow_loc)) ||
      (rc = dupload(h, &wd.trow_glob, W.trow_glob)) || (rc = dupload(h, &wd.tptr, W.tptr)) ||
      (rc = dupload(h, &wd.tsrc_leaf, W.tsrc_leaf)) || (rc = dupload(h, &wd.tsrc_ri, W.tsrc_ri)) ||
      (rc = dupload(h, &wd.tile_blk, W.tile_blk)) || (rc = dupload(h, &wd.tile_row0, W.tile_row0)) ||
      (rc = dupload(h, &wd.tile_ptr, W.tile_ptr)) || (rc = dupload(h, &wd.seg_leaf, W.seg_leaf)) ||
      (rc = dupload(h, &wd.seg_ri, W.seg_ri)))
    return rc;
  long long *d_offR, *d_offL, *d_mapoff;
  int *d_bsize, *d_nL, *d_lmap, *d_umap, *d_aoff;
  std::vector<long long> offR(P.offR.begin(), P.offR.end()), offL(P.offL.begin(), P.offL.end());
  std::vector<long long> mapoff(P.num_block + 1, 0);
  std::vector<int> lmap, umap;
  for (int p = 1; p < np; p++) {
    mapoff[p] = (long long)lmap.size();
    lmap.insert(lmap.end(), P.lmap[p].begin(), P.lmap[p].end());
    umap.insert(umap.end(), P.umap[p].begin(), P.umap[p].end());
  }
  if (lmap.empty()) lmap.push_back(-1), umap.push_back(-1);
  if ((rc = dupload(h, &d_offR, offR)) || (rc = dupload(h, &d_offL, offL)) ||
      (rc = dupload(h, &d_mapoff, mapoff)) || (rc = dupload(h, &d_bsize, P.block_size)) ||
      (rc = dupload(h, &d_nL, P.nL)) || (rc = dupload(h, &d_lmap, lmap)) || (rc = dupload(h, &d_umap, umap)) ||
      (rc = dupload(h, &d_aoff, aoff)))
    return rc;
  DenseDev dd;
  dd.offR = d_offR;
  dd.offL = d_offL;
  dd.mapoff = d_mapoff;
  dd.bsize = d_bsize;
  dd.nL = d_nL;
  dd.lmap = d_lmap;
  dd.umap = d_umap;
  dd.aoff = d_aoff;
  dd.nk = P.nk;
  if ((rc = dupload(h, &h->d_dd, std::vector<DenseDev>{dd}))) return rc;
  for (int r = 0; r < np; r++) {
    LeafDev &D = h->h_leaves[r];
    const Off &o = off[r];
    D.rowptr = d_rowptr + o.rp;
    D.diag = d_diag + o.rp;
    D.wcol = d_wcol + o.idx;
    D.levptr = d_levptr + o.lev;
    D.levrows = d_levrows + o.levr;
    D.ulevptr = d_ulevptr + o.ulev;
    D.ulevrows = d_ulevrows + o.ulevr;
    D.brow = d_brow + o.brow;
    D.bcol_glob = d_bcol_glob + o.bcol;
    D.bcol_blk = d_bcol_blk + o.bcol;
    D.bcol_loc = d_bcol_loc + o.bcol;
    D.ftask = d_ftask + o.task;
    D.btask = d_btask + o.task;
    D.ubnd = d_ubnd + o.row;
    D.dd = h->d_dd;
  }
  if ((rc = dalloc(h, &h->d_vec, 1)) || (rc = dalloc(h, &h->d_perm, 2)) ||
      (rc = dalloc(h, &h->d_prep, 1)) || (rc = dalloc(h, &h->d_ir, 1)) ||
      (rc = dalloc(h, &h->d_active, 1)) || (rc = dalloc(h, &h->d_norms, 1)))
    return rc;
  if (h->prec == SNU_FP32) {
    if ((rc = dupload(h, &h->d_arow, P.a_rowptr)) || (rc = dupload(h, &h->d_acol, P.a_colidx)))
      return rc;
    rc = alloc_values<float>(h, S_total);
    if (!rc) {
      for (int r = 0; r < np; r++) h->h_leaves[r].S = h->v32.S + off[r].S;
      rc = build_front_plan<float>(h);
    }
  } else {
    rc = alloc_values<double>(h, S_total);
    if (!rc) {
      for (int r = 0; r < np; r++) h->h_leaves[r].S = h->v64.S + off[r].S;
      rc = build_front_plan<double>(h);
    }
  }
  if (rc) return rc;
  if ((rc = dupload(h, &wd.leaves, h->h_leaves))) return rc;
  PrepDev pd;
  pd.arena = h->prec == SNU_FP32 ? (void *)h->v32.arena : (void *)h->v64.arena;
  pd.inv = h->prec == SNU_FP32 ? (void *)h->v32.inv : (void *)h->v64.inv;
  pd.arena_len = h->arena_len;
  pd.inv_len = h->inv_total;
  pd.aval = h->d_aval;
  pd.adest = h->d_adest;
  pd.nnz = P.nnz;
  HIPCHK(hipMemcpy(h->d_prep, &pd, sizeof(pd), hipMemcpyHostToDevice));
  return SNU_OK;
}

// Device bytes that upload() and a one-column solve take for this plan (the
// launch descriptor lists excluded).
size_t estimate_device_bytes(const snumat *h) {
  const Plan &P = h->P;
  const double vs = h->prec == SNU_FP32 ? 4.0 : 8.0;
  double S_total = 0, idx = 0, inv = 0, fronts = 0;
  const int fmax = h->prec == SNU_FP32 ? front_small_max<float>() : front_small_max<double>();
  for (const LeafPlan &L : P.leaves) {
    S_total += (double)L.nbr * L.nbc;
    idx += 2.0 * (L.core_n + 1) + L.lu_nnz + L.lev_ptr.size() + L.lev_rows.size() + L.ulev_ptr.size() +
           L.ulev_rows.size() + L.brow.size() + 3.0 * L.bcol.size() + 2.0 * L.nbr / 4 + 4;
    // fronts: values (larger ones with their diagonal-block inverses), slot /
    // position / extend-add tables, root row / column maps, descriptors
    fronts += (double)L.fentries;
    for (int f : L.ff) fronts += f > fmax ? 2.0 * 64 * 64 : 0.0;
    idx += 2.0 * L.fslot.size() + L.frel.size() + (double)(L.nbr + L.nbc) + 16.0 * L.ff.size();
  }
  S_total += fronts;
  for (int p = 1; p < P.np; p++) inv += 2.0 * P.block_size[p] * P.block_size[p];
  double bytes = vs * ((double)P.leaf_total + (double)P.dense_total + S_total + inv +
                       (double)inversion_tmp_size(P));
  bytes += 16.0 * P.nnz + 4.0 * (idx + P.n) + 48.0 * P.num_block;
  if (h->prec == SNU_FP32) bytes += 4.0 * (P.n + 1 + P.nnz);
  bytes += 2 * vs * P.n + 16.0 * P.n + (h->prec == SNU_FP32 ? 16.0 * P.n : 0.0);
  return (size_t)bytes;
}

void free_graph(FactorGraph &g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
}

void free_device(snumat *h) {
  free_graph(h->graph);
  free_gemv_plan(h->v64.sp);
  free_gemv_plan(h->v32.sp);
  free_dense_plan(h->v64.fact);
  free_dense_plan(h->v32.fact);
  free_dense_plan(h->v64.fronts.dense);
  free_dense_plan(h->v32.fronts.dense);
  for (void *p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  h->ev.clear();
  for (hipEvent_t e : h->pev) (void)hipEventDestroy(e);
  h->pev.clear();
  if (h->stream) (void)hipStreamDestroy(h->stream);
  h->stream = nullptr;
}

// ---------------------------------------------------------------------------
// executors
// ---------------------------------------------------------------------------
namespace {
// SNU_GEMM_SPLIT factorizations run their 64 x 64 GEMM tiles as bf16 splits
// too
constexpr bool kSplit64 = true;
template <typename T>
void run_launch(const Launch &L, const GemmProbT<T> *gp, const GemmSegT<T> *gs,
                const DiagProbT<T> *dp, const GemvProbT<T> *vp, const GemvSegT<T> *vs,
                T *part, hipStream_t st, bool split = false) {
  if constexpr (std::is_same<T, float>::value) {
    if (split && (L.kind == 5 || L.kind == 10 || L.kind == 11)) {
      launch_gemm128_split(gp + L.p0, L.np, gs, L.tiles, L.kind == 5 ? 0 : L.kind == 10 ? 1 : 2, st);
      return;
    }
    if (split && kSplit64 && (L.kind == 1 || L.kind == 2 || L.kind == 6)) {
      launch_gemm_split(gp + L.p0, L.np, gs, L.tiles, L.kind == 2 ? 0 : L.kind == 1 ? 1 : 2, st);
      return;
    }
  }
  switch (L.kind) {
    case 0: launch_diag_lu(dp + L.p0, L.np, st); break;
    case 1: launch_gemm(gp + L.p0, L.np, gs, L.tiles, 1, st); break;
    case 2: launch_gemm(gp + L.p0, L.np, gs, L.tiles, 0, st); break;
    case 3: launch_copy2d(gp + L.p0, L.np, gs, L.tiles, st); break;
    case 5: launch_gemm128(gp + L.p0, L.np, gs, L.tiles, 0, st); break;
    case 10: launch_gemm128(gp + L.p0, L.np, gs, L.tiles, 1, st); break;
    case 11: launch_gemm128(gp + L.p0, L.np, gs, L.tiles, 2, st); break;
    case 6: launch_gemm(gp + L.p0, L.np, gs, L.tiles, 2, st); break;
    case 7:
    case 8:
      if (L.wgs > 0) launch_gemv_split(vp + L.p0, L.np, vs, L.tiles, L.wgs, L.kind == 7 ? 1 : 0, part, st);
      else launch_gemv(vp + L.p0, L.np, vs, L.tiles, L.kind == 7 ? 1 : 0, st);
      break;
  }
}
}  // namespace

// ev (optional): 6 events marking prepare / leaf fronts / leaf Schur /
// assembly / dense; pev (optional, with profile): a begin/end pair per GEMM
// launch (the leaf fronts' launches, then the dense ones), their times summed
// into gemm_ms[0] (dense C -= AB), gemm_ms[1] (dense C = +-AB), gemm_ms[2]
// (the dense C -= AB launches on 128 x 128 tiles, part of gemm_ms[0]) and
// gemm_ms[3] / gemm_ms[4] (the fronts' C -= AB launches on 64 x 64 / 128 x
// 128 tiles).
// The launches of one factorization on st (no host synchronisation).
template <typename T>
int enqueue_factor(const FactorSet &fs, const DensePlan<T> &plan, hipStream_t st,
                   const std::vector<hipEvent_t> *ev, bool profile, std::vector<hipEvent_t> *pev) {
  const WorkDev &wd = *fs.wd;
  if (ev) HIPCHK(hipEventRecord((*ev)[0], st));
  launch_prepare<T>(fs.prep, fs.nmat, fs.max_len, fs.max_nnz, st);
  if (fs.ir) launch_rowabs_max(fs.ir, fs.nmat, fs.max_n, fs.anorm_out, st);
  if (ev) HIPCHK(hipEventRecord((*ev)[1], st));
  if (fs.scatter_only) {
    if (ev)
      for (int q = 2; q <= 5; q++) HIPCHK(hipEventRecord((*ev)[q], st));
    HIPCHK(hipGetLastError());
    return SNU_OK;
  }
  // leaf fronts, height by height: whole fronts in LDS; larger fronts
  // assembled, factored by the dense launches of their height, written back
  const FrontPlan<T> &fp = *(const FrontPlan<T> *)fs.fronts;
  int ngemm = 0;
  for (const Launch &L : fp.dense.launches) ngemm += is_gemm(L.kind);
  for (const Launch &L : plan.launches) ngemm += is_gemm(L.kind);
  const bool prof = profile && pev;
  if (prof && (int)pev->size() < 2 * ngemm) {
    for (hipEvent_t e : *pev) (void)hipEventDestroy(e);
    pev->assign(2 * ngemm, nullptr);
    for (auto &e : *pev) HIPCHK(hipEventCreate(&e));
  }
  int g = 0;
  auto run_timed = [&](const Launch &L, const DensePlan<T> &dp) -> int {
    const bool gm = prof && is_gemm(L.kind);
    if (gm) HIPCHK(hipEventRecord((*pev)[2 * g], st));
    run_launch<T>(L, dp.d_gp, dp.d_gs, dp.d_dp, nullptr, nullptr, nullptr, st, fs.split);
    if (gm) HIPCHK(hipEventRecord((*pev)[2 * g++ + 1], st));
    return SNU_OK;
  };
  size_t li = 0;
  const size_t nl = fp.dense.launches.size();
  for (int hh = 0; hh < (int)fp.levels.size(); hh++) {
    const FrontLevel &lv = fp.levels[hh];
    launch_front_small<T>(fp.d_small + lv.small0, lv.nsmall_lo, lv.max_f_lo, fs.eps, st);
    launch_front_small<T>(fp.d_small + lv.small0 + lv.nsmall_lo, lv.nsmall - lv.nsmall_lo, lv.max_f, fs.eps, st);
    launch_front_assemble<T>(fp.d_large, fp.d_work + lv.work0, fp.d_wrange, lv.nwork, st);
    for (; li < nl && front_key_height(fp.dense.launches[li].key) == hh; li++) {
      int rc = run_timed(fp.dense.launches[li], fp.dense);
      if (rc) return rc;
    }
    launch_front_writeback<T>(fp.d_large, fp.d_work + lv.work0, fp.d_wrange, lv.nwork, st);
  }
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord((*ev)[2], st));
  launch_front_schur<T>(fp.d_ls, fp.d_tiles, (int)fp.tiles.size(), st);
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord((*ev)[3], st));
  launch_assemble_schur<T>(wd.ntile, wd.tile_blk, wd.tile_row0, wd.tile_ptr, wd.seg_leaf, wd.seg_ri,
                           wd.leaves, st);
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord((*ev)[4], st));
  for (const Launch &L : plan.launches) {
    int rc = run_timed(L, plan);
    if (rc) return rc;
  }
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord((*ev)[5], st));
  return SNU_OK;
}

template <typename T>
int run_factor(const FactorSet &fs, const DensePlan<T> &plan, hipStream_t st,
               const std::vector<hipEvent_t> *ev, bool profile, std::vector<hipEvent_t> *pev,
               double *gemm_ms, FactorGraph *graph) {
  const WorkDev &wd = *fs.wd;
  int rc;
  if (graph && !profile) {
    // replay the captured launch sequence (captured on first use)
    if (!graph->exec) {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      rc = enqueue_factor<T>(fs, plan, st, nullptr, false, nullptr);
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return set_error(SNU_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&graph->exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        graph->exec = nullptr;
        return set_error(SNU_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
      }
    }
    if (ev) HIPCHK(hipEventRecord((*ev)[0], st));
    HIPCHK(hipGraphLaunch(graph->exec, st));
    if (ev) {
      for (int q = 1; q <= 5; q++) HIPCHK(hipEventRecord((*ev)[q], st));
    }
  } else if ((rc = enqueue_factor<T>(fs, plan, st, ev, profile, pev))) {
    return rc;
  }
  HIPCHK(hipStreamSynchronize(st));
  if (profile && pev && gemm_ms) {
    HIPCHK(hipStreamSynchronize(st));
    for (int q = 0; q < 5; q++) gemm_ms[q] = 0;
    int g = 0;
    const FrontPlan<T> &fp = *(const FrontPlan<T> *)fs.fronts;
    for (const Launch &L : fp.dense.launches) {
      if (!is_gemm(L.kind)) continue;
      const double ms = ev_ms((*pev)[2 * g], (*pev)[2 * g + 1]);
      if (L.kind == 2) gemm_ms[3] += ms;
      if (L.kind == 5) gemm_ms[4] += ms;
      g++;
    }
    for (const Launch &L : plan.launches) {
      if (!is_gemm(L.kind)) continue;
      const double ms = ev_ms((*pev)[2 * g], (*pev)[2 * g + 1]);
      gemm_ms[is_assign_gemm(L.kind) ? 1 : 0] += ms;
      if (L.kind == 5) gemm_ms[2] += ms;
      g++;
    }
  }
  return SNU_OK;
}

template <typename T>
void run_sweep(const SolveSet &ss, T *bp, T *cbuf, const GemvPlan<T> &plan, hipStream_t st) {
  const WorkDev &wd = *ss.wd;
  launch_permute_in(ss.perm, ss.nmat, ss.vec, bp, ss.max_elems, st);
  launch_leaf_solve(wd.leaves, wd.nleaves, ss.vec, bp, wd.max_leaf, wd.max_lev, ss.max_rhs, false, st);
  launch_leaf_fwd_sep(wd.leaves, wd.fwg_leaf, wd.fwg_first, wd.nfwg, 4, ss.vec, bp, cbuf, ss.cb_ld,
                      st);
  launch_assemble_rhs(wd.ntarget, wd.trow_glob, wd.tptr, wd.tsrc_leaf, wd.tsrc_ri, wd.leaves,
                      ss.vec, cbuf, ss.cb_ld, bp, ss.max_rhs, st);
  for (const Launch &L : plan.fwd)
    run_launch<T>(L, nullptr, nullptr, nullptr, plan.d_vp, plan.d_vs, plan.d_part, st);
  for (const Launch &L : plan.bwd)
    run_launch<T>(L, nullptr, nullptr, nullptr, plan.d_vp, plan.d_vs, plan.d_part, st);
  launch_leaf_solve(wd.leaves, wd.nleaves, ss.vec, bp, wd.max_leaf, wd.max_lev, ss.max_rhs, true, st);
  launch_permute_out(ss.perm, ss.nmat, ss.vec, bp, ss.max_elems, st);
}

// FP32 factors, FP64 refinement over a set of matrices: x = sweep(b), then
// x += sweep(b - A x) for the matrices whose backward error is above tol and
// still at least halving per step, at most max_steps corrections each.
int run_refined(const RefineSet &rs, float *bp, float *cbuf, const GemvPlan<float> &plan,
                hipStream_t st, int *steps_out, std::vector<double> *berr_mat) {
  const int nm = rs.nmat, ns = rs.nslots;
  std::vector<unsigned long long> u(3 * (size_t)ns);
  std::vector<double> bn(ns), rn(ns), xn(ns);
  auto fetch = [&](int area, std::vector<double> &out) -> int {
    HIPCHK(hipMemcpyAsync(u.data(), rs.norms + (size_t)area * ns, sizeof(unsigned long long) * ns,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < ns; i++) std::memcpy(&out[i], &u[i], sizeof(double));
    return SNU_OK;
  };
  int rc;
  launch_colmax(rs.ir, nm, rs.max_n, 0, rs.norms, ns, st);
  if ((rc = fetch(0, bn))) return rc;
  run_sweep<float>(rs.sweep_b, bp, cbuf, plan, st);
  std::vector<double> prev(nm, INFINITY), berr(nm, INFINITY);
  std::vector<int> active(nm, 0), steps(nm, 0);
  int max_steps_taken = 0;
  for (;;) {
    launch_residual(rs.ir, nm, rs.max_n, st);
    launch_colmax(rs.ir, nm, rs.max_n, 1, rs.norms + ns, ns, st);
    launch_colmax(rs.ir, nm, rs.max_n, 2, rs.norms + 2 * (size_t)ns, ns, st);
    HIPCHK(hipGetLastError());
    if ((rc = fetch(1, rn)) || (rc = fetch(2, xn))) return rc;
    std::fill(berr.begin(), berr.end(), 0.0);
    for (int i = 0; i < ns; i++) {
      const int m = (*rs.slot_mat)[i];
      const double den = (*rs.anorm)[m] * xn[i] + bn[i];
      double e = den > 0 ? rn[i] / den : (rn[i] > 0 ? INFINITY : 0.0);
      if (!(e == e)) e = INFINITY;
      berr[m] = std::max(berr[m], e);
    }
    int nact = 0;
    for (int m = 0; m < nm; m++) {
      active[m] = berr[m] > rs.tol && steps[m] < rs.max_steps && berr[m] < 0.5 * prev[m];
      nact += active[m];
    }
    if (!nact) break;
    for (int m = 0; m < nm; m++)
      if (active[m]) prev[m] = berr[m], steps[m]++, max_steps_taken = std::max(max_steps_taken, steps[m]);
    HIPCHK(hipMemcpyAsync(rs.d_active, active.data(), sizeof(int) * nm, hipMemcpyHostToDevice, st));
    run_sweep<float>(rs.sweep_r, bp, cbuf, plan, st);
    launch_update(rs.ir, nm, rs.max_elems, rs.d_active, st);
    HIPCHK(hipGetLastError());
  }
  if (steps_out) *steps_out = max_steps_taken;
  if (berr_mat) *berr_mat = berr;
  double worst = 0;
  for (double e : berr) worst = std::max(worst, e);
  if (!(worst <= 1e-12))
    return set_error(SNU_ERR_NOT_CONVERGED, "FP32 refinement stalled at backward error " +
                                                std::to_string(worst) + " after " +
                                                std::to_string(max_steps_taken) + " steps");
  return SNU_OK;
}

#define SNU_ENGINE_INST(T)                                                                    \
  template void build_solve_plan<T>(const snumat *, T *, T *, int, GemvPlan<T> &);            \
  template int upload_gemv_plan<T>(GemvPlan<T> &, hipStream_t);                               \
  template size_t layout_gemv_plan<T>(GemvPlan<T> &);                                         \
  template void free_gemv_plan<T>(GemvPlan<T> &);                                             \
  template int upload_dense_plan<T>(DensePlan<T> &);                                          \
  template void free_dense_plan<T>(DensePlan<T> &);                                           \
  template void merge_dense<T>(const std::vector<const DensePlan<T> *> &, DensePlan<T> &);    \
  template void merge_gemv<T>(const std::vector<const GemvPlan<T> *> &, GemvPlan<T> &);       \
  template int run_factor<T>(const FactorSet &, const DensePlan<T> &, hipStream_t,            \
                             const std::vector<hipEvent_t> *, bool, std::vector<hipEvent_t> *, \
                             double *, FactorGraph *);                                        \
  template void run_sweep<T>(const SolveSet &, T *, T *, const GemvPlan<T> &, hipStream_t);
SNU_ENGINE_INST(double)
SNU_ENGINE_INST(float)
#undef SNU_ENGINE_INST

}  // namespace snu
