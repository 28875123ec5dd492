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
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

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

// waves per leaf workgroup: the LDS of one workgroup holds the leaf's row
// completion flags and one workspace row (wsz values) per wave
// waves of a leaf workgroup: as many workspaces as the LDS holds, at most
// cap (leaf rows: 16, FP32 workspaces being half the size; separator rows:
// 16 in FP32, 8 in FP64, whose kernel needs more than 128 registers per lane)
int waves_for(int wsz, int vsize, int max_leaf, int cap) {
  int fit = (device_lds_limit() - leaf_flag_bytes(max_leaf)) / std::max(1, wsz * vsize);
  return std::min(cap, fit);
}

int leaf_workspace(const Plan &P) {
  int w = 1;
  for (const LeafPlan &L : P.leaves) w = std::max(w, L.leaf_size + L.nbc);
  return (w + 1) & ~1;
}
bool leaf_fits_lds(const Plan &P, int prec) {
  int ml = 0;
  for (const LeafPlan &L : P.leaves) ml = std::max(ml, L.leaf_size);
  const long long vs = prec == SNU_FP32 ? 4 : 8;
  return leaf_flag_bytes(ml) + (long long)leaf_workspace(P) * vs <= device_lds_limit();
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
template <typename T>
void tail_gemm_list(const std::vector<LeafDev> &leaves, std::vector<GemmProbT<T>> &gp,
                    std::vector<GemmSegT<T>> &gs, int &tiles) {
  double flops = 0;
  for (const LeafDev &L : leaves)
    if (L.nt > 0 && L.nbr > 0 && L.nbc > 0)
      add_gemm(gp, gs, tiles, (T *)L.S, L.nbc, L.nbc, L.nbr,
               {GemmSegT<T>{(const T *)L.tail_u, (const T *)L.tail_l, L.nbc, L.nt, L.nt}}, flops, 64);
}
template void tail_gemm_list<double>(const std::vector<LeafDev> &, std::vector<GemmProbT<double>> &,
                                     std::vector<GemmSegT<double>> &, int &);
template void tail_gemm_list<float>(const std::vector<LeafDev> &, std::vector<GemmProbT<float>> &,
                                    std::vector<GemmSegT<float>> &, int &);
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
// holds enough of those tiles to fill the device twice over (fewer would
// leave CUs idle that 64 x 64 tiles keep busy); the rest take 64 x 64 tiles
// (launch kind 5 and kind 2).
constexpr int kMin128Tiles = 512;
template <typename T>
struct GemmBatch {
  struct Pend {
    T *C;
    long long ldc;
    int M, N;
    std::vector<GemmSegT<T>> segs;
  };
  std::vector<Pend> pend;
  void add(T *C, long long ldc, int M, int N, const std::vector<GemmSegT<T>> &segs) {
    if (M > 0 && N > 0) pend.push_back(Pend{C, ldc, M, N, segs});
  }
  void flush(std::vector<GemmProbT<T>> &gp, std::vector<GemmSegT<T>> &gs, std::vector<Launch> &out,
             long long key) {
    const int prec = std::is_same<T, float>::value ? SNU_FP32 : SNU_FP64;
    long long big = 0;
    for (const Pend &q : pend)
      if (tile_for(prec, q.M, q.N) == 128) big += (long long)((q.M + 127) / 128) * ((q.N + 127) / 128);
    const bool use128 = big >= kMin128Tiles;
    std::vector<GemmProbT<T>> p[2];
    int tiles[2] = {0, 0};
    double flops[2] = {0, 0};
    for (const Pend &q : pend) {
      const int t = use128 && tile_for(prec, q.M, q.N) == 128 ? 1 : 0;
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
  const int prec = std::is_same<T, float>::value ? SNU_FP32 : SNU_FP64;
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
        DiagProb dp;
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
      // trailing update of D
      GemmBatch<T> tu;
      for (int i = lo; i <= hi; i++) {
        int s = P.block_size[i];
        if (s <= k) continue;
        int nb = std::min(64, s - k);
        T *Ccol = R(i) + k + (long long)(k + nb) * s;  // U12 rows
        const T *L21 = R(i) + (k + nb) + (long long)k * s;
        tu.add(R(i) + (k + nb) + (long long)(k + nb) * s, s, s - k - nb, s - k - nb,
               {GemmSeg{L21, Ccol, s, s, nb}});
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
          if (tile_for(prec, hh, hb) == 128) big += 2LL * ((hh + 127) / 128) * ((hb + 127) / 128);
        }
      }
      const int W = big >= kMin128Tiles ? 128 : 64;
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
    // L^-1 Up per W-row strip (rows [0, m0 + W) of Up), over the
    // structurally nonzero row / column runs, into scratch; then copied back.
    // W = 128 (128 x 128 tiles) when the autotuner prefers that tile for the
    // level's panels and they hold enough such tiles to fill the device
    {
      long long big = 0;
      for (int i = lo; i <= hi; i++) {
        const int s = P.block_size[i], a = P.anc_sum[i];
        for (auto rr : nz_runs(a, [&](int x0, int x1) { return P.anyL(i, x0, x1); }))
          if (tile_for(prec, rr.second - rr.first, s) == 128)
            big += (long long)((rr.second - rr.first + 127) / 128) * ((s + 127) / 128);
        for (auto cr : nz_runs(a, [&](int x0, int x1) { return P.anyU(i, x0, x1); }))
          if (tile_for(prec, s, cr.second - cr.first) == 128)
            big += (long long)((s + 127) / 128) * ((cr.second - cr.first + 127) / 128);
      }
      const int W = big >= kMin128Tiles ? 128 : 64;
      Launch pm{W == 128 ? 10 : 1, (int)h_gp.size(), 0, 0, 0, fkey(height, 60000000LL)};
      std::vector<GemmProbT<T>> cp;
      Launch cl{3, 0, 0, 0, 0, fkey(height, 60000001LL)};
      double cfl = 0;
      long long toff = 0;
      for (int i = lo; i <= hi; i++) {
        const int s = P.block_size[i], a = P.anc_sum[i];
        for (auto rr : nz_runs(a, [&](int x0, int x1) { return P.anyL(i, x0, x1); })) {
          const int m = rr.second - rr.first;
          T *dst = VV.tmp + toff;
          toff += (long long)m * s;
          T *src = Lp(i) + rr.first;
          for (int n0 = 0; n0 < s; n0 += W) {
            const int nt = std::min(W, s - n0), kk = n0 + nt;
            pm.np += add_gemm(h_gp, h_gs, pm.tiles, dst + (long long)n0 * m, m, m, nt,
                              {GemmSeg{src, Ui(i) + (long long)n0 * s, a, s, kk}}, pm.flops, W);
          }
          cl.np += add_gemm(cp, h_gs, cl.tiles, src, a, m, s, {GemmSeg{dst, dst, m, m, 1}}, cfl);
        }
        for (auto cr : nz_runs(a, [&](int x0, int x1) { return P.anyU(i, x0, x1); })) {
          const int nc = cr.second - cr.first;
          T *dst = VV.tmp + toff;
          toff += (long long)s * nc;
          T *src = R(i) + (long long)(s + cr.first) * s;
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
    // separators below it (one grouped launch per level).  Emitted per 64 x 64
    // target tile, with only the contributors whose L-panel rows and U-panel
    // columns under that tile hold structural nonzeros (skipped products are
    // exactly zero).
    GemmBatch<T> sc;
    for (int p = 1; p < lo; p++) {
      int lp = 31 - __builtin_clz(p);
      int first = p << (l - lp), last = ((p + 1) << (l - lp)) - 1;
      int sp = P.block_size[p], ap = P.anc_sum[p];
      for (int pass = 0; pass < 2; pass++) {  // 0: R_p = [D_p | Up_p], 1: Lp_p
        const int M = pass == 0 ? sp : ap, N = pass == 0 ? sp + ap : sp;
        const int tm = (M + 63) / 64, tn = (N + 63) / 64;
        // contributor set of every 64 x 64 target tile
        std::vector<std::vector<int>> who((size_t)tm * tn);
        for (int a = 0; a < tm; a++)
          for (int b = 0; b < tn; b++) {
            const int m0 = 64 * a, n0 = 64 * b, mt = std::min(64, M - m0), nt = std::min(64, N - n0);
            for (int i = first; i <= last; i++) {
              const int off = P.anc_off(i, p);
              const int r0 = off + (pass == 0 ? 0 : sp) + m0, c0 = off + n0;
              if (P.anyL(i, r0, r0 + mt) && P.anyU(i, c0, c0 + nt)) who[(size_t)a * tn + b].push_back(i);
            }
          }
        // one problem per 64 x 64 target tile with its contributor set, or
        // per 128 x 128 block when its four tiles share one set (the tile
        // kernel is then chosen per size bucket), emitted in strips of 8
        // tile rows, two tile columns at a time: consecutive tiles (which
        // one XCD runs together) share A rows (same tile row and
        // contributor) and B columns (same tile column), so their operands
        // meet in that XCD's L2
        auto segs_of = [&](const std::vector<int> &w, int m0, int n0) {
          std::vector<GemmSeg> segs;
          for (int i : w) {
            const int si = P.block_size[i], ai = P.anc_sum[i], off = P.anc_off(i, p);
            const int r0 = off + (pass == 0 ? 0 : sp) + m0, c0 = off + n0;
            segs.push_back(GemmSeg{Lp(i) + r0, R(i) + (long long)(si + c0) * si, ai, si, si});
          }
          return segs;
        };
        auto emit = [&](int a, int b, int ta, int tb) {
          const std::vector<int> &w = who[(size_t)a * tn + b];
          const int m0 = 64 * a, n0 = 64 * b;
          const int mt = std::min(64 * ta, M - m0), nt = std::min(64 * tb, N - n0);
          T *C = pass == 0 ? R(p) + m0 + (long long)n0 * sp : Lp(p) + m0 + (long long)n0 * ap;
          sc.add(C, pass == 0 ? sp : ap, mt, nt, segs_of(w, m0, n0));
        };
        auto nonempty = [&](int a, int b) { return a < tm && b < tn && !who[(size_t)a * tn + b].empty(); };
        constexpr int kStrip = 8;
        for (int a0 = 0; a0 < tm; a0 += kStrip)
          for (int b = 0; b < tn; b += 2)
            for (int a = a0; a < std::min(tm, a0 + kStrip); a += 2) {
              const std::vector<int> &w = who[(size_t)a * tn + b];
              if (nonempty(a, b) && nonempty(a + 1, b) && nonempty(a, b + 1) && nonempty(a + 1, b + 1) &&
                  who[(size_t)(a + 1) * tn + b] == w && who[(size_t)a * tn + b + 1] == w &&
                  who[(size_t)(a + 1) * tn + b + 1] == w) {
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
    for (int i = lo; i <= hi; i++) {
      const int s = P.block_size[i], a = P.anc_sum[i];
      for (auto rr : nz_runs(a, [&](int x0, int x1) { return P.anyL(i, x0, x1); }))
        panels += (long long)(rr.second - rr.first) * s;
      for (auto cr : nz_runs(a, [&](int x0, int x1) { return P.anyU(i, x0, x1); }))
        panels += (long long)(cr.second - cr.first) * s;
    }
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
          int r0 = P.anc_off(i, p) + m0;
          if (!P.anyL(i, r0, r0 + mt)) continue;
          segs.push_back(GemvSeg{Lp(i) + r0, yseg(i), P.anc_sum[i], n, P.block_size[i]});
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
        for (auto cr : nz_runs(P.block_size[a], [&](int x0, int x1) { return P.anyU(i, off + x0, off + x1); }))
          segs.push_back(GemvSeg{R(i) + (long long)(s + off + cr.first) * s, bseg(a) + cr.first, s, n,
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
// The leaf-tail GEMM problems of one matrix (leaf descriptors complete).
template <typename T>
int upload_tail_gemm(snumat *h) {
  std::vector<GemmProbT<T>> gp;
  std::vector<GemmSegT<T>> gs;
  int tiles = 0, rc;
  tail_gemm_list<T>(h->h_leaves, gp, gs, tiles);
  WorkDev &wd = h->wd;
  GemmProbT<T> *dgp = nullptr;
  GemmSegT<T> *dgs = nullptr;
  if (!gp.empty() && ((rc = dupload(h, &dgp, gp)) || (rc = dupload(h, &dgs, gs)))) return rc;
  wd.tail_gp = dgp;
  wd.tail_gs = dgs;
  wd.tail_np = (int)gp.size();
  wd.tail_tiles = tiles;
  return SNU_OK;
}
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

int upload(snumat *h) {
  Plan &P = h->P;
  int rc;
  HIPCHK(hipSetDevice(h->device));
  if (h->prec == SNU_FP32) autotune_tiles<float>();
  else autotune_tiles<double>();
  const int np = P.np;
  // ---- leaf arrays (concatenated; descriptors hold pointers into them) ----
  std::vector<int> rowptr, diag, levptr, levrows, ulevptr, ulevrows;
  std::vector<int> brow, bcol_glob, bcol_blk, bcol_loc, wcol, ftask, btask, ubnd, sep_tail;
  const int vsz = h->prec == SNU_FP32 ? (int)sizeof(float) : (int)sizeof(double);
  int max_nt = 0;
  struct Off { size_t rp, idx, lev, levr, ulev, ulevr, brow, bcol, task, row; long long S, tu, tl, tp; };
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
    // dense tail: the rows after the leaf's split (its own top separator),
    // when they fit the LDS of the tail kernels
    int nt = L.split_b > 0 && L.nbr > 0 ? L.leaf_size - L.split_b : 0;
    if (nt > kMaxLeafTail || (long long)128 * nt * vsz > device_lds_limit() ||
        (long long)nt * (nt + 1) / 2 * vsz > device_lds_limit())
      nt = 0;
    D.t0 = L.leaf_size - nt;
    D.nt = nt;
    max_nt = std::max(max_nt, nt);
    o.tu = S_total;
    S_total += (long long)nt * L.nbc;
    o.tl = S_total;
    S_total += (long long)L.nbr * nt;
    o.tp = S_total;
    S_total += (long long)nt * (nt + 1) / 2;
    for (int ri = 0; ri < L.nbr; ri++) {
      const int i = L.brow[ri];
      int q = L.lu_rowptr[i];
      while (q < L.lu_rowptr[i + 1] && L.lu_colidx[q] < D.t0) q++;
      sep_tail.push_back(q);
    }
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
    D.split_a = L.split_a;
    D.split_b = L.split_b;
    D.gstart = P.block_start[L.leaf];
    D.cb_off = cb;
    D.vec = 0;
    cb += L.nbr;
    maxw = std::max(maxw, L.leaf_size + L.nbc);
  }
  h->max_leaf_w = maxw;
  h->wd.cb_total = cb;
  h->wd.wsz = (maxw + 1) & ~1;
  h->wd.max_leaf = 0;
  for (int r = 0; r < np; r++) h->wd.max_leaf = std::max(h->wd.max_leaf, P.leaves[r].leaf_size);
  h->wd.max_lev = 0;
  for (int r = 0; r < np; r++)
    h->wd.max_lev = std::max(h->wd.max_lev, (int)std::max(P.leaves[r].lev_ptr.size(), P.leaves[r].ulev_ptr.size()) - 1);
  const int vsize = h->prec == SNU_FP32 ? (int)sizeof(float) : (int)sizeof(double);
  h->wd.waves = waves_for(h->wd.wsz, vsize, h->wd.max_leaf, 16);
  {
    // more leaves than CUs: narrower workgroups that keep every leaf
    // resident at once (k per CU) when they still have 6 or more waves
    // (measured on 512 leaves: 7 resident waves beat 15 waves in two
    // rounds, 4 resident waves lose to 7 in two rounds)
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0 && np > cus) {
      const int k = (np + cus - 1) / cus;
      const int per = ((device_lds_limit() - leaf_flag_bytes(h->wd.max_leaf)) / k -
                       leaf_flag_bytes(h->wd.max_leaf)) / std::max(1, h->wd.wsz * vsize);
      if (per >= 6 && per < h->wd.waves) h->wd.waves = per;
    }
  }
  h->wd.sep_wsz = h->wd.wsz;
  h->wd.sep_waves = waves_for(h->wd.sep_wsz, vsize, 0, vsize == 4 ? 16 : 8);
  if (h->wd.waves < 1)
    return set_error(SNU_ERR_OOM, "leaf workspace of " + std::to_string(maxw) +
                                      " values (the widest leaf's rows plus its boundary columns) exceeds the " +
                                      std::to_string(device_lds_limit()) +
                                      " bytes of LDS of one workgroup: use more leaves (np)");
  // work lists: leaf separator-row workgroups, forward boundary workgroups,
  // and boundary targets (separator rows receiving leaf contributions, each
  // with its sources in ascending rank order)
  WorkLists &W = h->wl;
  W = WorkLists();
  for (int r = 0; r < np; r++)
    for (int f = 0; f < P.leaves[r].nbr; f += h->wd.sep_waves) W.wg_leaf.push_back(r), W.wg_first.push_back(f);
  h->wd.max_nt = max_nt;
  for (int r = 0; r < np; r++)
    if (h->h_leaves[r].nt > 0)
      for (int f = 0; f < P.leaves[r].nbr; f += kTailRows) W.twg_leaf.push_back(r), W.twg_first.push_back(f);
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
  int *d_brow, *d_bcol_glob, *d_bcol_blk, *d_bcol_loc, *d_ftask, *d_btask, *d_ubnd, *d_sep_tail;
  if ((rc = dalloc(h, &h->d_aval, (size_t)P.nnz))) return rc;
  if ((rc = dupload(h, &h->d_adest, std::vector<long long>(P.a_dest.begin(), P.a_dest.end())))) return rc;
  if ((rc = dupload(h, &d_rowptr, rowptr)) || (rc = dupload(h, &d_diag, diag)) ||
      (rc = dupload(h, &d_wcol, wcol)) || (rc = dupload(h, &d_levptr, levptr)) ||
      (rc = dupload(h, &d_levrows, levrows)) || (rc = dupload(h, &d_ulevptr, ulevptr)) ||
      (rc = dupload(h, &d_ulevrows, ulevrows)) || (rc = dupload(h, &d_brow, brow)) ||
      (rc = dupload(h, &d_bcol_glob, bcol_glob)) || (rc = dupload(h, &d_bcol_blk, bcol_blk)) ||
      (rc = dupload(h, &d_bcol_loc, bcol_loc)) || (rc = dupload(h, &h->d_order, P.order)) ||
      (rc = dupload(h, &d_ftask, ftask)) || (rc = dupload(h, &d_btask, btask)) ||
      (rc = dupload(h, &d_ubnd, ubnd)) || (rc = dupload(h, &d_sep_tail, sep_tail)))
    return rc;
  WorkDev &wd = h->wd;
  wd.nleaves = np;
  wd.nwg = (int)W.wg_leaf.size();
  wd.nfwg = (int)W.fwg_leaf.size();
  wd.ntwg = (int)W.twg_leaf.size();
  wd.ntarget = (int)W.trow_glob.size();
  wd.ntile = (int)W.tile_blk.size();
  if ((rc = dupload(h, &wd.wg_leaf, W.wg_leaf)) || (rc = dupload(h, &wd.wg_first, W.wg_first)) ||
      (rc = dupload(h, &wd.twg_leaf, W.twg_leaf)) || (rc = dupload(h, &wd.twg_first, W.twg_first)) ||
      (rc = dupload(h, &wd.fwg_leaf, W.fwg_leaf)) || (rc = dupload(h, &wd.fwg_first, W.fwg_first)) ||
      (rc = dupload(h, &wd.trow_blk, W.trow_blk)) || (rc = dupload(h, &wd.trow_loc, W.trow_loc)) ||
      (rc = dupload(h, &wd.trow_glob, W.trow_glob)) || (rc = dupload(h, &wd.tptr, W.tptr)) ||
      (rc = dupload(h, &wd.tsrc_leaf, W.tsrc_leaf)) || (rc = dupload(h, &wd.tsrc_ri, W.tsrc_ri)) ||
      (rc = dupload(h, &wd.tile_blk, W.tile_blk)) || (rc = dupload(h, &wd.tile_row0, W.tile_row0)) ||
      (rc = dupload(h, &wd.tile_ptr, W.tile_ptr)) || (rc = dupload(h, &wd.seg_leaf, W.seg_leaf)) ||
      (rc = dupload(h, &wd.seg_ri, W.seg_ri)))
    return rc;
  long long *d_offR, *d_offL;
  int *d_bsize, *d_anc_sum, *d_aoff;
  std::vector<long long> offR(P.offR.begin(), P.offR.end()), offL(P.offL.begin(), P.offL.end());
  if ((rc = dupload(h, &d_offR, offR)) || (rc = dupload(h, &d_offL, offL)) ||
      (rc = dupload(h, &d_bsize, P.block_size)) || (rc = dupload(h, &d_anc_sum, P.anc_sum)) ||
      (rc = dupload(h, &d_aoff, aoff)))
    return rc;
  DenseDev dd;
  dd.offR = d_offR;
  dd.offL = d_offL;
  dd.bsize = d_bsize;
  dd.anc_sum = d_anc_sum;
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
    D.sep_tail = d_sep_tail + o.brow;
    D.dd = h->d_dd;
  }
  if ((rc = dalloc(h, &wd.stall, 1)) || (rc = dalloc(h, &wd.half_done, np)) ||
      (rc = dalloc(h, &h->d_vec, 1)) || (rc = dalloc(h, &h->d_perm, 2)) ||
      (rc = dalloc(h, &h->d_prep, 1)) || (rc = dalloc(h, &h->d_ir, 1)) ||
      (rc = dalloc(h, &h->d_active, 1)) || (rc = dalloc(h, &h->d_norms, 1)))
    return rc;
  if (h->prec == SNU_FP32) {
    if ((rc = dupload(h, &h->d_arow, P.a_rowptr)) || (rc = dupload(h, &h->d_acol, P.a_colidx)))
      return rc;
    rc = alloc_values<float>(h, S_total);
    if (!rc) {
      for (int r = 0; r < np; r++) {
        h->h_leaves[r].S = h->v32.S + off[r].S;
        h->h_leaves[r].tail_u = h->v32.S + off[r].tu;
        h->h_leaves[r].tail_l = h->v32.S + off[r].tl;
        h->h_leaves[r].tail_p = h->v32.S + off[r].tp;
      }
      rc = upload_tail_gemm<float>(h);
    }
  } else {
    rc = alloc_values<double>(h, S_total);
    if (!rc) {
      for (int r = 0; r < np; r++) {
        h->h_leaves[r].S = h->v64.S + off[r].S;
        h->h_leaves[r].tail_u = h->v64.S + off[r].tu;
        h->h_leaves[r].tail_l = h->v64.S + off[r].tl;
        h->h_leaves[r].tail_p = h->v64.S + off[r].tp;
      }
      rc = upload_tail_gemm<double>(h);
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
  double S_total = 0, idx = 0, inv = 0;
  for (const LeafPlan &L : P.leaves) {
    S_total += (double)L.nbr * L.nbc;
    idx += 2.0 * (L.core_n + 1) + L.lu_nnz + L.lev_ptr.size() + L.lev_rows.size() + L.ulev_ptr.size() +
           L.ulev_rows.size() + L.brow.size() + 3.0 * L.bcol.size() + 2.0 * L.nbr / 4 + 4;
  }
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
template <typename T>
void run_launch(const Launch &L, const GemmProbT<T> *gp, const GemmSegT<T> *gs,
                const DiagProbT<T> *dp, const GemvProbT<T> *vp, const GemvSegT<T> *vs,
                T *part, hipStream_t st) {
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

// ev (optional): 6 events marking prepare / leaf rows / leaf sep / assembly /
// dense; pev (optional, with profile): a begin/end pair per GEMM launch,
// their times summed into gemm_ms[0] (C -= AB), gemm_ms[1] (C = +-AB) and
// gemm_ms[2] (the C -= AB launches on 128 x 128 tiles, part of gemm_ms[0]).
// The launches of one factorization on st (no host synchronisation).
template <typename T>
int enqueue_factor(const FactorSet &fs, const DensePlan<T> &plan, hipStream_t st,
                   const std::vector<hipEvent_t> *ev, bool profile, std::vector<hipEvent_t> *pev) {
  const WorkDev &wd = *fs.wd;
  if (ev) HIPCHK(hipEventRecord((*ev)[0], st));
  launch_prepare<T>(fs.prep, fs.nmat, fs.max_len, fs.max_nnz, st);
  if (fs.ir) launch_rowabs_max(fs.ir, fs.nmat, fs.max_n, fs.anorm_out, st);
  if (ev) HIPCHK(hipEventRecord((*ev)[1], st));
  launch_zero_i32(wd.stall, 1, st);
  launch_leaf_rows<T>(wd.leaves, wd.nleaves, wd.wsz, wd.waves, wd.max_leaf, fs.eps, wd.stall,
                      wd.half_done, st);
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord((*ev)[2], st));
  launch_leaf_sep<T>(wd.leaves, wd.wg_leaf, wd.wg_first, wd.nwg, wd.sep_wsz, wd.sep_waves, st);
  launch_leaf_tail<T>(wd.leaves, wd.nleaves, wd.twg_leaf, wd.twg_first, wd.ntwg, kTailRows,
                      (const GemmProbT<T> *)wd.tail_gp, wd.tail_np, (const GemmSegT<T> *)wd.tail_gs,
                      wd.tail_tiles, wd.max_nt, st);
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord((*ev)[3], st));
  launch_assemble_schur<T>(wd.ntile, wd.tile_blk, wd.tile_row0, wd.tile_ptr, wd.seg_leaf, wd.seg_ri,
                           wd.leaves, st);
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord((*ev)[4], st));
  int ngemm = 0;
  for (const Launch &L : plan.launches) ngemm += is_gemm(L.kind);
  if (profile && pev && (int)pev->size() < 2 * ngemm) {
    for (hipEvent_t e : *pev) (void)hipEventDestroy(e);
    pev->assign(2 * ngemm, nullptr);
    for (auto &e : *pev) HIPCHK(hipEventCreate(&e));
  }
  int g = 0;
  for (const Launch &L : plan.launches) {
    const bool gm = is_gemm(L.kind);
    if (profile && pev && gm) HIPCHK(hipEventRecord((*pev)[2 * g], st));
    run_launch<T>(L, plan.d_gp, plan.d_gs, plan.d_dp, nullptr, nullptr, nullptr, st);
    if (profile && pev && gm) {
      HIPCHK(hipEventRecord((*pev)[2 * g + 1], st));
      g++;
    }
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
  int stalled = 0;
  HIPCHK(hipMemcpyAsync(&stalled, wd.stall, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (stalled) return set_error(SNU_ERR_HIP, "leaf factorization: a row's wait for its pivot rows timed out");
  if (profile && pev && gemm_ms) {
    HIPCHK(hipStreamSynchronize(st));
    gemm_ms[0] = gemm_ms[1] = gemm_ms[2] = 0;
    int g = 0;
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
