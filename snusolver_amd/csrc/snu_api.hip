// C-ABI entry points and the device-side orchestration of construct ->
// factorize -> solve on one MI355X.
//
// Phase map (reference -> here):
//   SnuMat::SnuMat (SnuMat.cu:67-112)           -> snumat_construct
//   core_run (kernel_sparse.cu:1011-1037)        -> scatter + k_leaf_rows +
//                                                   k_leaf_sep + k_assemble_schur
//   dense level loop (SnuMat_factorize.cu:258-284) -> per level: blocked
//        getrf/trsm (k_diag_lu + k_gemm ASSIGN/SUB), then one grouped k_gemm
//        for all Schur updates of the level (K-concatenated over siblings:
//        replaces reduction_gpu, SnuMat_factorize.cu:78-124)
//   forward solve fused in the reference (core_trsm/core_update_b,
//        trsm_Lxb/gemm on b)                    -> snumat_solve, forward half
//   backward solve (SnuMat_factorize.cu:286-345) -> snumat_solve, backward half
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <type_traits>
#include <string>
#include <vector>

#include "snu_internal.h"
#include "snu_kernels.h"

namespace snu {
static thread_local std::string g_err;
int set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace snu

using namespace snu;

#define HIPCHK(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess)                                                  \
      return set_error(SNU_ERR_HIP, std::string("HIP error ") +            \
                                        hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace {

struct Launch {
  int kind;  // 0 diag LU, 1 gemm assign, 2 gemm sub, 3 tri fwd, 4 tri bwd
  int p0, np, tiles;
  double flops;
};

struct SolvePlan {
  int nrhs = -1;
  std::vector<GemmProb> gp;
  std::vector<GemmSeg> gs;
  std::vector<TriProb> tp;
  std::vector<Launch> fwd, bwd;
  GemmProb *d_gp = nullptr;
  GemmSeg *d_gs = nullptr;
  TriProb *d_tp = nullptr;
};

}  // namespace

struct snumat {
  Plan P;
  snu_opts_t opts;
  bool on_device = false;
  bool factored = false;
  int device = -1;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  size_t dev_bytes = 0;
  // values
  double *d_arena = nullptr;  // leaf LU values, then dense blocks
  long long arena_len = 0;
  double *d_aval = nullptr;
  long long *d_adest = nullptr;
  // leaves
  std::vector<LeafDev> h_leaves;
  LeafDev *d_leaves = nullptr;
  int *d_rowptr = nullptr, *d_diag = nullptr, *d_wcol = nullptr;
  int *d_levptr = nullptr, *d_levrows = nullptr, *d_ulevptr = nullptr, *d_ulevrows = nullptr;
  int *d_brow = nullptr, *d_bcol_glob = nullptr, *d_bcol_blk = nullptr, *d_bcol_loc = nullptr;
  double *d_S = nullptr;
  int *d_wg_leaf = nullptr, *d_wg_first = nullptr;
  int nwg = 0, waves = 1, wsz = 0, sep_waves = 4;
  int *d_fwg_leaf = nullptr, *d_fwg_first = nullptr;
  int nfwg = 0;
  // boundary targets
  int ntarget = 0;
  int *d_trow_blk = nullptr, *d_trow_loc = nullptr, *d_trow_glob = nullptr;
  int *d_tptr = nullptr, *d_tsrc_leaf = nullptr, *d_tsrc_ri = nullptr;
  int cb_total = 0;
  // dense
  long long *d_offR = nullptr, *d_offL = nullptr;
  int *d_bsize = nullptr, *d_anc_sum = nullptr, *d_aoff = nullptr;
  double *d_inv = nullptr;
  std::vector<long long> inv_off;
  std::vector<DiagProb> h_dp;
  std::vector<GemmProb> h_gp;
  std::vector<GemmSeg> h_gs;
  std::vector<Launch> fact;
  DiagProb *d_dp = nullptr;
  GemmProb *d_gp = nullptr;
  GemmSeg *d_gs = nullptr;
  // solve
  int *d_order = nullptr;
  double *d_bp = nullptr, *d_cbuf = nullptr;
  double *d_bx = nullptr;  // staging for host-pointer solves
  int nrhs_cap = 0;
  SolvePlan sp;
  // timing
  bool profile = false;
  std::vector<hipEvent_t> ev, pev;
  double t[16] = {0};
};

namespace {

template <class T>
int dalloc(snumat *h, T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  size_t bytes = count * sizeof(T);
  hipError_t e = hipMalloc((void **)p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(SNU_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed (" +
                                      hipGetErrorString(e) + ")");
  }
  h->allocs.push_back(*p);
  h->dev_bytes += bytes;
  return SNU_OK;
}
template <class T>
int dupload(snumat *h, T **p, const std::vector<T> &v) {
  int rc = dalloc(h, p, v.size());
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return SNU_OK;
}

void free_device(snumat *h) {
  if (h->sp.d_gp) (void)hipFree(h->sp.d_gp);
  if (h->sp.d_gs) (void)hipFree(h->sp.d_gs);
  if (h->sp.d_tp) (void)hipFree(h->sp.d_tp);
  h->sp = SolvePlan();
  for (void *p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  h->ev.clear();
  for (hipEvent_t e : h->pev) (void)hipEventDestroy(e);
  h->pev.clear();
  if (h->stream) (void)hipStreamDestroy(h->stream);
  h->stream = nullptr;
}

int add_gemm(std::vector<GemmProb> &gp, std::vector<GemmSeg> &gs, int &tiles,
             double *C, long long ldc, int M, int N,
             const std::vector<GemmSeg> &segs, double &flops) {
  if (M <= 0 || N <= 0) return 0;
  GemmProb p;
  p.C = C;
  p.ldc = ldc;
  p.M = M;
  p.N = N;
  p.seg0 = (int)gs.size();
  p.nseg = 0;
  for (const GemmSeg &s : segs)
    if (s.K > 0) {
      gs.push_back(s);
      p.nseg++;
      flops += 2.0 * M * N * s.K;
    }
  if (p.nseg == 0) return 0;
  p.tiles_n = (N + 63) / 64;
  p.tile0 = tiles;
  tiles += ((M + 63) / 64) * p.tiles_n;
  gp.push_back(p);
  return 1;
}

// Host-built launch descriptors of the dense factorization (structure only).
void build_dense_plan(snumat *h) {
  const Plan &P = h->P;
  double *dense = h->d_arena + P.leaf_total;
  auto R = [&](int p) { return dense + P.offR[p]; };
  auto Lp = [&](int p) { return dense + P.offL[p]; };
  h->h_dp.clear();
  h->h_gp.clear();
  h->h_gs.clear();
  h->fact.clear();
  for (int l = P.max_level - 1; l >= 0; l--) {
    int lo = 1 << l, hi = std::min((2 << l) - 1, P.np - 1);
    int smax = 0;
    for (int i = lo; i <= hi; i++) smax = std::max(smax, P.block_size[i]);
    for (int k = 0; k < smax; k += 64) {
      Launch d{0, (int)h->h_dp.size(), 0, 0, 0};
      Launch tr{1, (int)h->h_gp.size(), 0, 0, 0};
      for (int i = lo; i <= hi; i++) {
        int s = P.block_size[i];
        if (s <= k) continue;
        int nb = std::min(64, s - k);
        DiagProb dp;
        dp.D = R(i) + k + (long long)k * s;
        dp.ld = s;
        dp.nb = nb;
        dp.Linv = h->d_inv + h->inv_off[i] + (long long)(k / 64) * 8192;
        dp.Uinv = dp.Linv + 4096;
        h->h_dp.push_back(dp);
        d.np++;
      }
      if (d.np) h->fact.push_back(d);
      // L21 <- L21 U11^-1 (rows of D below the block, all rows of Lp)
      for (int i = lo; i <= hi; i++) {
        int s = P.block_size[i], a = P.anc_sum[i];
        if (s <= k) continue;
        int nb = std::min(64, s - k);
        const double *Uinv = h->d_inv + h->inv_off[i] + (long long)(k / 64) * 8192 + 4096;
        double *X1 = R(i) + (k + nb) + (long long)k * s;
        tr.np += add_gemm(h->h_gp, h->h_gs, tr.tiles, X1, s, s - k - nb, nb,
                          {GemmSeg{X1, Uinv, s, 64, nb}}, tr.flops);
        double *X2 = Lp(i) + (long long)k * a;
        tr.np += add_gemm(h->h_gp, h->h_gs, tr.tiles, X2, a, a, nb,
                          {GemmSeg{X2, Uinv, a, 64, nb}}, tr.flops);
      }
      if (tr.np) h->fact.push_back(tr);
      // U12 <- L11^-1 U12 (row block of R right of the diagonal block)
      Launch tl{1, (int)h->h_gp.size(), 0, 0, 0};
      for (int i = lo; i <= hi; i++) {
        int s = P.block_size[i], a = P.anc_sum[i];
        if (s <= k) continue;
        int nb = std::min(64, s - k);
        const double *Linv = h->d_inv + h->inv_off[i] + (long long)(k / 64) * 8192;
        double *X = R(i) + k + (long long)(k + nb) * s;
        tl.np += add_gemm(h->h_gp, h->h_gs, tl.tiles, X, s, nb, s + a - k - nb,
                          {GemmSeg{Linv, X, 64, s, nb}}, tl.flops);
      }
      if (tl.np) h->fact.push_back(tl);
      // trailing updates inside the separator's own R and Lp
      Launch tu{2, (int)h->h_gp.size(), 0, 0, 0};
      for (int i = lo; i <= hi; i++) {
        int s = P.block_size[i], a = P.anc_sum[i];
        if (s <= k) continue;
        int nb = std::min(64, s - k);
        double *Ccol = R(i) + k + (long long)(k + nb) * s;  // U12 rows
        tu.np += add_gemm(h->h_gp, h->h_gs, tu.tiles, R(i) + (k + nb) + (long long)(k + nb) * s, s,
                          s - k - nb, s + a - k - nb,
                          {GemmSeg{R(i) + (k + nb) + (long long)k * s, Ccol, s, s, nb}}, tu.flops);
        tu.np += add_gemm(h->h_gp, h->h_gs, tu.tiles, Lp(i) + (long long)(k + nb) * a, a, a,
                          s - k - nb, {GemmSeg{Lp(i) + (long long)k * a, Ccol, a, s, nb}},
                          tu.flops);
      }
      if (tu.np) h->fact.push_back(tu);
    }
    // Schur updates of every ancestor, K-concatenated over the level's
    // separators below it (one grouped launch per level)
    Launch sc{2, (int)h->h_gp.size(), 0, 0, 0};
    for (int p = 1; p < lo; p++) {
      int lp = 31 - __builtin_clz(p);
      int first = p << (l - lp), last = ((p + 1) << (l - lp)) - 1;
      int sp = P.block_size[p], ap = P.anc_sum[p];
      std::vector<GemmSeg> segR, segL;
      for (int i = first; i <= last; i++) {
        int s = P.block_size[i], a = P.anc_sum[i];
        int off = P.anc_off(i, p);
        const double *Ai = Lp(i) + off;
        const double *Bi = R(i) + (long long)(s + off) * s;
        segR.push_back(GemmSeg{Ai, Bi, a, s, s});
        segL.push_back(GemmSeg{Ai + sp, Bi, a, s, s});
      }
      sc.np += add_gemm(h->h_gp, h->h_gs, sc.tiles, R(p), sp, sp, sp + ap, segR, sc.flops);
      sc.np += add_gemm(h->h_gp, h->h_gs, sc.tiles, Lp(p), ap, ap, sp, segL, sc.flops);
    }
    if (sc.np) h->fact.push_back(sc);
  }
}

int build_solve_plan(snumat *h, int nrhs) {
  SolvePlan &S = h->sp;
  if (S.nrhs == nrhs) return SNU_OK;
  if (S.d_gp) (void)hipFree(S.d_gp);
  if (S.d_gs) (void)hipFree(S.d_gs);
  if (S.d_tp) (void)hipFree(S.d_tp);
  S = SolvePlan();
  const Plan &P = h->P;
  double *dense = h->d_arena + P.leaf_total;
  const long long n = P.n;
  auto R = [&](int p) { return dense + P.offR[p]; };
  auto Lp = [&](int p) { return dense + P.offL[p]; };
  auto bseg = [&](int p) { return h->d_bp + P.block_start[p]; };
  for (int l = P.max_level - 1; l >= 0; l--) {
    int lo = 1 << l, hi = std::min((2 << l) - 1, P.np - 1);
    Launch tf{3, (int)S.tp.size(), 0, 0, 0};
    for (int i = lo; i <= hi; i++) {
      if (P.block_size[i] == 0) continue;
      S.tp.push_back(TriProb{R(i), P.block_size[i], P.block_size[i], h->d_inv + h->inv_off[i], bseg(i), n});
      tf.np++;
    }
    if (tf.np) S.fwd.push_back(tf);
    Launch g{2, (int)S.gp.size(), 0, 0, 0};
    for (int p = 1; p < lo; p++) {
      int lp = 31 - __builtin_clz(p);
      int first = p << (l - lp), last = ((p + 1) << (l - lp)) - 1;
      std::vector<GemmSeg> segs;
      for (int i = first; i <= last; i++)
        segs.push_back(GemmSeg{Lp(i) + P.anc_off(i, p), bseg(i), P.anc_sum[i], n, P.block_size[i]});
      g.np += add_gemm(S.gp, S.gs, g.tiles, bseg(p), n, P.block_size[p], nrhs, segs, g.flops);
    }
    if (g.np) S.fwd.push_back(g);
  }
  for (int l = 0; l < P.max_level; l++) {
    int lo = 1 << l, hi = std::min((2 << l) - 1, P.np - 1);
    Launch g{2, (int)S.gp.size(), 0, 0, 0};
    for (int i = lo; i <= hi; i++) {
      int s = P.block_size[i];
      std::vector<GemmSeg> segs;
      for (int a = i / 2; a >= 1; a /= 2)
        segs.push_back(GemmSeg{R(i) + (long long)(s + P.anc_off(i, a)) * s, bseg(a), s, n, P.block_size[a]});
      g.np += add_gemm(S.gp, S.gs, g.tiles, bseg(i), n, s, nrhs, segs, g.flops);
    }
    if (g.np) S.bwd.push_back(g);
    Launch tb{4, (int)S.tp.size(), 0, 0, 0};
    for (int i = lo; i <= hi; i++) {
      if (P.block_size[i] == 0) continue;
      S.tp.push_back(TriProb{R(i), P.block_size[i], P.block_size[i], h->d_inv + h->inv_off[i], bseg(i), n});
      tb.np++;
    }
    if (tb.np) S.bwd.push_back(tb);
  }
  auto up = [&](auto **dp, const auto &v) -> int {
    using T = typename std::remove_reference<decltype(v)>::type::value_type;
    size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    HIPCHK(hipMalloc((void **)dp, bytes));
    if (!v.empty()) HIPCHK(hipMemcpy(*dp, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return SNU_OK;
  };
  int rc;
  if ((rc = up(&S.d_gp, S.gp)) || (rc = up(&S.d_gs, S.gs)) || (rc = up(&S.d_tp, S.tp))) return rc;
  S.nrhs = nrhs;
  return SNU_OK;
}

int upload(snumat *h) {
  Plan &P = h->P;
  int rc;
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  const int np = P.np;
  // ---- leaf arrays ----
  std::vector<int> rowptr, diag, levptr, levrows, ulevptr, ulevrows;
  std::vector<int> brow, bcol_glob, bcol_blk, bcol_loc;
  long long idx_total = 0, S_total = 0;
  h->h_leaves.resize(np);
  int maxw = 1;
  int cb = 0;
  for (int r = 0; r < np; r++) {
    const LeafPlan &L = P.leaves[r];
    LeafDev &D = h->h_leaves[r];
    D.val_off = P.leaf_off[r];
    D.idx_off = idx_total;
    idx_total += L.lu_nnz;
    D.S_off = S_total;
    S_total += (long long)L.nbr * L.nbc;
    D.rp_off = (int)rowptr.size();
    D.core_n = L.core_n;
    D.leaf_size = L.leaf_size;
    D.nbr = L.nbr;
    D.nbc = L.nbc;
    rowptr.insert(rowptr.end(), L.lu_rowptr.begin(), L.lu_rowptr.end());
    diag.insert(diag.end(), L.lu_diag.begin(), L.lu_diag.end());
    diag.push_back(0);
    D.lev_off = (int)levptr.size();
    D.nlev = (int)L.lev_ptr.size() - 1;
    levptr.insert(levptr.end(), L.lev_ptr.begin(), L.lev_ptr.end());
    D.levrows_off = (int)levrows.size();
    levrows.insert(levrows.end(), L.lev_rows.begin(), L.lev_rows.end());
    D.ulev_off = (int)ulevptr.size();
    D.nulev = (int)L.ulev_ptr.size() - 1;
    ulevptr.insert(ulevptr.end(), L.ulev_ptr.begin(), L.ulev_ptr.end());
    D.ulevrows_off = (int)ulevrows.size();
    ulevrows.insert(ulevrows.end(), L.ulev_rows.begin(), L.ulev_rows.end());
    D.brow_off = (int)brow.size();
    brow.insert(brow.end(), L.brow.begin(), L.brow.end());
    D.bcol_off = (int)bcol_glob.size();
    for (int c : L.bcol) {
      int k = 1;
      while (c >= L.cstart[k + 1]) k++;
      int b = L.path[k];
      bcol_blk.push_back(b);
      bcol_loc.push_back(c - L.cstart[k]);
      bcol_glob.push_back(P.block_start[b] + c - L.cstart[k]);
    }
    D.gstart = P.block_start[L.leaf];
    D.cb_off = cb;
    cb += L.nbr;
    maxw = std::max(maxw, L.leaf_size + L.nbc);
  }
  h->cb_total = cb;
  // LDS workspace sizing: waves per workgroup limited by the LDS per block
  int lds = device_lds_limit();
  h->wsz = (maxw + 1) & ~1;
  int fit = lds / (h->wsz * (int)sizeof(double));
  if (fit < 1)
    return set_error(SNU_ERR_OOM, "leaf workspace (" + std::to_string(maxw) +
                                      " columns) exceeds the LDS of one workgroup");
  h->waves = std::min(4, fit);
  h->sep_waves = h->waves;
  std::vector<int> wcol(idx_total);
  for (int r = 0; r < np; r++)
    std::copy(P.leaves[r].wcol.begin(), P.leaves[r].wcol.end(), wcol.begin() + h->h_leaves[r].idx_off);
  // leaf-separator workgroups and boundary targets
  std::vector<int> wg_leaf, wg_first, fwg_leaf, fwg_first;
  for (int r = 0; r < np; r++)
    for (int f = 0; f < P.leaves[r].nbr; f += h->sep_waves) wg_leaf.push_back(r), wg_first.push_back(f);
  for (int r = 0; r < np; r++)
    for (int f = 0; f < P.leaves[r].nbr; f += 4) fwg_leaf.push_back(r), fwg_first.push_back(f);
  h->nwg = (int)wg_leaf.size();
  h->nfwg = (int)fwg_leaf.size();
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
  std::vector<int> trow_glob, trow_blk, trow_loc, tptr(1, 0), tsrc_leaf, tsrc_ri;
  std::vector<int> blk_of(P.n);
  for (int b = 1; b <= P.num_block; b++)
    for (int i = P.block_start[b]; i < P.block_start[b] + P.block_size[b]; i++) blk_of[i] = b;
  for (size_t s = 0; s < src.size(); s++) {
    if (s == 0 || src[s].R != src[s - 1].R) {
      if (s) tptr.push_back((int)s);
      trow_glob.push_back(src[s].R);
      trow_blk.push_back(blk_of[src[s].R]);
      trow_loc.push_back(src[s].R - P.block_start[blk_of[src[s].R]]);
    }
    tsrc_leaf.push_back(src[s].leaf);
    tsrc_ri.push_back(src[s].ri);
  }
  if (!src.empty()) tptr.push_back((int)src.size());
  h->ntarget = (int)trow_glob.size();
  // ---- dense tables ----
  std::vector<int> aoff((P.num_block + 1) * P.nk, 0);
  for (int p = 1; p <= P.num_block; p++)
    for (int d = 1; (p >> d) >= 1; d++) aoff[p * P.nk + d] = P.anc_off(p, p >> d);
  h->inv_off.assign(P.num_block + 1, 0);
  long long inv_total = 0;
  for (int p = 1; p < np; p++) {
    h->inv_off[p] = inv_total;
    inv_total += (long long)((P.block_size[p] + 63) / 64) * 8192;
  }
  // ---- allocations ----
  h->arena_len = P.leaf_total + P.dense_total;
  if ((rc = dalloc(h, &h->d_arena, (size_t)h->arena_len))) return rc;
  if ((rc = dalloc(h, &h->d_aval, (size_t)P.nnz))) return rc;
  if ((rc = dupload(h, &h->d_adest, std::vector<long long>(P.a_dest.begin(), P.a_dest.end())))) return rc;
  if ((rc = dupload(h, &h->d_leaves, h->h_leaves))) return rc;
  if ((rc = dupload(h, &h->d_rowptr, rowptr)) || (rc = dupload(h, &h->d_diag, diag)) ||
      (rc = dupload(h, &h->d_wcol, wcol)) || (rc = dupload(h, &h->d_levptr, levptr)) ||
      (rc = dupload(h, &h->d_levrows, levrows)) || (rc = dupload(h, &h->d_ulevptr, ulevptr)) ||
      (rc = dupload(h, &h->d_ulevrows, ulevrows)) || (rc = dupload(h, &h->d_brow, brow)) ||
      (rc = dupload(h, &h->d_bcol_glob, bcol_glob)) || (rc = dupload(h, &h->d_bcol_blk, bcol_blk)) ||
      (rc = dupload(h, &h->d_bcol_loc, bcol_loc)) || (rc = dupload(h, &h->d_wg_leaf, wg_leaf)) ||
      (rc = dupload(h, &h->d_wg_first, wg_first)) || (rc = dupload(h, &h->d_fwg_leaf, fwg_leaf)) ||
      (rc = dupload(h, &h->d_fwg_first, fwg_first)) || (rc = dupload(h, &h->d_trow_blk, trow_blk)) ||
      (rc = dupload(h, &h->d_trow_loc, trow_loc)) || (rc = dupload(h, &h->d_trow_glob, trow_glob)) ||
      (rc = dupload(h, &h->d_tptr, tptr)) || (rc = dupload(h, &h->d_tsrc_leaf, tsrc_leaf)) ||
      (rc = dupload(h, &h->d_tsrc_ri, tsrc_ri)) || (rc = dupload(h, &h->d_order, P.order)))
    return rc;
  if ((rc = dalloc(h, &h->d_S, (size_t)S_total))) return rc;
  std::vector<long long> offR(P.offR.begin(), P.offR.end()), offL(P.offL.begin(), P.offL.end());
  if ((rc = dupload(h, &h->d_offR, offR)) || (rc = dupload(h, &h->d_offL, offL)) ||
      (rc = dupload(h, &h->d_bsize, P.block_size)) || (rc = dupload(h, &h->d_anc_sum, P.anc_sum)) ||
      (rc = dupload(h, &h->d_aoff, aoff)))
    return rc;
  if ((rc = dalloc(h, &h->d_inv, (size_t)inv_total))) return rc;
  build_dense_plan(h);
  if ((rc = dupload(h, &h->d_dp, h->h_dp)) || (rc = dupload(h, &h->d_gp, h->h_gp)) ||
      (rc = dupload(h, &h->d_gs, h->h_gs)))
    return rc;
  return SNU_OK;
}

int ensure_rhs(snumat *h, int nrhs) {
  if (nrhs <= h->nrhs_cap) return SNU_OK;
  int rc;
  // these buffers are reallocated, so drop them from the arena list first
  auto drop = [&](void *p) {
    if (!p) return;
    auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
    if (it != h->allocs.end()) h->allocs.erase(it);
    (void)hipFree(p);
  };
  drop(h->d_bp);
  drop(h->d_cbuf);
  drop(h->d_bx);
  h->d_bp = h->d_cbuf = h->d_bx = nullptr;
  h->sp.nrhs = -1;  // plan holds d_bp pointers
  if ((rc = dalloc(h, &h->d_bp, (size_t)h->P.n * nrhs))) return rc;
  if ((rc = dalloc(h, &h->d_cbuf, (size_t)std::max(1, h->cb_total) * nrhs))) return rc;
  if ((rc = dalloc(h, &h->d_bx, (size_t)h->P.n * nrhs * 2))) return rc;
  h->nrhs_cap = nrhs;
  return SNU_OK;
}

void run_launch(snumat *h, const Launch &L, const GemmProb *gp, const GemmSeg *gs,
                const DiagProb *dp, const TriProb *tp, int nrhs) {
  switch (L.kind) {
    case 0: launch_diag_lu(dp + L.p0, L.np, h->stream); break;
    case 1: launch_gemm(gp + L.p0, L.np, gs, L.tiles, true, h->stream); break;
    case 2: launch_gemm(gp + L.p0, L.np, gs, L.tiles, false, h->stream); break;
    case 3: launch_tri_fwd(tp + L.p0, L.np, nrhs, h->stream); break;
    case 4: launch_tri_bwd(tp + L.p0, L.np, nrhs, h->stream); break;
  }
}

DenseDev dense_dev(snumat *h) {
  DenseDev d;
  d.offR = h->d_offR;
  d.offL = h->d_offL;
  d.bsize = h->d_bsize;
  d.anc_sum = h->d_anc_sum;
  d.aoff = h->d_aoff;
  d.nk = h->P.nk;
  return d;
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

}  // namespace

extern "C" void snu_default_opts(snu_opts_t *o) {
  std::memset(o, 0, sizeof(*o));
  o->np = 4;
  o->ordering = SNU_ORDER_ND;
  o->eps = 1e-15;
  o->precision = SNU_FP64;
  o->device = 0;
  o->seed = 0;
}

extern "C" const char *snu_last_error(void) { return g_err.c_str(); }

extern "C" int snu_initialize(int device) {
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0)
    return set_error(SNU_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= cnt) return set_error(SNU_ERR_NO_DEVICE, "device ordinal out of range");
  HIPCHK(hipSetDevice(device));
  return SNU_OK;
}

extern "C" int snumat_construct(const snu_csr_t *A, const snu_opts_t *opts,
                                snumat_t **out) {
  if (!out) return set_error(SNU_ERR_BAD_INPUT, "null output handle");
  *out = nullptr;
  snumat *h = new snumat();
  h->opts = *opts;
  int rc = analyze(A, opts, h->P);
  if (rc) {
    delete h;
    return rc;
  }
  if (opts->device >= 0) {
    rc = snu_initialize(opts->device);
    if (!rc) {
      h->device = opts->device;
      h->on_device = true;
      rc = upload(h);
    }
    if (!rc) {
      hipError_t e = hipMemcpy(h->d_aval, A->data, sizeof(double) * (size_t)A->nnz, hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = set_error(SNU_ERR_HIP, hipGetErrorString(e));
    }
    if (rc) {
      free_device(h);
      delete h;
      return rc;
    }
  }
  *out = h;
  return SNU_OK;
}

extern "C" int snumat_update_values(snumat_t *h, const double *values) {
  if (!h || !values) return set_error(SNU_ERR_BAD_INPUT, "null handle/values");
  if (!h->on_device) return set_error(SNU_ERR_NO_DEVICE, "handle was built host-only");
  HIPCHK(hipMemcpy(h->d_aval, values, sizeof(double) * (size_t)h->P.nnz, hipMemcpyHostToDevice));
  h->factored = false;
  return SNU_OK;
}

static int factorize_stage(snumat_t *h, int stage);
extern "C" int snumat_factorize(snumat_t *h) { return factorize_stage(h, 2); }
extern "C" int snumat_factorize_stage(snumat_t *h, int stage) {
  return factorize_stage(h, stage);
}
static int factorize_stage(snumat_t *h, int stage) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null handle");
  if (!h->on_device) return set_error(SNU_ERR_NO_DEVICE, "handle was built host-only");
  HIPCHK(hipSetDevice(h->device));
  Plan &P = h->P;
  hipStream_t st = h->stream;
  if (h->ev.empty()) {
    h->ev.resize(8);
    for (auto &e : h->ev) HIPCHK(hipEventCreate(&e));
  }
  HIPCHK(hipEventRecord(h->ev[0], st));
  launch_fill(h->d_arena, h->arena_len, st);
  launch_scatter(h->d_aval, h->d_adest, P.nnz, h->d_arena, st);
  HIPCHK(hipEventRecord(h->ev[1], st));
  launch_leaf_rows(h->d_leaves, P.np, h->d_rowptr, h->d_diag, h->d_wcol, h->d_levptr,
                   h->d_levrows, h->d_arena, h->wsz, h->waves, h->opts.eps, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(h->ev[2], st));
  launch_leaf_sep(h->d_leaves, h->d_wg_leaf, h->d_wg_first, h->nwg, h->d_rowptr, h->d_diag,
                  h->d_wcol, h->d_brow, h->d_arena, h->d_S, h->wsz, h->sep_waves, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(h->ev[3], st));
  launch_assemble_schur(h->ntarget, h->d_trow_blk, h->d_trow_loc, h->d_tptr, h->d_tsrc_leaf,
                        h->d_tsrc_ri, h->d_leaves, h->d_bcol_blk, h->d_bcol_loc, h->d_S,
                        h->d_arena + P.leaf_total, dense_dev(h), st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(h->ev[4], st));
  int ngemm = 0;
  if (stage >= 2) {
    if (h->profile && h->pev.empty()) {
      for (const Launch &L : h->fact)
        if (L.kind == 1 || L.kind == 2) ngemm++;
      h->pev.resize(2 * ngemm);
      for (auto &e : h->pev) HIPCHK(hipEventCreate(&e));
    }
    int g = 0;
    for (const Launch &L : h->fact) {
      bool gemm = L.kind == 1 || L.kind == 2;
      if (h->profile && gemm) HIPCHK(hipEventRecord(h->pev[2 * g], st));
      run_launch(h, L, h->d_gp, h->d_gs, h->d_dp, nullptr, 1);
      if (h->profile && gemm) {
        HIPCHK(hipEventRecord(h->pev[2 * g + 1], st));
        g++;
      }
    }
    ngemm = g;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(h->ev[5], st));
  HIPCHK(hipEventSynchronize(h->ev[5]));
  h->t[0] = ev_ms(h->ev[0], h->ev[1]);
  h->t[1] = ev_ms(h->ev[1], h->ev[2]);
  h->t[2] = ev_ms(h->ev[2], h->ev[3]);
  h->t[3] = ev_ms(h->ev[3], h->ev[4]);
  h->t[4] = ev_ms(h->ev[4], h->ev[5]);
  h->t[5] = ev_ms(h->ev[0], h->ev[5]);
  // t[7..9]: SUB GEMM launches (time when profiling, count, GFLOP);
  // t[10..12]: the same for the ASSIGN (trsm-by-inverse) launches
  for (int q = 7; q <= 12; q++) h->t[q] = 0;
  {
    int g = 0;
    for (const Launch &L : h->fact) {
      if (L.kind != 1 && L.kind != 2) continue;
      int base = L.kind == 2 ? 7 : 10;
      if (h->profile && stage >= 2) h->t[base] += ev_ms(h->pev[2 * g], h->pev[2 * g + 1]);
      h->t[base + 1] += 1;
      h->t[base + 2] += L.flops * 1e-9;
      g++;
    }
  }
  h->factored = stage >= 2;
  return SNU_OK;
}

static int solve_dev(snumat *h, const double *d_b, long long ldb, double *d_x,
                     long long ldx, int nrhs) {
  Plan &P = h->P;
  hipStream_t st = h->stream;
  int rc = ensure_rhs(h, nrhs);
  if (rc) return rc;
  if ((rc = build_solve_plan(h, nrhs))) return rc;
  const int n = P.n;
  HIPCHK(hipEventRecord(h->ev[6], st));
  launch_permute_in(d_b, ldb, h->d_order, h->d_bp, n, nrhs, st);
  launch_leaf_fwd(h->d_leaves, P.np, h->d_rowptr, h->d_diag, h->d_wcol, h->d_levptr,
                  h->d_levrows, h->d_arena, h->d_bp, n, nrhs, st);
  launch_leaf_fwd_sep(h->d_leaves, h->d_fwg_leaf, h->d_fwg_first, h->nfwg, 4, h->d_rowptr,
                      h->d_wcol, h->d_brow, h->d_arena, h->d_bp, h->d_cbuf, n, nrhs,
                      h->cb_total, st);
  launch_assemble_rhs(h->ntarget, h->d_trow_glob, h->d_tptr, h->d_tsrc_leaf, h->d_tsrc_ri,
                      h->d_leaves, h->d_cbuf, h->cb_total, h->d_bp, n, nrhs, st);
  for (const Launch &L : h->sp.fwd) run_launch(h, L, h->sp.d_gp, h->sp.d_gs, nullptr, h->sp.d_tp, nrhs);
  for (const Launch &L : h->sp.bwd) run_launch(h, L, h->sp.d_gp, h->sp.d_gs, nullptr, h->sp.d_tp, nrhs);
  launch_leaf_bwd(h->d_leaves, P.np, h->d_rowptr, h->d_diag, h->d_wcol, h->d_ulevptr,
                  h->d_ulevrows, h->d_bcol_glob, h->d_arena, h->d_bp, n, nrhs, st);
  launch_permute_out(h->d_bp, h->d_order, d_x, ldx, n, nrhs, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(h->ev[7], st));
  HIPCHK(hipEventSynchronize(h->ev[7]));
  h->t[6] = ev_ms(h->ev[6], h->ev[7]);
  return SNU_OK;
}

extern "C" int snumat_solve_device(snumat_t *h, const double *d_b, double *d_x,
                                   int nrhs, int ld) {
  if (!h || !d_b || !d_x || nrhs < 1) return set_error(SNU_ERR_BAD_INPUT, "bad solve arguments");
  if (!h->factored) return set_error(SNU_ERR_NOT_FACTORED, "solve before factorize");
  if (ld < h->P.n) return set_error(SNU_ERR_BAD_INPUT, "ld < n");
  HIPCHK(hipSetDevice(h->device));
  return solve_dev(h, d_b, ld, d_x, ld, nrhs);
}

extern "C" int snumat_solve(snumat_t *h, const double *b, double *x, int nrhs,
                            int ld) {
  if (!h || !b || !x || nrhs < 1) return set_error(SNU_ERR_BAD_INPUT, "bad solve arguments");
  if (!h->factored) return set_error(SNU_ERR_NOT_FACTORED, "solve before factorize");
  const int n = h->P.n;
  if (ld < n) return set_error(SNU_ERR_BAD_INPUT, "ld < n");
  HIPCHK(hipSetDevice(h->device));
  int rc = ensure_rhs(h, nrhs);
  if (rc) return rc;
  double *db = h->d_bx, *dx = h->d_bx + (size_t)n * nrhs;
  HIPCHK(hipMemcpy2DAsync(db, sizeof(double) * n, b, sizeof(double) * ld, sizeof(double) * n, nrhs,
                          hipMemcpyHostToDevice, h->stream));
  rc = solve_dev(h, db, n, dx, n, nrhs);
  if (rc) return rc;
  HIPCHK(hipMemcpy2DAsync(x, sizeof(double) * ld, dx, sizeof(double) * n, sizeof(double) * n, nrhs,
                          hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  // report Inf/NaN in the solution as a numeric failure (zero pivot in a
  // dense block: the reference ignores getrf's info, src/kernel_cuda.cu:67)
  for (int j = 0; j < nrhs; j++)
    for (int i = 0; i < n; i++)
      if (!std::isfinite(x[(size_t)j * ld + i]))
        return set_error(SNU_ERR_NUMERIC, "non-finite solution (zero pivot in the factors?)");
  return SNU_OK;
}

extern "C" int snumat_destroy(snumat_t *h) {
  if (!h) return SNU_OK;
  if (h->on_device) {
    (void)hipSetDevice(h->device);
    free_device(h);
  }
  delete h;
  return SNU_OK;
}

extern "C" int snu_solve_csr(const snu_csr_t *A, const double *b, double *x,
                             const snu_opts_t *opts) {
  snu_opts_t o;
  if (opts) o = *opts;
  else snu_default_opts(&o);
  snumat_t *h = nullptr;
  int rc = snumat_construct(A, &o, &h);
  if (rc) return rc;
  rc = snumat_factorize(h);
  if (!rc) rc = snumat_solve(h, b, x, 1, A->n);
  snumat_destroy(h);
  return rc;
}

// ---------------- introspection ----------------
extern "C" int snumat_info(const snumat_t *h, int64_t *info) {
  if (!h || !info) return set_error(SNU_ERR_BAD_INPUT, "null");
  const Plan &P = h->P;
  std::memset(info, 0, 16 * sizeof(int64_t));
  info[0] = P.n;
  info[1] = P.nnz;
  info[2] = P.np;
  info[3] = P.num_block;
  info[4] = P.max_level;
  info[5] = P.leaf_total;
  info[6] = P.dense_total;
  int mc = 0, ml = 0, mr = 0, mcol = 0;
  for (const LeafPlan &L : P.leaves) {
    mc = std::max(mc, L.core_n);
    ml = std::max(ml, L.leaf_size);
    mr = std::max(mr, L.nbr);
    mcol = std::max(mcol, L.nbc);
  }
  info[7] = mc;
  info[8] = ml;
  info[9] = mr;
  info[10] = mcol;
  info[11] = h->on_device;
  info[12] = h->factored;
  info[13] = (int64_t)h->dev_bytes;
  info[14] = h->waves;
  info[15] = h->ntarget;
  return SNU_OK;
}
extern "C" int snumat_blocks(const snumat_t *h, int *bs, int *sz) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null");
  for (int e = 1; e <= h->P.num_block; e++) bs[e - 1] = h->P.block_start[e], sz[e - 1] = h->P.block_size[e];
  return SNU_OK;
}
extern "C" int snumat_ordering(const snumat_t *h, int *order, int *sizes) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null");
  std::copy(h->P.order.begin(), h->P.order.end(), order);
  std::copy(h->P.sizes.begin(), h->P.sizes.end(), sizes);
  return SNU_OK;
}
extern "C" int snumat_permuted(const snumat_t *h, int *rp, int *ci, int *pm) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null");
  std::copy(h->P.pa_rowptr.begin(), h->P.pa_rowptr.end(), rp);
  std::copy(h->P.pa_colidx.begin(), h->P.pa_colidx.end(), ci);
  std::copy(h->P.perm_map.begin(), h->P.perm_map.end(), pm);
  return SNU_OK;
}
extern "C" int snumat_leaf_sizes(const snumat_t *h, int rank, int *s4) {
  if (!h || rank < 0 || rank >= h->P.np) return set_error(SNU_ERR_BAD_INPUT, "bad rank");
  const LeafPlan &L = h->P.leaves[rank];
  s4[0] = L.leaf_size;
  s4[1] = L.core_n;
  s4[2] = L.core_nnz;
  s4[3] = L.lu_nnz;
  return SNU_OK;
}
extern "C" int snumat_leaf_struct(const snumat_t *h, int rank, int *rp, int *ci,
                                  int *dg, int *bias, int *cmap) {
  if (!h || rank < 0 || rank >= h->P.np) return set_error(SNU_ERR_BAD_INPUT, "bad rank");
  const LeafPlan &L = h->P.leaves[rank];
  std::copy(L.lu_rowptr.begin(), L.lu_rowptr.end(), rp);
  std::copy(L.lu_colidx.begin(), L.lu_colidx.end(), ci);
  std::copy(L.lu_diag.begin(), L.lu_diag.end(), dg);
  std::copy(L.lu_bias.begin(), L.lu_bias.end(), bias);
  std::copy(L.core_map.begin(), L.core_map.end(), cmap);
  return SNU_OK;
}
extern "C" int snumat_leaf_values(const snumat_t *h, int rank, double *out) {
  if (!h || rank < 0 || rank >= h->P.np) return set_error(SNU_ERR_BAD_INPUT, "bad rank");
  if (!h->factored) return set_error(SNU_ERR_NOT_FACTORED, "not factored");
  const LeafPlan &L = h->P.leaves[rank];
  HIPCHK(hipMemcpy(out, h->d_arena + h->P.leaf_off[rank], sizeof(double) * L.lu_nnz,
                   hipMemcpyDeviceToHost));
  return SNU_OK;
}
extern "C" int snumat_sep_diag(const snumat_t *h, int e, double *out) {
  if (!h || e < 1 || e >= h->P.np) return set_error(SNU_ERR_BAD_INPUT, "bad separator");
  if (!h->factored) return set_error(SNU_ERR_NOT_FACTORED, "not factored");
  int s = h->P.block_size[e];
  HIPCHK(hipMemcpy(out, h->d_arena + h->P.leaf_total + h->P.offR[e], sizeof(double) * (size_t)s * s,
                   hipMemcpyDeviceToHost));
  return SNU_OK;
}
extern "C" int snumat_sep_block(const snumat_t *h, int i, int j, double *out) {
  const Plan &P = h->P;
  if (!h || i < 1 || j < 1 || i >= P.np || j >= P.np) return set_error(SNU_ERR_BAD_INPUT, "bad block");
  int x = std::max(i, j), y = std::min(i, j);
  while (x > y) x /= 2;
  if (x != y) return set_error(SNU_ERR_BAD_INPUT, "blocks not on one root path");
  int si = P.block_size[i], sj = P.block_size[j];
  std::vector<double> arena((size_t)P.dense_total);
  HIPCHK(hipMemcpy(arena.data(), h->d_arena + P.leaf_total, sizeof(double) * arena.size(),
                   hipMemcpyDeviceToHost));
  for (int c = 0; c < sj; c++)
    for (int r = 0; r < si; r++) out[r + (size_t)c * si] = arena[P.sep_index(i, r, j, c)];
  return SNU_OK;
}
extern "C" int snumat_timings(const snumat_t *h, double *t16) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null");
  std::memcpy(t16, h->t, sizeof(h->t));
  return SNU_OK;
}
extern "C" int snumat_counts(const snumat_t *h, double *c4) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null");
  c4[0] = h->P.f_dense;
  c4[1] = h->P.f_leaf;
  c4[2] = h->P.b_leaf;
  c4[3] = h->P.b_solve;
  return SNU_OK;
}

// ---------------- kernel-level test hooks (host pointers) ----------------
// C = C - A B (assign = 0) or C = A B (assign = 1) with one segment, through
// the grouped MFMA kernel; col-major, host arrays.
extern "C" int snu_test_gemm(int M, int N, int K, const double *A, int lda,
                             const double *B, int ldb, double *C, int ldc,
                             int assign) {
  if (M < 0 || N < 0 || K < 0 || lda < std::max(1, M) || ldb < std::max(1, K) || ldc < std::max(1, M))
    return set_error(SNU_ERR_BAD_INPUT, "bad gemm shape");
  double *dA, *dB, *dC;
  size_t na = (size_t)lda * std::max(K, 1), nb = (size_t)ldb * std::max(N, 1), nc = (size_t)ldc * std::max(N, 1);
  HIPCHK(hipMalloc((void **)&dA, na * 8));
  HIPCHK(hipMalloc((void **)&dB, nb * 8));
  HIPCHK(hipMalloc((void **)&dC, nc * 8));
  HIPCHK(hipMemcpy(dA, A, na * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dB, B, nb * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dC, C, nc * 8, hipMemcpyHostToDevice));
  std::vector<GemmProb> gp;
  std::vector<GemmSeg> gs;
  int tiles = 0;
  double fl = 0;
  add_gemm(gp, gs, tiles, dC, ldc, M, N, {GemmSeg{dA, dB, lda, ldb, K}}, fl);
  GemmProb *dgp;
  GemmSeg *dgs;
  HIPCHK(hipMalloc((void **)&dgp, sizeof(GemmProb) * std::max<size_t>(1, gp.size())));
  HIPCHK(hipMalloc((void **)&dgs, sizeof(GemmSeg) * std::max<size_t>(1, gs.size())));
  if (!gp.empty()) {
    HIPCHK(hipMemcpy(dgp, gp.data(), sizeof(GemmProb) * gp.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dgs, gs.data(), sizeof(GemmSeg) * gs.size(), hipMemcpyHostToDevice));
    launch_gemm(dgp, (int)gp.size(), dgs, tiles, assign != 0, nullptr);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(C, dC, nc * 8, hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  (void)hipFree(dgp);
  (void)hipFree(dgs);
  return SNU_OK;
}

// In-place no-pivot LU of an nb x nb block (nb <= 64) and the inverses of its
// triangular factors (64 x 64, ld 64), through k_diag_lu.
extern "C" int snu_test_diag_lu(int nb, double *D, int ld, double *Linv,
                                double *Uinv) {
  if (nb < 1 || nb > 64 || ld < nb) return set_error(SNU_ERR_BAD_INPUT, "bad block");
  double *dD, *dI;
  HIPCHK(hipMalloc((void **)&dD, sizeof(double) * ld * nb));
  HIPCHK(hipMalloc((void **)&dI, sizeof(double) * 8192));
  HIPCHK(hipMemcpy(dD, D, sizeof(double) * ld * nb, hipMemcpyHostToDevice));
  DiagProb p{dD, ld, nb, dI, dI + 4096};
  DiagProb *dp;
  HIPCHK(hipMalloc((void **)&dp, sizeof(DiagProb)));
  HIPCHK(hipMemcpy(dp, &p, sizeof(DiagProb), hipMemcpyHostToDevice));
  launch_diag_lu(dp, 1, nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(D, dD, sizeof(double) * ld * nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(Linv, dI, sizeof(double) * 4096, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(Uinv, dI + 4096, sizeof(double) * 4096, hipMemcpyDeviceToHost));
  (void)hipFree(dD);
  (void)hipFree(dI);
  (void)hipFree(dp);
  return SNU_OK;
}

// ---------------- device memory plumbing (bench, tests) ----------------
extern "C" int snu_dev_alloc(size_t bytes, void **p) {
  if (!p) return set_error(SNU_ERR_BAD_INPUT, "null");
  HIPCHK(hipMalloc(p, bytes ? bytes : 1));
  return SNU_OK;
}
extern "C" int snu_dev_free(void *p) {
  HIPCHK(hipFree(p));
  return SNU_OK;
}
extern "C" int snu_dev_copy(void *dst, const void *src, size_t bytes, int kind) {
  hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                                                                  : hipMemcpyDeviceToDevice;
  HIPCHK(hipMemcpy(dst, src, bytes, k));
  return SNU_OK;
}
extern "C" int snu_dev_sync(void) {
  HIPCHK(hipDeviceSynchronize());
  return SNU_OK;
}
extern "C" int snumat_set_profiling(snumat_t *h, int on) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null");
  h->profile = on != 0;
  return SNU_OK;
}
