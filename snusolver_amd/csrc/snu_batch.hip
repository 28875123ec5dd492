// Batches of independent matrices factorized and solved together: every
// step runs as one launch across all of them.  The leaves of all matrices
// share the leaf launches; separators of equal tree height (counted from
// the leaves, so trees of different depths align at the bottom) share one
// grouped launch per step of the dense level loop; solves pack the
// right-hand sides of all matrices, each with its own column count, into one
// buffer and run one launch per sweep step and tree height.
//
// No reference counterpart (the reference factorizes one matrix per MPI
// job); each matrix's arithmetic is unchanged, so its factors and solutions
// equal those of a standalone factorization and solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "snu_engine.h"

using namespace snu;

struct snubatch {
  std::vector<snumat *> mats;
  int prec = SNU_FP64, device = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  // factor set
  WorkDev wd;
  PrepDev *d_prep = nullptr;
  IrDev *d_ir_a = nullptr;  // |A| norm tables (FP32)
  unsigned long long *d_anorm = nullptr;
  long long max_len = 0, max_nnz = 0;
  int max_n = 0;
  DensePlan<double> f64;
  DensePlan<float> f32;
  FrontPlan<double> fr64;  // the members' leaf fronts, merged per height
  FrontPlan<float> fr32;
  // solves (sized for the last right-hand-side pattern)
  std::vector<int> pattern;  // nrhs per matrix of the current solve plan
  long long vec_cap = 0;
  int cb_cap = 0, slot_cap = 0;
  void *bp = nullptr, *ys = nullptr, *cbuf = nullptr;
  double *r = nullptr, *dx = nullptr;
  unsigned long long *norms = nullptr;
  int *d_active = nullptr;
  VecDev *d_vec = nullptr;
  PermDev *d_perm = nullptr;  // 2 x nmat: b -> x, then r -> dx
  IrDev *d_ir = nullptr;
  GemvPlan<double> s64;
  GemvPlan<float> s32;
  std::vector<VecDev> h_vec;
  std::vector<double> anorm;
  int ir_steps = 0;
  std::vector<double> berr;
  double *stage = nullptr;  // host-pointer solves: packed b then x
  size_t stage_cap = 0;
  bool factor_plan = false;  // merged dense plan built (on first factorize)
  bool use_graph = false;
  FactorGraph graph;
  std::vector<hipEvent_t> ev;
  double t[16] = {0};
};

namespace {

int balloc(snubatch *b, void **p, size_t bytes) {
  *p = nullptr;
  hipError_t e = hipMalloc(p, std::max<size_t>(1, bytes));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(SNU_ERR_OOM, std::string("batch allocation failed: ") + hipGetErrorString(e));
  }
  b->allocs.push_back(*p);
  return SNU_OK;
}
template <class T>
int bupload(snubatch *b, T **p, const std::vector<T> &v) {
  int rc = balloc(b, (void **)p, v.size() * sizeof(T));
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return SNU_OK;
}
void bfree(snubatch *b, void *p) {
  if (!p) return;
  auto it = std::find(b->allocs.begin(), b->allocs.end(), p);
  if (it != b->allocs.end()) b->allocs.erase(it);
  (void)hipFree(p);
}
void free_batch(snubatch *b) {
  free_graph(b->graph);
  free_dense_plan(b->f64);
  free_dense_plan(b->f32);
  free_dense_plan(b->fr64.dense);
  free_dense_plan(b->fr32.dense);
  free_gemv_plan(b->s64);
  free_gemv_plan(b->s32);
  for (void *p : b->allocs) (void)hipFree(p);
  b->allocs.clear();
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->stream = nullptr;
}

template <typename T>
int build_factor_set(snubatch *b) {
  std::vector<const DensePlan<T> *> parts;
  for (snumat *h : b->mats) parts.push_back(&vals<T>(h).fact);
  DensePlan<T> &out = std::is_same<T, double>::value ? (DensePlan<T> &)b->f64 : (DensePlan<T> &)b->f32;
  merge_dense(parts, out);
  return upload_dense_plan(out);
}

// Solve plan and buffers for a right-hand-side pattern (columns per matrix;
// 0 skips a matrix).
template <typename T>
int ensure_pattern(snubatch *b, const std::vector<int> &nrhs) {
  if (b->pattern == nrhs) return SNU_OK;
  const int nm = (int)b->mats.size();
  long long tot = 0;
  int cbt = 0, slots = 0;
  for (int m = 0; m < nm; m++) {
    tot += (long long)b->mats[m]->P.n * nrhs[m];
    cbt = std::max(cbt, nrhs[m]);
    slots += nrhs[m];
  }
  int rc;
  if (tot > b->vec_cap) {
    bfree(b, b->bp);
    bfree(b, b->ys);
    bfree(b, b->r);
    bfree(b, b->dx);
    b->bp = b->ys = nullptr;
    b->r = b->dx = nullptr;
    if ((rc = balloc(b, &b->bp, sizeof(T) * tot)) || (rc = balloc(b, &b->ys, sizeof(T) * tot)))
      return rc;
    if (b->prec == SNU_FP32 &&
        ((rc = balloc(b, (void **)&b->r, sizeof(double) * tot)) ||
         (rc = balloc(b, (void **)&b->dx, sizeof(double) * tot))))
      return rc;
    b->vec_cap = tot;
  }
  if (cbt > b->cb_cap) {
    bfree(b, b->cbuf);
    b->cbuf = nullptr;
    if ((rc = balloc(b, &b->cbuf, sizeof(T) * (size_t)std::max(1, b->wd.cb_total) * cbt))) return rc;
    b->cb_cap = cbt;
  }
  if (slots > b->slot_cap) {
    bfree(b, b->norms);
    b->norms = nullptr;
    if ((rc = balloc(b, (void **)&b->norms, sizeof(unsigned long long) * 3 * (size_t)slots))) return rc;
    b->slot_cap = slots;
  }
  b->h_vec.assign(nm, VecDev{0, 0, 0});
  long long base = 0;
  for (int m = 0; m < nm; m++) {
    b->h_vec[m] = VecDev{base, b->mats[m]->P.n, nrhs[m]};
    base += (long long)b->mats[m]->P.n * nrhs[m];
  }
  HIPCHK(hipMemcpyAsync(b->d_vec, b->h_vec.data(), sizeof(VecDev) * nm, hipMemcpyHostToDevice, b->stream));
  GemvPlan<T> &out = std::is_same<T, double>::value ? (GemvPlan<T> &)b->s64 : (GemvPlan<T> &)b->s32;
  std::vector<GemvPlan<T>> parts(nm);
  std::vector<const GemvPlan<T> *> ptrs;
  T *bp = (T *)b->bp, *ys = (T *)b->ys;
  for (int m = 0; m < nm; m++) {
    if (!nrhs[m]) continue;
    build_solve_plan<T>(b->mats[m], bp + b->h_vec[m].base, ys + b->h_vec[m].base, nrhs[m], parts[m]);
    ptrs.push_back(&parts[m]);
  }
  merge_gemv(ptrs, out);  // keeps out's device arrays, which grow only
  if ((rc = upload_gemv_plan(out, b->stream))) return rc;
  b->pattern = nrhs;
  return SNU_OK;
}

template <typename T>
int batch_factorize(snubatch *b) {
  FactorSet fs;
  fs.prep = b->d_prep;
  fs.nmat = (int)b->mats.size();
  fs.max_len = b->max_len;
  fs.max_nnz = b->max_nnz;
  fs.ir = b->prec == SNU_FP32 ? b->d_ir_a : nullptr;
  fs.anorm_out = b->d_anorm;
  fs.max_n = b->max_n;
  fs.wd = &b->wd;
  fs.fronts = std::is_same<T, double>::value ? (const void *)&b->fr64 : (const void *)&b->fr32;
  fs.eps = b->mats[0]->opts.eps;
  fs.split = b->mats[0]->opts.fp32_gemm == SNU_GEMM_SPLIT;
  const DensePlan<T> &plan = std::is_same<T, double>::value ? (const DensePlan<T> &)b->f64
                                                            : (const DensePlan<T> &)b->f32;
  int rc = run_factor<T>(fs, plan, b->stream, &b->ev, false, nullptr, nullptr,
                         b->use_graph ? &b->graph : nullptr);
  if (rc) return rc;
  HIPCHK(hipEventSynchronize(b->ev[5]));
  for (int q = 0; q < 5; q++) b->t[q] = ev_ms(b->ev[q], b->ev[q + 1]);
  b->t[5] = ev_ms(b->ev[0], b->ev[5]);
  const int nm = (int)b->mats.size();
  b->anorm.assign(nm, 0.0);
  if (b->prec == SNU_FP32) {
    std::vector<unsigned long long> u(nm);
    HIPCHK(hipMemcpy(u.data(), b->d_anorm, sizeof(unsigned long long) * nm, hipMemcpyDeviceToHost));
    for (int m = 0; m < nm; m++) std::memcpy(&b->anorm[m], &u[m], sizeof(double));
  }
  for (int m = 0; m < nm; m++) {
    b->mats[m]->factored = true;
    b->mats[m]->fb64_current = false;  // an FP64 twin refactorizes on its next use
    b->mats[m]->anorm = b->anorm[m];
  }
  return SNU_OK;
}

template <typename T>
int batch_solve(snubatch *b, const double *const *d_b, double *const *d_x, const int *nrhs,
                const int *ld) {
  const int nm = (int)b->mats.size();
  std::vector<int> pat(nrhs, nrhs + nm);
  int rc = ensure_pattern<T>(b, pat);
  if (rc) return rc;
  hipStream_t st = b->stream;
  std::vector<PermDev> pd(2 * (size_t)nm);
  long long max_elems = 0;
  int max_rhs = 0, max_n = 0;
  for (int m = 0; m < nm; m++) {
    const snumat *h = b->mats[m];
    const int n = h->P.n;
    const long long base = b->h_vec[m].base;
    pd[m] = PermDev{h->d_order, d_b[m], d_x[m], (long long)ld[m], (long long)ld[m], n, m};
    pd[nm + m] = PermDev{h->d_order, b->r ? b->r + base : nullptr, b->dx ? b->dx + base : nullptr,
                         (long long)n, (long long)n, n, m};
    max_elems = std::max(max_elems, (long long)n * nrhs[m]);
    max_rhs = std::max(max_rhs, nrhs[m]);
    max_n = std::max(max_n, n);
  }
  HIPCHK(hipMemcpyAsync(b->d_perm, pd.data(), sizeof(PermDev) * pd.size(), hipMemcpyHostToDevice, st));
  SolveSet ss;
  ss.perm = b->d_perm;
  ss.vec = b->d_vec;
  ss.nmat = nm;
  ss.max_elems = max_elems;
  ss.max_rhs = max_rhs;
  ss.wd = &b->wd;
  ss.cb_ld = b->wd.cb_total;
  HIPCHK(hipEventRecord(b->ev[6], st));
  if (b->prec == SNU_FP32) {
    std::vector<IrDev> ir(nm);
    std::vector<int> slot_mat;
    for (int m = 0; m < nm; m++) {
      const snumat *h = b->mats[m];
      const long long base = b->h_vec[m].base;
      IrDev &e = ir[m];
      std::memset(&e, 0, sizeof(e));
      e.rowptr = h->d_arow;
      e.colidx = h->d_acol;
      e.val = h->d_aval;
      e.b = d_b[m];
      e.x = d_x[m];
      e.r = b->r + base;
      e.dx = b->dx + base;
      e.ldb = e.ldx = ld[m];
      e.ldr = e.lddx = h->P.n;
      e.n = h->P.n;
      e.nrhs = nrhs[m];
      e.norm0 = (int)slot_mat.size();
      for (int c = 0; c < nrhs[m]; c++) slot_mat.push_back(m);
    }
    HIPCHK(hipMemcpyAsync(b->d_ir, ir.data(), sizeof(IrDev) * nm, hipMemcpyHostToDevice, st));
    const snu_opts_t &o = b->mats[0]->opts;
    RefineSet rs;
    rs.sweep_b = ss;
    rs.sweep_r = ss;
    rs.sweep_r.perm = b->d_perm + nm;
    rs.ir = b->d_ir;
    rs.nmat = nm;
    rs.max_n = max_n;
    rs.max_elems = max_elems;
    rs.norms = b->norms;
    rs.nslots = (int)slot_mat.size();
    rs.slot_mat = &slot_mat;
    b->anorm.assign(nm, 0.0);  // |A| of each member, from its last factorization
    for (int m = 0; m < nm; m++) b->anorm[m] = b->mats[m]->anorm;
    rs.anorm = &b->anorm;
    rs.d_active = b->d_active;
    rs.max_steps = o.ir_max_steps > 0 ? o.ir_max_steps : o.ir_max_steps < 0 ? 0 : 10;
    rs.tol = o.ir_tol > 0 ? o.ir_tol : 1e-15;
    rc = run_refined(rs, (float *)b->bp, (float *)b->cbuf, b->s32, st, &b->ir_steps, &b->berr);
    if (rc && rc != SNU_ERR_NOT_CONVERGED) return rc;
    b->t[15] = 0;
    if (rc == SNU_ERR_NOT_CONVERGED) {
      // members whose refinement stalled above 1e-12 are answered by FP64
      // factors of the same matrix (each member's FP64 twin); the batch's
      // own work is complete (run_refined has read the final norms back)
      for (int m = 0; m < nm; m++) {
        if (!nrhs[m] || b->berr[m] <= 1e-12) continue;
        if ((rc = fp64_fallback_solve(b->mats[m], d_b[m], ld[m], d_x[m], ld[m], nrhs[m]))) return rc;
        b->t[15] += 1;
      }
      HIPCHK(hipSetDevice(b->device));
      rc = SNU_OK;
    }
  } else {
    run_sweep<double>(ss, (double *)b->bp, (double *)b->cbuf, b->s64, st);
    b->ir_steps = 0;
    b->berr.assign(nm, 0.0);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->ev[7], st));
  HIPCHK(hipEventSynchronize(b->ev[7]));
  b->t[6] = ev_ms(b->ev[6], b->ev[7]);
  b->t[13] = b->ir_steps;
  double worst = 0;
  for (double e : b->berr) worst = std::max(worst, e);
  b->t[14] = worst;
  return rc;
}

}  // namespace

extern "C" int snubatch_create(snumat_t *const *mats, int count, snubatch_t **out) {
  if (!out || !mats || count < 1) return set_error(SNU_ERR_BAD_INPUT, "bad batch arguments");
  *out = nullptr;
  for (int m = 0; m < count; m++) {
    if (!mats[m] || !mats[m]->on_device) return set_error(SNU_ERR_BAD_INPUT, "batch member not on a device");
    if (mats[m]->prec != mats[0]->prec || mats[m]->device != mats[0]->device ||
        (mats[m]->prec == SNU_FP32 && mats[m]->opts.fp32_gemm != mats[0]->opts.fp32_gemm))
      return set_error(SNU_ERR_BAD_INPUT, "batch members must share device, precision and FP32 GEMM form");
    for (int q = 0; q < m; q++)
      if (mats[q] == mats[m]) return set_error(SNU_ERR_BAD_INPUT, "matrix listed twice in a batch");
  }
  snubatch *b = new snubatch();
  b->mats.assign(mats, mats + count);
  b->prec = mats[0]->prec;
  b->device = mats[0]->device;
  auto fail = [&](int rc) {
    free_batch(b);
    delete b;
    return rc;
  };
  hipError_t he = hipSetDevice(b->device);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (he != hipSuccess) return fail(set_error(SNU_ERR_HIP, hipGetErrorString(he)));
  b->ev.resize(8);
  for (auto &e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(set_error(SNU_ERR_HIP, "event creation failed"));
  // concatenated leaves and work lists
  std::vector<LeafDev> leaves;
  WorkLists W;
  W.tptr.assign(1, 0);
  W.tile_ptr.assign(1, 0);
  int max_leaf = 0;
  for (snumat *h : b->mats) max_leaf = std::max(max_leaf, h->wd.max_leaf);
  int cb = 0;
  std::vector<PrepDev> prep;
  std::vector<IrDev> ira;
  for (int m = 0; m < count; m++) {
    snumat *h = b->mats[m];
    const int lbase = (int)leaves.size(), sbase = (int)W.tsrc_leaf.size();
    for (LeafDev L : h->h_leaves) {
      L.cb_off += cb;
      L.vec = m;
      leaves.push_back(L);
    }
    for (int r = 0; r < h->P.np; r++) {
      const int nbr = h->P.leaves[r].nbr;
      for (int f = 0; f < nbr; f += 4) W.fwg_leaf.push_back(lbase + r), W.fwg_first.push_back(f);
    }
    const WorkLists &w = h->wl;
    W.trow_blk.insert(W.trow_blk.end(), w.trow_blk.begin(), w.trow_blk.end());
    W.trow_loc.insert(W.trow_loc.end(), w.trow_loc.begin(), w.trow_loc.end());
    W.trow_glob.insert(W.trow_glob.end(), w.trow_glob.begin(), w.trow_glob.end());
    for (size_t t = 1; t < w.tptr.size(); t++) W.tptr.push_back(sbase + w.tptr[t]);
    for (int s : w.tsrc_leaf) W.tsrc_leaf.push_back(lbase + s);
    W.tsrc_ri.insert(W.tsrc_ri.end(), w.tsrc_ri.begin(), w.tsrc_ri.end());
    const int gbase = (int)W.seg_leaf.size();
    W.tile_blk.insert(W.tile_blk.end(), w.tile_blk.begin(), w.tile_blk.end());
    W.tile_row0.insert(W.tile_row0.end(), w.tile_row0.begin(), w.tile_row0.end());
    for (size_t t = 1; t < w.tile_ptr.size(); t++) W.tile_ptr.push_back(gbase + w.tile_ptr[t]);
    for (int s : w.seg_leaf) W.seg_leaf.push_back(lbase + s);
    W.seg_ri.insert(W.seg_ri.end(), w.seg_ri.begin(), w.seg_ri.end());
    cb += h->wd.cb_total;
    PrepDev pd;
    pd.arena = b->prec == SNU_FP32 ? (void *)h->v32.arena : (void *)h->v64.arena;
    pd.inv = b->prec == SNU_FP32 ? (void *)h->v32.inv : (void *)h->v64.inv;
    pd.arena_len = h->arena_len;
    pd.inv_len = h->inv_total;
    pd.aval = h->d_aval;
    pd.adest = h->d_adest;
    pd.nnz = h->P.nnz;
    prep.push_back(pd);
    IrDev e;
    std::memset(&e, 0, sizeof(e));
    e.rowptr = h->d_arow;
    e.colidx = h->d_acol;
    e.val = h->d_aval;
    e.n = h->P.n;
    ira.push_back(e);
    b->max_len = std::max(b->max_len, std::max(h->arena_len, h->inv_total));
    b->max_nnz = std::max(b->max_nnz, (long long)h->P.nnz);
    b->max_n = std::max(b->max_n, h->P.n);
  }
  WorkDev &wd = b->wd;
  wd.nleaves = (int)leaves.size();
  wd.nfwg = (int)W.fwg_leaf.size();
  wd.ntarget = (int)W.trow_glob.size();
  wd.ntile = (int)W.tile_blk.size();
  wd.cb_total = cb;
  wd.max_leaf = max_leaf;
  wd.max_lev = 0;
  for (snumat *h : b->mats) wd.max_lev = std::max(wd.max_lev, h->wd.max_lev);
  int rc;
  if ((rc = bupload(b, &wd.leaves, leaves)) || (rc = bupload(b, &wd.fwg_leaf, W.fwg_leaf)) ||
      (rc = bupload(b, &wd.fwg_first, W.fwg_first)) || (rc = bupload(b, &wd.trow_blk, W.trow_blk)) ||
      (rc = bupload(b, &wd.trow_loc, W.trow_loc)) || (rc = bupload(b, &wd.trow_glob, W.trow_glob)) ||
      (rc = bupload(b, &wd.tptr, W.tptr)) || (rc = bupload(b, &wd.tsrc_leaf, W.tsrc_leaf)) ||
      (rc = bupload(b, &wd.tsrc_ri, W.tsrc_ri)) || (rc = bupload(b, &b->d_prep, prep)) ||
      (rc = bupload(b, &wd.tile_blk, W.tile_blk)) || (rc = bupload(b, &wd.tile_row0, W.tile_row0)) ||
      (rc = bupload(b, &wd.tile_ptr, W.tile_ptr)) || (rc = bupload(b, &wd.seg_leaf, W.seg_leaf)) ||
      (rc = bupload(b, &wd.seg_ri, W.seg_ri)) ||
      (rc = bupload(b, &b->d_ir_a, ira)))
    return fail(rc);
  // the members' leaf fronts, merged per height (descriptors point into the
  // members' own arrays)
  {
    auto fronts = [&](auto tag) -> int {
      using T = decltype(tag);
      FrontPlan<T> &fp = std::is_same<T, double>::value ? (FrontPlan<T> &)b->fr64 : (FrontPlan<T> &)b->fr32;
      std::vector<const FrontPlan<T> *> parts;
      for (snumat *h : b->mats) parts.push_back(&vals<T>(h).fronts);
      merge_fronts(parts, fp);
      int r;
      if ((r = upload_dense_plan(fp.dense)) || (r = bupload(b, &fp.d_small, fp.small)) ||
          (r = bupload(b, &fp.d_large, fp.large)) || (r = bupload(b, &fp.d_work, fp.work)) || (r = bupload(b, &fp.d_wrange, fp.wrange)) ||
          (r = bupload(b, &fp.d_ls, fp.ls)) || (r = bupload(b, &fp.d_tiles, fp.tiles)))
        return r;
      return SNU_OK;
    };
    if ((rc = b->prec == SNU_FP32 ? fronts(float()) : fronts(double()))) return fail(rc);
  }
  if ((rc = balloc(b, (void **)&b->d_anorm, sizeof(unsigned long long) * count)) ||
      (rc = balloc(b, (void **)&b->d_active, sizeof(int) * count)) ||
      (rc = balloc(b, (void **)&b->d_vec, sizeof(VecDev) * count)) ||
      (rc = balloc(b, (void **)&b->d_perm, sizeof(PermDev) * 2 * count)) ||
      (rc = balloc(b, (void **)&b->d_ir, sizeof(IrDev) * count)))
    return fail(rc);
  *out = b;
  return SNU_OK;
}

extern "C" int snubatch_factorize(snubatch_t *b) {
  if (!b) return set_error(SNU_ERR_BAD_INPUT, "null batch");
  HIPCHK(hipSetDevice(b->device));
  if (!b->factor_plan) {
    int rc = b->prec == SNU_FP32 ? build_factor_set<float>(b) : build_factor_set<double>(b);
    if (rc) return rc;
    b->factor_plan = true;
  }
  for (snumat *h : b->mats) h->factored = false;
  return b->prec == SNU_FP32 ? batch_factorize<float>(b) : batch_factorize<double>(b);
}

extern "C" int snubatch_solve_device(snubatch_t *b, const double *const *d_b, double *const *d_x,
                                     const int *nrhs, const int *ld) {
  if (!b || !d_b || !d_x || !nrhs || !ld) return set_error(SNU_ERR_BAD_INPUT, "bad batch solve arguments");
  const int nm = (int)b->mats.size();
  for (int m = 0; m < nm; m++) {
    if (nrhs[m] < 0 || (nrhs[m] > 0 && (!d_b[m] || !d_x[m] || ld[m] < b->mats[m]->P.n || d_b[m] == d_x[m])))
      return set_error(SNU_ERR_BAD_INPUT, "bad batch solve arguments for matrix " + std::to_string(m));
    if (nrhs[m] > 0 && !b->mats[m]->factored)
      return set_error(SNU_ERR_NOT_FACTORED, "batch member not factored");
  }
  HIPCHK(hipSetDevice(b->device));
  return b->prec == SNU_FP32 ? batch_solve<float>(b, d_b, d_x, nrhs, ld)
                             : batch_solve<double>(b, d_b, d_x, nrhs, ld);
}

extern "C" int snubatch_solve(snubatch_t *b, const double *const *hb, double *const *hx,
                              const int *nrhs, const int *ld) {
  if (!b || !hb || !hx || !nrhs || !ld) return set_error(SNU_ERR_BAD_INPUT, "bad batch solve arguments");
  const int nm = (int)b->mats.size();
  HIPCHK(hipSetDevice(b->device));
  size_t tot = 0;
  for (int m = 0; m < nm; m++) {
    if (nrhs[m] < 0 || (nrhs[m] > 0 && (!hb[m] || !hx[m] || ld[m] < b->mats[m]->P.n)))
      return set_error(SNU_ERR_BAD_INPUT, "bad batch solve arguments for matrix " + std::to_string(m));
    tot += (size_t)b->mats[m]->P.n * nrhs[m];
  }
  if (2 * tot > b->stage_cap) {
    bfree(b, b->stage);
    b->stage = nullptr;
    b->stage_cap = 0;
    int arc = balloc(b, (void **)&b->stage, sizeof(double) * 2 * std::max<size_t>(1, tot));
    if (arc) return arc;
    b->stage_cap = 2 * std::max<size_t>(1, tot);
  }
  double *stage = b->stage;
  std::vector<const double *> db(nm, nullptr);
  std::vector<double *> dx(nm, nullptr);
  std::vector<int> dld(nm);
  size_t off = 0;
  int rc = SNU_OK;
  for (int m = 0; m < nm && !rc; m++) {
    const int n = b->mats[m]->P.n;
    dld[m] = n;
    if (!nrhs[m]) continue;
    double *pb = stage + off, *px = stage + tot + off;
    off += (size_t)n * nrhs[m];
    if (hipMemcpy2DAsync(pb, sizeof(double) * n, hb[m], sizeof(double) * ld[m], sizeof(double) * n,
                         nrhs[m], hipMemcpyHostToDevice, b->stream) != hipSuccess)
      rc = set_error(SNU_ERR_HIP, "batch rhs upload failed");
    db[m] = pb;
    dx[m] = px;
  }
  if (!rc) rc = snubatch_solve_device(b, db.data(), dx.data(), nrhs, dld.data());
  if (!rc || rc == SNU_ERR_NOT_CONVERGED) {
    for (int m = 0; m < nm; m++) {
      if (!nrhs[m]) continue;
      const int n = b->mats[m]->P.n;
      if (hipMemcpy2DAsync(hx[m], sizeof(double) * ld[m], dx[m], sizeof(double) * n, sizeof(double) * n,
                           nrhs[m], hipMemcpyDeviceToHost, b->stream) != hipSuccess) {
        rc = set_error(SNU_ERR_HIP, "batch solution download failed");
        break;
      }
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) rc = set_error(SNU_ERR_HIP, "batch solve sync failed");
  }
  return rc;
}

extern "C" int snubatch_timings(const snubatch_t *b, double *t16) {
  if (!b || !t16) return set_error(SNU_ERR_BAD_INPUT, "null");
  std::memcpy(t16, b->t, sizeof(b->t));
  return SNU_OK;
}

extern "C" int snubatch_launches(const snubatch_t *b, int *factor_launches, int *solve_launches) {
  if (!b) return set_error(SNU_ERR_BAD_INPUT, "null");
  if (factor_launches)
    *factor_launches = (int)(b->prec == SNU_FP32 ? b->f32.launches.size() : b->f64.launches.size());
  if (solve_launches)
    *solve_launches = (int)(b->prec == SNU_FP32 ? b->s32.fwd.size() + b->s32.bwd.size()
                                                : b->s64.fwd.size() + b->s64.bwd.size());
  return SNU_OK;
}

extern "C" int snubatch_use_graph(snubatch_t *b, int on) {
  if (!b) return set_error(SNU_ERR_BAD_INPUT, "null");
  b->use_graph = on != 0;
  return SNU_OK;
}

extern "C" int snubatch_destroy(snubatch_t *b) {
  if (!b) return SNU_OK;
  (void)hipSetDevice(b->device);
  free_batch(b);
  delete b;
  return SNU_OK;
}
