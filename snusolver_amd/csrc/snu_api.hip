// C-ABI entry points of the MI355X SnuSolver path: construct -> factorize
// -> solve on one GPU, one handle per matrix (the batch and serving entry
// points live in snu_batch.hip).
//
//   SnuMat::SnuMat (SnuMat.cu:67-112)  -> snumat_construct
//   SnuMat::solve (SnuMat_factorize.cu:213-357) -> snumat_factorize +
//                                         snumat_solve
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "snu_engine.h"

namespace snu {
static thread_local std::string g_err;
int set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace snu

using namespace snu;

namespace {

// The handle's stream and events, created on its first standalone use
// (members of a batch run on the batch's stream).
int ensure_stream(snumat *h) {
  if (!h->stream) HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  if (h->ev.empty()) {
    h->ev.resize(8);
    for (auto &e : h->ev) HIPCHK(hipEventCreate(&e));
  }
  return SNU_OK;
}

template <class T>
int halloc(snumat *h, T **p, size_t count) {
  return device_alloc(h, (void **)p, std::max<size_t>(1, count) * sizeof(T));
}
void drop_alloc(snumat *h, void *p) {
  if (!p) return;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
  if (it != h->allocs.end()) h->allocs.erase(it);
  (void)hipFree(p);
}

// Solve vectors for nrhs columns (grown on demand; the solve plan points at
// them, so it is rebuilt after a reallocation).
int ensure_rhs(snumat *h, int nrhs) {
  if (nrhs <= h->nrhs_cap) return SNU_OK;
  int rc;
  const size_t n = (size_t)h->P.n;
  auto vec = [&](auto &VV) -> int {
    drop_alloc(h, VV.bp);
    drop_alloc(h, VV.cbuf);
    drop_alloc(h, VV.ys);
    VV.bp = VV.cbuf = VV.ys = nullptr;
    VV.plan_nrhs = -1;
    int r;
    if ((r = halloc(h, &VV.bp, n * nrhs)) || (r = halloc(h, &VV.ys, n * nrhs)) ||
        (r = halloc(h, &VV.cbuf, (size_t)std::max(1, h->wd.cb_total) * nrhs)))
      return r;
    return SNU_OK;
  };
  drop_alloc(h, h->d_bx);
  drop_alloc(h, h->d_r);
  drop_alloc(h, h->d_dx);
  drop_alloc(h, h->d_colnorms);
  h->d_bx = h->d_r = h->d_dx = nullptr;
  h->d_colnorms = nullptr;
  if ((rc = h->prec == SNU_FP32 ? vec(h->v32) : vec(h->v64))) return rc;
  if ((rc = halloc(h, &h->d_bx, n * nrhs * 2))) return rc;
  if (h->prec == SNU_FP32 &&
      ((rc = halloc(h, &h->d_r, n * nrhs)) || (rc = halloc(h, &h->d_dx, n * nrhs)) ||
       (rc = halloc(h, &h->d_colnorms, 3 * (size_t)nrhs))))
    return rc;
  h->nrhs_cap = nrhs;
  return SNU_OK;
}

template <typename T>
int ensure_plan(snumat *h, int nrhs) {
  Values<T> &VV = vals<T>(h);
  if (VV.plan_nrhs == nrhs) return SNU_OK;
  build_solve_plan<T>(h, VV.bp, VV.ys, nrhs, VV.sp);  // device arrays grow only
  int rc = upload_gemv_plan(VV.sp, h->stream);
  if (rc) return rc;
  VV.plan_nrhs = nrhs;
  return SNU_OK;
}

IrDev ir_entry(snumat *h, const double *b, long long ldb, double *x, long long ldx, int nrhs) {
  IrDev e;
  std::memset(&e, 0, sizeof(e));
  e.rowptr = h->d_arow;
  e.colidx = h->d_acol;
  e.val = h->d_aval;
  e.b = b;
  e.x = x;
  e.r = h->d_r;
  e.dx = h->d_dx;
  e.ldb = ldb;
  e.ldx = ldx;
  e.ldr = h->P.n;
  e.lddx = h->P.n;
  e.n = h->P.n;
  e.nrhs = nrhs;
  e.norm0 = 0;
  return e;
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
  o->ir_max_steps = 10;
  o->ir_tol = 1e-15;
  o->fp32_gemm = SNU_GEMM_SPLIT;
}

extern "C" const char *snu_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Scratch of the allocation-free solve (snumat_solve_device_scratch): one
// fixed array per device, held from the library's load until the process
// ends.  In host code `scratch` names a host-side stand-in, not the device
// array: host code uses only the per-device address that
// hipGetSymbolAddress returns (scratch_base), never the name.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchBytes = size_t(256) << 20;  // 256 MiB per device
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_base[kMaxDevices] = {};

// Address of this device's scratch (looked up once per device, with that
// device current); the caller has made `device` current.
int scratch_base(int device, unsigned char **out) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(SNU_ERR_NO_DEVICE, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_base[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(SNU_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                        hipGetErrorString(e));
    g_scratch_base[device] = p;
  }
  *out = (unsigned char *)g_scratch_base[device];
  return SNU_OK;
}
}  // namespace

extern "C" int snu_initialize(int device) {
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0)
    return set_error(SNU_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= cnt) return set_error(SNU_ERR_NO_DEVICE, "device ordinal out of range");
  HIPCHK(hipSetDevice(device));
  unsigned char *base = nullptr;
  return scratch_base(device, &base);
}

extern "C" int snumat_construct(const snu_csr_t *A, const snu_opts_t *opts,
                                snumat_t **out) {
  if (!out) return set_error(SNU_ERR_BAD_INPUT, "null output handle");
  *out = nullptr;
  if (!opts) return set_error(SNU_ERR_BAD_INPUT, "null options");
  if (opts->fp32_gemm != SNU_GEMM_SPLIT && opts->fp32_gemm != SNU_GEMM_F32)
    return set_error(SNU_ERR_BAD_INPUT, "fp32_gemm must be SNU_GEMM_SPLIT or SNU_GEMM_F32");
  snumat *h = new snumat();
  h->opts = *opts;
  h->prec = opts->precision;
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
    if (!rc && h->prec == SNU_FP32) {
      IrDev e = ir_entry(h, nullptr, 0, nullptr, 0, 0);
      hipError_t he = hipMemcpy(h->d_ir, &e, sizeof(e), hipMemcpyHostToDevice);
      if (he != hipSuccess) rc = set_error(SNU_ERR_HIP, hipGetErrorString(he));
    }
    if (!rc) rc = ensure_stream(h);
    if (rc) {
      free_device(h);
      delete h;
      return rc;
    }
  }
  *out = h;
  return SNU_OK;
}

extern "C" int snumat_bytes_needed(const snumat_t *h, size_t *bytes) {
  if (!h || !bytes) return set_error(SNU_ERR_BAD_INPUT, "null");
  *bytes = estimate_device_bytes(h);
  return SNU_OK;
}

extern "C" int snumat_attach(snumat_t *h, int device, const double *values) {
  if (!h || !values) return set_error(SNU_ERR_BAD_INPUT, "null handle/values");
  if (h->on_device) return set_error(SNU_ERR_BAD_INPUT, "handle is already on a device");
  int rc = snu_initialize(device);
  if (rc) return rc;
  h->device = device;
  h->opts.device = device;
  h->on_device = true;
  rc = upload(h);
  if (!rc) {
    hipError_t e = hipMemcpy(h->d_aval, values, sizeof(double) * (size_t)h->P.nnz, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = set_error(SNU_ERR_HIP, hipGetErrorString(e));
  }
  if (!rc && h->prec == SNU_FP32) {
    IrDev e = ir_entry(h, nullptr, 0, nullptr, 0, 0);
    hipError_t he = hipMemcpy(h->d_ir, &e, sizeof(e), hipMemcpyHostToDevice);
    if (he != hipSuccess) rc = set_error(SNU_ERR_HIP, hipGetErrorString(he));
  }
  if (!rc) rc = ensure_stream(h);
  if (rc) {
    free_device(h);
    h->on_device = false;
    h->device = -1;
  }
  return rc;
}

// Release every device allocation of a handle and return it to the
// host-only state of a construct with device < 0 (the symbolic plan, the
// options and the graph and profiling switches kept), ready for
// snumat_attach.  The FP64 twin of an FP32 handle is destroyed.
extern "C" int snumat_detach(snumat_t *h) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null handle");
  if (!h->on_device) return SNU_OK;
  if (h->fb64) {
    snumat_destroy(h->fb64);
    h->fb64 = nullptr;
  }
  HIPCHK(hipSetDevice(h->device));
  if (h->stream) HIPCHK(hipStreamSynchronize(h->stream));
  free_device(h);
  Plan P = std::move(h->P);
  snu_opts_t o = h->opts;
  const int prec = h->prec;
  const bool prof = h->profile, graph = h->use_graph;
  h->~snumat();
  new (h) snumat();
  h->P = std::move(P);
  h->opts = o;
  h->opts.device = -1;
  h->prec = prec;
  h->profile = prof;
  h->use_graph = graph;
  return SNU_OK;
}

extern "C" int snumat_update_values(snumat_t *h, const double *values) {
  if (!h || !values) return set_error(SNU_ERR_BAD_INPUT, "null handle/values");
  if (!h->on_device) return set_error(SNU_ERR_NO_DEVICE, "handle was built host-only");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipMemcpy(h->d_aval, values, sizeof(double) * (size_t)h->P.nnz, hipMemcpyHostToDevice));
  h->factored = false;
  h->staged = false;
  return SNU_OK;
}

namespace {
template <typename T>
int factorize_values(snumat *h, int stage) {
  Values<T> &VV = vals<T>(h);
  FactorSet fs;
  fs.prep = h->d_prep;
  fs.nmat = 1;
  fs.max_len = std::max(h->arena_len, h->inv_total);
  fs.max_nnz = h->P.nnz;
  fs.ir = h->prec == SNU_FP32 ? h->d_ir : nullptr;
  fs.anorm_out = h->d_norms;
  fs.max_n = h->P.n;
  fs.wd = &h->wd;
  fs.fronts = &vals<T>(h).fronts;
  fs.eps = h->opts.eps;
  fs.split = h->opts.fp32_gemm == SNU_GEMM_SPLIT;
  fs.scatter_only = stage == 0;
  static const DensePlan<T> empty;
  double gms[5] = {0, 0, 0, 0, 0};
  int rc = run_factor<T>(fs, stage >= 2 ? VV.fact : empty, h->stream, &h->ev, h->profile && stage >= 2,
                         &h->pev, gms, h->use_graph && stage >= 2 ? &h->graph : nullptr);
  if (rc) return rc;
  HIPCHK(hipEventSynchronize(h->ev[5]));
  if (h->prec == SNU_FP32) {
    unsigned long long a = 0;
    HIPCHK(hipMemcpy(&a, h->d_norms, sizeof(a), hipMemcpyDeviceToHost));
    std::memcpy(&h->anorm, &a, sizeof(a));
  }
  for (int q = 0; q < 5; q++) h->t[q] = ev_ms(h->ev[q], h->ev[q + 1]);
  h->t[5] = ev_ms(h->ev[0], h->ev[5]);
  // t[7..9]: C -= AB GEMM launches (time when profiling, count, GFLOP);
  // t[10..12]: the same for the C = +-AB (triangular solve by inverse) ones
  for (int q = 7; q <= 12; q++) h->t[q] = 0;
  for (double &v : h->gsplit) v = 0;
  if (stage >= 2) {
    for (const Launch &L : VV.fact.launches) {
      if (!is_gemm(L.kind)) continue;
      int base = is_assign_gemm(L.kind) ? 10 : 7;
      h->t[base + 1] += 1;
      h->t[base + 2] += L.flops * 1e-9;
      if (L.kind == 2 || L.kind == 5) {
        const int q = L.kind == 5 ? 3 : 0;
        h->gsplit[q + 1] += 1;
        h->gsplit[q + 2] += L.flops * 1e-9;
      }
    }
    // the per-tile split also counts the leaf fronts' C -= AB launches (the
    // same kernel instantiations)
    for (const Launch &L : VV.fronts.dense.launches)
      if (L.kind == 2 || L.kind == 5) {
        const int q = L.kind == 5 ? 3 : 0;
        h->gsplit[q + 1] += 1;
        h->gsplit[q + 2] += L.flops * 1e-9;
      }
    if (h->profile) {
      h->t[7] = gms[0];
      h->t[10] = gms[1];
      h->gsplit[0] = gms[0] - gms[2] + gms[3];
      h->gsplit[3] = gms[2] + gms[4];
    }
  }
  h->factored = stage >= 2;
  h->staged = true;
  return SNU_OK;
}

int factorize_stage(snumat_t *h, int stage) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null handle");
  if (stage < 0 || stage > 2) return set_error(SNU_ERR_BAD_INPUT, "stage must be 0, 1 or 2");
  if (!h->on_device) return set_error(SNU_ERR_NO_DEVICE, "handle was built host-only");
  HIPCHK(hipSetDevice(h->device));
  int rc = ensure_stream(h);
  if (rc) return rc;
  h->factored = false;
  h->fb64_current = false;
  return h->prec == SNU_FP32 ? factorize_values<float>(h, stage) : factorize_values<double>(h, stage);
}

// FP64 twin of an FP32 handle (same plan and values, FP64 arena), factorized
// for the handle's current values.
int fp64_fallback(snumat *h) {
  if (!h->fb64) {
    snumat *f = new snumat();
    f->P = h->P;
    f->opts = h->opts;
    f->opts.precision = SNU_FP64;
    f->prec = SNU_FP64;
    f->device = h->device;
    f->on_device = true;
    int rc = upload(f);
    if (rc) {
      free_device(f);
      delete f;
      return rc;
    }
    h->fb64 = f;
    h->fb64_current = false;
  }
  if (!h->fb64_current) {
    HIPCHK(hipMemcpy(h->fb64->d_aval, h->d_aval, sizeof(double) * (size_t)h->P.nnz,
                     hipMemcpyDeviceToDevice));
    int rc = factorize_stage(h->fb64, 2);
    if (rc) return rc;
    h->fb64_current = true;
  }
  return SNU_OK;
}

int solve_dev(snumat *h, const double *d_b, long long ldb, double *d_x, long long ldx, int nrhs);
}  // namespace

namespace snu {
int solve_core(snumat *h, const double *d_b, long long ldb, double *d_x, long long ldx, int nrhs,
               const SolveBuffers &B, bool fallback) {
  hipStream_t st = h->stream;
  h->t[15] = 0;
  const int n = h->P.n;
  VecDev vd{0, n, nrhs};
  PermDev pd[2];
  std::memset(pd, 0, sizeof(pd));
  pd[0] = PermDev{h->d_order, d_b, d_x, ldb, ldx, n, 0};
  pd[1] = PermDev{h->d_order, B.r, B.dx, (long long)n, (long long)n, n, 0};
  HIPCHK(hipMemcpyAsync(h->d_vec, &vd, sizeof(vd), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(h->d_perm, pd, sizeof(pd), hipMemcpyHostToDevice, st));
  SolveSet ss;
  ss.perm = h->d_perm;
  ss.vec = h->d_vec;
  ss.nmat = 1;
  ss.max_elems = (long long)n * nrhs;
  ss.max_rhs = nrhs;
  ss.wd = &h->wd;
  ss.cb_ld = h->wd.cb_total;
  HIPCHK(hipEventRecord(h->ev[6], st));
  int rc = SNU_OK;
  if (h->prec == SNU_FP32) {
    std::vector<int> slot_mat(nrhs, 0);
    std::vector<double> anorm{h->anorm};
    IrDev e = ir_entry(h, d_b, ldb, d_x, ldx, nrhs);
    e.r = B.r;
    e.dx = B.dx;
    HIPCHK(hipMemcpyAsync(h->d_ir, &e, sizeof(e), hipMemcpyHostToDevice, st));
    RefineSet rs;
    rs.sweep_b = ss;
    rs.sweep_r = ss;
    rs.sweep_r.perm = h->d_perm + 1;
    rs.ir = h->d_ir;
    rs.nmat = 1;
    rs.max_n = n;
    rs.max_elems = (long long)n * nrhs;
    rs.norms = B.colnorms;
    rs.nslots = nrhs;
    rs.slot_mat = &slot_mat;
    rs.anorm = &anorm;
    rs.d_active = h->d_active;
    rs.max_steps = h->opts.ir_max_steps > 0 ? h->opts.ir_max_steps : h->opts.ir_max_steps < 0 ? 0 : 10;
    rs.tol = h->opts.ir_tol > 0 ? h->opts.ir_tol : 1e-15;
    std::vector<double> berr;
    rc = run_refined(rs, (float *)B.bp, (float *)B.cbuf, *(const GemvPlan<float> *)B.plan, st, &h->ir_steps,
                     &berr);
    h->ir_berr = berr.empty() ? 0 : berr[0];
    if (rc == SNU_ERR_NOT_CONVERGED && fallback) {
      // refinement stalled above 1e-12: solve again with FP64 factors of
      // the same matrix (t[15] = 1 reports it)
      if ((rc = fp64_fallback(h))) return rc;
      if ((rc = solve_dev(h->fb64, d_b, ldb, d_x, ldx, nrhs))) return rc;
      HIPCHK(hipSetDevice(h->device));
      h->t[15] = 1;
    }
    if (rc && rc != SNU_ERR_NOT_CONVERGED) return rc;
  } else {
    run_sweep<double>(ss, (double *)B.bp, (double *)B.cbuf, *(const GemvPlan<double> *)B.plan, st);
    h->ir_steps = 0;
    h->ir_berr = 0;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(h->ev[7], st));
  HIPCHK(hipEventSynchronize(h->ev[7]));
  h->t[6] = ev_ms(h->ev[6], h->ev[7]);
  h->t[13] = h->ir_steps;
  h->t[14] = h->ir_berr;
  return rc;
}
}  // namespace snu

namespace {
int solve_dev(snumat *h, const double *d_b, long long ldb, double *d_x, long long ldx, int nrhs) {
  int rc = ensure_stream(h);
  if (rc) return rc;
  if ((rc = ensure_rhs(h, nrhs))) return rc;
  SolveBuffers B;
  if (h->prec == SNU_FP32) {
    if ((rc = ensure_plan<float>(h, nrhs))) return rc;
    B.bp = h->v32.bp;
    B.cbuf = h->v32.cbuf;
    B.plan = &h->v32.sp;
  } else {
    if ((rc = ensure_plan<double>(h, nrhs))) return rc;
    B.bp = h->v64.bp;
    B.cbuf = h->v64.cbuf;
    B.plan = &h->v64.sp;
  }
  B.r = h->d_r;
  B.dx = h->d_dx;
  B.colnorms = h->d_colnorms;
  return solve_core(h, d_b, ldb, d_x, ldx, nrhs, B, true);
}
}  // namespace

namespace snu {
int fp64_fallback_solve(snumat *h, const double *d_b, long long ldb, double *d_x, long long ldx, int nrhs) {
  HIPCHK(hipSetDevice(h->device));
  int rc = fp64_fallback(h);
  if (rc) return rc;
  return solve_dev(h->fb64, d_b, ldb, d_x, ldx, nrhs);
}
}  // namespace snu

extern "C" int snumat_factorize(snumat_t *h) { return factorize_stage(h, 2); }
extern "C" int snumat_factorize_stage(snumat_t *h, int stage) {
  return factorize_stage(h, stage);
}

extern "C" int snumat_solve_device(snumat_t *h, const double *d_b, double *d_x,
                                   int nrhs, int ld) {
  if (!h || !d_b || !d_x || nrhs < 1) return set_error(SNU_ERR_BAD_INPUT, "bad solve arguments");
  if (!h->factored) return set_error(SNU_ERR_NOT_FACTORED, "solve before factorize");
  if (ld < h->P.n) return set_error(SNU_ERR_BAD_INPUT, "ld < n");
  if (d_b == d_x) return set_error(SNU_ERR_BAD_INPUT, "b and x must not alias");
  HIPCHK(hipSetDevice(h->device));
  return solve_dev(h, d_b, ld, d_x, ld, nrhs);
}

namespace {
// The solve of snumat_solve_device over buffers carved from the device's
// scratch: permuted right-hand sides, dense-solve vectors, boundary
// contributions, (SNU_FP32) residual, correction and column norms, then the
// dense-solve descriptors and split-K partial sums.  Every buffer is written
// before it is read within the call (the column norms are zeroed by their
// own kernel), so no clearing is needed; the descriptors are copied in on
// every call, as another handle's call may have overwritten them.
template <typename T>
int solve_scratch(snumat *h, const double *d_b, double *d_x, int nrhs, int ld) {
  unsigned char *base = nullptr;
  int rc = scratch_base(h->device, &base);
  if (rc) return rc;
  const size_t n = (size_t)h->P.n;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  const size_t o_bp = carve(sizeof(T) * n * nrhs);
  const size_t o_ys = carve(sizeof(T) * n * nrhs);
  const size_t o_cb = carve(sizeof(T) * (size_t)std::max(1, h->wd.cb_total) * nrhs);
  size_t o_r = 0, o_dx = 0, o_cn = 0;
  if (h->prec == SNU_FP32) {
    o_r = carve(sizeof(double) * n * nrhs);
    o_dx = carve(sizeof(double) * n * nrhs);
    o_cn = carve(sizeof(double) * 3 * (size_t)nrhs);
  }
  Values<T> &VV = vals<T>(h);
  GemvPlan<T> &g = VV.sp_scratch;
  if (VV.scratch_nrhs != nrhs) {
    VV.scratch_nrhs = -1;
    g = GemvPlan<T>();  // device pointers point into the scratch: nothing to free
    build_solve_plan<T>(h, (T *)(base + o_bp), (T *)(base + o_ys), nrhs, g);
  }
  const size_t need_part = layout_gemv_plan(g);
  const size_t o_vp = carve(sizeof(GemvProbT<T>) * std::max<size_t>(1, g.vp.size()));
  const size_t o_vs = carve(sizeof(GemvSegT<T>) * std::max<size_t>(1, g.vs.size()));
  const size_t o_part = carve(sizeof(T) * std::max<size_t>(1, need_part));
  if (off > kScratchBytes)
    return set_error(SNU_ERR_BAD_INPUT, "allocation-free solve needs " + std::to_string(off) +
                                         " bytes of scratch for n = " + std::to_string(n) +
                                         ", nrhs = " + std::to_string(nrhs) + "; the limit is " +
                                         std::to_string(kScratchBytes) +
                                         " (use snumat_solve_device, or fewer columns per call)");
  VV.scratch_nrhs = nrhs;
  g.d_vp = (GemvProbT<T> *)(base + o_vp);
  g.d_vs = (GemvSegT<T> *)(base + o_vs);
  g.d_part = (T *)(base + o_part);
  g.cap_vp = g.vp.size();
  g.cap_vs = g.vs.size();
  g.cap_part = need_part;
  if (!g.vp.empty())
    HIPCHK(hipMemcpyAsync(g.d_vp, g.vp.data(), g.vp.size() * sizeof(g.vp[0]), hipMemcpyHostToDevice,
                          h->stream));
  if (!g.vs.empty())
    HIPCHK(hipMemcpyAsync(g.d_vs, g.vs.data(), g.vs.size() * sizeof(g.vs[0]), hipMemcpyHostToDevice,
                          h->stream));
  SolveBuffers B;
  B.bp = base + o_bp;
  B.cbuf = base + o_cb;
  B.plan = &g;
  if (h->prec == SNU_FP32) {
    B.r = (double *)(base + o_r);
    B.dx = (double *)(base + o_dx);
    B.colnorms = (unsigned long long *)(base + o_cn);
  }
  // no FP64 twin here: building one allocates (a stalled refinement
  // returns SNU_ERR_NOT_CONVERGED with x = the last iterate)
  return solve_core(h, d_b, ld, d_x, ld, nrhs, B, false);
}
}  // namespace

extern "C" int snumat_solve_device_scratch(snumat_t *h, const double *d_b, double *d_x,
                                           int nrhs, int ld) {
  if (!h || !d_b || !d_x || nrhs < 1) return set_error(SNU_ERR_BAD_INPUT, "bad solve arguments");
  if (!h->on_device) return set_error(SNU_ERR_NO_DEVICE, "handle was built host-only");
  if (!h->factored) return set_error(SNU_ERR_NOT_FACTORED, "solve before factorize");
  if (ld < h->P.n) return set_error(SNU_ERR_BAD_INPUT, "ld < n");
  if (d_b == d_x) return set_error(SNU_ERR_BAD_INPUT, "b and x must not alias");
  if (!h->stream || h->ev.size() < 8)
    return set_error(SNU_ERR_BAD_INPUT, "handle has no stream (construct it on a device)");
  HIPCHK(hipSetDevice(h->device));
  return h->prec == SNU_FP32 ? solve_scratch<float>(h, d_b, d_x, nrhs, ld)
                             : solve_scratch<double>(h, d_b, d_x, nrhs, ld);
}

extern "C" int snumat_solve(snumat_t *h, const double *b, double *x, int nrhs,
                            int ld) {
  if (!h || !b || !x || nrhs < 1) return set_error(SNU_ERR_BAD_INPUT, "bad solve arguments");
  if (!h->factored) return set_error(SNU_ERR_NOT_FACTORED, "solve before factorize");
  const int n = h->P.n;
  if (ld < n) return set_error(SNU_ERR_BAD_INPUT, "ld < n");
  HIPCHK(hipSetDevice(h->device));
  int rc = ensure_stream(h);
  if (rc) return rc;
  if ((rc = ensure_rhs(h, nrhs))) return rc;
  double *db = h->d_bx, *dx = h->d_bx + (size_t)n * nrhs;
  HIPCHK(hipMemcpy2DAsync(db, sizeof(double) * n, b, sizeof(double) * ld, sizeof(double) * n, nrhs,
                          hipMemcpyHostToDevice, h->stream));
  rc = solve_dev(h, db, n, dx, n, nrhs);
  if (rc && rc != SNU_ERR_NOT_CONVERGED) return rc;
  const int solve_rc = rc;
  HIPCHK(hipMemcpy2DAsync(x, sizeof(double) * ld, dx, sizeof(double) * n, sizeof(double) * n, nrhs,
                          hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  // report Inf/NaN in the solution as a numeric failure (zero pivot in a
  // dense block: the reference ignores getrf's info, src/kernel_cuda.cu:67)
  for (int j = 0; j < nrhs; j++)
    for (int i = 0; i < n; i++)
      if (!std::isfinite(x[(size_t)j * ld + i]))
        return set_error(SNU_ERR_NUMERIC, "non-finite solution (zero pivot in the factors?)");
  return solve_rc;
}

extern "C" int snumat_destroy(snumat_t *h) {
  if (!h) return SNU_OK;
  if (h->fb64) {
    snumat_destroy(h->fb64);
    h->fb64 = nullptr;
  }
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
  if (!A || !b || !x) return set_error(SNU_ERR_BAD_INPUT, "null matrix or vectors");
  if (o.device < 0) return set_error(SNU_ERR_BAD_INPUT, "snu_solve_csr needs a device");
  int rc = snu_initialize(o.device);
  if (rc) return rc;
  snumat_t *h = nullptr;
  if ((rc = snumat_construct(A, &o, &h))) return rc;
  rc = snumat_factorize(h);
  if (!rc) rc = snumat_solve(h, b, x, 1, A->n);
  snumat_destroy(h);
  return rc;
}

// ---------------- introspection ----------------
// count values of the factor arena from offset off, widened to FP64
static int read_values(const snumat_t *h, long long off, size_t count, double *out) {
  if (h->prec == SNU_FP32) {
    std::vector<float> tmp(count);
    HIPCHK(hipMemcpy(tmp.data(), h->v32.arena + off, sizeof(float) * count, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < count; i++) out[i] = tmp[i];
  } else {
    HIPCHK(hipMemcpy(out, h->v64.arena + off, sizeof(double) * count, hipMemcpyDeviceToHost));
  }
  return SNU_OK;
}
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
  int64_t nfr = 0;
  for (const LeafPlan &L : P.leaves) nfr += (int64_t)L.ff.size();
  info[14] = nfr;
  info[15] = h->wd.ntarget;
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
extern "C" int snumat_leaf_front_sizes(const snumat_t *h, int rank, int *s4) {
  if (!h || !s4 || rank < 0 || rank >= h->P.np) return set_error(SNU_ERR_BAD_INPUT, "bad rank");
  const LeafPlan &L = h->P.leaves[rank];
  s4[0] = (int)L.ff.size();
  s4[1] = (int)L.fidx.size();
  s4[2] = (int)L.fslot.size();
  s4[3] = (int)L.fchild.size();
  return SNU_OK;
}
extern "C" int snumat_leaf_fronts(const snumat_t *h, int rank, int *front5, int *idx_ptr, int *idx,
                                  int *rel, int *slot_ptr, int *slot, int *pos, int *child_ptr,
                                  int *child, int *srow, int *scol) {
  if (!h || rank < 0 || rank >= h->P.np) return set_error(SNU_ERR_BAD_INPUT, "bad rank");
  const LeafPlan &L = h->P.leaves[rank];
  const size_t nf = L.ff.size();
  for (size_t f = 0; f < nf; f++) {
    front5[5 * f + 0] = L.fk[L.fkp[f]];  // first pivot
    front5[5 * f + 1] = L.fnk[f];
    front5[5 * f + 2] = L.ff[f];
    front5[5 * f + 3] = L.fpar[f];
    front5[5 * f + 4] = L.fheight[f];
  }
  std::copy(L.fidx_ptr.begin(), L.fidx_ptr.end(), idx_ptr);
  std::copy(L.fidx.begin(), L.fidx.end(), idx);
  std::copy(L.frel.begin(), L.frel.end(), rel);
  std::copy(L.fslot_ptr.begin(), L.fslot_ptr.end(), slot_ptr);
  std::copy(L.fslot.begin(), L.fslot.end(), slot);
  std::copy(L.fpos.begin(), L.fpos.end(), pos);
  std::copy(L.fchild_ptr.begin(), L.fchild_ptr.end(), child_ptr);
  std::copy(L.fchild.begin(), L.fchild.end(), child);
  std::copy(L.fsrow.begin(), L.fsrow.end(), srow);
  std::copy(L.fscol.begin(), L.fscol.end(), scol);
  return SNU_OK;
}
extern "C" int snumat_leaf_front_pivots(const snumat_t *h, int rank, int *k_ptr, int *k) {
  if (!h || !k_ptr || !k || rank < 0 || rank >= h->P.np) return set_error(SNU_ERR_BAD_INPUT, "bad rank");
  const LeafPlan &L = h->P.leaves[rank];
  std::copy(L.fkp.begin(), L.fkp.end(), k_ptr);
  std::copy(L.fk.begin(), L.fk.end(), k);
  return SNU_OK;
}
extern "C" int snumat_leaf_values(const snumat_t *h, int rank, double *out) {
  if (!h || rank < 0 || rank >= h->P.np) return set_error(SNU_ERR_BAD_INPUT, "bad rank");
  if (!h->factored && !h->staged) return set_error(SNU_ERR_NOT_FACTORED, "not factored");
  const LeafPlan &L = h->P.leaves[rank];
  return read_values(h, h->P.leaf_off[rank], (size_t)L.lu_nnz, out);
}
extern "C" int snumat_sep_diag(const snumat_t *h, int e, double *out) {
  if (!h || e < 1 || e >= h->P.np) return set_error(SNU_ERR_BAD_INPUT, "bad separator");
  if (!h->factored) return set_error(SNU_ERR_NOT_FACTORED, "not factored");
  int s = h->P.block_size[e];
  return read_values(h, h->P.leaf_total + h->P.offR[e], (size_t)s * s, out);
}
extern "C" int snumat_sep_block(const snumat_t *h, int i, int j, double *out) {
  if (!h || !out) return set_error(SNU_ERR_BAD_INPUT, "null handle or output");
  const Plan &P = h->P;
  if (i < 1 || j < 1 || i >= P.np || j >= P.np) return set_error(SNU_ERR_BAD_INPUT, "bad block");
  if (!h->on_device) return set_error(SNU_ERR_NO_DEVICE, "handle was built host-only");
  int x = std::max(i, j), y = std::min(i, j);
  while (x > y) x /= 2;
  if (x != y) return set_error(SNU_ERR_BAD_INPUT, "blocks not on one root path");
  const int si = P.block_size[i], sj = P.block_size[j];
  if (!si || !sj) return SNU_OK;
  // Rows (U panel) or columns (L panel) of the block that fall in panel tiles
  // not stored are zero.  The stored ones come in runs, each a strided view
  // of the arena: one 2-D copy per run.
  std::fill(out, out + (size_t)si * sj, 0.0);
  HIPCHK(hipSetDevice(h->device));
  auto copy = [&](long long src, long long ld, int rows, int cols, double *dst) -> int {
    if (h->prec == SNU_FP32) {
      std::vector<float> tmp((size_t)rows * cols);
      HIPCHK(hipMemcpy2D(tmp.data(), sizeof(float) * rows, h->v32.arena + P.leaf_total + src, sizeof(float) * ld,
                         sizeof(float) * rows, cols, hipMemcpyDeviceToHost));
      for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++) dst[r + (size_t)c * si] = tmp[r + (size_t)c * rows];
    } else {
      HIPCHK(hipMemcpy2D(dst, sizeof(double) * si, h->v64.arena + P.leaf_total + src, sizeof(double) * ld,
                         sizeof(double) * rows, cols, hipMemcpyDeviceToHost));
    }
    return SNU_OK;
  };
  int rc;
  if (i == j) return copy(P.offR[i], si, si, si, out);
  if (i > j) {  // columns of the U panel of i: runs of stored columns
    const int o = P.anc_off(i, j);
    for (int c = 0; c < sj;) {
      if (P.umap[i][o + c] < 0) { c++; continue; }
      int e = c + 1;
      while (e < sj && P.umap[i][o + e] == P.umap[i][o + c] + (e - c)) e++;
      if ((rc = copy(P.sep_index(i, 0, j, c), si, si, e - c, out + (size_t)c * si))) return rc;
      c = e;
    }
    return SNU_OK;
  }
  // rows of the L panel of j: runs of stored rows
  const int o = P.anc_off(j, i);
  for (int r = 0; r < si;) {
    if (P.lmap[j][o + r] < 0) { r++; continue; }
    int e = r + 1;
    while (e < si && P.lmap[j][o + e] == P.lmap[j][o + r] + (e - r)) e++;
    if ((rc = copy(P.sep_index(i, r, j, 0), P.nL[j], e - r, sj, out + r))) return rc;
    r = e;
  }
  return SNU_OK;
}
extern "C" int snumat_timings(const snumat_t *h, double *t16) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null");
  std::memcpy(t16, h->t, sizeof(h->t));
  return SNU_OK;
}
extern "C" int snumat_gemm_split(const snumat_t *h, double *out6) {
  if (!h || !out6) return set_error(SNU_ERR_BAD_INPUT, "null");
  std::memcpy(out6, h->gsplit, sizeof(h->gsplit));
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

namespace {
// Work the launch plans execute (zero tiles skipped), as opposed to the
// nominal dense counts of snumat_counts.
template <typename T>
void executed_work(const snumat *h, double *w) {
  const Values<T> &VV = vals<T>(h);
  double fl = 0, nl = 0, nd = 0;
  for (const Launch &L : VV.fact.launches) {
    nl += 1;
    if (L.kind == 0) {
      nd += 1;
      for (int q = 0; q < L.np; q++) {
        const double nb = VV.fact.dp[L.p0 + q].nb;
        fl += 4.0 / 3.0 * nb * nb * nb;  // LU (2/3 nb^3) + both triangular inverses
      }
    } else if (is_gemm(L.kind)) {
      fl += L.flops;
    }
  }
  GemvPlan<T> sp;
  build_solve_plan<T>(h, nullptr, nullptr, 1, sp);
  double gb = 0;
  for (const std::vector<Launch> *ls : {&sp.fwd, &sp.bwd})
    for (const Launch &L : *ls) gb += L.flops;  // matrix bytes read by the launch
  double lb = 0;
  for (const LeafPlan &L : h->P.leaves) lb += (double)(sizeof(T) + 4) * L.lu_nnz;
  w[0] = fl;
  w[1] = nl;
  w[2] = nd;
  w[3] = gb + lb + 2.0 * sizeof(T) * h->P.n;
  w[4] = gb;
  w[5] = lb;
  w[6] = (double)(sp.fwd.size() + sp.bwd.size());
  w[7] = 0;
}
}  // namespace

extern "C" int snumat_work(const snumat_t *h, double *w8) {
  if (!h || !w8) return set_error(SNU_ERR_BAD_INPUT, "null");
  if (!h->on_device) return set_error(SNU_ERR_NO_DEVICE, "handle was built host-only");
  if (h->prec == SNU_FP32) executed_work<float>(h, w8);
  else executed_work<double>(h, w8);
  return SNU_OK;
}

// ---------------- kernel-level test hooks (host pointers) ----------------
namespace {
template <typename T>
int test_gemm(int M, int N, int K, const T *A, int lda, const T *B, int ldb, T *C, int ldc,
              int assign) {
  if (M < 0 || N < 0 || K < 0 || lda < std::max(1, M) || ldb < std::max(1, K) || ldc < std::max(1, M) ||
      assign < 0 || assign > (std::is_same<T, float>::value ? 11 : 5))
    return set_error(SNU_ERR_BAD_INPUT, "bad gemm shape");
  T *dA, *dB, *dC;
  size_t na = (size_t)lda * std::max(K, 1), nb = (size_t)ldb * std::max(N, 1), nc = (size_t)ldc * std::max(N, 1);
  HIPCHK(hipMalloc((void **)&dA, na * sizeof(T)));
  HIPCHK(hipMalloc((void **)&dB, nb * sizeof(T)));
  HIPCHK(hipMalloc((void **)&dC, nc * sizeof(T)));
  HIPCHK(hipMemcpy(dA, A, na * sizeof(T), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dB, B, nb * sizeof(T), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dC, C, nc * sizeof(T), hipMemcpyHostToDevice));
  std::vector<GemmProbT<T>> gp;
  std::vector<GemmSegT<T>> gs;
  int tiles = 0;
  double fl = 0;
  add_gemm<T>(gp, gs, tiles, dC, ldc, M, N, {GemmSegT<T>{dA, dB, lda, ldb, K}}, fl,
              assign == 2 || (assign >= 4 && assign <= 8) ? 128 : 64);
  GemmProbT<T> *dgp;
  GemmSegT<T> *dgs;
  HIPCHK(hipMalloc((void **)&dgp, sizeof(GemmProbT<T>) * std::max<size_t>(1, gp.size())));
  HIPCHK(hipMalloc((void **)&dgs, sizeof(GemmSegT<T>) * std::max<size_t>(1, gs.size())));
  if (!gp.empty()) {
    HIPCHK(hipMemcpy(dgp, gp.data(), sizeof(GemmProbT<T>) * gp.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dgs, gs.data(), sizeof(GemmSegT<T>) * gs.size(), hipMemcpyHostToDevice));
    if (assign >= 6) {  // FP32 only (checked above)
      if constexpr (std::is_same<T, float>::value) {
        if (assign <= 8) launch_gemm128_split(dgp, (int)gp.size(), dgs, tiles, assign - 6, nullptr);
        else launch_gemm_split(dgp, (int)gp.size(), dgs, tiles, assign - 9, nullptr);
      }
    } else if (assign == 2 || assign >= 4)
      launch_gemm128<T>(dgp, (int)gp.size(), dgs, tiles, assign == 4 ? 1 : assign == 5 ? 2 : 0, nullptr);
    else launch_gemm<T>(dgp, (int)gp.size(), dgs, tiles, assign == 1 ? 1 : assign == 3 ? 2 : 0, nullptr);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(C, dC, nc * sizeof(T), hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  (void)hipFree(dgp);
  (void)hipFree(dgs);
  return SNU_OK;
}
template <typename T>
int test_diag_lu(int nb, T *D, int ld, T *Linv, T *Uinv) {
  if (nb < 1 || nb > 64 || ld < nb) return set_error(SNU_ERR_BAD_INPUT, "bad block");
  T *dD, *dI;
  HIPCHK(hipMalloc((void **)&dD, sizeof(T) * ld * nb));
  HIPCHK(hipMalloc((void **)&dI, sizeof(T) * 8192));
  HIPCHK(hipMemcpy(dD, D, sizeof(T) * ld * nb, hipMemcpyHostToDevice));
  DiagProbT<T> p{dD, ld, nb, dI, dI + 4096, 64};
  DiagProbT<T> *dp;
  HIPCHK(hipMalloc((void **)&dp, sizeof(p)));
  HIPCHK(hipMemcpy(dp, &p, sizeof(p), hipMemcpyHostToDevice));
  launch_diag_lu<T>(dp, 1, nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(D, dD, sizeof(T) * ld * nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(Linv, dI, sizeof(T) * 4096, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(Uinv, dI + 4096, sizeof(T) * 4096, hipMemcpyDeviceToHost));
  (void)hipFree(dD);
  (void)hipFree(dI);
  (void)hipFree(dp);
  return SNU_OK;
}
}  // namespace

// C = C - A B (assign = 0), C = A B (1), C -= A B with 128 x 128 tiles (2),
// C = -A B (3), C = A B with 128 x 128 tiles (4), C = -A B with 128 x 128
// tiles (5), one segment, through the grouped MFMA kernel; col-major.  FP32
// also: C -= A B (6), C = A B (7), C = -A B (8) through the bf16-split
// 128 x 128 kernel, and (9, 10, 11) through its 64 x 64 form.
extern "C" int snu_test_gemm(int M, int N, int K, const double *A, int lda,
                             const double *B, int ldb, double *C, int ldc,
                             int assign) {
  return test_gemm<double>(M, N, K, A, lda, B, ldb, C, ldc, assign);
}
extern "C" int snu_test_gemm_f32(int M, int N, int K, const float *A, int lda,
                                 const float *B, int ldb, float *C, int ldc,
                                 int assign) {
  return test_gemm<float>(M, N, K, A, lda, B, ldb, C, ldc, assign);
}

// In-place no-pivot LU of an nb x nb block (nb <= 64) and the inverses of its
// triangular factors (64 x 64, ld 64), through k_diag_lu.
extern "C" int snu_test_diag_lu(int nb, double *D, int ld, double *Linv,
                                double *Uinv) {
  return test_diag_lu<double>(nb, D, ld, Linv, Uinv);
}
extern "C" int snu_test_diag_lu_f32(int nb, float *D, int ld, float *Linv,
                                    float *Uinv) {
  return test_diag_lu<float>(nb, D, ld, Linv, Uinv);
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
extern "C" int snumat_use_graph(snumat_t *h, int on) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null");
  h->use_graph = on != 0;
  return SNU_OK;
}
extern "C" int snumat_set_profiling(snumat_t *h, int on) {
  if (!h) return set_error(SNU_ERR_BAD_INPUT, "null");
  h->profile = on != 0;
  return SNU_OK;
}

// Microbenchmark of the grouped GEMM kernels on device-resident operands:
// C (M x N) -= A (M x K) B (K x N) split into nseg equal K segments, tile 64
// or 128, FP64 (prec SNU_FP64) or FP32 (SNU_FP32); returns the average kernel
// time over reps launches (HIP events).
namespace {
template <typename T>
int bench_gemm(int M, int N, int K, int nseg, int tile, int kb, int depth, bool split, int reps,
               double *ms_out) {
  T *dA, *dB, *dC;
  HIPCHK(hipMalloc((void **)&dA, sizeof(T) * (size_t)M * K));
  HIPCHK(hipMalloc((void **)&dB, sizeof(T) * (size_t)K * N));
  HIPCHK(hipMalloc((void **)&dC, sizeof(T) * (size_t)M * N));
  launch_fill(dA, (long long)M * K, nullptr);
  launch_fill(dB, (long long)K * N, nullptr);
  launch_fill(dC, (long long)M * N, nullptr);
  std::vector<GemmSegT<T>> segs;
  int kseg = (K + nseg - 1) / nseg;
  for (int k0 = 0; k0 < K; k0 += kseg)
    segs.push_back(GemmSegT<T>{dA + (size_t)k0 * M, dB + k0, M, K, std::min(kseg, K - k0)});
  std::vector<GemmProbT<T>> gp;
  std::vector<GemmSegT<T>> gs;
  int tiles = 0;
  double fl = 0;
  add_gemm<T>(gp, gs, tiles, dC, M, M, N, segs, fl, tile);
  GemmProbT<T> *dgp;
  GemmSegT<T> *dgs;
  HIPCHK(hipMalloc((void **)&dgp, sizeof(GemmProbT<T>) * gp.size()));
  HIPCHK(hipMalloc((void **)&dgs, sizeof(GemmSegT<T>) * gs.size()));
  HIPCHK(hipMemcpy(dgp, gp.data(), sizeof(GemmProbT<T>) * gp.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dgs, gs.data(), sizeof(GemmSegT<T>) * gs.size(), hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto run = [&]() {
    if constexpr (std::is_same<T, float>::value) {
      if (split) {
        if (tile == 128) launch_gemm128_split(dgp, 1, dgs, tiles, 0, nullptr);
        else launch_gemm_split(dgp, 1, dgs, tiles, 0, nullptr);
        return;
      }
    }
    launch_gemm_variant<T>(dgp, 1, dgs, tiles, tile, kb, depth, nullptr);
  };
  run();
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; r++) run();
  HIPCHK(hipEventRecord(e1, nullptr));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_out = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  (void)hipFree(dgp);
  (void)hipFree(dgs);
  return SNU_OK;
}
}  // namespace
extern "C" int snu_bench_gemm(int M, int N, int K, int nseg, int tile, int prec, int reps,
                              double *ms_out) {
  // tile: 64 or 128 (K chunk 16), 6432 / 12832 (K chunk 32), 642 / 1282
  // (K chunk 16, loads two chunks ahead), 1283 / 643 (FP32: through bf16
  // splits)
  int t = tile, kb = 16, depth = 1;
  const bool split = tile == 1283 || tile == 643;
  if (tile == 6432 || tile == 12832) t = tile / 100, kb = 32;
  if (tile == 642 || tile == 1282) t = tile / 10, depth = 2;
  if (split) t = tile / 10;
  if (M < 1 || N < 1 || K < 1 || nseg < 1 || reps < 1 || (t != 64 && t != 128) ||
      (prec != SNU_FP64 && prec != SNU_FP32) || (split && prec != SNU_FP32) || !ms_out)
    return set_error(SNU_ERR_BAD_INPUT, "bad gemm bench shape");
  return prec == SNU_FP32 ? bench_gemm<float>(M, N, K, nseg, t, kb, depth, split, reps, ms_out)
                          : bench_gemm<double>(M, N, K, nseg, t, kb, depth, false, reps, ms_out);
}

// Average time (ms) of a plain device copy of `bytes` (multiple of 16) over
// reps launches: the known-size kernel against which HBM byte counters are
// scaled.
extern "C" int snu_bench_copy(size_t bytes, int reps, double *ms_out) {
  if (bytes < 16 || bytes % 16 || reps < 1 || !ms_out) return set_error(SNU_ERR_BAD_INPUT, "bad copy bench");
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, bytes));
  HIPCHK(hipMalloc(&b, bytes));
  launch_copy(a, b, (long long)bytes, nullptr);
  HIPCHK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; r++) launch_copy(a, b, (long long)bytes, nullptr);
  HIPCHK(hipEventRecord(e1, nullptr));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_out = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return SNU_OK;
}

// Tile table of the grouped C -= AB GEMM per size bucket of min(M, N)
// ([128, 512), [512, 2048), >= 2048; below 128 always 64), as measured by the
// autotuner (run now if it has not run in this process).
extern "C" int snu_gemm_tiles(int device, int precision, int *tiles3) {
  if (!tiles3 || (precision != SNU_FP64 && precision != SNU_FP32))
    return set_error(SNU_ERR_BAD_INPUT, "bad arguments");
  int rc = snu_initialize(device);
  if (rc) return rc;
  if (precision == SNU_FP32) autotune_tiles<float>();
  else autotune_tiles<double>();
  tile_table(precision, tiles3);
  return SNU_OK;
}
