// Host-side engine of the MI355X SnuSolver path: device state of one matrix
// (struct snumat), the launch lists of its factorization and solves, and the
// executors that run a set of matrices (one, or a batch) with one launch per
// step across all of them.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>
#include <vector>

#include "snu_internal.h"
#include "snu_kernels.h"

#define HIPCHK(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess)                                                  \
      return snu::set_error(SNU_ERR_HIP, std::string("HIP error ") +       \
                                             hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace snu {

// One grouped launch.  key orders the launches of one matrix; launches of
// several matrices with equal (key, kind) are independent and are merged
// into one launch.
struct Launch {
  int kind;  // 0 diag LU; GEMM (64 tile) 1: C = AB, 2: C -= AB, 6: C = -AB;
             // GEMM (128 tile) 5: C -= AB, 10: C = AB, 11: C = -AB; 3: copy
             // C = A (GEMM descriptors, 64 x 64 tiles); GEMV 7: y = Ax, 8: y -= Ax
  int p0, np, tiles;
  double flops;
  long long key;
  int wgs = 0;          // GEMV: workgroups of the split-K partial pass (0: unsplit)
  long long part = 0;   // GEMV: partial-sum values of the launch
};
inline bool is_gemm(int kind) {
  return kind == 1 || kind == 2 || kind == 5 || kind == 6 || kind == 10 || kind == 11;
}
// C = +-AB launches (products by triangular inverses) as opposed to C -= AB
inline bool is_assign_gemm(int kind) { return kind == 1 || kind == 6 || kind == 10 || kind == 11; }

// Launch list with its descriptors (device copies once uploaded).
template <typename T>
struct DensePlan {
  std::vector<DiagProbT<T>> dp;
  std::vector<GemmProbT<T>> gp;
  std::vector<GemmSegT<T>> gs;
  std::vector<Launch> launches;
  DiagProbT<T> *d_dp = nullptr;
  GemmProbT<T> *d_gp = nullptr;
  GemmSegT<T> *d_gs = nullptr;
};
template <typename T>
struct GemvPlan {
  std::vector<GemvProbT<T>> vp;
  std::vector<GemvSegT<T>> vs;
  std::vector<Launch> fwd, bwd;
  GemvProbT<T> *d_vp = nullptr;
  GemvSegT<T> *d_vs = nullptr;
  T *d_part = nullptr;            // split-K partial sums (launches run in turn)
  size_t cap_vp = 0, cap_vs = 0, cap_part = 0;  // device capacities (grow-only)
};

// Multifrontal leaf factorization of a launch set (one matrix, or a batch
// merged): per height of the leaves' front trees, the fronts the LDS kernel
// takes whole and the larger fronts' 64-column work items (front index into
// `large`); the larger fronts' blocked partial LU as a keyed launch list
// (key: height, step within it); the tiles of the boundary Schur blocks.
struct FrontLevel {
  int small0 = 0, nsmall = 0, max_f = 0;
  int nsmall_lo = 0, max_f_lo = 0;  // the first nsmall_lo small fronts have order <= kFrontLo
  int work0 = 0, nwork = 0;
};
template <typename T>
struct FrontPlan {
  std::vector<FrontDev> small, large;
  std::vector<int4> work;   // larger fronts: {front, first column, ranges offset, 0}
  std::vector<int> wrange;  // per work item: its owned-slot range, then per child its column range
  std::vector<FrontLevel> levels;
  DensePlan<T> dense;
  std::vector<LeafSchurDev> ls;
  std::vector<int4> tiles;
  FrontDev *d_small = nullptr, *d_large = nullptr;
  int4 *d_work = nullptr;
  int *d_wrange = nullptr;
  LeafSchurDev *d_ls = nullptr;
  int4 *d_tiles = nullptr;
};
inline long long front_key(int height, int step, int kind) {
  return (long long)height * 100000000LL + step * 8LL + kind;
}
inline int front_key_height(long long key) { return (int)(key / 100000000LL); }

// Value-typed device state of one matrix (FP64, or FP32 refined in FP64).
template <typename T>
struct Values {
  T *arena = nullptr;  // leaf LU values, then dense blocks
  T *S = nullptr;      // boundary Schur blocks
  T *inv = nullptr;    // per separator: full L^-1 then U^-1 (s x s each)
  T *tmp = nullptr;    // products of the triangular inversion
  T *farena = nullptr; // leaf fronts (each f x f; larger fronts then 2 x 64 x 64 of inverses)
  FrontPlan<T> fronts;
  DensePlan<T> fact;
  // standalone solves
  T *bp = nullptr, *ys = nullptr, *cbuf = nullptr;
  int plan_nrhs = -1;
  GemvPlan<T> sp;
  // allocation-free solves: the same plan over buffers in the library's
  // static per-device scratch (its device pointers point into the scratch
  // and are never freed)
  GemvPlan<T> sp_scratch;
  int scratch_nrhs = -1;
};

// Host copies of the per-leaf work lists (kept for batching).
struct WorkLists {
  std::vector<int> fwg_leaf, fwg_first;  // forward boundary-row workgroups
  std::vector<int> trow_blk, trow_loc, trow_glob, tptr, tsrc_leaf, tsrc_ri;  // targets
  std::vector<int> tile_blk, tile_row0, tile_ptr, seg_leaf, seg_ri;          // assembly tiles
};

// Device copies of the work lists of a launch set (one matrix or a batch).
struct WorkDev {
  LeafDev *leaves = nullptr;
  int nleaves = 0;
  int *fwg_leaf = nullptr, *fwg_first = nullptr, nfwg = 0;
  int *trow_blk = nullptr, *trow_loc = nullptr, *trow_glob = nullptr, *tptr = nullptr;
  int *tsrc_leaf = nullptr, *tsrc_ri = nullptr, ntarget = 0;
  int *tile_blk = nullptr, *tile_row0 = nullptr, *tile_ptr = nullptr;
  int *seg_leaf = nullptr, *seg_ri = nullptr, ntile = 0;
  int cb_total = 0;
  int max_leaf = 0;  // rows of the largest leaf
  int max_lev = 0;   // most level sets of a leaf (forward or backward)
};

// A factorization's launch sequence captured once into a HIP graph and
// replayed (no memset nodes: buffers are cleared by kernels).
struct FactorGraph {
  hipGraphExec_t exec = nullptr;
};

}  // namespace snu

struct snumat {
  snu::Plan P;
  snu_opts_t opts;
  int prec = SNU_FP64;
  bool on_device = false;
  bool factored = false;
  bool staged = false;  // a factorize_stage ran on the current values
  int device = -1;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  size_t dev_bytes = 0;
  long long arena_len = 0;
  double *d_aval = nullptr;  // A's values in the caller's CSR order (FP64)
  long long *d_adest = nullptr;
  int *d_arow = nullptr, *d_acol = nullptr;  // A's pattern (refinement residuals)
  int *d_order = nullptr;
  // leaves and work lists
  std::vector<snu::LeafDev> h_leaves;  // vec = 0, cb_off local
  snu::WorkLists wl;
  snu::WorkDev wd;
  int max_leaf_w = 0;  // widest leaf workspace (leaf_size + nbc)
  // dense tables
  snu::DenseDev *d_dd = nullptr;
  std::vector<long long> inv_off;
  long long inv_total = 0, tmp_total = 0;
  snu::Values<double> v64;
  snu::Values<float> v32;
  // standalone solves
  double *d_bx = nullptr;  // staging for host-pointer solves
  double *d_r = nullptr, *d_dx = nullptr;  // refinement residual / correction
  unsigned long long *d_norms = nullptr;     // |A|inf (bits of a double)
  unsigned long long *d_colnorms = nullptr;  // refinement column norms (3 x nrhs)
  int *d_active = nullptr;
  snu::VecDev *d_vec = nullptr;
  snu::PermDev *d_perm = nullptr;
  snu::PrepDev *d_prep = nullptr;
  snu::IrDev *d_ir = nullptr;
  double anorm = 0;
  int nrhs_cap = 0;
  int ir_steps = 0;
  double ir_berr = 0;
  // timing
  bool profile = false;
  bool use_graph = false;
  snu::FactorGraph graph;
  std::vector<hipEvent_t> ev, pev;
  double t[16] = {0};
  // C -= AB launches by tile: 64 x 64 (time ms, launches, GFLOP), then 128 x 128
  double gsplit[6] = {0};
  // SNU_FP32: FP64 factorization of the same matrix on the same device, built
  // and factorized on the first solve whose refinement stalls above 1e-12,
  // refactorized after every new factorization of this handle
  snumat *fb64 = nullptr;
  bool fb64_current = false;
};

namespace snu {

template <typename T>
Values<T> &vals(snumat *h) {
  if constexpr (std::is_same<T, double>::value) return h->v64;
  else return h->v32;
}
template <typename T>
const Values<T> &vals(const snumat *h) {
  if constexpr (std::is_same<T, double>::value) return h->v64;
  else return h->v32;
}

int upload(snumat *h);
// The multifrontal plan of one uploaded matrix (value arrays allocated).
template <typename T>
int build_front_plan(snumat *h);
// Merge the front plans of several matrices into out's host tables (per
// height: their fronts side by side; dense launches merged by key).
template <typename T>
void merge_fronts(const std::vector<const FrontPlan<T> *> &in, FrontPlan<T> &out);
// GEMM tile (64 or 128) for a C block of M x N, by precision; autotune_tiles
// measures the table once per process (on the current device)
template <typename T>
void autotune_tiles();
int tile_for(int prec, int M, int N);
int tile_table(int prec, int *out3);
void free_device(snumat *h);
size_t estimate_device_bytes(const snumat *h);
int device_alloc(snumat *h, void **p, size_t bytes);
float ev_ms(hipEvent_t a, hipEvent_t b);

template <typename T>
int upload_list(T **p, const std::vector<T> &v);

// Append problem C (M x N) -= / = sum_seg A_seg B_seg to a grouped GEMM
// launch with tb x tb tiles; returns 1 if a problem was added.
template <typename T>
int add_gemm(std::vector<GemmProbT<T>> &gp, std::vector<GemmSegT<T>> &gs, int &tiles,
             T *C, long long ldc, int M, int N, const std::vector<GemmSegT<T>> &segs,
             double &flops, int tb = 64);

// GEMV plan of one matrix's dense solves into the packed vectors bp / ys at
// this matrix's segment (ld = n, nrhs columns); keys order fwd / bwd.
template <typename T>
void build_solve_plan(const snumat *h, T *bp, T *ys, int nrhs, GemvPlan<T> &out);
// Copy the plan's descriptors to the device on stream st, growing the
// device arrays only when they are too small (the host vectors must stay
// unchanged until st has passed the copy).
template <typename T>
int upload_gemv_plan(GemvPlan<T> &g, hipStream_t st = nullptr);
// The split-K layout of the plan's launches (no device work); returns the
// partial-sum values its largest launch needs.
template <typename T>
size_t layout_gemv_plan(GemvPlan<T> &g);
template <typename T>
void free_gemv_plan(GemvPlan<T> &g);
template <typename T>
int upload_dense_plan(DensePlan<T> &d);
template <typename T>
void free_dense_plan(DensePlan<T> &d);

// Merge the launch lists of several matrices: launches with equal (key,
// kind) become one launch over the concatenated problems.
template <typename T>
void merge_dense(const std::vector<const DensePlan<T> *> &in, DensePlan<T> &out);
template <typename T>
void merge_gemv(const std::vector<const GemvPlan<T> *> &in, GemvPlan<T> &out);

// Executors over a launch set.
struct FactorSet {
  const PrepDev *prep;
  int nmat;
  long long max_len, max_nnz;
  const IrDev *ir;           // FP32 only: |A| norms into anorm_out
  unsigned long long *anorm_out;
  int max_n;
  const WorkDev *wd;
  const void *fronts;  // FrontPlan of the set's value type
  double eps;
  bool split = false;  // FP32: GEMM tiles through bf16 splits (SNU_GEMM_SPLIT)
  bool scatter_only = false;  // A's values into place, nothing factored
};
// graph (optional, not with profile): capture the launches on the first
// call, replay them afterwards; the phase events then all mark the end.
template <typename T>
int run_factor(const FactorSet &fs, const DensePlan<T> &plan, hipStream_t st,
               const std::vector<hipEvent_t> *ev, bool profile, std::vector<hipEvent_t> *pev,
               double *gemm_ms, FactorGraph *graph = nullptr);
void free_graph(FactorGraph &g);
struct SolveSet {
  const PermDev *perm;
  const VecDev *vec;
  int nmat;
  long long max_elems;
  int max_rhs;
  const WorkDev *wd;
  int cb_ld;
};
template <typename T>
void run_sweep(const SolveSet &ss, T *bp, T *cbuf, const GemvPlan<T> &plan, hipStream_t st);
struct RefineSet {
  SolveSet sweep_b;  // permutation tables b -> x (the caller's vectors)
  SolveSet sweep_r;  // permutation tables r -> dx
  const IrDev *ir;
  int nmat, max_n;
  long long max_elems;
  unsigned long long *norms;  // 3 areas (b, r, x) of nslots
  int nslots;
  const std::vector<int> *slot_mat;   // matrix of each column slot
  const std::vector<double> *anorm;   // |A|inf per matrix
  int *d_active;                      // nmat flags
  int max_steps;
  double tol;
};
int run_refined(const RefineSet &rs, float *bp, float *cbuf, const GemvPlan<float> &plan,
                hipStream_t st, int *steps_out, std::vector<double> *berr_mat);

// Buffers of one standalone solve: the permuted right-hand sides and the
// boundary contributions (value type of the handle), the dense-solve launch
// plan with its device descriptors, and (SNU_FP32) the FP64 residual /
// correction and the column norms of the refinement.
struct SolveBuffers {
  void *bp = nullptr, *cbuf = nullptr;
  const void *plan = nullptr;  // GemvPlan<T> of the handle's value type
  double *r = nullptr, *dx = nullptr;
  unsigned long long *colnorms = nullptr;
};
// Solve with device pointers (caller's b / x, column-major) on the given
// buffers, on the handle's stream; returns when the work is done.  fallback:
// a refinement that stalls above 1e-12 is answered by the handle's FP64 twin
// (built on first need); otherwise SNU_ERR_NOT_CONVERGED is returned with x
// holding the last iterate.
int solve_core(snumat *h, const double *d_b, long long ldb, double *d_x, long long ldx, int nrhs,
               const SolveBuffers &B, bool fallback);
// x = A^-1 b with FP64 factors of an SNU_FP32 handle's matrix (its FP64
// twin: same plan, FP64 arena, the handle's current values; built and
// factorized on first need, refactorized after the handle's next
// factorization); returns when the work is done.
int fp64_fallback_solve(snumat *h, const double *d_b, long long ldb, double *d_x, long long ldx, int nrhs);

}  // namespace snu
