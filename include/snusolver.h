/*
 * snusolver.h -- C-ABI of the MI355X-native SnuSolver factor-and-solve path.
 *
 * One process, one GPU, one handle per matrix.  Plain pointers and sizes only.
 * Every entry point returns an int status (SNU_OK == 0) and never exits the
 * process; snu_last_error() returns the message of the last failure on the
 * calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   snu_initialize          <- void initialize()            src/snusolver.h:5
 *                              (src/solver.cu:35-49)
 *   snumat_construct        <- SnuMat::SnuMat(csr_matrix, double *b, ...)
 *                              src/SnuMat.cu:67-112 (ordering, structure,
 *                              allocation; b is supplied at solve time here)
 *   snumat_factorize +      <- SnuMat::solve(double *x)
 *   snumat_solve               src/SnuMat_factorize.cu:213-357 (factorize with
 *                              fused forward solve, then backward solve);
 *                              split here so factors stay resident
 *   snu_csr_t               <- csr_matrix  src/matrix.h:9-13 (same layout)
 *   snu_solve_csr           <- void solve(csr_matrix, double *b, double *x)
 *                              src/snusolver.h:6 (src/solver.cu:51-89)
 * The C++ shims with the reference's exact initialize()/solve() signatures
 * are declared in include/snusolver_compat.hpp.
 */
#ifndef SNUSOLVER_AMD_H
#define SNUSOLVER_AMD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SNU_OK 0
#define SNU_ERR_BAD_INPUT -1     /* malformed matrix, options or ordering */
#define SNU_ERR_STRUCTURE -2     /* ordering violates the separator property */
#define SNU_ERR_ZERO_DIAG -3     /* structurally zero diagonal in a leaf */
#define SNU_ERR_DUPLICATE -4     /* duplicate (row, col) entries */
#define SNU_ERR_NO_DEVICE -5     /* no HIP device / HIP init failed */
#define SNU_ERR_HIP -6           /* a HIP runtime call failed */
#define SNU_ERR_OOM -7           /* device allocation or HBM budget exceeded */
#define SNU_ERR_NOT_FACTORED -8  /* solve before factorize */
#define SNU_ERR_NUMERIC -9       /* Inf/NaN in the factors (zero pivot) */
#define SNU_ERR_NOT_CONVERGED -10

/* ordering kinds */
#define SNU_ORDER_ND 0    /* project nested dissection of A + A^T (the role of
                             ParMETIS_V3_NodeND): multilevel or level-set
                             separators, the smaller at each bisection */
#define SNU_ORDER_GRID 1  /* geometric nested dissection of an nx*ny*nz grid */
#define SNU_ORDER_GIVEN 2 /* caller's order[n] / sizes[2np-1] (ParMETIS layout) */
#define SNU_ORDER_ND_LEVELSET 3 /* nested dissection by BFS level-set
                                   separators (simpler, for comparison) */

/* precision kinds */
#define SNU_FP64 0 /* FP64 factors and solves (the reference's arithmetic) */
#define SNU_FP32 1 /* FP32 factors and sweeps, FP64 iterative refinement
                      (no reference counterpart) */

/* Same field layout as the reference's csr_matrix (src/matrix.h:9-13):
 * int32 n, m, nnz; rowptr[n+1]; colidx[nnz] (0-based; any column order within
 * a row, duplicates rejected); double data[nnz]. */
typedef struct {
  int n, m, nnz;
  int *rowptr, *colidx;
  double *data;
} snu_csr_t;

typedef struct {
  int np;          /* leaves of the separator tree, power of two >= 1 */
  int ordering;    /* SNU_ORDER_* */
  int grid[3];     /* nx, ny, nz for SNU_ORDER_GRID (nz = 1 for 2D) */
  const int *order;/* SNU_ORDER_GIVEN: order[i] = new index of vertex i */
  const int *sizes;/* SNU_ORDER_GIVEN: 2np-1 block sizes, ParMETIS layout */
  double eps;      /* leaf pivot clamp, reference src/param.h:3 (1e-15) */
  int precision;   /* SNU_FP64 or SNU_FP32 */
  int device;      /* HIP device ordinal */
  uint64_t seed;   /* ordering seed (the ND is deterministic for a seed) */
  int ir_max_steps;/* SNU_FP32: refinement steps per solve (0: default 10;
                      negative: none, the FP32 sweep alone).  A snumat solve
                      whose refinement ends above 1e-12 backward error is solved
                      again with FP64 factors of the same matrix (built and
                      factorized on first need, kept until the next
                      factorize; about the memory of an FP64 handle) */
  double ir_tol;   /* SNU_FP32: target backward error per column (1e-15) */
  int fp32_gemm;   /* SNU_FP32: arithmetic of the factorization's dense GEMM
                      tiles.  SNU_GEMM_SPLIT (default): each FP32 operand
                      as three bf16 pieces (h + m + l == x exactly), the six
                      products of weight >= 2^-16 on the bf16 MFMA, FP32
                      accumulation: FP32-accurate (terms left out <= about
                      2^-23 |a||b| per product).  SNU_GEMM_F32: the FP32 MFMA.
                      The reference's FP32 path is cuBLAS SGEMM. */
} snu_opts_t;

#define SNU_GEMM_SPLIT 0
#define SNU_GEMM_F32 1

typedef struct snumat snumat_t;

/* Fills opts with the defaults: np = 4, ND ordering, eps = 1e-15, FP64,
 * ir_max_steps = 10, ir_tol = 1e-15, fp32_gemm = SNU_GEMM_SPLIT. */
void snu_default_opts(snu_opts_t *opts);

/* Selects the device and checks that it is usable (<- initialize()). */
int snu_initialize(int device);

/* Ordering, symbolic structure, device layout and upload of A's values.
 * Host-only when device < 0 (structure analysis only; no GPU needed). */
int snumat_construct(const snu_csr_t *A, const snu_opts_t *opts,
                     snumat_t **out);

/* Device bytes an upload of this handle takes (value arrays, structure and
 * one-column solve buffers; an estimate from the symbolic structure, valid
 * for host-only handles too).  No reference counterpart: used by callers that
 * admit matrices under a memory budget. */
int snumat_bytes_needed(const snumat_t *h, size_t *bytes);

/* Move a host-only handle (constructed with device < 0) onto a device with
 * A's values (same pattern as at construct).  No reference counterpart. */
int snumat_attach(snumat_t *h, int device, const double *values);

/* Release every device allocation of a handle, keeping its symbolic plan:
 * the handle is then host-only, as constructed with device < 0, and can be
 * attached again (snumat_attach) without a new analysis.  An FP32 handle's
 * FP64 twin is destroyed.  No reference counterpart (serving under a memory
 * budget parks evicted matrices this way). */
int snumat_detach(snumat_t *h);

/* Numeric factorization on the GPU from the resident copy of A's values. */
int snumat_factorize(snumat_t *h);

/* Replace A's values (same pattern) before the next factorize. */
int snumat_update_values(snumat_t *h, const double *values);

/* Solve A x = b for nrhs right-hand sides, column-major with leading
 * dimension ld >= n.  Host pointers.
 * SNU_FP32 handles: x0 from the FP32 factors, then x += A_32^-1 (b - A x)
 * with the residual in FP64 until every column's backward error
 * |b - Ax|inf / (|A|inf |x|inf + |b|inf) <= ir_tol, or until a step stops
 * halving it; returns SNU_ERR_NOT_CONVERGED if it is then still above 1e-12
 * (x holds the last iterate). */
int snumat_solve(snumat_t *h, const double *b, double *x, int nrhs, int ld);

/* Same with device pointers (resident in HBM, no PCIe in the path). */
int snumat_solve_device(snumat_t *h, const double *d_b, double *d_x, int nrhs,
                        int ld);

/* Allocation-free form of snumat_solve_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every working buffer of the solve (permuted right-hand sides,
 * dense-solve vectors and descriptors, split-K partial sums, and for
 * SNU_FP32 handles the FP64 residual, correction and column norms) comes
 * from one 256 MiB scratch array per device that the library holds until
 * the process ends; the call makes no hipMalloc / hipFree.  Needs a handle
 * constructed on a device and factorized.  A call whose buffers exceed the
 * scratch (about 16 n nrhs bytes for FP64 handles, 24 n nrhs for SNU_FP32)
 * returns SNU_ERR_BAD_INPUT with the size in the message.  All calls on one
 * device share the scratch: calls that overlap on one device (other host
 * threads, other handles) race and must be serialised by the caller.
 * SNU_FP32 handles refine as snumat_solve_device does but have no FP64
 * fallback here (building it would allocate): a stalled refinement returns
 * SNU_ERR_NOT_CONVERGED with x holding the last iterate. */
int snumat_solve_device_scratch(snumat_t *h, const double *d_b, double *d_x,
                                int nrhs, int ld);

int snumat_destroy(snumat_t *h);

/* ---------------- batches (no reference counterpart) ----------------
 * A batch runs the factorization and the solves of several constructed
 * matrices (same device, precision and, in FP32, fp32_gemm) as one launch
 * per step across all of them: the leaves of all matrices share the leaf
 * launches, separators of
 * equal tree height (counted from the leaves) share one grouped launch per
 * dense step.  Each matrix's arithmetic is that of its standalone
 * factorization; after snubatch_factorize every member is factored and may
 * also be solved on its own.  The batch does not own the matrices, which
 * must outlive it. */
typedef struct snubatch snubatch_t;
int snubatch_create(snumat_t *const *mats, int count, snubatch_t **out);
int snubatch_factorize(snubatch_t *b);
/* Per matrix m: nrhs[m] columns (0 skips it), b[m] / x[m] column-major with
 * leading dimension ld[m] >= n_m.  Device pointers; all right-hand sides are
 * packed into one sweep (one launch per step and tree height). */
int snubatch_solve_device(snubatch_t *b, const double *const *d_b,
                          double *const *d_x, const int *nrhs, const int *ld);
/* Same with host pointers. */
int snubatch_solve(snubatch_t *b, const double *const *hb, double *const *hx,
                   const int *nrhs, const int *ld);
/* t[0..5] factor phases as snumat_timings, t[6] last solve, t[13] the most
 * refinement steps of any member (SNU_FP32), t[14] the largest final
 * backward error of the refinement, t[15] the number of members whose
 * refinement stalled above 1e-12 and that were answered by FP64 factors of
 * the same matrix instead (each member's FP64 twin, built on first need, as
 * snumat_solve does for a single handle). */
int snubatch_timings(const snubatch_t *b, double *t16);
/* number of dense factor launches and dense solve launches of the batch */
int snubatch_launches(const snubatch_t *b, int *factor_launches,
                      int *solve_launches);
/* Replay the factorization's launch sequence from a HIP graph captured on
 * the next factorize (launch overhead only; same kernels and results). */
int snubatch_use_graph(snubatch_t *b, int on);
int snubatch_destroy(snubatch_t *b);

/* ---------------- serving (no reference counterpart) ----------------
 * A server keeps factorizations resident in HBM under budget_bytes and
 * answers solve requests from any thread.  `slots` worker threads each pack
 * every pending request (up to max_cols columns) into one batched solve --
 * several requests for one matrix become columns of one multi-rhs solve,
 * different matrices share one launch per sweep step and tree height -- and
 * slots run concurrently on their own streams.  snuserve_add analyses a
 * matrix on the host, evicts idle least-recently-used matrices until it
 * fits the budget, then uploads and factorizes it (one precision per
 * server).  Each answer equals the matrix's standalone solve of the same
 * columns to rounding (the sums of the sparse solves are grouped by the
 * column count packed per matrix) and the refinement tolerance. */
typedef struct snuserve snuserve_t;
int snuserve_create(int device, size_t budget_bytes, int slots, int max_cols,
                    snuserve_t **out);
int snuserve_add(snuserve_t *s, const snu_csr_t *A, const snu_opts_t *opts,
                 int *id);
int snuserve_remove(snuserve_t *s, int id);
/* Queue x = A_id^-1 b (host pointers, column-major, ld >= n); b and x must
 * stay valid until snuserve_wait returns for the ticket. */
int snuserve_submit(snuserve_t *s, int id, const double *b, double *x,
                    int nrhs, int ld, long long *ticket);
/* Block until the request is answered; returns its status. */
int snuserve_wait(snuserve_t *s, long long ticket, double *latency_ms);
/* st[0] requests answered, st[1] columns, st[2] packed solves, st[3] p50
 * latency ms, st[4] p99 ms, st[5] mean ms, st[6] max ms, st[7] evictions,
 * st[8] admissions, st[9] resident matrices, st[10] bytes reserved, st[11]
 * wall ms first submit -> last answer, st[12] summed slot busy ms; reset
 * clears the latency record. */
int snuserve_stats(snuserve_t *s, double *st16, int reset);
int snuserve_destroy(snuserve_t *s);

/* Reference-style one shot: construct + factorize + solve + destroy
 * (<- void solve(csr_matrix, double *b, double *x)). */
int snu_solve_csr(const snu_csr_t *A, const double *b, double *x,
                  const snu_opts_t *opts);

const char *snu_last_error(void);

/* Hash of the sources the library was built from (16 hex digits). */
const char *snu_build_id(void);

/* ---------------- introspection (tests, bench) ---------------- */
/* info[0..15]: n, nnz, np, num_block, max_level, total leaf LU nnz,
 * dense entries, max core_n, max leaf_size, max boundary rows, max boundary
 * cols, on device (0/1), factored (0/1), device bytes held, leaf fronts
 * (all leaves), boundary target rows. */
int snumat_info(const snumat_t *h, int64_t *info16);
/* block_start/block_size of e = 1..2np-1, written at [e-1]. */
int snumat_blocks(const snumat_t *h, int *block_start, int *block_size);
/* order[n] and sizes[2np-1] actually used. */
int snumat_ordering(const snumat_t *h, int *order, int *sizes);
/* PAP^T rowptr/colidx and perm_map (original nnz -> merged COO position). */
int snumat_permuted(const snumat_t *h, int *pa_rowptr, int *pa_colidx,
                    int *perm_map);
/* leaf of rank r (heap block 2np-1-r): sizes[4] = leaf_size, core_n,
 * core_nnz, lu_nnz. */
int snumat_leaf_sizes(const snumat_t *h, int rank, int *sizes4);
int snumat_leaf_struct(const snumat_t *h, int rank, int *lu_rowptr,
                       int *lu_colidx, int *lu_diag, int *lu_bias,
                       int *core_map);
/* Multifrontal structure of the leaf of rank r (test hook): sizes4 = fronts,
 * B entries, owned LU slots, child links; then per front {first pivot,
 * pivots, order, parent front (-1: root), height}, B lists (idx_ptr / idx,
 * core indices), each B entry's index in the parent front (rel), owned LU
 * slots with their column-major front positions, children, and for root
 * fronts each B entry's boundary row / column (-1: none). */
int snumat_leaf_front_sizes(const snumat_t *h, int rank, int *sizes4);
int snumat_leaf_fronts(const snumat_t *h, int rank, int *front5, int *idx_ptr,
                       int *idx, int *rel, int *slot_ptr, int *slot, int *pos,
                       int *child_ptr, int *child, int *srow, int *scol);
/* the pivots K of each front (ascending within a front; k_ptr has
 * fronts + 1 entries, k the leaf size). */
int snumat_leaf_front_pivots(const snumat_t *h, int rank, int *k_ptr, int *k);
/* leaf LU values after factorize (copied back from the device); after
 * snumat_factorize_stage(h, 0), A's leaf values in the LU layout. */
int snumat_leaf_values(const snumat_t *h, int rank, double *lu_data);
/* dense diagonal block of separator e (1 <= e < np), s_e x s_e col-major. */
int snumat_sep_diag(const snumat_t *h, int e, double *out);
/* debug: run the factorization up to a stage (0 = A's values scattered into
 * place only, 1 = also the leaves and the boundary Schur assembly,
 * 2 = everything) and read a dense separator block (i, j) of one root path,
 * s_i x s_j col-major. */
int snumat_factorize_stage(snumat_t *h, int stage);
int snumat_sep_block(const snumat_t *h, int i, int j, double *out);
/* phase timings of the last factorize / solve in ms (GPU events):
 * t[0] scatter, t[1] leaf fronts (the leaves' LU), t[2] the leaves'
 * boundary Schur blocks gathered from their root fronts, t[3] Schur assembly, t[4] dense levels, t[5] factorize total,
 * t[6] solve total; dense GEMM launches of the factorization:
 * t[7] C -= AB kernel time (sum over launches, needs snumat_set_profiling),
 * t[8] its launches, t[9] its algorithmic GFLOP (2 M N K per problem);
 * t[10..12] the same for the C = AB (triangular solve by block inverse)
 * launches; t[13] refinement steps of the last solve (SNU_FP32), t[14] its
 * final backward error (max over columns), t[15] 1 when that solve fell back
 * to FP64 factors (refinement stalled above 1e-12), else 0. */
int snumat_timings(const snumat_t *h, double *t16);
/* the C -= AB launches of the last factorize (leaf fronts and dense
 * levels) split by tile: out[0..2] the 64 x 64 ones (kernel time in ms with
 * snumat_set_profiling, launches, GFLOP), out[3..5] the 128 x 128 ones.  Not
 * part of the reference interface (profiling aid). */
int snumat_gemm_split(const snumat_t *h, double *out6);
/* algorithmic counts from the symbolic structure: c[0] F_dense (flops),
 * c[1] F_leaf (flops), c[2] leaf factor bytes, c[3] solve bytes per rhs. */
int snumat_counts(const snumat_t *h, double *c4);
/* work the launch plans execute (structurally zero 64-tiles skipped):
 * w[0] dense factor flops (2 M N K per GEMM problem, 4/3 nb^3 per diagonal
 * block LU with its triangular inverses), w[1] dense factor launches, w[2]
 * of which diagonal-block LU launches, w[3] bytes of one solve sweep per
 * right-hand side (w[4] dense blocks and inverses read by the GEMVs, w[5]
 * leaf LU values and indices, plus the solution vector read and written),
 * w[6] dense solve launches.  Device handles only. */
int snumat_work(const snumat_t *h, double *w8);

/* ---------------- host utilities ---------------- */
/* Synthetic grid operators (seeded splitmix64 stream, one draw per
 * off-diagonal entry in row-major CSR order, x fastest in the grid).
 * stencil: 5, 9 (2D), 7, 27 (3D).  recipe 0: convection-diffusion with
 * off-diagonals -1 + 0.25 (u - 0.5) and the given diagonal; recipe 1:
 * anisotropic 1:10:100 axis weights, +-10% perturbation, diagonal 1.01 x the
 * off-diagonal row sum. */
int snu_gen_grid_size(int nx, int ny, int nz, int stencil, int *n, int *nnz);
int snu_gen_grid(int nx, int ny, int nz, int stencil, int recipe, double diag,
                 uint64_t seed, int *rowptr, int *colidx, double *val);
/* x_true ~ U[-1, 1) (seed), column-major n x nrhs; b = A x_true. */
int snu_gen_rhs(const snu_csr_t *A, int nrhs, uint64_t seed, double *x_true,
                double *b);
/* Orderings in the ParMETIS contract (order[n], sizes[2np-1]). */
int snu_order_grid(int nx, int ny, int nz, int np, int *order, int *sizes);
int snu_order_nd(int n, const int *rowptr, const int *colidx, int np,
                 uint64_t seed, int *order, int *sizes);
/* method 0 (= snu_order_nd): at every bisection the smaller of the
 * multilevel vertex separator and the BFS level-set separator; 1: level-set
 * separators only; 2: multilevel separators only. */
int snu_order_nd_ex(int n, const int *rowptr, const int *colidx, int np,
                    uint64_t seed, int method, int *order, int *sizes);
/* Matrix Market: read header (n, m, nnz after symmetric expansion), then
 * data into caller arrays (CSR, rows column-sorted)
 * (<- read_matrix / coo_to_csr, src/read_matrix.cpp:6-42,50-107). */
int snu_mtx_size(const char *path, int *n, int *m, int *nnz, int *symmetric);
int snu_mtx_read(const char *path, int *rowptr, int *colidx, double *val);
/* Same with flags: SNU_MTX_FILE_ORDER keeps each row's entries in file
 * order (the reference's coo_to_csr), SNU_MTX_NO_EXPAND reads only the
 * stored triangle of symmetric files (the reference expands nothing,
 * src/read_matrix.cpp:73-104).  Both together read exactly what the
 * reference's read_matrix reads.  Pass the same flags to both calls. */
#define SNU_MTX_FILE_ORDER 1
#define SNU_MTX_NO_EXPAND 2
int snu_mtx_size_ex(const char *path, int flags, int *n, int *m, int *nnz,
                    int *symmetric);
int snu_mtx_read_ex(const char *path, int flags, int *rowptr, int *colidx,
                    double *val);
/* Dense n x 1 vector (a right-hand side b): "array" or one-column
 * "coordinate" file (absent coordinate entries are 0). */
int snu_mtx_read_vector(const char *path, int n, double *x);
int snu_mtx_write_vector(const char *path, const double *x, int n);
/* relative backward error ||Ax-b||_inf / (||A||_inf ||x||_inf + ||b||_inf) */
double snu_backward_error(const snu_csr_t *A, const double *x,
                          const double *b);

/* ---------------- device memory plumbing ----------------
 * For callers that keep b / x resident in HBM (kind 1: host->device,
 * 2: device->host, 3: device->device). */
int snu_dev_alloc(size_t bytes, void **p);
int snu_dev_free(void *p);
int snu_dev_copy(void *dst, const void *src, size_t bytes, int kind);
int snu_dev_sync(void);
/* Time every dense GEMM launch of the next factorizations with HIP events on
 * the handle's stream (timings t[7]); adds event records between launches. */
int snumat_set_profiling(snumat_t *h, int on);
/* Factorize by replaying a HIP graph of the launch sequence, captured on the
 * next factorize (no reference counterpart; results unchanged; phase timings
 * then report the total only). */
int snumat_use_graph(snumat_t *h, int on);

/* ---------------- kernel-level test hooks (host pointers) ----------------
 * Not part of the reference interface: they expose single kernels of this
 * library to the parity tests. */
/* assign = 0: C -= A B (64 x 64 tiles), 2: C -= A B (128 x 128 tiles),
 * 1: C = A B, 3: C = -A B, 4: C = A B (128 x 128 tiles), 5: C = -A B
 * (128 x 128 tiles); col-major.  snu_test_gemm_f32 also takes 6: C -= A B,
 * 7: C = A B, 8: C = -A B through the FP32 bf16-split kernel (128 x 128
 * tiles, as SNU_GEMM_SPLIT) and 9, 10, 11 the same on 64 x 64 tiles. */
int snu_test_gemm(int M, int N, int K, const double *A, int lda,
                  const double *B, int ldb, double *C, int ldc, int assign);
int snu_test_gemm_f32(int M, int N, int K, const float *A, int lda,
                      const float *B, int ldb, float *C, int ldc, int assign);
/* average time (ms) of the grouped C -= A B kernel on device-resident
 * operands, K split into nseg concatenated segments; tile 64 or 128 (K chunk
 * 16), 6432 / 12832 (K chunk 32) or 642 / 1282 (K chunk 16, loads two chunks
 * ahead), or 1283 / 643 (SNU_FP32 only: 128 x 128 / 64 x 64 through bf16
 * splits); prec
 * SNU_FP64 or SNU_FP32. */
int snu_bench_gemm(int M, int N, int K, int nseg, int tile, int prec, int reps,
                   double *ms_out);
/* tiles3: 64 or 128, the tile of the grouped C -= A B kernel per size bucket
 * of min(M, N) ([128, 512), [512, 2048), >= 2048), chosen by the
 * autotuner's timings on this device (run once per process). */
int snu_gemm_tiles(int device, int precision, int *tiles3);
/* average time (ms) of a plain device-to-device copy kernel of `bytes`
 * (multiple of 16): the known-size reference for HBM byte counters. */
int snu_bench_copy(size_t bytes, int reps, double *ms_out);
/* in-place no-pivot LU of an nb x nb block (nb <= 64); Linv/Uinv: 64 x 64. */
int snu_test_diag_lu(int nb, double *D, int ld, double *Linv, double *Uinv);
int snu_test_diag_lu_f32(int nb, float *D, int ld, float *Linv, float *Uinv);

#ifdef __cplusplus
}
#endif
#endif
