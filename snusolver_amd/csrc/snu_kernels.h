// Device-side descriptors and kernel launchers of the MI355X SnuSolver path.
// Value-carrying kernels are templated on the value type T (double for the
// FP64 path, float for the FP32 factorization refined in FP64); launchers are
// instantiated for both in snu_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace snu {

// Single-copy separator storage, panels tile-compressed:
//   R_p = [D_p | Up_p] : s_p x (s_p + nU_p), ld s_p
//   Lp_p              : nL_p x s_p, ld nL_p
// lmap / umap (from mapoff[p], A_p entries, ancestors bottom-up): stored row
// of Lp_p / stored column of Up_p of each ancestor index, -1 if not stored.
struct DenseDev {
  const long long *offR, *offL, *mapoff;  // per heap block
  const int *bsize, *nL;                  // per heap block
  const int *lmap, *umap;
  const int *aoff;                        // (block * nk + distance) -> ancestor offset
  int nk;
};

// One leaf (one reference rank) of one matrix: device pointers into the
// owning matrix's arrays, so that leaves of several matrices can share a
// launch.
struct LeafDev {
  const int *rowptr, *diag, *wcol;       // LU rowptr (core_n + 1), diag, workspace column per slot
  const int *levptr, *levrows;           // factor / forward level sets
  const int *ulevptr, *ulevrows;         // backward level sets
  const int *ftask, *btask;              // level-ordered row tasks {row, p0, p1, diag}, fwd / bwd
  const int *ubnd;                       // per leaf row: first U entry on a boundary column
  const int *brow;                       // boundary rows (core-local)
  const int *bcol_glob, *bcol_blk, *bcol_loc;  // boundary columns: permuted index, block, offset
  void *vals;                            // leaf LU values (value type of the launch)
  void *S;                               // boundary Schur block, nbr x nbc row-major
  void *dense;                           // the matrix's dense separator arena
  const DenseDev *dd;                    // the matrix's separator tables
  int core_n, leaf_size, nbr, nbc, nlev, nulev;
  int gstart;                            // permuted index of the first leaf row
  int cb_off;                            // rows of the rhs contribution buffer
  int vec;                               // the matrix's slot in the launch's VecDev table
  int pad;
};
// Multifrontal leaf factorization.  A front is an f x f column-major matrix
// whose first nk rows / columns are its pivots; after its partial LU the
// trailing (f - nk) x (f - nk) block holds its update for the parent.
struct FrontChild {
  const void *U;     // the child's update block (its front at (nk, nk)), ld
  const int *rel;    // per entry of the child's B: index in this front (ascending)
  int ld, nb;        // the child's front order, |B| of the child
};
struct FrontDev {
  void *F;                  // the front's values (f x f, column-major)
  void *vals;               // the leaf's LU values (owned slots are read, then written)
  const int *slot, *pos;    // owned LU slots and their positions in F (ascending positions)
  const FrontChild *child;  // children in ascending order (their updates are added in it)
  int f, nk, nslot, nchild;
};
// Boundary Schur block of one leaf from the updates of its root fronts: S
// (nbr x nbc, row-major) = sum over roots of U[arow[r], acol[c]] (entries
// whose row or column the root lacks, -1, add nothing).
struct FrontRoot {
  const void *U;
  const int *arow, *acol;   // per S row / column: index in the root's B (-1)
  int ld;
};
struct LeafSchurDev {
  void *S;
  const FrontRoot *root;
  int nbr, nbc, nroot;
};

// Segment of a packed right-hand-side buffer: columns of length n at
// base + c ld, c < nrhs.
struct VecDev {
  long long base;
  int ld, nrhs;
};
// Caller-side vectors of one matrix for the (un)permutation.
struct PermDev {
  const int *order;
  const double *b;
  double *x;
  long long ldb, ldx;
  int n, vec;
};
// Factor set-up of one matrix.
struct PrepDev {
  void *arena, *inv;
  long long arena_len, inv_len;
  const double *aval;
  const long long *adest;
  long long nnz;
};
// FP64 refinement of one matrix: A in its CSR order, column-major vectors.
struct IrDev {
  const int *rowptr, *colidx;
  const double *val;
  const double *b;
  double *x, *r, *dx;
  long long ldb, ldx, ldr, lddx;
  int n, nrhs;
  int norm0;  // first slot of this matrix's columns in the norm arrays
  int pad;
};

template <typename T>
struct GemmSegT {
  const T *A, *B;
  long long lda, ldb;
  int K;
};
template <typename T>
struct GemmProbT {
  T *C;
  long long ldc;
  int M, N;
  int seg0, nseg;
  int tiles_n, tile0;
};
template <typename T>
struct DiagProbT {
  T *D;
  long long ld;
  int nb;
  T *Linv, *Uinv;  // nb x nb diagonal blocks of the full inverses
  long long ldi;
  T eps;           // pivots of magnitude below eps become +-eps (0: none)
};
// y (M x nrhs) = / -= sum_seg A_seg (M x K) x_seg (K x nrhs); tri: 0 dense,
// 1 lower triangular single segment (columns k <= row), 2 upper (k >= row).
template <typename T>
struct GemvSegT {
  const T *A, *x;
  long long lda, ldx;
  int K;
};
template <typename T>
struct GemvProbT {
  T *y;
  long long ldy;
  int M, seg0, nseg, tri, tile0, nrhs;
  // split K: each segment in pieces of kGemvPiece columns, one workgroup per
  // (64-row tile, piece) writing partial sums at poff; wg0 numbers the
  // problem's first workgroup in its launch
  int parts, wg0;
  long long poff;
};
constexpr int kGemvPiece = 128;
using GemmSeg = GemmSegT<double>;
using GemmProb = GemmProbT<double>;
using DiagProb = DiagProbT<double>;
using GemvSeg = GemvSegT<double>;
using GemvProb = GemvProbT<double>;

int device_lds_limit();

template <typename T>
void launch_fill(T *p, long long n, hipStream_t st);
// zero arena / inverses, scatter A's values (one matrix per table entry)
template <typename T>
void launch_prepare(const PrepDev *tab, int nmat, long long max_len, long long max_nnz,
                    hipStream_t st);
void launch_zero_i32(int *p, int n, hipStream_t st);
// dst = src, bytes a multiple of 16
void launch_copy(const void *src, void *dst, long long bytes, hipStream_t st);
// boundary Schur rows into the dense blocks, per tile of 64 separator rows
template <typename T>
void launch_assemble_schur(int ntile, const int *tile_blk, const int *tile_row0,
                           const int *tile_ptr, const int *seg_leaf, const int *seg_ri,
                           const LeafDev *leaves, hipStream_t st);
// mode 0: C -= sum A B; 1: C = A B; 2: C = -A B
// C = A (first segment) per problem, 64 x 64 tiles numbered as the GEMM's
template <typename T>
void launch_copy2d(const GemmProbT<T> *probs, int nprob, const GemmSegT<T> *segs, int total_tiles,
                   hipStream_t st);
template <typename T>
void launch_gemm(const GemmProbT<T> *probs, int nprob, const GemmSegT<T> *segs,
                 int total_tiles, int mode, hipStream_t st);
template <typename T>
void launch_gemm128(const GemmProbT<T> *probs, int nprob, const GemmSegT<T> *segs,
                    int total_tiles, int mode, hipStream_t st);
// FP32 128 x 128 tiles through three-piece bf16 splits on the bf16 MFMA
// (FP32-accurate; modes as launch_gemm128, same tile counts)
void launch_gemm128_split(const GemmProbT<float> *probs, int nprob, const GemmSegT<float> *segs,
                          int total_tiles, int mode, hipStream_t st);
// the same on 64 x 64 tiles (modes and tile counts of launch_gemm)
void launch_gemm_split(const GemmProbT<float> *probs, int nprob, const GemmSegT<float> *segs,
                       int total_tiles, int mode, hipStream_t st);
template <typename T>
void launch_gemm_variant(const GemmProbT<T> *probs, int nprob, const GemmSegT<T> *segs,
                         int total_tiles, int tile, int kb, int depth, hipStream_t st);
// mode 0: y -= A x; 1: y = A x
// Split-K GEMV: partial sums per (tile, piece) into part, then a fixed-order
// sum per row applied to y (mode 1: y = sum, 0: y -= sum).
template <typename T>
void launch_gemv_split(const GemvProbT<T> *probs, int nprob, const GemvSegT<T> *segs,
                       int total_tiles, int total_wgs, int mode, T *part, hipStream_t st);
template <typename T>
void launch_gemv(const GemvProbT<T> *probs, int nprob, const GemvSegT<T> *segs,
                 int total_tiles, int mode, hipStream_t st);
template <typename T>
void launch_diag_lu(const DiagProbT<T> *probs, int nprob, hipStream_t st);
// max_elems: the largest n x nrhs of the matrices
template <typename T>
void launch_permute_in(const PermDev *pd, int nmat, const VecDev *vd, T *bp,
                       long long max_elems, hipStream_t st);
template <typename T>
void launch_permute_out(const PermDev *pd, int nmat, const VecDev *vd, const T *bp,
                        long long max_elems, hipStream_t st);
template <typename T>
void launch_leaf_fwd(const LeafDev *leaves, int nleaves, const VecDev *vd, T *bp, hipStream_t st);
template <typename T>
void launch_leaf_fwd_sep(const LeafDev *leaves, const int *wg_leaf, const int *wg_first, int nwg,
                         int waves, const VecDev *vd, const T *bp, T *cbuf, int cb_ld,
                         hipStream_t st);
template <typename T>
void launch_assemble_rhs(int ntarget, const int *trow_glob, const int *tptr,
                         const int *tsrc_leaf, const int *tsrc_ri, const LeafDev *leaves,
                         const VecDev *vd, const T *cbuf, int cb_ld, T *bp, int max_rhs,
                         hipStream_t st);
template <typename T>
void launch_leaf_bwd(const LeafDev *leaves, int nleaves, const VecDev *vd, T *bp, hipStream_t st);
// Leaf forward (bwd false) or backward solve with the solution in LDS;
// leaves beyond the LDS fall back to launch_leaf_fwd / launch_leaf_bwd.
template <typename T>
void launch_leaf_solve(const LeafDev *leaves, int nleaves, const VecDev *vd, T *bp, int max_leaf,
                       int max_lev, int max_rhs, bool bwd, hipStream_t st);

// ---- multifrontal leaf factorization ----
// Fronts held whole in LDS (f <= max_f): assembly, partial LU with the
// pivot clamp, write-back of the owned slots, update block to global memory.
template <typename T>
void launch_front_small(const FrontDev *fr, int nfront, int max_f, double eps, hipStream_t st);
// Largest front order the LDS kernel takes for this value type.
template <typename T>
int front_small_max();
// Larger fronts, in global memory: assembly (zero, owned slots, children's
// updates) and write-back of the owned slots after the dense partial LU, one
// workgroup per (front, 64-column chunk) work item.
template <typename T>
void launch_front_assemble(const FrontDev *fr, const int4 *work, const int *wrange, int nwork, hipStream_t st);
template <typename T>
void launch_front_writeback(const FrontDev *fr, const int4 *work, const int *wrange, int nwork, hipStream_t st);
// Boundary Schur blocks from the root fronts, one workgroup per 64 x 64 S
// tile (leaf, first row, first column).
template <typename T>
void launch_front_schur(const LeafSchurDev *ls, const int4 *tiles, int ntile, hipStream_t st);

// ---- FP64 iterative refinement (one matrix per table entry) ----
void launch_residual(const IrDev *tab, int nmat, int max_n, hipStream_t st);
// which: 0 b, 1 r, 2 x; out (nout slots) is zeroed first
void launch_colmax(const IrDev *tab, int nmat, int max_n, int which, unsigned long long *out,
                   int nout, hipStream_t st);
void launch_rowabs_max(const IrDev *tab, int nmat, int max_n, unsigned long long *out,
                       hipStream_t st);
// x += dx where active[matrix] != 0
void launch_update(const IrDev *tab, int nmat, long long max_elems, const int *active,
                   hipStream_t st);

}  // namespace snu
