// Device-side descriptors and kernel launchers of the MI355X SnuSolver path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace snu {

// One leaf (one reference rank): offsets into the concatenated device arrays.
struct LeafDev {
  long long val_off;  // leaf LU values in the value arena
  long long idx_off;  // workspace-column array (one entry per LU slot)
  long long S_off;    // boundary Schur block, nbr x nbc row-major
  int rp_off;         // rowptr (core_n + 1) and diag (core_n) arrays
  int core_n, leaf_size, nbr, nbc;
  int lev_off, nlev, levrows_off;     // factor / forward level sets
  int ulev_off, nulev, ulevrows_off;  // backward level sets
  int brow_off, bcol_off;             // boundary row / col arrays
  int gstart;                         // permuted index of the first leaf row
  int cb_off;                         // rhs contribution rows (x nrhs)
};

// Single-copy separator storage (see the design notes):
//   R_p = [D_p | Up_p] : s_p x (s_p + A_p), ld s_p, ancestors bottom-up
//   Lp_p              : A_p x s_p, ld A_p
struct DenseDev {
  const long long *offR, *offL;  // per heap block
  const int *bsize, *anc_sum;    // per heap block
  const int *aoff;               // (block * nk + distance) -> ancestor offset
  int nk;
};

struct GemmSeg {
  const double *A, *B;
  long long lda, ldb;
  int K;
};
struct GemmProb {
  double *C;
  long long ldc;
  int M, N;
  int seg0, nseg;
  int tiles_n, tile0;
};
struct DiagProb {
  double *D;
  long long ld;
  int nb;
  double *Linv, *Uinv;  // 64 x 64 each, ld 64
};
struct TriProb {
  const double *D;   // s x s LU factors, ld s
  long long ld;
  int s;
  const double *inv;  // per 64-block: [Linv | Uinv], 2 * 4096 doubles each
  double *b;          // rhs rows of this separator, ld ldb
  long long ldb;
};

int device_lds_limit();

void launch_fill(double *p, long long n, hipStream_t st);
void launch_scatter(const double *aval, const long long *dest, int nnz,
                    double *arena, hipStream_t st);
void launch_leaf_rows(const LeafDev *leaves, int nleaves, const int *rowptr,
                      const int *diag, const int *wcol, const int *levptr,
                      const int *levrows, double *arena, int wsz, int waves,
                      double eps, hipStream_t st);
void launch_leaf_sep(const LeafDev *leaves, const int *wg_leaf,
                     const int *wg_first, int nwg, const int *rowptr,
                     const int *diag, const int *wcol, const int *brow,
                     double *arena, double *S, int wsz, int waves,
                     hipStream_t st);
void launch_assemble_schur(int ntarget, const int *trow_blk,
                           const int *trow_loc, const int *tptr,
                           const int *tsrc_leaf, const int *tsrc_ri,
                           const LeafDev *leaves, const int *bcol_blk,
                           const int *bcol_loc, const double *S, double *dense,
                           DenseDev dd, hipStream_t st);
void launch_gemm(const GemmProb *probs, int nprob, const GemmSeg *segs,
                 int total_tiles, bool assign, hipStream_t st);
void launch_diag_lu(const DiagProb *probs, int nprob, hipStream_t st);
void launch_tri_fwd(const TriProb *probs, int nprob, int nrhs, hipStream_t st);
void launch_tri_bwd(const TriProb *probs, int nprob, int nrhs, hipStream_t st);
void launch_permute_in(const double *b, long long ldb, const int *order,
                       double *bp, int n, int nrhs, hipStream_t st);
void launch_permute_out(const double *bp, const int *order, double *x,
                        long long ldx, int n, int nrhs, hipStream_t st);
void launch_leaf_fwd(const LeafDev *leaves, int nleaves, const int *rowptr,
                     const int *diag, const int *wcol, const int *levptr,
                     const int *levrows, const double *arena, double *bp,
                     int n, int nrhs, hipStream_t st);
void launch_leaf_fwd_sep(const LeafDev *leaves, const int *wg_leaf,
                         const int *wg_first, int nwg, int waves,
                         const int *rowptr, const int *wcol, const int *brow,
                         const double *arena, const double *bp, double *cbuf,
                         int n, int nrhs, int cb_ld, hipStream_t st);
void launch_assemble_rhs(int ntarget, const int *trow_glob, const int *tptr,
                         const int *tsrc_leaf, const int *tsrc_ri,
                         const LeafDev *leaves, const double *cbuf, int cb_ld,
                         double *bp, int n, int nrhs, hipStream_t st);
void launch_leaf_bwd(const LeafDev *leaves, int nleaves, const int *rowptr,
                     const int *diag, const int *wcol, const int *ulevptr,
                     const int *ulevrows, const int *bcol_glob,
                     const double *arena, double *bp, int n, int nrhs,
                     hipStream_t st);

}  // namespace snu
