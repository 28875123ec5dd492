/*
 * snu_oracle.h -- CPU restatement of SnuSolver's construct -> factorize -> solve
 * path (the reference's MKL host path, offlvl = -1), used ONLY as test
 * infrastructure: by tests/, by __graft_entry__.smoke() as the checker, and by
 * bench.py's cpu_baseline leg.  The product library never links or calls it.
 *
 * Every function follows a reference file:line, cited at its definition in
 * snu_oracle.c.  Deviations from the reference (all documented there):
 *   - the MPI ranks collapse to one process; per-leaf work is identical, the
 *     per-rank replicated dense blocks become one copy per block, so Schur
 *     contributions are summed in leaf order instead of the pairwise MPI tree;
 *   - the per-leaf MM_buf (dense_row^2) is compressed to the rows/cols that the
 *     leaf actually touches (untouched entries of MM_buf stay exactly 0).
 */
#ifndef SNU_ORACLE_H
#define SNU_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle_t;

/* Structure phase (integer work; SnuMat_construct.cu, SnuMat_malloc.cu,
 * kernel_sparse.cu:368-687).  Returns 0 on success, <0 on bad input. */
int oracle_create(int n, int nnz, const int *rowptr, const int *colidx,
                  const double *val, int np, const int *order,
                  const int *sizes, oracle_t **out);
void oracle_destroy(oracle_t *o);

/* Full numeric path: leaves (kernel_sparse.cu:1011-1037), dense levels with
 * MKL semantics (SnuMat_factorize.cu:221-244), backward solves
 * (SnuMat_factorize.cu:300-345) and the un-permute (:350-353).
 * nthreads: host threads for the per-leaf work (leaf-parallel, like one rank
 * per leaf); the dense calls use MKL's own threading. */
int oracle_factor_solve(oracle_t *o, const double *b, double *x, int nthreads);

/* Debug: leaves + assembly only (dense blocks hold the assembled values),
 * and a copy of the single-copy dense block (i, j), s_i x s_j col-major. */
int oracle_assemble_only(oracle_t *o, const double *b, int nthreads);
void oracle_sep_block(oracle_t *o, int i, int j, double *out);

/* Only the per-leaf numeric factorization (core_run up to core_numfact_v2). */
int oracle_leaf_factor(oracle_t *o, int nthreads);
/* Only the value scatter that starts core_run: A's leaf values into the LU
 * layout (zeros at fill positions), nothing factored. */
void oracle_leaf_init(oracle_t *o);

/* Which dense backend ran: 1 = MKL (libmkl_rt), 0 = the project's own loops. */
int oracle_dense_backend(void);
/* Force the fallback dense loops (for testing them against MKL). */
void oracle_force_fallback(int on);

/* Phase timings of the last oracle_factor_solve, seconds:
 * [0]=leaf numeric+fwd+Schur, [1]=assembly, [2]=dense levels, [3]=backward. */
void oracle_last_times(double *t4);

/* ---- integer structure getters (for bit-exact parity tests) ---- */
int oracle_num_block(const oracle_t *o);
/* block_start/block_size for e = 1..2np-1 (global permuted order). */
void oracle_blocks(const oracle_t *o, int *block_start, int *block_size);
/* PAP^T (n+1 rowptr, nnz colidx) and perm_map[nnz] (original nnz -> position
 * in the block-merged COO layout). */
void oracle_permuted(const oracle_t *o, int *pa_rowptr, int *pa_colidx,
                     int *perm_map);
/* Per leaf (rank r owns leaf 2np-1-r): sizes then arrays. */
void oracle_leaf_sizes(const oracle_t *o, int rank, int *leaf_size,
                       int *core_n, int *core_nnz, int *lu_nnz);
void oracle_leaf_struct(const oracle_t *o, int rank, int *lu_rowptr,
                        int *lu_colidx, int *lu_diag, int *lu_bias,
                        int *core_map);
/* LU values of the leaf after oracle_leaf_factor / oracle_factor_solve. */
void oracle_leaf_values(const oracle_t *o, int rank, double *lu_data);

#ifdef __cplusplus
}
#endif
#endif
