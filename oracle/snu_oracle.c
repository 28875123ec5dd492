/*
 * snu_oracle.c -- CPU restatement of SnuSolver's construct -> factorize ->
 * solve path, MKL host semantics (reference compiled with offlvl = -1).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product and
 * the CPU baseline timed by bench.py; the product never links or calls it.
 *
 * Each function cites the reference file:line it restates.  What follows the
 * reference exactly (same operations in the same order, no FMA contraction):
 *   calculate_block, _who, permutate, construct (COO merge layout, perm_map),
 *   core_preprocess, core_symbfact, core_map, LU_bias, CSC transpose,
 *   core_numfact_v2, core_trsm, core_update_b, core_GEMM (per element:
 *   the same subtraction sequence), the dense level loop and its MKL calls,
 *   the dense and leaf backward solves, the gather / un-permute.
 * What differs (one process holds what the reference spreads over MPI ranks):
 *   - one copy per dense block instead of one per rank, so the sibling
 *     contributions (core_reduction / reduction, SnuMat_factorize.cu:29-77,
 *     kernel_sparse.cu:54-90) are added in ascending rank order;
 *   - each leaf's MM_buf is held compressed to the separator rows with a
 *     non-empty L_SL row and the separator columns with a non-empty U_LS
 *     column; every other MM_buf entry of the reference stays exactly 0.
 */
#define _GNU_SOURCE
#include "snu_oracle.h"

#include <dlfcn.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static const double kEps = 1e-15; /* src/param.h:3 */

/* ------------------------------------------------------------------------ */
/* Dense backend: MKL through libmkl_rt (kernel_mkl.cpp:16-103), or loops.  */
/* ------------------------------------------------------------------------ */
enum {
  kColMajor = 102, kNoTrans = 111, kUpper = 121, kLower = 122,
  kNonUnit = 131, kUnit = 132, kLeft = 141, kRight = 142
};
typedef int (*getrfnp_fn)(int, int, int, double *, int);
typedef void (*trsm_fn)(int, int, int, int, int, int, int, double,
                        const double *, int, double *, int);
typedef void (*gemm_fn)(int, int, int, int, int, int, double, const double *,
                        int, const double *, int, double, double *, int);
static getrfnp_fn p_getrfnp;
static trsm_fn p_trsm;
static gemm_fn p_gemm;
static int g_backend = -1;
static int g_force_fallback = 0;
static double g_times[4];

static void load_backend(void) {
  if (g_backend >= 0) return;
  g_backend = 0;
  const char *names[] = {"libmkl_rt.so.1", "libmkl_rt.so",
                         "/opt/conda/lib/libmkl_rt.so.1", NULL};
  for (int i = 0; names[i]; i++) {
    void *h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    p_getrfnp = (getrfnp_fn)dlsym(h, "LAPACKE_mkl_dgetrfnp");
    p_trsm = (trsm_fn)dlsym(h, "cblas_dtrsm");
    p_gemm = (gemm_fn)dlsym(h, "cblas_dgemm");
    if (p_getrfnp && p_trsm && p_gemm) {
      g_backend = 1;
      return;
    }
  }
}
int oracle_dense_backend(void) {
  load_backend();
  return g_force_fallback ? 0 : g_backend;
}
void oracle_force_fallback(int on) { g_force_fallback = on; }
void oracle_last_times(double *t4) { memcpy(t4, g_times, sizeof(g_times)); }

/* Fallback loops, textbook forms of the same four operations. */
static void fb_getrfnp(int n, int m, double *a, int lda) {
  int mn = n < m ? n : m;
  for (int k = 0; k < mn; k++) {
    double piv = a[k + (size_t)k * lda];
    for (int i = k + 1; i < n; i++) a[i + (size_t)k * lda] /= piv;
    for (int j = k + 1; j < m; j++) {
      double u = a[k + (size_t)j * lda];
      for (int i = k + 1; i < n; i++)
        a[i + (size_t)j * lda] -= a[i + (size_t)k * lda] * u;
    }
  }
}

/* snusolver_LU: LAPACKE_mkl_dgetrfnp(col, n, m, A, n)  kernel_mkl.cpp:16-31 */
static void d_lu(int n, int m, double *a) {
  if (!n || !m) return;
  if (oracle_dense_backend()) p_getrfnp(kColMajor, n, m, a, n);
  else fb_getrfnp(n, m, a, n);
}
/* snusolver_trsm_Lxb: B(n x m) <- L^-1 B, unit lower  kernel_mkl.cpp:32-47 */
static void d_trsm_Lxb(const double *L, int n, int m, double *b) {
  if (!n || !m) return;
  if (oracle_dense_backend()) {
    p_trsm(kColMajor, kLeft, kLower, kNoTrans, kUnit, n, m, 1.0, L, n, b, n);
    return;
  }
  for (int j = 0; j < m; j++) {
    double *x = b + (size_t)j * n;
    for (int k = 0; k < n; k++)
      for (int i = k + 1; i < n; i++) x[i] -= L[i + (size_t)k * n] * x[k];
  }
}
/* snusolver_trsm_xUb: B(n x m) <- B U^-1, U m x m  kernel_mkl.cpp:48-64 */
static void d_trsm_xUb(const double *U, int n, int m, double *b) {
  if (!n || !m) return;
  if (oracle_dense_backend()) {
    p_trsm(kColMajor, kRight, kUpper, kNoTrans, kNonUnit, n, m, 1.0, U, m, b,
           n);
    return;
  }
  for (int j = 0; j < m; j++) {
    double *xj = b + (size_t)j * n;
    for (int k = 0; k < j; k++) {
      double u = U[k + (size_t)j * m];
      const double *xk = b + (size_t)k * n;
      for (int i = 0; i < n; i++) xj[i] -= xk[i] * u;
    }
    double d = U[j + (size_t)j * m];
    for (int i = 0; i < n; i++) xj[i] /= d;
  }
}
/* snusolver_trsm_Uxb: B(n x m) <- U^-1 B  kernel_mkl.cpp:66-82 */
static void d_trsm_Uxb(const double *U, int n, int m, double *b) {
  if (!n || !m) return;
  if (oracle_dense_backend()) {
    p_trsm(kColMajor, kLeft, kUpper, kNoTrans, kNonUnit, n, m, 1.0, U, n, b,
           n);
    return;
  }
  for (int j = 0; j < m; j++) {
    double *x = b + (size_t)j * n;
    for (int k = n - 1; k >= 0; k--) {
      x[k] /= U[k + (size_t)k * n];
      for (int i = 0; i < k; i++) x[i] -= U[i + (size_t)k * n] * x[k];
    }
  }
}
/* snusolver_gemm: C(m x n) -= A(m x k) B(k x n)  kernel_mkl.cpp:83-103 */
static void d_gemm(int m, int k, int n, const double *A, const double *B,
                   double *C) {
  if (!m || !n || !k) return;
  if (oracle_dense_backend()) {
    p_gemm(kColMajor, kNoTrans, kNoTrans, m, n, k, -1.0, A, m, B, k, 1.0, C,
           m);
    return;
  }
  for (int j = 0; j < n; j++)
    for (int p = 0; p < k; p++) {
      double bv = B[p + (size_t)j * k];
      for (int i = 0; i < m; i++) C[i + (size_t)j * m] -= A[i + (size_t)p * m] * bv;
    }
}

/* ------------------------------------------------------------------------ */
/* Data model                                                               */
/* ------------------------------------------------------------------------ */
typedef struct {
  int rank, leaf, leaf_size, core_n, dense_row, npath;
  int *path;   /* path[k] = leaf >> k (all_parents, SnuMat.cu:56-64) */
  int *cstart; /* core-local block start of path[k]; npath+1 entries */
  int core_nnz, coo_off;
  int *core_rowptr, *core_colidx, *core_map;
  int LU_nnz;
  int *LU_rowptr, *LU_colidx, *LU_diag, *LU_bias;
  int *LU_colptr_trans, *LU_rowidx_trans, *LU_trans_map, *LU_map;
  double *LU_data;
  /* compressed MM_buf */
  int nbr, nbc;
  int *brow;     /* core rows >= leaf_size with a non-empty L row */
  int *bcol_inv; /* (core col - leaf_size) -> compressed col or -1 */
  int *bcol;     /* compressed col -> core col */
  double *S;     /* nbr x nbc, row major */
  double *bl;    /* core-local right-hand side */
} oleaf_t;

struct oracle {
  int n, nnz, np, num_block, max_level, nk;
  int *block_start, *block_size, *who, *level; /* 0..num_block */
  int *order; /* order[i] = permuted index of vertex i (ParMETIS contract) */
  int *pa_rowptr, *pa_colidx, *perm_map;
  int *coo_r, *coo_c;
  double *coo_val;
  int *gnnz, *goff; /* per grid id: nnz and first COO position */
  int *mstart, *msize;
  oleaf_t *leaves;
  double **dblk; /* per grid id: single-copy dense block */
};

/* grid id of the COO block (i, j); SnuMat.h:55-65 (U[x][k] / L[x][k]) */
static int grid_id(const struct oracle *o, int i, int j) {
  int x = i > j ? i : j, y = i < j ? i : j, k = 0;
  for (int jj = x; jj > y; jj /= 2) k++;
  return (x * o->nk + k) * 2 + (i >= j ? 0 : 1);
}

static int ilog2(int v) {
  int l = 0;
  while (v > 1) v >>= 1, l++;
  return l;
}

/* dfs_block_order + calculate_block  SnuMat_construct.cu:21-38 */
static void dfs_block_order(int *v, int *cnt, int cur, int parts) {
  if (parts > 1) {
    dfs_block_order(v, cnt, cur * 2 + 1, parts / 2);
    dfs_block_order(v, cnt, cur * 2, parts / 2);
  }
  v[(*cnt)++] = cur;
}

/* _who  SnuMat.cu:24-30 */
static int who_of(int np, int block) {
  while (block >= np + np) block /= 2;
  while (block < np) block = block + block + 1;
  return np + np - block - 1;
}

typedef struct {
  int key, idx;
} kv_t;
static int kv_cmp(const void *a, const void *b) {
  const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}
static int int_cmp(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return x < y ? -1 : (x > y);
}

/* permutate  SnuMat_construct.cu:40-80 */
static void permutate(struct oracle *o, const int *rowptr, const int *colidx,
                      const int *order) {
  int n = o->n;
  int *np_ = o->pa_rowptr, *ni = o->pa_colidx;
  np_[0] = 0;
  for (int i = 0; i < n; i++) np_[order[i] + 1] = rowptr[i + 1] - rowptr[i];
  int max_row = 0;
  for (int i = 0; i < n; i++)
    if (np_[i + 1] > max_row) max_row = np_[i + 1];
  for (int i = 0; i < n; i++) np_[i + 1] += np_[i];
  kv_t *v = (kv_t *)malloc(sizeof(kv_t) * (max_row + 1));
  for (int i = 0; i < n; i++) {
    int bs = rowptr[i + 1] - rowptr[i];
    for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
      v[j - rowptr[i]].key = order[colidx[j]], v[j - rowptr[i]].idx = j;
    qsort(v, bs, sizeof(kv_t), kv_cmp);
    for (int j = np_[order[i]], jj = 0; j < np_[order[i] + 1]; j++, jj++) {
      ni[j] = order[colidx[v[jj].idx]];
      o->perm_map[v[jj].idx] = j;
    }
  }
  free(v);
}

/* construct (clearGrid/setGrid/addElement)  SnuMat_construct.cu:81-159 */
static int construct(struct oracle *o) {
  int n = o->n, nb = o->num_block;
  int *block = (int *)malloc(sizeof(int) * (n ? n : 1));
  int *cmap = (int *)malloc(sizeof(int) * (o->nnz ? o->nnz : 1));
  for (int b = 1; b <= nb; b++)
    for (int i = o->block_start[b]; i < o->block_start[b] + o->block_size[b]; i++)
      block[i] = b;
  /* separator property: every entry couples blocks on one root path */
  for (int r = 0; r < n; r++)
    for (int p = o->pa_rowptr[r]; p < o->pa_rowptr[r + 1]; p++) {
      int x = block[r], y = block[o->pa_colidx[p]];
      if (x < y) { int t = x; x = y; y = t; }
      while (x > y) x /= 2;
      if (x != y) { free(block); free(cmap); return -2; }
    }
  int ng = (nb + 1) * o->nk * 2;
  memset(o->gnnz, 0, sizeof(int) * ng);
  for (int r = 0; r < n; r++)
    for (int p = o->pa_rowptr[r]; p < o->pa_rowptr[r + 1]; p++)
      o->gnnz[grid_id(o, block[r], block[o->pa_colidx[p]])]++;
  /* setGrid order: for i descending, (i,i) then (i,ii),(ii,i) upwards */
  int bias = 0;
  int *gend = (int *)malloc(sizeof(int) * ng);
  for (int i = nb; i >= 1; i--) {
    o->mstart[i] = bias;
    int g = grid_id(o, i, i);
    o->goff[g] = bias; bias += o->gnnz[g]; gend[g] = bias;
    for (int ii = i / 2; ii; ii /= 2) {
      g = grid_id(o, i, ii);
      o->goff[g] = bias; bias += o->gnnz[g]; gend[g] = bias;
      g = grid_id(o, ii, i);
      o->goff[g] = bias; bias += o->gnnz[g]; gend[g] = bias;
    }
    o->msize[i] = bias - o->mstart[i];
  }
  for (int r = n - 1; r >= 0; r--)
    for (int idx = o->pa_rowptr[r + 1] - 1; idx >= o->pa_rowptr[r]; idx--) {
      int c = o->pa_colidx[idx];
      int g = grid_id(o, block[r], block[c]);
      int pos = --gend[g];
      o->coo_r[pos] = r;
      o->coo_c[pos] = c;
      cmap[idx] = pos;
    }
  for (int i = 0; i < o->nnz; i++) o->perm_map[i] = cmap[o->perm_map[i]];
  free(gend);
  free(cmap);
  free(block);
  return 0;
}

/* core_symbfact  kernel_sparse.cu:368-524 */
static void core_symbfact(oleaf_t *L) {
  int core_n = L->core_n, leaf_size = L->leaf_size;
  size_t cap = (size_t)L->core_nnz * 4 + 16;
  int *new_rowptr = (int *)malloc(sizeof(int) * (core_n + 1));
  int *new_colidx = (int *)malloc(sizeof(int) * cap);
  char *vst = (char *)calloc(core_n + 1, 1);
  int *Uptr = (int *)malloc(sizeof(int) * (core_n + 1));
  int *Lj = (int *)malloc(sizeof(int) * (core_n + 1));
  int *Uj = (int *)malloc(sizeof(int) * (core_n + 1));
  size_t cur = 0;
  new_rowptr[0] = 0;
  for (int j = 0; j < core_n; j++) {
    int cntL = 0, cntU = 0;
    for (int i = L->core_rowptr[j]; i < L->core_rowptr[j + 1]; i++) {
      int h = L->core_colidx[i];
      if (h < j) Lj[cntL++] = h, vst[h] = 1;
      else Uj[cntU++] = h, vst[h] = 1;
    }
    for (int idx = 0; idx < cntL; idx++) {
      int i = Lj[idx];
      for (int uidx = Uptr[i]; uidx < new_rowptr[i + 1]; uidx++) {
        int h = new_colidx[uidx];
        if (j >= leaf_size && h >= leaf_size) break; /* :448 */
        if (h < j) {
          if (!vst[h]) Lj[cntL++] = h, vst[h] = 1;
        } else {
          if (!vst[h]) Uj[cntU++] = h, vst[h] = 1;
        }
      }
    }
    qsort(Lj, cntL, sizeof(int), int_cmp);
    qsort(Uj, cntU, sizeof(int), int_cmp);
    if (cur + cntL + cntU > cap) {
      while (cur + cntL + cntU > cap) cap <<= 1;
      new_colidx = (int *)realloc(new_colidx, sizeof(int) * cap);
    }
    for (int i = 0; i < cntL; i++) new_colidx[cur++] = Lj[i];
    Uptr[j] = (int)cur;
    for (int i = 0; i < cntU; i++) new_colidx[cur++] = Uj[i];
    new_rowptr[j + 1] = (int)cur;
    for (int i = 0; i < cntL; i++) vst[Lj[i]] = 0;
    for (int i = 0; i < cntU; i++) vst[Uj[i]] = 0;
  }
  L->LU_colidx = new_colidx;
  L->LU_rowptr = new_rowptr;
  L->LU_diag = Uptr;
  L->LU_nnz = (int)cur;
  free(Lj);
  free(Uj);
  free(vst);
}

/* core_preprocess  kernel_sparse.cu:526-687 (block_start :527-531, core CSR
 * :533-573, core_map :576-589, LU_bias :591-597, CSC :622-645). */
static int core_preprocess(struct oracle *o, oleaf_t *L) {
  int leaf = L->leaf;
  L->npath = 0;
  for (int i = leaf; i >= 1; i /= 2) L->npath++;
  L->path = (int *)malloc(sizeof(int) * L->npath);
  L->cstart = (int *)malloc(sizeof(int) * (L->npath + 1));
  int sum = 0;
  for (int k = 0; k < L->npath; k++) {
    L->path[k] = leaf >> k;
    L->cstart[k] = sum;
    sum += o->block_size[leaf >> k];
  }
  L->cstart[L->npath] = sum;
  L->core_n = sum;
  L->leaf_size = o->block_size[leaf];
  L->dense_row = sum - L->leaf_size;
  L->coo_off = o->mstart[leaf];
  L->core_nnz = o->msize[leaf];
  int core_n = L->core_n;
  /* core_blocks: (leaf,leaf), (leaf,p1), (p1,leaf), (leaf,p2), ... */
  int nblk = 1 + 2 * (L->npath - 1);
  int *bi = (int *)malloc(sizeof(int) * nblk), *bj = (int *)malloc(sizeof(int) * nblk);
  int *bki = (int *)malloc(sizeof(int) * nblk), *bkj = (int *)malloc(sizeof(int) * nblk);
  bi[0] = bj[0] = leaf; bki[0] = bkj[0] = 0;
  for (int k = 1, t = 1; k < L->npath; k++) {
    bi[t] = leaf; bj[t] = leaf >> k; bki[t] = 0; bkj[t] = k; t++;
    bi[t] = leaf >> k; bj[t] = leaf; bki[t] = k; bkj[t] = 0; t++;
  }
  L->core_rowptr = (int *)calloc(core_n + 1, sizeof(int));
  L->core_colidx = (int *)malloc(sizeof(int) * (L->core_nnz + 1));
  L->core_map = (int *)malloc(sizeof(int) * (L->core_nnz + 1));
  for (int e = 0; e < nblk; e++) {
    int g = grid_id(o, bi[e], bj[e]);
    for (int p = 0; p < o->gnnz[g]; p++) {
      int q = o->goff[g] + p;
      L->core_rowptr[o->coo_r[q] - o->block_start[bi[e]] + L->cstart[bki[e]]]++;
    }
  }
  for (int i = 0; i < core_n; i++) L->core_rowptr[i + 1] += L->core_rowptr[i];
  for (int e = nblk - 1; e >= 0; e--) {
    int g = grid_id(o, bi[e], bj[e]);
    for (int p = o->gnnz[g] - 1; p >= 0; p--) {
      int q = o->goff[g] + p;
      int r = o->coo_r[q] - o->block_start[bi[e]] + L->cstart[bki[e]];
      int c = o->coo_c[q] - o->block_start[bj[e]] + L->cstart[bkj[e]];
      L->core_colidx[--L->core_rowptr[r]] = c;
    }
  }
  core_symbfact(L);
  int *cur = (int *)malloc(sizeof(int) * (core_n + 1));
  for (int i = 0; i < core_n; i++) cur[i] = L->LU_rowptr[i];
  for (int e = 0; e < nblk; e++) {
    int g = grid_id(o, bi[e], bj[e]);
    for (int p = 0; p < o->gnnz[g]; p++) {
      int q = o->goff[g] + p;
      int r = o->coo_r[q] - o->block_start[bi[e]] + L->cstart[bki[e]];
      int c = o->coo_c[q] - o->block_start[bj[e]] + L->cstart[bkj[e]];
      while (L->LU_colidx[cur[r]] != c) cur[r]++;
      L->core_map[q - L->coo_off] = cur[r];
    }
  }
  free(cur);
  free(bi); free(bj); free(bki); free(bkj);
  int ls = L->leaf_size;
  for (int i = 0; i < ls; i++) /* structurally non-zero diagonal required */
    if (L->LU_diag[i] >= L->LU_rowptr[i + 1] || L->LU_colidx[L->LU_diag[i]] != i)
      return -3;
  L->LU_bias = (int *)malloc(sizeof(int) * (ls + 1));
  for (int i = 0; i < ls; i++) {
    int ptr = L->LU_rowptr[i];
    while (ptr < L->LU_rowptr[i + 1] && L->LU_colidx[ptr] < ls) ptr++;
    L->LU_bias[i] = ptr;
  }
  int nz = L->LU_nnz;
  L->LU_colptr_trans = (int *)calloc(core_n + 1, sizeof(int));
  L->LU_rowidx_trans = (int *)malloc(sizeof(int) * (nz + 1));
  L->LU_trans_map = (int *)malloc(sizeof(int) * (nz + 1));
  L->LU_map = (int *)malloc(sizeof(int) * (nz + 1));
  for (int r = 0; r < core_n; r++)
    for (int p = L->LU_rowptr[r]; p < L->LU_rowptr[r + 1]; p++)
      L->LU_colptr_trans[L->LU_colidx[p]]++;
  for (int i = 0; i < core_n; i++) L->LU_colptr_trans[i + 1] += L->LU_colptr_trans[i];
  for (int r = core_n - 1; r >= 0; r--)
    for (int p = L->LU_rowptr[r]; p < L->LU_rowptr[r + 1]; p++) {
      int c = L->LU_colidx[p];
      L->LU_rowidx_trans[--L->LU_colptr_trans[c]] = r;
      L->LU_trans_map[L->LU_colptr_trans[c]] = p;
      L->LU_map[p] = L->LU_colptr_trans[c];
    }
  L->LU_data = (double *)calloc(nz + 1, sizeof(double));
  /* compressed MM_buf maps */
  int nsep = core_n - ls;
  L->brow = (int *)malloc(sizeof(int) * (nsep + 1));
  L->bcol_inv = (int *)malloc(sizeof(int) * (nsep + 1));
  L->bcol = (int *)malloc(sizeof(int) * (nsep + 1));
  L->nbr = 0;
  for (int r = ls; r < core_n; r++)
    if (L->LU_rowptr[r + 1] > L->LU_rowptr[r]) L->brow[L->nbr++] = r;
  for (int c = 0; c < nsep; c++) L->bcol_inv[c] = 0;
  for (int r = 0; r < ls; r++)
    for (int p = L->LU_bias[r]; p < L->LU_rowptr[r + 1]; p++)
      L->bcol_inv[L->LU_colidx[p] - ls] = 1;
  L->nbc = 0;
  for (int c = 0; c < nsep; c++) {
    if (L->bcol_inv[c]) { L->bcol[L->nbc] = c + ls; L->bcol_inv[c] = L->nbc++; }
    else L->bcol_inv[c] = -1;
  }
  L->S = NULL;
  L->bl = (double *)malloc(sizeof(double) * (core_n + 1));
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Numeric leaf work                                                        */
/* ------------------------------------------------------------------------ */
/* value scatter of core_run  kernel_sparse.cu:1012-1016 */
static void leaf_init_values(const struct oracle *o, oleaf_t *L) {
  memset(L->LU_data, 0, sizeof(double) * L->LU_nnz);
  for (int i = 0; i < L->core_nnz; i++)
    L->LU_data[L->core_map[i]] = o->coo_val[L->coo_off + i];
}

/* core_numfact_v2  kernel_sparse.cu:740-781 */
static void core_numfact_v2(oleaf_t *L) {
  double *LU_data = L->LU_data;
  double *LU_buf = (double *)calloc(L->core_n + 1, sizeof(double));
  for (int r = 0; r < L->leaf_size; r++) {
    double tmp = LU_data[L->LU_diag[r]];
    if (tmp < kEps && tmp > -kEps) {
      if (tmp < 0) tmp = -kEps;
      else tmp = kEps;
    }
    LU_data[L->LU_diag[r]] = tmp;
    for (int p = L->LU_diag[r] + 1; p < L->LU_rowptr[r + 1]; p++)
      LU_buf[L->LU_colidx[p]] = LU_data[p];
    int Lbound = L->LU_map[L->LU_diag[r]] + 1;
    int Rbound = L->LU_colptr_trans[r + 1];
    for (int cp = Lbound; cp < Rbound; cp++) {
      int nxt_row = L->LU_rowidx_trans[cp];
      int nxt_ptr = L->LU_trans_map[cp];
      LU_data[nxt_ptr] /= tmp;
      double dd = LU_data[nxt_ptr];
      int end = L->LU_rowptr[nxt_row + 1];
      for (int it = nxt_ptr + 1; it < end; it++)
        LU_data[it] -= dd * LU_buf[L->LU_colidx[it]];
    }
    for (int p = L->LU_diag[r] + 1; p < L->LU_rowptr[r + 1]; p++)
      LU_buf[L->LU_colidx[p]] = 0;
  }
  free(LU_buf);
}

/* core_trsm  kernel_sparse.cu:915-927 */
static void core_trsm(oleaf_t *L, double *b) {
  for (int r = 0; r < L->leaf_size; r++) {
    double tmp = 0;
    for (int p = L->LU_rowptr[r]; p < L->LU_diag[r]; p++)
      tmp += L->LU_data[p] * b[L->LU_colidx[p]];
    b[r] -= tmp;
  }
}
/* core_update_b  kernel_sparse.cu:935-944 */
static void core_update_b(oleaf_t *L, double *b) {
  for (int r = L->leaf_size; r < L->core_n; r++) {
    double tmp = 0;
    for (int p = L->LU_rowptr[r]; p < L->LU_rowptr[r + 1]; p++)
      tmp += L->LU_data[p] * b[L->LU_colidx[p]];
    b[r] -= tmp;
  }
}
/* core_GEMM  kernel_sparse.cu:953-971, into the compressed MM_buf */
static void core_GEMM(oleaf_t *L) {
  int ls = L->leaf_size;
  size_t sz = (size_t)L->nbr * L->nbc;
  L->S = (double *)calloc(sz ? sz : 1, sizeof(double));
  for (int ri = 0; ri < L->nbr; ri++) {
    int r = L->brow[ri];
    double *Srow = L->S + (size_t)ri * L->nbc;
    for (int p1 = L->LU_rowptr[r]; p1 < L->LU_rowptr[r + 1]; p1++) {
      int r2 = L->LU_colidx[p1];
      double val = L->LU_data[p1];
      for (int p2 = L->LU_bias[r2]; p2 < L->LU_rowptr[r2 + 1]; p2++) {
        int c = L->LU_colidx[p2];
        Srow[L->bcol_inv[c - ls]] -= val * L->LU_data[p2];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Creation                                                                 */
/* ------------------------------------------------------------------------ */
int oracle_create(int n, int nnz, const int *rowptr, const int *colidx,
                  const double *val, int np, const int *order,
                  const int *sizes, oracle_t **out) {
  *out = NULL;
  if (n < 0 || nnz < 0 || np < 1 || (np & (np - 1))) return -1;
  if (rowptr[0] != 0 || rowptr[n] != nnz) return -1;
  for (int i = 0; i < nnz; i++)
    if (colidx[i] < 0 || colidx[i] >= n) return -1;
  struct oracle *o = (struct oracle *)calloc(1, sizeof(struct oracle));
  o->n = n; o->nnz = nnz; o->np = np;
  o->num_block = 2 * np - 1;
  o->max_level = ilog2(np);
  o->nk = o->max_level + 1;
  int nb = o->num_block;
  /* order must be a permutation, sizes must sum to n */
  {
    char *seen = (char *)calloc(n + 1, 1);
    long long s = 0;
    int bad = 0;
    for (int i = 0; i < n && !bad; i++) {
      if (order[i] < 0 || order[i] >= n || seen[order[i]]) bad = 1;
      else seen[order[i]] = 1;
    }
    for (int i = 0; i < nb; i++) {
      if (sizes[i] < 0) bad = 1;
      s += sizes[i];
    }
    free(seen);
    if (bad || s != n) { free(o); return -1; }
  }
  o->block_start = (int *)calloc(nb + 2, sizeof(int));
  o->block_size = (int *)calloc(nb + 2, sizeof(int));
  o->who = (int *)calloc(nb + 2, sizeof(int));
  o->level = (int *)calloc(nb + 2, sizeof(int));
  {
    int *bo = (int *)malloc(sizeof(int) * (nb + 1)), cnt = 0, cur = 0;
    dfs_block_order(bo, &cnt, 1, np);
    for (int t = 0; t < cnt; t++) {
      int e = bo[t];
      o->block_start[e] = cur;
      o->block_size[e] = sizes[np + np - 1 - e];
      cur += o->block_size[e];
    }
    free(bo);
  }
  for (int e = 1; e <= nb; e++) { o->who[e] = who_of(np, e); o->level[e] = ilog2(e); }
  o->pa_rowptr = (int *)malloc(sizeof(int) * (n + 1));
  o->pa_colidx = (int *)malloc(sizeof(int) * (nnz + 1));
  o->perm_map = (int *)malloc(sizeof(int) * (nnz + 1));
  permutate(o, rowptr, colidx, order);
  o->order = (int *)malloc(sizeof(int) * (n + 1));
  memcpy(o->order, order, sizeof(int) * n);
  for (int r = 0; r < n; r++) /* duplicate (r, c) entries are rejected */
    for (int p = o->pa_rowptr[r] + 1; p < o->pa_rowptr[r + 1]; p++)
      if (o->pa_colidx[p] == o->pa_colidx[p - 1]) {
        free(o->order); free(o->pa_rowptr); free(o->pa_colidx);
        free(o->perm_map); free(o->block_start); free(o->block_size);
        free(o->who); free(o->level); free(o);
        return -4;
      }
  int ng = (nb + 1) * o->nk * 2;
  o->gnnz = (int *)calloc(ng, sizeof(int));
  o->goff = (int *)calloc(ng, sizeof(int));
  o->mstart = (int *)calloc(nb + 2, sizeof(int));
  o->msize = (int *)calloc(nb + 2, sizeof(int));
  o->coo_r = (int *)malloc(sizeof(int) * (nnz + 1));
  o->coo_c = (int *)malloc(sizeof(int) * (nnz + 1));
  o->coo_val = (double *)malloc(sizeof(double) * (nnz + 1));
  int rc = construct(o);
  if (rc) { o->leaves = NULL; oracle_destroy(o); return rc; }
  for (int i = 0; i < nnz; i++) o->coo_val[o->perm_map[i]] = val[i]; /* :177 */
  o->leaves = (oleaf_t *)calloc(np, sizeof(oleaf_t));
  for (int r = 0; r < np; r++) {
    oleaf_t *L = &o->leaves[r];
    L->rank = r;
    L->leaf = np + np - 1 - r; /* SnuMat_distribute_structure.cu:142 */
    rc = core_preprocess(o, L);
    if (rc) { oracle_destroy(o); return rc; }
  }
  o->dblk = (double **)calloc(ng, sizeof(double *));
  *out = o;
  return 0;
}

static void free_leaf(oleaf_t *L) {
  free(L->path); free(L->cstart); free(L->core_rowptr); free(L->core_colidx);
  free(L->core_map); free(L->LU_rowptr); free(L->LU_colidx); free(L->LU_diag);
  free(L->LU_bias); free(L->LU_colptr_trans); free(L->LU_rowidx_trans);
  free(L->LU_trans_map); free(L->LU_map); free(L->LU_data); free(L->brow);
  free(L->bcol_inv); free(L->bcol); free(L->S); free(L->bl);
}

static void free_dense(struct oracle *o) {
  if (!o->dblk) return;
  int ng = (o->num_block + 1) * o->nk * 2;
  for (int g = 0; g < ng; g++) { free(o->dblk[g]); o->dblk[g] = NULL; }
}

void oracle_destroy(oracle_t *o) {
  if (!o) return;
  if (o->leaves)
    for (int r = 0; r < o->np; r++) free_leaf(&o->leaves[r]);
  free(o->leaves);
  free_dense(o);
  free(o->dblk);
  free(o->block_start); free(o->block_size); free(o->who); free(o->level);
  free(o->order); free(o->pa_rowptr); free(o->pa_colidx); free(o->perm_map);
  free(o->gnnz); free(o->goff); free(o->mstart); free(o->msize);
  free(o->coo_r); free(o->coo_c); free(o->coo_val);
  free(o);
}

/* ------------------------------------------------------------------------ */
/* Leaf-parallel driver (one host thread stands in for one rank)            */
/* ------------------------------------------------------------------------ */
typedef struct {
  struct oracle *o;
  const double *pb;
  int next, mode;
  pthread_mutex_t mu;
} pool_t;

static void leaf_numeric(struct oracle *o, oleaf_t *L, const double *pb,
                         int mode) {
  leaf_init_values(o, L);
  core_numfact_v2(L);
  if (mode == 0) return;
  /* core-local b: leaf rows from perm_b, separator rows start at 0 */
  int ls = L->leaf_size, bs = o->block_start[L->leaf];
  for (int r = 0; r < ls; r++) L->bl[r] = pb[bs + r];
  for (int r = ls; r < L->core_n; r++) L->bl[r] = 0.0;
  core_trsm(L, L->bl);
  core_update_b(L, L->bl);
  core_GEMM(L);
}

static void *pool_worker(void *arg) {
  pool_t *P = (pool_t *)arg;
  for (;;) {
    pthread_mutex_lock(&P->mu);
    int r = P->next++;
    pthread_mutex_unlock(&P->mu);
    if (r >= P->o->np) break;
    leaf_numeric(P->o, &P->o->leaves[r], P->pb, P->mode);
  }
  return NULL;
}

static void run_leaves(struct oracle *o, const double *pb, int mode,
                       int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > o->np) nthreads = o->np;
  pool_t P;
  P.o = o; P.pb = pb; P.next = 0; P.mode = mode;
  pthread_mutex_init(&P.mu, NULL);
  if (nthreads == 1) {
    pool_worker(&P);
  } else {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, pool_worker, &P);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
  }
  pthread_mutex_destroy(&P.mu);
}

void oracle_leaf_init(oracle_t *o) {
  for (int rk = 0; rk < o->np; rk++) leaf_init_values(o, &o->leaves[rk]);
}

int oracle_leaf_factor(oracle_t *o, int nthreads) {
  run_leaves(o, NULL, 0, nthreads);
  return 0;
}

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* single-copy dense block of separator pair (i, j); rows x cols, col-major */
static double *blk(struct oracle *o, int i, int j) {
  int g = grid_id(o, i, j);
  if (!o->dblk[g]) {
    size_t sz = (size_t)o->block_size[i] * o->block_size[j];
    o->dblk[g] = (double *)calloc(sz ? sz : 1, sizeof(double));
  }
  return o->dblk[g];
}


/* Target of a separator-separator entry (R, C) of PAP^T (global permuted
 * indices, both in separator blocks on one root path): the single-copy dense
 * block that holds it and the column-major offset inside it. */
static double *sep_target(struct oracle *o, int bi, int bj, int lr, int lc) {
  double *B = blk(o, bi, bj);
  return B + lr + (size_t)lc * o->block_size[bi];
}

static int factor_solve_stage(oracle_t *o, const double *b, double *x,
                              int nthreads, int stage);
int oracle_factor_solve(oracle_t *o, const double *b, double *x,
                        int nthreads) {
  return factor_solve_stage(o, b, x, nthreads, 2);
}
int oracle_assemble_only(oracle_t *o, const double *b, int nthreads) {
  return factor_solve_stage(o, b, NULL, nthreads, 1);
}
void oracle_sep_block(oracle_t *o, int i, int j, double *out) {
  size_t sz = (size_t)o->block_size[i] * o->block_size[j];
  memcpy(out, blk(o, i, j), sizeof(double) * sz);
}
static int factor_solve_stage(oracle_t *o, const double *b, double *x,
                              int nthreads, int stage) {
  load_backend();
  int n = o->n, np = o->np, nb = o->num_block;
  double t0 = now_s();
  double *pb = (double *)malloc(sizeof(double) * (n + 1));
  for (int i = 0; i < n; i++) pb[o->order[i]] = b[i]; /* construct.cu:183-186 */

  /* core_run for every leaf: numfact_v2, core_trsm, core_update_b, core_GEMM
   * (kernel_sparse.cu:1011-1037) */
  run_leaves(o, pb, 1, nthreads);
  double t1 = now_s();

  /* Assembly of the separator blocks: A_SS entries (set_LU,
   * SnuMat_distribute_data.cu:19-56) then the leaf Schur contributions
   * (core_reduction + core_update, kernel_sparse.cu:54-90,979-994) and the
   * separator rows of each leaf's b (core_update_b result), in rank order. */
  free_dense(o);
  for (int i = 1; i < np; i++) { /* blocks of every separator exist */
    blk(o, i, i);
    for (int j = i / 2; j; j /= 2) { blk(o, i, j); blk(o, j, i); }
  }
  for (int i = np - 1; i >= 1; i--) {
    int pairs[64][2], np_ = 0;
    pairs[np_][0] = i; pairs[np_][1] = i; np_++;
    for (int ii = i / 2; ii; ii /= 2) {
      pairs[np_][0] = i; pairs[np_][1] = ii; np_++;
      pairs[np_][0] = ii; pairs[np_][1] = i; np_++;
    }
    for (int t = 0; t < np_; t++) {
      int bi = pairs[t][0], bj = pairs[t][1];
      int g = grid_id(o, bi, bj);
      for (int p = 0; p < o->gnnz[g]; p++) {
        int q = o->goff[g] + p;
        *sep_target(o, bi, bj, o->coo_r[q] - o->block_start[bi],
                    o->coo_c[q] - o->block_start[bj]) += o->coo_val[q];
      }
    }
  }
  for (int rk = 0; rk < np; rk++) {
    oleaf_t *L = &o->leaves[rk];
    int ls = L->leaf_size;
    /* map core row -> (block, local) */
    for (int ri = 0; ri < L->nbr; ri++) {
      int r = L->brow[ri];
      int kr = 1;
      while (r >= L->cstart[kr + 1]) kr++;
      int bi = L->path[kr], lr = r - L->cstart[kr];
      const double *Srow = L->S + (size_t)ri * L->nbc;
      for (int cj = 0; cj < L->nbc; cj++) {
        int c = L->bcol[cj];
        int kc = 1;
        while (c >= L->cstart[kc + 1]) kc++;
        int bj = L->path[kc], lc = c - L->cstart[kc];
        *sep_target(o, bi, bj, lr, lc) += Srow[cj];
      }
    }
    int bsl = o->block_start[L->leaf];
    for (int r = 0; r < ls; r++) pb[bsl + r] = L->bl[r];
    for (int k = 1; k < L->npath; k++) {
      int bk = L->path[k], gs = o->block_start[bk];
      for (int r = L->cstart[k]; r < L->cstart[k + 1]; r++)
        pb[gs + r - L->cstart[k]] += L->bl[r];
    }
    free(L->S);
    L->S = NULL;
  }
  double t2 = now_s();
  if (stage == 1) {
    free(pb);
    return 0;
  }

  /* Dense levels bottom-up, MKL path  SnuMat_factorize.cu:221-244 */
  for (int l = o->max_level - 1; l >= 0; l--) {
    for (int i = 1 << l; i < (2 << l) && i < np; i++) {
      int si = o->block_size[i];
      double *Dii = blk(o, i, i);
      double *bi = pb + o->block_start[i];
      d_lu(si, si, Dii);
      d_trsm_Lxb(Dii, si, 1, bi);
      for (int j = i / 2; j; j /= 2) {
        d_trsm_xUb(Dii, o->block_size[j], si, blk(o, j, i));
        d_trsm_Lxb(Dii, si, o->block_size[j], blk(o, i, j));
      }
      for (int j1 = i / 2; j1; j1 /= 2)
        for (int j2 = i / 2; j2; j2 /= 2)
          d_gemm(o->block_size[j1], si, o->block_size[j2], blk(o, j1, i),
                 blk(o, i, j2), blk(o, j1, j2));
      for (int j = i / 2; j; j /= 2)
        d_gemm(o->block_size[j], si, 1, blk(o, j, i), bi,
               pb + o->block_start[j]);
    }
  }
  double t3 = now_s();
  /* Dense backward top-down  SnuMat_factorize.cu:300-310 */
  for (int l = 0; l < o->max_level; l++) {
    for (int i = 1 << l; i < (2 << l) && i < np; i++) {
      int si = o->block_size[i];
      double *bi = pb + o->block_start[i];
      for (int j = i / 2; j >= 1; j /= 2)
        d_gemm(si, o->block_size[j], 1, blk(o, i, j), pb + o->block_start[j],
               bi);
      d_trsm_Uxb(blk(o, i, i), si, 1, bi);
    }
  }
  /* Leaf backward  SnuMat_factorize.cu:312-345 */
  for (int rk = 0; rk < np; rk++) {
    oleaf_t *L = &o->leaves[rk];
    int ls = L->leaf_size;
    double *bl = L->bl;
    int bsl = o->block_start[L->leaf];
    for (int r = 0; r < ls; r++) bl[r] = pb[bsl + r];
    for (int k = 1; k < L->npath; k++) {
      int gs = o->block_start[L->path[k]];
      for (int r = L->cstart[k]; r < L->cstart[k + 1]; r++)
        bl[r] = pb[gs + r - L->cstart[k]];
    }
    for (int r = 0; r < ls; r++)
      for (int p = L->LU_rowptr[r + 1] - 1; p >= L->LU_bias[r]; p--)
        bl[r] -= L->LU_data[p] * bl[L->LU_colidx[p]];
    for (int r = ls - 1; r >= 0; r--) {
      for (int p = L->LU_diag[r] + 1; p < L->LU_bias[r]; p++)
        bl[r] -= L->LU_data[p] * bl[L->LU_colidx[p]];
      bl[r] /= L->LU_data[L->LU_diag[r]];
    }
    for (int r = 0; r < ls; r++) pb[bsl + r] = bl[r];
  }
  /* gather + un-permute  SnuMat_factorize.cu:350-353 */
  for (int i = 0; i < n; i++) x[i] = pb[o->order[i]];
  double t4 = now_s();
  free(pb);
  g_times[0] = t1 - t0;
  g_times[1] = t2 - t1;
  g_times[2] = t3 - t2;
  g_times[3] = t4 - t3;
  (void)nb;
  return 0;
}

/* ---- getters ---- */
int oracle_num_block(const oracle_t *o) { return o->num_block; }
void oracle_blocks(const oracle_t *o, int *block_start, int *block_size) {
  for (int e = 1; e <= o->num_block; e++) {
    block_start[e - 1] = o->block_start[e];
    block_size[e - 1] = o->block_size[e];
  }
}
void oracle_permuted(const oracle_t *o, int *pa_rowptr, int *pa_colidx,
                     int *perm_map) {
  memcpy(pa_rowptr, o->pa_rowptr, sizeof(int) * (o->n + 1));
  memcpy(pa_colidx, o->pa_colidx, sizeof(int) * o->nnz);
  memcpy(perm_map, o->perm_map, sizeof(int) * o->nnz);
}
void oracle_leaf_sizes(const oracle_t *o, int rank, int *leaf_size,
                       int *core_n, int *core_nnz, int *lu_nnz) {
  const oleaf_t *L = &o->leaves[rank];
  *leaf_size = L->leaf_size;
  *core_n = L->core_n;
  *core_nnz = L->core_nnz;
  *lu_nnz = L->LU_nnz;
}
void oracle_leaf_struct(const oracle_t *o, int rank, int *lu_rowptr,
                        int *lu_colidx, int *lu_diag, int *lu_bias,
                        int *core_map) {
  const oleaf_t *L = &o->leaves[rank];
  memcpy(lu_rowptr, L->LU_rowptr, sizeof(int) * (L->core_n + 1));
  memcpy(lu_colidx, L->LU_colidx, sizeof(int) * L->LU_nnz);
  memcpy(lu_diag, L->LU_diag, sizeof(int) * L->core_n);
  memcpy(lu_bias, L->LU_bias, sizeof(int) * L->leaf_size);
  memcpy(core_map, L->core_map, sizeof(int) * L->core_nnz);
}
void oracle_leaf_values(const oracle_t *o, int rank, double *lu_data) {
  const oleaf_t *L = &o->leaves[rank];
  memcpy(lu_data, L->LU_data, sizeof(double) * L->LU_nnz);
}
