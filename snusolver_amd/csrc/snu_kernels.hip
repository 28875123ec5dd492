// HIP kernels for gfx950 (MI355X): leaf sparse LU, boundary Schur assembly,
// blocked no-pivot dense LU on MFMA, and the level-scheduled solves.
//
// Wave size 64 throughout.  The leaf kernels are bit-exact with the
// reference's core_numfact_v2 (kernel_sparse.cu:740-781): every LU entry
// receives the same sequence of `a - l*u` updates (ascending pivot order, no
// FMA contraction) and the same correctly rounded divisions; the reference's
// right-looking column sweep is evaluated here row by row (up-looking), which
// applies the identical per-entry sequence.
#include <hip/hip_runtime.h>

#include "snu_kernels.h"

namespace snu {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int rl(int v, int l) {
  return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ double rld(double v, int l) {
  long long x = __double_as_longlong(v);
  int lo = rl((int)(x & 0xffffffffLL), l);
  int hi = rl((int)(x >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// elementwise
// ---------------------------------------------------------------------------
__global__ void k_fill(double *p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0;
}
void launch_fill(double *p, long long n, hipStream_t st) {
  if (n <= 0) return;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
}

__global__ void k_scatter(const double *aval, const long long *dest, int nnz,
                          double *arena) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) arena[dest[i]] = aval[i];
}
void launch_scatter(const double *aval, const long long *dest, int nnz,
                    double *arena, hipStream_t st) {
  if (nnz <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3((nnz + 255) / 256), dim3(256), 0, st,
                     aval, dest, nnz, arena);
}

// ---------------------------------------------------------------------------
// leaf factorization
// ---------------------------------------------------------------------------
// Factor one core row i with one wave.  w is the wave's LDS workspace,
// indexed by workspace column (leaf columns, then the leaf's boundary
// columns), all zero on entry and on exit.
//   LEAF_ROW: L part ends at diag[i]; the diagonal is clamped to +-eps
//             (kernel_sparse.cu:742-749).
//   otherwise a separator row: every entry is an L entry, and the updates
//             that land on boundary columns form the row's Schur
//             contribution -sum_k L(r,k) U(k,c) (kernel_sparse.cu:953-971).
template <bool LEAF_ROW>
__device__ void factor_row(int i, const int *__restrict__ rowptr,
                           const int *__restrict__ diag,
                           const int *__restrict__ wcol,
                           double *__restrict__ val, double *w, int lane,
                           double eps) {
#pragma clang fp contract(off)
  const int p0 = rowptr[i], p1 = rowptr[i + 1];
  const int pd = LEAF_ROW ? diag[i] : p1;
  for (int p = p0 + lane; p < p1; p += 64) w[wcol[p]] = val[p];
  wave_lds_fence();
  for (int pb = p0; pb < pd; pb += 64) {
    const int cnt = min(64, pd - pb);
    int k = 0, ks = 0, ke = 0;
    double piv = 1.0;
    if (lane < cnt) {
      k = wcol[pb + lane];
      int kd = diag[k];
      ks = kd + 1;
      ke = rowptr[k + 1];
      piv = val[kd];
    }
    int cn = 0;
    double un = 0.0;
    {
      int s0 = rl(ks, 0), e0 = rl(ke, 0), q = s0 + lane;
      if (q < e0) cn = wcol[q], un = val[q];
    }
    for (int t = 0; t < cnt; t++) {
      const int kt = rl(k, t), st = rl(ks, t), et = rl(ke, t);
      const double pt = rld(piv, t);
      const int c = cn;
      const double u = un;
      if (t + 1 < cnt) {  // prefetch the next pivot row's first chunk
        int s1 = rl(ks, t + 1), e1 = rl(ke, t + 1), q1 = s1 + lane;
        cn = 0;
        un = 0.0;
        if (q1 < e1) cn = wcol[q1], un = val[q1];
      }
      const double l = w[kt] / pt;
      int q = st + lane;
      if (q < et) {
        double prod = l * u;
        w[c] = w[c] - prod;
      }
      for (q += 64; q < et; q += 64) {
        int c2 = wcol[q];
        double prod = l * val[q];
        w[c2] = w[c2] - prod;
      }
      if (lane == 0) w[kt] = l;
      wave_lds_fence();
    }
  }
  if (LEAF_ROW) {
    if (lane == 0) {
      double d = w[i];
      if (d < eps && d > -eps) d = d < 0 ? -eps : eps;
      w[i] = d;
    }
    wave_lds_fence();
  }
  for (int p = p0 + lane; p < p1; p += 64) {
    int c = wcol[p];
    val[p] = w[c];
    w[c] = 0.0;
  }
  wave_lds_fence();
}

// One workgroup per leaf; rows of a level are spread over the waves, levels
// are separated by workgroup barriers.
__global__ __launch_bounds__(256) void k_leaf_rows(
    const LeafDev *__restrict__ leaves, const int *__restrict__ rowptr,
    const int *__restrict__ diag, const int *__restrict__ wcol,
    const int *__restrict__ levptr, const int *__restrict__ levrows,
    double *__restrict__ arena, int wsz, double eps) {
  extern __shared__ double lds[];
  const LeafDev L = leaves[blockIdx.x];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int W = blockDim.x >> 6;
  double *w = lds + (size_t)wave * wsz;
  for (int c = lane; c < L.leaf_size + L.nbc; c += 64) w[c] = 0.0;
  wave_lds_fence();
  const int *rp = rowptr + L.rp_off;
  const int *dg = diag + L.rp_off;
  const int *wc = wcol + L.idx_off;
  double *v = arena + L.val_off;
  const int *lp = levptr + L.lev_off;
  const int *lr = levrows + L.levrows_off;
  for (int lev = 0; lev < L.nlev; lev++) {
    for (int t = lp[lev] + wave; t < lp[lev + 1]; t += W)
      factor_row<true>(lr[t], rp, dg, wc, v, w, lane, eps);
    __syncthreads();
  }
}
void launch_leaf_rows(const LeafDev *leaves, int nleaves, const int *rowptr,
                      const int *diag, const int *wcol, const int *levptr,
                      const int *levrows, double *arena, int wsz, int waves,
                      double eps, hipStream_t st) {
  if (nleaves <= 0) return;
  size_t lds = (size_t)wsz * waves * sizeof(double);
  (void)hipFuncSetAttribute((const void *)k_leaf_rows,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_leaf_rows, dim3(nleaves), dim3(64 * waves), lds, st,
                     leaves, rowptr, diag, wcol, levptr, levrows, arena, wsz,
                     eps);
}

// Separator rows of the leaves: each wave factors one boundary row (its L_SL
// row) and writes its Schur contribution row into S.
__global__ __launch_bounds__(256) void k_leaf_sep(
    const LeafDev *__restrict__ leaves, const int *__restrict__ wg_leaf,
    const int *__restrict__ wg_first, const int *__restrict__ rowptr,
    const int *__restrict__ diag, const int *__restrict__ wcol,
    const int *__restrict__ brow, double *__restrict__ arena,
    double *__restrict__ S, int wsz) {
  extern __shared__ double lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const LeafDev L = leaves[wg_leaf[blockIdx.x]];
  const int ri = wg_first[blockIdx.x] + wave;
  if (ri >= L.nbr) return;
  double *w = lds + (size_t)wave * wsz;
  const int ls = L.leaf_size;
  for (int c = lane; c < ls + L.nbc; c += 64) w[c] = 0.0;
  wave_lds_fence();
  const int r = brow[L.brow_off + ri];
  factor_row<false>(r, rowptr + L.rp_off, diag + L.rp_off, wcol + L.idx_off,
                    arena + L.val_off, w, lane, 0.0);
  double *Srow = S + L.S_off + (long long)ri * L.nbc;
  for (int j = lane; j < L.nbc; j += 64) {
    Srow[j] = w[ls + j];
    w[ls + j] = 0.0;
  }
}
void launch_leaf_sep(const LeafDev *leaves, const int *wg_leaf,
                     const int *wg_first, int nwg, const int *rowptr,
                     const int *diag, const int *wcol, const int *brow,
                     double *arena, double *S, int wsz, int waves,
                     hipStream_t st) {
  if (nwg <= 0) return;
  size_t lds = (size_t)wsz * waves * sizeof(double);
  (void)hipFuncSetAttribute((const void *)k_leaf_sep,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_leaf_sep, dim3(nwg), dim3(64 * waves), lds, st, leaves,
                     wg_leaf, wg_first, rowptr, diag, wcol, brow, arena, S,
                     wsz);
}

// Slot of separator entry (bi, lr) x (bj, lc) in the dense arena.
__device__ __forceinline__ long long sep_index(const DenseDev &d, int bi,
                                               int lr, int bj, int lc) {
  if (bi == bj) return d.offR[bi] + lr + (long long)lc * d.bsize[bi];
  if (bi > bj) {  // row block deeper: U panel of bi
    int dist = (31 - __clz(bi)) - (31 - __clz(bj));
    int off = d.aoff[bi * d.nk + dist];
    return d.offR[bi] + lr + (long long)(d.bsize[bi] + off + lc) * d.bsize[bi];
  }
  int dist = (31 - __clz(bj)) - (31 - __clz(bi));
  int off = d.aoff[bj * d.nk + dist];
  return d.offL[bj] + off + lr + (long long)lc * d.anc_sum[bj];
}

// One wave per target separator row; contributions are added leaf by leaf
// in ascending rank order (deterministic; no atomics).
__global__ __launch_bounds__(256) void k_assemble_schur(
    int ntarget, const int *__restrict__ trow_blk,
    const int *__restrict__ trow_loc, const int *__restrict__ tptr,
    const int *__restrict__ tsrc_leaf, const int *__restrict__ tsrc_ri,
    const LeafDev *__restrict__ leaves, const int *__restrict__ bcol_blk,
    const int *__restrict__ bcol_loc, const double *__restrict__ S,
    double *dense, DenseDev dd) {
  const int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= ntarget) return;
  const int bi = trow_blk[t], lr = trow_loc[t];
  for (int s = tptr[t]; s < tptr[t + 1]; s++) {
    const LeafDev L = leaves[tsrc_leaf[s]];
    const double *Srow = S + L.S_off + (long long)tsrc_ri[s] * L.nbc;
    for (int j = lane; j < L.nbc; j += 64) {
      double v = Srow[j];
      if (v != 0.0) {
        long long idx = sep_index(dd, bi, lr, bcol_blk[L.bcol_off + j],
                                  bcol_loc[L.bcol_off + j]);
        dense[idx] += v;
      }
    }
    __threadfence();
  }
}
void launch_assemble_schur(int ntarget, const int *trow_blk,
                           const int *trow_loc, const int *tptr,
                           const int *tsrc_leaf, const int *tsrc_ri,
                           const LeafDev *leaves, const int *bcol_blk,
                           const int *bcol_loc, const double *S, double *dense,
                           DenseDev dd, hipStream_t st) {
  if (ntarget <= 0) return;
  hipLaunchKernelGGL(k_assemble_schur, dim3((ntarget + 3) / 4), dim3(256), 0,
                     st, ntarget, trow_blk, trow_loc, tptr, tsrc_leaf, tsrc_ri,
                     leaves, bcol_blk, bcol_loc, S, dense, dd);
}

// ---------------------------------------------------------------------------
// dense: grouped GEMM on MFMA f64 16x16x4
// ---------------------------------------------------------------------------
// Workgroup tile 64 x 64, 4 waves each 32 x 32 (2 x 2 MFMA blocks), K chunk
// 16 staged through LDS.  SUB: C -= sum_seg A_seg B_seg (K concatenated over
// the segments, in order).  ASSIGN: C = A B with one segment; C may alias A
// (N <= 64) or B (M <= 64) because each workgroup reads its whole A row
// panel / B column panel before it writes.
// MFMA operands are swapped (A-operand <- B tile, B-operand <- A tile) so
// that the 16 lanes of a group hold 16 consecutive rows of one column of C
// (coalesced epilogue).  v_mfma_f64_16x16x4f64 layouts on gfx950 (checked
// by a lane-layout probe): A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D[i][j] in lane j + 16 (i % 4), register i / 4.
constexpr int TB = 64, KB = 16, PADA = 4;
template <bool ASSIGN>
__global__ __launch_bounds__(256) void k_gemm(const GemmProb *__restrict__ probs,
                                              int nprob,
                                              const GemmSeg *__restrict__ segs) {
  __shared__ double As[KB][TB + PADA];
  __shared__ double Bs[KB][TB + PADA];
  const int tile = blockIdx.x;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemmProb P = probs[lo];
  const int lt = tile - P.tile0;
  const int m0 = (lt / P.tiles_n) * TB, n0 = (lt % P.tiles_n) * TB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int l16 = lane & 15, lk = lane >> 4;
  d4 acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
  for (int sg = 0; sg < P.nseg; sg++) {
    const GemmSeg S = segs[P.seg0 + sg];
    for (int k0 = 0; k0 < S.K; k0 += KB) {
#pragma unroll
      for (int e = tid; e < TB * KB; e += 256) {
        int m = e & 63, k = e >> 6, gm = m0 + m, gk = k0 + k;
        As[k][m] = (gm < P.M && gk < S.K) ? S.A[gm + (long long)gk * S.lda] : 0.0;
      }
#pragma unroll
      for (int e = tid; e < TB * KB; e += 256) {
        int k = e & 15, n = e >> 4, gk = k0 + k, gn = n0 + n;
        Bs[k][n] = (gk < S.K && gn < P.N) ? S.B[gk + (long long)gn * S.ldb] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < KB; kk += 4) {
        double av[2], bv[2];
        for (int b = 0; b < 2; b++) bv[b] = Bs[kk + lk][wn + b * 16 + l16];
        for (int a = 0; a < 2; a++) av[a] = As[kk + lk][wm + a * 16 + l16];
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[b], av[a],
                                                             acc[a][b], 0, 0, 0);
      }
      __syncthreads();
    }
  }
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int r = 0; r < 4; r++) {
        int m = m0 + wm + a * 16 + l16;
        int n = n0 + wn + b * 16 + lk + 4 * r;
        if (m < P.M && n < P.N) {
          double *c = P.C + m + (long long)n * P.ldc;
          if (ASSIGN) *c = acc[a][b][r];
          else *c = *c - acc[a][b][r];
        }
      }
}
void launch_gemm(const GemmProb *probs, int nprob, const GemmSeg *segs,
                 int total_tiles, bool assign, hipStream_t st) {
  if (total_tiles <= 0 || nprob <= 0) return;
  if (assign)
    hipLaunchKernelGGL(k_gemm<true>, dim3(total_tiles), dim3(256), 0, st,
                       probs, nprob, segs);
  else
    hipLaunchKernelGGL(k_gemm<false>, dim3(total_tiles), dim3(256), 0, st,
                       probs, nprob, segs);
}

// ---------------------------------------------------------------------------
// dense: diagonal block LU (no pivoting, as cusolverDnDgetrf with a null
// ipiv / LAPACKE_mkl_dgetrfnp) and the inverses of its triangular factors.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag_lu(const DiagProb *__restrict__ probs) {
  __shared__ double a[64 * 64];   // col-major, ld 64
  __shared__ double xs[64 * 64];  // inverse being built, col-major
  const DiagProb P = probs[blockIdx.x];
  const int nb = P.nb, tid = threadIdx.x;
  for (int e = tid; e < 64 * 64; e += 256) {
    int r = e & 63, c = e >> 6;
    a[e] = (r < nb && c < nb) ? P.D[r + (long long)c * P.ld] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    const double piv = a[j + j * 64];
    for (int r = j + 1 + tid; r < nb; r += 256) a[r + j * 64] /= piv;
    __syncthreads();
    const int w = nb - j - 1;
    for (int e = tid; e < w * w; e += 256) {
      int r = j + 1 + e % w, c = j + 1 + e / w;
      a[r + c * 64] -= a[r + j * 64] * a[j + c * 64];
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += 256) {
    int r = e % nb, c = e / nb;
    P.D[r + (long long)c * P.ld] = a[r + c * 64];
  }
  // U^-1, one column per thread
  if (tid < 64) {
    const int c = tid;
    double *x = xs + c * 64;
    for (int r = 0; r < 64; r++) x[r] = 0.0;
    if (c < nb) {
      x[c] = 1.0 / a[c + c * 64];
      for (int r = c - 1; r >= 0; r--) {
        double s = 0.0;
        for (int t = r + 1; t <= c; t++) s += a[r + t * 64] * x[t];
        x[r] = -s / a[r + r * 64];
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < 64 * 64; e += 256) P.Uinv[e] = xs[e];
  __syncthreads();
  // L^-1 (unit lower)
  if (tid < 64) {
    const int c = tid;
    double *x = xs + c * 64;
    for (int r = 0; r < 64; r++) x[r] = 0.0;
    if (c < nb) {
      x[c] = 1.0;
      for (int r = c + 1; r < nb; r++) {
        double s = 0.0;
        for (int t = c; t < r; t++) s += a[r + t * 64] * x[t];
        x[r] = -s;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < 64 * 64; e += 256) P.Linv[e] = xs[e];
}
void launch_diag_lu(const DiagProb *probs, int nprob, hipStream_t st) {
  if (nprob <= 0) return;
  hipLaunchKernelGGL(k_diag_lu, dim3(nprob), dim3(256), 0, st, probs);
}

// ---------------------------------------------------------------------------
// dense triangular solves of one separator's diagonal block (one workgroup
// per separator), blocked by 64 with the stored diagonal-block inverses.
// ---------------------------------------------------------------------------
constexpr int TRI_THREADS = 512;
__global__ __launch_bounds__(TRI_THREADS) void k_tri_fwd(const TriProb *__restrict__ probs,
                                                         int nrhs) {
  __shared__ double y[64];
  const TriProb P = probs[blockIdx.x];
  const int tid = threadIdx.x;
  for (int col = 0; col < nrhs; col++) {
    double *b = P.b + (long long)col * P.ldb;
    for (int k0 = 0; k0 < P.s; k0 += 64) {
      const int nbk = min(64, P.s - k0);
      const double *Li = P.inv + (long long)(k0 / 64) * 8192;
      if (tid < nbk) {
        double s = 0.0;
        for (int c = 0; c <= tid; c++) s += Li[tid + c * 64] * b[k0 + c];
        y[tid] = s;
      }
      __syncthreads();
      if (tid < nbk) b[k0 + tid] = y[tid];
      for (int r = k0 + nbk + tid; r < P.s; r += TRI_THREADS) {
        double s = 0.0;
        const double *Dr = P.D + r + (long long)k0 * P.ld;
        for (int c = 0; c < nbk; c++) s += Dr[(long long)c * P.ld] * y[c];
        b[r] -= s;
      }
      __syncthreads();
    }
  }
}
__global__ __launch_bounds__(TRI_THREADS) void k_tri_bwd(const TriProb *__restrict__ probs,
                                                         int nrhs) {
  __shared__ double y[64];
  const TriProb P = probs[blockIdx.x];
  const int tid = threadIdx.x;
  const int nblk = (P.s + 63) / 64;
  for (int col = 0; col < nrhs; col++) {
    double *b = P.b + (long long)col * P.ldb;
    for (int kb = nblk - 1; kb >= 0; kb--) {
      const int k0 = kb * 64, nbk = min(64, P.s - k0);
      const double *Ui = P.inv + (long long)kb * 8192 + 4096;
      if (tid < nbk) {
        double s = 0.0;
        for (int c = tid; c < nbk; c++) s += Ui[tid + c * 64] * b[k0 + c];
        y[tid] = s;
      }
      __syncthreads();
      if (tid < nbk) b[k0 + tid] = y[tid];
      for (int r = tid; r < k0; r += TRI_THREADS) {
        double s = 0.0;
        const double *Dr = P.D + r + (long long)k0 * P.ld;
        for (int c = 0; c < nbk; c++) s += Dr[(long long)c * P.ld] * y[c];
        b[r] -= s;
      }
      __syncthreads();
    }
  }
}
void launch_tri_fwd(const TriProb *probs, int nprob, int nrhs, hipStream_t st) {
  if (nprob <= 0) return;
  hipLaunchKernelGGL(k_tri_fwd, dim3(nprob), dim3(TRI_THREADS), 0, st, probs,
                     nrhs);
}
void launch_tri_bwd(const TriProb *probs, int nprob, int nrhs, hipStream_t st) {
  if (nprob <= 0) return;
  hipLaunchKernelGGL(k_tri_bwd, dim3(nprob), dim3(TRI_THREADS), 0, st, probs,
                     nrhs);
}

// ---------------------------------------------------------------------------
// solve: permutation, leaf triangular solves, boundary rhs assembly
// ---------------------------------------------------------------------------
__global__ void k_permute_in(const double *b, long long ldb, const int *order,
                             double *bp, int n, int nrhs) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * nrhs) return;
  int col = (int)(i / n), r = (int)(i % n);
  bp[order[r] + (long long)col * n] = b[r + col * ldb];
}
__global__ void k_permute_out(const double *bp, const int *order, double *x,
                              long long ldx, int n, int nrhs) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * nrhs) return;
  int col = (int)(i / n), r = (int)(i % n);
  x[r + col * ldx] = bp[order[r] + (long long)col * n];
}
void launch_permute_in(const double *b, long long ldb, const int *order,
                       double *bp, int n, int nrhs, hipStream_t st) {
  long long tot = (long long)n * nrhs;
  hipLaunchKernelGGL(k_permute_in, dim3((unsigned)((tot + 255) / 256)), dim3(256),
                     0, st, b, ldb, order, bp, n, nrhs);
}
void launch_permute_out(const double *bp, const int *order, double *x,
                        long long ldx, int n, int nrhs, hipStream_t st) {
  long long tot = (long long)n * nrhs;
  hipLaunchKernelGGL(k_permute_out, dim3((unsigned)((tot + 255) / 256)), dim3(256),
                     0, st, bp, order, x, ldx, n, nrhs);
}

// y_L = L_LL^-1 b_L (unit lower), level by level, one wave per row.
__global__ __launch_bounds__(256) void k_leaf_fwd(
    const LeafDev *__restrict__ leaves, const int *__restrict__ rowptr,
    const int *__restrict__ diag, const int *__restrict__ wcol,
    const int *__restrict__ levptr, const int *__restrict__ levrows,
    const double *__restrict__ arena, double *bp, int n, int nrhs) {
  const LeafDev L = leaves[blockIdx.x];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int W = blockDim.x >> 6;
  const int *rp = rowptr + L.rp_off;
  const int *dg = diag + L.rp_off;
  const int *wc = wcol + L.idx_off;
  const double *v = arena + L.val_off;
  const int *lp = levptr + L.lev_off;
  const int *lr = levrows + L.levrows_off;
  for (int lev = 0; lev < L.nlev; lev++) {
    for (int t = lp[lev] + wave; t < lp[lev + 1]; t += W) {
      const int i = lr[t];
      for (int col = 0; col < nrhs; col++) {
        double *b = bp + (long long)col * n + L.gstart;
        double s = 0.0;
        for (int p = rp[i] + lane; p < dg[i]; p += 64) s += v[p] * b[wc[p]];
        s = wave_sum(s);
        if (lane == 0) b[i] -= s;
      }
    }
    __syncthreads();
  }
}
void launch_leaf_fwd(const LeafDev *leaves, int nleaves, const int *rowptr,
                     const int *diag, const int *wcol, const int *levptr,
                     const int *levrows, const double *arena, double *bp,
                     int n, int nrhs, hipStream_t st) {
  if (nleaves <= 0) return;
  hipLaunchKernelGGL(k_leaf_fwd, dim3(nleaves), dim3(256), 0, st, leaves,
                     rowptr, diag, wcol, levptr, levrows, arena, bp, n, nrhs);
}

// Boundary rows: cbuf[ri] = sum_k L_SL(r, k) y_k (kernel_sparse.cu:935-944).
__global__ __launch_bounds__(256) void k_leaf_fwd_sep(
    const LeafDev *__restrict__ leaves, const int *__restrict__ wg_leaf,
    const int *__restrict__ wg_first, const int *__restrict__ rowptr,
    const int *__restrict__ wcol, const int *__restrict__ brow,
    const double *__restrict__ arena, const double *__restrict__ bp,
    double *__restrict__ cbuf, int n, int nrhs, int cb_ld) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const LeafDev L = leaves[wg_leaf[blockIdx.x]];
  const int ri = wg_first[blockIdx.x] + wave;
  if (ri >= L.nbr) return;
  const int r = brow[L.brow_off + ri];
  const int *rp = rowptr + L.rp_off;
  const int *wc = wcol + L.idx_off;
  const double *v = arena + L.val_off;
  for (int col = 0; col < nrhs; col++) {
    const double *b = bp + (long long)col * n + L.gstart;
    double s = 0.0;
    for (int p = rp[r] + lane; p < rp[r + 1]; p += 64) s += v[p] * b[wc[p]];
    s = wave_sum(s);
    if (lane == 0) cbuf[L.cb_off + ri + (long long)col * cb_ld] = s;
  }
}
void launch_leaf_fwd_sep(const LeafDev *leaves, const int *wg_leaf,
                         const int *wg_first, int nwg, int waves,
                         const int *rowptr, const int *wcol, const int *brow,
                         const double *arena, const double *bp, double *cbuf,
                         int n, int nrhs, int cb_ld, hipStream_t st) {
  if (nwg <= 0) return;
  hipLaunchKernelGGL(k_leaf_fwd_sep, dim3(nwg), dim3(64 * waves), 0, st,
                     leaves, wg_leaf, wg_first, rowptr, wcol, brow, arena, bp,
                     cbuf, n, nrhs, cb_ld);
}

// b_R -= contributions of the leaves touching row R, in ascending rank.
__global__ void k_assemble_rhs(int ntarget, const int *__restrict__ trow_glob,
                               const int *__restrict__ tptr,
                               const int *__restrict__ tsrc_leaf,
                               const int *__restrict__ tsrc_ri,
                               const LeafDev *__restrict__ leaves,
                               const double *__restrict__ cbuf, int cb_ld,
                               double *bp, int n, int nrhs) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)ntarget * nrhs) return;
  const int t = (int)(i % ntarget), col = (int)(i / ntarget);
  double *b = bp + (long long)col * n + trow_glob[t];
  double acc = *b;
  for (int s = tptr[t]; s < tptr[t + 1]; s++)
    acc -= cbuf[leaves[tsrc_leaf[s]].cb_off + tsrc_ri[s] + (long long)col * cb_ld];
  *b = acc;
}
void launch_assemble_rhs(int ntarget, const int *trow_glob, const int *tptr,
                         const int *tsrc_leaf, const int *tsrc_ri,
                         const LeafDev *leaves, const double *cbuf, int cb_ld,
                         double *bp, int n, int nrhs, hipStream_t st) {
  long long tot = (long long)ntarget * nrhs;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_assemble_rhs, dim3((unsigned)((tot + 255) / 256)),
                     dim3(256), 0, st, ntarget, trow_glob, tptr, tsrc_leaf,
                     tsrc_ri, leaves, cbuf, cb_ld, bp, n, nrhs);
}

// x_L = U_LL^-1 (y_L - U_LS x_S), level by level (SnuMat_factorize.cu:331-344).
__global__ __launch_bounds__(256) void k_leaf_bwd(
    const LeafDev *__restrict__ leaves, const int *__restrict__ rowptr,
    const int *__restrict__ diag, const int *__restrict__ wcol,
    const int *__restrict__ ulevptr, const int *__restrict__ ulevrows,
    const int *__restrict__ bcol_glob, const double *__restrict__ arena,
    double *bp, int n, int nrhs) {
  const LeafDev L = leaves[blockIdx.x];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int W = blockDim.x >> 6;
  const int ls = L.leaf_size;
  const int *rp = rowptr + L.rp_off;
  const int *dg = diag + L.rp_off;
  const int *wc = wcol + L.idx_off;
  const int *bg = bcol_glob + L.bcol_off;
  const double *v = arena + L.val_off;
  const int *lp = ulevptr + L.ulev_off;
  const int *lr = ulevrows + L.ulevrows_off;
  for (int lev = 0; lev < L.nulev; lev++) {
    for (int t = lp[lev] + wave; t < lp[lev + 1]; t += W) {
      const int i = lr[t];
      for (int col = 0; col < nrhs; col++) {
        double *b = bp + (long long)col * n;
        double s = 0.0;
        for (int p = dg[i] + 1 + lane; p < rp[i + 1]; p += 64) {
          int c = wc[p];
          long long g = c < ls ? (long long)L.gstart + c : (long long)bg[c - ls];
          s += v[p] * b[g];
        }
        s = wave_sum(s);
        if (lane == 0) b[L.gstart + i] = (b[L.gstart + i] - s) / v[dg[i]];
      }
    }
    __syncthreads();
  }
}
void launch_leaf_bwd(const LeafDev *leaves, int nleaves, const int *rowptr,
                     const int *diag, const int *wcol, const int *ulevptr,
                     const int *ulevrows, const int *bcol_glob,
                     const double *arena, double *bp, int n, int nrhs,
                     hipStream_t st) {
  if (nleaves <= 0) return;
  hipLaunchKernelGGL(k_leaf_bwd, dim3(nleaves), dim3(256), 0, st, leaves,
                     rowptr, diag, wcol, ulevptr, ulevrows, bcol_glob, arena,
                     bp, n, nrhs);
}

int device_lds_limit() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 65536;
  return (int)prop.sharedMemPerBlock;
}

}  // namespace snu
