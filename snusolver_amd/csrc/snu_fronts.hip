// Kernels of the multifrontal leaf factorization (gfx950, wave64).
//
// A leaf's no-pivot LU (the reference's core_numfact_v2 and core_GEMM,
// src/kernel_sparse.cu:740-781,953-971) runs as dense partial factorizations
// of frontal matrices, level by level of the leaves' front trees, every leaf
// of every matrix of a launch set sharing each launch:
//   small fronts (the whole front fits the LDS): k_front_small does
//     everything in one workgroup per front;
//   larger fronts live in global memory: k_front_assemble, then the blocked
//     partial LU on the dense kernels (k_diag_lu with the pivot clamp, the
//     triangular solves and the update on the MFMA k_gemm), then
//     k_front_writeback;
//   k_front_schur gathers the root fronts' updates into each leaf's boundary
//     Schur block S, which k_assemble_schur adds into the separator blocks.
// Every front position receives its contributions in a fixed order (owned
// slots, then the children in ascending order), so results do not depend on
// scheduling.
#include <hip/hip_runtime.h>

#include "snu_kernels.h"

namespace snu {

extern __shared__ __align__(16) char front_lds[];

// One workgroup per front, the whole front in LDS (f^2 values).
template <typename T>
__global__ __launch_bounds__(256) void k_front_small(const FrontDev *__restrict__ fr, T eps) {
  T *F = (T *)front_lds;
  const FrontDev D = fr[blockIdx.x];
  const int f = D.f, nk = D.nk, n2 = f * f, tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < n2; i += nt) F[i] = T(0);
  __syncthreads();
  const T *vin = (const T *)D.vals;
  for (int q = tid; q < D.nslot; q += nt) F[D.pos[q]] = vin[D.slot[q]];
  __syncthreads();
  // children's updates, one child after another (fixed summation order)
  // (lane = row within 64-row groups, wave = column modulo the waves)
  const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  for (int c = 0; c < D.nchild; c++) {
    const FrontChild C = D.child[c];
    const T *U = (const T *)C.U;
    const int nb = C.nb;
    for (int b = wv; b < nb; b += nw) {
      const int cb = C.rel[b] * f;
      for (int a = lane; a < nb; a += 64) F[C.rel[a] + cb] += U[a + (long long)b * C.ld];
    }
    __syncthreads();
  }
  // partial LU of the nk pivots (right-looking; each entry receives its
  // updates in ascending pivot order, as in the reference's elimination)
  for (int p = 0; p < nk; p++) {
    T d = F[p + p * f];
    if (d < eps && d > -eps) d = d < T(0) ? -eps : eps;
    const int m = f - p - 1;
    T *col = F + p * f + p + 1;
    for (int i = tid; i < m; i += nt) col[i] = col[i] / d;
    if (tid == 0) F[p + p * f] = d;
    __syncthreads();
    for (int c = p + 1 + wv; c < f; c += nw) {
      const T u = F[p + c * f];
      T *dst = F + c * f + p + 1;
      for (int r = lane; r < m; r += 64) dst[r] -= col[r] * u;
    }
    __syncthreads();
  }
  // factors into the owned slots, the update block to global memory
  T *vout = (T *)D.vals;
  for (int q = tid; q < D.nslot; q += nt) vout[D.slot[q]] = F[D.pos[q]];
  T *G = (T *)D.F;
  for (int c = nk + wv; c < f; c += nw)
    for (int r = nk + lane; r < f; r += 64) G[r + (long long)c * f] = F[r + c * f];
}

// Largest front factored whole in LDS (order): above it a front goes
// through HBM and the dense kernels.  Measured 64 / 80 / 96 / 128 (FP64) on
// the 64^3, 96^3 and batch workloads: smaller LDS fronts keep more
// workgroups per CU, and the larger ones are faster on MFMA.
template <typename T>
int front_small_max() {
  return sizeof(T) == 8 ? 64 : 96;
}

template <typename T>
void launch_front_small(const FrontDev *fr, int nfront, int max_f, double eps, hipStream_t st) {
  if (nfront <= 0) return;
  const size_t lds = (size_t)max_f * max_f * sizeof(T);
  (void)hipFuncSetAttribute((const void *)k_front_small<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
  hipLaunchKernelGGL(k_front_small<T>, dim3(nfront), dim3(256), lds, st, fr, (T)eps);
}

// Columns [c0, c0 + 64) of one front in global memory: zeroed, the owned
// slots placed, the children's update columns that land there added (child
// after child).  Within one workgroup the phases are separated by barriers;
// different work items touch different columns.  The work item's ranges
// (its owned slots, each child's columns in the chunk) come from the host.
template <typename T>
__global__ __launch_bounds__(256) void k_front_assemble(const FrontDev *__restrict__ fr,
                                                        const int4 *__restrict__ work,
                                                        const int *__restrict__ wrange) {
  const int4 w = work[blockIdx.x];
  const FrontDev D = fr[w.x];
  const int *rg = wrange + w.z;
  const int f = D.f, c0 = w.y, c1 = min(f, c0 + 64), tid = threadIdx.x, nt = blockDim.x;
  T *F = (T *)D.F;
  const long long base = (long long)c0 * f, cnt = (long long)(c1 - c0) * f;
  for (long long i = tid; i < cnt; i += nt) F[base + i] = T(0);
  __syncthreads();
  const T *vin = (const T *)D.vals;
  const int q0 = rg[0], q1 = rg[1];
  for (int q = q0 + tid; q < q1; q += nt) F[D.pos[q]] = vin[D.slot[q]];
  __syncthreads();
  // children: wave w takes the child's columns b0 + w, b0 + w + 4, ...,
  // its lanes the rows (reads of U run down its columns)
  const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  for (int c = 0; c < D.nchild; c++) {
    const int b0 = rg[2 + 2 * c], b1 = rg[3 + 2 * c];
    if (b0 == b1) continue;  // uniform: no column of this child lands here
    const FrontChild C = D.child[c];
    const T *U = (const T *)C.U;
    const int nb = C.nb;
    for (int b = b0 + wv; b < b1; b += nw) {
      T *Fc = F + (long long)C.rel[b] * f;
      const T *Ub = U + (long long)b * C.ld;
      for (int a = lane; a < nb; a += 64) Fc[C.rel[a]] += Ub[a];
    }
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_front_writeback(const FrontDev *__restrict__ fr,
                                                         const int4 *__restrict__ work,
                                                         const int *__restrict__ wrange) {
  const int4 w = work[blockIdx.x];
  const FrontDev D = fr[w.x];
  const T *F = (const T *)D.F;
  T *vout = (T *)D.vals;
  const int q0 = wrange[w.z], q1 = wrange[w.z + 1];
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) vout[D.slot[q]] = F[D.pos[q]];
}

template <typename T>
void launch_front_assemble(const FrontDev *fr, const int4 *work, const int *wrange, int nwork, hipStream_t st) {
  if (nwork <= 0) return;
  hipLaunchKernelGGL(k_front_assemble<T>, dim3(nwork), dim3(256), 0, st, fr, work, wrange);
}
template <typename T>
void launch_front_writeback(const FrontDev *fr, const int4 *work, const int *wrange, int nwork, hipStream_t st) {
  if (nwork <= 0) return;
  hipLaunchKernelGGL(k_front_writeback<T>, dim3(nwork), dim3(256), 0, st, fr, work, wrange);
}

// One 64 x 64 tile of a leaf's S per workgroup: accumulated column by column
// (reads of the roots' updates run down their columns), then written row by
// row through LDS (S is row-major).
template <typename T>
__global__ __launch_bounds__(256) void k_front_schur(const LeafSchurDev *__restrict__ ls,
                                                     const int4 *__restrict__ tiles) {
  __shared__ T t[64][65];
  const int4 w = tiles[blockIdx.x];
  const LeafSchurDev L = ls[w.x];
  const int r0 = w.y, c0 = w.z, tid = threadIdx.x;
  const int nr = min(64, L.nbr - r0), nc = min(64, L.nbc - c0);
  for (int i = tid; i < 64 * 64; i += 256) t[i & 63][i >> 6] = T(0);
  for (int k = 0; k < L.nroot; k++) {
    const FrontRoot R = L.root[k];
    const T *U = (const T *)R.U;
    for (int i = tid; i < 64 * 64; i += 256) {
      const int r = i & 63, c = i >> 6;
      if (r < nr && c < nc) {
        const int a = R.arow[r0 + r], b = R.acol[c0 + c];
        if (a >= 0 && b >= 0) t[r][c] += U[a + (long long)b * R.ld];
      }
    }
  }
  __syncthreads();
  T *S = (T *)L.S;
  for (int i = tid; i < 64 * 64; i += 256) {
    const int r = i >> 6, c = i & 63;
    if (r < nr && c < nc) S[(long long)(r0 + r) * L.nbc + c0 + c] = t[r][c];
  }
}

template <typename T>
void launch_front_schur(const LeafSchurDev *ls, const int4 *tiles, int ntile, hipStream_t st) {
  if (ntile <= 0) return;
  hipLaunchKernelGGL(k_front_schur<T>, dim3(ntile), dim3(256), 0, st, ls, tiles);
}

#define SNU_FRONT_INST(T)                                                                      \
  template int front_small_max<T>();                                                           \
  template void launch_front_small<T>(const FrontDev *, int, int, double, hipStream_t);        \
  template void launch_front_assemble<T>(const FrontDev *, const int4 *, const int *, int, hipStream_t);  \
  template void launch_front_writeback<T>(const FrontDev *, const int4 *, const int *, int, hipStream_t); \
  template void launch_front_schur<T>(const LeafSchurDev *, const int4 *, int, hipStream_t);
SNU_FRONT_INST(double)
SNU_FRONT_INST(float)
#undef SNU_FRONT_INST

}  // namespace snu
