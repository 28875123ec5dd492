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

#include <algorithm>
#include <type_traits>

#include "snu_kernels.h"

#include <utility>

namespace snu {

typedef double d4 __attribute__((ext_vector_type(4)));
// Pointers read from launch descriptors are generic; casting them to the
// global address space makes the compiler emit global_load/global_store
// (flat accesses also count against lgkmcnt, which would make every LDS wait
// drain the global prefetch as well).
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ gcdouble *gptr(const double *p) { return (gcdouble *)p; }
__device__ __forceinline__ gdouble *gptr(double *p) { return (gdouble *)p; }
typedef __attribute__((address_space(1))) const float gcfloat;
typedef __attribute__((address_space(1))) float gfloat;
__device__ __forceinline__ gcfloat *gptr(const float *p) { return (gcfloat *)p; }
__device__ __forceinline__ gfloat *gptr(float *p) { return (gfloat *)p; }

__device__ __forceinline__ int rl(int v, int l) {
  return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ double rld(double v, int l) {
  long long x = __double_as_longlong(v);
  int lo = rl((int)(x & 0xffffffffLL), l);
  int hi = rl((int)(x >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float rld(float v, int l) {
  return __int_as_float(rl(__float_as_int(v), l));
}
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}
// Workgroup barrier for LDS-only communication: waits for this wave's LDS
// (and scalar) operations, not for its global loads, which stay in flight
// across the barrier (a __syncthreads would drain them).
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// elementwise
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = T(0);
}
template <typename T>
void launch_fill(T *p, long long n, hipStream_t st) {
  if (n <= 0) return;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
}


// Slot of separator entry (bi, lr) x (bj, lc) in the dense arena, -1 when
// it lies in a panel tile that is not stored.
__device__ __forceinline__ long long sep_index(const DenseDev &d, int bi,
                                               int lr, int bj, int lc) {
  if (bi == bj) return d.offR[bi] + lr + (long long)lc * d.bsize[bi];
  if (bi > bj) {  // row block deeper: U panel of bi
    int dist = (31 - __clz(bi)) - (31 - __clz(bj));
    int c = d.umap[d.mapoff[bi] + d.aoff[bi * d.nk + dist] + lc];
    return c < 0 ? -1 : d.offR[bi] + lr + (long long)(d.bsize[bi] + c) * d.bsize[bi];
  }
  int dist = (31 - __clz(bj)) - (31 - __clz(bi));
  int r = d.lmap[d.mapoff[bj] + d.aoff[bj * d.nk + dist] + lr];
  return r < 0 ? -1 : d.offL[bj] + r + (long long)lc * d.nL[bj];
}

// One workgroup per assembly tile: 64 consecutive rows of one separator
// block.  For each contributing leaf (ascending rank) and each chunk of 64 of
// its boundary columns, the S block is staged through LDS (rows read along
// S's rows) and added into the dense block down its columns, where the 64
// rows of the tile are contiguous: both sides coalesced.  Every element
// receives its leaves' contributions in ascending rank order from one thread
// (deterministic; no atomics); zero contributions are skipped.
constexpr int kAsmWaves = 16;  // waves of an assembly workgroup
template <typename T>
__global__ __launch_bounds__(64 * kAsmWaves) void k_assemble_schur(
    int ntile, const int *__restrict__ tile_blk, const int *__restrict__ tile_row0,
    const int *__restrict__ tile_ptr, const int *__restrict__ seg_leaf,
    const int *__restrict__ seg_ri, const LeafDev *__restrict__ leaves) {
  __shared__ T st[64][65];
  const int tile = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int bi = tile_blk[tile], row0 = tile_row0[tile];
  for (int sg = tile_ptr[tile]; sg < tile_ptr[tile + 1]; sg++) {
    const LeafDev L = leaves[seg_leaf[sg]];
    const DenseDev dd = *L.dd;
    T *dense = (T *)L.dense;
    const T *S = (const T *)L.S;
    const int *ri = seg_ri + (long long)sg * 64;
    for (int c0 = 0; c0 < L.nbc; c0 += 64) {
      // stage: wave w reads tile rows w, w + kAsmWaves, ... (64 columns each)
      for (int rr = wave; rr < 64; rr += kAsmWaves) {
        const int r = ri[rr];
        T v = T(0);
        if (r >= 0 && c0 + lane < L.nbc) v = S[(long long)r * L.nbc + c0 + lane];
        st[rr][lane] = v;
      }
      __syncthreads();
      // add: wave w takes columns w, w + kAsmWaves, ...; lane = tile row.
      // The columns of one leaf go to distinct dense columns, so all targets
      // of a lane are read before any is written (one round trip)
      const int nc = min(64, L.nbc - c0);
      constexpr int NJ = 64 / kAsmWaves;
      // (every structurally nonzero entry of S lies in a stored tile; an
      // index outside them, which the analysis rules out, is skipped)
      T v[NJ], old[NJ];
      long long idx[NJ];
#pragma unroll
      for (int j = 0; j < NJ; j++) {
        const int cc = wave + kAsmWaves * j;
        v[j] = cc < nc ? st[lane][cc] : T(0);
        idx[j] = -1;
        if (v[j] != T(0)) {
          idx[j] = sep_index(dd, bi, row0 + lane, L.bcol_blk[c0 + cc], L.bcol_loc[c0 + cc]);
          if (idx[j] >= 0) old[j] = dense[idx[j]];
        }
      }
#pragma unroll
      for (int j = 0; j < NJ; j++)
        if (idx[j] >= 0) dense[idx[j]] = old[j] + v[j];
      __syncthreads();
    }
  }
}
template <typename T>
void launch_assemble_schur(int ntile, const int *tile_blk, const int *tile_row0,
                           const int *tile_ptr, const int *seg_leaf, const int *seg_ri,
                           const LeafDev *leaves, hipStream_t st) {
  if (ntile <= 0) return;
  hipLaunchKernelGGL(k_assemble_schur<T>, dim3(ntile), dim3(64 * kAsmWaves), 0, st, ntile, tile_blk,
                     tile_row0, tile_ptr, seg_leaf, seg_ri, leaves);
}

// ---------------------------------------------------------------------------
// dense: grouped GEMM on MFMA (FP64 v_mfma_f64_16x16x4f64, FP32
// v_mfma_f32_16x16x4f32)
// ---------------------------------------------------------------------------
// Operands are swapped (first operand <- B tile, second <- A tile): the
// product is D = B^T A^T, so the 16 lanes of a row group hold 16 consecutive
// rows of one column of C (coalesced epilogue).  Lane layouts (checked by
// lane-layout probes and the GEMM parity tests): A[i][k] in lane i + 16k,
// B[k][j] in lane j + 16k; FP64 D[i][j] in lane j + 16 (i mod 4), register
// i / 4; FP32 D[i][j] in lane j + 16 (i / 4), register i mod 4.
//
// Workgroup tile (32 BMB) x (32 BNB), 4 waves in 2 x 2, each (16 BMB) x
// (16 BNB); K chunk 16; double-buffered LDS fed from a register prefetch that
// stays in flight across the MFMA block (one barrier per chunk).
// MODE 0: C -= sum_seg A_seg B_seg (K concatenated over the segments).
// MODE 1 / 2: C = A B / C = -A B (one segment); C may alias A (N <= tile
// width) or B (M <= tile height): a workgroup reads its whole A row panel /
// B column panel before its epilogue writes.
// Workgroup ids are remapped so that each XCD (ids go round-robin over the 8
// XCDs) works on runs of consecutive tiles, which share operand panels in
// that XCD's L2.
// Chunks of kXcdChunk consecutive tiles go to the XCDs in turn (XCD x runs
// chunks x, x + 8, x + 16, ...): neighbouring tiles still share an L2, and
// tiles of unequal K (problems of one launch differ) spread evenly over the
// XCDs instead of one XCD receiving a whole run of long tiles.  The last
// partial round of 8 chunks keeps the identity order.
constexpr int kXcdChunk = 16;
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  const int full = nb / (8 * kXcdChunk) * (8 * kXcdChunk);
  if (bid >= full) return bid;
  const int x = bid & 7, i = bid >> 3;
  return ((i / kXcdChunk) * 8 + x) * kXcdChunk + (i % kXcdChunk);
}
// row_ror within rows of 16 lanes reads a valid lane for every lane, so the
// destination's old value is never kept (no zeroing needed)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, false);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
// DEPTH: chunks of A and B in flight in registers ahead of the chunk being
// multiplied (1: the next chunk's loads overlap one MFMA block; 2: loads are
// issued two chunks ahead, alternating two register stages).
template <typename T, int MODE, int BMB, int BNB, int KB, int DEPTH = 1>
__global__ __launch_bounds__(256) void k_gemm(const GemmProbT<T> *__restrict__ probs,
                                              int nprob,
                                              const GemmSegT<T> *__restrict__ segs) {
  constexpr bool F64 = std::is_same<T, double>::value;
  constexpr int BM = 32 * BMB, BN = 32 * BNB;
  constexpr int NA = BM * KB / 256, NB = BN * KB / 256;
  // LDS tiles k-contiguous per row of A / column of B.  A lane reads GR
  // consecutive k of its row per access: FP32 128 x 128 tile 4 (16 bytes),
  // FP32 64 x 64 2, FP64 64 x 64 2 (16 bytes), FP64 128 x 128 1 (more
  // operands in registers would cost the second wave per SIMD).  Rows are
  // padded by GR elements: the row stride is then an odd multiple of the
  // access width modulo 32 banks, so the accesses of consecutive rows that
  // the LDS serves together (8 of 16 bytes, 16 of 8) hit disjoint banks.
  constexpr int GR = F64 ? (BMB * BNB >= 16 ? 1 : 2) : (BMB * BNB >= 16 ? 4 : 2);
  constexpr int KP = KB + GR;
  __shared__ __align__(16) T As[2][BM][KP];
  __shared__ __align__(16) T Bs[2][BN][KP];
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  int lo = 0, hi = nprob - 1;
  if (nprob == (int)gridDim.x) lo = hi = tile;  // one tile per problem
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemmProbT<T> P = probs[lo];
  // grouped order: consecutive tile ids walk down a strip of GM tile rows
  // before moving right, so the tiles in flight on one XCD share both their
  // A and their B panels in its L2
  constexpr int GM = 8;
  const int lt = tile - P.tile0;
  const int tiles_m = (P.M + BM - 1) / BM;
  const int grp = lt / (GM * P.tiles_n), first = grp * GM;
  const int gsz = min(GM, tiles_m - first), li = lt - grp * GM * P.tiles_n;
  const int m0 = (first + li % gsz) * BM, n0 = (li / gsz) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 16 * BMB, wn = (wave >> 1) * 16 * BNB;
  const int l16 = lane & 15, lk = lane >> 4;
  // one 16x16x4 product per 16 x 16 block, 4 output registers
  f4 accf[BMB][BNB];
  d4 accd[BMB][BNB];
#pragma unroll
  for (int a = 0; a < BMB; a++)
#pragma unroll
    for (int b = 0; b < BNB; b++) {
      accf[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      accd[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    }
  T ra[NA], rb[NB], ra2[NA], rb2[NB];
  // Loads are unconditional and unmasked with addresses clamped into the
  // operands, so they stay in flight across the MFMA block: a clamped row of
  // A (column of B) only feeds a row (column) of the tile that is never
  // stored; the K tail is zeroed in LDS after the store.  Element i of a
  // thread's A chunk is row m0 + tid % BM, column kk0 + NA (tid / BM) + i
  // (NA consecutive k: one vector store into LDS); of its B chunk row
  // kk0 + tid % KB, column n0 + tid / KB + i (256 / KB): on full chunks (no
  // K tail) and full tiles (no N edge), the uniform case, the addresses are
  // one base plus multiples of one stride.
  static_assert(256 % BM == 0 && 256 % KB == 0, "thread-to-element map");
  constexpr int BSTEP = 256 / KB;
  const int am = tid % BM, ak = tid / BM, bk = tid % KB, bn = tid / KB;
  auto load_into = [&](int sgi, int kk0, T (&xa)[NA], T (&xb)[NB]) {
    const GemmSegT<T> S = segs[P.seg0 + sgi];
    const auto pa = gptr(S.A) + min(m0 + am, P.M - 1);
    if (kk0 + KB <= S.K) {
      const auto p = pa + (long long)(kk0 + ak * NA) * S.lda;
#pragma unroll
      for (int i = 0; i < NA; i++) xa[i] = p[i * S.lda];
    } else {
#pragma unroll
      for (int i = 0; i < NA; i++) xa[i] = pa[(long long)min(kk0 + ak * NA + i, S.K - 1) * S.lda];
    }
    const auto pb = gptr(S.B) + min(kk0 + bk, S.K - 1);
    if (n0 + BN <= P.N) {
      const auto p = pb + (long long)(n0 + bn) * S.ldb;
      const long long st = (long long)BSTEP * S.ldb;
#pragma unroll
      for (int i = 0; i < NB; i++) xb[i] = p[i * st];
    } else {
#pragma unroll
      for (int i = 0; i < NB; i++) xb[i] = pb[(long long)min(n0 + bn + i * BSTEP, P.N - 1) * S.ldb];
    }
  };
  auto load = [&](int sgi, int kk0) { load_into(sgi, kk0, ra, rb); };
  // vectors of GR consecutive k
  typedef T vec_t __attribute__((ext_vector_type(GR)));
  static_assert(NA % GR == 0 && (KB / 4) % GR == 0, "vector LDS accesses");
  auto store_from = [&](int buf, int kvalid, const T (&xa)[NA], const T (&xb)[NB]) {
#pragma unroll
    for (int v = 0; v < NA / GR; v++) {
      vec_t w;
#pragma unroll
      for (int j = 0; j < GR; j++) w[j] = xa[v * GR + j];
      *(vec_t *)&As[buf][am][ak * NA + v * GR] = w;
    }
#pragma unroll
    for (int i = 0; i < NB; i++) Bs[buf][bn + i * BSTEP][bk] = xb[i];
    if (kvalid < KB && bk >= kvalid) {  // last chunk of a segment whose K % KB != 0
#pragma unroll
      for (int i = 0; i < NB; i++) Bs[buf][bn + i * BSTEP][bk] = T(0);
    }
  };
  auto store = [&](int buf, int kvalid) { store_from(buf, kvalid, ra, rb); };
  auto advance = [&](int &sgi, int &kk0) {
    kk0 += KB;
    while (sgi < P.nseg && kk0 >= segs[P.seg0 + sgi].K) {
      sgi++;
      kk0 = 0;
    }
  };
  // One K chunk from LDS buffer buf.  The chunk's k indices are permuted
  // between the MFMA steps (A and B alike, so the sum is the same): at step
  // kq, lanes 16 k' .. 16 k' + 15 hold k = (KB / 4) k' + kq, so a lane's
  // operands for GR consecutive steps are GR consecutive values of its row
  // of A (column of B): one LDS read per row block per GR steps.
  constexpr int KQ = KB / 4;
  auto compute = [&](int buf) {
#pragma unroll
    for (int g0 = 0; g0 < KQ; g0 += GR) {
      vec_t bv[BNB], av[BMB];
#pragma unroll
      for (int b = 0; b < BNB; b++) bv[b] = *(const vec_t *)&Bs[buf][wn + b * 16 + l16][KQ * lk + g0];
#pragma unroll
      for (int a = 0; a < BMB; a++) av[a] = *(const vec_t *)&As[buf][wm + a * 16 + l16][KQ * lk + g0];
#pragma unroll
      for (int j = 0; j < GR; j++) {
        if constexpr (F64) {
#pragma unroll
          for (int a = 0; a < BMB; a++)
#pragma unroll
            for (int b = 0; b < BNB; b++)
              accd[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[b][j], av[a][j], accd[a][b], 0, 0, 0);
        } else {
#pragma unroll
          for (int a = 0; a < BMB; a++)
#pragma unroll
            for (int b = 0; b < BNB; b++)
              accf[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[b][j], av[a][j], accf[a][b], 0, 0, 0);
        }
      }
    }
  };
  // element (a, b, r) of this lane's accumulators: its place in C and value
  auto elem = [&](int a, int b, int r, int &m, int &n, T &v) __attribute__((always_inline)) {
    if constexpr (F64) {
      // D = B^T A^T: register r of lane l holds C[l % 16][4 r + l / 16]
      m = m0 + wm + a * 16 + l16;
      n = n0 + wn + b * 16 + 4 * r + lk;
      v = accd[a][b][r];
    } else {
      // D = B^T A^T: register r of lane l holds C[l % 16][4 (l / 16) + r]
      m = m0 + wm + a * 16 + l16;
      n = n0 + wn + b * 16 + 4 * lk + r;
      v = accf[a][b][r];
    }
  };
  // C -= AB: after the K loop the C values of a group of block rows are all
  // loaded before any is stored, so they are in flight together (one round
  // trip per group, not one per element; groups of 2 block rows keep the
  // registers of the large tile in bounds)
  constexpr int AG = BMB > 2 ? 2 : BMB;
  T cold[BMB][BNB][4];
  auto load_c = [&](int a0) __attribute__((always_inline)) {
    if constexpr (MODE == 0) {
#pragma unroll
      for (int a = a0; a < a0 + AG; a++)
#pragma unroll
        for (int b = 0; b < BNB; b++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            int m, n;
            T v;
            elem(a, b, r, m, n, v);
            cold[a][b][r] = (m < P.M && n < P.N) ? gptr(P.C)[m + (long long)n * P.ldc] : T(0);
          }
    }
  };
  int sg = 0, k0 = 0;
  if (P.nseg == 0) return;
  if constexpr (DEPTH == 1) {
    load(sg, k0);
    store(0, segs[P.seg0].K - k0);
    __syncthreads();
    int buf = 0;
    for (;;) {
      int nsg = sg, nk0 = k0;
      advance(nsg, nk0);
      const bool more = nsg < P.nseg;
      if (more) load(nsg, nk0);
      compute(buf);
      if (!more) break;
      store(buf ^ 1, segs[P.seg0 + nsg].K - nk0);
      __syncthreads();
      buf ^= 1;
      sg = nsg;
      k0 = nk0;
    }
  } else {
    // chunk c is multiplied from LDS while chunk c+1 sits in one register
    // stage (stored after the MFMA block) and chunk c+2 is in flight into
    // the other; the stages alternate, so no register copies wait on loads
    load(sg, k0);
    store(0, segs[P.seg0].K - k0);
    __syncthreads();
    int s1 = sg, k1 = k0;
    advance(s1, k1);
    int s2 = s1, k2 = k1;
    if (s1 < P.nseg) {
      load_into(s1, k1, ra, rb);
      advance(s2, k2);
      if (s2 < P.nseg) load_into(s2, k2, ra2, rb2);
    }
    int buf = 0;
    for (;;) {
      // stage a holds (s1, k1), stage b holds (s2, k2)
      compute(buf);
      if (s1 >= P.nseg) break;
      store_from(buf ^ 1, segs[P.seg0 + s1].K - k1, ra, rb);
      __syncthreads();
      buf ^= 1;
      int s3 = s2, k3 = k2;
      if (s3 < P.nseg) {
        advance(s3, k3);
        if (s3 < P.nseg) load_into(s3, k3, ra, rb);
      }
      compute(buf);
      if (s2 >= P.nseg) break;
      store_from(buf ^ 1, segs[P.seg0 + s2].K - k2, ra2, rb2);
      __syncthreads();
      buf ^= 1;
      s1 = s3;
      k1 = k3;
      s2 = s3;
      k2 = k3;
      if (s2 < P.nseg) {
        advance(s2, k2);
        if (s2 < P.nseg) load_into(s2, k2, ra2, rb2);
      }
    }
  }
#pragma unroll
  for (int a0 = 0; a0 < BMB; a0 += AG) {
    load_c(a0);
#pragma unroll
    for (int a = a0; a < a0 + AG; a++)
#pragma unroll
      for (int b = 0; b < BNB; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int m, n;
          T v;
          elem(a, b, r, m, n, v);
          if (m < P.M && n < P.N) {
            auto c = gptr(P.C) + m + (long long)n * P.ldc;
            if (MODE == 1) *c = v;
            else if (MODE == 2) *c = -v;
            else *c = cold[a][b][r] - v;
          }
        }
  }
}
// ---------------------------------------------------------------------------
// FP32 grouped GEMM on the bf16 MFMA (v_mfma_f32_16x16x32_bf16; the
// SNU_GEMM_SPLIT form of the FP32 factorization's GEMM launches)
// ---------------------------------------------------------------------------
// Every FP32 operand x is split into three bf16 pieces with x = h + m + l
// exactly: h = bf16(x), m = bf16(x - h), l = bf16(x - h - m), each rounded
// to nearest even.  x - h and x - h - m are exact in FP32, |m| <= 2^-8 |x|
// and |l| <= 2^-16 |x|, and x - h - m has at most 8 significant bits, so l
// is exact as well.  C accumulates in FP32 the six products of pieces of
// weight >= 2^-16, hh + hm + mh + hl + mm + lh (each bf16 product is exact
// in FP32); the three left out (ml, lm, ll) total at most about
// 2^-23 |a| |b|, two units of FP32's own rounding of one product.  Per
// 16 x 16 x 32 block that is six bf16 MFMAs (16 cycles each) in place of
// eight 16x16x4 f32 MFMAs (32 cycles each).
// Workgroup tile (32 BMB) x (32 BNB) (128 x 128 or 64 x 64), 4 waves in
// 2 x 2; K chunk 32 (one bf16 MFMA deep).  LDS holds one chunk, pieces per
// row of A and per column of B as [h 0..31 | m 0..31 | l 0..31 | pad 8] bf16
// (52 dwords: the 16-byte reads of 8 consecutive rows fall on disjoint
// banks); the next chunk is in flight in registers during the MFMA block,
// and several workgroups share a CU (2 of the large tile, 4 of the small),
// so one splits and stores while another multiplies.
// Tiles, problems, modes, aliasing and the XCD order are those of k_gemm
// with the same tile size (the same tile counts).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float fl2_t __attribute__((ext_vector_type(2)));
constexpr int kSplitRow = 104;  // bf16 per LDS row: 3 pieces x 32 + 8
// pieces of two FP32 values as packed bf16 pairs (element 0 in the low half)
__device__ __forceinline__ void split_pair(float x0, float x1, unsigned &h, unsigned &m, unsigned &l) {
  const fl2_t x = fl2_t{x0, x1};
  const bf16x2_t hb = __builtin_convertvector(x, bf16x2_t);
  const fl2_t r = x - __builtin_convertvector(hb, fl2_t);
  const bf16x2_t mb = __builtin_convertvector(r, bf16x2_t);
  const fl2_t r2 = r - __builtin_convertvector(mb, fl2_t);
  const bf16x2_t lb = __builtin_convertvector(r2, bf16x2_t);
  h = __builtin_bit_cast(unsigned, hb);
  m = __builtin_bit_cast(unsigned, mb);
  l = __builtin_bit_cast(unsigned, lb);
}
template <int MODE, int BMB, int BNB>
__global__ __launch_bounds__(256, BMB * BNB >= 16 ? 2 : 4) void k_gemm_split(
    const GemmProbT<float> *__restrict__ probs, int nprob, const GemmSegT<float> *__restrict__ segs) {
  constexpr int BM = 32 * BMB, BN = 32 * BNB, KB = 32;
  constexpr int NA = BM / 8, NBC = BN / 16;  // A: NA consecutive k per thread; B: NBC columns
  constexpr int RW = kSplitRow / 2;          // dwords per LDS row
  __shared__ __align__(16) unsigned As[BM * RW];
  __shared__ __align__(16) unsigned Bs[BN * RW];
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  int lo = 0, hi = nprob - 1;
  if (nprob == (int)gridDim.x) lo = hi = tile;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemmProbT<float> P = probs[lo];
  constexpr int GM = 8;
  const int lt = tile - P.tile0;
  const int tiles_m = (P.M + BM - 1) / BM;
  const int grp = lt / (GM * P.tiles_n), first = grp * GM;
  const int gsz = min(GM, tiles_m - first), li = lt - grp * GM * P.tiles_n;
  const int m0 = (first + li % gsz) * BM, n0 = (li / gsz) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 16 * BMB, wn = (wave >> 1) * 16 * BNB;
  const int l16 = lane & 15, lk = lane >> 4;
  f4 acc[BMB][BNB];
#pragma unroll
  for (int a = 0; a < BMB; a++)
#pragma unroll
    for (int b = 0; b < BNB; b++) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  // A chunk: row m0 + am, k = NA ak + 0..NA-1 (coalesced over rows); B
  // chunk: columns n0 + bc + 16 i (i < NBC), k = 2 bp, 2 bp + 1 (a lane
  // pair's two loads cover 8 bytes, 16 lanes one 128-byte column segment).
  // Addresses are clamped into the operands as in k_gemm; the K tail of B is
  // zeroed before the split.
  static_assert(NA % 8 == 0 && 256 % BM == 0, "thread-to-element map");
  const int am = tid % BM, ak = tid / BM, bp = tid & 15, bc = tid >> 4;
  float xa[NA], xb[2 * NBC];
  auto load = [&](int sgi, int kk0) {
    const GemmSegT<float> S = segs[P.seg0 + sgi];
    const auto pa = gptr(S.A) + min(m0 + am, P.M - 1);
    if (kk0 + KB <= S.K) {
      const auto p = pa + (long long)(kk0 + NA * ak) * S.lda;
#pragma unroll
      for (int i = 0; i < NA; i++) xa[i] = p[i * S.lda];
    } else {
#pragma unroll
      for (int i = 0; i < NA; i++) xa[i] = pa[(long long)min(kk0 + NA * ak + i, S.K - 1) * S.lda];
    }
    const auto pb = gptr(S.B);
    if (kk0 + KB <= S.K && n0 + BN <= P.N) {
      const auto p = pb + kk0 + 2 * bp + (long long)(n0 + bc) * S.ldb;
      const long long st = 16 * S.ldb;
#pragma unroll
      for (int i = 0; i < NBC; i++) {
        xb[2 * i] = p[i * st];
        xb[2 * i + 1] = p[i * st + 1];
      }
    } else {
      const int k0c = min(kk0 + 2 * bp, S.K - 1), k1c = min(kk0 + 2 * bp + 1, S.K - 1);
#pragma unroll
      for (int i = 0; i < NBC; i++) {
        const auto p = pb + (long long)min(n0 + bc + 16 * i, P.N - 1) * S.ldb;
        xb[2 * i] = p[k0c];
        xb[2 * i + 1] = p[k1c];
      }
    }
  };
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  auto store = [&](int kvalid) {
    unsigned h[NA / 2], m[NA / 2], l[NA / 2];
#pragma unroll
    for (int j = 0; j < NA / 2; j++) split_pair(xa[2 * j], xa[2 * j + 1], h[j], m[j], l[j]);
    unsigned *ra = As + am * RW + (NA / 2) * ak;
#pragma unroll
    for (int v = 0; v < NA / 8; v++) {
      *(u4 *)(ra + 4 * v) = u4{h[4 * v], h[4 * v + 1], h[4 * v + 2], h[4 * v + 3]};
      *(u4 *)(ra + 16 + 4 * v) = u4{m[4 * v], m[4 * v + 1], m[4 * v + 2], m[4 * v + 3]};
      *(u4 *)(ra + 32 + 4 * v) = u4{l[4 * v], l[4 * v + 1], l[4 * v + 2], l[4 * v + 3]};
    }
    const bool z0 = 2 * bp >= kvalid, z1 = 2 * bp + 1 >= kvalid;
#pragma unroll
    for (int i = 0; i < NBC; i++) {
      unsigned hb, mb, lb;
      split_pair(z0 ? 0.f : xb[2 * i], z1 ? 0.f : xb[2 * i + 1], hb, mb, lb);
      unsigned *rb = Bs + (bc + 16 * i) * RW + bp;
      rb[0] = hb;
      rb[16] = mb;
      rb[32] = lb;
    }
  };
  auto advance = [&](int &sgi, int &kk0) {
    kk0 += KB;
    while (sgi < P.nseg && kk0 >= segs[P.seg0 + sgi].K) {
      sgi++;
      kk0 = 0;
    }
  };
  // lane l16 + 16 lk supplies row (column) l16 of its block with k slots
  // 8 lk .. 8 lk + 7 of every piece, for both operands alike
  auto compute = [&]() {
    bf16x8_t av[BMB][3];
#pragma unroll
    for (int a = 0; a < BMB; a++)
#pragma unroll
      for (int p = 0; p < 3; p++)
        av[a][p] = __builtin_bit_cast(bf16x8_t, *(const u4 *)(As + (wm + a * 16 + l16) * RW + 16 * p + 4 * lk));
#pragma unroll
    for (int b = 0; b < BNB; b++) {
      bf16x8_t bv[3];
#pragma unroll
      for (int p = 0; p < 3; p++)
        bv[p] = __builtin_bit_cast(bf16x8_t, *(const u4 *)(Bs + (wn + b * 16 + l16) * RW + 16 * p + 4 * lk));
#pragma unroll
      for (int a = 0; a < BMB; a++) {
        // small terms first: l h, m m, h l, m h, h m, h h
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[2], av[a][0], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[1], av[a][1], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[0], av[a][2], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[1], av[a][0], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[0], av[a][1], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[0], av[a][0], acc[a][b], 0, 0, 0);
      }
    }
  };
  if (P.nseg == 0) return;
  int sg = 0, k0 = 0;
  load(sg, k0);
  store(segs[P.seg0].K - k0);
  __syncthreads();
  for (;;) {
    int nsg = sg, nk0 = k0;
    advance(nsg, nk0);
    const bool more = nsg < P.nseg;
    if (more) load(nsg, nk0);
    compute();
    if (!more) break;
    __syncthreads();
    store(segs[P.seg0 + nsg].K - nk0);
    __syncthreads();
    sg = nsg;
    k0 = nk0;
  }
  // D = B^T A^T: register r of lane l holds C[l % 16][4 (l / 16) + r]
  constexpr int AG = BMB > 2 ? 2 : BMB;
#pragma unroll
  for (int a0 = 0; a0 < BMB; a0 += AG) {
    float cold[AG][BNB][4];
    if constexpr (MODE == 0) {
#pragma unroll
      for (int a = 0; a < AG; a++)
#pragma unroll
        for (int b = 0; b < BNB; b++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int m = m0 + wm + (a0 + a) * 16 + l16, n = n0 + wn + b * 16 + 4 * lk + r;
            cold[a][b][r] = (m < P.M && n < P.N) ? gptr(P.C)[m + (long long)n * P.ldc] : 0.f;
          }
    }
#pragma unroll
    for (int a = 0; a < AG; a++)
#pragma unroll
      for (int b = 0; b < BNB; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int m = m0 + wm + (a0 + a) * 16 + l16, n = n0 + wn + b * 16 + 4 * lk + r;
          if (m < P.M && n < P.N) {
            auto c = gptr(P.C) + m + (long long)n * P.ldc;
            const float v = acc[a0 + a][b][r];
            if (MODE == 1) *c = v;
            else if (MODE == 2) *c = -v;
            else *c = cold[a][b][r] - v;
          }
        }
  }
}
void launch_gemm128_split(const GemmProbT<float> *probs, int nprob, const GemmSegT<float> *segs,
                          int total_tiles, int mode, hipStream_t st) {
  if (total_tiles <= 0 || nprob <= 0) return;
  if (mode == 1)
    hipLaunchKernelGGL((k_gemm_split<1, 4, 4>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, segs);
  else if (mode == 2)
    hipLaunchKernelGGL((k_gemm_split<2, 4, 4>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, segs);
  else
    hipLaunchKernelGGL((k_gemm_split<0, 4, 4>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, segs);
}
void launch_gemm_split(const GemmProbT<float> *probs, int nprob, const GemmSegT<float> *segs,
                       int total_tiles, int mode, hipStream_t st) {
  if (total_tiles <= 0 || nprob <= 0) return;
  if (mode == 1)
    hipLaunchKernelGGL((k_gemm_split<1, 2, 2>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, segs);
  else if (mode == 2)
    hipLaunchKernelGGL((k_gemm_split<2, 2, 2>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, segs);
  else
    hipLaunchKernelGGL((k_gemm_split<0, 2, 2>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, segs);
}

// Grouped 2-D copy over GEMM descriptors: C (ldc) = A of the problem's first
// segment (lda), one workgroup per 64 x 64 tile, rows contiguous per wave.
template <typename T>
__global__ __launch_bounds__(256) void k_copy2d(const GemmProbT<T> *__restrict__ probs, int nprob,
                                                const GemmSegT<T> *__restrict__ segs) {
  const int tile = blockIdx.x;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemmProbT<T> P = probs[lo];
  const GemmSegT<T> S = segs[P.seg0];
  const int lt = tile - P.tile0;
  const int m0 = (lt / P.tiles_n) * 64, n0 = (lt % P.tiles_n) * 64;
  const int r = threadIdx.x & 63;
  if (m0 + r >= P.M) return;
  const T *src = S.A + (m0 + r) + (long long)n0 * S.lda;
  T *dst = P.C + (m0 + r) + (long long)n0 * P.ldc;
  const int nc = min(64, P.N - n0);
  for (int c = threadIdx.x >> 6; c < nc; c += 4) dst[(long long)c * P.ldc] = src[(long long)c * S.lda];
}
template <typename T>
void launch_copy2d(const GemmProbT<T> *probs, int nprob, const GemmSegT<T> *segs, int total_tiles,
                   hipStream_t st) {
  if (total_tiles <= 0 || nprob <= 0) return;
  hipLaunchKernelGGL(k_copy2d<T>, dim3(total_tiles), dim3(256), 0, st, probs, nprob, segs);
}
// K chunk of the production launches: 32 in FP32 (the 128 x 128 tile
// 110 -> 122 TF/s at 4096^3, the 64 x 64 tile even with 16), 16 in FP64
// (32 costs the FP64 tiles their second workgroup per CU)
template <typename T>
constexpr int gemm_kb() {
  return sizeof(T) == 4 ? 32 : 16;
}
template <typename T>
void launch_gemm(const GemmProbT<T> *probs, int nprob, const GemmSegT<T> *segs,
                 int total_tiles, int mode, hipStream_t st) {
  if (total_tiles <= 0 || nprob <= 0) return;
  // loads one chunk ahead (two ahead measured slower in both precisions
  // once the accumulators stay in VGPRs)
  constexpr int KB = gemm_kb<T>();
  if (mode == 1)
    hipLaunchKernelGGL((k_gemm<T, 1, 2, 2, KB, 1>), dim3(total_tiles), dim3(256), 0, st, probs, nprob,
                       segs);
  else if (mode == 2)
    hipLaunchKernelGGL((k_gemm<T, 2, 2, 2, KB, 1>), dim3(total_tiles), dim3(256), 0, st, probs, nprob,
                       segs);
  else
    hipLaunchKernelGGL((k_gemm<T, 0, 2, 2, KB, 1>), dim3(total_tiles), dim3(256), 0, st, probs, nprob,
                       segs);
}
// 128 x 128 tiles, mode 0 (C -= AB), 1 (C = AB) or 2 (C = -AB); loads one
// chunk ahead (with accumulators in VGPRs the large tile is fastest that way
// in both precisions)
template <typename T>
void launch_gemm128(const GemmProbT<T> *probs, int nprob, const GemmSegT<T> *segs,
                    int total_tiles, int mode, hipStream_t st) {
  if (total_tiles <= 0 || nprob <= 0) return;
  constexpr int KB = gemm_kb<T>();
  if (mode == 1)
    hipLaunchKernelGGL((k_gemm<T, 1, 4, 4, KB, 1>), dim3(total_tiles), dim3(256), 0, st, probs, nprob,
                       segs);
  else if (mode == 2)
    hipLaunchKernelGGL((k_gemm<T, 2, 4, 4, KB, 1>), dim3(total_tiles), dim3(256), 0, st, probs, nprob,
                       segs);
  else
    hipLaunchKernelGGL((k_gemm<T, 0, 4, 4, KB, 1>), dim3(total_tiles), dim3(256), 0, st, probs, nprob,
                       segs);
}
// C -= AB variants for tuning: tile 64 or 128, K chunk 16 or 32, register
// prefetch depth 1 or 2 (the tile count is that of the tile size)
template <typename T>
void launch_gemm_variant(const GemmProbT<T> *probs, int nprob, const GemmSegT<T> *segs,
                         int total_tiles, int tile, int kb, int depth, hipStream_t st) {
  if (total_tiles <= 0 || nprob <= 0) return;
  const dim3 g(total_tiles), b(256);
  if (tile == 128) {
    if (depth == 2) hipLaunchKernelGGL((k_gemm<T, 0, 4, 4, 16, 2>), g, b, 0, st, probs, nprob, segs);
    else if (kb == 32) hipLaunchKernelGGL((k_gemm<T, 0, 4, 4, 32, 1>), g, b, 0, st, probs, nprob, segs);
    else hipLaunchKernelGGL((k_gemm<T, 0, 4, 4, 16, 1>), g, b, 0, st, probs, nprob, segs);
  } else {
    if (depth == 2) hipLaunchKernelGGL((k_gemm<T, 0, 2, 2, 16, 2>), g, b, 0, st, probs, nprob, segs);
    else if (kb == 32) hipLaunchKernelGGL((k_gemm<T, 0, 2, 2, 32, 1>), g, b, 0, st, probs, nprob, segs);
    else hipLaunchKernelGGL((k_gemm<T, 0, 2, 2, 16, 1>), g, b, 0, st, probs, nprob, segs);
  }
}

// ---------------------------------------------------------------------------
// dense: grouped GEMV (the dense solves)
// ---------------------------------------------------------------------------
// One workgroup per 64 rows of a problem; lane = row, the 4 waves take
// interleaved columns and are combined in LDS (fixed order: deterministic).
// A is read once, coalesced down its columns; x through scalar loads.
// nrhs columns are processed in groups of 16 (A is re-read per group).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_gemv(const GemvProbT<T> *__restrict__ probs,
                                              int nprob,
                                              const GemvSegT<T> *__restrict__ segs) {
  constexpr int CG = 16;  // right-hand-side columns per pass over A
  __shared__ T red[4][CG][64];
  const int tile = blockIdx.x;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemvProbT<T> P = probs[lo];
  const int nrhs = P.nrhs;
  const int m0 = (tile - P.tile0) * 64;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int m = m0 + lane, mc = min(m, P.M - 1);
  for (int c0 = 0; c0 < nrhs; c0 += CG) {
    const int nc = min(CG, nrhs - c0);
    T acc[CG];
#pragma unroll
    for (int c = 0; c < CG; c++) acc[c] = T(0);
    for (int sg = 0; sg < P.nseg; sg++) {
      const GemvSegT<T> S = segs[P.seg0 + sg];
      int kb = 0, ke = S.K;
      if (P.tri == 1) ke = min(S.K, m0 + 64);
      if (P.tri == 2) kb = m0;
      auto A = gptr(S.A) + mc;
      auto X = gptr(S.x) + (long long)c0 * S.ldx;
      int k = kb + w;
      // NF columns of A in flight per wave (the loads are independent of the
      // accumulation), added in groups of four: 8 in FP64, 16 in FP32 (the
      // same bytes in flight)
      constexpr int NF = sizeof(T) == 4 ? 16 : 8;
      for (; k + 4 * (NF - 1) < ke; k += 4 * NF) {
        T a[NF];
#pragma unroll
        for (int j = 0; j < NF; j++) a[j] = A[(long long)(k + 4 * j) * S.lda];
#pragma unroll
        for (int c = 0; c < CG; c++)
          if (c < nc) {
            auto xc = X + (long long)c * S.ldx + k;
#pragma unroll
            for (int j = 0; j < NF; j += 4)
              acc[c] += a[j] * xc[4 * j] + a[j + 1] * xc[4 * j + 4] + a[j + 2] * xc[4 * j + 8] +
                        a[j + 3] * xc[4 * j + 12];
          }
      }
      for (; k + 12 < ke; k += 16) {
        T a0 = A[(long long)k * S.lda], a1 = A[(long long)(k + 4) * S.lda];
        T a2 = A[(long long)(k + 8) * S.lda], a3 = A[(long long)(k + 12) * S.lda];
#pragma unroll
        for (int c = 0; c < CG; c++)
          if (c < nc) {
            auto xc = X + (long long)c * S.ldx;
            acc[c] += a0 * xc[k] + a1 * xc[k + 4] + a2 * xc[k + 8] + a3 * xc[k + 12];
          }
      }
      for (; k < ke; k += 4) {
        T a0 = A[(long long)k * S.lda];
#pragma unroll
        for (int c = 0; c < CG; c++)
          if (c < nc) acc[c] += a0 * X[(long long)c * S.ldx + k];
      }
    }
#pragma unroll
    for (int c = 0; c < CG; c++) red[w][c][lane] = acc[c];
    __syncthreads();
    if (w == 0 && m < P.M) {
      for (int c = 0; c < nc; c++) {
        T v = red[0][c][lane] + red[1][c][lane] + red[2][c][lane] + red[3][c][lane];
        auto y = gptr(P.y) + m + (long long)(c0 + c) * P.ldy;
        if (MODE == 1) *y = v;
        else *y = *y - v;
      }
    }
    __syncthreads();
  }
}
// Split-K GEMV, pass 1: workgroup (problem, tile lt, piece) sums
// A[rows of lt][k] x[k] over k in [piece * kGemvPiece, (piece + 1) *
// kGemvPiece) of every segment (within the triangle for tri problems), in
// the order of the unsplit kernel, and stores the 64 x nrhs partial block.
// Pass 2 adds the pieces in ascending order.  The decomposition depends on
// the problem alone, so a problem gives the same bits in any launch.
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_part(const GemvProbT<T> *__restrict__ probs,
                                                   int nprob,
                                                   const GemvSegT<T> *__restrict__ segs,
                                                   T *__restrict__ part) {
  constexpr int CG = 16;
  __shared__ T red[4][CG][64];
  const int b = blockIdx.x;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].wg0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const GemvProbT<T> P = probs[lo];
  const int local = b - P.wg0, lt = local / P.parts, pc = local % P.parts;
  const int nrhs = P.nrhs;
  const int m0 = lt * 64;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int m = m0 + lane, mc = min(m, P.M - 1);
  T *out = part + P.poff + (long long)(lt * P.parts + pc) * nrhs * 64;
  for (int c0 = 0; c0 < nrhs; c0 += CG) {
    const int nc = min(CG, nrhs - c0);
    T acc[CG];
#pragma unroll
    for (int c = 0; c < CG; c++) acc[c] = T(0);
    for (int sg = 0; sg < P.nseg; sg++) {
      const GemvSegT<T> S = segs[P.seg0 + sg];
      int kb = 0, ke = S.K;
      if (P.tri == 1) ke = min(S.K, m0 + 64);
      if (P.tri == 2) kb = m0;
      kb = max(kb, pc * kGemvPiece);
      ke = min(ke, (pc + 1) * kGemvPiece);
      auto A = gptr(S.A) + mc;
      auto X = gptr(S.x) + (long long)c0 * S.ldx;
      int k = kb + w;
      // NF columns of A in flight per wave (the loads are independent of the
      // accumulation), added in groups of four: 8 in FP64, 16 in FP32 (the
      // same bytes in flight)
      constexpr int NF = sizeof(T) == 4 ? 16 : 8;
      for (; k + 4 * (NF - 1) < ke; k += 4 * NF) {
        T a[NF];
#pragma unroll
        for (int j = 0; j < NF; j++) a[j] = A[(long long)(k + 4 * j) * S.lda];
#pragma unroll
        for (int c = 0; c < CG; c++)
          if (c < nc) {
            auto xc = X + (long long)c * S.ldx + k;
#pragma unroll
            for (int j = 0; j < NF; j += 4)
              acc[c] += a[j] * xc[4 * j] + a[j + 1] * xc[4 * j + 4] + a[j + 2] * xc[4 * j + 8] +
                        a[j + 3] * xc[4 * j + 12];
          }
      }
      for (; k + 12 < ke; k += 16) {
        T a0 = A[(long long)k * S.lda], a1 = A[(long long)(k + 4) * S.lda];
        T a2 = A[(long long)(k + 8) * S.lda], a3 = A[(long long)(k + 12) * S.lda];
#pragma unroll
        for (int c = 0; c < CG; c++)
          if (c < nc) {
            auto xc = X + (long long)c * S.ldx;
            acc[c] += a0 * xc[k] + a1 * xc[k + 4] + a2 * xc[k + 8] + a3 * xc[k + 12];
          }
      }
      for (; k < ke; k += 4) {
        T a0 = A[(long long)k * S.lda];
#pragma unroll
        for (int c = 0; c < CG; c++)
          if (c < nc) acc[c] += a0 * X[(long long)c * S.ldx + k];
      }
    }
#pragma unroll
    for (int c = 0; c < CG; c++) red[w][c][lane] = acc[c];
    __syncthreads();
    if (w == 0)
      for (int c = 0; c < nc; c++)
        out[(c0 + c) * 64 + lane] = red[0][c][lane] + red[1][c][lane] + red[2][c][lane] + red[3][c][lane];
    __syncthreads();
  }
}
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_gemv_sum(const GemvProbT<T> *__restrict__ probs, int nprob,
                                                  const T *__restrict__ part) {
  const int tile = blockIdx.x;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemvProbT<T> P = probs[lo];
  const int lt = tile - P.tile0, m = lt * 64 + (threadIdx.x & 63);
  if (m >= P.M) return;
  const T *in = part + P.poff + (long long)lt * P.parts * P.nrhs * 64 + (threadIdx.x & 63);
  for (int c = threadIdx.x >> 6; c < P.nrhs; c += 4) {
    T v = in[c * 64];
    for (int pc = 1; pc < P.parts; pc++) v += in[((long long)pc * P.nrhs + c) * 64];
    auto y = gptr(P.y) + m + (long long)c * P.ldy;
    if (MODE == 1) *y = v;
    else *y = *y - v;
  }
}
template <typename T>
void launch_gemv_split(const GemvProbT<T> *probs, int nprob, const GemvSegT<T> *segs,
                       int total_tiles, int total_wgs, int mode, T *part, hipStream_t st) {
  if (total_tiles <= 0 || nprob <= 0) return;
  hipLaunchKernelGGL(k_gemv_part<T>, dim3(total_wgs), dim3(256), 0, st, probs, nprob, segs, part);
  if (mode == 1)
    hipLaunchKernelGGL((k_gemv_sum<T, 1>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, part);
  else
    hipLaunchKernelGGL((k_gemv_sum<T, 0>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, part);
}
template <typename T>
void launch_gemv(const GemvProbT<T> *probs, int nprob, const GemvSegT<T> *segs,
                 int total_tiles, int mode, hipStream_t st) {
  if (total_tiles <= 0 || nprob <= 0) return;
  if (mode == 1)
    hipLaunchKernelGGL((k_gemv<T, 1>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, segs);
  else
    hipLaunchKernelGGL((k_gemv<T, 0>), dim3(total_tiles), dim3(256), 0, st, probs, nprob, segs);
}

// ---------------------------------------------------------------------------
// dense: diagonal block LU (no pivoting, as cusolverDnDgetrf with a null
// ipiv / LAPACKE_mkl_dgetrfnp) and the inverses of its triangular factors.
// ---------------------------------------------------------------------------
// One workgroup of 8 waves per block, thread (r, g) holding row r, columns
// g, g + 8, ..., g + 56 of three 64 x 64 matrices in registers (the block is
// padded with a unit diagonal): the factors being formed, X = L^-1 and
// Y = U^-1.
// Pivot step j (one barrier): column j of the factors and of Y are published
// in LDS (by the wave holding them); row j of the factors (final: U row j)
// and of X is needed by each wave only on its own columns, which lane j of
// the same wave holds, so it is read from that lane's registers (no LDS
// round trip: the single-lane LDS stores of the row cost more than all the
// arithmetic of a step); then with one division m = (r > j ? A_rj : Y_rj) /
// U_jj per thread,
//   rows r > j:  A_rc -= m U_jc (c > j), L_rj = m,  X_rc -= m X_jc (c <= j)
//   rows r <= j: Y_rj = m,  Y_rc -= m U_jc (c > j)
// which are right-looking LU, forward elimination L X = I (X_j is final at
// step j) and Y U = I column by column (Y_rj final once divided by U_jj).
// Pivot step J (a template parameter, so that whether column g + 8q is left
// of, at or right of J is a compile-time fact for most (J, q) and a uniform
// select for the rest; runtime steps make the compiler shuffle the register
// arrays between branches).
template <typename T>
__device__ __forceinline__ T lane_value(T v, int lane) {
  if constexpr (std::is_same<T, double>::value) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
  } else {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
  }
}
template <typename T, int J>
__device__ __forceinline__ void diag_step(T (&x)[8], T (&xl)[8], T (&xu)[8], int r, int g,
                                          T (*lb)[64], T (*yb)[64], T eps) {
  constexpr int b = J & 1;
  if (g == (J & 7)) {
    lb[b][r] = x[J >> 3];
    yb[b][r] = xu[J >> 3];
  }
  // row J on this wave's columns, from lane J
  T uq[8], xq[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    if (8 * q + 7 > J) uq[q] = lane_value(x[q], J);
    if (8 * q <= J) xq[q] = lane_value(xl[q], J);
  }
  __syncthreads();
  // keep the scheduler from interleaving steps (register pressure)
  __builtin_amdgcn_sched_barrier(0);
  const bool below = r > J;
  // lb[b][J]: column J at row J, the pivot U_JJ (clamped to +-eps when
  // smaller in magnitude: the leaf pivots' rule; eps = 0 leaves it as is)
  T piv = lb[b][J];
  if (piv < eps && piv > -eps) piv = piv < T(0) ? -eps : eps;
  const T m = (below ? lb[b][r] : yb[b][r]) / piv;
  // multipliers zeroed on the lanes an update does not apply to (x - 0 u is
  // x exactly)
  const T mb = below ? m : T(0), ma = below ? T(0) : m;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    if (8 * q + 7 < J) {  // left of J for every wave
      xl[q] = xl[q] - mb * xq[q];
    } else if (8 * q > J) {  // right of J for every wave
      x[q] = x[q] - mb * uq[q];
      xu[q] = xu[q] - ma * uq[q];
    } else {  // column g + 8q is left of, at or right of J depending on g
      const int c = g + 8 * q;
      const bool right = c > J, at = c == J;
      const T nx = x[q] - mb * uq[q];
      const T ny = xu[q] - ma * uq[q];
      const T nl = xl[q] - mb * xq[q];
      x[q] = right ? nx : (at ? (below ? m : (r == J ? piv : x[q])) : x[q]);
      xu[q] = right ? ny : (at && !below ? m : xu[q]);
      xl[q] = right ? xl[q] : nl;
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  // opaque to the optimizer: without this it carries facts about earlier
  // steps' values forward and the register pressure grows with every step
#pragma unroll
  for (int q = 0; q < 8; q++) asm volatile("" : "+v"(x[q]), "+v"(xl[q]), "+v"(xu[q]));
}
template <typename T, int... J>
__device__ __forceinline__ void diag_steps(std::integer_sequence<int, J...>, T (&x)[8],
                                           T (&xl)[8], T (&xu)[8], int r, int g, T (*lb)[64],
                                           T (*yb)[64], T eps) {
  (diag_step<T, J>(x, xl, xu, r, g, lb, yb, eps), ...);
}
template <typename T>
__global__ __launch_bounds__(512) void k_diag_lu(const DiagProbT<T> *__restrict__ probs) {
  __shared__ T lb[2][64], yb[2][64];
  const DiagProbT<T> P = probs[blockIdx.x];
  const int nb = P.nb, tid = threadIdx.x;
  const int r = tid & 63, g = __builtin_amdgcn_readfirstlane(tid >> 6);
  T x[8], xl[8], xu[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int c = g + 8 * q;
    x[q] = (r < nb && c < nb) ? P.D[r + (long long)c * P.ld] : (r == c ? T(1) : T(0));
    xl[q] = xu[q] = r == c ? T(1) : T(0);
  }
  // all 64 steps run: past nb the padding is an identity block, whose steps
  // change nothing (zero multipliers)
  diag_steps<T>(std::make_integer_sequence<int, 64>{}, x, xl, xu, r, g, lb, yb, P.eps);
  // the descriptor is read again (through a pointer the compiler cannot
  // equate with the first), so no store address stays live across the steps
  const DiagProbT<T> *pq = probs;
  asm volatile("" : "+s"(pq));
  const DiagProbT<T> Q = pq[blockIdx.x];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int c = g + 8 * q;
    if (r < Q.nb && c < Q.nb) {
      Q.D[r + (long long)c * Q.ld] = x[q];
      Q.Linv[r + (long long)c * Q.ldi] = xl[q];
      Q.Uinv[r + (long long)c * Q.ldi] = xu[q];
    }
  }
}
template <typename T>
void launch_diag_lu(const DiagProbT<T> *probs, int nprob, hipStream_t st) {
  if (nprob <= 0) return;
  hipLaunchKernelGGL(k_diag_lu<T>, dim3(nprob), dim3(512), 0, st, probs);
}

// ---------------------------------------------------------------------------
// solve: permutation, leaf triangular solves, boundary rhs assembly
// ---------------------------------------------------------------------------
// One matrix per blockIdx.y: bp[vbase + order[r] + col ld] = b[r + col ldb].
template <typename T>
__global__ void k_permute_in(const PermDev *__restrict__ pd, const VecDev *__restrict__ vd,
                             T *bp) {
  const PermDev P = pd[blockIdx.y];
  const VecDev V = vd[P.vec];
  const long long tot = (long long)P.n * V.nrhs;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const int col = (int)(i / P.n), r = (int)(i % P.n);
    bp[V.base + P.order[r] + (long long)col * V.ld] = (T)P.b[r + col * P.ldb];
  }
}
template <typename T>
__global__ void k_permute_out(const PermDev *__restrict__ pd, const VecDev *__restrict__ vd,
                              const T *bp) {
  const PermDev P = pd[blockIdx.y];
  const VecDev V = vd[P.vec];
  const long long tot = (long long)P.n * V.nrhs;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const int col = (int)(i / P.n), r = (int)(i % P.n);
    P.x[r + col * P.ldx] = (double)bp[V.base + P.order[r] + (long long)col * V.ld];
  }
}
static unsigned perm_blocks(long long max_elems) {
  long long b = (max_elems + 255) / 256;
  return (unsigned)std::max(1LL, std::min(b, 2048LL));
}
template <typename T>
void launch_permute_in(const PermDev *pd, int nmat, const VecDev *vd, T *bp,
                       long long max_elems, hipStream_t st) {
  if (nmat <= 0 || max_elems <= 0) return;
  hipLaunchKernelGGL(k_permute_in<T>, dim3(perm_blocks(max_elems), nmat), dim3(256), 0, st, pd,
                     vd, bp);
}
template <typename T>
void launch_permute_out(const PermDev *pd, int nmat, const VecDev *vd, const T *bp,
                        long long max_elems, hipStream_t st) {
  if (nmat <= 0 || max_elems <= 0) return;
  hipLaunchKernelGGL(k_permute_out<T>, dim3(perm_blocks(max_elems), nmat), dim3(256), 0, st,
                     pd, vd, bp);
}

// Sparse dot products of the leaf solves with several right-hand sides: a
// wave serves CW = the largest power of two <= min(nrhs, 16) columns at once,
// KL = 64 / CW lanes per column striding over the row's entries, partial
// sums combined within the KL lanes.  CW = 1 (one rhs) is the plain wave
// reduction.  The rounding of a column's sum depends on CW, i.e. on nrhs.
__device__ __forceinline__ int col_width(int nrhs) {
  const int m = min(nrhs, 16);
  return 1 << (31 - __clz(max(m, 1)));
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_v(T v) {
  if constexpr (sizeof(T) == 8) return dpp_d<CTRL>(v);
  else return dpp_f<CTRL>(v);
}
// Sum over each group of kl consecutive lanes (kl a power of two <= 64),
// valid in the group's first lane: DPP within rows of 16 lanes (quad
// butterflies, then rotations that bring lanes +4 and +8 of a row to its
// front), read-lane sums across rows.
template <typename T>
__device__ __forceinline__ T group_sum(T v, int kl) {
  if (kl >= 2) v += dpp_v<0xB1>(v);   // quad_perm 1,0,3,2
  if (kl >= 4) v += dpp_v<0x4E>(v);   // quad_perm 2,3,0,1
  if (kl >= 8) v += dpp_v<0x12C>(v);  // row_ror:12: lane i <- lane i + 4
  if (kl >= 16) v += dpp_v<0x128>(v); // row_ror:8: lane i <- lane i + 8
  if (kl == 32) {
    const T lo = rld(v, 0) + rld(v, 16), hi = rld(v, 32) + rld(v, 48);
    v = (threadIdx.x & 63) < 32 ? lo : hi;
  } else if (kl == 64) {
    v = (rld(v, 0) + rld(v, 16)) + (rld(v, 32) + rld(v, 48));
  }
  return v;
}

// y_L = L_LL^-1 b_L (unit lower), level by level; the (row, column group)
// pairs of a level are spread over the waves.
template <typename T>
__global__ __launch_bounds__(256) void k_leaf_fwd(const LeafDev *__restrict__ leaves,
                                                  const VecDev *__restrict__ vd, T *bp) {
  const LeafDev L = leaves[blockIdx.x];
  const VecDev V = vd[L.vec];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int W = blockDim.x >> 6;
  const int *rp = L.rowptr, *dg = L.diag, *wc = L.wcol;
  const T *v = (const T *)L.vals;
  const int *lp = L.levptr, *lr = L.levrows;
  const int cw = col_width(V.nrhs), kl = 64 / cw, ng = (V.nrhs + cw - 1) / cw;
  const int sub = lane % kl, cl = lane / kl;
  for (int lev = 0; lev < L.nlev; lev++) {
    const int r0 = lp[lev], nr = lp[lev + 1] - r0;
    for (int t = wave; t < nr * ng; t += W) {
      const int i = lr[r0 + t % nr], col = (t / nr) * cw + cl;
      const bool on = col < V.nrhs;
      T *b = bp + V.base + (long long)min(col, V.nrhs - 1) * V.ld + L.gstart;
      T s = T(0);
      for (int p = rp[i] + sub; p < dg[i]; p += kl) s += v[p] * b[wc[p]];
      s = group_sum(s, kl);
      if (on && sub == 0) b[i] -= s;
    }
    __syncthreads();
  }
}
template <typename T>
void launch_leaf_fwd(const LeafDev *leaves, int nleaves, const VecDev *vd, T *bp, hipStream_t st) {
  if (nleaves <= 0) return;
  hipLaunchKernelGGL(k_leaf_fwd<T>, dim3(nleaves), dim3(256), 0, st, leaves, vd, bp);
}

// Boundary rows: cbuf[ri] = sum_k L_SL(r, k) y_k (kernel_sparse.cu:935-944).
template <typename T>
__global__ __launch_bounds__(256) void k_leaf_fwd_sep(const LeafDev *__restrict__ leaves,
                                                      const int *__restrict__ wg_leaf,
                                                      const int *__restrict__ wg_first,
                                                      const VecDev *__restrict__ vd,
                                                      const T *__restrict__ bp,
                                                      T *__restrict__ cbuf, int cb_ld) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const LeafDev L = leaves[wg_leaf[blockIdx.x]];
  const VecDev V = vd[L.vec];
  const int ri = wg_first[blockIdx.x] + wave;
  if (ri >= L.nbr) return;
  const int r = L.brow[ri];
  const int *rp = L.rowptr, *wc = L.wcol;
  const T *v = (const T *)L.vals;
  const int cw = col_width(V.nrhs), kl = 64 / cw;
  const int sub = lane % kl, cl = lane / kl;
  for (int c0 = 0; c0 < V.nrhs; c0 += cw) {
    const int col = c0 + cl;
    const bool on = col < V.nrhs;
    const T *b = bp + V.base + (long long)min(col, V.nrhs - 1) * V.ld + L.gstart;
    T s = T(0);
    for (int p = rp[r] + sub; p < rp[r + 1]; p += kl) s += v[p] * b[wc[p]];
    s = group_sum(s, kl);
    if (on && sub == 0) cbuf[L.cb_off + ri + (long long)col * cb_ld] = s;
  }
}
template <typename T>
void launch_leaf_fwd_sep(const LeafDev *leaves, const int *wg_leaf, const int *wg_first, int nwg,
                         int waves, const VecDev *vd, const T *bp, T *cbuf, int cb_ld,
                         hipStream_t st) {
  if (nwg <= 0) return;
  hipLaunchKernelGGL(k_leaf_fwd_sep<T>, dim3(nwg), dim3(64 * waves), 0, st, leaves, wg_leaf,
                     wg_first, vd, bp, cbuf, cb_ld);
}

// b_R -= contributions of the leaves touching row R, in ascending rank
// (one thread per target row and rhs column; the target's matrix is that of
// its sources).
template <typename T>
__global__ void k_assemble_rhs(int ntarget, const int *__restrict__ trow_glob,
                               const int *__restrict__ tptr,
                               const int *__restrict__ tsrc_leaf,
                               const int *__restrict__ tsrc_ri,
                               const LeafDev *__restrict__ leaves,
                               const VecDev *__restrict__ vd,
                               const T *__restrict__ cbuf, int cb_ld, T *bp, int max_rhs) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)ntarget * max_rhs) return;
  const int t = (int)(i % ntarget), col = (int)(i / ntarget);
  const VecDev V = vd[leaves[tsrc_leaf[tptr[t]]].vec];
  if (col >= V.nrhs) return;
  T *b = bp + V.base + (long long)col * V.ld + trow_glob[t];
  T acc = *b;
  for (int s = tptr[t]; s < tptr[t + 1]; s++)
    acc -= cbuf[leaves[tsrc_leaf[s]].cb_off + tsrc_ri[s] + (long long)col * cb_ld];
  *b = acc;
}
template <typename T>
void launch_assemble_rhs(int ntarget, const int *trow_glob, const int *tptr,
                         const int *tsrc_leaf, const int *tsrc_ri, const LeafDev *leaves,
                         const VecDev *vd, const T *cbuf, int cb_ld, T *bp, int max_rhs,
                         hipStream_t st) {
  long long tot = (long long)ntarget * max_rhs;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_assemble_rhs<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     ntarget, trow_glob, tptr, tsrc_leaf, tsrc_ri, leaves, vd, cbuf, cb_ld, bp,
                     max_rhs);
}

// x_L = U_LL^-1 (y_L - U_LS x_S), level by level (SnuMat_factorize.cu:331-344);
// the (row, rhs column) pairs of a level are spread over the waves.
template <typename T>
__global__ __launch_bounds__(256) void k_leaf_bwd(const LeafDev *__restrict__ leaves,
                                                  const VecDev *__restrict__ vd, T *bp) {
  const LeafDev L = leaves[blockIdx.x];
  const VecDev V = vd[L.vec];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int W = blockDim.x >> 6;
  const int ls = L.leaf_size;
  const int *rp = L.rowptr, *dg = L.diag, *wc = L.wcol, *bg = L.bcol_glob;
  const T *v = (const T *)L.vals;
  const int *lp = L.ulevptr, *lr = L.ulevrows;
  const int cw = col_width(V.nrhs), kl = 64 / cw, ng = (V.nrhs + cw - 1) / cw;
  const int sub = lane % kl, cl = lane / kl;
  for (int lev = 0; lev < L.nulev; lev++) {
    const int r0 = lp[lev], nr = lp[lev + 1] - r0;
    for (int t = wave; t < nr * ng; t += W) {
      const int i = lr[r0 + t % nr], col = (t / nr) * cw + cl;
      const bool on = col < V.nrhs;
      T *b = bp + V.base + (long long)min(col, V.nrhs - 1) * V.ld;
      T s = T(0);
      for (int p = dg[i] + 1 + sub; p < rp[i + 1]; p += kl) {
        int c = wc[p];
        long long g = c < ls ? (long long)L.gstart + c : (long long)bg[c - ls];
        s += v[p] * b[g];
      }
      s = group_sum(s, kl);
      if (on && sub == 0) b[L.gstart + i] = (b[L.gstart + i] - s) / v[dg[i]];
    }
    __syncthreads();
  }
}
template <typename T>
void launch_leaf_bwd(const LeafDev *leaves, int nleaves, const VecDev *vd, T *bp, hipStream_t st) {
  if (nleaves <= 0) return;
  hipLaunchKernelGGL(k_leaf_bwd<T>, dim3(nleaves), dim3(256), 0, st, leaves, vd, bp);
}

// Leaf triangular solves with the leaf's solution segment in LDS (used when
// every leaf of the launch is resident at once): the
// level offsets and the level-ordered row tasks {row, first entry, end,
// diagonal} are staged in LDS once, then per level every (row, column) task
// gathers from LDS, so a level costs one round of entry loads instead of a
// dependent chain of index, entry and vector loads.  Backward: the boundary
// part of every U row (entries on separator columns, whose x is final) is
// subtracted for all rows at once before the sweep, which then reads only
// leaf columns.  Columns go in batches of cmax.
// LDS: [level offsets | tasks at task_off | cmax x ls values at xs_off].
// Threads per leaf: the waves take a level's rows in turn.
extern __shared__ __align__(16) char lds_raw[];
constexpr int kLeafSolveThreads = 1024;
template <typename T, bool BWD>
__global__ __launch_bounds__(kLeafSolveThreads) void k_leaf_solve(const LeafDev *__restrict__ leaves,
                                                    const VecDev *__restrict__ vd, T *bp,
                                                    int cmax, int task_off, int xs_off) {
  const LeafDev L = leaves[blockIdx.x];
  const VecDev V = vd[L.vec];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, W = blockDim.x >> 6;
  const int ls = L.leaf_size, nlev = BWD ? L.nulev : L.nlev;
  int *lps = (int *)lds_raw;
  int4 *tq = (int4 *)(lds_raw + task_off);
  T *xs = (T *)(lds_raw + xs_off);
  const int *lp = BWD ? L.ulevptr : L.levptr;
  const int4 *tk = (const int4 *)(BWD ? L.btask : L.ftask);
  for (int e = tid; e <= nlev; e += blockDim.x) lps[e] = lp[e];
  for (int e = tid; e < ls; e += blockDim.x) tq[e] = tk[e];
  const T *v = (const T *)L.vals;
  const int *wc = L.wcol;
  for (int c0 = 0; c0 < V.nrhs; c0 += cmax) {
    const int nc = min(cmax, V.nrhs - c0);
    __syncthreads();
    for (int e = tid; e < nc * ls; e += blockDim.x) {
      const int c = e / ls, i = e - c * ls;
      xs[e] = bp[V.base + (long long)(c0 + c) * V.ld + L.gstart + i];
    }
    __syncthreads();
    const int cw = col_width(nc), kl = 64 / cw, ng = (nc + cw - 1) / cw;
    const int sub = lane % kl, cl = lane / kl;
    if (BWD) {
      for (int t = wave; t < ls * ng; t += W) {
        const int i = t % ls, col = (t / ls) * cw + cl, cc = min(col, nc - 1);
        const T *b = bp + V.base + (long long)(c0 + cc) * V.ld;
        T s = T(0);
        const int pe = L.rowptr[i + 1];
        for (int p = L.ubnd[i] + sub; p < pe; p += 8 * kl) {
          // eight chunks of loads in flight (clamped addresses, masked use)
          T vv[8];
          int gg[8];
#pragma unroll
          for (int j = 0; j < 8; j++) {
            const int pc = min(p + j * kl, pe - 1);
            vv[j] = v[pc];
            gg[j] = L.bcol_glob[wc[pc] - ls];
          }
#pragma unroll
          for (int j = 0; j < 8; j++)
            if (p + j * kl < pe) s += vv[j] * b[gg[j]];
        }
        s = group_sum(s, kl);
        if (col < nc && sub == 0) xs[cc * ls + i] -= s;
      }
      __syncthreads();
    }
    // the first chunk of each wave's first task of the next level is loaded
    // while the current level is solved
    auto first_chunk = [&](int lev, T &pv, int &pw) {
      pv = T(0);
      pw = 0;
      if (lev >= nlev) return;
      const int r0 = lps[lev], nr = lps[lev + 1] - r0;
      if (wave >= nr * ng) return;
      const int4 q = tq[r0 + wave % nr];
      const int p = q.y + sub;
      if (p < q.z) pv = v[p], pw = wc[p];
    };
    // one level; (cv, cwi) hold this level's first chunk, (nv, nw) receive
    // the next level's (two register sets alternate: no copy waits on loads)
    auto level = [&](int lev, T cv, int cwi, T &nv, int &nw) {
      const int r0 = lps[lev], nr = lps[lev + 1] - r0;
      first_chunk(lev + 1, nv, nw);
      for (int t = wave; t < nr * ng; t += W) {
        const int4 q = tq[r0 + t % nr];
        const int col = (t / nr) * cw + cl, cc = min(col, nc - 1);
        const T *x = xs + cc * ls;
        T s = T(0);
        int p = q.y + sub;
        if (t == wave) {  // first chunk already in registers
          if (p < q.z) s = cv * x[cwi];
          p += kl;
        }
        for (; p < q.z; p += 8 * kl) {
          // eight chunks of loads in flight (clamped addresses, masked use)
          T vv[8];
          int ww[8];
#pragma unroll
          for (int j = 0; j < 8; j++) {
            const int pc = min(p + j * kl, q.z - 1);
            vv[j] = v[pc];
            ww[j] = wc[pc];
          }
#pragma unroll
          for (int j = 0; j < 8; j++)
            if (p + j * kl < q.z) s += vv[j] * x[ww[j]];
        }
        s = group_sum(s, kl);
        if (col < nc && sub == 0) {
          T *xi = xs + cc * ls + q.x;
          if (BWD) *xi = (*xi - s) / v[q.w];
          else *xi -= s;
        }
      }
      lds_barrier();
    };
    T va, vb;
    int wa, wb;
    first_chunk(0, va, wa);
    for (int lev = 0; lev < nlev; lev += 2) {
      level(lev, va, wa, vb, wb);
      if (lev + 1 < nlev) level(lev + 1, vb, wb, va, wa);
    }
    for (int e = tid; e < nc * ls; e += blockDim.x) {
      const int c = e / ls, i = e - c * ls;
      bp[V.base + (long long)(c0 + c) * V.ld + L.gstart + i] = xs[e];
    }
  }
}
template <typename T>
void launch_leaf_solve(const LeafDev *leaves, int nleaves, const VecDev *vd, T *bp, int max_leaf,
                       int max_lev, int max_rhs, bool bwd, hipStream_t st) {
  if (nleaves <= 0) return;
  constexpr int kLds = 160 * 1024;
  const int task_off = ((max_lev + 1) * 4 + 15) & ~15;
  const int xs_off = task_off + max_leaf * 16;
  const int col_bytes = max_leaf * (int)sizeof(T);
  if (xs_off + col_bytes > kLds || max_rhs < 1) {  // leaves too large for LDS
    if (bwd) launch_leaf_bwd(leaves, nleaves, vd, bp, st);
    else launch_leaf_fwd(leaves, nleaves, vd, bp, st);
    return;
  }
  const int cmax = std::max(1, std::min(std::min(max_rhs, 16), (kLds - xs_off) / std::max(col_bytes, 1)));
  const size_t lds = (size_t)xs_off + (size_t)cmax * col_bytes;
  // many leaves (batches): the LDS sized for the largest would leave most
  // of them waiting for a CU; the global-memory kernels keep them resident.
  // That pays for single-column FP64 sweeps only (512-matrix batch: 4.50 ms
  // against 4.67 with the LDS kernel); FP32 sweeps (13.5 against 14.2 ms)
  // and several columns (96^3 over 512 leaves, 16 columns: 34.7 against 55.2
  // ms) are faster queued on the LDS kernel
  const void *kern = bwd ? (const void *)k_leaf_solve<T, true> : (const void *)k_leaf_solve<T, false>;
  (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int dev = 0, cus = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kLeafSolveThreads, lds);
  if (sizeof(T) == 8 && max_rhs == 1 && (long long)nleaves > (long long)cus * std::max(per_cu, 1)) {
    if (bwd) launch_leaf_bwd(leaves, nleaves, vd, bp, st);
    else launch_leaf_fwd(leaves, nleaves, vd, bp, st);
    return;
  }
  if (bwd)
    hipLaunchKernelGGL((k_leaf_solve<T, true>), dim3(nleaves), dim3(kLeafSolveThreads), lds, st, leaves, vd, bp,
                       cmax, task_off, xs_off);
  else
    hipLaunchKernelGGL((k_leaf_solve<T, false>), dim3(nleaves), dim3(kLeafSolveThreads), lds, st, leaves, vd, bp,
                       cmax, task_off, xs_off);
}

// ---------------------------------------------------------------------------
// FP64 iterative refinement around the FP32 factors (original CSR order)
// ---------------------------------------------------------------------------
// One matrix per blockIdx.y throughout; FP64 arithmetic.
// r = b - A x with 8 lanes per row.
__global__ __launch_bounds__(256) void k_residual(const IrDev *__restrict__ tab) {
  const IrDev M = tab[blockIdx.y];
  const int sub = threadIdx.x & 7;
  for (long long row = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 3; row < M.n;
       row += ((long long)gridDim.x * blockDim.x) >> 3) {
    const int p0 = M.rowptr[row], p1 = M.rowptr[row + 1];
    for (int c = 0; c < M.nrhs; c++) {
      const double *xc = M.x + (long long)c * M.ldx;
      double s = 0.0;
      for (int p = p0 + sub; p < p1; p += 8) s += M.val[p] * xc[M.colidx[p]];
      s += __shfl_xor(s, 1, 8);
      s += __shfl_xor(s, 2, 8);
      s += __shfl_xor(s, 4, 8);
      if (sub == 0) M.r[row + (long long)c * M.ldr] = M.b[row + (long long)c * M.ldb] - s;
    }
  }
}
static unsigned ir_blocks(long long rows_x_lanes) {
  long long b = (rows_x_lanes + 255) / 256;
  return (unsigned)std::max(1LL, std::min(b, 1024LL));
}
void launch_residual(const IrDev *tab, int nmat, int max_n, hipStream_t st) {
  if (nmat <= 0 || max_n <= 0) return;
  hipLaunchKernelGGL(k_residual, dim3(ir_blocks(8LL * max_n), nmat), dim3(256), 0, st, tab);
}

__device__ __forceinline__ double block_max(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  v = red[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); i++) v = fmax(v, red[i]);
  return v;
}
// out[norm0 + c] = max_i |v[i + c ld]| for v = b (which 0), r (1) or x (2).
// Nonnegative doubles order like their bit patterns, so an integer atomic
// max on the bits is exact.
__global__ __launch_bounds__(256) void k_colmax(const IrDev *__restrict__ tab, int which,
                                                unsigned long long *out) {
  __shared__ double red[4];
  const IrDev M = tab[blockIdx.y];
  const double *v = which == 0 ? M.b : which == 1 ? M.r : M.x;
  const long long ld = which == 0 ? M.ldb : which == 1 ? M.ldr : M.ldx;
  for (int c = 0; c < M.nrhs; c++) {
    double m = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M.n;
         i += (long long)gridDim.x * blockDim.x)
      m = fmax(m, fabs(v[i + c * ld]));
    m = block_max(m, red);
    if (threadIdx.x == 0) atomicMax(out + M.norm0 + c, (unsigned long long)__double_as_longlong(m));
  }
}
// Plain streaming copy (16 B per lane per iteration): the reference kernel
// for scaling the HBM byte counters.
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src,
                                              double2 *__restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
void launch_copy(const void *src, void *dst, long long bytes, hipStream_t st) {
  const long long n = bytes / 16;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, (const double2 *)src, (double2 *)dst, n);
}
__global__ void k_zero_i32(int *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
void launch_zero_i32(int *p, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_zero_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}
__global__ void k_zero_u64(unsigned long long *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}
void launch_colmax(const IrDev *tab, int nmat, int max_n, int which, unsigned long long *out,
                   int nout, hipStream_t st) {
  if (nout > 0)
    hipLaunchKernelGGL(k_zero_u64, dim3((nout + 255) / 256), dim3(256), 0, st, out, nout);
  if (nmat <= 0 || max_n <= 0) return;
  hipLaunchKernelGGL(k_colmax, dim3(ir_blocks(max_n), nmat), dim3(256), 0, st, tab, which, out);
}
// out[matrix] = max_i sum_j |A_ij|
__global__ __launch_bounds__(256) void k_rowabs_max(const IrDev *__restrict__ tab,
                                                    unsigned long long *out) {
  __shared__ double red[4];
  const IrDev M = tab[blockIdx.y];
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M.n;
       i += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int p = M.rowptr[i]; p < M.rowptr[i + 1]; p++) s += fabs(M.val[p]);
    m = fmax(m, s);
  }
  m = block_max(m, red);
  if (threadIdx.x == 0) atomicMax(out + blockIdx.y, (unsigned long long)__double_as_longlong(m));
}
void launch_rowabs_max(const IrDev *tab, int nmat, int max_n, unsigned long long *out,
                       hipStream_t st) {
  if (nmat <= 0) return;
  hipLaunchKernelGGL(k_zero_u64, dim3((nmat + 255) / 256), dim3(256), 0, st, out, nmat);
  if (max_n <= 0) return;
  hipLaunchKernelGGL(k_rowabs_max, dim3(ir_blocks(max_n), nmat), dim3(256), 0, st, tab, out);
}
// x += dx for the matrices whose flag is set
__global__ void k_update(const IrDev *__restrict__ tab, const int *__restrict__ active) {
  const IrDev M = tab[blockIdx.y];
  if (!active[blockIdx.y]) return;
  const long long tot = (long long)M.n * M.nrhs;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const long long c = i / M.n, r = i % M.n;
    M.x[r + c * M.ldx] += M.dx[r + c * M.lddx];
  }
}
void launch_update(const IrDev *tab, int nmat, long long max_elems, const int *active,
                   hipStream_t st) {
  if (nmat <= 0 || max_elems <= 0) return;
  hipLaunchKernelGGL(k_update, dim3(ir_blocks(max_elems), nmat), dim3(256), 0, st, tab, active);
}

// Factor set-up for a set of matrices (one per blockIdx.y): zero the value
// arena and the inverse blocks, then scatter A's values into their slots.
template <typename T>
__global__ void k_prep_fill(const PrepDev *__restrict__ tab) {
  const PrepDev M = tab[blockIdx.y];
  T *a = (T *)M.arena, *v = (T *)M.inv;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M.arena_len; i += stride)
    a[i] = T(0);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M.inv_len; i += stride)
    v[i] = T(0);
}
template <typename T>
__global__ void k_prep_scatter(const PrepDev *__restrict__ tab) {
  const PrepDev M = tab[blockIdx.y];
  T *a = (T *)M.arena;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M.nnz;
       i += (long long)gridDim.x * blockDim.x)
    a[M.adest[i]] = (T)M.aval[i];
}
template <typename T>
void launch_prepare(const PrepDev *tab, int nmat, long long max_len, long long max_nnz,
                    hipStream_t st) {
  if (nmat <= 0) return;
  auto blocks = [](long long n) {
    return (unsigned)std::max(1LL, std::min((n + 255) / 256, 4096LL));
  };
  hipLaunchKernelGGL(k_prep_fill<T>, dim3(blocks(max_len), nmat), dim3(256), 0, st, tab);
  hipLaunchKernelGGL(k_prep_scatter<T>, dim3(blocks(max_nnz), nmat), dim3(256), 0, st, tab);
}

int device_lds_limit() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 65536;
  return (int)prop.sharedMemPerBlock;
}


// ---- explicit instantiations: FP64 and FP32 value paths ----
#define SNU_INSTANTIATE(T)                                                                    \
  template void launch_fill<T>(T *, long long, hipStream_t);                                  \
  template void launch_prepare<T>(const PrepDev *, int, long long, long long, hipStream_t);   \
  template void launch_assemble_schur<T>(int, const int *, const int *, const int *,          \
                                         const int *, const int *, const LeafDev *,           \
                                         hipStream_t);                                        \
  template void launch_copy2d<T>(const GemmProbT<T> *, int, const GemmSegT<T> *, int, hipStream_t); \
  template void launch_gemm<T>(const GemmProbT<T> *, int, const GemmSegT<T> *, int, int,     \
                               hipStream_t);                                                  \
  template void launch_gemm128<T>(const GemmProbT<T> *, int, const GemmSegT<T> *, int, int,  \
                                  hipStream_t);                                               \
  template void launch_gemm_variant<T>(const GemmProbT<T> *, int, const GemmSegT<T> *, int,  \
                                       int, int, int, hipStream_t);                           \
  template void launch_gemv_split<T>(const GemvProbT<T> *, int, const GemvSegT<T> *, int, int,  \
                                     int, T *, hipStream_t);                                    \
  template void launch_gemv<T>(const GemvProbT<T> *, int, const GemvSegT<T> *, int, int,     \
                               hipStream_t);                                                  \
  template void launch_diag_lu<T>(const DiagProbT<T> *, int, hipStream_t);                   \
  template void launch_permute_in<T>(const PermDev *, int, const VecDev *, T *, long long,    \
                                     hipStream_t);                                            \
  template void launch_permute_out<T>(const PermDev *, int, const VecDev *, const T *,        \
                                      long long, hipStream_t);                                \
  template void launch_leaf_fwd<T>(const LeafDev *, int, const VecDev *, T *, hipStream_t);   \
  template void launch_leaf_fwd_sep<T>(const LeafDev *, const int *, const int *, int, int,   \
                                       const VecDev *, const T *, T *, int, hipStream_t);     \
  template void launch_assemble_rhs<T>(int, const int *, const int *, const int *,            \
                                       const int *, const LeafDev *, const VecDev *,          \
                                       const T *, int, T *, int, hipStream_t);                \
  template void launch_leaf_bwd<T>(const LeafDev *, int, const VecDev *, T *, hipStream_t);   \
  template void launch_leaf_solve<T>(const LeafDev *, int, const VecDev *, T *, int, int, int,  \
                                     bool, hipStream_t);
SNU_INSTANTIATE(double)
SNU_INSTANTIATE(float)
#undef SNU_INSTANTIATE

}  // namespace snu
