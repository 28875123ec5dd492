// Synthetic grid operators and right-hand sides (seeded, reproducible).
//
// RNG: splitmix64; a uniform draw is u = (next() >> 11) * 2^-53 in [0, 1).
// Vertex index v = x + nx * (y + ny * z).  Rows are generated in v order and
// columns ascending; one draw per off-diagonal entry, in that order.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "snu_internal.h"

namespace snu {

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double u01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

// Neighbour offsets of a stencil, sorted so that the column index of a
// neighbour grows with the offset's position (z, then y, then x).
static int stencil_offsets(int stencil, int off[27][3]) {
  int k = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        int nzc = (dx != 0) + (dy != 0) + (dz != 0);
        bool keep = false;
        if (stencil == 5) keep = dz == 0 && nzc <= 1;
        else if (stencil == 9) keep = dz == 0;
        else if (stencil == 7) keep = nzc <= 1;
        else if (stencil == 27) keep = true;
        if (keep) off[k][0] = dx, off[k][1] = dy, off[k][2] = dz, k++;
      }
  return k;
}

static bool valid_stencil(int nx, int ny, int nz, int stencil) {
  if (nx < 1 || ny < 1 || nz < 1) return false;
  if (stencil == 5 || stencil == 9) return nz == 1;
  return stencil == 7 || stencil == 27;
}

}  // namespace snu

using namespace snu;

extern "C" int snu_gen_grid_size(int nx, int ny, int nz, int stencil, int *n,
                                 int *nnz) {
  if (!valid_stencil(nx, ny, nz, stencil)) return set_error(SNU_ERR_BAD_INPUT, "bad grid/stencil");
  int off[27][3];
  int k = stencil_offsets(stencil, off);
  long long cnt = 0;
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        for (int t = 0; t < k; t++) {
          int X = x + off[t][0], Y = y + off[t][1], Z = z + off[t][2];
          if (X >= 0 && X < nx && Y >= 0 && Y < ny && Z >= 0 && Z < nz) cnt++;
        }
  long long nn = (long long)nx * ny * nz;
  if (nn > 2147483647LL || cnt > 2147483647LL) return set_error(SNU_ERR_BAD_INPUT, "grid too large for int32 indices");
  *n = (int)nn;
  *nnz = (int)cnt;
  return SNU_OK;
}

extern "C" int snu_gen_grid(int nx, int ny, int nz, int stencil, int recipe,
                            double diag, uint64_t seed, int *rowptr,
                            int *colidx, double *val) {
  if (!valid_stencil(nx, ny, nz, stencil) || (recipe != 0 && recipe != 1))
    return set_error(SNU_ERR_BAD_INPUT, "bad grid/stencil/recipe");
  int off[27][3];
  int k = stencil_offsets(stencil, off);
  const double axis_w[3] = {1.0, 10.0, 100.0};
  uint64_t s = seed;
  int p = 0;
  rowptr[0] = 0;
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        int v = x + nx * (y + ny * z);
        int dpos = -1;
        double rowsum = 0.0;
        for (int t = 0; t < k; t++) {
          int X = x + off[t][0], Y = y + off[t][1], Z = z + off[t][2];
          if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
          int c = X + nx * (Y + ny * Z);
          colidx[p] = c;
          if (c == v) {
            dpos = p;
            val[p] = 0.0;
          } else if (recipe == 0) {
            val[p] = -1.0 + 0.25 * (u01(s) - 0.5);
          } else {
            // anisotropic: face neighbour along axis a couples with w_a;
            // an edge/corner neighbour with the mean of its axes' weights
            // divided by the number of axes involved.
            int na = 0;
            double wsum = 0.0;
            for (int a = 0; a < 3; a++)
              if (off[t][a]) na++, wsum += axis_w[a];
            double w = (wsum / na) / na;
            val[p] = -w * (1.0 + 0.1 * (2.0 * u01(s) - 1.0));
            rowsum += std::fabs(val[p]);
          }
          p++;
        }
        val[dpos] = recipe == 0 ? diag : 1.01 * rowsum;
        rowptr[v + 1] = p;
      }
  return SNU_OK;
}

extern "C" int snu_gen_rhs(const snu_csr_t *A, int nrhs, uint64_t seed,
                           double *x_true, double *b) {
  if (!A || nrhs < 1) return set_error(SNU_ERR_BAD_INPUT, "bad rhs request");
  uint64_t s = seed;
  int n = A->n;
  for (int j = 0; j < nrhs; j++)
    for (int i = 0; i < n; i++) x_true[(size_t)j * n + i] = 2.0 * u01(s) - 1.0;
  for (int j = 0; j < nrhs; j++)
    for (int i = 0; i < n; i++) {
      double sum = 0.0;
      for (int q = A->rowptr[i]; q < A->rowptr[i + 1]; q++)
        sum += A->data[q] * x_true[(size_t)j * n + A->colidx[q]];
      b[(size_t)j * n + i] = sum;
    }
  return SNU_OK;
}

extern "C" double snu_backward_error(const snu_csr_t *A, const double *x,
                                     const double *b) {
  double rmax = 0, amax = 0, xmax = 0, bmax = 0;
  for (int i = 0; i < A->n; i++) {
    double s = 0, a = 0;
    for (int q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) {
      s += A->data[q] * x[A->colidx[q]];
      a += std::fabs(A->data[q]);
    }
    rmax = std::fmax(rmax, std::fabs(s - b[i]));
    amax = std::fmax(amax, a);
    xmax = std::fmax(xmax, std::fabs(x[i]));
    bmax = std::fmax(bmax, std::fabs(b[i]));
  }
  double den = amax * xmax + bmax;
  return den > 0 ? rmax / den : rmax;
}
