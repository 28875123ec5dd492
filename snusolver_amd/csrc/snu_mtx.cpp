// Matrix Market I/O.  Counterpart of read_matrix / coo_to_csr
// (src/read_matrix.cpp:6-42,50-107) and mm_read_banner /
// mm_read_mtx_crd_size (src/mmio.cpp:89,171).
//
// Same as the reference for "coordinate real general" files: 1-based indices
// become 0-based, entries are bucketed into rows by a stable counting sort.
// Extensions (documented): "integer" and "pattern" fields (pattern values are
// 1.0), "symmetric" / "skew-symmetric" files are expanded, and each row is
// sorted by column (stable).  With SNU_MTX_FILE_ORDER the rows keep the
// file's entry order, as the reference's coo_to_csr does; with
// SNU_MTX_NO_EXPAND symmetric files are not expanded (only the stored
// triangle is read, as in the reference, src/read_matrix.cpp:73-104).
// Complex files are rejected, as in the reference (src/read_matrix.cpp:73-78).
// Dense vectors (b, x) are read from "array" files or from coordinate files
// with one column.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "snu_internal.h"

namespace snu {
namespace {

struct MtxHeader {
  int n = 0, m = 0, entries = 0;
  bool pattern = false, symmetric = false, skew = false;
};

int read_header(FILE *f, MtxHeader &h) {
  char line[4096];
  if (!fgets(line, sizeof line, f)) return set_error(SNU_ERR_BAD_INPUT, "mtx: empty file");
  std::string s(line);
  for (auto &c : s) c = (char)tolower((unsigned char)c);
  if (s.rfind("%%matrixmarket", 0) != 0) return set_error(SNU_ERR_BAD_INPUT, "mtx: missing banner");
  if (s.find("matrix") == std::string::npos || s.find("coordinate") == std::string::npos)
    return set_error(SNU_ERR_BAD_INPUT, "mtx: only 'matrix coordinate' is supported");
  if (s.find("complex") != std::string::npos) return set_error(SNU_ERR_BAD_INPUT, "mtx: complex matrices are not supported");
  if (s.find("hermitian") != std::string::npos) return set_error(SNU_ERR_BAD_INPUT, "mtx: hermitian not supported");
  h.pattern = s.find("pattern") != std::string::npos;
  h.skew = s.find("skew-symmetric") != std::string::npos;
  h.symmetric = !h.skew && s.find("symmetric") != std::string::npos;
  do {
    if (!fgets(line, sizeof line, f)) return set_error(SNU_ERR_BAD_INPUT, "mtx: missing size line");
  } while (line[0] == '%' || line[0] == '\n' || line[0] == '\r');
  if (sscanf(line, "%d %d %d", &h.n, &h.m, &h.entries) != 3 || h.n < 0 || h.m < 0 || h.entries < 0)
    return set_error(SNU_ERR_BAD_INPUT, "mtx: bad size line");
  return SNU_OK;
}

int read_all(const char *path, MtxHeader &h, std::vector<int> &I,
             std::vector<int> &J, std::vector<double> &V, bool expand) {
  FILE *f = fopen(path, "r");
  if (!f) return set_error(SNU_ERR_BAD_INPUT, std::string("mtx: cannot open ") + path);
  int rc = read_header(f, h);
  if (rc) { fclose(f); return rc; }
  I.reserve((size_t)h.entries * (expand && (h.symmetric || h.skew) ? 2 : 1));
  for (int e = 0; e < h.entries; e++) {
    int i, j;
    double v = 1.0;
    int got = h.pattern ? fscanf(f, "%d %d", &i, &j) : fscanf(f, "%d %d %lg", &i, &j, &v);
    if (got != (h.pattern ? 2 : 3)) { fclose(f); return set_error(SNU_ERR_BAD_INPUT, "mtx: truncated entry list"); }
    if (i < 1 || i > h.n || j < 1 || j > h.m) { fclose(f); return set_error(SNU_ERR_BAD_INPUT, "mtx: index out of range"); }
    I.push_back(i - 1); J.push_back(j - 1); V.push_back(v);
    if (expand && (h.symmetric || h.skew) && i != j) {
      I.push_back(j - 1); J.push_back(i - 1); V.push_back(h.skew ? -v : v);
    }
  }
  fclose(f);
  return SNU_OK;
}

}  // namespace
}  // namespace snu

using namespace snu;


extern "C" int snu_mtx_size_ex(const char *path, int flags, int *n, int *m, int *nnz,
                               int *symmetric) {
  FILE *f = fopen(path, "r");
  if (!f) return set_error(SNU_ERR_BAD_INPUT, std::string("mtx: cannot open ") + path);
  MtxHeader h;
  int rc = read_header(f, h);
  if (rc) { fclose(f); return rc; }
  long long cnt = h.entries;
  if ((h.symmetric || h.skew) && !(flags & SNU_MTX_NO_EXPAND)) {
    // count off-diagonal entries to size the expansion
    cnt = 0;
    for (int e = 0; e < h.entries; e++) {
      int i, j;
      double v;
      int got = h.pattern ? fscanf(f, "%d %d", &i, &j) : fscanf(f, "%d %d %lg", &i, &j, &v);
      if (got != (h.pattern ? 2 : 3)) { fclose(f); return set_error(SNU_ERR_BAD_INPUT, "mtx: truncated entry list"); }
      cnt += (i != j) ? 2 : 1;
    }
  }
  fclose(f);
  if (cnt > 2147483647LL) return set_error(SNU_ERR_BAD_INPUT, "mtx: too many entries for int32");
  *n = h.n; *m = h.m; *nnz = (int)cnt;
  if (symmetric) *symmetric = (h.symmetric || h.skew) ? 1 : 0;
  return SNU_OK;
}

extern "C" int snu_mtx_read_ex(const char *path, int flags, int *rowptr, int *colidx,
                               double *val) {
  MtxHeader h;
  std::vector<int> I, J;
  std::vector<double> V;
  int rc = read_all(path, h, I, J, V, !(flags & SNU_MTX_NO_EXPAND));
  if (rc) return rc;
  int nz = (int)I.size();
  // stable counting sort into rows (coo_to_csr, src/read_matrix.cpp:6-42)
  std::vector<int> cnt(h.n + 1, 0);
  for (int e = 0; e < nz; e++) cnt[I[e] + 1]++;
  for (int i = 0; i < h.n; i++) cnt[i + 1] += cnt[i];
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  std::vector<int> ci(nz);
  std::vector<double> cv(nz);
  for (int e = 0; e < nz; e++) {
    int d = pos[I[e]]++;
    ci[d] = J[e];
    cv[d] = V[e];
  }
  if (flags & SNU_MTX_FILE_ORDER) {
    std::copy(ci.begin(), ci.end(), colidx);
    std::copy(cv.begin(), cv.end(), val);
  } else {  // column-sorted rows (stable)
    std::vector<int> idx;
    for (int i = 0; i < h.n; i++) {
      int a = cnt[i], b = cnt[i + 1];
      idx.resize(b - a);
      for (int t = 0; t < b - a; t++) idx[t] = a + t;
      std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return ci[x] < ci[y]; });
      for (int t = 0; t < b - a; t++) colidx[a + t] = ci[idx[t]], val[a + t] = cv[idx[t]];
    }
  }
  for (int i = 0; i <= h.n; i++) rowptr[i] = cnt[i];
  return SNU_OK;
}

extern "C" int snu_mtx_size(const char *path, int *n, int *m, int *nnz,
                            int *symmetric) {
  return snu_mtx_size_ex(path, 0, n, m, nnz, symmetric);
}

extern "C" int snu_mtx_read(const char *path, int *rowptr, int *colidx,
                            double *val) {
  return snu_mtx_read_ex(path, 0, rowptr, colidx, val);
}

extern "C" int snu_mtx_read_vector(const char *path, int n, double *x) {
  FILE *f = fopen(path, "r");
  if (!f) return set_error(SNU_ERR_BAD_INPUT, std::string("mtx: cannot open ") + path);
  char line[4096];
  if (!fgets(line, sizeof line, f)) { fclose(f); return set_error(SNU_ERR_BAD_INPUT, "mtx: empty file"); }
  std::string s(line);
  for (auto &c : s) c = (char)tolower((unsigned char)c);
  if (s.rfind("%%matrixmarket", 0) != 0) { fclose(f); return set_error(SNU_ERR_BAD_INPUT, "mtx: missing banner"); }
  if (s.find("complex") != std::string::npos || s.find("pattern") != std::string::npos) {
    fclose(f);
    return set_error(SNU_ERR_BAD_INPUT, "mtx: a vector must be real or integer");
  }
  const bool array = s.find("array") != std::string::npos;
  do {
    if (!fgets(line, sizeof line, f)) { fclose(f); return set_error(SNU_ERR_BAD_INPUT, "mtx: missing size line"); }
  } while (line[0] == '%' || line[0] == '\n' || line[0] == '\r');
  int rows = 0, cols = 0, entries = 0;
  const int want = array ? 2 : 3;
  if (sscanf(line, "%d %d %d", &rows, &cols, &entries) < want || rows != n || cols != 1) {
    fclose(f);
    return set_error(SNU_ERR_BAD_INPUT, "mtx: expected an n x 1 vector of length " + std::to_string(n));
  }
  if (array) {
    for (int i = 0; i < n; i++)
      if (fscanf(f, "%lg", &x[i]) != 1) { fclose(f); return set_error(SNU_ERR_BAD_INPUT, "mtx: truncated vector"); }
  } else {
    std::fill(x, x + n, 0.0);
    for (int e = 0; e < entries; e++) {
      int i, j;
      double v;
      if (fscanf(f, "%d %d %lg", &i, &j, &v) != 3) { fclose(f); return set_error(SNU_ERR_BAD_INPUT, "mtx: truncated vector"); }
      if (i < 1 || i > n || j != 1) { fclose(f); return set_error(SNU_ERR_BAD_INPUT, "mtx: index out of range"); }
      x[i - 1] = v;
    }
  }
  fclose(f);
  return SNU_OK;
}

extern "C" int snu_mtx_write_vector(const char *path, const double *x,
                                    int n) {
  FILE *f = fopen(path, "w");
  if (!f) return set_error(SNU_ERR_BAD_INPUT, std::string("cannot write ") + path);
  fprintf(f, "%%%%MatrixMarket matrix array real general\n%d 1\n", n);
  for (int i = 0; i < n; i++) fprintf(f, "%.17g\n", x[i]);
  fclose(f);
  return SNU_OK;
}
