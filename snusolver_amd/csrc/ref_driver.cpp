// The reference's driver flow (src/driver.cpp:30-63) written against the
// C++ compatibility header, to show that a caller of the reference's
// initialize() / solve(csr_matrix, double*, double*) (src/snusolver.h:5-6)
// links against this library unchanged.  Reads the matrix the way the
// reference's read_matrix does (file order within rows, no symmetric
// expansion), draws b ~ U[0, 1) (std::mt19937 as the reference, but seeded),
// calls initialize() and solve(), and prints check()'s max |Ax - b|
// (src/driver.cpp:12-28).
//
// usage: snu_ref_driver <matrix.mtx> [--b b.mtx] [--out x.mtx] [--seed S]
// The leaf count comes from SNU_NP (default 4), the device from SNU_DEVICE.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/snusolver_compat.hpp"

static void check(csr_matrix A, double *b, double *x) {
  double M = 0;
  for (int i = 0; i < A.n; i++) {
    double sum = 0.0;
    for (int p = A.rowptr[i]; p < A.rowptr[i + 1]; p++) sum += A.data[p] * x[A.colidx[p]];
    sum -= b[i];
    if (M < sum) M = sum;
    if (M < -sum) M = -sum;
  }
  std::printf("max |Ax-b| = %e\n", M);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "Usage: %s [matrix-market-filename] [--b b.mtx] [--out x.mtx] [--seed S]\n", argv[0]);
    return 1;
  }
  const char *bfile = nullptr, *out = nullptr;
  unsigned seed = 0;
  for (int i = 2; i < argc; i++) {
    if (!std::strcmp(argv[i], "--b") && i + 1 < argc) bfile = argv[++i];
    else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
    else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) seed = (unsigned)std::strtoul(argv[++i], nullptr, 10);
  }
  initialize();
  const int flags = SNU_MTX_FILE_ORDER | SNU_MTX_NO_EXPAND;
  int n, m, nnz, sym;
  if (snu_mtx_size_ex(argv[1], flags, &n, &m, &nnz, &sym)) {
    std::fprintf(stderr, "%s\n", snu_last_error());
    return 1;
  }
  std::vector<int> rp(n + 1), ci(nnz);
  std::vector<double> v(nnz);
  if (snu_mtx_read_ex(argv[1], flags, rp.data(), ci.data(), v.data())) {
    std::fprintf(stderr, "%s\n", snu_last_error());
    return 1;
  }
  std::printf("Read matrix done!\n");
  csr_matrix A{n, m, nnz, rp.data(), ci.data(), v.data()};
  std::vector<double> b(n), x(n, 0.0);
  if (bfile) {
    if (snu_mtx_read_vector(bfile, n, b.data())) {
      std::fprintf(stderr, "%s\n", snu_last_error());
      return 1;
    }
  } else {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<> dis(0.0, 1.0);
    for (int i = 0; i < n; i++) b[i] = dis(gen);
  }
  solve(A, b.data(), x.data());
  check(A, b.data(), x.data());
  if (out && snu_mtx_write_vector(out, x.data(), n)) {
    std::fprintf(stderr, "%s\n", snu_last_error());
    return 1;
  }
  return 0;
}
