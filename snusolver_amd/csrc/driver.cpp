// Driver: .mtx in, x out.  Counterpart of the reference's driver
// (src/driver.cpp:30-63): reads the matrix, draws b, solves, checks.
// Differences: b = A x_true with x_true ~ U[-1,1) from a seeded stream
// (the reference draws b from an unseeded std::random_device, :48-53), x is
// written to a file, and the check prints the normalised backward error
// ||Ax-b||_inf / (||A||_inf ||x||_inf + ||b||_inf) besides max|Ax-b|.
//
// usage: snu_driver <matrix.mtx> [--np N] [--out x.mtx] [--seed S]
//                   [--device D] [--grid nx ny nz] [--nrhs K] [--b b.mtx]
//                   [--fp32] [--fp32-mfma] [--file-order]
// --b reads one right-hand side from a Matrix Market vector file instead of
// drawing it; --file-order reads the matrix as the reference's read_matrix
// does (file order within rows, no symmetric expansion); --fp32 factors in
// FP32 (GEMM tiles as bf16 splits) and refines in FP64; --fp32-mfma does the
// same with the GEMM tiles on the FP32 MFMA.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/snusolver.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "Usage: %s [matrix-market-filename] [--np N] [--out file] [--seed S] [--device D] "
                 "[--grid nx ny nz] [--nrhs K] [--b file] [--fp32] [--fp32-mfma] [--file-order]\n", argv[0]);
    return 1;
  }
  snu_opts_t o;
  snu_default_opts(&o);
  const char *out = nullptr, *bfile = nullptr;
  unsigned long long seed = 1;
  int nrhs = 1, flags = 0;
  for (int i = 2; i < argc; i++) {
    if (!std::strcmp(argv[i], "--np") && i + 1 < argc) o.np = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
    else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
    else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) o.device = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--nrhs") && i + 1 < argc) nrhs = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--b") && i + 1 < argc) bfile = argv[++i];
    else if (!std::strcmp(argv[i], "--fp32")) o.precision = SNU_FP32;
    else if (!std::strcmp(argv[i], "--fp32-mfma")) o.precision = SNU_FP32, o.fp32_gemm = SNU_GEMM_F32;
    else if (!std::strcmp(argv[i], "--file-order")) flags = SNU_MTX_FILE_ORDER | SNU_MTX_NO_EXPAND;
    else if (!std::strcmp(argv[i], "--grid") && i + 3 < argc) {
      o.ordering = SNU_ORDER_GRID;
      o.grid[0] = std::atoi(argv[++i]);
      o.grid[1] = std::atoi(argv[++i]);
      o.grid[2] = std::atoi(argv[++i]);
    }
  }
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
  std::printf("Read matrix done! n=%d nnz=%d\n", n, nnz);
  snu_csr_t A{n, m, nnz, rp.data(), ci.data(), v.data()};
  std::vector<double> xt((size_t)n * nrhs), b((size_t)n * nrhs), x((size_t)n * nrhs);
  if (bfile) {
    if (nrhs != 1) {
      std::fprintf(stderr, "--b reads one right-hand side (--nrhs 1)\n");
      return 1;
    }
    if (snu_mtx_read_vector(bfile, n, b.data())) {
      std::fprintf(stderr, "%s\n", snu_last_error());
      return 1;
    }
  } else {
    snu_gen_rhs(&A, nrhs, seed, xt.data(), b.data());
  }
  auto t0 = std::chrono::steady_clock::now();
  snumat_t *h = nullptr;
  int rc = snumat_construct(&A, &o, &h);
  auto t1 = std::chrono::steady_clock::now();
  if (!rc) rc = snumat_factorize(h);
  auto t2 = std::chrono::steady_clock::now();
  if (!rc) rc = snumat_solve(h, b.data(), x.data(), nrhs, n);
  auto t3 = std::chrono::steady_clock::now();
  if (rc) {
    std::fprintf(stderr, "snusolver error %d: %s\n", rc, snu_last_error());
    snumat_destroy(h);
    return 2;
  }
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  std::printf("construct %.3f ms, factorize %.3f ms, solve %.3f ms\n", ms(t0, t1), ms(t1, t2), ms(t2, t3));
  for (int j = 0; j < nrhs; j++) {
    double M = 0;
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int q = rp[i]; q < rp[i + 1]; q++) s += v[q] * x[(size_t)j * n + ci[q]];
      s -= b[(size_t)j * n + i];
      M = std::fmax(M, std::fabs(s));
    }
    std::printf("rhs %d: max |Ax-b| = %e, backward error = %e\n", j, M,
                snu_backward_error(&A, x.data() + (size_t)j * n, b.data() + (size_t)j * n));
  }
  if (out) snu_mtx_write_vector(out, x.data(), n);
  snumat_destroy(h);
  return 0;
}
