// C++ shims with the reference's initialize()/solve() signatures
// (src/snusolver.h:5-6), forwarding to the C-ABI.
#include "../../include/snusolver_compat.hpp"

#include <cstdio>
#include <cstdlib>

static int env_int(const char *name, int def) {
  const char *v = std::getenv(name);
  return v ? std::atoi(v) : def;
}

void initialize() {
  int rc = snu_initialize(env_int("SNU_DEVICE", 0));
  if (rc) std::fprintf(stderr, "snusolver: initialize failed: %s\n", snu_last_error());
}

void solve(csr_matrix A_csr, double *b, double *x) {
  snu_csr_t A{A_csr.n, A_csr.m, A_csr.nnz, A_csr.rowptr, A_csr.colidx, A_csr.data};
  snu_opts_t o;
  snu_default_opts(&o);
  o.np = env_int("SNU_NP", 4);
  o.device = env_int("SNU_DEVICE", 0);
  int rc = snu_solve_csr(&A, b, x, &o);
  if (rc) std::fprintf(stderr, "snusolver: solve failed (%d): %s\n", rc, snu_last_error());
}
