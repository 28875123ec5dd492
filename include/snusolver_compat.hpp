// snusolver_compat.hpp -- C++ shims with the reference's exact entry points,
// so a driver written against src/snusolver.h:5-6 links unchanged:
//   void initialize();                              (src/solver.cu:35-49)
//   void solve(csr_matrix A, double *b, double *x); (src/solver.cu:51-89)
// csr_matrix has the field layout of src/matrix.h:9-13.  Options (np, device,
// ordering) come from the environment: SNU_NP (default 4), SNU_DEVICE
// (default 0).  Errors print snu_last_error() and leave x untouched.
#ifndef SNUSOLVER_AMD_COMPAT_HPP
#define SNUSOLVER_AMD_COMPAT_HPP
#include "snusolver.h"

#ifndef SNUSOLVER_MATRIX
typedef struct csr_matrix {
  int n, m, nnz;
  int *rowptr, *colidx;
  double *data;
} csr_matrix;
#endif

void initialize();
void solve(csr_matrix A_csr, double *b, double *x);
#endif
