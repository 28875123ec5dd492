"""Writes the golden fixtures of this directory (run from the repo root).

Independent of both the product and the oracle: scipy reads the Matrix
Market file and solves with SuperLU (partial pivoting).

  matrix_20_csr.npz   data/matrix_20.mtx of the reference, read by
                      scipy.io.mmread -> CSR (the reference's only fixture)
  matrix_20_x.npz     b = A x_true (x_true ~ U[-1,1), splitmix64 seed 1),
                      x = spsolve(A, b)
  c1.npz              20x20 5-point operator of the C1 recipe, x_true/b
                      (seed 1) and x = spsolve(A, b)
"""
import os
import sys

import numpy as np
import scipy.io
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def csr_of(path):
    A = sp.csr_matrix(scipy.io.mmread(path))
    A.sort_indices()
    return A


def main():
    import snusolver_amd as S

    A = csr_of(os.path.join(HERE, "matrix_20.mtx"))
    np.savez(os.path.join(HERE, "matrix_20_csr.npz"), rowptr=A.indptr.astype(np.int32),
             colidx=A.indices.astype(np.int32), data=A.data, n=A.shape[0])
    Ac = S.CSR(A.shape[0], A.indptr, A.indices, A.data)
    xt, b = S.gen_rhs(Ac, 1, 1)
    x = spla.spsolve(A.tocsc(), b[:, 0])
    np.savez(os.path.join(HERE, "matrix_20_x.npz"), x_true=xt[:, 0], b=b[:, 0], x=x)

    C1 = S.gen_grid(20, 20, 1, 5, 0, 4.1, 0)
    xt, b = S.gen_rhs(C1, 1, 1)
    x = spla.spsolve(C1.to_scipy().tocsc(), b[:, 0])
    np.savez(os.path.join(HERE, "c1.npz"), rowptr=C1.rowptr, colidx=C1.colidx, data=C1.data,
             x_true=xt[:, 0], b=b[:, 0], x=x)


if __name__ == "__main__":
    main()
