"""Host utilities: synthetic workload generators (recipes of the named
configurations), right-hand sides, Matrix Market reader/writer extensions."""
import os

import numpy as np
import pytest


def test_c1_c2_sizes_and_recipe(S):
    A = S.gen_grid(20, 20, 1, 5, 0, 4.1, 0)
    assert (A.n, A.nnz) == (400, 1920)
    A2, grid, np_ = S.config_matrix("c2")
    assert (A2.n, A2.nnz, np_) == (262144, 1810432, 128)
    for M, dg in ((A, 4.1), (A2, 6.1)):
        rows = np.repeat(np.arange(M.n), np.diff(M.rowptr))
        diag = M.data[rows == M.colidx]
        off = M.data[rows != M.colidx]
        assert np.all(diag == dg)
        assert off.min() >= -1.125 and off.max() <= -0.875  # -1 + 0.25 (u - 0.5)


def test_c3_size_and_diagonal(S):
    n, nnz = 96 ** 3, 23393656
    import ctypes as C
    a, b = C.c_int(), C.c_int()
    assert S.lib().snu_gen_grid_size(96, 96, 96, 27, C.byref(a), C.byref(b)) == 0
    assert (a.value, b.value) == (n, nnz)
    A = S.gen_grid(8, 7, 6, 27, 1, 0.0, 0)
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    off = np.where(rows != A.colidx, np.abs(A.data), 0.0)
    rowsum = np.bincount(rows, weights=off, minlength=A.n)
    diag = A.data[rows == A.colidx]
    assert np.allclose(diag, 1.01 * rowsum, rtol=1e-15)
    # face couplings along x, y, z carry 1:10:100 weights within +-10%
    r = 3 + 8 * (3 + 7 * 3)
    cols = A.colidx[A.rowptr[r]:A.rowptr[r + 1]]
    vals = A.data[A.rowptr[r]:A.rowptr[r + 1]]
    for d, w in ((1, 1.0), (8, 10.0), (56, 100.0)):
        v = -vals[cols == r + d][0]
        assert 0.9 * w <= v <= 1.1 * w


def test_rhs_is_A_times_xtrue(S):
    A = S.gen_grid(9, 9, 1, 9, 0, 8.1, 2)
    xt, b = S.gen_rhs(A, 3, 7)
    assert xt.shape == (81, 3) and np.all(np.abs(xt) <= 1.0)
    assert np.allclose(A.matvec(xt), b, rtol=0, atol=1e-13)
    xt2, _ = S.gen_rhs(A, 3, 7)
    assert np.array_equal(xt, xt2)  # seeded


def test_mtx_symmetric_pattern_extensions(S, tmp_path):
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n% c\n3 3 4\n1 1 2.0\n2 1 -1.0\n"
                 "3 2 -0.5\n3 3 4.0\n")
    A = S.read_mtx(str(p))
    assert A.nnz == 6
    D = A.to_scipy().toarray()
    assert np.allclose(D, D.T) and D[0, 1] == -1.0 and D[1, 1] == 0.0
    q = tmp_path / "p.mtx"
    q.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 3\n1 1\n2 2\n1 2\n")
    B = S.read_mtx(str(q))
    assert np.array_equal(B.colidx, [0, 1, 1]) and np.all(B.data == 1.0)  # rows column-sorted
    c = tmp_path / "c.mtx"
    c.write_text("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n")
    with pytest.raises(S.SnuError):
        S.read_mtx(str(c))
    t = tmp_path / "t.mtx"
    t.write_text("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n")
    with pytest.raises(S.SnuError):
        S.read_mtx(str(t))


def test_vector_writer_roundtrip(S, tmp_path):
    x = np.array([1.0, -2.5e-300, 3.141592653589793])
    p = str(tmp_path / "x.mtx")
    assert S.lib().snu_mtx_write_vector(p.encode(), x.ctypes.data_as(__import__("ctypes").c_void_p), 3) == 0
    lines = open(p).read().split("\n")
    assert lines[0].startswith("%%MatrixMarket matrix array real general")
    assert np.array_equal(np.array([float(v) for v in lines[2:5]]), x)
