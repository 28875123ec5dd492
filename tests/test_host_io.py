"""Host side of the boundary (no GPU): CSR rows in any column order,
Matrix Market reading modes (sorted / file order, symmetric expansion on or
off), vector files, and the drivers' usage messages."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _shuffle_rows(S, A, seed):
    rng = np.random.default_rng(seed)
    ci, v = A.colidx.copy(), A.data.copy()
    for i in range(A.n):
        a, b = A.rowptr[i], A.rowptr[i + 1]
        p = rng.permutation(b - a)
        ci[a:b], v[a:b] = ci[a:b][p], v[a:b][p]
    return S.CSR(A.n, A.rowptr.copy(), ci, v)


def test_unsorted_rows_same_structure(S):
    """Rows in any column order give the structure of the sorted matrix:
    same ordering, same PAP^T, and perm_map sends every entry (in the
    caller's order) to the slot of its (row, col)."""
    A = S.gen_grid(12, 11, 1, 9, 0, 8.1, 2)
    U = _shuffle_rows(S, A, 5)
    assert any(np.any(np.diff(U.colidx[U.rowptr[i]:U.rowptr[i + 1]]) < 0) for i in range(U.n))
    with S.SnuMat(A, 4, device=-1) as M, S.SnuMat(U, 4, device=-1) as N:
        oa, sa = M.ordering()
        ou, su = N.ordering()
        assert np.array_equal(oa, ou) and np.array_equal(sa, su)
        rpa, cia, pma = M.permuted()
        rpu, ciu, pmu = N.permuted()
        assert np.array_equal(rpa, rpu) and np.array_equal(cia, ciu)
        # entry (r, c) of either matrix lands on the same merged slot
        slot = {}
        rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
        for k in range(A.nnz):
            slot[(rows[k], A.colidx[k])] = pma[k]
        rows_u = np.repeat(np.arange(U.n), np.diff(U.rowptr))
        for k in range(U.nnz):
            assert pmu[k] == slot[(rows_u[k], U.colidx[k])]


def test_unsorted_duplicate_rejected(S):
    A = S.gen_grid(6, 6, 1, 5, 0, 4.1, 0)
    ci = A.colidx.copy()
    a, b = A.rowptr[7], A.rowptr[8]
    ci[a:b] = ci[a:b][::-1]
    ci[a] = ci[b - 1]  # duplicate inside an unsorted row
    D = S.CSR(A.n, A.rowptr, ci, A.data)
    with pytest.raises(S.SnuError) as e:
        S.SnuMat(D, 4, device=-1)
    assert e.value.code == -4


def test_fp32_gemm_option_checked(S):
    """fp32_gemm takes SNU_GEMM_SPLIT or SNU_GEMM_F32 only (host-only
    construct: the check needs no GPU)."""
    import ctypes as C
    A = S.gen_grid(6, 6, 1, 5, 0, 4.1, 0)
    for form in ("split", "f32"):
        with S.SnuMat(A, 4, device=-1, precision="fp32", fp32_gemm=form):
            pass
    with pytest.raises(ValueError):
        S.SnuMat(A, 4, device=-1, precision="fp32", fp32_gemm="bf16")
    o = S._Opts()
    S.lib().snu_default_opts(C.byref(o))
    assert o.fp32_gemm == S.GEMM_SPLIT
    o.np, o.device, o.precision, o.fp32_gemm = 4, -1, S.FP32, 7
    csr = S._Csr(A.n, A.n, A.nnz, A.rowptr.ctypes.data_as(C.POINTER(C.c_int)),
                 A.colidx.ctypes.data_as(C.POINTER(C.c_int)), A.data.ctypes.data_as(C.POINTER(C.c_double)))
    h = C.c_void_p()
    rc = S.lib().snumat_construct(C.byref(csr), C.byref(o), C.byref(h))
    assert rc == -1 and not h.value, rc
    assert b"fp32_gemm" in S.lib().snu_last_error()


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_mtx_reading_modes(S):
    text = ("%%MatrixMarket matrix coordinate real symmetric\n% comment\n3 3 4\n"
            "3 1 -1.5\n1 1 4\n2 2 5\n3 3 6\n")
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "s.mtx")
        _write(p, text)
        E = S.read_mtx(p)  # expanded, sorted
        assert E.nnz == 5
        assert np.array_equal(E.to_scipy().toarray(), [[4, 0, -1.5], [0, 5, 0], [-1.5, 0, 6]])
        R = S.read_mtx(p, file_order=True, expand=False)  # the reference's reading
        assert R.nnz == 4
        assert list(R.rowptr) == [0, 1, 2, 4]
        assert list(R.colidx) == [0, 1, 0, 2] and list(R.data) == [4, 5, -1.5, 6]
        g = os.path.join(d, "g.mtx")
        _write(g, "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 2 7\n1 1 1\n2 2 2\n")
        F = S.read_mtx(g, file_order=True)
        assert list(F.colidx) == [1, 0, 1] and list(F.data) == [7, 1, 2]
        Q = S.read_mtx(g)
        assert list(Q.colidx) == [0, 1, 1] and list(Q.data) == [1, 7, 2]


def test_vector_files(S):
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "x.mtx")
        x = np.array([1.25, -3.0, 1e-300, 7.0])
        S.write_vector(p, x)
        assert np.array_equal(S.read_vector(p, 4), x)
        c = os.path.join(d, "c.mtx")
        _write(c, "%%MatrixMarket matrix coordinate real general\n4 1 2\n2 1 5.5\n4 1 -1\n")
        assert np.array_equal(S.read_vector(c, 4), [0, 5.5, 0, -1])
        with pytest.raises(S.SnuError):
            S.read_vector(p, 5)


def test_golden_matrix_file_order(S):
    """matrix_20.mtx read in file order has the same entries as the sorted
    reading (rows permuted within themselves)."""
    p = os.path.join(GOLD, "matrix_20.mtx")
    A, F = S.read_mtx(p), S.read_mtx(p, file_order=True, expand=False)
    assert np.array_equal(A.rowptr, F.rowptr)
    assert np.array_equal(A.to_scipy().toarray(), F.to_scipy().toarray())


@pytest.mark.parametrize("name", ["snu_driver", "snu_ref_driver"])
def test_drivers_usage(name):
    r = subprocess.run([os.path.join(ROOT, "snusolver_amd", name)], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr


def test_wide_leaf_fronts(S):
    """A single leaf of 40,000 rows (2D 5-point grid, np = 1) is factored as
    fronts: the pivots are partitioned among the fronts, the largest front is
    of the order of the widest separator (not of the leaf), and more leaves
    give smaller fronts."""
    A = S.gen_grid(200, 200, 1, 5, 0, 4.1, 0)
    biggest = []
    for np_ in (1, 4):
        with S.SnuMat(A, np_, grid=(200, 200, 1), device=-1) as M:
            f = M.leaf_fronts(0)
            ls = M.leaf_sizes(0)[0]
            assert np.array_equal(np.sort(f["k"]), np.arange(ls))
            assert np.all(np.diff(f["k_ptr"]) == f["nk"])
            biggest.append(int(f["f"].max()))
            assert biggest[-1] <= 3 * 200
    assert biggest[0] >= biggest[1]
