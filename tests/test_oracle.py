"""Pins the CPU oracle (the checker) against what fixes its answers
independently of it:

* golden vectors written by scipy (tests/golden, make_golden.py): the
  reference's data/matrix_20.mtx read by scipy.io.mmread, and SuperLU
  (partial pivoting) solutions of matrix_20 and C1;
* brute force on tiny cores: boolean dense elimination (textbook fill rule,
  with the reference's restriction of separator rows to leaf columns) for the
  symbolic pattern, dense no-pivot Doolittle for the leaf LU values, the
  +-eps pivot clamp special case;
* closed forms: the post-order block layout for np = 2, 4;
* library routines: MKL's getrfnp/trsm/gemm against the oracle's own loops.
A plausible mistake (dropped term, wrong sign/index, transposed operand) in
the oracle fails at least one of these.
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _csr(S, d):
    return S.CSR(int(len(d["rowptr"]) - 1), d["rowptr"], d["colidx"], d["data"])


def test_matrix20_reader_matches_scipy_golden(S):
    g = np.load(os.path.join(GOLD, "matrix_20_csr.npz"))
    A = S.read_mtx(os.path.join(GOLD, "matrix_20.mtx"))
    assert A.n == 20 and A.nnz == 65
    assert np.array_equal(A.rowptr, g["rowptr"])
    assert np.array_equal(A.colidx, g["colidx"])
    assert np.array_equal(A.data, g["data"])


def test_c1_generator_golden(S):
    g = np.load(os.path.join(GOLD, "c1.npz"))
    A = S.gen_grid(20, 20, 1, 5, 0, 4.1, 0)
    assert A.n == 400 and A.nnz == 1920
    for k in ("rowptr", "colidx", "data"):
        assert np.array_equal(getattr(A, k), g[k])
    xt, b = S.gen_rhs(A, 1, 1)
    assert np.array_equal(xt[:, 0], g["x_true"]) and np.array_equal(b[:, 0], g["b"])


@pytest.mark.parametrize("np_,kind", [(4, "grid"), (4, "nd"), (2, "nd"), (8, "grid"), (1, "grid")])
def test_oracle_c1_vs_superlu(S, oracle_mod, np_, kind):
    g = np.load(os.path.join(GOLD, "c1.npz"))
    A = _csr(S, g)
    order, sizes = S.order_grid(20, 20, 1, np_) if kind == "grid" else S.order_nd(A, np_)
    x = oracle_mod.Oracle(A, np_, order, sizes).factor_solve(g["b"])
    assert np.max(np.abs(x - g["x"])) / np.max(np.abs(g["x"])) < 1e-12
    assert S.backward_error(A, x, g["b"]) < 1e-14


@pytest.mark.parametrize("np_", [2, 4, 8])
def test_oracle_matrix20_vs_superlu(S, oracle_mod, np_):
    A = S.read_mtx(os.path.join(GOLD, "matrix_20.mtx"))
    g = np.load(os.path.join(GOLD, "matrix_20_x.npz"))
    order, sizes = S.order_nd(A, np_)
    x = oracle_mod.Oracle(A, np_, order, sizes).factor_solve(g["b"])
    assert np.max(np.abs(x - g["x"])) / np.max(np.abs(g["x"])) < 1e-12


def test_block_layout_closed_form(S, oracle_mod):
    # np = 4: post-order with child 2e+1 first -> 7, 6, 3, 5, 4, 2, 1;
    # block e has size sizes[2np-1-e]
    A = S.gen_grid(7, 1, 1, 5, 0, 4.1, 0)  # n = 7: one vertex per block
    order = np.arange(7, dtype=np.int32)
    sizes = np.ones(7, np.int32)
    n = 7
    A = S.CSR(n, np.arange(n + 1), np.arange(n), np.full(n, 2.0))  # diagonal
    O = oracle_mod.Oracle(A, 4, order, sizes)
    bs, sz = O.blocks()
    post = [7, 6, 3, 5, 4, 2, 1]
    for pos, e in enumerate(post):
        assert bs[e - 1] == pos and sz[e - 1] == 1
    sizes = np.array([3, 1, 2, 0, 5, 4, 6], np.int32)  # leaf0..3, sep(3), sep(2), top
    n = int(sizes.sum())
    A = S.CSR(n, np.arange(n + 1), np.arange(n), np.full(n, 2.0))
    O = oracle_mod.Oracle(A, 4, np.arange(n, dtype=np.int32), sizes)
    bs, sz = O.blocks()
    expect_size = {e: sizes[7 - e] for e in range(1, 8)}
    start = 0
    for e in post:
        assert bs[e - 1] == start and sz[e - 1] == expect_size[e]
        start += expect_size[e]
    # np = 2: 3, 2, 1
    sizes = np.array([2, 3, 1], np.int32)
    A = S.CSR(6, np.arange(7), np.arange(6), np.full(6, 2.0))
    bs, sz = oracle_mod.Oracle(A, 2, np.arange(6, dtype=np.int32), sizes).blocks()
    assert list(bs) == [5, 2, 0] and list(sz) == [1, 3, 2]


def _dense(A):
    D = np.zeros((A.n, A.n))
    for i in range(A.n):
        for q in range(A.rowptr[i], A.rowptr[i + 1]):
            D[i, A.colidx[q]] = A.data[q]
    return D


def _core_dense(S, O, A, order, rank, np_):
    """Dense core matrix of a leaf (leaf rows, then ancestor separators
    bottom-up; separator-separator block left at zero) and leaf size."""
    bs, sz = O.blocks()
    leaf = 2 * np_ - 1 - rank
    path = []
    e = leaf
    while e >= 1:
        path.append(e)
        e //= 2
    idx = np.concatenate([np.arange(bs[e - 1], bs[e - 1] + sz[e - 1]) for e in path]).astype(int)
    D = _dense(A)
    P = np.zeros_like(D)
    P[np.ix_(order, order)] = D  # PAP^T: PA[order[i], order[j]] = A[i, j]
    C = P[np.ix_(idx, idx)].copy()
    ls = int(sz[leaf - 1])
    C[ls:, ls:] = 0.0
    return C, ls


def _cases(S):
    A = S.gen_grid(9, 8, 1, 5, 0, 4.1, 3)
    yield A, 4, S.order_grid(9, 8, 1, 4)
    A = S.gen_grid(5, 4, 4, 7, 0, 6.1, 2)
    yield A, 4, S.order_nd(A, 4)
    A = S.gen_grid(5, 5, 3, 27, 1, 0.0, 1)
    yield A, 2, S.order_grid(5, 5, 3, 2)
    A = S.read_mtx(os.path.join(GOLD, "matrix_20.mtx"))
    yield A, 4, S.order_nd(A, 4)


def test_symbolic_vs_boolean_elimination(S, oracle_mod):
    for A, np_, (order, sizes) in _cases(S):
        O = oracle_mod.Oracle(A, np_, order, sizes)
        for rank in range(np_):
            C, ls = _core_dense(S, O, A, order, rank, np_)
            n = C.shape[0]
            pat = C != 0
            for k in range(ls):  # textbook fill; separator rows keep leaf columns only
                rows = [i for i in range(k + 1, n) if pat[i, k]]
                cols = [c for c in range(k + 1, n) if pat[k, c]]
                for i in rows:
                    for c in cols:
                        if i < ls or c < ls:
                            pat[i, c] = True
            rp, ci, dg, bias, _ = O.leaf_struct(rank)
            for i in range(n):
                got = set(ci[rp[i]:rp[i + 1]].tolist())
                exp = set(np.nonzero(pat[i])[0].tolist())
                assert got == exp, (rank, i)
                assert list(ci[rp[i]:rp[i + 1]]) == sorted(got)
                if i < ls:
                    assert ci[dg[i]] == i
                    assert bias[i] == rp[i] + np.count_nonzero(ci[rp[i]:rp[i + 1]] < ls)


def test_leaf_lu_vs_dense_doolittle(S, oracle_mod):
    for A, np_, (order, sizes) in _cases(S):
        O = oracle_mod.Oracle(A, np_, order, sizes)
        O.leaf_factor()
        for rank in range(np_):
            C, ls = _core_dense(S, O, A, order, rank, np_)
            n = C.shape[0]
            W = C.copy()
            for k in range(ls):  # right-looking, no pivoting, +-eps clamp
                d = W[k, k]
                if -1e-15 < d < 1e-15:
                    W[k, k] = d = -1e-15 if d < 0 else 1e-15
                W[k + 1:, k] /= d
                W[k + 1:, k + 1:] -= np.outer(W[k + 1:, k], W[k, k + 1:])
            rp, ci, dg, bias, _ = O.leaf_struct(rank)
            v = O.leaf_values(rank)
            for i in range(n):
                for q in range(rp[i], rp[i + 1]):
                    ref = W[i, ci[q]]
                    assert abs(v[q] - ref) <= 1e-13 * max(1.0, abs(ref)), (rank, i, ci[q])


def test_leaf_pivot_clamp_zero(S, oracle_mod):
    # a leaf whose second pivot is exactly 0 after elimination:
    # [[1, 1], [1, 1]] -> U(1,1) = 0 -> clamped to +1e-15 (kernel_sparse.cu:742-749)
    rp = np.array([0, 2, 4, 5], np.int32)
    ci = np.array([0, 1, 0, 1, 2], np.int32)
    va = np.array([1.0, 1.0, 1.0, 1.0, 3.0])
    A = S.CSR(3, rp, ci, va)
    order = np.array([0, 1, 2], np.int32)
    sizes = np.array([2, 1, 0], np.int32)  # leaf0 = {0,1}, leaf1 = {2}, top empty
    O = oracle_mod.Oracle(A, 2, order, sizes)
    O.leaf_factor()
    v = O.leaf_values(0)
    r, c, dg, _, _ = O.leaf_struct(0)
    assert v[dg[1]] == 1e-15
    # tiny negative pivot: [[1, 1], [1, 1 - 2^-52]] -> U11 = -2^-52 -> -1e-15
    va2 = np.array([1.0, 1.0, 1.0, 1.0 - 2.0 ** -52, 3.0])
    O = oracle_mod.Oracle(S.CSR(3, rp, ci, va2), 2, order, sizes)
    O.leaf_factor()
    assert O.leaf_values(0)[dg[1]] == -1e-15


def test_dense_fallback_matches_mkl(S, oracle_mod):
    if oracle_mod.dense_backend() != 1:
        pytest.skip("MKL runtime not present")
    A = S.gen_grid(12, 11, 10, 7, 0, 6.1, 4)
    order, sizes = S.order_grid(12, 11, 10, 8)
    _, b = S.gen_rhs(A, 1, 1)
    x_mkl = oracle_mod.Oracle(A, 8, order, sizes).factor_solve(b[:, 0])
    oracle_mod.force_fallback(True)
    try:
        x_fb = oracle_mod.Oracle(A, 8, order, sizes).factor_solve(b[:, 0])
    finally:
        oracle_mod.force_fallback(False)
    assert np.max(np.abs(x_mkl - x_fb)) / np.max(np.abs(x_mkl)) < 1e-12
    xs = spla.spsolve(A.to_scipy().tocsc(), b[:, 0])
    assert np.max(np.abs(x_mkl - xs)) / np.max(np.abs(xs)) < 1e-12


def test_oracle_3d_cases_vs_superlu(S, oracle_mod):
    for A, np_, (order, sizes) in _cases(S):
        _, b = S.gen_rhs(A, 1, 1)
        x = oracle_mod.Oracle(A, np_, order, sizes).factor_solve(b[:, 0])
        xs = spla.spsolve(A.to_scipy().tocsc(), b[:, 0])
        assert np.max(np.abs(x - xs)) / np.max(np.abs(xs)) < 1e-11
        assert S.backward_error(A, x, b[:, 0]) < 1e-14
