"""Integer structures of the product's construct phase, bit-exact against the
oracle (block layout, PAP^T, perm_map, per-leaf symbolic LU, LU_bias,
core_map), plus invariants and the error paths of construct.  Host only: the
handles are built with device = -1, no GPU needed."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cases(S):
    A = S.gen_grid(20, 20, 1, 5, 0, 4.1, 0)
    yield "c1_grid", A, 4, S.order_grid(20, 20, 1, 4)
    yield "c1_nd", A, 4, S.order_nd(A, 4)
    M20 = S.read_mtx(os.path.join(GOLD, "matrix_20.mtx"))
    for np_ in (1, 2, 4, 8):
        yield f"m20_np{np_}", M20, np_, S.order_nd(M20, np_)
    A = S.gen_grid(17, 13, 1, 9, 0, 8.1, 2)
    yield "2d9", A, 8, S.order_grid(17, 13, 1, 8)
    A = S.gen_grid(11, 9, 8, 7, 0, 6.1, 0)
    yield "3d7_nd", A, 16, S.order_nd(A, 16, seed=3)
    A = S.gen_grid(9, 8, 7, 27, 1, 0.0, 0)
    yield "3d27", A, 32, S.order_grid(9, 8, 7, 32)


def test_structures_bitexact_vs_oracle(S, oracle_mod):
    for name, A, np_, (order, sizes) in _cases(S):
        O = oracle_mod.Oracle(A, np_, order, sizes)
        M = S.SnuMat(A, np_, order=order, sizes=sizes, device=-1)
        for a, b in zip(M.blocks(), O.blocks()):
            assert np.array_equal(a, b), name
        for a, b in zip(M.permuted(), O.permuted()):
            assert np.array_equal(a, b), name
        for r in range(np_):
            assert M.leaf_sizes(r) == O.leaf_sizes(r), (name, r)
            for a, b in zip(M.leaf_struct(r), O.leaf_struct(r)):
                assert np.array_equal(a, b), (name, r)
        M.close()


def test_permutation_invariants(S):
    A = S.gen_grid(12, 10, 1, 9, 0, 8.1, 1)
    order, sizes = S.order_grid(12, 10, 1, 8)
    M = S.SnuMat(A, 8, order=order, sizes=sizes, device=-1)
    rp, ci, pm = M.permuted()
    assert np.array_equal(np.sort(order), np.arange(A.n))
    assert sizes.sum() == A.n
    assert np.array_equal(np.sort(pm), np.arange(A.nnz))  # perm_map is a permutation
    for i in range(A.n):  # PA[order[i], order[j]] = A[i, j], rows sorted
        row = ci[rp[order[i]]:rp[order[i] + 1]]
        assert np.all(np.diff(row) > 0)
        assert set(row.tolist()) == set(order[A.colidx[A.rowptr[i]:A.rowptr[i + 1]]].tolist())
    # separator property: each entry couples blocks on one root path
    bs, sz = M.blocks()
    blk = np.zeros(A.n, int)
    for e in range(1, 16):
        blk[bs[e - 1]:bs[e - 1] + sz[e - 1]] = e
    for r in range(A.n):
        for c in ci[rp[r]:rp[r + 1]]:
            x, y = max(blk[r], blk[c]), min(blk[r], blk[c])
            while x > y:
                x //= 2
            assert x == y


@pytest.mark.parametrize("np_", [1, 2, 4, 8, 16])
def test_orderings_contract(S, np_):
    for nx, ny, nz, st in [(20, 20, 1, 5), (9, 7, 5, 7), (6, 6, 6, 27), (31, 3, 1, 9)]:
        A = S.gen_grid(nx, ny, nz, st, 0, 6.1, 0)
        for order, sizes in (S.order_grid(nx, ny, nz, np_), S.order_nd(A, np_)):
            assert len(sizes) == 2 * np_ - 1 and sizes.min() >= 0 and sizes.sum() == A.n
            assert np.array_equal(np.sort(order), np.arange(A.n))
            S.SnuMat(A, np_, order=order, sizes=sizes, device=-1).close()  # separator property holds


def test_construct_error_paths(S):
    A = S.gen_grid(6, 6, 1, 5, 0, 4.1, 0)
    # np not a power of two
    with pytest.raises(S.SnuError) as e:
        S.SnuMat(A, 3, device=-1)
    assert e.value.code == -1
    # duplicate entry
    rp = np.array([0, 2, 3], np.int32)
    B = S.CSR(2, rp, np.array([0, 0, 1], np.int32), np.array([1.0, 2.0, 3.0]))
    with pytest.raises(S.SnuError) as e:
        S.SnuMat(B, 1, device=-1)
    assert e.value.code == -4
    # unsorted rows are accepted (the reference's reader keeps file order)
    B = S.CSR(2, rp, np.array([1, 0, 1], np.int32), np.array([1.0, 2.0, 3.0]))
    S.SnuMat(B, 1, device=-1).close()
    # structurally zero diagonal in a leaf
    B = S.CSR(2, np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32), np.array([1.0, 1.0]))
    with pytest.raises(S.SnuError) as e:
        S.SnuMat(B, 1, device=-1)
    assert e.value.code == -3
    # ordering that is not a permutation
    order = np.zeros(A.n, np.int32)
    with pytest.raises(S.SnuError) as e:
        S.SnuMat(A, 2, order=order, sizes=np.array([18, 12, 6], np.int32), device=-1)
    assert e.value.code == -1
    # separator property violated (natural order, arbitrary sizes)
    with pytest.raises(S.SnuError) as e:
        S.SnuMat(A, 2, order=np.arange(A.n, dtype=np.int32), sizes=np.array([18, 12, 6], np.int32),
                 device=-1)
    assert e.value.code == -2
    # grid dims that do not match n
    with pytest.raises(S.SnuError) as e:
        S.SnuMat(A, 2, grid=(5, 6, 1), device=-1)
    assert e.value.code == -1


def test_host_only_handle_refuses_numerics(S):
    A = S.gen_grid(6, 6, 1, 5, 0, 4.1, 0)
    M = S.SnuMat(A, 2, device=-1)
    with pytest.raises(S.SnuError) as e:
        M.factorize()
    assert e.value.code == -5
    M.close()


def test_dense_layout_accounting(S):
    A = S.gen_grid(16, 16, 16, 7, 0, 6.1, 0)
    M = S.SnuMat(A, 16, grid=(16, 16, 16), device=-1)
    bs, sz = M.blocks()
    total = diag = 0
    for p in range(1, 16):  # uncompressed: sum over separators of s^2 + 2 s A
        a = 0
        q = p // 2
        while q:
            a += sz[q - 1]
            q //= 2
        total += sz[p - 1] ** 2 + 2 * sz[p - 1] * a
        diag += sz[p - 1] ** 2
    info = M.info()
    # the panels keep only the 64-wide tiles of each ancestor holding a
    # structural nonzero: between the diagonal blocks alone and the full
    # panels, and strictly below the latter on a 3-D grid (far separators
    # touch few of the ancestors' tiles)
    assert diag < info["dense_entries"] < total
    assert info["leaf_lu_nnz"] == sum(M.leaf_sizes(r)[3] for r in range(16))
    M.close()
