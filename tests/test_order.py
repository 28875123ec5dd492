"""Nested-dissection orderings (host, no GPU): the ParMETIS order/sizes
contract for every method, determinism for a seed, and quality (fill and
flops from the symbolic structure) against the simpler level-set ND on an
irregular graph and against the geometric ND on a 7-point grid."""
import numpy as np
import pytest


def knn_graph(S, n, k, seed):
    """Irregular planar-like graph: k nearest neighbours of random points in
    the unit square, symmetrised, as a diagonally dominant matrix."""
    import scipy.sparse as sp
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    P = rng.random((n, 2))
    _, nb = cKDTree(P).query(P, k + 1)
    rows = np.repeat(np.arange(n), k)
    M = sp.coo_matrix((np.ones(n * k), (rows, nb[:, 1:].ravel())), shape=(n, n)).tocsr()
    M = ((M + M.T) > 0).astype(float).tocsr()
    M.setdiag(0)
    M.eliminate_zeros()
    M = (sp.diags(np.asarray(M.sum(1)).ravel() + 1.0) - M).tocsr()
    M.sort_indices()
    return S.CSR(n, M.indptr, M.indices, M.data)


def work(S, A, np_, order, sizes):
    with S.SnuMat(A, np_, order=order, sizes=sizes, device=-1) as M:  # separator property checked here
        c, i = M.counts(), M.info()
    return i["leaf_lu_nnz"] + i["dense_entries"], c["f_dense"] + c["f_leaf"]


@pytest.mark.parametrize("method", ["default", "levelset", "multilevel"])
def test_contract_and_determinism(S, method):
    A = knn_graph(S, 3000, 6, 1)
    o1, z1 = S.order_nd(A, 8, 5, method)
    o2, z2 = S.order_nd(A, 8, 5, method)
    assert np.array_equal(o1, o2) and np.array_equal(z1, z2)
    assert np.array_equal(np.sort(o1), np.arange(A.n)) and z1.sum() == A.n and len(z1) == 15
    work(S, A, 8, o1, z1)


def test_multilevel_beats_levelset_on_irregular_graph(S):
    A = knn_graph(S, 20000, 6, 0)
    fl = {m: work(S, A, 16, *S.order_nd(A, 16, 0, m)) for m in ("levelset", "multilevel", "default")}
    print({m: f"fill {f:.3g} flops {w:.3g}" for m, (f, w) in fl.items()})
    assert fl["multilevel"][1] < 0.5 * fl["levelset"][1]
    assert fl["default"][1] <= 1.1 * fl["multilevel"][1]


def test_default_nd_on_7pt_grid(S):
    """On a 7-point grid the level sets of a breadth-first search run
    diagonally and are smaller than axis planes: the default ND, which keeps
    the smaller separator, needs fewer flops than the geometric ND."""
    A = S.gen_grid(24, 24, 24, 7, 0, 6.1, 0)
    g = work(S, A, 16, *S.order_grid(24, 24, 24, 16))
    d = work(S, A, 16, *S.order_nd(A, 16, 0))
    print(f"geometric: fill {g[0]:.3g} flops {g[1]:.3g}; default ND: fill {d[0]:.3g} flops {d[1]:.3g}")
    assert d[1] < g[1]
