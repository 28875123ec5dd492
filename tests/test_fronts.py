"""Multifrontal structure of the leaves (host side, no GPU).

The leaf factorization runs as dense partial factorizations of frontal
matrices.  These tests execute that structure with numpy, front by front
(assembly of the owned LU slots from A, extend-add of the children's
updates, partial no-pivot LU with the reference's pivot clamp, write-back of
the owned slots) and check the result against the oracle's leaf LU values
(the reference's row-by-row elimination, src/kernel_sparse.cu:740-781), and
the root fronts' updates against the boundary Schur block -L_SL U_LS
computed from the oracle's factors (src/kernel_sparse.cu:953-971).  They also
check the structural invariants: every leaf pivot in exactly one front, every
LU slot owned by exactly one front, children before parents, extend-add
indices inside the parent front.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-12


def _cases(S):
    out = []
    A = S.gen_grid(20, 20, 1, 5, 0, 4.1, 0)
    out.append(("c1_grid", A, 4, S.order_grid(20, 20, 1, 4)))
    A = S.gen_grid(12, 10, 9, 7, 0, 6.1, 0)
    out.append(("3d7_np8", A, 8, S.order_grid(12, 10, 9, 8)))
    A = S.gen_grid(10, 9, 8, 27, 1, 0.0, 0)
    out.append(("3d27_np8", A, 8, S.order_grid(10, 9, 8, 8)))
    M20 = S.read_mtx(os.path.join(GOLD, "matrix_20.mtx"))
    out.append(("matrix_20", M20, 4, S.order_nd(M20, 4)))
    A = S.gen_grid(14, 14, 14, 7, 0, 6.1, 5)
    out.append(("3d7_nd_np16", A, 16, S.order_nd(A, 16)))
    A = S.gen_grid(40, 40, 1, 5, 0, 4.1, 7)
    out.append(("2d5_np1", A, 1, S.order_grid(40, 40, 1, 1)))
    return out


def _leaf_initial_values(S, M, A, r, order):
    """A's values at the leaf's LU slots (zero at fill positions)."""
    np_ = M.np
    bs, bsz = M.blocks()
    leaf = 2 * np_ - 1 - r
    path = []
    e = leaf
    while e >= 1:
        path.append(e)
        e //= 2
    gmap = np.concatenate([np.arange(bs[p - 1], bs[p - 1] + bsz[p - 1]) for p in path])
    coo = sp.csr_matrix((A.data, A.colidx, A.rowptr), shape=(A.n, A.n)).tocoo()
    inv = np.asarray(order)
    PA = sp.csr_matrix((coo.data, (inv[coo.row], inv[coo.col])), shape=(A.n, A.n))
    rp, ci, _, _, _ = M.leaf_struct(r)
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return np.asarray(PA[gmap[rows], gmap[ci]]).ravel().astype(np.float64), rows, ci


def _execute(Fs, vals, eps=1e-15):
    """Numpy multifrontal executor over the exported structure: returns the
    factored slot values and the root fronts' (index list, update) pairs."""
    nf = len(Fs["k0"])
    upd, roots = {}, []
    for f in range(nf):
        nk, n = int(Fs["nk"][f]), int(Fs["f"][f])
        F = np.zeros(n * n)
        sl = slice(Fs["slot_ptr"][f], Fs["slot_ptr"][f + 1])
        F[Fs["pos"][sl]] = vals[Fs["slot"][sl]]
        F = F.reshape(n, n, order="F")
        for c in Fs["child"][Fs["child_ptr"][f]:Fs["child_ptr"][f + 1]]:
            rel = Fs["rel"][Fs["idx_ptr"][c]:Fs["idx_ptr"][c + 1]]
            F[np.ix_(rel, rel)] += upd.pop(int(c))
        for p in range(nk):
            d = F[p, p]
            if -eps < d < eps:
                d = -eps if d < 0 else eps
            F[p, p] = d
            F[p + 1:, p] /= d
            F[p + 1:, p + 1:] -= np.outer(F[p + 1:, p], F[p, p + 1:])
        vals[Fs["slot"][sl]] = F.reshape(-1, order="F")[Fs["pos"][sl]]
        if Fs["parent"][f] >= 0:
            upd[f] = F[nk:, nk:].copy()
        else:
            roots.append((f, F[nk:, nk:].copy()))
    assert not upd, "an update was never consumed"
    return vals, roots


def test_fronts_structure_and_values_vs_oracle(S, oracle_mod):
    for name, A, np_, (order, sizes) in _cases(S):
        M = S.SnuMat(A, np_, order=order, sizes=sizes, device=-1)
        O = oracle_mod.Oracle(A, np_, order, sizes)
        O.leaf_factor()
        for r in range(np_):
            ls, cn, _, lnz = M.leaf_sizes(r)
            Fs = M.leaf_fronts(r)
            nf = len(Fs["k0"])
            # pivots partitioned among the fronts, ascending within each,
            # fronts ordered by their largest pivot
            kp, K = Fs["k_ptr"], Fs["k"]
            assert np.array_equal(np.sort(K), np.arange(ls)) and np.all(np.diff(kp) == Fs["nk"])
            assert all(np.all(np.diff(K[kp[f]:kp[f + 1]]) > 0) for f in range(nf))
            assert np.all(np.diff(K[kp[1:] - 1]) > 0) and np.array_equal(Fs["k0"], K[kp[:-1]])
            # every LU slot owned once; children precede their parents
            assert np.array_equal(np.sort(Fs["slot"]), np.arange(lnz))
            par = Fs["parent"]
            assert np.all((par < 0) | (par > np.arange(nf)))
            for f in range(nf):
                if par[f] >= 0:
                    rel = Fs["rel"][Fs["idx_ptr"][f]:Fs["idx_ptr"][f + 1]]
                    assert np.all((rel >= 0) & (rel < Fs["f"][par[f]])) and np.all(np.diff(rel) > 0)
            vals, rows, cols = _leaf_initial_values(S, M, A, r, order)
            got, roots = _execute(Fs, vals)
            ref = O.leaf_values(r)
            den = np.maximum(np.abs(ref), np.max(np.abs(ref)) * 1e-16)
            rel_err = float(np.max(np.abs(got - ref) / den))
            assert rel_err <= TOL, (name, r, rel_err)
            # boundary Schur block from the roots vs -L_SL U_LS of the oracle's factors
            rp, ci, dg, _, _ = M.leaf_struct(r)
            lo = (rows >= ls) & (cols < ls)
            up = (rows < ls) & (cols >= ls)
            Lsl = sp.csr_matrix((ref[lo], (rows[lo] - ls, cols[lo])), shape=(cn - ls, ls))
            Uls = sp.csr_matrix((ref[up], (rows[up], cols[up] - ls)), shape=(ls, cn - ls))
            Sref = -(Lsl @ Uls).toarray()
            Sgot = np.zeros_like(Sref)
            for f, U in roots:
                idx = Fs["idx"][Fs["idx_ptr"][f]:Fs["idx_ptr"][f + 1]] - ls
                Sgot[np.ix_(idx, idx)] += U
            if Sref.size:
                scale = max(np.max(np.abs(Sref)), 1e-300)
                assert np.max(np.abs(Sgot - Sref)) <= 1e-13 * scale, (name, r)
        O.close()
        M.close()
