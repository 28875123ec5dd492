"""GPU parity: the HIP path against the CPU oracle, through the C-ABI.

Bars: integer structures bit-exact (checked on the CPU side too); leaf LU
values bit-exact (the kernels apply the reference's per-entry update sequence
without FMA contraction); solution x within 1e-9 relative (max-norm) of the
oracle, relative backward error <= 1e-12.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REL_TOL = 1e-9
BERR_TOL = 1e-12


def _cases(S):
    out = []
    A = S.gen_grid(20, 20, 1, 5, 0, 4.1, 0)
    out.append(("c1_grid", A, 4, S.order_grid(20, 20, 1, 4)))
    out.append(("c1_nd", A, 4, S.order_nd(A, 4)))
    M20 = S.read_mtx(os.path.join(GOLD, "matrix_20.mtx"))
    out.append(("matrix_20", M20, 4, S.order_nd(M20, 4)))
    out.append(("matrix_20_np2", M20, 2, S.order_nd(M20, 2)))
    A = S.gen_grid(33, 17, 1, 9, 0, 8.1, 3)
    out.append(("2d9_np8", A, 8, S.order_grid(33, 17, 1, 8)))
    A = S.gen_grid(12, 10, 9, 7, 0, 6.1, 0)
    out.append(("3d7_np8", A, 8, S.order_grid(12, 10, 9, 8)))
    A = S.gen_grid(10, 9, 8, 27, 1, 0.0, 0)
    out.append(("3d27_np8", A, 8, S.order_grid(10, 9, 8, 8)))
    A = S.gen_grid(14, 14, 14, 7, 0, 6.1, 5)
    out.append(("3d7_nd_np16", A, 16, S.order_nd(A, 16)))
    A = S.gen_grid(24, 24, 24, 7, 0, 6.1, 0)
    out.append(("3d7_24_np32", A, 32, S.order_grid(24, 24, 24, 32)))
    A = S.gen_grid(40, 40, 1, 5, 0, 4.1, 7)
    out.append(("2d5_np1", A, 1, S.order_grid(40, 40, 1, 1)))
    return out


@pytest.fixture(scope="module")
def cases(S):
    return _cases(S)


def test_leaf_lu_bitexact(S, oracle_mod, cases):
    for name, A, np_, (order, sizes) in cases:
        print("case", name, flush=True)
        O = oracle_mod.Oracle(A, np_, order, sizes)
        O.leaf_factor()
        with S.SnuMat(A, np_, order=order, sizes=sizes, device=0) as M:
            M.factorize()
            for r in range(np_):
                g, o = M.leaf_values(r), O.leaf_values(r)
                assert np.array_equal(g, o), (name, r, np.max(np.abs(g - o)))


def test_solution_vs_oracle(S, oracle_mod, cases):
    for name, A, np_, (order, sizes) in cases:
        print("case", name, flush=True)
        _, b = S.gen_rhs(A, 1, 1)
        b = b[:, 0]
        xo = oracle_mod.Oracle(A, np_, order, sizes).factor_solve(b)
        with S.SnuMat(A, np_, order=order, sizes=sizes, device=0) as M:
            x = M.factorize().solve(b)
        rel = np.max(np.abs(x - xo)) / np.max(np.abs(xo))
        berr = S.backward_error(A, x, b)
        print(f"  rel diff vs oracle {rel:.2e}  backward error {berr:.2e}")
        assert rel <= REL_TOL, (name, rel)
        assert berr <= BERR_TOL, (name, berr)


def test_multi_rhs_columns(S):
    A = S.gen_grid(16, 16, 12, 7, 0, 6.1, 0)
    order, sizes = S.order_grid(16, 16, 12, 16)
    _, B = S.gen_rhs(A, 7, 1)
    with S.SnuMat(A, 16, order=order, sizes=sizes, device=0) as M:
        M.factorize()
        X = M.solve(B)
        for j in range(B.shape[1]):
            xj = M.solve(B[:, j])
            assert np.max(np.abs(X[:, j] - xj)) <= 1e-12 * np.max(np.abs(xj))
        assert S.backward_error(A, X, B) <= BERR_TOL


def test_deterministic(S):
    A = S.gen_grid(20, 20, 20, 7, 0, 6.1, 0)
    order, sizes = S.order_grid(20, 20, 20, 16)
    _, b = S.gen_rhs(A, 1, 1)
    with S.SnuMat(A, 16, order=order, sizes=sizes, device=0) as M:
        x1 = M.factorize().solve(b[:, 0])
        x2 = M.factorize().solve(b[:, 0])
    assert np.array_equal(x1, x2)


def test_gemm_kernel_vs_numpy(S):
    import ctypes as C
    L = S.lib()
    rng = np.random.default_rng(0)
    for (M, N, K) in [(1, 1, 1), (64, 64, 16), (65, 63, 17), (130, 7, 200), (5, 200, 64), (257, 129, 3)]:
        A = np.asfortranarray(rng.standard_normal((M, K)))
        B = np.asfortranarray(rng.standard_normal((K, N)))
        C0 = np.asfortranarray(rng.standard_normal((M, N)))
        for assign in (0, 1):
            Cg = C0.copy(order="F")
            rc = L.snu_test_gemm(M, N, K, A.ctypes.data_as(C.c_void_p), M, B.ctypes.data_as(C.c_void_p), K,
                                 Cg.ctypes.data_as(C.c_void_p), M, assign)
            assert rc == 0
            ref = A @ B if assign else C0 - A @ B
            err = np.max(np.abs(Cg - ref)) / max(1.0, np.max(np.abs(ref)))
            assert err < 1e-13, (M, N, K, assign, err)


def test_diag_lu_kernel_vs_doolittle(S):
    import ctypes as C
    L = S.lib()
    rng = np.random.default_rng(1)
    for nb in (1, 2, 17, 64):
        A = rng.standard_normal((nb, nb)) + nb * np.eye(nb)
        D = np.asfortranarray(A.copy())
        Li = np.zeros(4096)
        Ui = np.zeros(4096)
        rc = L.snu_test_diag_lu(nb, D.ctypes.data_as(C.c_void_p), nb, Li.ctypes.data_as(C.c_void_p),
                                Ui.ctypes.data_as(C.c_void_p))
        assert rc == 0
        Lf = np.tril(D, -1) + np.eye(nb)
        Uf = np.triu(D)
        assert np.max(np.abs(Lf @ Uf - A)) <= 1e-13 * np.max(np.abs(A))
        Li = Li.reshape(64, 64, order="F")[:nb, :nb]
        Ui = Ui.reshape(64, 64, order="F")[:nb, :nb]
        assert np.max(np.abs(Li @ Lf - np.eye(nb))) < 1e-12
        assert np.max(np.abs(Ui @ Uf - np.eye(nb))) < 1e-12
