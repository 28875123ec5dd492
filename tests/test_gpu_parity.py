"""GPU parity: the HIP path against the CPU oracle, through the C-ABI.

Bars: integer structures bit-exact (checked on the CPU side too); leaf LU
values within 1e-12 of the oracle's, relative to the leaf's largest factor
entry (the leaves are factored as dense fronts on the GPU: same factors,
other rounding); solution x within 1e-9 relative (max-norm) of the oracle,
relative backward error <= 1e-12.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REL_TOL = 1e-9
BERR_TOL = 1e-12
LEAF_TOL = 1e-12


def leaf_rel(g, o):
    """max |g - o| over the leaf's LU values, relative to max |o|."""
    return float(np.max(np.abs(g - o)) / max(np.max(np.abs(o)), 1e-300))


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


def test_leaf_lu_vs_oracle(S, oracle_mod, cases):
    for name, A, np_, (order, sizes) in cases:
        print("case", name, flush=True)
        O = oracle_mod.Oracle(A, np_, order, sizes)
        O.leaf_factor()
        with S.SnuMat(A, np_, order=order, sizes=sizes, device=0) as M:
            M.factorize()
            for r in range(np_):
                g, o = M.leaf_values(r), O.leaf_values(r)
                assert leaf_rel(g, o) <= LEAF_TOL, (name, r, leaf_rel(g, o))


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
    for (M, N, K) in [(1, 1, 1), (64, 64, 16), (65, 63, 17), (130, 7, 200), (5, 200, 64), (257, 129, 3),
                      (128, 128, 16), (300, 260, 100), (129, 383, 64), (1000, 999, 33)]:
        A = np.asfortranarray(rng.standard_normal((M, K)))
        B = np.asfortranarray(rng.standard_normal((K, N)))
        C0 = np.asfortranarray(rng.standard_normal((M, N)))
        for assign in (0, 1, 2, 3, 4, 5):
            Cg = C0.copy(order="F")
            rc = L.snu_test_gemm(M, N, K, A.ctypes.data_as(C.c_void_p), M, B.ctypes.data_as(C.c_void_p), K,
                                 Cg.ctypes.data_as(C.c_void_p), M, assign)
            assert rc == 0
            ref = A @ B if assign in (1, 4) else (-(A @ B) if assign in (3, 5) else C0 - A @ B)
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


def test_gemm_kernel_f32_vs_numpy(S):
    """FP32 MFMA GEMM (16x16x4 f32, FP32 accumulation) against an FP64
    product of the same FP32 operands, elementwise within the standard
    dot-product rounding bound (K + 2) u32 (|A||B| + |C|)."""
    import ctypes as C
    L = S.lib()
    rng = np.random.default_rng(2)
    u32 = 2.0 ** -24
    for (M, N, K) in [(1, 1, 1), (64, 64, 16), (65, 63, 17), (130, 7, 200), (5, 200, 64),
                      (128, 128, 16), (300, 260, 100), (129, 383, 64), (1000, 999, 33)]:
        A = np.asfortranarray(rng.standard_normal((M, K)).astype(np.float32))
        B = np.asfortranarray(rng.standard_normal((K, N)).astype(np.float32))
        C0 = np.asfortranarray(rng.standard_normal((M, N)).astype(np.float32))
        A64, B64, C64 = A.astype(np.float64), B.astype(np.float64), C0.astype(np.float64)
        absprod = np.abs(A64) @ np.abs(B64)
        for assign in (0, 1, 2, 3, 4, 5):
            Cg = C0.copy(order="F")
            rc = L.snu_test_gemm_f32(M, N, K, A.ctypes.data_as(C.c_void_p), M, B.ctypes.data_as(C.c_void_p), K,
                                     Cg.ctypes.data_as(C.c_void_p), M, assign)
            assert rc == 0
            P = A64 @ B64
            ref = P if assign in (1, 4) else (-P if assign in (3, 5) else C64 - P)
            bound = (K + 2) * u32 * (absprod + (np.abs(C64) if assign in (0, 2) else 0)) + 1e-30
            assert np.all(np.abs(Cg - ref) <= bound), (M, N, K, assign, np.max(np.abs(Cg - ref) / bound))


def test_gemm_kernel_split_vs_numpy(S):
    """FP32 GEMM through three-piece bf16 splits on the bf16 MFMA (the
    SNU_GEMM_SPLIT form of the FP32 factorization's GEMM tiles, 128 x 128
    and 64 x 64)
    against an FP64 product of the same FP32 operands, elementwise within
    (K + 4) u32 (|A||B| + |C|): the FP32 dot-product bound plus two units for
    the piece products left out.  A bf16-accurate product (one piece) misses
    this bound by about 2^15, and a mismatched k order of the two operands by
    far more.  Operands span 2^+-20 so the splits see every exponent range."""
    import ctypes as C
    L = S.lib()
    rng = np.random.default_rng(4)
    u32 = 2.0 ** -24
    for (M, N, K) in [(1, 1, 1), (128, 128, 32), (65, 63, 17), (130, 7, 200), (5, 200, 64),
                      (257, 129, 3), (300, 260, 100), (129, 383, 64), (1000, 999, 33), (256, 256, 1024)]:
        A = rng.standard_normal((M, K)) * 2.0 ** rng.integers(-20, 21, (M, K))
        B = rng.standard_normal((K, N)) * 2.0 ** rng.integers(-20, 21, (K, N))
        A = np.asfortranarray(A.astype(np.float32))
        B = np.asfortranarray(B.astype(np.float32))
        C0 = np.asfortranarray(rng.standard_normal((M, N)).astype(np.float32))
        A64, B64, C64 = A.astype(np.float64), B.astype(np.float64), C0.astype(np.float64)
        absprod = np.abs(A64) @ np.abs(B64)
        P = A64 @ B64
        for assign in (6, 7, 8, 9, 10, 11):  # 128 x 128, then 64 x 64 tiles
            Cg = C0.copy(order="F")
            rc = L.snu_test_gemm_f32(M, N, K, A.ctypes.data_as(C.c_void_p), M, B.ctypes.data_as(C.c_void_p), K,
                                     Cg.ctypes.data_as(C.c_void_p), M, assign)
            assert rc == 0
            ref = P if assign in (7, 10) else (-P if assign in (8, 11) else C64 - P)
            bound = (K + 4) * u32 * (absprod + (np.abs(C64) if assign in (6, 9) else 0)) + 1e-30
            assert np.all(np.abs(Cg - ref) <= bound), (M, N, K, assign, np.max(np.abs(Cg - ref) / bound))
    # the FP64 hook has no split form
    z = np.zeros(1)
    assert L.snu_test_gemm(1, 1, 1, z.ctypes.data_as(C.c_void_p), 1, z.ctypes.data_as(C.c_void_p), 1,
                           z.ctypes.data_as(C.c_void_p), 1, 6) != 0


def test_diag_lu_kernel_f32(S):
    import ctypes as C
    L = S.lib()
    rng = np.random.default_rng(3)
    for nb in (1, 2, 17, 64):
        A = (rng.standard_normal((nb, nb)) + nb * np.eye(nb)).astype(np.float32)
        D = np.asfortranarray(A.copy())
        Li = np.zeros(4096, np.float32)
        Ui = np.zeros(4096, np.float32)
        rc = L.snu_test_diag_lu_f32(nb, D.ctypes.data_as(C.c_void_p), nb, Li.ctypes.data_as(C.c_void_p),
                                    Ui.ctypes.data_as(C.c_void_p))
        assert rc == 0
        Lf = np.tril(D, -1).astype(np.float64) + np.eye(nb)
        Uf = np.triu(D).astype(np.float64)
        assert np.max(np.abs(Lf @ Uf - A)) <= 1e-5 * np.max(np.abs(A))
        Li = Li.reshape(64, 64, order="F")[:nb, :nb].astype(np.float64)
        Ui = Ui.reshape(64, 64, order="F")[:nb, :nb].astype(np.float64)
        assert np.max(np.abs(Li @ Lf - np.eye(nb))) < 1e-4
        assert np.max(np.abs(Ui @ Uf - np.eye(nb))) < 1e-4


def test_fp32_refined_vs_oracle(S, oracle_mod, cases):
    """FP32 factors + FP64 refinement: same bars as the FP64 path (x within
    1e-9 of the FP64 oracle, backward error <= 1e-12), in <= 10 steps."""
    for name, A, np_, (order, sizes) in cases:
        print("case", name, flush=True)
        _, b = S.gen_rhs(A, 1, 1)
        b = b[:, 0]
        xo = oracle_mod.Oracle(A, np_, order, sizes).factor_solve(b)
        with S.SnuMat(A, np_, order=order, sizes=sizes, device=0, precision="fp32") as M:
            x = M.factorize().solve(b)
            t = M.timings()
        rel = np.max(np.abs(x - xo)) / np.max(np.abs(xo))
        berr = S.backward_error(A, x, b)
        print(f"  steps {t['ir_steps']:.0f}  rel diff vs oracle {rel:.2e}  backward error {berr:.2e}")
        assert rel <= REL_TOL, (name, rel)
        assert berr <= BERR_TOL, (name, berr)
        assert t["ir_steps"] <= 10 and t["ir_berr"] <= BERR_TOL


def test_fp32_refined_multi_rhs(S):
    A = S.gen_grid(16, 16, 12, 7, 0, 6.1, 0)
    order, sizes = S.order_grid(16, 16, 12, 16)
    _, B = S.gen_rhs(A, 9, 1)
    with S.SnuMat(A, 16, order=order, sizes=sizes, device=0) as M64:
        X64 = M64.factorize().solve(B)
    with S.SnuMat(A, 16, order=order, sizes=sizes, device=0, precision="fp32") as M:
        M.factorize()
        X = M.solve(B)
        for j in (0, 4, 8):
            xj = M.solve(B[:, j])
            assert np.max(np.abs(X[:, j] - xj)) <= 1e-9 * np.max(np.abs(xj))
    assert S.backward_error(A, X, B) <= BERR_TOL
    assert np.max(np.abs(X - X64)) <= REL_TOL * np.max(np.abs(X64))


def test_fp32_stalled_refinement_falls_back_to_fp64(S):
    """Refinement that cannot reach 1e-12 (forced from the host: no
    correction steps allowed) is answered by FP64 factors of the same matrix:
    the result equals the FP64 handle's bitwise, t[15] reports the fallback,
    and a following factorize + solve of new values refactorizes the FP64
    twin as well (same values scaled: x scales inversely)."""
    A = S.gen_grid(20, 20, 10, 7, 0, 6.1, 3)
    order, sizes = S.order_grid(20, 20, 10, 8)
    _, b = S.gen_rhs(A, 1, 1)
    b = b[:, 0]
    with S.SnuMat(A, 8, order=order, sizes=sizes, device=0) as M64:
        x64 = M64.factorize().solve(b)
    with S.SnuMat(A, 8, order=order, sizes=sizes, device=0, precision="fp32", ir_max_steps=-1) as M:
        x = M.factorize().solve(b)
        t = M.timings()
        assert t["fp64_fallback"] == 1 and t["ir_steps"] == 0 and t["ir_berr"] > 1e-12
        assert np.array_equal(x, x64)
        assert S.backward_error(A, x, b) <= BERR_TOL
        M.update_values(2.0 * A.data)
        x2 = M.factorize().solve(b)
        assert M.timings()["fp64_fallback"] == 1
        assert np.max(np.abs(2.0 * x2 - x64)) <= 1e-12 * np.max(np.abs(x64))
    with S.SnuMat(A, 8, order=order, sizes=sizes, device=0, precision="fp32") as M:
        x = M.factorize().solve(b)
        assert M.timings()["fp64_fallback"] == 0


def _batch_members(S):
    """Uneven matrices: 2D 9-point grids of several sides and leaf counts
    (np = 1 included), two 3D 7-point grids."""
    out = []
    for i, (side, np_) in enumerate([(12, 1), (17, 2), (24, 4), (40, 4), (33, 8), (64, 16), (20, 1), (50, 8)]):
        A = S.gen_grid(side, side, 1, 9, 0, 8.1, i)
        out.append((A, np_, S.order_grid(side, side, 1, np_)))
    for i, (d, np_) in enumerate([(10, 4), (14, 8)]):
        A = S.gen_grid(d, d, d, 7, 0, 6.1, 20 + i)
        out.append((A, np_, S.order_grid(d, d, d, np_)))
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_equals_standalone(S, precision):
    """A batch factorizes and solves every member with the arithmetic of its
    standalone run: the solutions are bitwise equal, with uneven right-hand
    side counts (0 skips a member) in one packed solve."""
    mem = _batch_members(S)
    rng = np.random.default_rng(7)
    counts = [1, 3, 0, 2, 1, 5, 1, 0, 2, 1]
    bs = [None if k == 0 else rng.standard_normal((A.n, k)) for (A, _, _), k in zip(mem, counts)]
    ref = []
    for (A, np_, (order, sizes)), b in zip(mem, bs):
        with S.SnuMat(A, np_, order=order, sizes=sizes, device=0, precision=precision) as M:
            ref.append(None if b is None else M.factorize().solve(b))
    mats = [S.SnuMat(A, np_, order=order, sizes=sizes, device=0, precision=precision)
            for (A, np_, (order, sizes)) in mem]
    with S.Batch(mats) as B:
        B.factorize()
        xs = B.solve(bs)
        for (A, _, _), b, x, r in zip(mem, bs, xs, ref):
            if b is None:
                assert x is None
                continue
            assert np.array_equal(x, r)
            assert S.backward_error(A, x, b) <= BERR_TOL
        # members are factored after the batch and solve on their own too
        x5 = mats[5].solve(bs[5])
        assert np.array_equal(x5, ref[5])
        fl, sl = B.launches()
        print(f"batch of {len(mats)}: {fl} dense factor launches, {sl} dense solve launches")
    for M in mats:
        M.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_serve_alone_vs_mix(S, precision):
    """Requests against several resident matrices, with uneven column counts,
    submitted together and packed by two slots: every answer equals the
    matrix's standalone solve of the same columns within 1e-12 relative for
    FP64 (the leaf solves' partial sums are grouped by the matrix's column
    count in the pack) and 1e-10 for FP32 + refinement (whose step count also
    depends on the columns packed together); backward error <= 1e-12."""
    mem = _batch_members(S)[2:7]
    rng = np.random.default_rng(11)
    reqs = []
    for q in range(30):
        m = int(rng.integers(len(mem)))
        k = int(2 ** rng.integers(0, 4))
        reqs.append((m, rng.standard_normal((mem[m][0].n, k))))
    alone = []
    for m, b in reqs:
        A, np_, (order, sizes) = mem[m]
        with S.SnuMat(A, np_, order=order, sizes=sizes, device=0, precision=precision) as M:
            alone.append(M.factorize().solve(b))
    with S.Server(budget_bytes=64 << 30, slots=2, max_cols=64) as srv:
        ids = [srv.add(A, np_, order=order, sizes=sizes, precision=precision) for A, np_, (order, sizes) in mem]
        tickets = [srv.submit(ids[m], b) for m, b in reqs]
        for (m, b), t, xa in zip(reqs, tickets, alone):
            x, lat = srv.wait(t)
            A = mem[m][0]
            assert S.backward_error(A, x, b) <= BERR_TOL
            tol = 1e-12 if precision == "fp64" else 1e-10
            assert np.max(np.abs(x - xa)) <= tol * np.max(np.abs(xa))
        st = srv.stats()
        print(f"served {st['requests']:.0f} requests in {st['packed_solves']:.0f} packed solves, "
              f"p50 {st['p50_ms']:.2f} ms p99 {st['p99_ms']:.2f} ms")
        assert st["requests"] == len(reqs)
        assert st["packed_solves"] <= len(reqs)


def test_serve_admission_evicts_lru(S):
    """Admission under a budget parks the least recently used idle matrix;
    a request for a parked matrix admits it again (parking the then least
    recently used one) and is answered."""
    mats = []
    for i, side in enumerate((48, 50, 52)):
        A = S.gen_grid(side, side, 1, 5, 0, 4.1, i)
        mats.append((A, S.order_grid(side, side, 1, 4)))
    need = []
    for A, (order, sizes) in mats:
        with S.SnuMat(A, 4, order=order, sizes=sizes, device=-1, precision="fp32") as M:
            nb = C_size()
            S.lib().snumat_bytes_needed(M._h, C_ref(nb))
            need.append(nb.value)
    budget = need[0] + need[1] + need[2] // 2  # room for two of the three
    with S.Server(budget_bytes=budget, slots=1) as srv:
        a = srv.add(mats[0][0], 4, order=mats[0][1][0], sizes=mats[0][1][1])
        b = srv.add(mats[1][0], 4, order=mats[1][1][0], sizes=mats[1][1][1])
        x, _ = srv.wait(srv.submit(a, np.ones(mats[0][0].n)))  # a is now the most recent
        c = srv.add(mats[2][0], 4, order=mats[2][1][0], sizes=mats[2][1][1])
        st = srv.stats()
        assert st["evictions"] == 1 and st["resident"] == 2 and st["parked"] == 1
        y, _ = srv.wait(srv.submit(c, np.ones(mats[2][0].n)))
        assert S.backward_error(mats[2][0], y, np.ones(mats[2][0].n)) <= BERR_TOL
        # b was parked; a is now the least recently used and is parked for it
        z, _ = srv.wait(srv.submit(b, np.ones(mats[1][0].n)))
        assert S.backward_error(mats[1][0], z, np.ones(mats[1][0].n)) <= BERR_TOL
        st = srv.stats()
        assert st["evictions"] == 2 and st["readmissions"] == 1
        assert st["resident"] == 2 and st["parked"] == 1 and st["bytes_reserved"] <= budget
        with pytest.raises(S.SnuError):
            srv.submit(99, np.ones(mats[1][0].n))  # never admitted


def C_size():
    import ctypes
    return ctypes.c_size_t()


def C_ref(v):
    import ctypes
    return ctypes.byref(v)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_detach_attach_refactorizes_bitwise(S, precision):
    """A handle detached (device memory released, plan kept) and attached
    again factorizes and solves to the same bits; a host-only handle
    attached later gives the same bits too."""
    A = S.gen_grid(14, 12, 10, 7, 0, 6.1, 3)
    order, sizes = S.order_nd(A, 8, 0)
    _, b = S.gen_rhs(A, 2, 1)
    with S.SnuMat(A, 8, order=order, sizes=sizes, device=0, precision=precision) as M:
        x0 = M.factorize().solve(b)
        assert M.info()["on_device"]
        M.detach()
        assert not M.info()["on_device"]
        with pytest.raises(S.SnuError):
            M.factorize()
        x1 = M.attach(0).factorize().solve(b)
    with S.SnuMat(A, 8, order=order, sizes=sizes, device=-1, precision=precision) as H:
        x2 = H.attach(0).factorize().solve(b)
    assert np.array_equal(x0, x1) and np.array_equal(x0, x2)
    assert S.backward_error(A, x0, b) <= BERR_TOL


def test_graph_replays_equal_eager(S):
    """Factorizations replayed from a captured HIP graph give the eager
    results on every replay, with eager work and synchronisations between
    the replays (single matrix and batch)."""
    A = S.gen_grid(16, 16, 12, 7, 0, 6.1, 0)
    order, sizes = S.order_grid(16, 16, 12, 16)
    _, b = S.gen_rhs(A, 2, 1)
    with S.SnuMat(A, 16, order=order, sizes=sizes, device=0) as M:
        x0 = M.factorize().solve(b)
        M.use_graph(True)
        for rep in range(3):
            x = M.factorize().solve(b)
            assert np.array_equal(x, x0), rep
            with S.SnuMat(A, 16, order=order, sizes=sizes, device=0) as E:  # eager work in between
                assert np.array_equal(E.factorize().solve(b), x0)
    mem = _batch_members(S)[:6]
    bs = [S.gen_rhs(A, 1, 3)[1][:, 0] for A, _, _ in mem]
    for precision in ("fp64", "fp32"):
        mats = [S.SnuMat(A, np_, order=o, sizes=z, device=0, precision=precision) for A, np_, (o, z) in mem]
        with S.Batch(mats) as B:
            ref = B.factorize().solve(bs)
            B.use_graph(True)
            for rep in range(3):
                xs = B.factorize().solve(bs)
                for x, r in zip(xs, ref):
                    assert np.array_equal(x, r), (precision, rep)
        for M in mats:
            M.close()
