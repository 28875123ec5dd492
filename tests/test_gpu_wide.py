"""GPU parity of the wider paths against the oracle, through the C-ABI.

Every answer of the multi-column, batched and served solves is compared with
the oracle's solution of the same column (the reference path has one
right-hand side, so the oracle solves each column on its own), and the
full-size 64^3 case and a 512-leaf case exercise the production launch
shapes (128 x 128 Schur tiles, the narrow resident leaf workgroups).

Bars as in test_gpu_parity: x within 1e-9 relative (max norm) of the oracle,
relative backward error <= 1e-12; leaf LU values within 1e-12 (normwise per
leaf); stage-1 separator
blocks (A_SS plus the leaf Schur contributions) within 1e-13 of the oracle's
(the leaf fronts' updates are summed on MFMA) with the same zero pattern.
"""
import os
import subprocess
import tempfile
import threading
import time

import numpy as np
import pytest

from test_gpu_parity import LEAF_TOL, _batch_members, leaf_rel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
REL_TOL = 1e-9
BERR_TOL = 1e-12
NT = min(16, os.cpu_count() or 1)  # oracle threads: the host share of one GPU


def _rel(x, xo):
    return float(np.max(np.abs(x - xo)) / np.max(np.abs(xo)))


def _vs_oracle(S, oracle_mod, A, np_, order, sizes, X, B, name):
    O = oracle_mod.Oracle(A, np_, order, sizes)
    XO = O.factor_solve_columns(B, NT)
    O.close()
    rel = _rel(X, XO)
    berr = S.backward_error(A, X, B)
    print(f"  {name}: rel diff vs oracle {rel:.2e}  backward error {berr:.2e}", flush=True)
    assert rel <= REL_TOL, (name, rel)
    assert berr <= BERR_TOL, (name, berr)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_multi_rhs_columns_vs_oracle(S, oracle_mod, precision):
    """Seven right-hand sides in one solve: every column against the oracle."""
    A = S.gen_grid(16, 16, 12, 7, 0, 6.1, 0)
    order, sizes = S.order_grid(16, 16, 12, 16)
    _, B = S.gen_rhs(A, 7, 1)
    with S.SnuMat(A, 16, order=order, sizes=sizes, device=0, precision=precision) as M:
        X = M.factorize().solve(B)
    _vs_oracle(S, oracle_mod, A, 16, order, sizes, X, B, f"multi-rhs {precision}")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_members_vs_oracle(S, oracle_mod, precision):
    """Uneven batch (np 1..16, 2D and 3D, uneven column counts, skipped
    members): every member's columns against the oracle."""
    mem = _batch_members(S)
    rng = np.random.default_rng(7)
    counts = [1, 3, 0, 2, 1, 5, 1, 0, 2, 1]
    bs = [None if k == 0 else rng.standard_normal((A.n, k)) for (A, _, _), k in zip(mem, counts)]
    mats = [S.SnuMat(A, np_, order=order, sizes=sizes, device=0, precision=precision)
            for (A, np_, (order, sizes)) in mem]
    try:
        with S.Batch(mats) as B:
            xs = B.factorize().solve(bs)
    finally:
        for M in mats:
            M.close()
    for q, ((A, np_, (order, sizes)), b, x) in enumerate(zip(mem, bs, xs)):
        if b is None:
            assert x is None
            continue
        _vs_oracle(S, oracle_mod, A, np_, order, sizes, x, b, f"member {q}")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_c4_recipe_vs_oracle(S, oracle_mod, precision):
    """The batched workload's recipe (2D 9-point convection-diffusion,
    diagonal 8.1, sides log-uniform in [16, 192] from seed 2, matrix i's
    coefficients from seed i, leaves = power of two nearest n / 2048, one
    right-hand side each from seed 3): 64 of its 512 matrices (every 8th,
    sides spread over the whole range) factorized and solved as one batch,
    every member against the oracle."""
    mem = S.config_batch(512)[::8]
    bs = [S.gen_rhs(A, 1, 3)[1][:, 0] for A, _, _ in mem]
    mats = [S.SnuMat(A, np_, grid=(side, side, 1), device=0, precision=precision) for A, side, np_ in mem]
    try:
        with S.Batch(mats) as B:
            xs = B.factorize().solve(bs)
            fl, sl = B.launches()
        orders = [M.ordering() for M in mats]
    finally:
        for M in mats:
            M.close()
    print(f"batch of {len(mem)} (sides {min(s for _, s, _ in mem)}..{max(s for _, s, _ in mem)}): "
          f"{fl} dense factor launches, {sl} dense solve launches", flush=True)
    worst = 0.0
    for (A, side, np_), (order, sizes), b, x in zip(mem, orders, bs, xs):
        O = oracle_mod.Oracle(A, np_, order, sizes)
        xo = O.factor_solve(b, NT)
        O.close()
        rel = _rel(x, xo)
        worst = max(worst, rel)
        assert rel <= REL_TOL, (side, np_, rel)
        assert S.backward_error(A, x, b) <= BERR_TOL, (side, np_)
    print(f"  worst rel diff vs oracle {worst:.2e}")


def _serve_mix(S, rng):
    """Serving-shaped mix at test size: 3D 7-point grids (sides 10..20) and 2D
    5-point grids (sides 48..128), leaves as in the batched workload."""
    out = []
    for kind in (0, 1):
        for _ in range(4):
            if kind == 0:
                d = int(rng.integers(10, 21))
                A = S.gen_grid(d, d, d, 7, 0, 6.1, int(rng.integers(1 << 30)))
                g = (d, d, d)
            else:
                d = int(rng.integers(48, 129))
                A = S.gen_grid(d, d, 1, 5, 0, 4.1, int(rng.integers(1 << 30)))
                g = (d, d, 1)
            np_ = S.batch_np(A.n, 512)
            out.append((A, np_, S.order_grid(*g, np_)))
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_serve_stream_vs_oracle(S, oracle_mod, precision):
    """A serving stream: eight resident matrices (3D 7-point and 2D 5-point),
    requests with 2^k columns (k uniform in 0..4) arriving at exponential
    gaps while earlier ones are on the GPU, two new factorizations admitted
    in the middle of the stream; two slots pack the pending requests.  Every
    answered column against the oracle."""
    rng = np.random.default_rng(4)
    mix = _serve_mix(S, rng)
    extra = _serve_mix(S, np.random.default_rng(40))[3:5]
    arr = np.random.default_rng(5)
    with S.Server(budget_bytes=64 << 30, slots=2, max_cols=256) as srv:
        ids = [srv.add(A, np_, order=o, sizes=z, precision=precision) for A, np_, (o, z) in mix]
        known = dict(zip(ids, mix))
        lock = threading.Lock()

        def admit(entry):
            A, np_, (o, z) = entry
            mid = srv.add(A, np_, order=o, sizes=z, precision=precision)
            with lock:
                known[mid] = entry

        tickets, adds = [], []
        for q in range(30):
            time.sleep(float(arr.exponential(0.004)))
            if q in (10, 20):
                th = threading.Thread(target=admit, args=(extra[len(adds)],))
                th.start()
                adds.append(th)
            with lock:
                live = list(known.keys())
            mid = live[int(arr.integers(len(live)))]
            A = known[mid][0]
            b = arr.standard_normal((A.n, 1 << int(arr.integers(0, 5))))
            tickets.append((mid, b, srv.submit(mid, b)))
        answers = [(mid, b, srv.wait(t)[0]) for mid, b, t in tickets]
        for th in adds:
            th.join()
        st = srv.stats()
        print(f"served {st['requests']:.0f} requests in {st['packed_solves']:.0f} packed solves, "
              f"p50 {st['p50_ms']:.2f} ms p99 {st['p99_ms']:.2f} ms, {st['admissions']:.0f} admissions",
              flush=True)
        assert st["requests"] == len(tickets)
    for q, (mid, b, x) in enumerate(answers):
        A, np_, (o, z) = known[mid]
        _vs_oracle(S, oracle_mod, A, np_, o, z, x, b, f"request {q} (matrix {mid}, {b.shape[1]} cols)")


def _bytes_needed(S, A, np_, order, sizes, precision):
    import ctypes
    with S.SnuMat(A, np_, order=order, sizes=sizes, device=-1, precision=precision) as M:
        nb = ctypes.c_size_t()
        S._check(S.lib().snumat_bytes_needed(M._h, ctypes.byref(nb)))
        return nb.value


def test_serve_under_memory_pressure_vs_oracle(S, oracle_mod):
    """A working set larger than the budget: eight matrices, a budget that
    holds about 40% of their bytes, requests for all of them; the server
    parks idle least-recently-used matrices to admit others and admits a
    parked matrix again when a request for it arrives.  Evictions and
    re-admissions must happen; every answer against the oracle."""
    mix = _serve_mix(S, np.random.default_rng(7))
    need = [_bytes_needed(S, A, np_, o, z, "fp32") for A, np_, (o, z) in mix]
    budget = max(int(0.4 * sum(need)), max(need) + 1)
    arr = np.random.default_rng(8)
    with S.Server(budget_bytes=budget, slots=2, max_cols=256) as srv:
        ids = [srv.add(A, np_, order=o, sizes=z, precision="fp32") for A, np_, (o, z) in mix]
        tickets = []
        for q in range(40):
            key = int(arr.integers(len(mix)))
            b = arr.standard_normal((mix[key][0].n, 1 << int(arr.integers(0, 4))))
            tickets.append((key, b, srv.submit(ids[key], b)))
        answers = [(key, b, srv.wait(t)[0]) for key, b, t in tickets]
        st = srv.stats()
        print(f"budget {budget / 1e6:.1f} MB of {sum(need) / 1e6:.1f} MB: {st['admissions']:.0f} admissions, "
              f"{st['evictions']:.0f} evictions, {st['readmissions']:.0f} re-admissions, "
              f"{st['requests']:.0f} requests, p99 {st['p99_ms']:.2f} ms", flush=True)
        assert st["requests"] == len(tickets)
        assert st["evictions"] >= 3 and st["readmissions"] >= 3
        assert st["bytes_reserved"] <= budget
    for q, (key, b, x) in enumerate(answers):
        A, np_, (o, z) = mix[key]
        _vs_oracle(S, oracle_mod, A, np_, o, z, x, b, f"request {q} (matrix {key}, {b.shape[1]} cols)")


@pytest.mark.parametrize("ordering", ["benchmark", "geometric"])
def test_c2_full_size_vs_oracle(S, oracle_mod, ordering):
    """64^3 7-point, 128 leaves, with the benchmark's ordering (the graph
    ND) and with the geometric ND: every leaf's LU against the oracle's, x
    against the oracle; the factorization runs the 128 x 128 Schur
    launches."""
    A, grid, np_ = S.config_matrix("c2")
    if ordering == "benchmark":
        order, sizes, _ = S.config_order("c2", A, grid, np_)
    else:
        order, sizes = S.order_grid(*grid, np_)
    _, b = S.gen_rhs(A, 1, 1)
    b = b[:, 0]
    O = oracle_mod.Oracle(A, np_, order, sizes)
    t0 = time.perf_counter()
    xo = O.factor_solve(b, NT)
    print(f"  oracle factor+solve {time.perf_counter() - t0:.1f} s on {NT} threads", flush=True)
    with S.SnuMat(A, np_, order=order, sizes=sizes, device=0) as M:
        x = M.factorize().solve(b)
        split = M.gemm_split()
        bad = [r for r in range(np_) if leaf_rel(M.leaf_values(r), O.leaf_values(r)) > LEAF_TOL]
    O.close()
    print(f"  128x128 Schur launches {split['t128_launches']:.0f}, 64x64 {split['t64_launches']:.0f}")
    assert not bad, bad
    assert split["t128_launches"] > 0
    rel = _rel(x, xo)
    berr = S.backward_error(A, x, b)
    print(f"  rel diff vs oracle {rel:.2e}  backward error {berr:.2e}")
    assert rel <= REL_TOL and berr <= BERR_TOL


@pytest.mark.parametrize("fp32_gemm", ["split", "f32"])
def test_c2_full_size_fp32_vs_oracle(S, oracle_mod, fp32_gemm):
    """64^3 7-point, 128 leaves, benchmark ordering, FP32 factors + FP64
    refinement with both forms of the FP32 GEMM tiles (bf16 splits, FP32
    MFMA): x against the oracle's FP64 solution."""
    A, grid, np_ = S.config_matrix("c2")
    order, sizes, _ = S.config_order("c2", A, grid, np_)
    _, b = S.gen_rhs(A, 1, 1)
    b = b[:, 0]
    O = oracle_mod.Oracle(A, np_, order, sizes)
    xo = O.factor_solve(b, NT)
    O.close()
    with S.SnuMat(A, np_, order=order, sizes=sizes, device=0, precision="fp32", ir_tol=1e-12,
                  fp32_gemm=fp32_gemm) as M:
        x = M.factorize().solve(b)
        split = M.gemm_split()
        steps = M.timings()["ir_steps"]
    print(f"  128x128 launches {split['t128_launches']:.0f}, refinement steps {steps}")
    assert split["t128_launches"] > 0
    rel = _rel(x, xo)
    berr = S.backward_error(A, x, b)
    print(f"  rel diff vs oracle {rel:.2e}  backward error {berr:.2e}")
    assert rel <= REL_TOL and berr <= BERR_TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_many_leaves_27pt_vs_oracle(S, oracle_mod, precision):
    """27-point anisotropic operator (the precision workload's recipe) on a
    48^3 grid with 512 leaves, more leaves than CUs.  Leaf LU against the
    oracle's (FP64), x against the oracle in FP64 and FP32 + refinement."""
    A = S.gen_grid(48, 48, 48, 27, 1, 0.0, 0)
    order, sizes = S.order_grid(48, 48, 48, 512)
    _, B = S.gen_rhs(A, 3, 1)
    b = B[:, 0]
    O = oracle_mod.Oracle(A, 512, order, sizes)
    xo = O.factor_solve(b, NT)
    XO = np.column_stack([xo, O.factor_solve_columns(B[:, 1:], NT)])
    with S.SnuMat(A, 512, order=order, sizes=sizes, device=0, precision=precision) as M:
        x = M.factorize().solve(b)
        # several columns over more leaves than fit on the CUs at once: the
        # LDS leaf solves queue (one column in FP64 takes the global-memory
        # leaf solves)
        X = M.solve(B)
        info = M.info()
        bad = []
        if precision == "fp64":
            bad = [r for r in range(512) if leaf_rel(M.leaf_values(r), O.leaf_values(r)) > LEAF_TOL]
    O.close()
    print(f"  leaf fronts {info['leaf_fronts']}")
    assert not bad, bad
    rel = _rel(x, xo)
    berr = S.backward_error(A, x, b)
    print(f"  rel diff vs oracle {rel:.2e}  backward error {berr:.2e}")
    assert rel <= REL_TOL and berr <= BERR_TOL
    rel3, berr3 = _rel(X, XO), S.backward_error(A, X, B)
    print(f"  3 columns: rel diff vs oracle {rel3:.2e}  backward error {berr3:.2e}")
    assert rel3 <= REL_TOL and berr3 <= BERR_TOL


def test_stage0_scatter_bitexact(S, oracle_mod):
    """A's separator entries scattered into the dense blocks (stage 0: no
    leaf, no assembly) land bit for bit where the ordering puts them: block
    (p, q) holds A[i, j] at (order[i] - start_p, order[j] - start_q) for every
    entry with order[i] in separator p and order[j] in separator q, and zero
    elsewhere (the tile-compressed storage included); block ranges from the
    oracle's layout.  The leaves' LU arrays hold A's leaf values exactly where
    the oracle's value scatter puts them, zeros at the fill positions."""
    cases = [((20, 20, 20), 7, 0, 6.1, 16), ((33, 17, 1), 9, 0, 8.1, 8), ((10, 9, 8), 27, 1, 0.0, 8)]
    for grid, st, rec, dg, np_ in cases:
        A = S.gen_grid(*grid, st, rec, dg, 0)
        order, sizes = S.order_grid(*grid, np_)
        O = oracle_mod.Oracle(A, np_, order, sizes)
        bs, sz = O.blocks()
        O.leaf_init()
        leaves_o = [O.leaf_values(r) for r in range(np_)]
        O.close()
        order = np.asarray(order)
        rn = order[np.repeat(np.arange(A.n), np.diff(A.rowptr))]
        cn = order[A.colidx]
        nblocks = nonzero = 0
        with S.SnuMat(A, np_, order=order, sizes=sizes, device=0) as M:
            M.factorize_stage(0)
            for i in range(1, np_):
                anc = []
                a = i // 2
                while a >= 1:
                    anc.append(a)
                    a //= 2
                for p, q in [(i, i)] + [(i, a) for a in anc] + [(a, i) for a in anc]:
                    r0, r1 = bs[p - 1], bs[p - 1] + sz[p - 1]
                    c0, c1 = bs[q - 1], bs[q - 1] + sz[q - 1]
                    sel = (rn >= r0) & (rn < r1) & (cn >= c0) & (cn < c1)
                    e = np.zeros((sz[p - 1], sz[q - 1]))
                    e[rn[sel] - r0, cn[sel] - c0] = A.data[sel]
                    g = M.sep_block(p, q)
                    assert g.shape == e.shape and np.array_equal(g, e), (grid, p, q)
                    nblocks += 1
                    nonzero += int(sel.sum())
            for r in range(np_):
                assert np.array_equal(M.leaf_values(r), leaves_o[r]), (grid, r)
        print(f"  {grid} {st}-point np={np_}: {nblocks} blocks and {np_} leaves bit-equal, "
              f"{nonzero} separator entries of A", flush=True)
        assert nonzero > 0


def test_stage1_blocks_vs_oracle(S, oracle_mod):
    """After the leaves and the boundary Schur assembly (stage 1) every dense
    separator block equals the oracle's assembled block (A_SS entries, then
    each leaf's contribution in ascending rank) within 1e-13 relative, with
    the same zero pattern (the leaves' fronts sum the contributions in
    another order than the oracle's row-by-row elimination)."""
    cases = [((20, 20, 20), 7, 0, 6.1, 16), ((33, 17, 1), 9, 0, 8.1, 8), ((10, 9, 8), 27, 1, 0.0, 8),
             ((24, 24, 24), 7, 0, 6.1, 32)]
    for grid, st, rec, dg, np_ in cases:
        A = S.gen_grid(*grid, st, rec, dg, 0)
        order, sizes = S.order_grid(*grid, np_)
        O = oracle_mod.Oracle(A, np_, order, sizes)
        O.assemble_only(None, NT)
        exact = total = 0
        with S.SnuMat(A, np_, order=order, sizes=sizes, device=0) as M:
            M.factorize_stage(1)
            for i in range(1, np_):
                anc = []
                a = i // 2
                while a >= 1:
                    anc.append(a)
                    a //= 2
                for p, q in [(i, i)] + [(i, a) for a in anc] + [(a, i) for a in anc]:
                    g, o = M.sep_block(p, q), O.sep_block(p, q)
                    total += 1
                    exact += bool(np.array_equal(g, o))
                    assert np.array_equal(g == 0, o == 0), (grid, p, q)
                    scale = max(1.0, float(np.max(np.abs(o)))) if o.size else 1.0
                    err = float(np.max(np.abs(g - o))) / scale if o.size else 0.0
                    assert err <= 1e-13, (grid, p, q, err)
        O.close()
        print(f"  {grid} {st}-point np={np_}: {total} blocks, {exact} bit-equal", flush=True)


def _write_vec(path, v):
    with open(path, "w") as f:
        f.write("%%%%MatrixMarket matrix array real general\n%d 1\n" % len(v))
        for x in v:
            f.write("%.17g\n" % x)


def _read_vec(path, n):
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith("%")]
    assert lines[0].split()[:2] == [str(n), "1"]
    return np.array([float(ln) for ln in lines[1:1 + n]])


def _shuffled_mtx(src, dst, seed):
    """The same matrix with its entry lines in random order: read in file
    order, its CSR rows are not column-sorted."""
    with open(src) as f:
        lines = f.readlines()
    k = 0
    while lines[k].startswith("%"):
        k += 1
    body = lines[k + 1:]
    np.random.default_rng(seed).shuffle(body)
    with open(dst, "w") as f:
        f.writelines(lines[:k + 1] + body)


def test_reference_driver_through_compat_shims(S, oracle_mod):
    """The reference's driver flow (initialize(); solve(csr_matrix, b, x))
    linked against the C++ shims: matrix_20.mtx with the golden b, read in
    file order (also with shuffled entry lines: unsorted rows), solved on
    the GPU; x against the SuperLU golden solution and the oracle."""
    drv = os.path.join(ROOT, "snusolver_amd", "snu_ref_driver")
    gold = np.load(os.path.join(GOLD, "matrix_20_x.npz"))
    mtx = os.path.join(GOLD, "matrix_20.mtx")
    A = S.read_mtx(mtx)
    order, sizes = S.order_nd(A, 4)
    xo = oracle_mod.Oracle(A, 4, order, sizes).factor_solve(gold["b"])
    with tempfile.TemporaryDirectory() as d:
        bpath, xpath, shuf = (os.path.join(d, f) for f in ("b.mtx", "x.mtx", "shuffled.mtx"))
        _write_vec(bpath, gold["b"])
        _shuffled_mtx(mtx, shuf, 3)
        env = dict(os.environ, SNU_NP="4", SNU_DEVICE="0")
        for path in (mtx, shuf):
            r = subprocess.run([drv, path, "--b", bpath, "--out", xpath], capture_output=True, text=True,
                               env=env, timeout=120)
            print(r.stdout.strip(), r.stderr.strip())
            assert r.returncode == 0 and "max |Ax-b|" in r.stdout and "failed" not in r.stderr
            x = _read_vec(xpath, A.n)
            assert _rel(x, gold["x"]) <= REL_TOL and _rel(x, xo) <= REL_TOL
            assert S.backward_error(A, x, gold["b"]) <= BERR_TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32-mfma"])
def test_driver_writes_x_vs_oracle(S, oracle_mod, precision):
    """snu_driver: .mtx in (b = A x_true, seed 1), x out; the written x
    against the oracle; also with b read from a file, in file order."""
    drv = os.path.join(ROOT, "snusolver_amd", "snu_driver")
    mtx = os.path.join(GOLD, "matrix_20.mtx")
    A = S.read_mtx(mtx)
    order, sizes = S.order_nd(A, 4)
    _, b = S.gen_rhs(A, 1, 1)
    b = b[:, 0]
    xo = oracle_mod.Oracle(A, 4, order, sizes).factor_solve(b)
    extra = [] if precision == "fp64" else ["--" + precision]
    with tempfile.TemporaryDirectory() as d:
        xpath, bpath = os.path.join(d, "x.mtx"), os.path.join(d, "b.mtx")
        _write_vec(bpath, b)
        for args in ([], ["--b", bpath, "--file-order"]):
            r = subprocess.run([drv, mtx, "--np", "4", "--out", xpath] + args + extra, capture_output=True,
                               text=True, timeout=120)
            print(r.stdout.strip(), r.stderr.strip())
            assert r.returncode == 0, r.stderr
            x = _read_vec(xpath, A.n)
            assert _rel(x, xo) <= REL_TOL
            assert S.backward_error(A, x, b) <= BERR_TOL


def test_wide_leaf_factorizes(S):
    """One leaf of 48,400 rows (np = 1): its fronts reach well beyond the
    LDS of a workgroup and go through the global-memory front path; the
    handle and the one-shot solve (the reference's solve()) both solve; x
    against SuperLU (scipy) and the backward error."""
    import ctypes as C

    import scipy.sparse.linalg as spla
    A = S.gen_grid(220, 220, 1, 5, 0, 4.1, 9)
    _, b = S.gen_rhs(A, 1, 1)
    b = b[:, 0]
    xs = spla.spsolve(A.to_scipy().tocsc(), b)
    with S.SnuMat(A, 1, grid=(220, 220, 1), device=0) as M:
        x = M.factorize().solve(b)
    assert _rel(x, xs) <= REL_TOL
    assert S.backward_error(A, x, b) <= BERR_TOL
    o = S._Opts()
    S.lib().snu_default_opts(C.byref(o))
    o.np = 1
    o.ordering = S.ORDER_GRID
    o.grid[0], o.grid[1], o.grid[2] = 220, 220, 1
    x = np.zeros(A.n)
    a = A.c()
    S._check(S.lib().snu_solve_csr(C.byref(a), S._dp(b), S._dp(x), C.byref(o)))
    assert _rel(x, xs) <= REL_TOL
    assert S.backward_error(A, x, b) <= BERR_TOL


def test_batch_and_serve_stalled_refinement_fall_back_to_fp64(S, oracle_mod):
    """FP32 members whose refinement cannot reach 1e-12 (forced from the
    host: no correction steps allowed) are answered by FP64 factors of the
    same matrix, in a batch and in the serving loop: each answer equals the
    FP64 handle's solve of the same columns and meets the oracle bars."""
    mem = _batch_members(S)[:5]
    rng = np.random.default_rng(11)
    counts = [1, 2, 0, 3, 1]
    bs = [None if k == 0 else rng.standard_normal((A.n, k)) for (A, _, _), k in zip(mem, counts)]
    ref = []
    for (A, np_, (o, z)), b in zip(mem, bs):
        if b is None:
            ref.append(None)
            continue
        with S.SnuMat(A, np_, order=o, sizes=z, device=0) as M64:
            ref.append(M64.factorize().solve(b))
    mats = [S.SnuMat(A, np_, order=o, sizes=z, device=0, precision="fp32", ir_max_steps=-1)
            for (A, np_, (o, z)) in mem]
    try:
        with S.Batch(mats) as B:
            xs = B.factorize().solve(bs)
            t = B.timings()
    finally:
        for M in mats:
            M.close()
    print(f"batch: {t['fp64_fallbacks']:.0f} members fell back, refinement backward error {t['ir_berr']:.2e}",
          flush=True)
    assert t["fp64_fallbacks"] == sum(1 for k in counts if k) and t["ir_berr"] > 1e-12
    for q, ((A, np_, (o, z)), b, x, x64) in enumerate(zip(mem, bs, xs, ref)):
        if b is None:
            assert x is None
            continue
        assert np.array_equal(x, x64), (q, float(np.max(np.abs(x - x64))))
        _vs_oracle(S, oracle_mod, A, np_, o, z, x, b, f"batch member {q} (FP64 fallback)")
    with S.Server(budget_bytes=16 << 30, slots=1, max_cols=64) as srv:
        ids = [srv.add(A, np_, order=o, sizes=z, precision="fp32", ir_max_steps=-1)
               for (A, np_, (o, z)) in mem]
        tickets = [(q, srv.submit(mid, b)) for q, (mid, b) in enumerate(zip(ids, bs)) if b is not None]
        for q, tk in tickets:
            x, _ = srv.wait(tk)
            A, np_, (o, z) = mem[q]
            rel = _rel(x, ref[q])
            assert rel <= 1e-12, (q, rel)
            _vs_oracle(S, oracle_mod, A, np_, o, z, x, bs[q], f"served request {q} (FP64 fallback)")
