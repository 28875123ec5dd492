"""The allocation-free solve (snumat_solve_device_scratch) against the
allocating one (snumat_solve_device), through the C-ABI.

Same kernels and launch plan, other buffers: the solutions must be bitwise
equal.  Two calls in a row on one device, the second on a smaller input
(another matrix, fewer columns), so the second call runs over scratch
contents the first one left.  The handle's device bytes must not change
across the scratch calls (it allocates nothing), and a request beyond the
scratch must be refused with an error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_both(S, M, B):
    n, k = B.shape
    d_b = S.DeviceBuffer(8 * n * k)
    d_x = S.DeviceBuffer(8 * n * k)
    try:
        d_b.upload(B)
        before = M.info()["device_bytes"]
        M.solve_device_scratch_ptr(d_b.ptr.value, d_x.ptr.value, k)
        S.device_sync()
        after = M.info()["device_bytes"]
        xs = np.zeros((n, k), order="F")
        d_x.download(xs)
        M.solve_device_ptr(d_b.ptr.value, d_x.ptr.value, k)
        S.device_sync()
        xa = np.zeros((n, k), order="F")
        d_x.download(xa)
    finally:
        d_b.free()
        d_x.free()
    return xs, xa, before, after


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_scratch_solve_equals_allocating_solve(S, prec):
    big = S.gen_grid(24, 24, 24, 7, 0, 6.1, 0)
    small = S.gen_grid(40, 40, 1, 5, 0, 4.1, 7)
    runs = [(big, (24, 24, 24), 32, 4), (small, (40, 40, 1), 4, 1)]
    for A, grid, np_, k in runs:
        print("case", prec, A.n, "nrhs", k, flush=True)
        _, B = S.gen_rhs(A, k, 1)
        with S.SnuMat(A, np_, grid=grid, device=0, precision=prec, ir_tol=1e-12) as M:
            M.factorize()
            xs, xa, before, after = _solve_both(S, M, B)
            assert before == after, "the scratch solve allocated device memory"
            assert np.array_equal(xs, xa), float(np.max(np.abs(xs - xa)))
            assert S.backward_error(A, xs, B) <= 1e-12


def test_scratch_solve_refuses_oversized_request(S):
    A = S.gen_grid(64, 64, 64, 7, 0, 6.1, 0)
    k = 128  # 2 x 8 B x 262,144 x 128 = 512 MiB of solve vectors: beyond the scratch
    with S.SnuMat(A, 128, grid=(64, 64, 64), device=0) as M:
        M.factorize()
        d_b = S.DeviceBuffer(8 * A.n * k)
        d_x = S.DeviceBuffer(8 * A.n * k)
        try:
            with pytest.raises(S.SnuError) as e:
                M.solve_device_scratch_ptr(d_b.ptr.value, d_x.ptr.value, k)
            assert "scratch" in str(e.value)
        finally:
            d_b.free()
            d_x.free()
        # the handle still solves afterwards through the scratch path
        _, B = S.gen_rhs(A, 1, 1)
        xs, xa, _, _ = _solve_both(S, M, B)
        assert np.array_equal(xs, xa)
