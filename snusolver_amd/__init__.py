"""MI355X-native SnuSolver factor-and-solve path (Python host mirror).

Thin ctypes layer over the C-ABI in ``include/snusolver.h`` (libsnusolver.so,
built in-tree).  It mirrors the reference's operator interface for this path:

* ``SnuMat(A, ...)``  -- construct: ordering, symbolic structure, device
  layout (reference ``SnuMat::SnuMat``, src/SnuMat.cu:67-112)
* ``SnuMat.factorize()`` / ``SnuMat.solve(b)`` -- the reference's fused
  ``SnuMat::solve(double *x)`` (src/SnuMat_factorize.cu:213-357), split so the
  factors stay resident in HBM
* ``solve(A, b)`` -- the reference's ``void solve(csr_matrix, double*,
  double*)`` (src/solver.cu:51-89)

There is no CPU fallback: every numeric call runs the HIP kernels and raises
``SnuError`` when the library or the GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

__all__ = [
    "SnuError", "CSR", "SnuMat", "lib", "gen_grid", "gen_rhs", "order_grid",
    "order_nd", "read_mtx", "backward_error", "solve", "config_matrix", "Batch",
    "config_batch", "batch_sides", "batch_np", "Server", "read_vector", "write_vector",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsnusolver.so")

ORDER_ND, ORDER_GRID, ORDER_GIVEN = 0, 1, 2
FP64, FP32 = 0, 1
GEMM_SPLIT, GEMM_F32 = 0, 1
_STATUS = {
    -1: "BAD_INPUT", -2: "STRUCTURE", -3: "ZERO_DIAG", -4: "DUPLICATE",
    -5: "NO_DEVICE", -6: "HIP", -7: "OOM", -8: "NOT_FACTORED", -9: "NUMERIC",
    -10: "NOT_CONVERGED",
}


class SnuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("nnz", C.c_int),
                ("rowptr", C.POINTER(C.c_int)), ("colidx", C.POINTER(C.c_int)),
                ("data", C.POINTER(C.c_double))]


class _Opts(C.Structure):
    _fields_ = [("np", C.c_int), ("ordering", C.c_int), ("grid", C.c_int * 3),
                ("order", C.POINTER(C.c_int)), ("sizes", C.POINTER(C.c_int)),
                ("eps", C.c_double), ("precision", C.c_int), ("device", C.c_int),
                ("seed", C.c_uint64), ("ir_max_steps", C.c_int), ("ir_tol", C.c_double),
                ("fp32_gemm", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    """Load libsnusolver.so (fails loudly if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SnuError(-5, f"{LIB_PATH} is missing; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.snu_last_error.restype = C.c_char_p
        L.snu_backward_error.restype = C.c_double
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise SnuError(rc, lib().snu_last_error().decode())


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class CSR:
    """CSR matrix with the reference's csr_matrix fields (src/matrix.h:9-13)."""
    n: int
    rowptr: np.ndarray
    colidx: np.ndarray
    data: np.ndarray

    def __post_init__(self):
        self.rowptr = np.ascontiguousarray(self.rowptr, dtype=np.int32)
        self.colidx = np.ascontiguousarray(self.colidx, dtype=np.int32)
        self.data = np.ascontiguousarray(self.data, dtype=np.float64)

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def c(self) -> _Csr:
        return _Csr(self.n, self.n, self.nnz, _ip(self.rowptr), _ip(self.colidx), _dp(self.data))

    def matvec(self, x: np.ndarray) -> np.ndarray:
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        x = np.asarray(x, dtype=np.float64)
        if x.ndim == 1:
            return np.bincount(rows, weights=self.data * x[self.colidx], minlength=self.n)
        return np.stack([self.matvec(x[:, j]) for j in range(x.shape[1])], axis=1)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.data, self.colidx, self.rowptr), shape=(self.n, self.n))


def gen_grid(nx: int, ny: int, nz: int, stencil: int, recipe: int = 0,
             diag: float = 4.1, seed: int = 0) -> CSR:
    """Seeded grid operator; see snu_gen.cpp for the recipe."""
    n, nnz = C.c_int(), C.c_int()
    _check(lib().snu_gen_grid_size(nx, ny, nz, stencil, C.byref(n), C.byref(nnz)))
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(nnz.value, np.int32)
    v = np.zeros(nnz.value, np.float64)
    _check(lib().snu_gen_grid(nx, ny, nz, stencil, recipe, C.c_double(diag), C.c_uint64(seed),
                              _ip(rp), _ip(ci), _dp(v)))
    return CSR(n.value, rp, ci, v)


def gen_rhs(A: CSR, nrhs: int = 1, seed: int = 1):
    """x_true ~ U[-1,1) (seeded), b = A x_true; both n x nrhs (Fortran order)."""
    xt = np.zeros((nrhs, A.n), np.float64)
    b = np.zeros((nrhs, A.n), np.float64)
    a = A.c()
    _check(lib().snu_gen_rhs(C.byref(a), nrhs, C.c_uint64(seed), _dp(xt), _dp(b)))
    return xt.T, b.T


def order_grid(nx: int, ny: int, nz: int, np_: int):
    order = np.zeros(nx * ny * nz, np.int32)
    sizes = np.zeros(2 * np_ - 1, np.int32)
    _check(lib().snu_order_grid(nx, ny, nz, np_, _ip(order), _ip(sizes)))
    return order, sizes


ND_METHODS = {"default": 0, "levelset": 1, "multilevel": 2}


def order_nd(A: CSR, np_: int, seed: int = 0, method: str = "default"):
    """Graph nested dissection of A + A^T in the ParMETIS order/sizes
    contract.  method "default": at every bisection the smaller of the
    multilevel and the level-set separator; "levelset" / "multilevel": one
    kind only."""
    order = np.zeros(A.n, np.int32)
    sizes = np.zeros(2 * np_ - 1, np.int32)
    _check(lib().snu_order_nd_ex(A.n, _ip(A.rowptr), _ip(A.colidx), np_, C.c_uint64(seed),
                                 ND_METHODS[method], _ip(order), _ip(sizes)))
    return order, sizes


MTX_FILE_ORDER, MTX_NO_EXPAND = 1, 2


def read_mtx(path: str, file_order: bool = False, expand: bool = True) -> CSR:
    """Matrix Market coordinate file -> CSR.  Default: symmetric files
    expanded, rows column-sorted.  file_order=True, expand=False reads what
    the reference's read_matrix reads (rows in file order, stored triangle
    only)."""
    flags = (MTX_FILE_ORDER if file_order else 0) | (0 if expand else MTX_NO_EXPAND)
    n, m, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    p = path.encode()
    _check(lib().snu_mtx_size_ex(p, flags, C.byref(n), C.byref(m), C.byref(nnz), C.byref(sym)))
    if n.value != m.value:
        raise SnuError(-1, "matrix is not square")
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(nnz.value, np.int32)
    v = np.zeros(nnz.value, np.float64)
    _check(lib().snu_mtx_read_ex(p, flags, _ip(rp), _ip(ci), _dp(v)))
    return CSR(n.value, rp, ci, v)


def read_vector(path: str, n: int) -> np.ndarray:
    """Dense n x 1 Matrix Market vector (array or one-column coordinate)."""
    x = np.zeros(n, np.float64)
    _check(lib().snu_mtx_read_vector(path.encode(), n, _dp(x)))
    return x


def write_vector(path: str, x: np.ndarray):
    x = np.ascontiguousarray(x, np.float64)
    _check(lib().snu_mtx_write_vector(path.encode(), _dp(x), len(x)))


class DeviceBuffer:
    """Device memory for right-hand sides / solutions kept resident in HBM."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        _check(lib().snu_dev_alloc(C.c_size_t(nbytes), C.byref(self.ptr)))

    def upload(self, arr: np.ndarray):
        a = np.asfortranarray(arr)  # column-major, as the C-ABI expects
        assert a.nbytes <= self.nbytes
        _check(lib().snu_dev_copy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1))

    def download(self, arr: np.ndarray):
        assert arr.flags.f_contiguous, "download into a column-major array"
        _check(lib().snu_dev_copy(arr.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(arr.nbytes), 2))

    def free(self):
        if self.ptr.value:
            lib().snu_dev_free(self.ptr)
            self.ptr = C.c_void_p()


def device_sync():
    _check(lib().snu_dev_sync())


def backward_error(A: CSR, x: np.ndarray, b: np.ndarray) -> float:
    """||Ax - b||_inf / (||A||_inf ||x||_inf + ||b||_inf), per column max."""
    x = np.asarray(x, np.float64)
    b = np.asarray(b, np.float64)
    if x.ndim == 2:
        return max(backward_error(A, x[:, j], b[:, j]) for j in range(x.shape[1]))
    r = A.matvec(x) - b
    anorm = np.max(np.bincount(np.repeat(np.arange(A.n), np.diff(A.rowptr)),
                               weights=np.abs(A.data), minlength=A.n))
    den = anorm * np.max(np.abs(x)) + np.max(np.abs(b))
    return float(np.max(np.abs(r)) / den) if den > 0 else float(np.max(np.abs(r)))


class SnuMat:
    """construct -> factorize -> solve on one GPU (device=-1: host analysis only)."""

    def __init__(self, A: CSR, np_: int = 4, ordering: str = "nd", grid=None,
                 order=None, sizes=None, device: int = 0, eps: float = 1e-15, seed: int = 0,
                 precision: str = "fp64", ir_max_steps: int = 10, ir_tol: float = 1e-15,
                 fp32_gemm: str = "split"):
        """precision "fp64": the reference's arithmetic; "fp32": FP32 factors
        and sweeps with FP64 iterative refinement (ir_max_steps, ir_tol).
        fp32_gemm ("fp32" only): "split", the factorization's GEMM tiles as
        three-piece bf16 splits on the bf16 MFMA (FP32-accurate), or "f32",
        the FP32 MFMA."""
        self.A = A
        o = _Opts()
        lib().snu_default_opts(C.byref(o))
        o.np = np_
        o.eps = eps
        o.device = device
        o.seed = seed
        if precision not in ("fp64", "fp32"):
            raise ValueError(precision)
        o.precision = FP32 if precision == "fp32" else FP64
        o.ir_max_steps = ir_max_steps
        o.ir_tol = ir_tol
        if fp32_gemm not in ("split", "f32"):
            raise ValueError(fp32_gemm)
        o.fp32_gemm = GEMM_SPLIT if fp32_gemm == "split" else GEMM_F32
        self.precision = precision
        self._keep = []
        if order is not None:
            o.ordering = ORDER_GIVEN
            order = np.ascontiguousarray(order, np.int32)
            sizes = np.ascontiguousarray(sizes, np.int32)
            self._keep += [order, sizes]
            o.order, o.sizes = _ip(order), _ip(sizes)
        elif ordering == "grid" or grid is not None:
            o.ordering = ORDER_GRID
            g = list(grid) + [1] * (3 - len(grid))
            o.grid[0], o.grid[1], o.grid[2] = g
        else:
            o.ordering = ORDER_ND
        self.np = np_
        self._h = C.c_void_p()
        a = A.c()
        _check(lib().snumat_construct(C.byref(a), C.byref(o), C.byref(self._h)))

    # -- lifecycle --
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().snumat_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- numerics --
    def factorize(self):
        _check(lib().snumat_factorize(self._h))
        return self

    def update_values(self, values: np.ndarray):
        v = np.ascontiguousarray(values, np.float64)
        _check(lib().snumat_update_values(self._h, _dp(v)))

    def solve(self, b: np.ndarray) -> np.ndarray:
        b = np.asarray(b, np.float64)
        one = b.ndim == 1
        B = np.asfortranarray(b.reshape(self.A.n, -1))
        X = np.zeros_like(B, order="F")
        _check(lib().snumat_solve(self._h, _dp(B), _dp(X), B.shape[1], self.A.n))
        return X[:, 0].copy() if one else X

    def solve_device_ptr(self, d_b: int, d_x: int, nrhs: int):
        _check(lib().snumat_solve_device(self._h, C.c_void_p(d_b), C.c_void_p(d_x), nrhs, self.A.n))

    def solve_device_scratch_ptr(self, d_b: int, d_x: int, nrhs: int):
        """snumat_solve_device_scratch: the same solve with every working
        buffer in the library's per-device scratch (no device allocation)."""
        _check(lib().snumat_solve_device_scratch(self._h, C.c_void_p(d_b), C.c_void_p(d_x), nrhs,
                                                 self.A.n))

    def set_profiling(self, on: bool):
        _check(lib().snumat_set_profiling(self._h, 1 if on else 0))

    def detach(self):
        """Release the handle's device memory, keeping its symbolic plan
        (host-only afterwards, as constructed with device=-1)."""
        _check(lib().snumat_detach(self._h))
        return self

    def attach(self, device: int = 0, values=None):
        """Move a host-only or detached handle onto a device with A's
        values (default: those it was constructed with)."""
        v = np.ascontiguousarray(self.A.data if values is None else values, np.float64)
        _check(lib().snumat_attach(self._h, device, _dp(v)))
        return self

    def use_graph(self, on: bool = True):
        """Factorize by replaying a HIP graph of the launch sequence (captured
        on the next factorize)."""
        _check(lib().snumat_use_graph(self._h, 1 if on else 0))
        return self

    # -- introspection --
    def info(self) -> dict:
        a = (C.c_int64 * 16)()
        _check(lib().snumat_info(self._h, a))
        keys = ["n", "nnz", "np", "num_block", "max_level", "leaf_lu_nnz", "dense_entries",
                "max_core_n", "max_leaf_size", "max_boundary_rows", "max_boundary_cols",
                "on_device", "factored", "device_bytes", "leaf_fronts", "boundary_targets"]
        return {k: int(a[i]) for i, k in enumerate(keys)}

    def blocks(self):
        nb = 2 * self.np - 1
        bs, sz = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
        _check(lib().snumat_blocks(self._h, _ip(bs), _ip(sz)))
        return bs, sz

    def ordering(self):
        order, sizes = np.zeros(self.A.n, np.int32), np.zeros(2 * self.np - 1, np.int32)
        _check(lib().snumat_ordering(self._h, _ip(order), _ip(sizes)))
        return order, sizes

    def permuted(self):
        rp = np.zeros(self.A.n + 1, np.int32)
        ci = np.zeros(self.A.nnz, np.int32)
        pm = np.zeros(self.A.nnz, np.int32)
        _check(lib().snumat_permuted(self._h, _ip(rp), _ip(ci), _ip(pm)))
        return rp, ci, pm

    def leaf_sizes(self, rank: int):
        s = np.zeros(4, np.int32)
        _check(lib().snumat_leaf_sizes(self._h, rank, _ip(s)))
        return [int(v) for v in s]

    def leaf_struct(self, rank: int):
        ls, cn, cnz, lnz = self.leaf_sizes(rank)
        rp, ci = np.zeros(cn + 1, np.int32), np.zeros(lnz, np.int32)
        dg, bias = np.zeros(cn, np.int32), np.zeros(ls, np.int32)
        cm = np.zeros(cnz, np.int32)
        _check(lib().snumat_leaf_struct(self._h, rank, _ip(rp), _ip(ci), _ip(dg), _ip(bias), _ip(cm)))
        return rp, ci, dg, bias, cm

    def leaf_fronts(self, rank: int) -> dict:
        """Multifrontal structure of a leaf (test hook): per front first
        pivot k0, pivots nk, order f, parent, height; B lists, extend-add
        indices, owned LU slots and positions, children, root-front boundary
        rows / cols."""
        s = np.zeros(4, np.int32)
        _check(lib().snumat_leaf_front_sizes(self._h, rank, _ip(s)))
        nf, ni, ns, nc = (int(v) for v in s)
        a = {"front": np.zeros(5 * nf, np.int32), "idx_ptr": np.zeros(nf + 1, np.int32),
             "idx": np.zeros(ni, np.int32), "rel": np.zeros(ni, np.int32),
             "slot_ptr": np.zeros(nf + 1, np.int32), "slot": np.zeros(ns, np.int32),
             "pos": np.zeros(ns, np.int32), "child_ptr": np.zeros(nf + 1, np.int32),
             "child": np.zeros(nc, np.int32), "srow": np.zeros(ni, np.int32), "scol": np.zeros(ni, np.int32)}
        keys = ["front", "idx_ptr", "idx", "rel", "slot_ptr", "slot", "pos", "child_ptr", "child", "srow", "scol"]
        _check(lib().snumat_leaf_fronts(self._h, rank, *[_ip(a[k]) for k in keys]))
        fr = a.pop("front").reshape(nf, 5)
        a.update(k0=fr[:, 0], nk=fr[:, 1], f=fr[:, 2], parent=fr[:, 3], height=fr[:, 4])
        a["k_ptr"] = np.zeros(nf + 1, np.int32)
        a["k"] = np.zeros(int(fr[:, 1].sum()), np.int32)
        _check(lib().snumat_leaf_front_pivots(self._h, rank, _ip(a["k_ptr"]), _ip(a["k"])))
        return a

    def leaf_values(self, rank: int) -> np.ndarray:
        lnz = self.leaf_sizes(rank)[3]
        v = np.zeros(lnz, np.float64)
        _check(lib().snumat_leaf_values(self._h, rank, _dp(v)))
        return v

    def sep_diag(self, e: int) -> np.ndarray:
        s = int(self.blocks()[1][e - 1])
        out = np.zeros(s * s, np.float64)
        _check(lib().snumat_sep_diag(self._h, e, _dp(out)))
        return out.reshape(s, s, order="F")

    def factorize_stage(self, stage: int):
        """Debug: stage 0 scatters A's values into place only (the separator
        blocks then hold A_SS); stage 1 also runs the leaves and the boundary
        Schur assembly (A_SS plus the leaf contributions)."""
        _check(lib().snumat_factorize_stage(self._h, stage))
        return self

    def sep_block(self, i: int, j: int) -> np.ndarray:
        """Dense separator block (i, j) (i, j on one root path), s_i x s_j."""
        sz = self.blocks()[1]
        si, sj = int(sz[i - 1]), int(sz[j - 1])
        out = np.zeros(si * sj, np.float64)
        _check(lib().snumat_sep_block(self._h, i, j, _dp(out)))
        return out.reshape(si, sj, order="F")

    def timings(self) -> dict:
        t = (C.c_double * 16)()
        _check(lib().snumat_timings(self._h, t))
        keys = ["scatter_ms", "leaf_fronts_ms", "leaf_schur_ms", "assemble_ms", "dense_ms",
                "factor_ms", "solve_ms", "gemm_sub_ms", "gemm_sub_launches", "gemm_sub_gflop",
                "gemm_assign_ms", "gemm_assign_launches", "gemm_assign_gflop",
                "ir_steps", "ir_berr", "fp64_fallback"]
        return {k: float(t[i]) for i, k in enumerate(keys)}

    def gemm_split(self) -> dict:
        """C -= AB launches of the last factorize (leaf fronts and dense
        levels) by tile size (kernel ms when profiling, launches, GFLOP)."""
        t = (C.c_double * 6)()
        _check(lib().snumat_gemm_split(self._h, t))
        return {"t64_ms": t[0], "t64_launches": t[1], "t64_gflop": t[2],
                "t128_ms": t[3], "t128_launches": t[4], "t128_gflop": t[5]}

    def work(self) -> dict:
        """Work the launch plans execute (structurally zero tiles skipped)."""
        w = (C.c_double * 8)()
        _check(lib().snumat_work(self._h, w))
        return {"dense_flops": w[0], "dense_launches": w[1], "diag_lu_launches": w[2],
                "solve_bytes": w[3], "solve_dense_bytes": w[4], "solve_leaf_bytes": w[5],
                "solve_launches": w[6]}

    def counts(self) -> dict:
        c = (C.c_double * 4)()
        _check(lib().snumat_counts(self._h, c))
        return {"f_dense": c[0], "f_leaf": c[1], "b_leaf": c[2], "b_solve": c[3]}


class Batch:
    """Several constructed matrices factorized and solved together, one
    launch per step across all of them (leaves of all matrices together,
    separators of equal tree height together).  Each matrix's arithmetic is
    that of its standalone factorization and solve."""

    def __init__(self, mats):
        self.mats = list(mats)
        arr = (C.c_void_p * len(self.mats))(*[m._h.value for m in self.mats])
        self._h = C.c_void_p()
        _check(lib().snubatch_create(arr, len(self.mats), C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().snubatch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def factorize(self):
        _check(lib().snubatch_factorize(self._h))
        return self

    def use_graph(self, on: bool = True):
        _check(lib().snubatch_use_graph(self._h, 1 if on else 0))
        return self

    def solve(self, bs):
        """bs: one array per matrix (n or n x k; None skips the matrix).
        Returns the solutions in the same shapes."""
        k = len(self.mats)
        Bs, Xs, outs = [], [], []
        nrhs = (C.c_int * k)()
        ld = (C.c_int * k)()
        pb = (C.c_void_p * k)()
        px = (C.c_void_p * k)()
        for i, (M, b) in enumerate(zip(self.mats, bs)):
            ld[i] = M.A.n
            if b is None:
                nrhs[i] = 0
                Bs.append(None)
                Xs.append(None)
                continue
            b = np.asarray(b, np.float64)
            B = np.asfortranarray(b.reshape(M.A.n, -1))
            X = np.zeros_like(B, order="F")
            nrhs[i] = B.shape[1]
            pb[i] = B.ctypes.data
            px[i] = X.ctypes.data
            Bs.append(B)
            Xs.append((X, b.ndim == 1))
        _check(lib().snubatch_solve(self._h, pb, px, nrhs, ld))
        for x in Xs:
            outs.append(None if x is None else (x[0][:, 0].copy() if x[1] else x[0]))
        return outs

    def timings(self) -> dict:
        t = (C.c_double * 16)()
        _check(lib().snubatch_timings(self._h, t))
        keys = ["scatter_ms", "leaf_fronts_ms", "leaf_schur_ms", "assemble_ms", "dense_ms",
                "factor_ms", "solve_ms"]
        out = {k: float(t[i]) for i, k in enumerate(keys)}
        out["ir_steps"] = float(t[13])
        out["ir_berr"] = float(t[14])
        out["fp64_fallbacks"] = float(t[15])
        return out

    def launches(self):
        f, s_ = C.c_int(), C.c_int()
        _check(lib().snubatch_launches(self._h, C.byref(f), C.byref(s_)))
        return f.value, s_.value


class Server:
    """Serving loop: resident factorizations under an HBM byte budget,
    solve requests packed across matrices by `slots` worker threads."""

    def __init__(self, budget_bytes: int, slots: int = 2, max_cols: int = 256, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().snuserve_create(device, C.c_size_t(int(budget_bytes)), slots, max_cols,
                                     C.byref(self._h)))
        self._keep = {}

    def add(self, A: CSR, np_: int, grid=None, precision: str = "fp32", order=None, sizes=None,
            ir_max_steps: int = 10, ir_tol: float = 1e-15) -> int:
        o = _Opts()
        lib().snu_default_opts(C.byref(o))
        o.np = np_
        o.precision = FP32 if precision == "fp32" else FP64
        o.ir_max_steps = ir_max_steps
        o.ir_tol = ir_tol
        keep = []
        if order is not None:
            o.ordering = ORDER_GIVEN
            order = np.ascontiguousarray(order, np.int32)
            sizes = np.ascontiguousarray(sizes, np.int32)
            keep += [order, sizes]
            o.order, o.sizes = _ip(order), _ip(sizes)
        elif grid is not None:
            o.ordering = ORDER_GRID
            g = list(grid) + [1] * (3 - len(grid))
            o.grid[0], o.grid[1], o.grid[2] = g
        a = A.c()
        mid = C.c_int()
        _check(lib().snuserve_add(self._h, C.byref(a), C.byref(o), C.byref(mid)))
        return mid.value

    def remove(self, mid: int):
        _check(lib().snuserve_remove(self._h, mid))

    def submit(self, mid: int, b: np.ndarray):
        """Queue a solve; returns a ticket holding the (b, x) buffers."""
        b = np.asarray(b, np.float64)
        B = np.asfortranarray(b.reshape(b.shape[0], -1))
        X = np.zeros_like(B, order="F")
        t = C.c_longlong()
        _check(lib().snuserve_submit(self._h, mid, _dp(B), _dp(X), B.shape[1], B.shape[0], C.byref(t)))
        self._keep[t.value] = (B, X, b.ndim == 1)
        return t.value

    def wait(self, ticket: int):
        """Block until answered; returns (x, latency_ms)."""
        lat = C.c_double()
        rc = lib().snuserve_wait(self._h, C.c_longlong(ticket), C.byref(lat))
        B, X, one = self._keep.pop(ticket)
        _check(rc)
        return (X[:, 0].copy() if one else X), lat.value

    def stats(self, reset: bool = False) -> dict:
        st = (C.c_double * 16)()
        _check(lib().snuserve_stats(self._h, st, 1 if reset else 0))
        keys = ["requests", "columns", "packed_solves", "p50_ms", "p99_ms", "mean_ms", "max_ms",
                "evictions", "admissions", "resident", "bytes_reserved", "wall_ms", "slot_busy_ms",
                "readmissions", "parked"]
        return {k: float(st[i]) for i, k in enumerate(keys)}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().snuserve_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def solve(A: CSR, b: np.ndarray, np_: int = 4, **kw) -> np.ndarray:
    """One-shot construct + factorize + solve (reference ``solve``)."""
    with SnuMat(A, np_, **kw) as M:
        return M.factorize().solve(b)


def batch_sides(count: int = 512, lo: int = 16, hi: int = 192, seed: int = 2):
    """Grid sides of the batched workload: log-uniform in [lo, hi] from
    numpy's default generator (PCG64) with the given seed, rounded."""
    rng = np.random.default_rng(seed)
    return [int(round(v)) for v in np.exp(rng.uniform(np.log(lo), np.log(hi), count))]


def batch_np(n: int, rows_per_leaf: int = 2048) -> int:
    """Leaves per matrix: the power of two nearest to n / rows_per_leaf (>= 1)."""
    r = max(1.0, n / rows_per_leaf)
    return 1 << int(round(np.log2(r)))


def config_batch(count: int = 512, seed: int = 2):
    """The batched workload: count 2D 9-point convection-diffusion matrices
    (diagonal 8.1; matrix i draws its coefficients from seed i) with sides
    from batch_sides; returns a list of (A, side, np)."""
    out = []
    for i, side in enumerate(batch_sides(count, seed=seed)):
        A = gen_grid(side, side, 1, 9, 0, 8.1, i)
        out.append((A, side, batch_np(A.n)))
    return out


def config_matrix(name: str):
    """Matrices of the named workloads: returns (A, grid dims, np)."""
    if name == "c1":      # 20x20 2D 5-point, diag 4.1, seed 0; single process, np = 4
        return gen_grid(20, 20, 1, 5, 0, 4.1, 0), (20, 20, 1), 4
    if name == "c2":      # 64^3 7-point, diag 6.1, seed 0; leaves of ~2k rows
        return gen_grid(64, 64, 64, 7, 0, 6.1, 0), (64, 64, 64), 128
    if name == "c3":      # 96^3 27-point anisotropic 1:10:100, seed 0
        return gen_grid(96, 96, 96, 27, 1, 0.0, 0), (96, 96, 96), 512
    raise ValueError(name)


def config_order(name: str, A: CSR, grid, np_: int):
    """Ordering of the named workloads: (order, sizes, description).  The
    one with the fewer factorization flops: on the 7-point grids the graph
    ND (its breadth-first level sets run diagonally and are smaller than
    axis planes: 64^3 2.4 against 6.1 nominal TFLOP), on the 27-point grid the
    geometric ND (127 against 133 TFLOP for the graph ND)."""
    if name == "c2":
        o, z = order_nd(A, np_, 0)
        return o, z, "graph ND (multilevel / level-set, smaller separator per bisection), seed 0"
    o, z = order_grid(*grid, np_)
    return o, z, "geometric ND"

