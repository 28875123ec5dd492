"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product never imports
it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """Reference restatement: structure at construction, then factor_solve."""

    def __init__(self, A, np_: int, order, sizes):
        self.A = A
        self.np = np_
        self.order = np.ascontiguousarray(order, np.int32)
        self.sizes = np.ascontiguousarray(sizes, np.int32)
        self._h = C.c_void_p()
        rc = lib().oracle_create(A.n, A.nnz, _ip(A.rowptr), _ip(A.colidx), _dp(A.data), np_,
                                 _ip(self.order), _ip(self.sizes), C.byref(self._h))
        if rc != 0:
            raise RuntimeError(f"oracle_create failed: {rc}")

    def close(self):
        if self._h and self._h.value:
            lib().oracle_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def factor_solve(self, b, nthreads: int = 1) -> np.ndarray:
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        rc = lib().oracle_factor_solve(self._h, _dp(b), _dp(x), nthreads)
        if rc != 0:
            raise RuntimeError(f"oracle_factor_solve failed: {rc}")
        return x

    def factor_solve_columns(self, B, nthreads: int = 1) -> np.ndarray:
        """The reference path has one right-hand side: each column of B is
        solved by its own factor_solve."""
        B = np.asarray(B, np.float64)
        if B.ndim == 1:
            return self.factor_solve(B, nthreads)
        return np.stack([self.factor_solve(B[:, j], nthreads) for j in range(B.shape[1])], axis=1)

    def assemble_only(self, b=None, nthreads: int = 1):
        """Leaves and the assembly of the separator blocks only (stage 1)."""
        b = np.zeros(self.A.n) if b is None else np.ascontiguousarray(b, np.float64)
        rc = lib().oracle_assemble_only(self._h, _dp(b), nthreads)
        if rc != 0:
            raise RuntimeError(f"oracle_assemble_only failed: {rc}")

    def sep_block(self, i: int, j: int) -> np.ndarray:
        """Dense block (i, j) of the separator storage, s_i x s_j."""
        _, sz = self.blocks()
        si, sj = int(sz[i - 1]), int(sz[j - 1])
        out = np.zeros(si * sj, np.float64)
        lib().oracle_sep_block(self._h, i, j, _dp(out))
        return out.reshape(si, sj, order="F")

    def leaf_factor(self, nthreads: int = 1):
        lib().oracle_leaf_factor(self._h, nthreads)

    def leaf_init(self):
        """A's leaf values scattered into the LU layout, nothing factored."""
        lib().oracle_leaf_init(self._h)

    def times(self):
        t = (C.c_double * 4)()
        lib().oracle_last_times(t)
        return list(t)

    def blocks(self):
        nb = 2 * self.np - 1
        bs, sz = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
        lib().oracle_blocks(self._h, _ip(bs), _ip(sz))
        return bs, sz

    def permuted(self):
        rp = np.zeros(self.A.n + 1, np.int32)
        ci = np.zeros(self.A.nnz, np.int32)
        pm = np.zeros(self.A.nnz, np.int32)
        lib().oracle_permuted(self._h, _ip(rp), _ip(ci), _ip(pm))
        return rp, ci, pm

    def leaf_sizes(self, rank):
        v = [C.c_int() for _ in range(4)]
        lib().oracle_leaf_sizes(self._h, rank, *[C.byref(x) for x in v])
        return [x.value for x in v]

    def leaf_struct(self, rank):
        ls, cn, cnz, lnz = self.leaf_sizes(rank)
        rp, ci = np.zeros(cn + 1, np.int32), np.zeros(lnz, np.int32)
        dg, bias = np.zeros(cn, np.int32), np.zeros(ls, np.int32)
        cm = np.zeros(cnz, np.int32)
        lib().oracle_leaf_struct(self._h, rank, _ip(rp), _ip(ci), _ip(dg), _ip(bias), _ip(cm))
        return rp, ci, dg, bias, cm

    def leaf_values(self, rank):
        lnz = self.leaf_sizes(rank)[3]
        v = np.zeros(lnz, np.float64)
        lib().oracle_leaf_values(self._h, rank, _dp(v))
        return v


def dense_backend() -> int:
    return lib().oracle_dense_backend()


def force_fallback(on: bool):
    lib().oracle_force_fallback(1 if on else 0)
