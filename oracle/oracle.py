"""ctypes wrapper of the CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product (cup2d_amd) never imports it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

BLEN = 64
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_LL = ctypes.c_longlong
_I = ctypes.c_int
_D = ctypes.c_double


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    lib.orc_last_error.restype = ctypes.c_char_p
    lib.orc_pinv.argtypes = [_P]
    lib.orc_assemble.restype = _LL
    lib.orc_assemble.argtypes = [_P, _P, _P, _LL, _I, _I, _I, _I, _I, _P, _P, _P, _LL, _P]
    lib.orc_h2.argtypes = [_P, _LL, _I, _I, _D, _P]
    lib.orc_spmv.argtypes = [_LL, _P, _P, _P, _P, _P, _I, _LL, _P]
    lib.orc_true_residual.argtypes = [_LL, _P, _P, _P, _P, _P, _I, _LL, _P, _P]
    lib.orc_mean_remove.restype = _D
    lib.orc_mean_remove.argtypes = [_LL, _P, _P]
    lib.orc_bicgstab.argtypes = [_LL, _P, _P, _P, _P, _P, _P, _I, _LL, _P, _I, _D, _D, _I, _I, _D,
                                 _D, _P]
    lib.orc_block_cholesky_apply.argtypes = [_LL, _P, _P, _P, _I, _I, _P, _P]
    lib.orc_pipelined_bicgstab.argtypes = [_LL, _P, _P, _P, _P, _P, _P, _I, _I, _P, _P, _I, _LL,
                                           _P, _I, _D, _D, _I, _P]
    lib.orc_num_threads.restype = _I
    lib.orc_set_num_threads.argtypes = [_I]
    _lib = lib
    return lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


@dataclass
class Matrix:
    """FP64 CSR of the assembled operator plus its mean-constraint row."""

    row_ptr: np.ndarray  # int64, n+1
    col: np.ndarray      # int32
    val: np.ndarray      # float64
    mean_row: int        # -1 when no row is replaced

    @property
    def n(self):
        return self.row_ptr.shape[0] - 1

    @property
    def nnz(self):
        return self.col.shape[0]

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.row_ptr), shape=(self.n, self.n))


def num_threads():
    return _load().orc_num_threads()


def set_num_threads(n):
    _load().orc_set_num_threads(int(n))


def pinv() -> np.ndarray:
    """P_inv = -K^-1 (ExpAMRSolver.cpp:37-95), row-major 64x64."""
    P = np.empty((BLEN, BLEN), np.float64)
    _load().orc_pinv(P.ctypes.data)
    return P


def assemble(mesh, mean_constraint: int = 0) -> Matrix:
    """Restatement of ExpAMRSolver::getMat (ExpAMRSolver.cpp:180-313)."""
    lib = _load()
    lev, bx, by = _c(mesh.level, np.int32), _c(mesh.bx, np.int32), _c(mesh.by, np.int32)
    nb = lev.shape[0]
    n = nb * BLEN
    mr = np.zeros(1, np.int64)
    cap = 6 * n + 64
    for _ in range(2):
        rp = np.empty(n + 1, np.int64)
        col = np.empty(cap, np.int32)
        val = np.empty(cap, np.float64)
        nnz = lib.orc_assemble(lev.ctypes.data, bx.ctypes.data, by.ctypes.data, nb, mesh.bpdx,
                               mesh.bpdy, int(mesh.periodic_x), int(mesh.periodic_y),
                               int(mean_constraint), rp.ctypes.data, col.ctypes.data,
                               val.ctypes.data, cap, mr.ctypes.data)
        if nnz == -(2 ** 63):
            raise RuntimeError(lib.orc_last_error().decode())
        if nnz >= 0:
            return Matrix(rp, col[:nnz].copy(), val[:nnz].copy(), int(mr[0]))
        cap = -nnz
    raise RuntimeError("assemble: capacity negotiation failed")


def h2(mesh) -> np.ndarray:
    lev = _c(mesh.level, np.int32)
    out = np.empty(lev.shape[0], np.float64)
    _load().orc_h2(lev.ctypes.data, lev.shape[0], mesh.bpdx, mesh.bpdy, float(mesh.extent),
                   out.ctypes.data)
    return out


def spmv(A: Matrix, x, mean_mode=0, h2v=None) -> np.ndarray:
    x = _c(x, np.float64)
    y = np.empty(A.n, np.float64)
    h2v = _c(h2v if h2v is not None else np.zeros(A.n // BLEN), np.float64)
    _load().orc_spmv(A.n, A.row_ptr.ctypes.data, A.col.ctypes.data, A.val.ctypes.data,
                     x.ctypes.data, y.ctypes.data, int(mean_mode), A.mean_row, h2v.ctypes.data)
    return y


def true_residual(A: Matrix, b, x, mean_mode=0, h2v=None):
    """Returns dict(r2, rinf, b2, binf) of r = b - A x in long double."""
    b = _c(b, np.float64)
    x = _c(x, np.float64)
    h2v = _c(h2v if h2v is not None else np.zeros(A.n // BLEN), np.float64)
    out = np.empty(4, np.float64)
    _load().orc_true_residual(A.n, A.row_ptr.ctypes.data, A.col.ctypes.data, A.val.ctypes.data,
                              b.ctypes.data, x.ctypes.data, int(mean_mode), A.mean_row,
                              h2v.ctypes.data, out.ctypes.data)
    return {"r2": out[0], "rinf": out[1], "b2": out[2], "binf": out[3],
            "rel2": out[0] / out[2] if out[2] > 0 else 0.0,
            "relinf": out[1] / out[3] if out[3] > 0 else 0.0}


def mean_remove(h2v, x) -> np.ndarray:
    """x - sum(h2 x)/sum(h2) (ExpAMRSolver.cpp:383-415)."""
    x = np.array(x, dtype=np.float64, copy=True)
    h2v = _c(h2v, np.float64)
    _load().orc_mean_remove(h2v.shape[0], h2v.ctypes.data, x.ctypes.data)
    return x


def _stats(st):
    return {"iterations": int(st[0]), "restarts": int(st[1]), "converged": bool(st[2]),
            "err_init": float(st[3]), "err_opt": float(st[4]), "err_last": float(st[5])}


def bicgstab(A: Matrix, b, x0=None, mean_mode=0, h2v=None, max_iter=1000, tol=0.0, rel_tol=1e-10,
             max_restarts=100, norm="inf", P=None, eps=1e-100, breakdown_tol=1e-16):
    """Reference GPU BiCGSTAB (BiCGSTAB.cu:412-620) on the CPU.  Returns
    (x_opt, stats).  eps=1e-21 reproduces the reference's denominator guard
    exactly; the default is the CPU path's 1e-100 (AMRSolver.cpp:167)."""
    b = _c(b, np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    P = _c(pinv() if P is None else P, np.float64)
    h2v = _c(h2v if h2v is not None else np.zeros(A.n // BLEN), np.float64)
    st = np.zeros(8, np.float64)
    _load().orc_bicgstab(A.n, A.row_ptr.ctypes.data, A.col.ctypes.data, A.val.ctypes.data,
                         P.ctypes.data, b.ctypes.data, x.ctypes.data, int(mean_mode), A.mean_row,
                         h2v.ctypes.data, int(max_iter), float(tol), float(rel_tol),
                         int(max_restarts), 0 if norm == "inf" else 1, float(eps),
                         float(breakdown_tol), st.ctypes.data)
    return x, _stats(st)


def block_cholesky_apply(mesh, v) -> np.ndarray:
    """AMRSolver::getZ on every block (AMRSolver.cpp:13-74)."""
    v = _c(v, np.float64)
    out = np.empty_like(v)
    lev, bx, by = _c(mesh.level, np.int32), _c(mesh.bx, np.int32), _c(mesh.by, np.int32)
    _load().orc_block_cholesky_apply(lev.shape[0], lev.ctypes.data, bx.ctypes.data,
                                     by.ctypes.data, mesh.bpdx, mesh.bpdy, v.ctypes.data,
                                     out.ctypes.data)
    return out


def pipelined_bicgstab(A: Matrix, mesh, b, x0=None, mean_mode=0, h2v=None, max_iter=1000, tol=0.0,
                       rel_tol=1e-10, max_restarts=100):
    """Reference CPU pipelined BiCGSTAB (AMRSolver.cpp:149-490) on the
    explicit operator.  Returns (x, stats)."""
    b = _c(b, np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    lev, bx, by = _c(mesh.level, np.int32), _c(mesh.bx, np.int32), _c(mesh.by, np.int32)
    h2v = _c(h2v if h2v is not None else np.zeros(A.n // BLEN), np.float64)
    st = np.zeros(8, np.float64)
    _load().orc_pipelined_bicgstab(A.n, A.row_ptr.ctypes.data, A.col.ctypes.data,
                                   A.val.ctypes.data, lev.ctypes.data, bx.ctypes.data,
                                   by.ctypes.data, mesh.bpdx, mesh.bpdy, b.ctypes.data,
                                   x.ctypes.data, int(mean_mode), A.mean_row, h2v.ctypes.data,
                                   int(max_iter), float(tol), float(rel_tol), int(max_restarts),
                                   st.ctypes.data)
    return x, _stats(st)
