"""The exact sparse LDL^T of the escalated preconditioner (precond/sparse_chol.rs:216-449,
`SparseChol::try_factor` and `solve`) through the C ABI: host compute, for callers that factor a
reduced-system matrix themselves (the solver builds its own on escalation)."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


class SparseChol:
    """LDL^T of a symmetric PSD matrix from its lower triangle in CSR form (row i: columns <= i,
    ascending), scaled to unit diagonal, pivots <= 1e-10 grounded (the solve is then a reflexive
    generalized inverse). Indices [0, nlast) are eliminated last."""

    def __init__(self, n, ptr, col, val, nlast=0):
        L = lib()
        self.n = int(n)
        ptr = np.ascontiguousarray(ptr, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        if ptr.shape != (self.n + 1,) or col.shape != val.shape or col.shape[0] != int(ptr[-1]):
            raise _lib.FeregError(_lib.FH_INVALID, "SparseChol: ptr/col/val shapes do not match")
        h = ctypes.c_void_p()
        check(L.fh_sparse_chol_factor(self.n, ptr.ctypes.data_as(_lib.u64p), col.ctypes.data_as(_lib.u32p),
                                      val.ctypes.data_as(_lib.f64p), int(nlast), ctypes.byref(h)))
        self._h = h
        ln, gr, ht = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        fl = ctypes.c_double()
        check(L.fh_sparse_chol_info(h, ctypes.byref(ln), ctypes.byref(fl), ctypes.byref(gr), ctypes.byref(ht)))
        self.len, self.flops, self.grounded, self.height = int(ln.value), float(fl.value), int(gr.value), int(ht.value)

    @classmethod
    def from_dense(cls, a, nlast=0):
        a = np.asarray(a, dtype=np.float64)
        lo = np.tril(a)
        rows, cols = np.nonzero(lo)
        ptr = np.zeros(a.shape[0] + 1, np.uint64)
        np.add.at(ptr, rows + 1, 1)
        return cls(a.shape[0], np.cumsum(ptr), cols, lo[rows, cols], nlast)

    def solve(self, b):
        """x = A^+ b for b of shape (n,) or (n, nrhs)."""
        x = np.array(b, dtype=np.float64, order="C", copy=True)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        if x.shape[0] != self.n:
            raise _lib.FeregError(_lib.FH_INVALID, "SparseChol.solve: b has the wrong number of rows")
        check(lib().fh_sparse_chol_solve(self._h, x.ctypes.data_as(_lib.f64p), nrhs))
        return x

    def inverse_entries(self, rows, cols):
        """Entries of the generalized inverse at (rows[q], cols[q]) by selected inversion (exact on
        every pair that shares a structural entry of the matrix; 0 elsewhere)."""
        rows = np.ascontiguousarray(rows, dtype=np.uint32)
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        out = np.zeros(len(rows))
        check(lib().fh_sparse_chol_inverse_entries(self._h, len(rows), rows.ctypes.data_as(_lib.u32p),
                                                   cols.ctypes.data_as(_lib.u32p), out.ctypes.data_as(_lib.f64p)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().fh_sparse_chol_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
