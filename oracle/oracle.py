"""ctypes wrapper of the CPU oracle (oracle/fereg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's cpu_baseline leg,
always as the checker. The product path (fereg_amd) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfereg_oracle.so")
_lib = None

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
f64p = ctypes.POINTER(ctypes.c_double)


class OrInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("nterms", ctypes.c_uint64), ("g1", ctypes.c_uint64),
                ("ncells", ctypes.c_uint64), ("nb", ctypes.c_uint64), ("m", ctypes.c_uint64),
                ("m_small", ctypes.c_uint64), ("small_work", ctypes.c_uint64), ("fe1", ctypes.c_int64),
                ("norm_fe1", ctypes.c_int64), ("has_cell_obs_off", ctypes.c_int)]


def build():
    """Compiles the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_build.argtypes = [ctypes.POINTER(u32p), u64p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                               f64p, ctypes.POINTER(ctypes.c_void_p)]
        L.or_free.argtypes = [ctypes.c_void_p]
        L.or_threads.restype = ctypes.c_int
        L.or_get_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrInfo)]
        L.or_get_blocks.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), u64p, u64p,
                                    ctypes.POINTER(ctypes.c_int32)]
        L.or_get_structure.argtypes = [ctypes.c_void_p, u32p, u32p, u32p, u32p]
        L.or_get_level_cells.argtypes = [ctypes.c_void_p, ctypes.c_uint64, u32p, u32p]
        L.or_get_weight_state.argtypes = [ctypes.c_void_p, f64p, f64p, f64p, ctypes.POINTER(ctypes.c_uint8), f64p]
        L.or_set_weights.argtypes = [ctypes.c_void_p, f64p]
        L.or_apply.argtypes = [ctypes.c_void_p, f64p, ctypes.c_uint64, f64p]
        L.or_rhs.argtypes = [ctypes.c_void_p, f64p, ctypes.c_uint64, f64p, f64p, f64p, f64p]
        L.or_residualize_rows.argtypes = [ctypes.c_void_p, f64p, ctypes.c_uint64, f64p, f64p, f64p]
        L.or_small_apply.argtypes = [ctypes.c_void_p, f64p, ctypes.c_uint64, f64p]
        L.or_residualize.argtypes = [ctypes.c_void_p, f64p, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64,
                                     f64p, f64p, u64p, ctypes.POINTER(ctypes.c_int32)]
        _lib = L
    return _lib


def threads():
    """Threads of the oracle's OpenMP loops."""
    return lib().or_threads()


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class OracleSolver:
    """The oracle's absorber: structure, weight state and residualization on the CPU."""

    def __init__(self, codes, g, weights=None, small_limit=4096):
        L = lib()
        self._codes = [np.ascontiguousarray(c, dtype=np.uint32) for c in codes]
        n = len(self._codes[0]) if self._codes else (len(weights) if weights is not None else 0)
        self._g = np.ascontiguousarray(g, dtype=np.uint64)
        arr = (u32p * max(1, len(self._codes)))(*[_p(c, u32p) for c in self._codes])
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        h = ctypes.c_void_p()
        rc = L.or_build(arr, _p(self._g, u64p), len(self._codes), n, small_limit, _p(w, f64p), ctypes.byref(h))
        if rc != 0:
            raise ValueError(f"oracle build failed (rc {rc})")
        self.h = h
        info = OrInfo()
        L.or_get_info(h, ctypes.byref(info))
        self.info = info

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_free(self.h)
            self.h = None

    @property
    def n(self):
        return self.info.n

    def blocks(self):
        nb = self.info.nb
        term = np.zeros(nb, np.int64)
        g = np.zeros(nb, np.uint64)
        off = np.zeros(nb, np.uint64)
        small = np.zeros(nb, np.int32)
        lib().or_get_blocks(self.h, _p(term, ctypes.POINTER(ctypes.c_int64)), _p(g, u64p), _p(off, u64p),
                            _p(small, ctypes.POINTER(ctypes.c_int32)))
        return term, g, off, small

    def structure(self):
        i = self.info
        pi = np.zeros(i.n, np.uint32)
        seg = np.zeros(i.g1 + 1, np.uint32)
        cc = np.zeros(max(1, i.ncells * i.nb), np.uint32)
        coo = np.zeros(i.ncells + 1, np.uint32) if i.has_cell_obs_off else None
        lib().or_get_structure(self.h, _p(pi, u32p), _p(seg, u32p), _p(cc, u32p), _p(coo, u32p))
        return {"pi": pi, "seg_off": seg, "cell_codes": cc[: i.ncells * i.nb], "cell_obs_off": coo}

    def level_cells(self, b):
        _, g, _, _ = self.blocks()
        perm = np.zeros(max(1, self.info.ncells), np.uint32)
        off = np.zeros(int(g[b]) + 1, np.uint32)
        lib().or_get_level_cells(self.h, b, _p(perm, u32p), _p(off, u32p))
        return perm[: self.info.ncells], off

    def weight_state(self):
        i = self.info
        n1 = np.zeros(i.g1)
        nr = np.zeros(i.m)
        dg = np.zeros(i.m)
        red = np.zeros(i.m, np.uint8)
        sw = ctypes.c_double()
        lib().or_get_weight_state(self.h, _p(n1, f64p), _p(nr, f64p), _p(dg, f64p),
                                  _p(red, ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(sw))
        return {"n1": n1, "nr": nr, "diag": dg, "redundant": red.astype(bool), "sw": sw.value}

    def set_weights(self, w_rows):
        w = None if w_rows is None else np.ascontiguousarray(w_rows, dtype=np.float64)
        rc = lib().or_set_weights(self.h, _p(w, f64p))
        if rc != 0:
            raise ValueError("weights must be positive and finite")

    def apply(self, p, r):
        p = np.ascontiguousarray(p, dtype=np.float64)
        y = np.zeros(self.info.m * r)
        lib().or_apply(self.h, _p(p, f64p), r, _p(y, f64p))
        return y

    def rhs(self, rows, r):
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        i = self.info
        rhs = np.zeros(i.m * r)
        b1 = np.zeros(i.g1 * r)
        wsq = np.zeros(r)
        wsum = np.zeros(r)
        lib().or_rhs(self.h, _p(rows, f64p), r, _p(rhs, f64p), _p(b1, f64p), _p(wsq, f64p), _p(wsum, f64p))
        return rhs, b1, wsq, wsum

    def small_apply(self, rin, r):
        rin = np.ascontiguousarray(rin, dtype=np.float64)
        z = np.zeros(self.info.m_small * r)
        lib().or_small_apply(self.h, _p(rin, f64p), r, _p(z, f64p))
        return z

    def residualize(self, z, r, tol=1e-8, maxiter=16000, warm=None, want_coefs=False):
        """Residualizes the n x r row-order block z in place; returns (iterations, exact, coefs)."""
        assert z.dtype == np.float64 and z.flags.c_contiguous
        i = self.info
        coefs = np.zeros((i.m + i.g1) * r) if want_coefs else None
        wm = None if warm is None else np.ascontiguousarray(warm, dtype=np.float64)
        it = ctypes.c_uint64()
        ex = ctypes.c_int32()
        rc = lib().or_residualize(self.h, _p(z, f64p), r, tol, maxiter, _p(wm, f64p), _p(coefs, f64p),
                                  ctypes.byref(it), ctypes.byref(ex))
        if rc == 430:
            raise RuntimeError(f"the solver did not converge within {maxiter} iterations")
        if rc != 0:
            raise ValueError(f"oracle residualize failed (rc {rc})")
        return it.value, bool(ex.value), coefs
