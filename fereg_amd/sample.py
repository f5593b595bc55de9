"""Factor encoding and the estimation sample over the C ABI (factor.rs, sample.rs, fit/input.rs).

`factorize`, `combine`, `restrict` and `drop_singletons` mirror the reference's factor and
sample functions; `Sample` builds the estimation sample of the fit layer (missing values, zero
weights, interactions, singleton and fe-separation removal, dense codes over the kept rows), all
on the GPU. Inputs are numpy arrays (host entry points) or CUDA tensors (device entry points).
"""
import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from .absorber import _is_torch, _ptr


@dataclass
class Codes:
    """A pre-coded factor column (FactorCol::Codes): codes base .. base + nlevels - 1 are levels,
    any other code is missing."""
    codes: object
    nlevels: int
    base: int = 0


def _col_kind(v):
    if isinstance(v, Codes):
        return L.FH_COL["codes"]
    if _is_torch(v):
        import torch
        kinds = {torch.float64: "f64", torch.int32: "i32", torch.int64: "i64"}
        if v.dtype not in kinds:
            raise L.FeregError(L.FH_INVALID, f"factor column of dtype {v.dtype}: float64, int32 or int64 expected")
        return L.FH_COL[kinds[v.dtype]]
    a = np.asarray(v)
    if a.dtype == np.int32:
        return L.FH_COL["i32"]
    if np.issubdtype(a.dtype, np.integer):
        return L.FH_COL["i64"]
    return L.FH_COL["f64"]


class _Held:
    """Host copies (numpy) or device tensors kept alive for the duration of a call."""

    def __init__(self):
        self.items = []

    def keep(self, x):
        self.items.append(x)
        return x


def _factor_col(v, held, dev):
    kind = _col_kind(v)
    if isinstance(v, Codes):
        data = v.codes
        if dev:
            import torch
            data = held.keep(data.to(torch.int32).contiguous()) if data.dtype != torch.int32 else held.keep(data.contiguous())
            ptr = data.data_ptr()
        else:
            a = np.asarray(data)
            if a.size and (a.min() < 0 or a.max() >= 2 ** 32):
                raise L.FeregError(L.FH_INVALID, "pre-coded factor: codes in 0 .. 2^32 - 1 expected")
            data = held.keep(np.ascontiguousarray(a, dtype=np.uint32))
            ptr = data.ctypes.data
        return L.FactorCol(kind, 0, ptr, int(v.base), int(v.nlevels))
    if dev:
        t = held.keep(v.contiguous())
        if not t.is_cuda:
            raise L.FeregError(L.FH_INVALID, "factor column: a CUDA tensor expected")
        return L.FactorCol(kind, 0, t.data_ptr(), 0, 0)
    dt = {L.FH_COL["f64"]: np.float64, L.FH_COL["i32"]: np.int32, L.FH_COL["i64"]: np.int64}[kind]
    a = held.keep(np.ascontiguousarray(v, dtype=dt))
    return L.FactorCol(kind, 0, a.ctypes.data, 0, 0)


def _len(v):
    if isinstance(v, Codes):
        v = v.codes
    return int(v.shape[0]) if hasattr(v, "shape") else len(v)


def _is_dev(v):
    return _is_torch(v.codes if isinstance(v, Codes) else v)


def factorize(values, strict=False):
    """factor::factorize_col (factor.rs:307-345) / factorize (strict, factor.rs:44-76): dense
    codes in ascending value order; missing entries share level 0 when present. Returns
    (codes, nlevels, n_missing); codes is a uint32 array (host input) or an int32 CUDA tensor."""
    lib = L.lib()
    n = _len(values)
    held = _Held()
    dev = _is_dev(values)
    col = _factor_col(values, held, dev)
    nl, miss = ctypes.c_uint64(0), ctypes.c_uint64(0)
    if dev:
        import torch
        out = torch.empty(n, dtype=torch.int32, device=values.device if not isinstance(values, Codes) else values.codes.device)
        L.check(lib.fh_factorize_dev(ctypes.byref(col), n, int(strict), out.data_ptr(), ctypes.byref(nl), ctypes.byref(miss)))
    else:
        out = np.zeros(n, np.uint32)
        L.check(lib.fh_factorize(ctypes.byref(col), n, int(strict), _ptr(out, L.u32p), ctypes.byref(nl), ctypes.byref(miss)))
    return out, int(nl.value), int(miss.value)


def combine(parts, nlevels):
    """factor::combine (factor.rs:259-281): dense codes of the interaction, lexicographic order."""
    lib = L.lib()
    n = _len(parts[0])
    gl = np.ascontiguousarray(nlevels, dtype=np.uint64)
    nl = ctypes.c_uint64(0)
    if _is_torch(parts[0]):
        import torch
        ts = [p.to(torch.int32).contiguous() for p in parts]
        ptrs = (L.vp * len(ts))(*[t.data_ptr() for t in ts])
        out = torch.empty(n, dtype=torch.int32, device=parts[0].device)
        L.check(lib.fh_combine_dev(ptrs, _ptr(gl, L.u64p), len(ts), n, out.data_ptr(), ctypes.byref(nl)))
    else:
        arrs = [np.ascontiguousarray(p, dtype=np.uint32) for p in parts]
        ptrs = (L.vp * len(arrs))(*[a.ctypes.data for a in arrs])
        out = np.zeros(n, np.uint32)
        L.check(lib.fh_combine(ptrs, _ptr(gl, L.u64p), len(arrs), n, _ptr(out, L.u32p), ctypes.byref(nl)))
    return out, int(nl.value)


def restrict(codes, nlevels, kept):
    """factor::restrict (factor.rs:285-305): codes of the kept rows, renumbered densely."""
    lib = L.lib()
    nl = ctypes.c_uint64(0)
    if _is_torch(codes):
        import torch
        c = codes.to(torch.int32).contiguous()
        k = kept.to(torch.int32).contiguous()
        out = torch.empty(k.shape[0], dtype=torch.int32, device=codes.device)
        L.check(lib.fh_restrict_dev(c.data_ptr(), int(nlevels), c.shape[0], k.data_ptr(), k.shape[0], out.data_ptr(),
                                    ctypes.byref(nl)))
    else:
        c = np.ascontiguousarray(codes, dtype=np.uint32)
        k = np.ascontiguousarray(kept, dtype=np.uint32)
        out = np.zeros(k.shape[0], np.uint32)
        L.check(lib.fh_restrict(_ptr(c, L.u32p), int(nlevels), c.shape[0], _ptr(k, L.u32p), k.shape[0],
                                _ptr(out, L.u32p), ctypes.byref(nl)))
    return out, int(nl.value)


def drop_singletons(terms, alive=None, fw=None, y_sep=None, family="gaussian", singletons=True):
    """sample::drop_singletons (sample.rs:44-74) on host arrays: terms = [(codes, g), ...];
    returns (alive bool array, singletons dropped, separated dropped)."""
    lib = L.lib()
    n = _len(terms[0][0]) if terms else (0 if alive is None else len(alive))
    cs = [np.ascontiguousarray(c, dtype=np.uint32) for c, _ in terms]
    tarr = (L.Term * max(len(cs), 1))(*[L.Term(_ptr(c, L.u32p), int(g), 1) for c, (_, g) in zip(cs, terms)])
    al = np.ones(n, np.uint8) if alive is None else np.ascontiguousarray(alive, dtype=np.uint8).copy()
    fwa = None if fw is None else np.ascontiguousarray(fw, dtype=np.float64)
    ya = None if y_sep is None else np.ascontiguousarray(y_sep, dtype=np.float64)
    s, f = ctypes.c_uint64(0), ctypes.c_uint64(0)
    L.check(lib.fh_drop_singletons(tarr, len(cs), n, al.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                   None if fwa is None else _ptr(fwa, L.f64p), None if ya is None else _ptr(ya, L.f64p),
                                   L.FH_FAMILY[family], int(singletons), ctypes.byref(s), ctypes.byref(f)))
    return al.astype(bool), int(s.value), int(f.value)


class Sample:
    """The estimation sample (fit/input.rs:193-407; session.rs:458-490) built on the GPU.

    fe and cluster are lists of terms; a term is a factor column or a tuple of columns (an
    interaction, combined in order). A column object used in several places is one factor.
    missing=True is the fit layer (rows with a missing value or a zero weight are dropped and
    counted), missing=False the session (every row is used; a missing factor value is an error).
    """

    def __init__(self, y=None, x=(), fe=(), cluster=(), weights=None, weight_kind=None, offset=None, drop=None,
                 family="gaussian", missing=True, drop_singletons=True, separation=False, ctx=None):
        lib = L.lib()
        self._h = None
        terms = [t if isinstance(t, (tuple, list)) else (t,) for t in fe]
        cdims = [t if isinstance(t, (tuple, list)) else (t,) for t in cluster]
        cols, index = [], {}
        for t in terms + cdims:
            for c in t:
                if id(c) not in index:
                    index[id(c)] = len(cols)
                    cols.append(c)
        n = _len(y) if y is not None else (_len(cols[0]) if cols else 0)
        dev = any(_is_dev(c) for c in cols) or _is_torch(y)
        held = _Held()
        fcols = (L.FactorCol * max(len(cols), 1))(*[_factor_col(c, held, dev) for c in cols])
        for c in cols:
            if _len(c) != n:
                raise L.FeregError(L.FH_INVALID, "every factor column must have n rows")

        def lists(ts):
            f = np.array([index[id(c)] for t in ts for c in t], np.uint32)
            off = np.zeros(len(ts) + 1, np.uint32)
            off[1:] = np.cumsum([len(t) for t in ts])
            return held.keep(f), held.keep(off)

        tf, toff = lists(terms)
        cf, coff = lists(cdims)

        def fcol(v):
            if v is None:
                return None
            if dev:
                import torch
                if not _is_torch(v):
                    v = torch.as_tensor(np.asarray(v, np.float64), device="cuda")
                t = held.keep(v.to(torch.float64).contiguous())
                if t.shape[0] != n:
                    raise L.FeregError(L.FH_INVALID, "every input column must have n rows")
                return t.data_ptr()
            a = held.keep(np.ascontiguousarray(v, dtype=np.float64))
            if a.shape[0] != n:
                raise L.FeregError(L.FH_INVALID, "every input column must have n rows")
            return a.ctypes.data

        xs = list(x.T) if isinstance(x, np.ndarray) and x.ndim == 2 else list(x)
        xp = (L.vp * max(len(xs), 1))(*[fcol(c) for c in xs])
        if drop is not None:
            if dev:
                import torch
                dp = held.keep(torch.as_tensor(drop, device="cuda").to(torch.uint8).contiguous()).data_ptr()
            else:
                dp = held.keep(np.ascontiguousarray(drop, dtype=np.uint8)).ctypes.data
        else:
            dp = None
        spec = L.SampleSpec(n, fcol(y), xp, len(xs), fcol(weights), fcol(offset), dp, fcols, len(cols),
                            _ptr(tf, L.u32p), _ptr(toff, L.u32p), len(terms), _ptr(cf, L.u32p), _ptr(coff, L.u32p),
                            len(cdims), L.FH_WEIGHT[weight_kind], L.FH_FAMILY[family], int(missing),
                            int(drop_singletons), int(separation), 0)
        h = L.vp()
        fn = lib.fh_sample_build_dev if dev else lib.fh_sample_build
        L.check(fn(ctypes.byref(spec), L.ctx_ref(ctx), ctypes.byref(h)))
        self._h = h
        self.dev = dev
        self.nterms = len(terms)
        self.nclusters = len(cdims)
        cnt = L.SampleCounts()
        L.check(lib.fh_sample_get_counts(h, ctypes.byref(cnt)))
        self.counts = {k: int(getattr(cnt, k)) for k, _ in L.SampleCounts._fields_}
        self.nobs = self.counts["nobs"]
        self.n_input = self.counts["n_input"]

    def kept(self):
        """Kept input rows, ascending (uint32)."""
        out = np.zeros(self.nobs, np.uint32)
        L.check(L.lib().fh_sample_kept(self._h, _ptr(out, L.u32p)))
        return out

    def term(self, q):
        """(codes over the kept rows, levels) of absorbed term q, on the host."""
        out = np.zeros(self.nobs, np.uint32)
        g = ctypes.c_uint64(0)
        L.check(L.lib().fh_sample_term(self._h, q, _ptr(out, L.u32p), ctypes.byref(g)))
        return out, int(g.value)

    def cluster(self, q):
        out = np.zeros(self.nobs, np.uint32)
        g = ctypes.c_uint64(0)
        L.check(L.lib().fh_sample_cluster(self._h, q, _ptr(out, L.u32p), ctypes.byref(g)))
        return out, int(g.value)

    def term_dev(self, q):
        """(device pointer, levels) of term q's codes; valid while this sample lives."""
        g = ctypes.c_uint64(0)
        p = L.lib().fh_sample_term_dev(self._h, q, ctypes.byref(g))
        return p, int(g.value)

    def gather(self, col):
        """An input column (n) on the kept rows (nobs), on the host."""
        a = np.ascontiguousarray(col, dtype=np.float64)
        if a.shape[0] != self.n_input:
            raise L.FeregError(L.FH_INVALID, "the column must have the input's rows")
        out = np.zeros(self.nobs)
        L.check(L.lib().fh_sample_gather(self._h, _ptr(a, L.f64p), _ptr(out, L.f64p)))
        return out

    def close(self):
        if self._h:
            L.lib().fh_sample_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
