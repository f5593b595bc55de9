"""The Absorber boundary (crates/fereg-core/src/absorber.rs:62-97) over the HIP library.

Blocks are n x R row-major in `row_order` (the solver's observation order). Host blocks are
numpy float64 arrays; device blocks are torch tensors on the absorber's GPU (HBM-resident, no
copies across PCIe).
"""
import ctypes
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib as L


@dataclass
class SolveOpts:
    """absorber.rs:18-24 (+ the precision of this solve: None = the absorber's)."""
    tol: float = 1e-8
    maxiter: int = 16000
    want_coefs: bool = False
    precision: Optional[int] = None

    def c(self):
        return L.SolveOptsC(self.tol, self.maxiter, int(self.want_coefs),
                            -1 if self.precision is None else int(self.precision))


@dataclass
class FeCoefs:
    """absorber.rs:39-60: p x r row-major, p = m + g1 ([x; alpha1], solve.rs:7-9)."""
    p: int
    r: int
    data: np.ndarray

    def col(self, j):
        return self.data.reshape(self.p, self.r)[:, j].copy()

    def combine(self, c):
        c = np.asarray(c, float)
        assert len(c) == self.r, "combine: one weight per column"
        return FeCoefs(self.p, 1, self.data.reshape(self.p, self.r) @ c)


@dataclass
class SolveInfo:
    """absorber.rs:27-37."""
    iterations: int
    escalated: bool
    coefs: Optional[FeCoefs]
    exact: bool
    refinements: int = 0
    seconds: float = 0.0


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _dev_ready(*ts):
    """Contiguous forms of the torch device tensors handed to a `_dev` entry point, with torch's
    current stream drained. The library runs on the absorber's own stream, which does not wait for
    torch's: a copy into an input, a `.contiguous()` copy or a `torch.zeros` output queued there
    could otherwise still be running when the library reads or writes the memory. The caller keeps
    the returned tensors alive until the call returns (the library is synchronous to the host)."""
    import torch
    out = [None if t is None else t.contiguous() for t in ts]
    for d in {t.device for t in out if t is not None and t.is_cuda}:
        torch.cuda.current_stream(d).synchronize()
    return out


def _on_gpu(t, what):
    if not t.is_cuda:
        raise L.FeregError(L.FH_INVALID, f"{what}: a torch tensor must live on the GPU (device pointer)")


def _dev_f64(t, what):
    """A float64 device tensor of the values of `t` (a float tensor on the GPU)."""
    import torch
    if t is None:
        return None
    _on_gpu(t, what)
    if not t.is_floating_point():
        raise L.FeregError(L.FH_INVALID, f"{what}: floating-point values expected, got {t.dtype}")
    return t.to(torch.float64)


def _dev_codes(t, what):
    """A 32-bit device tensor of integer codes (int32 and uint32 pass as they are; other integer
    types are converted after checking that every code lies in 0 .. 2^31 - 1)."""
    import torch
    if t is None:
        return None
    _on_gpu(t, what)
    if t.dtype in (torch.int32, torch.uint32):
        return t
    if t.is_floating_point() or t.dtype == torch.bool or t.is_complex():
        raise L.FeregError(L.FH_INVALID, f"{what}: integer codes expected, got {t.dtype}")
    if t.numel() and (int(t.min()) < 0 or int(t.max()) >= 2 ** 31):
        raise L.FeregError(L.FH_INVALID, f"{what}: codes must lie in 0 .. 2^31 - 1")
    return t.to(torch.int32)


class Absorber:
    """solve::build / FeSolver::new (solve.rs:338-465) on the GPU.

    codes: list of per-term kept-order codes (uint32 numpy arrays, or torch uint32/int32 device
    tensors), g: levels per term, weights: kept-order weights or None.
    intercepts (per term, default all True) and slopes (per term, a list of kept-order slope columns,
    default none) describe varying slopes (fe#c.x, fe##c.x): any slope or a term without intercept
    makes this the varying-slopes solver (slopes.rs), which solves in FP64.
    """

    def __init__(self, codes, g, weights=None, small_limit=4096, precision=L.FH_PLAIN_FP64, bucket_log2=-1,
                 autotune=False, verbose=0, mixed_lead=-1, sparse_chol=True, joint=True, dense_limit=-1,
                 intercepts=None, slopes=None):
        lib = L.lib()
        self._keep = []
        ptr_arrays = []  # the per-term slope pointer arrays, alive through the create call
        n = None
        dev = any(_is_torch(c) for c in codes) or (weights is not None and _is_torch(weights))
        terms = (L.Term * max(1, len(codes)))()
        intercepts = [True] * len(codes) if intercepts is None else list(intercepts)
        slopes = [[] for _ in codes] if slopes is None else [list(s) for s in slopes]
        if len(intercepts) != len(codes) or len(slopes) != len(codes):
            raise L.FeregError(L.FH_INVALID, "intercepts and slopes need one entry per term")
        for q, (c, gq) in enumerate(zip(codes, g)):
            sl, xs = [], []
            for x in slopes[q]:
                if dev:
                    x = _dev_f64(x, "slopes").contiguous()
                    sl.append(ctypes.cast(x.data_ptr(), L.f64p))
                else:
                    x = np.ascontiguousarray(x, dtype=np.float64)
                    sl.append(_ptr(x, L.f64p))
                xs.append(x)
            self._keep.extend(xs)
            sarr = (L.f64p * max(1, len(sl)))(*sl)
            ptr_arrays.append(sarr)
            icpt = 1 if intercepts[q] else 0
            if dev:
                if not _is_torch(c):
                    raise L.FeregError(L.FH_INVALID, "codes: mix of torch device tensors and host arrays")
                c = _dev_codes(c, "codes").contiguous()
                self._keep.append(c)
                terms[q] = L.Term(ctypes.cast(c.data_ptr(), L.u32p), int(gq), icpt, len(sl), sarr)
                nq = c.numel()
            else:
                c = np.ascontiguousarray(c, dtype=np.uint32)
                self._keep.append(c)
                terms[q] = L.Term(_ptr(c, L.u32p), int(gq), icpt, len(sl), sarr)
                nq = len(c)
            if any((x.numel() if dev else len(x)) != nq for x in xs):
                raise L.FeregError(L.FH_INVALID, "slope column length differs from the kept sample")
            if n is None:
                n = nq
            elif n != nq:
                raise L.FeregError(L.FH_INVALID, "term codes length differs from the kept sample")
        if weights is not None:
            if n is None:
                n = len(weights)
            elif len(weights) != n:
                raise L.FeregError(L.FH_INVALID, "weight length differs from the kept sample")
        if n is None:
            raise L.FeregError(L.FH_INVALID, "no terms and no weights: the sample size is unknown")
        cfg = L.SolverConfig(small_limit, precision, bucket_log2, int(bool(autotune)), int(mixed_lead),
                             int(bool(sparse_chol)), int(bool(joint)), int(dense_limit))
        h = ctypes.c_void_p()
        if dev:
            wptr = None
            if weights is not None:
                if not _is_torch(weights):
                    raise L.FeregError(L.FH_INVALID, "weights: mix of torch device tensors and host arrays")
                weights = _dev_f64(weights, "weights").contiguous()
                self._keep.append(weights)
                wptr = weights.data_ptr()
            _dev_ready(*self._keep)
            L.check(lib.fh_absorber_create_dev(terms, len(codes), wptr, n, ctypes.byref(cfg), None, ctypes.byref(h)))
        else:
            w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
            L.check(lib.fh_absorber_create(terms, len(codes), None if w is None else _ptr(w, L.f64p), n,
                                           ctypes.byref(cfg), None, ctypes.byref(h)))
        self._h = h
        self._keep = []
        self._n = n
        self._weights = None if weights is None else (weights if dev else np.asarray(weights, float))
        info = L.StructureInfo()
        L.check(lib.fh_absorber_info(h, ctypes.byref(info)))
        self.info = info

    def plan(self, precision, r):
        """The S*p launch plan in use for this precision and column count:
        {"dense", "seg_lanes", "block_lanes", "tuned"}."""
        p = L.SpmvPlan()
        L.check(L.lib().fh_absorber_get_plan(self._h, int(precision), int(r), ctypes.byref(p)))
        return {"dense": int(p.dense), "seg_lanes": int(p.seg_lanes), "block_lanes": [int(v) for v in p.block_lanes],
                "block_mode": [int(v) for v in p.block_mode], "tuned": bool(p.tuned)}

    def set_plan(self, precision, r, plan):
        """Pins a plan (a dict as returned by plan(); lanes 0 = default, block_mode -1 = default)."""
        bl = list(plan.get("block_lanes", [])) + [0] * 9
        bm = list(plan.get("block_mode", [])) + [-1] * 9
        p = L.SpmvPlan(int(plan.get("dense", 1)), int(plan.get("seg_lanes", 0)), (ctypes.c_uint32 * 9)(*bl[:9]),
                       (ctypes.c_int32 * 9)(*bm[:9]), 1)
        L.check(L.lib().fh_absorber_set_plan(self._h, int(precision), int(r), ctypes.byref(p)))

    def time_spmv(self, r, precision, reps=20):
        """Milliseconds per S*p (HIP events) and the algorithmic bytes of one product."""
        ms, by = ctypes.c_double(0), ctypes.c_double(0)
        L.check(L.lib().fh_absorber_time_spmv(self._h, int(r), int(precision), int(reps), ctypes.byref(ms),
                                              ctypes.byref(by)))
        return ms.value, by.value

    def close(self):
        if getattr(self, "_h", None):
            L.lib().fh_absorber_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- trait
    def n(self):
        return self._n

    def row_order(self):
        p = L.lib().fh_absorber_row_order(self._h)
        if not p:
            L.check(L.FH_INTERNAL)
        return np.ctypeslib.as_array(p, shape=(self._n,)).copy()

    def weights(self):
        """Current weights in row order (None = all ones)."""
        return self._weights_rows

    _weights_rows = None

    def set_weights(self, w_rows):
        lib = L.lib()
        if w_rows is None:
            L.check(lib.fh_absorber_set_weights(self._h, None, None))
        elif _is_torch(w_rows):
            (w_rows,) = _dev_ready(_dev_f64(w_rows, "weights"))
            if w_rows.numel() != self._n:
                raise L.FeregError(L.FH_INVALID, "weights: one per row expected")
            L.check(lib.fh_absorber_set_weights_dev(self._h, w_rows.data_ptr(), None))
        else:
            w = np.ascontiguousarray(w_rows, dtype=np.float64)
            if len(w) != self._n:
                raise L.FeregError(L.FH_INVALID, "weights must be positive and finite")
            L.check(lib.fh_absorber_set_weights(self._h, _ptr(w, L.f64p), None))
        self._weights_rows = w_rows

    def coef_rows(self):
        return L.lib().fh_absorber_coef_rows(self._h)

    def residualize(self, z, r, opts=None, warm=None, ctx=None):
        """Residualizes the n x r row-major block z in place (solve.rs:1141-1311); ctx: an
        optional _lib.CallContext (interrupt and log callbacks)."""
        o = opts or SolveOpts()
        lib = L.lib()
        p = self.coef_rows()
        info = L.SolveInfoC()
        coefs = None
        if _is_torch(z):
            import torch
            _on_gpu(z, "residualize")
            if z.dtype != torch.float64 or not z.is_contiguous() or z.numel() != self._n * r:
                raise L.FeregError(L.FH_INVALID, "residualize: a contiguous float64 n x r block is residualized "
                                   "in place")
            cf = torch.zeros(p * r, dtype=torch.float64, device=z.device) if o.want_coefs else None
            wm = None
            if warm is not None:
                wm = warm.data if _is_torch(warm.data) else torch.as_tensor(warm.data, device=z.device)
                if warm.r != r or warm.p != p:
                    raise L.FeregError(L.FH_INVALID, "residualize: warm start has a different shape")
            _, wm, cf = _dev_ready(z, wm, cf)
            L.check(lib.fh_absorber_residualize_dev(self._h, z.data_ptr(), r, ctypes.byref(o.c()),
                                                    None if wm is None else wm.data_ptr(),
                                                    None if cf is None else cf.data_ptr(), ctypes.byref(info),
                                                    L.ctx_ref(ctx)))
            if cf is not None:
                coefs = FeCoefs(p, r, cf)
        else:
            if z.dtype != np.float64 or not z.flags.c_contiguous:
                raise L.FeregError(L.FH_INVALID, "residualize: a C-contiguous float64 block is residualized in place")
            if z.size != self._n * r:
                raise L.FeregError(L.FH_INVALID, "residualize: block size differs from n*R")
            cf = np.zeros(p * r) if o.want_coefs else None
            wm = None
            if warm is not None:
                if warm.r != r or warm.p != p:
                    raise L.FeregError(L.FH_INVALID, "residualize: warm start has a different shape")
                wm = np.ascontiguousarray(warm.data, dtype=np.float64)
            L.check(lib.fh_absorber_residualize(self._h, _ptr(z, L.f64p), r, ctypes.byref(o.c()),
                                                None if wm is None else _ptr(wm, L.f64p),
                                                None if cf is None else _ptr(cf, L.f64p), ctypes.byref(info),
                                                L.ctx_ref(ctx)))
            if cf is not None:
                coefs = FeCoefs(p, r, cf)
        return SolveInfo(info.iterations, bool(info.escalated), coefs, bool(info.exact), info.refinements,
                         info.seconds)

    def fit_rows(self, c: FeCoefs, col: int):
        out = np.zeros(self._n)
        data = np.ascontiguousarray(c.data, dtype=np.float64)
        L.check(L.lib().fh_absorber_fit_rows(self._h, _ptr(data, L.f64p), c.r, col, _ptr(out, L.f64p)))
        return out

    def term_values(self, c: FeCoefs, col: int) -> List[np.ndarray]:
        """Per term (per slot with varying slopes: each term's intercept, then its slopes) the
        level values (solve.rs:1319-1343, slopes.rs:2435-2464)."""
        nt = int(L.lib().fh_absorber_term_slots(self._h))
        lens = np.zeros(max(nt, 1), np.uint64)
        tot = self.coef_rows()
        out = np.zeros(max(tot, 1))
        data = np.ascontiguousarray(c.data, dtype=np.float64)
        L.check(L.lib().fh_absorber_term_values(self._h, _ptr(data, L.f64p), c.r, col, _ptr(out, L.f64p),
                                                _ptr(lens, L.u64p)))
        res, pos = [], 0
        for q in range(nt):
            k = int(lens[q])
            res.append(out[pos:pos + k].copy())
            pos += k
        return res

    # ---------------------------------------------------------------- introspection
    def blocks(self):
        nb = self.info.nb
        term = np.zeros(max(nb, 1), np.int64)
        g = np.zeros(max(nb, 1), np.uint64)
        off = np.zeros(max(nb, 1), np.uint64)
        small = np.zeros(max(nb, 1), np.int32)
        L.check(L.lib().fh_absorber_copy_blocks(self._h, _ptr(term, L.i64p), _ptr(g, L.u64p), _ptr(off, L.u64p),
                                                _ptr(small, L.i32p)))
        return term[:nb], g[:nb], off[:nb], small[:nb]

    def structure(self):
        i = self.info
        pi = np.zeros(i.n, np.uint32)
        seg = np.zeros(i.g1 + 1, np.uint32)
        cc = np.zeros(max(1, i.ncells * i.nb), np.uint32)
        coo = np.zeros(i.ncells + 1, np.uint32) if i.has_cell_obs_off else None
        L.check(L.lib().fh_absorber_copy_structure(self._h, _ptr(pi, L.u32p), _ptr(seg, L.u32p), _ptr(cc, L.u32p),
                                                   None if coo is None else _ptr(coo, L.u32p)))
        return {"pi": pi, "seg_off": seg, "cell_codes": cc[: i.ncells * i.nb], "cell_obs_off": coo}

    def level_cells(self, b):
        _, g, _, _ = self.blocks()
        perm = np.zeros(max(1, self.info.ncells), np.uint32)
        off = np.zeros(int(g[b]) + 1, np.uint32)
        L.check(L.lib().fh_absorber_copy_level_cells(self._h, b, _ptr(perm, L.u32p), _ptr(off, L.u32p)))
        return perm[: self.info.ncells], off

    def weight_state(self):
        i = self.info
        n1 = np.zeros(i.g1)
        nr = np.zeros(max(i.m, 1))
        dg = np.zeros(max(i.m, 1))
        red = np.zeros(max(i.m, 1), np.uint8)
        sw = ctypes.c_double()
        L.check(L.lib().fh_absorber_copy_weight_state(self._h, _ptr(n1, L.f64p), _ptr(nr, L.f64p), _ptr(dg, L.f64p),
                                                      red.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                                      ctypes.byref(sw)))
        return {"n1": n1, "nr": nr[: i.m], "diag": dg[: i.m], "redundant": red[: i.m].astype(bool), "sw": sw.value}

    def apply(self, p, r, precision=L.FH_PLAIN_FP64):
        """y = S p on the GPU (operator.rs:1375), host buffers m x r."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        y = np.zeros(self.info.m * r)
        L.check(L.lib().fh_absorber_apply(self._h, _ptr(p, L.f64p), r, _ptr(y, L.f64p), precision))
        return y
