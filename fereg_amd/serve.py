"""Serving over the library's continuous-batching engine (fh_server_* in include/fereg_hip.h).

`Server` admits regressions (OLS or Poisson PML) into a FIFO; the native engine runs up to
`slots` solves at once, each on its own stream, admitted by HBM footprint, and packs consecutive
small OLS requests into one solve (one absorber, one multi-RHS PCG over all of them). `poisson_stream`
draws the mixed request stream (log-uniform n, 2 or 3 FEs, K uniform, OLS/PPML mix) and
`run_stream` replays it with Poisson arrivals at a fixed rate and reports latency percentiles
and throughput.
"""
import ctypes
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import synth
from .absorber import _ptr


@dataclass
class Reply:
    coef: np.ndarray
    vcov: np.ndarray
    kept: np.ndarray
    rc: int
    queued_s: float
    service_s: float
    total_s: float
    iterations: int
    start_s: float = 0.0  # admission and completion, seconds since the server was created
    end_s: float = 0.0
    packed: int = 1  # requests in the solve this one ran in


class _Pending:
    """Keeps the request's arrays alive until its reply is collected."""

    def __init__(self, req, keep, beta, vcov, kept):
        self.req, self.keep, self.beta, self.vcov, self.kept = req, keep, beta, vcov, kept


class Server:
    def __init__(self, slots=8, hbm_budget=0, precision=L.FH_MIXED, pack=32):
        """pack: most consecutive small OLS requests (iid / hetero, no weights) solved as one packed
        absorber (1: every request alone)."""
        cfg = L.ServerConfig(int(slots), int(hbm_budget), int(precision), int(pack))
        h = ctypes.c_void_p()
        L.check(L.lib().fh_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._pending = {}
        self._lock = threading.Lock()

    def submit(self, model, codes, g, y, X, vcov="iid", cluster=None, tol=1e-8, maxiter=100):
        """Queues one regression (host arrays in kept order); returns its ticket."""
        codes = [np.ascontiguousarray(c, dtype=np.uint32) for c in codes]
        y = np.ascontiguousarray(y, dtype=np.float64)
        X = np.ascontiguousarray(X, dtype=np.float64)
        if X.ndim == 1:
            X = X[:, None]
        n, K = X.shape
        terms = (L.Term * len(codes))(*[L.Term(_ptr(c, L.u32p), int(gg), 1) for c, gg in zip(codes, g)])
        nest = np.zeros(max(len(codes), 1), np.int32)
        cl = None if cluster is None else np.ascontiguousarray(cluster, dtype=np.uint32)
        beta, V, kept = np.zeros(K), np.zeros(K * K), np.zeros(K, np.int32)
        req = L.Request(L.FH_MODEL[model], L.FH_VCOV[vcov], n, K, len(codes), terms, _ptr(y, L.f64p), _ptr(X, L.f64p),
                        None if cl is None else _ptr(cl, L.u32p), _ptr(nest, L.i32p), L.Ssc(1, 2, 1, 0), float(tol),
                        int(maxiter), _ptr(beta, L.f64p), _ptr(V, L.f64p), _ptr(kept, L.i32p))
        t = ctypes.c_uint64()
        L.check(L.lib().fh_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)))
        with self._lock:
            self._pending[t.value] = _Pending(req, (codes, y, X, cl, terms, nest), beta, V, kept)
        return t.value

    def wait(self, ticket):
        r = L.Response()
        L.check(L.lib().fh_server_wait(self._h, ctypes.c_uint64(ticket), ctypes.byref(r)))
        with self._lock:
            p = self._pending.pop(ticket)
        K = p.beta.size
        return Reply(p.beta, p.vcov.reshape(K, K), p.kept.astype(bool), r.rc, r.queued_s, r.service_s, r.total_s,
                     int(r.iterations), r.start_s, r.end_s, int(r.packed))

    def close(self):
        if getattr(self, "_h", None):
            L.lib().fh_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_stream(count, seed=7, n_lo=1e4, n_hi=2e6, ppml_share=0.2):
    """The request mix: n log-uniform in [n_lo, n_hi], T = 10 periods, F = I/10 firms, 2 or 3
    FEs (50/50), K uniform in 1..10, OLS / Poisson PML."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))) // 10 * 10
        specs.append({"n": max(n, 100), "q": int(rng.integers(2, 4)), "K": int(rng.integers(1, 11)),
                      "model": "ppml" if rng.uniform() < ppml_share else "ols", "seed": int(seed * 100003 + i)})
    return specs


def make_request(spec):
    """Synthetic panel of one request: codes, level counts, y, X."""
    d = synth.panel(spec["n"], T=10, F=max(1, spec["n"] // 100), design="simple", K=spec["K"],
                    model="poisson" if spec["model"] == "ppml" else "ols", seed=spec["seed"])
    if spec["q"] == 3:
        codes, g = [d["indiv"], d["firm"], d["year"]], d["g"]
    else:
        codes, g = [d["indiv"], d["year"]], [d["g"][0], d["g"][2]]
    return codes, g, d["y"], d["X"]


def run_stream(server, specs, rate, gen_threads=8, ahead=16, progress=None):
    """Submits the requests at Poisson arrival times (mean `rate` per second) while a collector
    thread takes the replies in submission order (releasing each request's arrays). Returns the
    replies, the wall time and the worst lateness of a submission behind its arrival time.
    progress = (every, fn) calls fn(done) every `every` replies."""
    import queue as _q
    rng = np.random.default_rng(1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(specs)))
    pool = ThreadPoolExecutor(gen_threads)
    futs = [None] * len(specs)
    for i in range(min(ahead, len(specs))):
        futs[i] = pool.submit(make_request, specs[i])
    tq = _q.Queue()
    replies = []

    def collect():
        while True:
            t = tq.get()
            if t is None:
                return
            replies.append(server.wait(t))
            if progress and len(replies) % progress[0] == 0:
                progress[1](len(replies))

    th = threading.Thread(target=collect)
    th.start()
    late = 0.0
    t0 = time.perf_counter()
    for i, spec in enumerate(specs):
        codes, g, y, X = futs[i].result()
        futs[i] = None
        if i + ahead < len(specs):
            futs[i + ahead] = pool.submit(make_request, specs[i + ahead])
        wait = t0 + arrivals[i] - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        else:
            late = max(late, -wait)
        vc = "hetero" if spec["model"] == "ppml" else "iid"
        tq.put(server.submit(spec["model"], codes, g, y, X, vcov=vc))
        del codes, g, y, X
    tq.put(None)
    th.join()
    wall = time.perf_counter() - t0
    pool.shutdown()
    return replies, wall, late


def run_burst(server, specs, reqs):
    """Submits already generated requests (make_request of each spec) all at once and waits for
    every reply: the engine's throughput with a full queue. Returns the replies and the wall time
    from the first submission to the last reply."""
    t0 = time.perf_counter()
    tickets = [server.submit(s["model"], *r, vcov="hetero" if s["model"] == "ppml" else "iid")
               for s, r in zip(specs, reqs)]
    replies = [server.wait(t) for t in tickets]
    return replies, time.perf_counter() - t0
