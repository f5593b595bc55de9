"""Estimation entry points over the C ABI (fit/mod.rs:591; fit/ols.rs:213; fit/glm.rs:362).

`fit` is the fit layer on raw columns: the estimation sample (missing values, zero weights,
factor encoding, interactions, singleton and fe-separation removal; fit/input.rs) is built on the
GPU, then `feols` / `fepois` run on the kept rows. `feols` and `fepois` take kept-sample arrays,
FE codes per term, and a variance request in fixest's conventions (fit/vcov.rs:278-382).
"""
import ctypes
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib as L
from .absorber import Absorber, SolveOpts, _dev_codes, _dev_f64, _dev_ready, _is_torch, _ptr


@dataclass
class FitResult:
    coef: np.ndarray
    vcov: np.ndarray
    kept: np.ndarray
    df_k: float
    nparams: float
    ssr: float
    deviance: float
    iterations: int
    pcg_iterations: int
    n_clusters: int
    se: np.ndarray = field(init=False)
    converged: bool = True
    nobs: Optional[int] = None
    obs: Optional[np.ndarray] = None
    dropped: Optional[dict] = None

    def __post_init__(self):
        self.se = np.sqrt(np.clip(np.diag(self.vcov), 0.0, None))


def nested_fixest(fe_codes, cluster_codes, fe_names=None, cluster_names=None):
    """fixest's K.fixef = "nonnested" test (fit/vcov.rs:115-136): by name first (a term that is a
    cluster variable, or an interaction containing one), then by value unless every cluster
    variable is itself an FE term."""
    nt = len(fe_codes)
    fe_names = fe_names or [None] * nt
    cl_names = cluster_names or []
    nested = [nm is not None and (nm in cl_names or ("^" in nm and any(f in cl_names for f in nm.split("^"))))
              for nm in fe_names]
    if not all(nested) and not (cl_names and all(c in fe_names for c in cl_names)):
        for q in range(nt):
            if not nested[q] and cluster_codes is not None:
                c = np.asarray(fe_codes[q])
                k = np.asarray(cluster_codes)
                g = int(c.max()) + 1
                first = np.full(g, -1, np.int64)
                order = np.arange(len(c))
                first[c[::-1]] = k[::-1]  # some cluster of each level
                nested[q] = bool(np.all(first[c] == k))
    return nested


def _host_f64(a, what, n=None):
    v = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and v.shape[0] != n:
        raise L.FeregError(L.FH_INVALID, f"{what}: {v.shape[0]} rows, the sample has {n}")
    return v


def _host_codes(a, what):
    v = np.asarray(a)
    if v.size and (not np.issubdtype(v.dtype, np.integer) or v.min() < 0 or v.max() >= 2 ** 32):
        raise L.FeregError(L.FH_INVALID, f"{what}: integer codes in 0 .. 2^32 - 1 expected")
    return np.ascontiguousarray(v, dtype=np.uint32)


def _cluster_dims(cluster):
    """A cluster argument as a list of dimensions (None, one array, or a list / tuple of arrays)."""
    if cluster is None:
        return []
    if isinstance(cluster, (list, tuple)):
        return list(cluster)
    return [cluster]


def _fit(kind_fn, y, X, codes, g, weights, vcov, cluster, nested, ssc, opts, absorber, extra_args, offset=None,
         ctx=None, intercepts=None, slopes=None):
    lib = L.lib()
    X = np.asarray(X, float) if not _is_torch(X) else X
    if not _is_torch(X) and X.ndim == 1:
        X = X[:, None]
    K = X.shape[1]
    if absorber is not None and weights is not None:
        raise L.FeregError(L.FH_INVALID, "weights belong to the absorber: pass them when it is built")
    dims = _cluster_dims(cluster)
    if len(dims) > 4:
        raise L.FeregError(L.FH_INVALID, "at most four cluster dimensions are supported")
    ab = absorber or Absorber(codes, g, weights=weights, intercepts=intercepts, slopes=slopes)
    kind = L.FH_VCOV[vcov]
    ssc = ssc or {}
    sscc = L.Ssc(int(ssc.get("K.adj", True)), L.FH_KFIXEF[ssc.get("K.fixef", "nonnested")], int(ssc.get("G.adj", True)), 0)
    nt = ab.info.nterms
    nest = np.zeros(max(nt, 1), np.int32)
    if nested is not None:
        nest[:nt] = np.asarray(nested, np.int32)
    beta = np.zeros(K)
    V = np.zeros(K * K)
    kept = np.zeros(K, np.int32)
    st = L.FitStats()
    o = (opts or SolveOpts(tol=1e-10)).c()
    poisson = kind_fn == "fh_ppml_fit"
    if _is_torch(y):
        held = _dev_ready(_dev_f64(y, "y"), _dev_f64(X, "X"), _dev_f64(offset, "offset"),
                          *[_dev_codes(c, "cluster") for c in dims])  # kept alive through the call
        cptr = [t.data_ptr() for t in held[3:]]
        fn = getattr(lib, kind_fn + "_multi_dev")
        args = [ab._h, held[0].data_ptr(), held[1].data_ptr(), K]
        if poisson:
            args.append(None if offset is None else held[2].data_ptr())
    else:
        n = ab.n()
        yv = _host_f64(y, "y", n)
        Xv = _host_f64(X, "X", n)
        cls = [_host_codes(c, "cluster") for c in dims]
        for c in cls:
            if c.shape[0] != n:
                raise L.FeregError(L.FH_INVALID, f"cluster: {c.shape[0]} rows, the sample has {n}")
        cptr = [c.ctypes.data for c in cls]
        fn = getattr(lib, kind_fn + "_multi")
        args = [ab._h, _ptr(yv, L.f64p), _ptr(Xv, L.f64p), K]
        if poisson:
            ov = None if offset is None else _host_f64(offset, "offset", n)
            args.append(None if ov is None else _ptr(ov, L.f64p))
    carr = (L.vp * max(len(cptr), 1))(*cptr)
    clc = L.Clusters(carr, len(cptr), L.FH_GDF[ssc.get("G.df", "min")], 0)
    args += [kind, ctypes.byref(clc) if cptr else None, _ptr(nest, L.i32p), ctypes.byref(sscc)]
    args += extra_args + [ctypes.byref(o), _ptr(beta, L.f64p), _ptr(V, L.f64p), _ptr(kept, L.i32p), ctypes.byref(st)]
    if poisson:
        args.append(L.ctx_ref(ctx))
    L.check(fn(*args))
    res = FitResult(beta, V.reshape(K, K), kept.astype(bool), st.df_k, st.nparams, st.ssr, st.deviance,
                    st.iterations, st.pcg_iterations, st.n_clusters)
    res.converged = bool(st.converged)
    if absorber is None:
        ab.close()
    return res


def feols(y, X, codes, g, weights=None, vcov="iid", cluster=None, nested=None, ssc=None, opts=None, absorber=None,
          intercepts=None, slopes=None):
    """OLS with absorbed FEs (fit/ols.rs:213-330) on the GPU. cluster: one code array or a list
    of up to four (multiway, fixest's ssc "G.df" rule). intercepts / slopes: per term, varying
    slopes as for Absorber (a model whose terms have no intercept keeps its constant among X,
    as fixest does)."""
    return _fit("fh_ols_fit", y, X, codes, g, weights, vcov, cluster, nested, ssc, opts, absorber, [],
                intercepts=intercepts, slopes=slopes)


def fepois(y, X, codes, g, weights=None, offset=None, vcov="hetero", cluster=None, nested=None, ssc=None, tol=1e-8,
           maxiter=100, opts=None, absorber=None, ctx=None, intercepts=None, slopes=None):
    """Poisson PML by IRLS (glm/irls.rs:367-913, fixest flavour) on the GPU. `weights` are prior
    weights (or the absorber's, when one is passed), `offset` enters the linear predictor."""
    return _fit("fh_ppml_fit", y, X, codes, g, weights, vcov, cluster, nested, ssc, opts, absorber,
                [ctypes.c_double(tol), ctypes.c_uint64(maxiter)], offset=offset, ctx=ctx, intercepts=intercepts,
                slopes=slopes)


def nested_by_value(fe_codes, cluster_codes):
    """Every level of each FE term within one cluster of some dimension (fit/vcov.rs:115-140 by
    value), on the host."""
    out = []
    for c in fe_codes:
        c = np.asarray(c, np.int64)
        g = int(c.max()) + 1 if len(c) else 0
        ok = False
        for k in cluster_codes:
            k = np.asarray(k, np.int64)
            lo = np.full(g, np.iinfo(np.int64).max)
            hi = np.full(g, np.iinfo(np.int64).min)
            np.minimum.at(lo, c, k)
            np.maximum.at(hi, c, k)
            if np.all(lo == hi):
                ok = True
                break
        out.append(ok)
    return out


def fit(y, X, fe, model="ols", vcov="iid", cluster=(), weights=None, offset=None, ssc=None, fe_names=None,
        cluster_names=None, tol=1e-8, maxiter=100, opts=None, ctx=None):
    """fit::fit (fit/mod.rs:591) in fixest's flavour on raw columns: fe and cluster are lists of
    terms (a factor column, or a tuple of columns for an interaction). Returns the FitResult of
    the kept sample with `nobs`, `obs` (kept input rows) and `dropped` (fit::Dropped counts)."""
    from .sample import Sample
    X = np.asarray(X, float)
    if X.ndim == 1:
        X = X[:, None]
    y = np.asarray(y, float)
    fam = {"ols": "gaussian", "poisson": "poisson"}[model]
    cl_terms = list(cluster) if vcov == "cluster" else []
    with Sample(y=y, x=list(X.T), fe=fe, cluster=cl_terms, weights=weights, offset=offset, family=fam,
                missing=True, drop_singletons=True, separation=fam != "gaussian", ctx=ctx) as smp:
        kept = smp.kept()
        terms = [smp.term(q) for q in range(smp.nterms)]
        clusters = [smp.cluster(q)[0] for q in range(smp.nclusters)]
        counts = dict(smp.counts)
    codes, g = [c for c, _ in terms], [gq for _, gq in terms]
    w = None if weights is None else np.asarray(weights, float)[kept]
    nested = None
    if vcov == "cluster":
        fe_names = fe_names or [None] * len(codes)
        cl_names = cluster_names or []
        by_name = [nm is not None and (nm in cl_names or ("^" in nm and any(f in cl_names for f in nm.split("^"))))
                   for nm in fe_names]
        if all(by_name) or (cl_names and all(c in fe_names for c in cl_names)):
            nested = by_name
        else:
            by_value = nested_by_value(codes, clusters)
            nested = [a or b for a, b in zip(by_name, by_value)]
    yk, Xk = y[kept], np.ascontiguousarray(X[kept])
    if model == "ols":
        res = feols(yk, Xk, codes, g, weights=w, vcov=vcov, cluster=clusters or None, nested=nested, ssc=ssc,
                    opts=opts)
    else:
        off = None if offset is None else np.asarray(offset, float)[kept]
        res = fepois(yk, Xk, codes, g, weights=w, offset=off, vcov=vcov, cluster=clusters or None, nested=nested,
                     ssc=ssc, tol=tol, maxiter=maxiter, opts=opts, ctx=ctx)
    res.nobs = int(counts["nobs"])
    res.obs = kept
    res.dropped = counts
    return res


@dataclass
class WildBootResult:
    stat: float
    p_val: float
    se: float
    point: float
    conf_int: tuple
    reps: int
    reps_feasible: int
    n_clusters: int
    full_enumeration: bool
    fe_not_projected: bool
    ssc: float
    seconds_bootstrap: float
    contraction_ms: float
    contraction_ops: float
    t_boot: np.ndarray
    fit: FitResult


def wildboot(y, X, cluster, R, r=0.0, reps=999, seed=1, level=0.95, pval_type="two-tailed", conf_int=True,
             tol=1e-6, maxiter=50, codes=None, g=None, nested=None, ssc=None, opts=None, absorber=None,
             boot_type=11):
    """Wild cluster bootstrap of R beta = r after OLS (wildboot/ols.rs:243-649): Rademacher
    weights from SplitMix64 streams (seed, draw), bootstrap clusters = error clusters, p-value and
    test-inverted interval. The draws x clusters contraction runs on int8 MFMA. boot_type 11 is
    WCR11; 13, 31 and 33 are MacKinnon, Nielsen and Webb's (2023) variants with the CRV3
    jackknife in the variance, the data generating process, or both (wildboot/ols.rs:651-1082)."""
    lib = L.lib()
    Xa = X if _is_torch(X) else np.asarray(X, float)
    if not _is_torch(Xa) and Xa.ndim == 1:
        Xa = Xa[:, None]
    K = Xa.shape[1]
    ab = absorber or Absorber(codes, g)
    ssc = ssc or {}
    sscc = L.Ssc(int(ssc.get("K.adj", True)), L.FH_KFIXEF[ssc.get("K.fixef", "nonnested")], int(ssc.get("G.adj", True)), 0)
    nt = ab.info.nterms
    nest = np.zeros(max(nt, 1), np.int32)
    if nested is not None:
        nest[:nt] = np.asarray(nested, np.int32)
    Rv = np.ascontiguousarray(np.asarray(R, float).reshape(-1))
    assert Rv.size == K, "R needs one entry per regressor"
    spec = L.WbootSpec(int(reps), int(seed), _ptr(Rv, L.f64p), float(r), float(level), float(tol), int(maxiter),
                       L.FH_PVAL[pval_type], int(bool(conf_int)), int(boot_type), 0)
    beta = np.zeros(K)
    V = np.zeros(K * K)
    kept = np.zeros(K, np.int32)
    res = L.WbootResult()
    tb = np.full(int(reps), np.nan)  # full enumeration never exceeds the requested reps
    o = (opts or SolveOpts(tol=1e-10)).c()
    if _is_torch(y):
        fn = lib.fh_ols_wildboot_dev
        held = _dev_ready(_dev_f64(y, "y"), _dev_f64(Xa, "X"), _dev_codes(cluster, "cluster"))  # kept alive
        args = [ab._h, held[0].data_ptr(), held[1].data_ptr(), K, held[2].data_ptr()]
    else:
        fn = lib.fh_ols_wildboot
        yv = np.ascontiguousarray(y, dtype=np.float64)
        Xv = np.ascontiguousarray(Xa, dtype=np.float64)
        cl = _host_codes(cluster, "cluster")
        args = [ab._h, _ptr(yv, L.f64p), _ptr(Xv, L.f64p), K, _ptr(cl, L.u32p)]
    L.check(fn(*args, _ptr(nest, L.i32p), ctypes.byref(sscc), ctypes.byref(o), ctypes.byref(spec), _ptr(beta, L.f64p),
               _ptr(V, L.f64p), _ptr(kept, L.i32p), ctypes.byref(res), _ptr(tb, L.f64p)))
    if absorber is None:
        ab.close()
    fit = FitResult(beta, V.reshape(K, K), kept.astype(bool), float("nan"), float("nan"), float("nan"), 0.0, 0, 0,
                    int(res.n_clusters))
    return WildBootResult(res.stat, res.p_val, res.se, res.point, (res.ci_lo, res.ci_hi), int(res.reps),
                          int(res.reps_feasible), int(res.n_clusters), bool(res.full_enumeration),
                          bool(res.fe_not_projected), res.ssc, res.seconds_bootstrap, res.contraction_ms,
                          res.contraction_ops, tb[:int(res.reps)].copy(), fit)
