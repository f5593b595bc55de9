"""reghdfe's estimation convention over the C ABI (Session::prepare + Session::ols, session.rs,
ols.rs; the reghdfe front end of python/fereg/_binding.py).

`reghdfe` builds the sample on the GPU (factor codes, interactions, singletons, frequency-weighted
counts), creates the absorber on the kept rows with the raw weights, and runs fh_reghdfe_fit:
standardized collinearity sweeps, the regression through the means for the constant, reghdfe's
degrees of freedom and small-sample factors, the multiway cluster PSD repair and the e() scalars.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .absorber import Absorber, _ptr
from .sample import Sample

STATUS = {0: "ok", 2: "absorbed", 3: "collinear"}


@dataclass
class ReghdfeResult:
    b: np.ndarray          # regressors, then _cons when reported
    V: np.ndarray
    status: np.ndarray     # per regressor: 0 kept, 2 absorbed by the fixed effects, 3 collinear
    dof_table: np.ndarray  # slots x (K, M, K - M, inexact, nested)
    scalars: dict
    sample: np.ndarray     # e(sample) over the input rows
    psd_fixed: bool = False
    se: np.ndarray = field(init=False)

    def __post_init__(self):
        self.se = np.sqrt(np.clip(np.diag(self.V), 0.0, None))


def _terms(v):
    return [t if isinstance(t, (tuple, list)) else (t,) for t in (v or [])]


def reghdfe(y, X, absorb=(), vce="unadjusted", cluster=(), weights=None, weight_kind=None, noconstant=False,
            singletons=True, dof=("pairwise", "clusters", "continuous"), tol=1e-8, maxiter=16000, precision=None,
            ctx=None):
    """reghdfe y X, absorb(absorb) vce(vce [cluster]) [weight_kind=weights] on the GPU.

    absorb / cluster: factor columns, or tuples of columns for interactions (a#b). dof: the
    dofadjustments() list ("firstpair", "pairwise", "clusters", "continuous"). Rows with a missing
    value (non-finite y, X or weights; a missing factor value) are left out, as Stata's sample
    marking does."""
    X = np.asarray(X, float)
    if X.ndim == 1:
        X = X[:, None]
    y = np.asarray(y, float)
    n, K = X.shape
    if vce not in L.FH_VCE:
        raise L.FeregError(L.FH_INVALID, f"unknown vce {vce!r}")
    if (weights is None) != (weight_kind is None):
        raise L.FeregError(L.FH_INVALID, "weights and weight_kind go together")
    cl_terms = _terms(cluster) if vce == "cluster" else []
    if vce == "cluster" and not cl_terms:
        raise L.FeregError(L.FH_INVALID, "vce(cluster) needs cluster variables")
    ab_terms = _terms(absorb)
    smp = Sample(y=y, x=list(X.T), fe=ab_terms, cluster=cl_terms, weights=weights, weight_kind=weight_kind,
                 missing=True, drop_singletons=singletons)
    try:
        kept = smp.kept()
        yk = y[kept]
        Xk = np.ascontiguousarray(X[kept])
        wk = None if weights is None else np.asarray(weights, float)[kept]
        terms = [smp.term(q) for q in range(smp.nterms)]
        clusters = [np.ascontiguousarray(smp.cluster(q)[0]) for q in range(smp.nclusters)]
        n_sing = smp.counts["singletons"]
        sample = np.zeros(n, bool)
        sample[kept] = True
    finally:
        smp.close()
    ab_codes = [c for c, _ in terms] or [np.zeros(len(kept), np.uint32)]
    ab_g = [g for _, g in terms] or [1]
    ab = Absorber(ab_codes, ab_g, weights=wk)
    try:
        spec = L.ReghdfeSpec()
        L.lib().fh_reghdfe_spec_default(ctypes.byref(spec))
        tarr = (L.Term * max(len(terms), 1))(*[L.Term(_ptr(c, L.u32p), int(g), 1) for c, g in terms])
        carr = (L.vp * max(len(clusters), 1))(*[c.ctypes.data for c in clusters])
        spec.vce = L.FH_VCE[vce]
        spec.weight_kind = L.FH_WEIGHT[weight_kind]
        spec.terms = tarr
        spec.nterms = len(terms)
        spec.clusters = carr
        spec.nclusters = len(clusters)
        spec.noconstant = int(bool(noconstant))
        opts = set(dof)
        spec.dof_firstpair = int("firstpair" in opts)
        spec.dof_pairwise = int("pairwise" in opts)
        spec.dof_clusters = int("clusters" in opts)
        spec.dof_continuous = int("continuous" in opts)
        spec.precision = -1 if precision is None else int(precision)
        spec.tol = float(tol)
        spec.maxiter = int(maxiter)
        spec.num_singletons = int(n_sing)
        k_all = K + (0 if noconstant else 1)
        b = np.zeros(K + 1)
        V = np.zeros((K + 1) * (K + 1))
        status = np.zeros(K, np.int32)
        nslots = max(len(terms), 1)
        dt = np.zeros(5 * nslots)
        st = L.ReghdfeStats()
        L.check(L.lib().fh_reghdfe_fit(ab._h, _ptr(np.ascontiguousarray(yk), L.f64p), _ptr(Xk, L.f64p), K,
                                       ctypes.byref(spec), _ptr(b, L.f64p), _ptr(V, L.f64p), _ptr(status, L.i32p),
                                       _ptr(dt, L.f64p), ctypes.byref(st), L.ctx_ref(ctx)))
    finally:
        ab.close()
    sc = {k: float(getattr(st, k)) for k in L.REGHDFE_SCALARS + L.REGHDFE_FIT}
    sc["ic"] = float(st.iterations)
    if vce == "cluster":
        for d in range(len(clusters)):
            sc[f"N_clust{d + 1}"] = float(st.N_clust_dim[d])
    else:
        del sc["N_clust"], sc["N_clustervars"]
    sc = {k: (None if isinstance(v, float) and np.isnan(v) else v) for k, v in sc.items()}
    return ReghdfeResult(b[:k_all].copy(), V[:k_all * k_all].reshape(k_all, k_all).copy(), status.astype(int),
                         dt.reshape(nslots, 5), sc, sample, bool(st.psd_fixed))
