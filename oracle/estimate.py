"""numpy restatement of the estimation steps on a residualized block (FP64).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and bench.py's cpu_baseline leg as the
checker. Follows fereg-core's fit layer (fixest conventions):
  gram          crates/fereg-core/src/stats.rs:195-251 (G = sum w z z')
  chol_collin   vcov.rs:124-128,163-211 (column-order Cholesky, pivot < tol excludes a column)
  regress       fit/design.rs:64-95 (coefficients from the Gram matrix)
  nparams       fit/ols.rs:27-37
  df_k          fit/vcov.rs:63-112 (K.adj, K.fixef none/full/nonnested)
  vcov          fit/vcov.rs:316-382 (iid, hetero, HC2/HC3, one- to multiway cluster)
  hat_values    leverage.rs:1-30 (the exact leverage of the full model, here by brute force on the
                explicit design: diag of W^1/2 X (X'WX)^+ X' W^1/2 over [X, every FE dummy])
  crv3          crv3.rs:1-104, fit/vcov.rs:369-397 (leave-one-cluster-out refits by dropping the
                cluster's rows, inclusion-exclusion over the cluster dimensions)
  poisson IRLS  glm/irls.rs:367-913 with glm/family.rs:103-145 (fixest flavour)
"""
import itertools

import numpy as np


def gram(rows, w=None):
    """G = sum_i w_i z_i z_i' over an n x R block (stats.rs:195-251)."""
    if w is None:
        return rows.T @ rows
    return (rows * w[:, None]).T @ rows


def chol_pivots(a, tol):
    """vcov.rs:163-188: Cholesky in column order; column j is excluded when its pivot < tol."""
    k = a.shape[0]
    r = np.zeros((k, k))
    excluded = np.zeros(k, bool)
    for j in range(k):
        rjj = a[j, j] - sum(r[i, j] ** 2 for i in range(j) if not excluded[i])
        if not (rjj >= tol):
            excluded[j] = True
            continue
        d = np.sqrt(rjj)
        r[j, j] = d
        for c in range(j + 1, k):
            v = a[c, j] - sum(r[i, c] * r[i, j] for i in range(j) if not excluded[i])
            r[j, c] = v / d
    return r, excluded


def chol_collin(a, tol):
    """vcov.rs:124-128: exclusions and (R'R)^-1 over the kept columns."""
    r, excluded = chol_pivots(a, tol)
    kept = np.flatnonzero(~excluded)
    rk = r[np.ix_(kept, kept)]
    ri = np.linalg.inv(rk) if len(kept) else rk
    return excluded, ri @ ri.T


def regress(g, target, design, tol=1e-9):
    """fit/design.rs:64-95: regression of combo `target` on combos `design` from the Gram g."""
    k = len(design)
    D = np.array(design, dtype=float).reshape(k, g.shape[0])
    block = D @ g @ D.T
    excluded, inv = chol_collin(block, tol)
    kept = np.flatnonzero(~excluded)
    xty = D[kept] @ g @ np.asarray(target, float)
    coef = inv @ xty
    resid = np.asarray(target, float) - coef @ D[kept]
    return excluded, kept, inv, coef, resid


def nparams(k, term_sizes):
    """fit/ols.rs:27-37: kept coefficients + sum of intercept-term levels - (#terms - 1)."""
    q = len(term_sizes)
    return k + sum(term_sizes) - max(q - 1, 0)


def df_k(kind, np_, term_sizes, nested, k_adj=True, k_fixef="nonnested", n_design=0):
    """fit/vcov.rs:63-112 (K.exact = FALSE)."""
    fe_params = (sum(term_sizes) - (len(term_sizes) - 1)) if term_sizes else 0
    kf = k_fixef
    if kind != "cluster" and kf == "nonnested":
        kf = "full"
    if kf == "none":
        return np_ - fe_params
    if kf == "full":
        return np_
    if not term_sizes:
        return np_
    n_nested = sum(nested)
    k = float(np_)
    if n_nested == len(term_sizes):
        k = k - fe_params + 1
    else:
        s = sum(g for g, x in zip(term_sizes, nested) if x)
        c = sum(1 for g, x in zip(term_sizes, nested) if x and g > 0)
        k -= s - c
    return max(k, n_design + 1)


def _cluster_meat(s, codes):
    codes = np.asarray(codes)
    _, dense = np.unique(codes, return_inverse=True)
    dense = dense.reshape(-1)
    G = int(dense.max()) + 1
    sg = np.stack([np.bincount(dense, weights=s[:, j], minlength=G) for j in range(s.shape[1])], axis=1)
    return sg.T @ sg, G


def _intersection(dims):
    return np.unique(np.stack([np.asarray(d) for d in dims], axis=1), axis=0, return_inverse=True)[1].reshape(-1)


def posdef_fix(V):
    """fixest mat_posdef_fix (fit/vcov.rs:514-534): eigenvalues <= 0 clipped at 1e-16."""
    if V.size == 0:
        return V
    lam, U = np.linalg.eigh(V)
    if np.all(lam > 0):
        return V
    return (U * np.maximum(lam, 1e-16)) @ U.T


def fe_design(codes, gs, n, intercept=False):
    """The explicit dummies of every absorbed term (and a constant when asked), n x sum(g)."""
    cols = [np.ones((n, 1))] if intercept else []
    for c, g in zip(codes, gs):
        d = np.zeros((n, g))
        d[np.arange(n), np.asarray(c)] = 1.0
        cols.append(d)
    return np.concatenate(cols, axis=1) if cols else np.zeros((n, 0))


def hat_values(design, w=None, row_weight=True):
    """h_i = w_i x_i' (X'WX)^+ x_i for the full design X (leverage.rs:1-30 defines the FE part the
    same way; with the regressors the FWL split h = h_FE + w x~'(X~'WX~)^-1 x~ holds): the
    diagonal of the weighted projection, by a pseudo-inverse of the explicit design.
    row_weight=False leaves the leading w_i out: fixest's OLS without absorbed terms
    (fit/vcov.rs:235, `weighted` is false for a t-inference fit with no fixed effect)."""
    sw = np.sqrt(w) if w is not None else np.ones(design.shape[0])
    a = design * sw[:, None]
    h = np.einsum("ij,ji->i", a, np.linalg.pinv(a, rcond=1e-12))
    return h if row_weight or w is None else h / w


def vcov(kind, bread, X, e, w, n, dfk, cluster=None, k_adj=True, g_adj=True, g_df="min", h=None):
    """fit/vcov.rs:316-382: iid, hetero, HC2 / HC3 (scores over sqrt(1 - h) or (1 - h) with the
    full model's leverage h, no small-sample factor: fit/vcov.rs:333-364, fixest flavour), CRV1
    over one or several cluster dimensions (inclusion-exclusion over the subsets of the
    dimensions, vcov.rs:480-558; G.df "min" or "conventional"), then fixest's PSD repair."""
    if kind == "iid":
        ssr = float(np.sum((w if w is not None else 1.0) * e * e))
        factor = (n - 1) / (n - dfk) if k_adj else 1.0
        return posdef_fix(bread * (ssr / (n - 1)) * factor)
    s = X * (e * (w if w is not None else 1.0))[:, None]
    if kind == "hetero":
        meat = s.T @ s
        factor = n / (n - dfk) if k_adj else 1.0
        return posdef_fix(factor * bread @ meat @ bread)
    if kind in ("hc2", "hc3"):
        if np.any(h > 1.0 - np.sqrt(np.finfo(float).eps)):
            raise ValueError("a leverage equals 1")
        d = (1.0 - h) if kind == "hc3" else np.sqrt(1.0 - h)
        sh = s / d[:, None]
        return posdef_fix(bread @ (sh.T @ sh) @ bread)
    if kind == "cluster":
        dims = cluster if isinstance(cluster, (list, tuple)) else [cluster]
        q = len(dims)
        conventional = g_adj and g_df == "conventional"
        meat = np.zeros((X.shape[1], X.shape[1]))
        gs = []
        for size in range(1, q + 1):
            sign = 1.0 if size % 2 else -1.0
            for sub in itertools.combinations(range(q), size):
                codes = dims[sub[0]] if size == 1 else _intersection([dims[d] for d in sub])
                m, G = _cluster_meat(s, codes)
                if size == 1:
                    gs.append(G)
                meat += sign * (G / (G - 1) if conventional else 1.0) * m
        factor = 1.0
        gmin = min(gs)
        if g_adj and g_df == "min":
            factor *= gmin / (gmin - 1)
        if k_adj:
            factor *= (n - 1) / (n - dfk)
        return posdef_fix(factor * bread @ meat @ bread)
    raise ValueError(kind)


def ols_from_block(rows, w=None, tol=1e-9):
    """OLS of column 0 on columns 1.. of a residualized block: (beta, bread, resid, kept)."""
    r = rows.shape[1]
    g = gram(rows, w)
    design = [np.eye(r)[j] for j in range(1, r)]
    excluded, kept, inv, coef, resid = regress(g, np.eye(r)[0], design, tol)
    e = rows @ resid
    return coef, inv, e, kept


def crv3(fit_beta, codes, y, X, w, dims, dfk, k_adj=True, g_adj=True, g_df="min", offset=None, slopes=None):
    """The CRV3 variance: for every non-empty subset of the cluster dimensions (clustered by the
    intersection) and every cluster g of it, beta_(-g) = fit_beta(codes, y, X, w, offset) on the
    rows outside g (codes re-densified, as a fresh fit sees them); V = factor sum sign
    scale(G) sum_g (beta_(-g) - beta)(beta_(-g) - beta)' with factor G_min/(G_min - 1)
    (n - 1)/(n - df_k) (crv3.rs:89-104, fit/vcov.rs:369-397), then fixest's PSD repair.
    slopes (per term, its slope columns): fit_beta then takes them, restricted to the same rows,
    after the codes."""
    n = len(y)

    def sub(keep):
        cs = []
        for c in codes:
            _, inv = np.unique(np.asarray(c)[keep], return_inverse=True)
            cs.append(inv.astype(np.uint32))
        return cs

    def fit(keep):
        rest = (y[keep], X[keep], None if w is None else w[keep], None if offset is None else offset[keep])
        if slopes is not None:
            return fit_beta(sub(keep), [[np.asarray(x)[keep] for x in xs] for xs in slopes], *rest)
        return fit_beta(sub(keep), *rest)

    beta = fit(np.ones(n, bool))
    k = len(beta)
    q = len(dims)
    conventional = g_adj and g_df == "conventional"
    meat = np.zeros((k, k))
    gs = []
    for size in range(1, q + 1):
        sign = 1.0 if size % 2 else -1.0
        for s in itertools.combinations(range(q), size):
            cl = np.asarray(dims[s[0]]) if size == 1 else _intersection([dims[d] for d in s])
            G = int(cl.max()) + 1
            m = np.zeros((k, k))
            for g in range(G):
                keep = cl != g
                if keep.all():
                    continue
                d = fit(keep) - beta
                m += np.outer(d, d)
            ng = len(np.unique(cl))
            if size == 1:
                gs.append(ng)
            meat += sign * (ng / (ng - 1) if conventional else 1.0) * m
    factor = 1.0
    if g_adj and g_df == "min":
        factor *= min(gs) / (min(gs) - 1)
    if k_adj:
        factor *= (n - 1) / (n - dfk)
    return beta, posdef_fix(factor * meat)
