"""numpy restatement of the estimation steps on a residualized block (FP64).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and bench.py's cpu_baseline leg as the
checker. Follows fereg-core's fit layer (fixest conventions):
  gram          crates/fereg-core/src/stats.rs:195-251 (G = sum w z z')
  chol_collin   vcov.rs:124-128,163-211 (column-order Cholesky, pivot < tol excludes a column)
  regress       fit/design.rs:64-95 (coefficients from the Gram matrix)
  nparams       fit/ols.rs:27-37
  df_k          fit/vcov.rs:63-112 (K.adj, K.fixef none/full/nonnested)
  vcov          fit/vcov.rs:316-382 (iid, hetero, one-way cluster)
  poisson IRLS  glm/irls.rs:367-913 with glm/family.rs:103-145 (fixest flavour)
"""
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


def vcov(kind, bread, X, e, w, n, dfk, cluster=None, k_adj=True, g_adj=True):
    """fit/vcov.rs:316-382: iid, hetero, one-way CRV1 (G.df = "min")."""
    if kind == "iid":
        ssr = float(np.sum((w if w is not None else 1.0) * e * e))
        factor = (n - 1) / (n - dfk) if k_adj else 1.0
        return bread * (ssr / (n - 1)) * factor
    s = X * (e * (w if w is not None else 1.0))[:, None]
    if kind == "hetero":
        meat = s.T @ s
        factor = n / (n - dfk) if k_adj else 1.0
        return factor * bread @ meat @ bread
    if kind == "cluster":
        codes = np.asarray(cluster)
        G = int(codes.max()) + 1
        sg = np.stack([np.bincount(codes, weights=s[:, j], minlength=G) for j in range(X.shape[1])], axis=1)
        Geff = len(np.unique(codes))
        meat = sg.T @ sg
        factor = 1.0
        if g_adj:
            factor *= Geff / (Geff - 1)
        if k_adj:
            factor *= (n - 1) / (n - dfk)
        return factor * bread @ meat @ bread
    raise ValueError(kind)


def ols_from_block(rows, w=None, tol=1e-9):
    """OLS of column 0 on columns 1.. of a residualized block: (beta, bread, resid, kept)."""
    r = rows.shape[1]
    g = gram(rows, w)
    design = [np.eye(r)[j] for j in range(1, r)]
    excluded, kept, inv, coef, resid = regress(g, np.eye(r)[0], design, tol)
    e = rows @ resid
    return coef, inv, e, kept
