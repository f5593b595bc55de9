"""numpy restatement of Poisson pseudo-maximum likelihood with absorbed fixed effects (FP64).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU IRLS; the product path never
imports it. The fixed effects are absorbed by the oracle's own solver (OracleSolver, the C
restatement), so the IRLS below sees only residualized blocks, as the reference's does.

Follows fereg-core's IRLS in fixest's flavour:
  initial values     glm/mod.rs:798-801           mu = y + 0.1, eta = ln mu
  deviance anchor    glm/irls.rs:432-447          dev0 at eta = 1e-5 (offset left out)
  fill               glm/irls.rs:467-493          w = tw mu, z = (eta - off) - 1 + y/mu, |v| <= ztol -> 0
  collinearity       glm/irls.rs:505-545          first iteration, weighted Gram, tol 1e-9
  coefficients       glm/irls.rs:459, 548-569     normal equations (invsym) while "fast", else QR
  step               glm/irls.rs:575-650          eta = z - e + off, mu = max(exp eta, 100 eps)
  halving            glm/irls.rs:693-710          fixest rule
  convergence        glm/irls.rs:824-827          |dev - dev_old| / (0.1 + |dev|) < tol at the final inner tol
  inner tolerance    glm/irls.rs:893-901          tol_in = max(min(eps^1.5, tol_in), highest)
  predict_eps        glm/irls.rs:105-141
  variance           fit/glm.rs:521-525           bread = (X~' W X~)^-1 (Cholesky), scores x~ e w
"""
import math

import numpy as np

from . import estimate
from .oracle import OracleSolver

MU_MIN = 100.0 * np.finfo(float).eps
ETA_ANCHOR = 1e-5
EPS1 = np.finfo(float).eps


def finish_deviance(raw, n):
    """glm/irls.rs:212-222."""
    d = raw
    if 2.0 * d / n < EPS1:
        d = 0.0
    if abs(d) <= EPS1:
        d = 0.0
    return 2.0 * d


def deviance_raw(mu, eta, y, tw, lny):
    """glm/irls.rs:268-270, summed."""
    w = np.ones_like(y) if tw is None else tw
    return float(np.sum(w * (mu - y) + np.where(y > 0, w * y * (lny - eta), 0.0)))


def predict_eps(hist, eps):
    """glm/irls.rs:105-141: cubic through (i, ln eps_i), i = 1..3, extrapolated to 4."""
    pos = sum(0 if math.isnan(v) else 1 for v in hist)
    abort = pos + 1 < 3
    le = math.log(eps) if eps > 0 else float("nan")
    if pos < 3:
        hist[pos] = le
    else:
        hist[0], hist[1], hist[2] = hist[1], hist[2], le
    if abort or any(math.isnan(v) for v in hist):
        return float("nan")
    A = np.array([[1.0, 1.0, 1.0], [2.0, 4.0, 8.0], [3.0, 9.0, 27.0]])
    c = np.linalg.solve(A, np.array(hist))
    return math.exp(4 * c[0] + 16 * c[1] + 64 * c[2])


def invsym(a, tol=1e-13):
    """vcov.rs:22-65: Goodnight sweeps in column order; a pivot <= tol x its diagonal drops the
    column (row and column of the result zero)."""
    k = a.shape[0]
    m = a.astype(float).copy()
    diag = np.diag(a).copy()
    for p in range(k):
        d = m[p, p]
        if math.isnan(d) or d <= tol * diag[p] or d <= 0:
            m[p, :] = 0.0
            m[:, p] = 0.0
            continue
        m[p, :] /= d
        for i in range(k):
            if i == p:
                continue
            b = m[i, p]
            if b == 0:
                continue
            m[i, :] -= b * m[p, :]
            m[i, p] = -b / d
        m[p, p] = 1.0 / d
    return 0.5 * (m + m.T)


def chol_collin_certain(a, tol, err):
    """vcov.rs:141-159: the chol_collin exclusions when column errors err cannot change them."""
    r, excluded = estimate.chol_pivots(a, tol)
    k = a.shape[0]
    pivots = np.array([a[j, j] - sum(r[i, j] ** 2 for i in range(j) if not excluded[i]) for j in range(k)])
    err = np.asarray(err, float) + 2.0 * np.sqrt(1e-11 * np.diag(a))
    b = np.zeros(k)
    for j in range(k):
        kept = [i for i in range(j) if not excluded[i]]
        for q in range(len(kept) - 1, -1, -1):
            i = kept[q]
            s = sum(r[i, l] * b[l] for l in kept[q + 1:])
            b[i] = (r[i, j] - s) / r[i, i]
        margin = err[j] + sum(abs(b[i]) * err[i] for i in kept)
        dist = math.sqrt(max(pivots[j], 0.0))
        stays = (dist + margin < math.sqrt(tol)) if excluded[j] else (dist - margin >= math.sqrt(tol))
        if not stays:
            return None
    return excluded


def qr_coefs(block, kept, w):
    """wls.rs:103-176: R of sqrt(w) [X_kept, z] (Householder, LAPACK) and back substitution with
    diagonal entries <= 1e-13 max |R_ii| taken as zero."""
    sw = np.sqrt(w)
    A = np.concatenate([block[:, 1 + np.asarray(kept)], block[:, :1]], axis=1) * sw[:, None]
    R = np.linalg.qr(A, mode="r")
    m = R.shape[0]
    k = m - 1
    dmax = max(abs(R[j, j]) for j in range(k))
    b = np.zeros(k)
    for j in range(k - 1, -1, -1):
        d = R[j, j]
        if abs(d) <= 1e-13 * dmax or d == 0:
            continue
        s = R[j, k] - sum(R[j, l] * b[l] for l in range(j + 1, k))
        b[j] = s / d
    return b


def ppml(codes, g, y, X, weights=None, offset=None, tol=1e-8, itol=1e-10, maxiter=100, vcov="hetero", cluster=None,
         nested=None, k_adj=True, g_adj=True, k_fixef="nonnested", slope_terms=None):
    """Poisson PML of y on X (kept order) with the FEs of `codes`; returns a dict with coef, vcov,
    se, deviance, iterations, converged, eta, mu, w (IRLS weights), kept, df_k, and pi (the row
    order of the internal arrays). slope_terms (codes, g, intercept, slope columns) replace
    codes / g by terms with varying slopes, absorbed by the explicit-design projection."""
    n, K = X.shape
    if slope_terms is not None:
        from . import slopes as _sl
        s = _sl.DenseSlopeSolver(slope_terms, n)
    else:
        s = OracleSolver(codes, g)
    pi = s.structure()["pi"]
    yr = np.asarray(y, float)[pi]
    xr = np.ascontiguousarray(np.asarray(X, float)[pi])
    tw = None if weights is None else np.asarray(weights, float)[pi]
    off = np.zeros(n) if offset is None else np.asarray(offset, float)[pi]
    if np.any(~np.isfinite(yr)) or np.any(yr < 0):
        raise ValueError("Poisson outcomes must be non-negative and finite")
    ztol = min(tol, 1e-12)
    lny = np.where(yr > 0, np.log(np.where(yr > 0, yr, 1.0)), 0.0)
    xtiny = bool(np.any((xr != 0) & (np.abs(xr) <= ztol)))
    mu = yr + 0.1
    eta = np.log(mu)
    dev0 = finish_deviance(deviance_raw(np.full(n, math.exp(ETA_ANCHOR)), np.full(n, ETA_ANCHOR), yr, tw, lny), n)
    highest = max(1e-12, min(itol, 0.1 * tol))
    tol_in = max(1e-4, tol)
    hist = [float("nan")] * 3
    eps = float("nan")
    dev_prev = float("nan")
    warm = None
    excluded = np.zeros(K, bool)
    kept = list(range(K))
    twv = np.ones(n) if tw is None else tw
    converged = False
    it = 0
    for it in range(1, maxiter + 1):
        predicted = predict_eps(hist, eps)
        fast = tol_in > tol * 11 and (math.isnan(predicted) or predicted > tol)
        w = twv * mu
        z = (eta - off) - 1.0 + yr / mu
        block = np.concatenate([z[:, None], xr], axis=1)
        block[:, 0] = np.where(np.abs(block[:, 0]) <= ztol, 0.0, block[:, 0])
        if xtiny:
            block[:, 1:] = np.where(np.abs(block[:, 1:]) <= ztol, 0.0, block[:, 1:])
        block = np.ascontiguousarray(block)
        s.set_weights(w)
        _, exact, coefs = s.residualize(block, K + 1, tol=tol_in, warm=warm, want_coefs=True)
        warm = coefs
        block[np.abs(block) <= ztol] = 0.0
        G = estimate.gram(block, w)
        if it == 1:
            raw = G[1:, 1:]
            if exact:
                excl = estimate.chol_pivots(raw, 1e-9)[1]
            else:
                err = 100.0 * tol_in * np.sqrt(np.maximum(np.diag(raw), 0.0))
                excl = chol_collin_certain(raw, 1e-9, err)
                if excl is None:
                    s.residualize(block, K + 1, tol=min(tol_in, highest))
                    block[np.abs(block) <= ztol] = 0.0
                    G = estimate.gram(block, w)
                    excl = estimate.chol_pivots(G[1:, 1:], 1e-9)[1]
            excluded = np.asarray(excl, bool)
            kept = [j for j in range(K) if not excluded[j]]
        if fast:
            sub = G[np.ix_([1 + j for j in kept], [1 + j for j in kept])]
            coef = invsym(sub) @ G[[1 + j for j in kept], 0]
        else:
            coef = qr_coefs(block, kept, w)
        beta = np.zeros(K)
        beta[kept] = coef
        e = block[:, 0] - block[:, 1:] @ beta
        eta_old = (ETA_ANCHOR + off) if it == 1 else eta.copy()
        eta = z - e + off
        mu = np.maximum(np.exp(eta), MU_MIN)
        dev = finish_deviance(deviance_raw(mu, eta, yr, tw, lny), n)
        old_dev = dev0 if it == 1 else dev_prev
        halved = False

        def evol(d):
            return d - old_dev

        small = math.isfinite(dev) and (abs(evol(dev)) < tol or abs(evol(dev)) / (0.1 + abs(dev)) < tol)
        if not small and (not math.isfinite(dev) or evol(dev) > 0):
            h = 0
            while (not math.isfinite(dev) or evol(dev) > 0) and not (it == 1 and math.isfinite(dev) and h >= 2) \
                    and h < maxiter:
                eta = 0.5 * (eta + eta_old)
                mu = np.maximum(np.exp(eta), MU_MIN)
                dev = finish_deviance(deviance_raw(mu, eta, yr, tw, lny), n)
                h += 1
            halved = True
        delta = old_dev - dev
        if not math.isnan(delta) and dev < 0.1 * delta:
            delta = dev
        if it > 1:
            eps = abs(delta) / max(min(dev, old_dev), 0.1)
        crit = math.inf if halved else abs(old_dev - dev) / (0.1 + abs(dev))
        if crit < tol and (exact or tol_in <= 1.1 * highest):
            converged = True
            break
        if it > 1:
            tol_in = max(min(eps * math.sqrt(eps), tol_in), highest)
        dev_prev = dev
    # variance at the last IRLS weights (fit/glm.rs:521-525)
    Xt = block[:, [1 + j for j in kept]]
    sub = G[np.ix_([1 + j for j in kept], [1 + j for j in kept])]
    _, bread = estimate.chol_collin(sub, np.finfo(float).tiny)
    if slope_terms is not None:
        np_ = _sl.nparams(len(kept), slope_terms)
        nest = nested if nested is not None else [False] * len(slope_terms)
        dfk = _sl.df_k(vcov, np_, slope_terms, nest, k_fixef=k_fixef, n_design=len(kept))
    else:
        sizes = list(g)
        np_ = len(kept) + (sum(sizes) - (len(sizes) - 1) if sizes else 0)
        nest = nested if nested is not None else [False] * len(sizes)
        dfk = estimate.df_k(vcov, np_, sizes, nest, k_adj=k_adj, k_fixef=k_fixef, n_design=len(kept))
    if vcov == "iid":
        V = bread * ((n - 1) / (n - dfk) if k_adj else 1.0)
    else:
        dims = [] if cluster is None else (cluster if isinstance(cluster, (list, tuple)) else [cluster])
        cl = [np.asarray(c)[pi] for c in dims] or None
        h = None
        if vcov in ("hc2", "hc3"):
            # the full model's leverage at the last IRLS weights (fit/glm.rs:588-599; a GLM keeps
            # the weight in the regressor part even without absorbed terms)
            if slope_terms is not None:
                D = _sl.slope_design(slope_terms, n)
            else:
                D = estimate.fe_design([np.asarray(c)[pi] for c in codes], g, n, intercept=not len(codes))
            h = estimate.hat_values(np.concatenate([xr[:, kept], D], axis=1), w)
        V = estimate.vcov(vcov, bread, Xt, e, w, n, dfk, cluster=cl, k_adj=k_adj, g_adj=g_adj, h=h)
    Vf = np.zeros((K, K))
    Vf[np.ix_(kept, kept)] = V
    return {"coef": beta, "vcov": Vf, "se": np.sqrt(np.maximum(np.diag(Vf), 0.0)), "deviance": dev,
            "iterations": it, "converged": converged, "eta": eta, "mu": mu, "w": w, "kept": ~excluded,
            "df_k": dfk, "pi": pi, "resid": e}
