"""numpy restatement of OLS with reghdfe 6.13.1's conventions (FP64).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU reghdfe path; the product
path never imports it. The fixed effects are absorbed by the oracle's own solver (OracleSolver,
the C restatement); a model without absorbed terms absorbs a one-level constant, as reghdfe does.

Follows fereg-core:
  dof                  dof.rs:64-182        nested slots, redundant levels, first-intercept
                                            exactness, pairwise mobility groups (connected
                                            components of the bipartite level graph)
  column statistics    stats.rs:10-44,154   unweighted sum x^2 and sd (n - 1, floored at 1e-3),
                                            weighted means and sums of squares (raw weights)
  weights              ols.rs:64-137        wv = w N_obs / sum w (aw, pw), w (fw); N = sum fw
  collinearity         ols.rs:31-56         stage 1 ss~ / ss <= min(1e-6, tol / 10); stage 2
                                            invsym on the sd-standardized Gram
  constant             vcov.rs:215-236      inverse extended by the regression through the means
  variance             ols.rs:404-500       unadjusted, robust (pweight + unadjusted -> robust),
                                            multiway cluster with q_cluster and the CGM PSD fix,
                                            all on the standardized scale
  F                    ols.rs:603-620       invsym_greedy of the slope block, tol 1e-9
  fit statistics       stats.rs:267-280
"""
import itertools
import math

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from .oracle import OracleSolver

INVSYM_TOL = 1e-13


def invsym(a, tol=INVSYM_TOL):
    """vcov.rs:22-65: Goodnight sweeps in column order; a pivot <= tol x its diagonal (or <= 0,
    NaN) drops the column. Returns (inverse, dropped flags)."""
    k = a.shape[0]
    m = a.astype(float).copy()
    diag = np.diag(a).copy()
    dropped = np.zeros(k, bool)
    for p in range(k):
        d = m[p, p]
        if math.isnan(d) or d <= tol * diag[p] or d <= 0:
            dropped[p] = True
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
    return 0.5 * (m + m.T), dropped


def invsym_greedy(a, tol):
    """vcov.rs:70-116: sweeps the unvisited pivot with the largest remaining diagonal."""
    k = a.shape[0]
    m = a.astype(float).copy()
    diag = np.diag(a).copy()
    visited = np.zeros(k, bool)
    dropped = np.zeros(k, bool)
    for _ in range(k):
        cand = [i for i in range(k) if not visited[i]]
        p = max(cand, key=lambda i: (m[i, i], -i))
        visited[p] = True
        d = m[p, p]
        if math.isnan(d) or d <= tol * diag[p] or d <= 0.0:
            dropped[p] = True
            m[p, :] = 0.0
            m[:, p] = 0.0
            continue
        m[p, :] /= d
        for i in range(k):
            if i == p:
                continue
            b = m[i, p]
            if b == 0.0:
                continue
            m[i, :] -= b * m[p, :]
            m[i, p] = -b / d
        m[p, p] = 1.0 / d
    m = 0.5 * (m + m.T)
    return m, dropped


def nested_within(codes, cluster):
    """Every level of the term maps to one cluster value (dof.rs:184-197)."""
    g = int(codes.max()) + 1
    lo = np.full(g, np.iinfo(np.int64).max)
    hi = np.full(g, np.iinfo(np.int64).min)
    np.minimum.at(lo, codes, cluster.astype(np.int64))
    np.maximum.at(hi, codes, cluster.astype(np.int64))
    return bool(np.all(lo == hi))


def mobility_groups(a, ga, b, gb):
    """Connected components of the bipartite graph of the two terms' levels (dof.rs:222-256)."""
    edges = coo_matrix((np.ones(len(a)), (a.astype(np.int64), ga + b.astype(np.int64))), shape=(ga + gb, ga + gb))
    ncomp, _ = connected_components(edges, directed=False)
    return ncomp


def estimate_dof(terms, clusters, n, firstpair=False, pairwise=True, use_clusters=True):
    """terms: [(codes, g)] (all with intercepts); clusters: [codes] of a cluster VCE.
    Returns dict(slots=[[K, M, exact, nested]], df_a, df_a_initial, df_a_redundant, df_a_nested)."""
    if not terms:
        terms = [(np.zeros(n, np.uint32), 1)]
    slots = [[g, 0, False, False] for _, g in terms]
    active = [True] * len(terms)
    nested_tot = 0
    if clusters and use_clusters:
        for c in clusters:
            for i, (codes, g) in enumerate(terms):
                if not active[i] or slots[i][2]:
                    continue
                if nested_within(codes, c):
                    slots[i][1] = g
                    slots[i][2] = slots[i][3] = True
                    nested_tot += g
                    active[i] = False
    remaining = sum(active)
    if remaining > 1 or (remaining > 0 and nested_tot > 0):
        skip = nested_tot == 0
        for i in range(len(terms)):
            if not active[i]:
                continue
            if skip:
                skip = False
                continue
            slots[i][1] = 1
    first = next((i for i in range(len(terms)) if active[i]), None)
    if first is not None:
        slots[first][2] = True
    if pairwise or firstpair:
        idx = [i for i in range(len(terms)) if active[i]]
        pairs = 0
        done = False
        for ia in range(len(idx)):
            for ib in range(ia + 1, len(idx)):
                qa, qb = idx[ia], idx[ib]
                pairs += 1
                m = mobility_groups(terms[qa][0], terms[qa][1], terms[qb][0], terms[qb][1])
                slots[qb][1] = max(slots[qb][1], m)
                if pairs == 1:
                    slots[qb][2] = True
                if firstpair:
                    done = True
                    break
            if done:
                break
    ini = sum(s[0] for s in slots)
    red = sum(s[1] for s in slots)
    return {"slots": slots, "df_a": ini - red, "df_a_initial": ini, "df_a_redundant": red, "df_a_nested": nested_tot}


def dof_table(d):
    return [v for s in d["slots"] for v in (s[0], s[1], s[0] - s[1], 0 if s[2] else 1, 1 if s[3] else 0)]


def _inv_ext(inv, means, n_c):
    """vcov.rs:215-236."""
    k = len(means)
    out = np.zeros((k + 1, k + 1))
    side = -(means @ inv)
    out[:k, :k] = inv
    out[:k, k] = side
    out[k, :k] = side
    out[k, k] = 1.0 / n_c - side @ means
    return out


def _clip_psd(v):
    lam, U = np.linalg.eigh(v)
    if not np.any(lam < 0):
        return None
    out = (U * np.maximum(lam, 0.0)) @ U.T
    return np.tril(out) + np.tril(out, -1).T


def fix_psd(v, report_constant):
    """vcov.rs:689-712 (Cameron, Gelbach and Miller)."""
    v = 0.5 * (v + v.T)
    k = v.shape[0]
    kb = k - 1 if report_constant else k
    backup = v[:kb, :kb].copy()
    fixed = _clip_psd(v)
    if fixed is None:
        return v, False
    v = fixed
    if report_constant:
        sub = _clip_psd(backup)
        v[:kb, :kb] = backup if sub is None else sub
    return v, True


def _cluster_meat(A, s, codes):
    _, dense = np.unique(codes, return_inverse=True)
    dense = dense.reshape(-1)
    G = int(dense.max()) + 1
    sg = np.zeros((G, A.shape[1]))
    np.add.at(sg, dense, A * s[:, None])
    return sg.T @ sg, G


def fit(y, X, terms, weights=None, weight_kind=None, vce="unadjusted", clusters=(), tol=1e-14, noconstant=False,
        num_singletons=0, dof_opts=(False, True, True, True)):
    """reghdfe OLS on the kept sample. terms: [(codes, g)]; clusters: [codes] (kept order).
    Returns dict(b, V, status, dof_table, scalars)."""
    y = np.asarray(y, float)
    X = np.asarray(X, float).reshape(len(y), -1)
    n_obs, nx = X.shape
    r = 1 + nx
    has_intercept = True  # every absorbed term here has an intercept; no terms absorb _cons
    report_constant = has_intercept and not noconstant
    w = None if weights is None else np.asarray(weights, float)
    sumw = float(np.sum(w)) if w is not None else float(n_obs)
    big_n = sumw if weight_kind == "fweight" else float(n_obs)
    if weight_kind in ("aweight", "pweight"):
        wv = w * (n_obs / sumw)
    elif weight_kind == "fweight":
        wv = w.copy()
    else:
        wv = None
    if vce == "unadjusted" and weight_kind == "pweight":
        vce = "robust"
    firstpair, pairwise, use_cl, _cont = dof_opts
    dof = estimate_dof(terms, list(clusters) if vce == "cluster" else [], n_obs, firstpair, pairwise, use_cl)

    raw = np.concatenate([y[:, None], X], axis=1)
    # column statistics (stats.rs:40-44, 154-184)
    sumsq = np.sum(raw * raw, axis=0)
    sd = np.maximum(np.std(raw, axis=0, ddof=1), 1e-3)
    wr = np.ones(n_obs) if w is None else w
    wmean = (wr @ raw) / np.sum(wr)
    wss = np.sum(wr[:, None] * (raw - wmean) ** 2, axis=0)
    wss_unc = np.sum(wr[:, None] * raw * raw, axis=0)
    scale = n_obs / sumw if weight_kind in ("aweight", "pweight") else 1.0
    tss_all = (wss if has_intercept else wss_unc) * scale

    # residualize every column (a model without terms absorbs a one-level constant)
    tcodes = [c for c, _ in terms] or [np.zeros(n_obs, np.uint32)]
    tg = [g for _, g in terms] or [1]
    s = OracleSolver(tcodes, tg, weights=w)
    pi = s.structure()["pi"]
    rows = np.ascontiguousarray(raw[pi])
    s.residualize(rows, r, tol=tol, maxiter=100000)
    z = np.empty_like(rows)
    z[pi] = rows
    wvv = np.ones(n_obs) if wv is None else wv
    g = (z * wvv[:, None]).T @ z
    ss_res = np.sum(z * z, axis=0)

    ctol = min(1e-6, tol / 10.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        absorbed = ss_res / sumsq <= ctol
    status = np.zeros(nx, np.int64)
    y_absorbed = nx >= 1 and bool(absorbed[0])
    if y_absorbed:
        z[:, 0] = 0.0
        g[0, :] = 0.0
        g[:, 0] = 0.0
    status[absorbed[1:]] = 2
    cand = [j for j in range(1, r) if status[j - 1] == 0]
    gs = g[np.ix_(cand, cand)] / np.outer(sd[cand], sd[cand])
    inv_c, drop = invsym(gs) if cand else (np.zeros((0, 0)), np.zeros(0, bool))
    for c, j in enumerate(cand):
        if drop[c]:
            status[j - 1] = 3
    keep_pos = [c for c in range(len(cand)) if not drop[c]]
    kept = [cand[c] for c in keep_pos]
    k = len(kept)
    inv_s = inv_c[np.ix_(keep_pos, keep_pos)]
    sd_y = sd[0]
    xy_s = np.array([g[j, 0] / (sd[j] * sd_y) for j in kept])
    b_s = inv_s @ xy_s if k else np.zeros(0)
    beta = b_s * sd_y / sd[kept] if k else np.zeros(0)
    e = z[:, 0] - (z[:, kept] @ beta if k else 0.0)
    rss = float(np.sum(wvv * e * e))
    rss_s = rss / (sd_y * sd_y)

    means_s = wmean[kept] / sd[kept]
    ybar_s = wmean[0] / sd_y
    b0_s = ybar_s - means_s @ b_s
    msize = k + (1 if report_constant else 0)
    n_c = float(n_obs) if weight_kind in ("aweight", "pweight") else sumw
    bread = _inv_ext(inv_s, means_s, n_c) if report_constant else inv_s
    b_full_s = np.concatenate([b_s, [b0_s]]) if report_constant else b_s

    df_a = float(dof["df_a"])
    df_m = float(k)
    df_r = big_n - df_a - df_m
    xbar = wmean[kept]
    A = z[:, kept] + (xbar if report_constant else 0.0)
    if report_constant:
        A = np.concatenate([A, np.ones((n_obs, 1))], axis=1)
    sc = np.concatenate([sd[kept], [1.0] if report_constant else []])
    std = np.outer(sc, sc) * sd_y * sd_y
    gmin = None
    gl = []
    if vce == "cluster":
        gl = [len(np.unique(c)) for c in clusters]
        gmin = float(min(gl))
    nested_adj = 1.0 if dof["df_a_nested"] > 0 else 0.0
    v_s = np.zeros((msize, msize))
    if msize:
        if vce == "unadjusted":
            v_s = bread * (rss_s / big_n) * (big_n / df_r)
        elif vce == "robust":
            if weight_kind is None:
                mult = e * e
            elif weight_kind == "fweight":
                mult = e * e * wvv
            else:
                mult = (e * wvv) ** 2
            meat = (A * mult[:, None]).T @ A / std
            v_s = bread @ meat @ bread
            v_s = 0.5 * (v_s + v_s.T) * (big_n / (big_n - df_a - df_m))
        elif vce == "cluster":
            sco = e * wvv
            meat = np.zeros((msize, msize))
            q = len(clusters)
            for size in range(1, q + 1):
                sign = 1.0 if size % 2 else -1.0
                for sub in itertools.combinations(range(q), size):
                    codes = clusters[sub[0]] if size == 1 else np.unique(
                        np.stack([clusters[d] for d in sub], 1), axis=0, return_inverse=True)[1].reshape(-1)
                    m, _ = _cluster_meat(A, sco, codes)
                    meat += sign * m
            meat /= std
            qc = (big_n - 1.0) / (big_n - nested_adj - df_m - df_a) * gmin / (gmin - 1.0)
            v_s = bread @ meat @ bread
            v_s = 0.5 * (v_s + v_s.T) * qc
            if q > 1:
                v_s, _ = fix_psd(v_s, report_constant)
        else:
            raise ValueError(vce)
    if vce == "cluster":
        df_r = min(df_r, gmin - 1.0)
    v_s = np.tril(v_s) + np.tril(v_s, -1).T
    f_stat = float("nan")
    if k > 0:
        inv_v, dv = invsym_greedy(v_s[:k, :k], 1e-9)
        if not dv.any():
            f_stat = float(b_s @ inv_v @ b_s) / df_m
            if not math.isfinite(f_stat):
                f_stat = float("nan")
    if np.isnan(v_s).any():
        v_s = np.zeros_like(v_s)
    sx = sc / sd_y
    b_raw = b_full_s / sx
    v_raw = v_s / np.outer(sx, sx)
    k_all = nx + (1 if report_constant else 0)
    pos = [j - 1 for j in kept] + ([nx] if report_constant else [])
    b_out = np.zeros(k_all)
    v_out = np.zeros((k_all, k_all))
    b_out[pos] = b_raw
    v_out[np.ix_(pos, pos)] = v_raw

    tss = tss_all[0]
    tss_within = g[0, 0]
    used_df_r = big_n - df_a - df_m - dof["df_a_nested"]
    ln2pi = math.log(2 * math.pi)
    hi = 1.0 if has_intercept else 0.0
    sc_out = {
        "N": big_n, "rank": df_m, "df_r": df_r, "df_m": df_m, "tss": tss, "tss_within": tss_within, "rss": rss,
        "mss": tss - rss, "rmse": math.sqrt(rss / used_df_r) if used_df_r != 0 else math.sqrt(rss),
        "F": f_stat, "ll": -0.5 * big_n * (1 + ln2pi + math.log(rss / big_n)) if rss > 0 else float("nan"),
        "ll_0": -0.5 * big_n * (1 + ln2pi + math.log(tss_within / big_n)) if tss_within > 0 else float("nan"),
        "r2": 1 - rss / tss, "r2_a": 1 - (rss / used_df_r) / (tss / (big_n - hi)),
        "r2_within": 1 - rss / tss_within if tss_within else float("nan"),
        "r2_a_within": 1 - (rss / used_df_r) / (tss_within / (used_df_r + df_m)) if tss_within else float("nan"),
        "sumweights": sumw, "converged": 1.0, "report_constant": 1.0 if report_constant else 0.0,
        "df_a": df_a, "df_a_initial": float(dof["df_a_initial"]), "df_a_redundant": float(dof["df_a_redundant"]),
        "df_a_nested": float(dof["df_a_nested"]), "num_singletons": float(num_singletons),
        "N_full": big_n + num_singletons,
    }
    if vce == "cluster":
        sc_out["N_clust"] = gmin
        sc_out["N_clustervars"] = float(len(gl))
        for i, v in enumerate(gl):
            sc_out[f"N_clust{i + 1}"] = float(v)
    return {"b": b_out, "V": v_out, "status": status, "dof_table": dof_table(dof), "scalars": sc_out,
            "y_absorbed": y_absorbed}
