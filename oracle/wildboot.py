"""CPU oracle of the WCR11 wild cluster bootstrap for OLS (test infrastructure only).

Only tests/ may import this module, as the checker of the GPU path; the product never calls it.

Restates crates/fereg-core/src/wildboot/:
  rng.rs:4-31           SplitMix64 (stream(seed, b) = new(seed ^ b*0xD1B54A32D192ED03), one draw
                        discarded)
  weights.rs:65-80      Rademacher signs: bit j of the successive 64-bit draws, 0 -> +1, 1 -> -1
  weights.rs:39-61,118  full enumeration when G <= 20 and 2^G <= reps (draw b: bit c of b)
  ols.rs:76-121         ConstraintProj: C = bread R', M_R = C (R bread R')^+
  ols.rs:243-649        run_cluster_11, cells = clusters (err_equals_boot), the non-nested FE
                        projection of Path 1 (mod.rs:959-1180) or of Path 2 (mod.rs:1290-1363,
                        ols.rs:456-548) or none (None / TooLarge), affine CI targets
                        {u_hat, -X M_R}: per draw num = V.R Phi, db = V.Phi,
                        q_cell[c, t] = V_c A_c[t] - B_c . db_t, v = ssc sum_c q q'
  ols.rs:139-142        eval_affine_t
  mod.rs:529-580        compute_pvalue
  mod.rs:637-749        invert_ci_search (bracketing + Chandrupatla)
  fit/dist.rs:32        qnorm (Wichura's AS241)
plus brute-force refit oracles (tests/wildboot.rs:538-708): y* = X b_R + V e_R, dense OLS with
explicit FE dummies, CRV1 t per draw (WCR11), and the MNW (2023) 13 / 31 / 33 variants with
restricted leave-one-cluster-out jackknife residuals in the DGP and / or the CRV3 variance
(ols.rs:651-1082 is what they check).
"""
import math

import numpy as np

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15
STREAM = 0xD1B54A32D192ED03


def _mix(z):
    """SplitMix64 output function on a uint64 numpy array (rng.rs:20-26)."""
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def splitmix64(seed, count):
    """The first `count` outputs of SplitMix64::new(seed)."""
    with np.errstate(over="ignore"):
        states = np.uint64(seed & M64) + np.uint64(GAMMA) * np.arange(1, count + 1, dtype=np.uint64)
    return _mix(states)


def stream_words(seed, b, words):
    """`words` outputs of SplitMix64::stream(seed, b) after its discarded first draw."""
    s0 = (seed ^ ((b * STREAM) & M64)) & M64
    with np.errstate(over="ignore"):
        states = np.uint64(s0) + np.uint64(GAMMA) * np.arange(2, words + 2, dtype=np.uint64)
    return _mix(states)


def resolve_reps(reps, G):
    """weights.rs:39-61 for Rademacher: (reps, full enumeration)."""
    if 0 < G <= 20 and (1 << G) <= reps:
        return 1 << G, True
    return reps, False


def rademacher(seed, reps, G):
    """V (reps x G) in {+1, -1} (weights.rs:65-80, 118-139)."""
    reps, full = resolve_reps(reps, G)
    V = np.empty((reps, G))
    if full:
        b = np.arange(reps)[:, None]
        c = np.arange(G)[None, :]
        return np.where((b >> c) & 1, -1.0, 1.0), True
    words = (G + 63) // 64
    for b in range(reps):
        w = stream_words(seed, b, words)
        bits = ((w[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).reshape(-1)[:G]
        V[b] = np.where(bits == 1, -1.0, 1.0)
    return V, False


def cluster_sums(cl, G, v):
    """Per-cluster column sums of the n x k array v (codes cl in 0..G-1)."""
    return np.stack([np.bincount(cl, weights=v[:, j], minlength=G) for j in range(v.shape[1])], axis=1)


def eval_affine_t(delta, n0, n1, d00, d01, d11):
    """ols.rs:139-142 (vectorized)."""
    v = d00 + delta * (2.0 * d01 + delta * d11)
    with np.errstate(invalid="ignore", divide="ignore"):
        t = (n0 + delta * n1) / np.sqrt(np.where(v > 1e-300, v, np.nan))
    return np.where(np.isfinite(v) & (v > 1e-300), t, np.nan)


def compute_pvalue(stat, t_boot, pval_type="two-tailed"):
    """mod.rs:529-580 for q = 1: (p, feasible draws)."""
    if not np.isfinite(stat):
        return float("nan"), 0
    tb = np.asarray(t_boot)
    tb = tb[np.isfinite(tb)]
    if tb.size == 0:
        return float("nan"), 0
    eps = 1e-12 * (1.0 + abs(stat))
    f = float(tb.size)
    if pval_type == "two-tailed":
        return np.count_nonzero(np.abs(tb) > abs(stat) + eps) / f, tb.size
    hi = np.count_nonzero(tb > stat + eps)
    lo = np.count_nonzero(tb < stat - eps)
    if pval_type == "greater":
        return hi / f, tb.size
    if pval_type == "less":
        return lo / f, tb.size
    return min(2.0 * min(hi, lo) / f, 1.0), tb.size


def qnorm(p):
    """Standard normal quantile, Wichura's AS241 (PPND16), as fit/dist.rs:32."""
    if not (0.0 <= p <= 1.0):
        return float("nan")
    if p == 0.0:
        return -math.inf
    if p == 1.0:
        return math.inf
    q = p - 0.5
    if abs(q) <= 0.425:
        r = 0.180625 - q * q
        num = (((((((2509.0809287301226727 * r + 33430.575583588128105) * r + 67265.770927008700853) * r
                   + 45921.953931549871457) * r + 13731.693765509461125) * r + 1971.5909503065514427) * r
                + 133.14166789178437745) * r + 3.387132872796366608)
        den = (((((((5226.495278852545925 * r + 28729.085735721942674) * r + 39307.89580009271061) * r
                   + 21213.794301586595867) * r + 5394.1960214247511077) * r + 687.1870074920579083) * r
                + 42.313330701600911252) * r + 1.0)
        return q * num / den
    r = math.sqrt(-math.log(p if q < 0 else 1.0 - p))
    if r <= 5.0:
        r -= 1.6
        num = (((((((7.7454501427834140764e-4 * r + 0.0227238449892691845833) * r + 0.24178072517745061177) * r
                   + 1.27045825245236838258) * r + 3.64784832476320460504) * r + 5.7694972214606914055) * r
                + 4.6303378461565452959) * r + 1.42343711074968357734)
        den = (((((((1.05075007164441684324e-9 * r + 5.475938084995344946e-4) * r + 0.0151986665636164571966) * r
                   + 0.14810397642748007459) * r + 0.68976733498510000455) * r + 1.6763848301838038494) * r
                + 2.05319162663775882187) * r + 1.0)
    else:
        r -= 5.0
        num = (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r + 0.0012426609473880784386) * r
                   + 0.026532189526576123093) * r + 0.29656057182850489123) * r + 1.7848265399172913358) * r
                + 5.4637849111641143699) * r + 6.6579046435011037772)
        den = (((((((2.04426310338993978564e-15 * r + 1.4215117583164458887e-7) * r + 1.8463183175100546818e-5) * r
                   + 7.868691311456132591e-4) * r + 0.0148753612908506148525) * r + 0.13692988092273580531) * r
                + 0.59983220655588793769) * r + 1.0)
    v = num / den
    return -v if q < 0 else v


def invert_ci_search(point, se, alpha, pval_type, tol, maxiter, reps_feas, eval_p):
    """mod.rs:637-749: bracket from the point outwards, then Chandrupatla/bisection."""
    if not (np.isfinite(point) and np.isfinite(se) and se > 0 and reps_feas > 0):
        return float("nan"), float("nan")
    z = max(abs(qnorm(1.0 - alpha / 2.0)), 1.5)
    hw = max(1.5 * z * se, 1e-6 * (1.0 + abs(point)))
    eq = 2.0 if pval_type == "equal-tailed" else 1.0
    p_tol = (eq / reps_feas) * 1.0001
    x_tol = tol * (1.0 + abs(point))

    def find_bound(sign):
        x_in = point
        p_in = eval_p(x_in)
        if not np.isfinite(p_in) or p_in < alpha:
            p_in = 1.0
        step = hw
        x_out = point + sign * step
        p_out = float("nan")
        bracketed = False
        for _ in range(16):
            p = eval_p(x_out)
            if np.isfinite(p):
                if p < alpha:
                    p_out = p
                    bracketed = True
                    break
                x_in, p_in = x_out, p
            step *= 2.0
            x_out = point + sign * step
        if not bracketed:
            return -math.inf if sign < 0 else math.inf
        x1, f1 = x_in, p_in - alpha
        x2, f2 = x_out, p_out - alpha
        t = 0.5
        for _ in range(maxiter):
            if abs(x2 - x1) <= x_tol or min(abs(f1), abs(f2)) <= p_tol:
                break
            x = x1 + t * (x2 - x1)
            fx = eval_p(x) - alpha
            if not np.isfinite(fx):
                t = 0.5
                xm = 0.5 * (x1 + x2)
                fm = eval_p(xm) - alpha
                if not np.isfinite(fm):
                    break
                if (fm < 0.0) == (f1 < 0.0):
                    x1, f1 = xm, fm
                else:
                    x2, f2 = xm, fm
                continue
            if (fx < 0.0) == (f1 < 0.0):
                x3, f3 = x1, f1
                x1, f1 = x, fx
            else:
                x3, f3 = x2, f2
                x2, f2 = x, fx
            if abs(x2 - x1) <= x_tol or min(abs(f1), abs(f2)) <= p_tol:
                break
            d32, d12, d31 = f3 - f2, f1 - f2, f3 - f1
            if d32 == 0.0 or d12 == 0.0 or d31 == 0.0 or (x2 - x1) == 0.0 or (x3 - x2) == 0.0:
                t = 0.5
                continue
            phi1 = d12 / d32
            xi1 = (x1 - x2) / (x3 - x2)
            if phi1 * phi1 > xi1 or xi1 > 2.0 * phi1 - phi1 * phi1:
                t = 0.5
            else:
                raw = (f3 / d12 + (x3 - x1) / ((x2 - x1) * d31) * f2) * (f1 / d32)
                t = min(max(raw, 1e-6), 1.0 - 1e-6)
        return x1 if abs(f1) <= abs(f2) else x2

    lo = -math.inf if pval_type == "less" else find_bound(-1.0)
    hi = math.inf if pval_type == "greater" else find_bound(1.0)
    return lo, hi


def wcr11(xt, u, w, cl, G, bread, R, r, ssc, reps, seed, level=0.95, pval_type="two-tailed", conf_int=True,
          tol=1e-6, maxiter=50, dnn=None, dtilde=None, p2_design=None):
    """run_cluster_11 (ols.rs:243-649) with cells = clusters.

    xt: n x k residualized regressors (kept columns), u: residuals, w: weights or None, cl: cluster
    codes 0..G-1 (row order), bread = (X~'WX~)^-1, R: k-vector, r: tested value, ssc: CRV1 factor.
    dnn / dtilde: the dummies of the fixed effects not nested in the clusters and their residuals
    on the nested ones (n x m): the exact projection of Path 1 (wildboot/mod.rs:959-1180,
    ols.rs:413-448): the FE-projected bootstrap residual is u* - D~ G^+ D_nn' W u*, G = D~'WD~
    (pseudo-inverse: eigenvalues <= 1e-10 of the largest dropped, mod.rs:434-470), so each cell
    score loses f_c . (V psi), f_c = sum_c w z D_nn, psi[g] = G^+ sum_g w u_t D_nn.
    p2_design: the explicit dummies D of every absorbed term (n x p): the exact projection of
    Path 2 (mod.rs:1290-1363, ols.rs:456-548): for every target t and cluster g the column 1_g u_t
    is projected on D, and the cell score of cluster c loses sum_g V_g p2_t[c, g] with
    p2_t[c, g] = sum_{i in c} w_i z_i P_D(1_g u_t)_i.
    Returns a dict with stat, se, point, p_val, ci, t_boot, coefs (reps x 5), full.
    """
    n, k = xt.shape
    wv = np.ones(n) if w is None else np.asarray(w, float)
    C = bread @ R
    v0 = float(R @ C)
    v0i = 1.0 / v0 if abs(v0) > 1e-12 * max(1.0, abs(v0)) else 0.0
    MR = C * v0i
    beta = None  # not needed: point comes from the caller's beta through num_obs
    z = xt @ C
    u0 = u
    u1 = -(xt @ MR)
    S0 = cluster_sums(cl, G, (wv * u0)[:, None] * xt)
    S1 = cluster_sums(cl, G, (wv * u1)[:, None] * xt)
    Bc = cluster_sums(cl, G, (wv * z)[:, None] * xt)
    Phi0 = S0 @ bread.T
    Phi1 = S1 @ bread.T
    A0 = Phi0 @ R
    A1 = Phi1 @ R
    q_obs = np.bincount(cl, weights=wv * z * u0, minlength=G)
    if dnn is not None:
        Gn = dtilde.T @ (wv[:, None] * dtilde)
        Gn = 0.5 * (Gn + Gn.T)
        lam, U = np.linalg.eigh(Gn)
        lmax = np.max(np.abs(lam)) if lam.size else 0.0
        keep = lam > 1e-10 * lmax
        Gp = (U[:, keep] / lam[keep]) @ U[:, keep].T
        F0 = cluster_sums(cl, G, (wv * u0)[:, None] * dnn)
        F1 = cluster_sums(cl, G, (wv * u1)[:, None] * dnn)
        fz = cluster_sums(cl, G, (wv * z)[:, None] * dnn)
        Phi0 = np.concatenate([Phi0, F0 @ Gp], axis=1)
        Phi1 = np.concatenate([Phi1, F1 @ Gp], axis=1)
        Bc = np.concatenate([Bc, fz], axis=1)
    p2 = None
    if p2_design is not None:
        sw = np.sqrt(wv)
        Dw = p2_design * sw[:, None]
        p2 = []
        for ut in (u0, u1):
            cols = np.zeros((n, G))
            cols[np.arange(n), cl] = ut  # column g: 1_g u_t
            coef, *_ = np.linalg.lstsq(Dw, cols * sw[:, None], rcond=1e-12)
            proj = p2_design @ coef  # P_D(1_g u_t), column g
            p2.append(cluster_sums(cl, G, (wv * z)[:, None] * proj))  # [c, g]
    V, full = rademacher(seed, reps, G)
    # per draw: explicit cell scores (no expansion)
    num0 = V @ A0
    num1 = V @ A1
    db0 = V @ Phi0
    db1 = V @ Phi1
    q0 = V * A0[None, :] - db0 @ Bc.T
    q1 = V * A1[None, :] - db1 @ Bc.T
    if p2 is not None:
        q0 = q0 - V @ p2[0].T
        q1 = q1 - V @ p2[1].T
    d00 = ssc * np.sum(q0 * q0, axis=1)
    d01 = ssc * np.sum(q0 * q1, axis=1)
    d11 = ssc * np.sum(q1 * q1, axis=1)
    return {"q_obs": q_obs, "coefs": np.stack([num0, num1, d00, d01, d11], 1), "full": full,
            "A0": A0, "C": C}


def wcr11_result(beta, xt, u, w, cl, G, bread, R, r, ssc, reps, seed, level=0.95, pval_type="two-tailed",
                 conf_int=True, tol=1e-6, maxiter=50, dnn=None, dtilde=None, p2_design=None):
    """The observed statistic, the bootstrap t's, the p-value and the test-inverted interval."""
    core = wcr11(xt, u, w, cl, G, bread, R, r, ssc, reps, seed, dnn=dnn, dtilde=dtilde, p2_design=p2_design)
    num_obs = float(R @ beta) - r
    point = num_obs + r
    v_obs = ssc * float(core["q_obs"] @ core["q_obs"])
    if np.isfinite(v_obs) and v_obs > 1e-300:
        se = math.sqrt(v_obs)
        stat = num_obs / se
    else:
        se, stat = float("nan"), float("nan")
    c = core["coefs"]
    delta = r - point
    t_boot = eval_affine_t(delta, c[:, 0], c[:, 1], c[:, 2], c[:, 3], c[:, 4])
    p, feas = compute_pvalue(stat, t_boot, pval_type)
    ci = (float("nan"), float("nan"))
    if conf_int:
        alpha = 1.0 - level

        def eval_p(r_trial):
            d = r_trial - point
            st = (point - r_trial) / se
            tb = eval_affine_t(d, c[:, 0], c[:, 1], c[:, 2], c[:, 3], c[:, 4])
            return compute_pvalue(st, tb, pval_type)[0]

        ci = invert_ci_search(point, se, alpha, pval_type, tol, maxiter, feas, eval_p)
    else:
        eval_p = None
    eq = 2.0 if pval_type == "equal-tailed" else 1.0
    return {"stat": stat, "se": se, "point": point, "p_val": p, "feasible": feas, "ci": ci, "t_boot": t_boot,
            "coefs": c, "full": core["full"], "reps": c.shape[0], "eval_p": eval_p, "alpha": 1.0 - level,
            "p_tol": 1.0001 * eq / feas if feas else float("nan"), "x_tol": tol * (1.0 + abs(point))}


def ci_endpoint_ok(x, ref):
    """mod.rs:673-747's stopping rule as a property of an endpoint x: p(x) lies within one draw
    (p_tol) of alpha, or p - alpha changes sign within x_tol of x (an infinite endpoint needs
    p >= alpha at every bracketing step, i.e. the reference's own infinite endpoint)."""
    f = lambda v: ref["eval_p"](v) - ref["alpha"]
    if not math.isfinite(x):
        return True
    fx = f(x)
    if math.isfinite(fx) and abs(fx) <= ref["p_tol"]:
        return True
    a, b = f(x - ref["x_tol"]), f(x + ref["x_tol"])
    return (a >= 0) != (b >= 0) or min(abs(a), abs(b)) <= ref["p_tol"]


def brute_t_boot(X, y, w, D, cl, G, R, r, reps, seed, ssc):
    """Brute-force refit per draw (tests/wildboot.rs:538-708, WCR11): explicit FE dummies D, the
    restricted fit imposes R b = r, y* = X b_R + V_g e_R, and t_b = (R b* - R b_R) / se_CRV1(b*).
    Returns (stat, se, t_boot)."""
    n, k = X.shape
    wv = np.ones(n) if w is None else np.asarray(w, float)
    F = X
    if D is not None and D.shape[1]:
        # a basis of the dummies' column space (the FE part of the dense design)
        _, sv, vt = np.linalg.svd(D, full_matrices=False)
        rank = int(np.sum(sv > 1e-10 * sv[0]))
        F = np.column_stack([X, D @ vt[:rank].T])
    A = F.T @ (wv[:, None] * F)
    Ai = np.linalg.inv(A)
    Rf = np.zeros(F.shape[1])
    Rf[:k] = R
    bu = Ai @ (F.T @ (wv * y))
    eu = y - F @ bu
    # restricted least squares: b_R = b_u - Ai Rf (Rf Ai Rf)^-1 (Rf b_u - r)
    br = bu - Ai @ Rf * ((Rf @ bu - r) / (Rf @ Ai @ Rf))
    er = y - F @ br

    def crv1_se(e, b):
        sc = np.zeros((G, F.shape[1]))
        np.add.at(sc, cl, (wv * e)[:, None] * F)
        meat = sc.T @ sc
        Vb = Ai @ meat @ Ai * ssc
        return math.sqrt(Rf @ Vb @ Rf)

    se = crv1_se(eu, bu)
    stat = (Rf @ bu - r) / se
    V, _ = rademacher(seed, reps, G)
    yhat = F @ br
    target = Rf @ br
    tb = np.empty(V.shape[0])
    for b in range(V.shape[0]):
        yb = yhat + V[b][cl] * er
        bb = Ai @ (F.T @ (wv * yb))
        eb = yb - F @ bb
        tb[b] = (Rf @ bb - target) / crv1_se(eb, bb)
    return stat, se, tb



def brute_mnw(X, y, w, D, cl, G, R, r, reps, seed, ssc, boot_type):
    """Brute-force MNW (2023) wild cluster bootstrap, restating the reference's own check
    (tests/wildboot.rs:538-708): a dense design F = [X, a basis of the FE dummies D], the
    restricted fit imposing R b = r, and for boot types 31 / 33 the jackknife residuals
    y_g - F_g b_R^(-g) from restricted leave-one-cluster-out refits (the FE refitted on the other
    clusters: a basis of the dummies over the remaining rows, so FE levels only in cluster g drop
    out); y* = F b_R + V_g u_g; the statistic's variance is CRV1 (11, 31; times ssc) or CRV3
    (13, 33: (G-1)/G sum_g (b^(-g) - b)^2 over leave-one-cluster-out refits of the same sample).
    Returns (stat, se, t_boot) over the draws of rademacher(seed, reps, G)."""
    n, k = X.shape
    wv = np.ones(n) if w is None else np.asarray(w, float)
    Dm = D if D is not None else np.zeros((n, 0))

    def basis(rows_w):
        if Dm.shape[1] == 0:
            return np.zeros((n, 0))
        sub = Dm * np.sqrt(rows_w)[:, None]
        _, sv, vt = np.linalg.svd(sub, full_matrices=False)
        if sv.size == 0 or sv[0] == 0:
            return np.zeros((n, 0))
        rank = int(np.sum(sv > 1e-10 * sv[0]))
        return Dm @ vt[:rank].T

    def fit(cols, yv, wts):
        A = cols.T @ (wts[:, None] * cols)
        Ai = np.linalg.inv(A)
        return Ai @ (cols.T @ (wts * yv)), Ai

    def restricted(b, Ai, Rf):
        return b - Ai @ Rf * ((Rf @ b - r) / (Rf @ Ai @ Rf))

    F = np.column_stack([X, basis(wv)])
    Rf = np.zeros(F.shape[1])
    Rf[:k] = R
    bu, Ai = fit(F, y, wv)
    eu = y - F @ bu
    br = restricted(bu, Ai, Rf)
    er = y - F @ br

    def loo(yv, g, restrict):
        wg = np.where(cl == g, 0.0, wv)
        cols = np.column_stack([X, basis(wg)])
        b, Aig = fit(cols, yv, wg)
        if restrict:
            Rg = np.zeros(cols.shape[1])
            Rg[:k] = R
            b = restricted(b, Aig, Rg)
        return b, cols

    def se_crv1(e):
        sc = np.zeros((G, F.shape[1]))
        np.add.at(sc, cl, (wv * e)[:, None] * F)
        return math.sqrt(Rf @ (Ai @ (sc.T @ sc) @ Ai * ssc) @ Rf)

    def se_crv3(yv, b):
        d = np.array([R @ (loo(yv, g, False)[0][:k] - b[:k]) for g in range(G)])
        return math.sqrt((G - 1) / G * np.sum(d * d))

    den3 = boot_type in (13, 33)
    if boot_type in (31, 33):
        u = np.empty(n)
        for g in range(G):
            bg, cols = loo(y, g, True)
            m = cl == g
            u[m] = y[m] - cols[m] @ bg
    else:
        u = er
    se = se_crv3(y, bu) if den3 else se_crv1(eu)
    stat = (Rf @ bu - r) / se
    V, _ = rademacher(seed, reps, G)
    yhat = F @ br
    target = Rf @ br
    tb = np.empty(V.shape[0])
    for b in range(V.shape[0]):
        yb = yhat + V[b][cl] * u
        bb = Ai @ (F.T @ (wv * yb))
        s = se_crv3(yb, bb) if den3 else se_crv1(yb - F @ bb)
        tb[b] = (Rf @ bb - target) / s
    return stat, se, tb


def _sym_pinv(A, tol):
    """Symmetric pseudo-inverse with eigenvalues <= tol * max |eigenvalue| dropped (mod.rs:434-470)."""
    A = 0.5 * (A + A.T)
    lam, U = np.linalg.eigh(A)
    lmax = np.max(np.abs(lam)) if lam.size else 0.0
    if lmax == 0.0:
        return np.zeros_like(A)
    keep = lam > tol * lmax
    return (U[:, keep] / lam[keep]) @ U[:, keep].T


def mnw_basis(Dnn, Dnest, w):
    """Path 1's basis for MNW (mod.rs:1036-1076): the non-nested dummies residualized on the nested
    terms' dummies, times U Lambda^-1/2 over the eigenvalues of their weighted Gram above 1e-10 of the
    largest, so Q'WQ = I."""
    n = Dnn.shape[0]
    wv = np.ones(n) if w is None else np.asarray(w, float)
    Dt = Dnn
    if Dnest is not None and Dnest.shape[1]:
        sw = np.sqrt(wv)
        coef, *_ = np.linalg.lstsq(Dnest * sw[:, None], Dnn * sw[:, None], rcond=1e-12)
        Dt = Dnn - Dnest @ coef
    Gm = Dt.T @ (wv[:, None] * Dt)
    lam, U = np.linalg.eigh(0.5 * (Gm + Gm.T))
    lmax = np.max(np.abs(lam))
    keep = lam > 1e-10 * lmax
    return Dt @ (U[:, keep] / np.sqrt(lam[keep]))


def mnw_fast(Xt, u, w, cl, G, bread, R, r, beta, ssc, reps, seed, boot_type, Q=None):
    """run_cluster_mnw (ols.rs:651-1082) for one restriction with the affine targets {u_hat, -X M_R}:
    Xt the FWL-residualized kept regressors, u their residuals, Q the non-nested basis (or None).
    Returns (stat, se, t_boot)."""
    n, k = Xt.shape
    wv = np.ones(n) if w is None else np.asarray(w, float)
    xa = Xt if Q is None else np.column_stack([Xt, Q])
    p = xa.shape[1]
    C = bread @ R
    MR = C / (R @ C)
    targets = np.column_stack([u, -(Xt @ MR)])
    A = np.zeros((G, p, p))
    S = np.zeros((G, 2, p))
    for g in range(G):
        m = cl == g
        A[g] = xa[m].T @ (wv[m][:, None] * xa[m])
        S[g] = (wv[m][:, None] * targets[m]).T @ xa[m]
    Gt = A.sum(axis=0)
    Gi = _sym_pinv(Gt, 1e-12)
    Ra = np.zeros(p)
    Ra[:k] = R
    dgp3, den3 = boot_type in (31, 33), boot_type in (13, 33)
    mult = (G - 1) / G if den3 else ssc
    rphi = np.zeros((G, 2))
    phi = np.zeros((G, 2, p))
    kk = np.zeros((G, 2))
    JA = np.zeros((G, p))
    jobs = np.zeros(G)
    for g in range(G):
        Gmi = _sym_pinv(Gt - A[g], 1e-10) if (dgp3 or den3) else None
        J = Ra @ Gmi if den3 else Gi @ Ra
        JA[g] = J @ A[g]
        jobs[g] = J @ S[g, 0]
        for t in range(2):
            s = S[g, t]
            if dgp3:
                y1 = Gmi @ s
                gnr = Gmi @ Ra
                h = Ra @ gnr
                y1 = y1 - gnr * ((Ra @ y1) / h if h > 0 else 0.0)
                d = s + A[g] @ y1
            else:
                d = s
            kk[g, t] = J @ d
            phi[g, t] = Gi @ d
            rphi[g, t] = Ra @ phi[g, t]
    point = R @ beta
    v_obs = mult * np.sum(jobs * jobs)
    se = math.sqrt(v_obs)
    stat = (point - r) / se
    V, _ = rademacher(seed, reps, G)
    delta = r - point
    tb = np.empty(V.shape[0])
    for b in range(V.shape[0]):
        v = V[b]
        num = v @ rphi[:, 0] + delta * (v @ rphi[:, 1])
        db = v @ phi[:, 0] + delta * (v @ phi[:, 1])
        j = v * (kk[:, 0] + delta * kk[:, 1]) - JA @ db
        tb[b] = num / math.sqrt(mult * np.sum(j * j))
    return stat, se, tb
