"""numpy restatement of the varying-slopes absorber (FP64, explicit design).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker; nothing in fereg_amd imports it.

Follows crates/fereg-core/src/slopes.rs (SlopeSolver, slopes.rs:1-33) through the definition it
solves, not its iteration: an absorbed term q with codes c_q, intercept flag a_q and slope
columns x_q1..x_qs contributes, per level l, the columns [c_q == l] (when a_q) and
[c_q == l] * x_qk. The residual of a block V is V - D (D'WD)^+ D'W V over the stacked columns D
of every term, the projection the reference's PCG converges to (its tests compare against the
same explicit-dummy projection, tests/slopes.rs:97-140, AbsorberKind::Dense). The projection does
not depend on which generalized inverse is used, so the oracle takes the minimum-norm one.

  slope_design   the explicit columns, term order, level-major (slopes.rs:6-7: a level of width b
                 owns b consecutive coordinates: intercept, then slopes)
  residualize    V - D (D'WD)^+ D'W V (weighted least squares by SVD)
  operator       the reduced matrix S of the non-eliminated terms after eliminating E exactly
                 (slopes.rs:13-17): S = D_R'W D_R - D_R'W H_E (H_E'W H_E)^+ H_E'W D_R
  elimination    E = the intercept-only term with the most levels (absorber.rs:100-108), else the
                 term with the most levels (slopes.rs:364-372)
  nparams        fixest nparams with slopes (fit/ols.rs:25-37): every term G (a + s), minus one
                 per intercept term beyond the first
  fe_params_df   the FE parameters df.K removes (fit/vcov.rs:63-69): intercept terms only

Pinned by tests/test_slopes_oracle.py on fixest 0.14.2's varying-slope fixtures (fit_fixest_fe.json
sl_*, fit_fixest_glm.json pois_sl*: coefficients, variance, df_k, nparams, residuals), the
closed form of a single term (per-level weighted regression), and the projection identities
(D'W e = 0, idempotence) on the reference's test panels (tests/slopes.rs:28-41).
"""
import numpy as np


def term_width(intercept, slopes):
    return (1 if intercept else 0) + len(slopes)


def slope_design(terms, n):
    """terms: list of (codes, g, intercept, [slope columns]). Returns D (n x sum g b) with the
    columns of term q at offsets off_q + l b_q + k (k = 0 the intercept when present)."""
    blocks = []
    for codes, g, icpt, slopes in terms:
        codes = np.asarray(codes, dtype=np.int64)
        b = term_width(icpt, slopes)
        d = np.zeros((n, g * b))
        rows = np.arange(n)
        k = 0
        if icpt:
            d[rows, codes * b] = 1.0
            k = 1
        for s, x in enumerate(slopes):
            d[rows, codes * b + k + s] = np.asarray(x, float)
        blocks.append(d)
    return np.concatenate(blocks, axis=1) if blocks else np.zeros((n, 0))


def residualize(V, D, w=None):
    """V - D (D'WD)^+ D'W V for an n x R block (weighted least squares, minimum norm)."""
    V = np.asarray(V, float)
    one = V.ndim == 1
    if one:
        V = V[:, None]
    sw = np.sqrt(w) if w is not None else np.ones(V.shape[0])
    A = D * sw[:, None]
    B = V * sw[:, None]
    coef, *_ = np.linalg.lstsq(A, B, rcond=1e-12)
    e = V - D @ coef
    return e[:, 0] if one else e


def elimination_index(terms):
    """E: the intercept-only term with the most levels (first on ties), else the term with the
    most levels (first on ties)."""
    best = None
    for q, (_, g, icpt, slopes) in enumerate(terms):
        if icpt and not slopes and (best is None or g > terms[best][1]):
            best = q
    if best is not None:
        return best
    best = 0
    for q, t in enumerate(terms):
        if t[1] > terms[best][1]:
            best = q
    return best


def operator(terms, n, w=None, order=None):
    """The reduced matrix S over the non-eliminated terms (in `order`, default term order without
    E), each term's coordinates level-major: S = A - C' M^+ C with A = D_R'W D_R,
    M = H_E'W H_E (block diagonal over E's levels) and C = H_E'W D_R."""
    e = elimination_index(terms)
    if order is None:
        order = [q for q in range(len(terms)) if q != e]
    w = np.ones(n) if w is None else np.asarray(w, float)
    DR = slope_design([terms[q] for q in order], n)
    HE = slope_design([terms[e]], n)
    A = (DR * w[:, None]).T @ DR
    C = (HE * w[:, None]).T @ DR
    M = (HE * w[:, None]).T @ HE
    return A - C.T @ np.linalg.pinv(M, rcond=1e-13, hermitian=True) @ C


def nparams(k, terms):
    """fixest nparams: k + sum_q G_q (a_q + s_q) - max(#intercept terms - 1, 0)."""
    n_icpt = sum(1 for t in terms if t[2])
    fe = sum(t[1] * term_width(t[2], t[3]) for t in terms)
    return k + fe - max(n_icpt - 1, 0)


def df_sizes(terms):
    """Level counts df.K sees per term: G for intercept terms, 0 for slope-only ones
    (fit/vcov.rs:64-68); the FE parameters are then sum - (#nonzero - 1)."""
    return [t[1] if t[2] else 0 for t in terms]


def df_k(kind, np_, terms, nested, k_fixef="nonnested", n_design=0):
    """fixest df.K with slope terms (fit/vcov.rs:63-112, K.exact = FALSE): only the intercept
    terms' levels are FE parameters; "nonnested" applies to clustered variances only."""
    sizes = df_sizes(terms)
    n_ok = sum(1 for g in sizes if g > 0)
    fe_params = sum(sizes) - (n_ok - 1) if n_ok > 0 else 0
    kf = k_fixef
    if kind not in ("cluster", "crv3") and kf == "nonnested":
        kf = "full"
    if kf == "none":
        return np_ - fe_params
    if kf == "full" or n_ok == 0:
        return np_
    n_nested = sum(1 for x in nested if x)
    k = float(np_)
    if n_nested == len(terms):
        k = k - fe_params + 1
    else:
        s = sum(g for g, x in zip(sizes, nested) if x)
        c = sum(1 for g, x in zip(sizes, nested) if x and g > 0)
        k -= s - c
    return max(k, n_design + 1)


class DenseSlopeSolver:
    """The explicit-design projection behind the solver interface the IRLS oracle drives
    (structure / set_weights / residualize). Like the reference's SlopeSolver it reports its solves
    as not exact (slopes.rs:2419), which the IRLS convergence rule reads."""

    def __init__(self, terms, n):
        self.D = slope_design(terms, n)
        self.n = n
        self.w = None

    def structure(self):
        return {"pi": np.arange(self.n)}

    def set_weights(self, w):
        self.w = None if w is None else np.asarray(w, float).copy()

    def residualize(self, block, r, tol=1e-8, warm=None, want_coefs=False):
        block[:, :r] = residualize(block[:, :r], self.D, self.w)
        return 1, False, None
