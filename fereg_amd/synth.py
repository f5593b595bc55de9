"""Synthetic panels of the benchmark configurations (seeded numpy generators).

Rows are individual-major (row = T*i + t). firm_it is uniform ("simple" graph) or
row mod F ("difficult" graph). alpha, gamma, delta, eps ~ N(0,1).
  OLS:     y = sum_k beta_k x_k + alpha_i + gamma_f + delta_t + eps, beta_k = 1,
           x_1 ~ N(0,1) (K = 1) or x_k = 0.3 (alpha_i + gamma_f) + N(0,1) (K > 1)
  Poisson: count ~ Poisson(exp(0.1 x1 + 0.3 (alpha_i + gamma_f + delta_t)))
"""
import numpy as np


def panel(n, T=10, F=None, design="simple", K=1, model="ols", seed=1, shuffle=False):
    rng = np.random.default_rng(seed)
    I = n // T
    n = I * T
    F = F if F is not None else max(1, I // 10)
    row = np.arange(n, dtype=np.int64)
    indiv = (row // T).astype(np.uint32)
    year = (row % T).astype(np.uint32)
    if design == "simple":
        firm = rng.integers(0, F, n).astype(np.uint32)
    elif design == "difficult":
        firm = (row % F).astype(np.uint32)
    else:
        raise ValueError(design)
    alpha = rng.standard_normal(I)
    gamma = rng.standard_normal(F)
    delta = rng.standard_normal(T)
    fe = alpha[indiv] + gamma[firm] + delta[year]
    if K == 1:
        X = rng.standard_normal((n, 1))
    else:
        X = 0.3 * (alpha[indiv] + gamma[firm])[:, None] + rng.standard_normal((n, K))
    if model == "ols":
        y = X.sum(axis=1) + fe + rng.standard_normal(n)
    elif model == "poisson":
        y = rng.poisson(np.exp(0.1 * X[:, 0] + 0.3 * fe)).astype(np.float64)
    else:
        raise ValueError(model)
    out = {"indiv": indiv, "firm": firm, "year": year, "y": y, "X": np.ascontiguousarray(X),
           "g": [I, F, T], "n": n}
    if shuffle:
        p = rng.permutation(n)
        for k in ("indiv", "firm", "year", "y"):
            out[k] = out[k][p]
        out["X"] = np.ascontiguousarray(out["X"][p])
    return out
