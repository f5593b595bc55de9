"""Fixture and design helpers shared by the tests (no GPU needed)."""
import csv
import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_csv(name):
    with open(os.path.join(GOLDEN, name)) as f:
        rd = csv.reader(f)
        names = [h.strip('"') for h in next(rd)]
        cols = [[] for _ in names]
        for row in rd:
            if not row:
                continue
            for c, v in zip(cols, row):
                c.append(math.nan if v in ("NA", "") else float(v))
    return {n: np.array(c) for n, c in zip(names, cols)}


def cases(name):
    return json.load(open(os.path.join(GOLDEN, name)))


def fixest_ols_cases(multiway=False):
    """fixest OLS fixtures the absorber path covers: at least one absorbed term; K.exact (fixest's
    reference-level search, fit/fixef.rs:300-430) only where it cannot change df_k (one iid
    case); with multiway=False, at most one cluster dimension."""
    out = []
    for c in cases("fixest_ols.json"):
        if not c["fe"]:
            continue
        if c["ssc"].get("K.exact") and c["name"] != "fe3_iid_exact":
            continue
        if len(c["cluster"]) > 1 and not multiway:
            continue
        out.append(c)
    return out


def dense_codes(*cols):
    """Dense level codes in ascending value order; several columns = their interaction
    (lexicographic), as factor.rs:38-76 and :255."""
    keys = np.stack(cols, axis=1)
    _, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    return inv.astype(np.uint32), int(inv.max()) + 1 if len(inv) else 0


def as_list(v):
    if v is None:
        return []
    return v if isinstance(v, list) else [v]


def case_inputs(case):
    """Kept-sample inputs of a fit fixture: FE codes, y, X, weights, cluster codes."""
    d = read_csv(case["data"])
    obs = np.array(case["expect"]["obs"], dtype=np.int64) - 1
    y = d[case["y"]][obs]
    X = np.stack([d[x][obs] for x in as_list(case["x"])], axis=1)
    codes, gs = [], []
    for t in case["fe"]:
        c, g = dense_codes(*[d[f][obs] for f in t["factors"]])
        codes.append(c)
        gs.append(g)
    w = d[case["weights"]][obs] if case.get("weights") else None
    off = d[case["offset"]][obs] if case.get("offset") else None
    clusters = []
    for cl in case.get("cluster", []):
        c, _ = dense_codes(*[d[f][obs] for f in cl])
        clusters.append(c)
    fe_cols = [[d[f][obs] for f in t["factors"]] for t in case["fe"]]
    return {"y": y, "X": X, "codes": codes, "g": gs, "w": w, "offset": off, "clusters": clusters,
            "fe_cols": fe_cols}


def nested_by_value(code, cluster):
    """Every level of the FE within one cluster (fit/vcov.rs:115-140 by value)."""
    g = int(code.max()) + 1
    first = np.full(g, -1, np.int64)
    ok = True
    for c, k in zip(code, cluster):
        if first[c] < 0:
            first[c] = k
        elif first[c] != k:
            ok = False
            break
    return ok


def splitmix_codes(n, g, seed):
    """Random codes in 0..g (numpy generator; designs of structure.rs:1187-1233 at our sizes)."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, g, n).astype(np.uint32)


def dummies(codes, g):
    D = np.zeros((len(codes), g))
    D[np.arange(len(codes)), codes] = 1.0
    return D


def dense_residualize(block, codes, gs, w=None):
    """Explicit-dummy projection v - D (D'WD)^+ D'W v (reference.rs:158-199 as a least-squares
    solve): the unique weighted residual of every column."""
    n = block.shape[0]
    if not codes:
        D = np.ones((n, 1))
    else:
        D = np.concatenate([dummies(c, g) for c, g in zip(codes, gs)], axis=1)
    sw = np.sqrt(w) if w is not None else np.ones(n)
    coef, *_ = np.linalg.lstsq(D * sw[:, None], block * sw[:, None], rcond=None)
    return block - D @ coef


def slope_inputs(case):
    """case_inputs plus the absorbed terms with their slopes: terms = [(codes, g, intercept,
    [slope columns])] in the fixture's term order; X gains fixest's constant column when no term
    has level intercepts."""
    inp = case_inputs(case)
    d = read_csv(case["data"])
    obs = np.array(case["expect"]["obs"], dtype=np.int64) - 1
    inp["terms"] = [(c, g, bool(t["intercept"]), [d[s][obs] for s in t["slopes"]])
                    for c, g, t in zip(inp["codes"], inp["g"], case["fe"])]
    if not any(t[2] for t in inp["terms"]):
        # no absorbed intercept: fixest's model keeps its constant as the first regressor
        inp["X"] = np.concatenate([np.ones((len(inp["y"]), 1)), inp["X"]], axis=1)
    return inp
