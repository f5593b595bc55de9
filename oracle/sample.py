"""numpy restatement of factor encoding and the estimation sample.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU factor layer and sample
builder; the product path never imports it.

Follows fereg-core:
  f64_key            factor.rs:24-29         order-preserving key; -0.0 and 0.0 share one
  factorize_col      factor.rs:307-366       dense rank of keys; missing (non-finite, INT_MIN,
                                             codes out of range) key 0, sorting first
  factorize (strict) factor.rs:44-76         a NaN is an error, infinities are levels
  combine            factor.rs:259-281       key = fold(acc * g_p + code_p), dense rank
  restrict           factor.rs:285-305       codes of the kept rows, densely renumbered
  drop_singletons    sample.rs:44-74         cyclic over terms until none changes; counts are
                                             weight sums with fweights; Poisson / logit
                                             separation (sample.rs:77-83)
  scan               fit/input.rs:193-289    missing-value and zero-weight scan, per-group counts
  build              fit/input.rs:301-345    factorize, interactions, singletons, kept rows
"""
import numpy as np

INT32_MIN = np.iinfo(np.int32).min
INT64_MIN = np.iinfo(np.int64).min


def f64_key(x):
    x = np.asarray(x, np.float64).copy()
    x[x == 0.0] = 0.0
    b = x.view(np.uint64)
    neg = (b >> np.uint64(63)) == 1
    return np.where(neg, ~b, b | np.uint64(1 << 63))


def column_keys(values, kind="f64", strict=False, base=0, nlevels=0):
    """Keys of a factor column and its missing mask (factor.rs:307-366)."""
    if kind == "f64":
        x = np.asarray(values, np.float64)
        miss = np.isnan(x) if strict else ~np.isfinite(x)
        k = np.where(miss, np.uint64(0), f64_key(np.where(miss, 0.0, x)))
    elif kind in ("i32", "i64"):
        x = np.asarray(values, np.int64)
        miss = x == (INT32_MIN if kind == "i32" else INT64_MIN)
        k = np.where(miss, np.uint64(0), (x ^ np.int64(INT64_MIN)).view(np.uint64))
    else:
        lv = np.asarray(values, np.int64) - int(base)
        miss = (lv < 0) | (lv >= int(nlevels))
        k = np.where(miss, np.uint64(0), (lv + 1).astype(np.uint64))
    return k, miss


def dense_rank(keys):
    u, inv = np.unique(np.asarray(keys, np.uint64), return_inverse=True)
    return inv.reshape(-1).astype(np.uint32), len(u)


def factorize(values, kind="f64", strict=False, base=0, nlevels=0):
    """(codes, levels, missing rows); strict raises on a NaN."""
    k, miss = column_keys(values, kind, strict, base, nlevels)
    if strict and miss.any():
        raise ValueError("missing values in factor")
    c, g = dense_rank(k)
    return c, g, int(miss.sum())


def combine(parts, g):
    radix = 1
    for gq in g:
        radix *= max(int(gq), 1)
        if radix >= 1 << 63:
            raise ValueError("interaction has too many levels")
    key = np.zeros(len(parts[0]), np.uint64)
    for c, gq in zip(parts, g):
        key = key * np.uint64(gq) + np.asarray(c, np.uint64)
    return dense_rank(key)


def restrict(codes, kept):
    return dense_rank(np.asarray(codes, np.uint64)[np.asarray(kept, np.int64)])


def _separated(family, ysum, count):
    if family == "poisson":
        return ysum == 0.0
    if family == "logit":
        return (ysum == 0.0) | (ysum == count)
    return np.zeros_like(ysum, bool)


def drop_singletons(terms, alive, fw=None, y=None, family="gaussian", singletons=True):
    """sample.rs:44-74 over terms = [(codes, g)]: alive (bool) is updated in place; returns
    (singletons, separated)."""
    nt = len(terms)
    s_tot = f_tot = 0
    if nt == 0 or (not singletons and y is None):
        return 0, 0
    remaining, i = nt, 0
    while remaining > 0:
        codes, g = terms[i % nt]
        i += 1
        live = alive
        w = np.ones(len(codes)) if fw is None else np.asarray(fw, float)
        cnt = np.bincount(codes[live], weights=w[live], minlength=g)
        drop_s = np.zeros(len(codes), bool)
        drop_f = np.zeros(len(codes), bool)
        if singletons:
            drop_s = live & (cnt[codes] == 1.0)
        if y is not None:
            yy = (np.asarray(y) != 0).astype(float) if family == "logit" else np.asarray(y, float)
            ys = np.bincount(codes[live], weights=(w * yy)[live], minlength=g)
            drop_f = live & ~drop_s & _separated(family, ys[codes], cnt[codes])
        s, f = int(drop_s.sum()), int(drop_f.sum())
        alive &= ~(drop_s | drop_f)
        s_tot += s
        f_tot += f
        remaining = nt - 1 if s + f > 0 else remaining - 1
    return s_tot, f_tot


def build(n, y=None, x=(), weights=None, offset=None, drop=None, factors=(), terms=(), clusters=(),
          weight_kind=None, family="gaussian", missing=True, singletons=True, separation=False):
    """The estimation sample. factors: list of (values, kind, base, nlevels); terms / clusters:
    lists of factor-index tuples. Returns dict(kept, terms=[(codes, g)], clusters, counts)."""
    used_fe = {f for t in terms for f in t}
    used_cl = {f for t in clusters for f in t}
    fc, fg, fmiss = {}, {}, {}
    for f in sorted(used_fe | used_cl):
        vals, kind, base, nl = factors[f]
        k, miss = column_keys(vals, kind, not missing, base, nl)
        if not missing and miss.any():
            raise ValueError("missing values in factor")
        fc[f], fg[f] = dense_rank(k)
        fmiss[f] = miss
    counts = dict(n_input=n, na=0, zero_weight=0, singletons=0, perfect_fit=0, na_y=0, na_x=0, na_fe=0,
                  na_cluster=0, na_weights=0, na_offset=0, na_extra=0)
    if missing:
        bits = {}
        bits["na_y"] = ~np.isfinite(y) if y is not None else np.zeros(n, bool)
        bits["na_x"] = np.zeros(n, bool)
        for c in x:
            bits["na_x"] |= ~np.isfinite(c)
        bits["na_fe"] = np.zeros(n, bool)
        for f in used_fe:
            bits["na_fe"] |= fmiss[f]
        bits["na_cluster"] = np.zeros(n, bool)
        for f in used_cl:
            bits["na_cluster"] |= fmiss[f]
        bits["na_weights"] = ~np.isfinite(weights) if weights is not None else np.zeros(n, bool)
        bits["na_offset"] = ~np.isfinite(offset) if offset is not None else np.zeros(n, bool)
        bits["na_extra"] = np.asarray(drop, bool) if drop is not None else np.zeros(n, bool)
        any_m = np.zeros(n, bool)
        for k, v in bits.items():
            counts[k] = int(v.sum())
            any_m |= v
        if weights is not None and np.any(np.asarray(weights) < 0):
            raise ValueError("weights must not be negative")
        zero = ~any_m & (np.asarray(weights) == 0.0) if weights is not None else np.zeros(n, bool)
        counts["na"] = int(any_m.sum())
        counts["zero_weight"] = int(zero.sum())
        alive = ~any_m & ~zero
    else:
        alive = np.ones(n, bool)
    uses = {}
    for t in list(terms) + list(clusters):
        for f in t:
            uses[f] = uses.get(f, 0) + 1

    def interact(t):
        if len(t) == 1:
            return fc[t[0]], fg[t[0]]
        return combine([fc[f] for f in t], [fg[f] for f in t])

    tc = [interact(t) for t in terms]
    cc = [interact(t) for t in clusters]
    fw = weights if weight_kind == "fweight" else None
    sep = separation and family != "gaussian"
    s, f = drop_singletons(tc, alive, fw, y if sep else None, family, singletons)
    counts["singletons"], counts["perfect_fit"] = s, f
    kept = np.flatnonzero(alive).astype(np.uint32)
    counts["nobs"] = len(kept)
    return {"kept": kept, "terms": [restrict(c, kept) for c, _ in tc], "clusters": [restrict(c, kept) for c, _ in cc],
            "counts": counts}
