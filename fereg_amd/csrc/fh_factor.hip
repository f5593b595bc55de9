// Factor encoding, interactions, restriction to a sample, cyclic singleton removal and the
// estimation sample on the GPU (factor.rs, sample.rs and fit/input.rs of the reference).
//
// Every encoding is a dense rank of 64-bit order-preserving keys: a row's code is the number of
// distinct keys below its own. Key ranges below max(2n, 2^20) are ranked through a presence table
// over the range and one exclusive scan of it; wider ranges by a stable radix sort of
// (key - min, row) pairs on just the bits the range needs, run-start flags and a scan. Both give
// the same codes, which depend on the multiset of keys only (as the reference's direct-table and
// sort paths do), whatever the launch shapes.
//
// Singleton removal counts each level's live rows with integer atomics (exact, so the order of
// the additions is immaterial) and drops the rows of the offending levels in the same pass that
// reads the counts; the host cycles over the terms until every term passes without a drop.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "fh_fit.hpp"
#include "fh_sort.hpp"

namespace fh {

namespace {

constexpr int FT = 256;
inline unsigned fgrid(size_t n) { return grid_for(n, FT * 4, 8192); }

// factor.rs:35-37: ranges below this are ranked through a table.
inline uint64_t table_limit(size_t n) { return std::max<uint64_t>(2 * (uint64_t)n, uint64_t(1) << 20); }

inline int bits_of(uint64_t range) { return range ? 64 - __builtin_clzll(range) : 0; }

struct KeyStats {
    unsigned long long lo, hi, missing, bad;
};

__global__ void k_stats_init(KeyStats* s) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s->lo = ~0ull;
        s->hi = 0ull;
        s->missing = 0ull;
        s->bad = 0ull;
    }
}

__device__ __forceinline__ void stats_add(unsigned long long lo, unsigned long long hi, unsigned long long miss,
                                          unsigned long long bad, KeyStats* s) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (unsigned long long)__shfl_xor(lo, o, WAVE));
        hi = max(hi, (unsigned long long)__shfl_xor(hi, o, WAVE));
        miss += __shfl_xor(miss, o, WAVE);
        bad += __shfl_xor(bad, o, WAVE);
    }
    if (lane_id() == 0) {
        atomicMin(&s->lo, lo);
        atomicMax(&s->hi, hi);
        if (miss) atomicAdd(&s->missing, miss);
        if (bad) atomicAdd(&s->bad, bad);
    }
}

// Order-preserving key of a double (factor.rs:24-29): -0.0 and 0.0 share one; no finite value
// or infinity maps to 0, the key of a missing entry.
__device__ __forceinline__ unsigned long long f64_key(double x) {
    if (x == 0.0) x = 0.0;
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

struct ColArg {
    const void* data;
    int64_t base;
    uint64_t nlevels;
};

// Keys of a factor column (factor.rs:307-366): missing entries (non-finite doubles, or NaN only
// when strict; the integer sentinels; codes outside base .. base + nlevels - 1) key 0, which sorts
// first; integers key as value ^ INT64_MIN, codes as level + 1.
template <int KIND>
__global__ void k_col_keys(ColArg c, size_t n, int strict, unsigned long long* __restrict__ keys, KeyStats* st) {
    unsigned long long lo = ~0ull, hi = 0ull, miss = 0ull;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        unsigned long long k;
        if constexpr (KIND == FH_COL_F64) {
            const double v = static_cast<const double*>(c.data)[i];
            const bool m = strict ? (v != v) : !isfinite(v);
            k = m ? 0ull : f64_key(v);
        } else if constexpr (KIND == FH_COL_I32) {
            const int32_t v = static_cast<const int32_t*>(c.data)[i];
            k = v == INT32_MIN ? 0ull : (unsigned long long)((int64_t)v ^ INT64_MIN);
        } else if constexpr (KIND == FH_COL_I64) {
            const int64_t v = static_cast<const int64_t*>(c.data)[i];
            k = v == INT64_MIN ? 0ull : (unsigned long long)(v ^ INT64_MIN);
        } else {
            const int64_t l = (int64_t)static_cast<const uint32_t*>(c.data)[i] - c.base;
            k = (l >= 0 && (uint64_t)l < c.nlevels) ? (unsigned long long)l + 1ull : 0ull;
        }
        keys[i] = k;
        miss += k == 0ull;
        lo = min(lo, k);
        hi = max(hi, k);
    }
    stats_add(lo, hi, miss, 0ull, st);
}

constexpr int COMBINE_MAX = 64;
struct Parts {
    const uint32_t* c[COMBINE_MAX];
    uint64_t g[COMBINE_MAX];
    int np;
};

// Lexicographic key of an interaction (factor.rs:259-281): sum over parts of code * radix.
__global__ void k_combine_keys(Parts p, size_t n, unsigned long long* __restrict__ keys, KeyStats* st) {
    unsigned long long lo = ~0ull, hi = 0ull, bad = 0ull;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        unsigned long long k = 0ull;
        for (int q = 0; q < p.np; q++) {
            const uint32_t v = p.c[q][i];
            bad += v >= p.g[q];
            k = k * p.g[q] + v;
        }
        keys[i] = k;
        lo = min(lo, k);
        hi = max(hi, k);
    }
    stats_add(lo, hi, 0ull, bad, st);
}

// keys[j] = codes[kept[j]] (factor.rs:285-305); out-of-range rows or codes are counted as bad.
__global__ void k_kept_keys(const uint32_t* __restrict__ codes, uint64_t g, const uint32_t* __restrict__ kept,
                            size_t nk, size_t n, unsigned long long* __restrict__ keys, KeyStats* st) {
    unsigned long long lo = ~0ull, hi = 0ull, bad = 0ull;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < nk; j += (size_t)gridDim.x * blockDim.x) {
        const uint32_t r = kept[j];
        unsigned long long k = 0ull;
        if (r >= n) {
            bad++;
        } else {
            const uint32_t c = codes[r];
            bad += c >= g;
            k = c;
        }
        keys[j] = k;
        lo = min(lo, k);
        hi = max(hi, k);
    }
    stats_add(lo, hi, 0ull, bad, st);
}

__global__ void k_mark(const unsigned long long* __restrict__ keys, size_t n, unsigned long long lo,
                       uint32_t* __restrict__ tab) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        tab[keys[i] - lo] = 1u;
}

__global__ void k_lookup(const unsigned long long* __restrict__ keys, size_t n, unsigned long long lo,
                         const uint32_t* __restrict__ rank, uint32_t* __restrict__ codes) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        codes[i] = rank[keys[i] - lo];
}

__global__ void k_shift_iota(unsigned long long* __restrict__ keys, size_t n, unsigned long long lo,
                             uint32_t* __restrict__ vals) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        keys[i] -= lo;
        vals[i] = (uint32_t)i;
    }
}

__global__ void k_run_start(const unsigned long long* __restrict__ keys, size_t n, uint32_t* __restrict__ flag) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// codes[row] = distinct keys before the sorted position's key (excl = exclusive scan of the
// run-start flags).
__global__ void k_scatter_rank(const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ vals,
                               const uint32_t* __restrict__ excl, size_t n, uint32_t* __restrict__ codes) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t start = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
        codes[vals[i]] = excl[i] + start - 1u;
    }
}

KeyStats read_stats(const KeyStats* d, hipStream_t st) {
    KeyStats h{};
    download(&h, d, sizeof(KeyStats), st);
    return h;
}

// ------------------------------------------------------------------ singleton removal

__global__ void k_level_counts(const uint32_t* __restrict__ codes, const uint8_t* __restrict__ alive, size_t n,
                               const double* __restrict__ y, uint32_t* __restrict__ count, uint32_t* __restrict__ pos) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (!alive[i]) continue;
        const uint32_t c = codes[i];
        atomicAdd(&count[c], 1u);
        if (y && y[i] != 0.0) atomicAdd(&pos[c], 1u);
    }
}

// Weighted counts and outcome sums (sample.rs:236-266): frequency weights, or a Poisson outcome
// with negative values. Sums of integer-valued weights and outcomes are exact in any order.
__global__ void k_level_sums(const uint32_t* __restrict__ codes, const uint8_t* __restrict__ alive, size_t n,
                             const double* __restrict__ fw, const double* __restrict__ y, int logit,
                             double* __restrict__ count, double* __restrict__ ysum) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (!alive[i]) continue;
        const uint32_t c = codes[i];
        const double w = fw ? fw[i] : 1.0;
        atomicAdd(&count[c], w);
        if (y) atomicAdd(&ysum[c], w * (logit ? (y[i] != 0.0 ? 1.0 : 0.0) : y[i]));
    }
}

// sample.rs:77-83
__device__ __forceinline__ bool separated(int family, double ysum, double count) {
    if (family == FH_FAMILY_POISSON) return ysum == 0.0;
    if (family == FH_FAMILY_LOGIT) return ysum == 0.0 || ysum == count;
    return false;
}

// Drops the live rows of singleton (verdict 1) and separated (verdict 2) levels; drops[0] and
// drops[1] count them (sample.rs:159-205).
template <class C>
__global__ void k_drop_rows(const uint32_t* __restrict__ codes, uint8_t* __restrict__ alive, size_t n,
                            const C* __restrict__ count, const C* __restrict__ pos, int singletons, int family,
                            unsigned long long* drops) {
    unsigned long long s = 0ull, f = 0ull;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (!alive[i]) continue;
        const uint32_t c = codes[i];
        const C k = count[c];
        if (singletons && k == (C)1) {
            alive[i] = 0;
            s++;
        } else if (pos && separated(family, (double)pos[c], (double)k)) {
            alive[i] = 0;
            f++;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, WAVE);
        f += __shfl_xor(f, o, WAVE);
    }
    if (lane_id() == 0) {
        if (s) atomicAdd(&drops[0], s);
        if (f) atomicAdd(&drops[1], f);
    }
}

__global__ void k_any_negative(const double* __restrict__ y, size_t n, uint32_t* flag) {
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bad |= !(y[i] >= 0.0);
    if (__ballot(bad) != 0 && lane_id() == 0) atomicOr(flag, 1u);
}

__global__ void k_check_codes(const uint32_t* __restrict__ codes, size_t n, uint64_t g, uint32_t* flag) {
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bad |= codes[i] >= g;
    if (__ballot(bad) != 0 && lane_id() == 0) atomicOr(flag, 1u);
}

// ------------------------------------------------------------------ the estimation sample

// Missing-value groups of a row (fit/input.rs:14-27): one bit per group of fh_sample_counts.
enum : uint32_t { M_Y = 1, M_X = 2, M_FE = 4, M_CL = 8, M_W = 16, M_OFF = 32, M_EXTRA = 64 };
constexpr int NGROUPS = 7;
constexpr int SCAN_MAXCOL = 32;
constexpr int SCAN_MAXFAC = 32;
struct ScanArgs {
    const double* x[SCAN_MAXCOL];  // y, offset and the regressors, each with its group bit
    uint32_t xbit[SCAN_MAXCOL];
    int nx;
    const uint32_t* fac[SCAN_MAXFAC];  // factor codes whose code 0 is the missing level
    uint32_t fbit[SCAN_MAXFAC];
    int nf;
    const double* w;
    const uint8_t* drop;
};

// counts: per group, then na, zero weight; flag: a negative weight (fit/input.rs:193-289).
__global__ void k_sample_scan(ScanArgs a, size_t n, uint8_t* __restrict__ alive, unsigned long long* counts,
                              uint32_t* neg) {
    unsigned long long c[NGROUPS + 2];
#pragma unroll
    for (int g = 0; g < NGROUPS + 2; g++) c[g] = 0ull;
    bool negative = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t m = 0;
        for (int j = 0; j < a.nx; j++)
            if (!isfinite(a.x[j][i])) m |= a.xbit[j];
        for (int j = 0; j < a.nf; j++)
            if (a.fac[j][i] == 0u) m |= a.fbit[j];
        double wi = 1.0;
        if (a.w) {
            wi = a.w[i];
            if (!isfinite(wi)) m |= M_W;
            negative |= wi < 0.0;
        }
        if (a.drop && a.drop[i]) m |= M_EXTRA;
#pragma unroll
        for (int g = 0; g < NGROUPS; g++) c[g] += (m >> g) & 1u;
        uint8_t keep = 0;
        if (m) c[NGROUPS]++;
        else if (wi == 0.0) c[NGROUPS + 1]++;
        else keep = 1;
        alive[i] = keep;
    }
#pragma unroll
    for (int g = 0; g < NGROUPS + 2; g++) {
        unsigned long long v = c[g];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
        if (lane_id() == 0 && v) atomicAdd(&counts[g], v);
    }
    if (__ballot(negative) != 0 && lane_id() == 0) atomicOr(neg, 1u);
}

__global__ void k_alive_flags(const uint8_t* __restrict__ alive, size_t n, uint32_t* __restrict__ f) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        f[i] = alive[i] ? 1u : 0u;
}

__global__ void k_compact_rows(const uint8_t* __restrict__ alive, const uint32_t* __restrict__ pos, size_t n,
                               uint32_t* __restrict__ kept) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (alive[i]) kept[pos[i]] = (uint32_t)i;
}

__global__ void k_gather_f64(const double* __restrict__ v, const uint32_t* __restrict__ rows, size_t nk,
                             double* __restrict__ out) {
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < nk; j += (size_t)gridDim.x * blockDim.x)
        out[j] = v[rows[j]];
}

__global__ void k_fill_u8(uint8_t* __restrict__ a, size_t n, uint8_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

}  // namespace

// ------------------------------------------------------------------ dense ranks

// codes[i] = rank of keys[i] among the distinct keys in [lo, hi]; returns their number. keys is
// scratch (the sort path shifts and reorders it).
uint64_t dense_rank(unsigned long long* keys, size_t n, unsigned long long lo, unsigned long long hi,
                    uint32_t* codes, hipStream_t st) {
    if (n == 0) return 0;
    StreamScope ss(st);
    const uint64_t range = hi - lo;
    DBuf<uint32_t> total(1);
    uint32_t nl = 0;
    if (range < table_limit(n)) {
        const size_t slots = (size_t)range + 1;
        DBuf<uint32_t> tab(slots);
        FH_HIP(hipMemsetAsync(tab.p, 0, slots * sizeof(uint32_t), st));
        hipLaunchKernelGGL(k_mark, dim3(fgrid(n)), dim3(FT), 0, st, keys, n, lo, tab.p);
        FH_CHECK_LAUNCH();
        exclusive_scan_u32(tab.p, tab.p, slots, total.p, st);
        hipLaunchKernelGGL(k_lookup, dim3(fgrid(n)), dim3(FT), 0, st, keys, n, lo, tab.p, codes);
        FH_CHECK_LAUNCH();
        download(&nl, total.p, sizeof(nl), st);
        return nl;
    }
    DBuf<uint32_t> vals(n), valt(n), flag(n);
    DBuf<uint64_t> kalt(n);
    hipLaunchKernelGGL(k_shift_iota, dim3(fgrid(n)), dim3(FT), 0, st, keys, n, lo, vals.p);
    FH_CHECK_LAUNCH();
    uint64_t* k64 = reinterpret_cast<uint64_t*>(keys);
    radix_sort_pairs(k64, vals.p, kalt.p, valt.p, n, bits_of(range), st);
    hipLaunchKernelGGL(k_run_start, dim3(fgrid(n)), dim3(FT), 0, st, keys, n, flag.p);
    FH_CHECK_LAUNCH();
    exclusive_scan_u32(flag.p, flag.p, n, total.p, st);
    hipLaunchKernelGGL(k_scatter_rank, dim3(fgrid(n)), dim3(FT), 0, st, keys, vals.p, flag.p, n, codes);
    FH_CHECK_LAUNCH();
    download(&nl, total.p, sizeof(nl), st);
    return nl;
}

// Dense codes of a factor column over its n rows (device data). Returns the levels; *missing =
// rows with a missing entry (code 0 when there are any).
uint64_t factorize_dev(const fh_factor_col& col, size_t n, bool strict, uint32_t* codes, uint64_t* missing,
                       hipStream_t st) {
    require(n < 0xFFFFFFFFull, "too many observations (limit 4294967294)");
    require(col.kind >= FH_COL_F64 && col.kind <= FH_COL_CODES, "unknown factor column kind");
    require(n == 0 || col.data != nullptr, "null factor data");
    if (missing) *missing = 0;
    if (n == 0) return 0;
    StreamScope ss(st);
    DBuf<unsigned long long> keys(n);
    DBuf<KeyStats> ks(1);
    hipLaunchKernelGGL(k_stats_init, dim3(1), dim3(1), 0, st, ks.p);
    const ColArg a{col.data, col.base, col.nlevels};
    const int s = strict ? 1 : 0;
    switch (col.kind) {
        case FH_COL_F64: hipLaunchKernelGGL(k_col_keys<FH_COL_F64>, dim3(fgrid(n)), dim3(FT), 0, st, a, n, s, keys.p, ks.p); break;
        case FH_COL_I32: hipLaunchKernelGGL(k_col_keys<FH_COL_I32>, dim3(fgrid(n)), dim3(FT), 0, st, a, n, s, keys.p, ks.p); break;
        case FH_COL_I64: hipLaunchKernelGGL(k_col_keys<FH_COL_I64>, dim3(fgrid(n)), dim3(FT), 0, st, a, n, s, keys.p, ks.p); break;
        default: hipLaunchKernelGGL(k_col_keys<FH_COL_CODES>, dim3(fgrid(n)), dim3(FT), 0, st, a, n, s, keys.p, ks.p); break;
    }
    FH_CHECK_LAUNCH();
    const KeyStats h = read_stats(ks.p, st);
    require(!(strict && h.missing), "missing values in factor");
    if (missing) *missing = h.missing;
    return dense_rank(keys.p, n, h.lo, h.hi, codes, st);
}

// Dense codes of the interaction of the device code arrays parts (factor.rs:259-281).
uint64_t combine_dev(const std::vector<const uint32_t*>& parts, const std::vector<uint64_t>& g, size_t n,
                     uint32_t* out, hipStream_t st) {
    require(!parts.empty() && parts.size() == g.size(), "combine: no factors", FH_INTERNAL);
    require(parts.size() <= (size_t)COMBINE_MAX, "combine: at most 64 factors");
    StreamScope ss(st);
    if (parts.size() == 1) {
        if (n) FH_HIP(hipMemcpyAsync(out, parts[0], n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        return g[0];
    }
    unsigned __int128 radix = 1;
    for (uint64_t gq : g) {
        radix *= (unsigned __int128)std::max<uint64_t>(gq, 1);
        require(radix < ((unsigned __int128)1 << 63), "interaction has too many levels");
    }
    if (n == 0) return 0;
    Parts p{};
    p.np = (int)parts.size();
    for (int q = 0; q < p.np; q++) {
        p.c[q] = parts[q];
        p.g[q] = g[q];
    }
    DBuf<unsigned long long> keys(n);
    DBuf<KeyStats> ks(1);
    hipLaunchKernelGGL(k_stats_init, dim3(1), dim3(1), 0, st, ks.p);
    hipLaunchKernelGGL(k_combine_keys, dim3(fgrid(n)), dim3(FT), 0, st, p, n, keys.p, ks.p);
    FH_CHECK_LAUNCH();
    const KeyStats h = read_stats(ks.p, st);
    require(h.bad == 0, "combine: a code is not below its level count");
    return dense_rank(keys.p, n, h.lo, h.hi, out, st);
}

// The codes of the rows kept (ascending input rows), renumbered densely (factor.rs:285-305).
uint64_t restrict_dev(const uint32_t* codes, uint64_t g, size_t n, const uint32_t* kept, size_t nk, uint32_t* out,
                      hipStream_t st) {
    if (nk == 0) return 0;
    StreamScope ss(st);
    DBuf<unsigned long long> keys(nk);
    DBuf<KeyStats> ks(1);
    hipLaunchKernelGGL(k_stats_init, dim3(1), dim3(1), 0, st, ks.p);
    hipLaunchKernelGGL(k_kept_keys, dim3(fgrid(nk)), dim3(FT), 0, st, codes, g, kept, nk, n, keys.p, ks.p);
    FH_CHECK_LAUNCH();
    const KeyStats h = read_stats(ks.p, st);
    require(h.bad == 0, "restrict: a kept row or a code is out of range");
    return dense_rank(keys.p, nk, h.lo, h.hi, out, st);
}

// Cyclic singleton (and fixest fe-separation) removal over the device term codes
// (sample.rs:44-74): alive (n bytes) is updated in place; returns {singletons, separated}.
std::pair<uint64_t, uint64_t> drop_singletons_dev(const std::vector<const uint32_t*>& codes,
                                                  const std::vector<uint64_t>& g, size_t n, uint8_t* alive,
                                                  const double* fw, const double* y_sep, int family, bool singletons,
                                                  const fh_ctx* ctx, hipStream_t st) {
    const size_t nt = codes.size();
    const bool sep = y_sep != nullptr;
    require(!sep || family == FH_FAMILY_POISSON || family == FH_FAMILY_LOGIT,
            "the separation rule needs a Poisson or logit family");
    std::pair<uint64_t, uint64_t> out{0, 0};
    if (nt == 0 || n == 0 || (!singletons && !sep)) return out;
    StreamScope ss(st);
    DBuf<uint32_t> flag(1);
    FH_HIP(hipMemsetAsync(flag.p, 0, sizeof(uint32_t), st));
    for (size_t q = 0; q < nt; q++) {
        require(codes[q] != nullptr, "null term codes");
        hipLaunchKernelGGL(k_check_codes, dim3(fgrid(n)), dim3(FT), 0, st, codes[q], n, g[q], flag.p);
    }
    FH_CHECK_LAUNCH();
    uint32_t hf = 0;
    download(&hf, flag.p, sizeof(hf), st);
    require(!hf, "a term code is not below its level count");
    // integer counts unless frequency weights or a Poisson outcome with negative values make the
    // sums real-valued (sample.rs:60-63)
    bool real = fw != nullptr;
    if (sep && family == FH_FAMILY_POISSON && !real) {
        FH_HIP(hipMemsetAsync(flag.p, 0, sizeof(uint32_t), st));
        hipLaunchKernelGGL(k_any_negative, dim3(fgrid(n)), dim3(FT), 0, st, y_sep, n, flag.p);
        FH_CHECK_LAUNCH();
        download(&hf, flag.p, sizeof(hf), st);
        real = hf != 0;
    }
    uint64_t gmax = 0;
    for (uint64_t gq : g) gmax = std::max(gmax, gq);
    DBuf<uint32_t> cnt, pos;
    DBuf<double> cntf, ysum;
    if (real) {
        cntf.alloc(gmax);
        if (sep) ysum.alloc(gmax);
    } else {
        cnt.alloc(gmax);
        if (sep) pos.alloc(gmax);
    }
    DBuf<unsigned long long> drops(2);
    unsigned long long hd[2];
    size_t remaining = nt, i = 0;
    while (remaining > 0) {
        check_interrupt(ctx);
        const size_t q = i % nt;
        i++;
        const uint64_t gq = g[q];
        FH_HIP(hipMemsetAsync(drops.p, 0, 2 * sizeof(unsigned long long), st));
        if (gq) {
            if (real) {
                FH_HIP(hipMemsetAsync(cntf.p, 0, gq * sizeof(double), st));
                if (sep) FH_HIP(hipMemsetAsync(ysum.p, 0, gq * sizeof(double), st));
                hipLaunchKernelGGL(k_level_sums, dim3(fgrid(n)), dim3(FT), 0, st, codes[q], alive, n, fw, y_sep,
                                   family == FH_FAMILY_LOGIT ? 1 : 0, cntf.p, sep ? ysum.p : nullptr);
                hipLaunchKernelGGL(k_drop_rows<double>, dim3(fgrid(n)), dim3(FT), 0, st, codes[q], alive, n, cntf.p,
                                   sep ? ysum.p : nullptr, singletons ? 1 : 0, family, drops.p);
            } else {
                FH_HIP(hipMemsetAsync(cnt.p, 0, gq * sizeof(uint32_t), st));
                if (sep) FH_HIP(hipMemsetAsync(pos.p, 0, gq * sizeof(uint32_t), st));
                hipLaunchKernelGGL(k_level_counts, dim3(fgrid(n)), dim3(FT), 0, st, codes[q], alive, n, y_sep, cnt.p,
                                   sep ? pos.p : nullptr);
                hipLaunchKernelGGL(k_drop_rows<uint32_t>, dim3(fgrid(n)), dim3(FT), 0, st, codes[q], alive, n, cnt.p,
                                   sep ? pos.p : nullptr, singletons ? 1 : 0, family, drops.p);
            }
            FH_CHECK_LAUNCH();
        }
        download(hd, drops.p, sizeof(hd), st);
        out.first += hd[0];
        out.second += hd[1];
        remaining = (hd[0] + hd[1] > 0) ? nt - 1 : remaining - 1;
    }
    return out;
}

}  // namespace fh

using namespace fh;

// The estimation sample: kept rows, dense term and cluster codes over them, and the counts.
struct fh_sample {
    int device = 0;
    hipStream_t st = nullptr;
    fh_sample_counts counts{};
    DBuf<uint32_t> kept;
    std::vector<DBuf<uint32_t>> terms, clusters;
    std::vector<uint64_t> term_g, cluster_g;
    ~fh_sample() {
        kept.release();
        terms.clear();
        clusters.clear();
    }
};

namespace {

// Device copy of a host array when the spec's arrays are on the host.
template <class T>
const T* on_dev(const T* p, size_t count, bool dev, std::vector<std::shared_ptr<DBuf<unsigned char>>>& hold,
                hipStream_t st) {
    if (!p || dev) return p;
    auto b = std::make_shared<DBuf<unsigned char>>(count * sizeof(T));
    if (count) upload(b->p, p, count * sizeof(T), st);
    hold.push_back(b);
    return reinterpret_cast<const T*>(b->p);
}

size_t col_bytes(int kind) { return kind == FH_COL_I32 || kind == FH_COL_CODES ? 4 : 8; }

void sample_build(const fh_sample_spec& s, const fh_ctx* ctx, fh_sample& out, bool dev) {
    const size_t n = s.n;
    hipStream_t st = out.st;
    require(n < 0xFFFFFFFFull, "too many observations (limit 4294967294)");
    require(s.nfactors == 0 || s.factors != nullptr, "null factor columns");
    require(s.nterms == 0 || (s.term_factors && s.term_off), "null term factor lists");
    require(s.nclusters == 0 || (s.cluster_factors && s.cluster_off), "null cluster factor lists");
    require(s.nclusters <= 4, "at most four cluster dimensions are supported");
    require(s.nx == 0 || s.x != nullptr, "null regressor columns");
    require(s.weight_kind >= FH_WEIGHT_NONE && s.weight_kind <= FH_WEIGHT_PWEIGHT, "unknown weight kind");
    require(s.weight_kind == FH_WEIGHT_NONE || s.weights != nullptr, "a weight kind without weights");
    require(s.family >= FH_FAMILY_GAUSSIAN && s.family <= FH_FAMILY_LOGIT, "unknown family");
    const bool sep = s.separation && s.family != FH_FAMILY_GAUSSIAN;
    require(!sep || s.y != nullptr, "the separation rule needs the outcome");
    auto check_list = [&](const uint32_t* f, const uint32_t* off, size_t cnt, const char* what) {
        for (size_t q = 0; q < cnt; q++) {
            require(off[q] < off[q + 1], std::string("an empty ") + what);
            for (uint32_t j = off[q]; j < off[q + 1]; j++)
                require(f[j] < s.nfactors, std::string("a ") + what + " references an invalid factor");
        }
    };
    check_list(s.term_factors, s.term_off, s.nterms, "fixed-effect term");
    check_list(s.cluster_factors, s.cluster_off, s.nclusters, "cluster dimension");
    std::vector<std::shared_ptr<DBuf<unsigned char>>> hold;
    fh_sample_counts& c = out.counts;
    c.n_input = n;

    // factor codes of the used factors
    std::vector<int> used_fe(s.nfactors, 0), used_cl(s.nfactors, 0);
    for (size_t q = 0; q < s.nterms; q++)
        for (uint32_t j = s.term_off[q]; j < s.term_off[q + 1]; j++) used_fe[s.term_factors[j]] = 1;
    for (size_t q = 0; q < s.nclusters; q++)
        for (uint32_t j = s.cluster_off[q]; j < s.cluster_off[q + 1]; j++) used_cl[s.cluster_factors[j]] = 1;
    std::vector<DBuf<uint32_t>> fcodes(s.nfactors);
    std::vector<uint64_t> fg(s.nfactors, 0), fmiss(s.nfactors, 0);
    for (size_t f = 0; f < s.nfactors; f++) {
        if (!used_fe[f] && !used_cl[f]) continue;
        fh_factor_col col = s.factors[f];
        require(col.kind >= FH_COL_F64 && col.kind <= FH_COL_CODES, "unknown factor column kind");
        col.data = on_dev(static_cast<const unsigned char*>(col.data), n * col_bytes(col.kind), dev, hold, st);
        fcodes[f].alloc(n);
        fg[f] = factorize_dev(col, n, !s.missing, fcodes[f].p, &fmiss[f], st);
        hold.clear();
    }
    log_line(ctx, 2, "fereg_hip: sample factorize");
    check_interrupt(ctx);

    // missing values and zero weights (fit/input.rs:193-289)
    const double* y = on_dev(s.y, n, dev, hold, st);
    const double* w = on_dev(s.weights, n, dev, hold, st);
    DBuf<uint8_t> alive(n);
    if (s.missing) {
        ScanArgs a{};
        auto add_col = [&](const double* p, uint32_t bit) {
            require(a.nx < SCAN_MAXCOL, "at most 30 regressors in the sample scan");
            a.x[a.nx] = p;
            a.xbit[a.nx++] = bit;
        };
        if (y) add_col(y, M_Y);
        if (s.offset) add_col(on_dev(s.offset, n, dev, hold, st), M_OFF);
        for (size_t j = 0; j < s.nx; j++) add_col(on_dev(s.x[j], n, dev, hold, st), M_X);
        for (size_t f = 0; f < s.nfactors; f++) {
            if (!fmiss[f]) continue;
            for (uint32_t bit : {M_FE, M_CL}) {
                if ((bit == M_FE && !used_fe[f]) || (bit == M_CL && !used_cl[f])) continue;
                require(a.nf < SCAN_MAXFAC, "too many factors with missing values");
                a.fac[a.nf] = fcodes[f].p;
                a.fbit[a.nf++] = bit;
            }
        }
        a.w = w;
        a.drop = on_dev(s.drop, n, dev, hold, st);
        DBuf<unsigned long long> cnt(NGROUPS + 2);
        DBuf<uint32_t> neg(1);
        FH_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), st));
        FH_HIP(hipMemsetAsync(neg.p, 0, sizeof(uint32_t), st));
        if (n) hipLaunchKernelGGL(k_sample_scan, dim3(fgrid(n)), dim3(FT), 0, st, a, n, alive.p, cnt.p, neg.p);
        FH_CHECK_LAUNCH();
        unsigned long long h[NGROUPS + 2];
        uint32_t hn = 0;
        download(h, cnt.p, sizeof(h), st);
        download(&hn, neg.p, sizeof(hn), st);
        require(!hn, "weights must not be negative");
        c.na_y = h[0];
        c.na_x = h[1];
        c.na_fe = h[2];
        c.na_cluster = h[3];
        c.na_weights = h[4];
        c.na_offset = h[5];
        c.na_extra = h[6];
        c.na = h[NGROUPS];
        c.zero_weight = h[NGROUPS + 1];
    } else if (n) {
        hipLaunchKernelGGL(k_fill_u8, dim3(fgrid(n)), dim3(FT), 0, st, alive.p, n, (uint8_t)1);
        FH_CHECK_LAUNCH();
    }
    log_line(ctx, 2, "fereg_hip: sample missing values");
    check_interrupt(ctx);

    // term and cluster codes over all rows: a factor used once alone keeps its codes
    std::vector<int> uses(s.nfactors, 0);
    for (size_t q = 0; q < s.nterms; q++)
        for (uint32_t j = s.term_off[q]; j < s.term_off[q + 1]; j++) uses[s.term_factors[j]]++;
    for (size_t q = 0; q < s.nclusters; q++)
        for (uint32_t j = s.cluster_off[q]; j < s.cluster_off[q + 1]; j++) uses[s.cluster_factors[j]]++;
    auto interact = [&](const uint32_t* f, uint32_t a0, uint32_t a1, uint64_t& g) {
        DBuf<uint32_t> r;
        if (a1 - a0 == 1 && uses[f[a0]] == 1) {
            r = std::move(fcodes[f[a0]]);
            g = fg[f[a0]];
        } else {
            std::vector<const uint32_t*> parts;
            std::vector<uint64_t> pg;
            for (uint32_t j = a0; j < a1; j++) {
                parts.push_back(fcodes[f[j]].p);
                pg.push_back(fg[f[j]]);
            }
            r.alloc(n);
            g = combine_dev(parts, pg, n, r.p, st);
        }
        for (uint32_t j = a0; j < a1; j++) uses[f[j]]--;
        return r;
    };
    std::vector<DBuf<uint32_t>> tcodes(s.nterms), ccodes(s.nclusters);
    std::vector<uint64_t> tg(s.nterms), cg(s.nclusters);
    for (size_t q = 0; q < s.nterms; q++) tcodes[q] = interact(s.term_factors, s.term_off[q], s.term_off[q + 1], tg[q]);
    for (size_t q = 0; q < s.nclusters; q++)
        ccodes[q] = interact(s.cluster_factors, s.cluster_off[q], s.cluster_off[q + 1], cg[q]);
    fcodes.clear();

    // singletons and separation (fit/input.rs:327-343; session.rs:469-473)
    std::vector<const uint32_t*> tp(s.nterms);
    for (size_t q = 0; q < s.nterms; q++) tp[q] = tcodes[q].p;
    const double* fw = s.weight_kind == FH_WEIGHT_FWEIGHT ? w : nullptr;
    const auto d = drop_singletons_dev(tp, tg, n, alive.p, fw, sep ? y : nullptr, s.family, s.drop_singletons != 0,
                                       ctx, st);
    c.singletons = d.first;
    c.perfect_fit = d.second;
    hold.clear();
    log_line(ctx, 2, "fereg_hip: sample singletons");

    // kept rows and the codes over them
    DBuf<uint32_t> flag(n), total(1);
    uint32_t nk = 0;
    if (n) {
        hipLaunchKernelGGL(k_alive_flags, dim3(fgrid(n)), dim3(FT), 0, st, alive.p, n, flag.p);
        FH_CHECK_LAUNCH();
        exclusive_scan_u32(flag.p, flag.p, n, total.p, st);
        download(&nk, total.p, sizeof(nk), st);
    }
    c.nobs = nk;
    require(nk >= 2, "fewer than 2 observations remain", FH_TOO_FEW);
    out.kept.alloc(nk);
    hipLaunchKernelGGL(k_compact_rows, dim3(fgrid(n)), dim3(FT), 0, st, alive.p, flag.p, n, out.kept.p);
    FH_CHECK_LAUNCH();
    for (size_t q = 0; q < s.nterms; q++) {
        DBuf<uint32_t> r(nk);
        out.term_g.push_back(restrict_dev(tcodes[q].p, tg[q], n, out.kept.p, nk, r.p, st));
        out.terms.push_back(std::move(r));
        tcodes[q].release();
    }
    for (size_t q = 0; q < s.nclusters; q++) {
        DBuf<uint32_t> r(nk);
        out.cluster_g.push_back(restrict_dev(ccodes[q].p, cg[q], n, out.kept.p, nk, r.p, st));
        out.clusters.push_back(std::move(r));
        ccodes[q].release();
    }
    stream_wait(st);
}

// A pool stream for the standalone entry points: their buffers are freed before it goes back.
struct JobStream {
    int device = 0;
    hipStream_t st = nullptr;
    JobStream() {
        FH_HIP(hipGetDevice(&device));
        st = stream_acquire();
    }
    ~JobStream() { stream_release(st, device); }
};

int factorize_entry(const fh_factor_col* col, size_t n, int32_t strict, uint32_t* codes, uint64_t* nlevels,
                    uint64_t* n_missing, bool dev) {
    return guard([&] {
        require(col && nlevels && (codes || n == 0), "null argument");
        JobStream js;
        {
            StreamScope ss(js.st);
            fh_factor_col c = *col;
            std::vector<std::shared_ptr<DBuf<unsigned char>>> hold;
            require(c.kind >= FH_COL_F64 && c.kind <= FH_COL_CODES, "unknown factor column kind");
            c.data = on_dev(static_cast<const unsigned char*>(c.data), n * col_bytes(c.kind), dev, hold, js.st);
            DBuf<uint32_t> dc;
            uint32_t* outc = codes;
            if (!dev) {
                dc.alloc(n);
                outc = dc.p;
            }
            uint64_t miss = 0;
            *nlevels = factorize_dev(c, n, strict != 0, outc, &miss, js.st);
            if (n_missing) *n_missing = miss;
            if (!dev) d2h(codes, dc.p, n, js.st);
            stream_wait(js.st);
        }
    });
}

int combine_entry(const uint32_t* const* parts, const uint64_t* nlevels, size_t nparts, size_t n, uint32_t* codes,
                  uint64_t* out_nlevels, bool dev) {
    return guard([&] {
        require(parts && nlevels && out_nlevels && (codes || n == 0), "null argument");
        require(nparts >= 1, "combine: no factors");
        JobStream js;
        {
            StreamScope ss(js.st);
            std::vector<std::shared_ptr<DBuf<unsigned char>>> hold;
            std::vector<const uint32_t*> p(nparts);
            std::vector<uint64_t> g(nlevels, nlevels + nparts);
            for (size_t q = 0; q < nparts; q++) {
                require(parts[q] != nullptr || n == 0, "null factor codes");
                p[q] = on_dev(parts[q], n, dev, hold, js.st);
            }
            DBuf<uint32_t> dc;
            uint32_t* outc = codes;
            if (!dev) {
                dc.alloc(n);
                outc = dc.p;
            }
            *out_nlevels = combine_dev(p, g, n, outc, js.st);
            if (!dev) d2h(codes, dc.p, n, js.st);
            stream_wait(js.st);
        }
    });
}

int restrict_entry(const uint32_t* codes, uint64_t nlevels, size_t n, const uint32_t* kept, size_t nkept,
                   uint32_t* out, uint64_t* out_nlevels, bool dev) {
    return guard([&] {
        require(out_nlevels && (nkept == 0 || (codes && kept && out)), "null argument");
        JobStream js;
        {
            StreamScope ss(js.st);
            std::vector<std::shared_ptr<DBuf<unsigned char>>> hold;
            const uint32_t* c = on_dev(codes, n, dev, hold, js.st);
            const uint32_t* k = on_dev(kept, nkept, dev, hold, js.st);
            DBuf<uint32_t> dc;
            uint32_t* o = out;
            if (!dev) {
                dc.alloc(nkept);
                o = dc.p;
            }
            *out_nlevels = restrict_dev(c, nlevels, n, k, nkept, o, js.st);
            if (!dev) d2h(out, dc.p, nkept, js.st);
            stream_wait(js.st);
        }
    });
}

int singletons_entry(const fh_term* terms, size_t nterms, size_t n, uint8_t* alive, const double* fw,
                     const double* y_sep, int32_t family, int32_t singletons, uint64_t* n_singletons,
                     uint64_t* n_separated, bool dev) {
    return guard([&] {
        require(alive || n == 0, "null alive flags");
        require(nterms == 0 || terms, "null terms");
        JobStream js;
        {
            StreamScope ss(js.st);
            std::vector<std::shared_ptr<DBuf<unsigned char>>> hold;
            std::vector<const uint32_t*> c(nterms);
            std::vector<uint64_t> g(nterms);
            for (size_t q = 0; q < nterms; q++) {
                c[q] = on_dev(terms[q].codes, n, dev, hold, js.st);
                g[q] = terms[q].g;
            }
            const double* f = on_dev(fw, n, dev, hold, js.st);
            const double* y = on_dev(y_sep, n, dev, hold, js.st);
            DBuf<uint8_t> da;
            uint8_t* a = alive;
            if (!dev && n) {
                da.alloc(n);
                h2d(da.p, alive, n, js.st);
                a = da.p;
            }
            const auto d = drop_singletons_dev(c, g, n, a, f, y, family, singletons != 0, nullptr, js.st);
            if (!dev && n) d2h(alive, da.p, n, js.st);
            stream_wait(js.st);
            if (n_singletons) *n_singletons = d.first;
            if (n_separated) *n_separated = d.second;
        }
    });
}

int sample_entry(const fh_sample_spec* s, const fh_ctx* ctx, fh_sample** out, bool dev) {
    return guard([&] {
        require(s && out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<fh_sample>();
        FH_HIP(hipGetDevice(&h->device));
        h->st = stream_acquire();
        const hipStream_t st = h->st;
        const int device = h->device;
        try {
            StreamScope ss(st);
            sample_build(*s, ctx, *h, dev);
        } catch (...) {
            // the buffers go back to the pool on their stream, then the stream
            (void)hipStreamSynchronize(st);
            {
                StreamScope ss(st);
                h.reset();
            }
            stream_release(st, device);
            throw;
        }
        *out = h.release();
    });
}

}  // namespace

extern "C" {

int fh_factorize(const fh_factor_col* col, size_t n, int32_t strict, uint32_t* codes, uint64_t* nlevels,
                 uint64_t* n_missing) {
    return factorize_entry(col, n, strict, codes, nlevels, n_missing, false);
}
int fh_factorize_dev(const fh_factor_col* col, size_t n, int32_t strict, uint32_t* codes, uint64_t* nlevels,
                     uint64_t* n_missing) {
    return factorize_entry(col, n, strict, codes, nlevels, n_missing, true);
}
int fh_combine(const uint32_t* const* parts, const uint64_t* nlevels, size_t nparts, size_t n, uint32_t* codes,
               uint64_t* out_nlevels) {
    return combine_entry(parts, nlevels, nparts, n, codes, out_nlevels, false);
}
int fh_combine_dev(const uint32_t* const* parts, const uint64_t* nlevels, size_t nparts, size_t n, uint32_t* codes,
                   uint64_t* out_nlevels) {
    return combine_entry(parts, nlevels, nparts, n, codes, out_nlevels, true);
}
int fh_restrict(const uint32_t* codes, uint64_t nlevels, size_t n, const uint32_t* kept, size_t nkept, uint32_t* out,
                uint64_t* out_nlevels) {
    return restrict_entry(codes, nlevels, n, kept, nkept, out, out_nlevels, false);
}
int fh_restrict_dev(const uint32_t* codes, uint64_t nlevels, size_t n, const uint32_t* kept, size_t nkept,
                    uint32_t* out, uint64_t* out_nlevels) {
    return restrict_entry(codes, nlevels, n, kept, nkept, out, out_nlevels, true);
}
int fh_drop_singletons(const fh_term* terms, size_t nterms, size_t n, uint8_t* alive, const double* fw,
                       const double* y_sep, int32_t family, int32_t singletons, uint64_t* n_singletons,
                       uint64_t* n_separated) {
    return singletons_entry(terms, nterms, n, alive, fw, y_sep, family, singletons, n_singletons, n_separated, false);
}
int fh_drop_singletons_dev(const fh_term* terms, size_t nterms, size_t n, uint8_t* alive, const double* fw,
                           const double* y_sep, int32_t family, int32_t singletons, uint64_t* n_singletons,
                           uint64_t* n_separated) {
    return singletons_entry(terms, nterms, n, alive, fw, y_sep, family, singletons, n_singletons, n_separated, true);
}

int fh_sample_build(const fh_sample_spec* s, const fh_ctx* ctx, fh_sample** out) {
    return sample_entry(s, ctx, out, false);
}
int fh_sample_build_dev(const fh_sample_spec* s, const fh_ctx* ctx, fh_sample** out) {
    return sample_entry(s, ctx, out, true);
}

void fh_sample_destroy(fh_sample* s) {
    if (!s) return;
    hipStream_t st = s->st;
    const int dev = s->device;
    DeviceGuard g(dev);
    (void)hipStreamSynchronize(st);
    {
        StreamScope ss(st);
        delete s;
    }
    stream_release(st, dev);
}

int fh_sample_get_counts(const fh_sample* s, fh_sample_counts* out) {
    return guard([&] {
        require(s && out, "null argument");
        *out = s->counts;
    });
}

const uint32_t* fh_sample_kept_dev(const fh_sample* s) { return s ? s->kept.p : nullptr; }

int fh_sample_kept(const fh_sample* s, uint32_t* rows) {
    return guard([&] {
        require(s && rows, "null argument");
        DeviceGuard g(s->device);
        d2h(rows, s->kept.p, s->counts.nobs, s->st);
        stream_wait(s->st);
    });
}

const uint32_t* fh_sample_term_dev(const fh_sample* s, size_t q, uint64_t* g) {
    if (!s || q >= s->terms.size()) return nullptr;
    if (g) *g = s->term_g[q];
    return s->terms[q].p;
}

const uint32_t* fh_sample_cluster_dev(const fh_sample* s, size_t q, uint64_t* g) {
    if (!s || q >= s->clusters.size()) return nullptr;
    if (g) *g = s->cluster_g[q];
    return s->clusters[q].p;
}

int fh_sample_term(const fh_sample* s, size_t q, uint32_t* codes, uint64_t* g) {
    return guard([&] {
        require(s && q < s->terms.size() && codes, "bad term index or null argument");
        DeviceGuard dg(s->device);
        d2h(codes, s->terms[q].p, s->counts.nobs, s->st);
        stream_wait(s->st);
        if (g) *g = s->term_g[q];
    });
}

int fh_sample_cluster(const fh_sample* s, size_t q, uint32_t* codes, uint64_t* g) {
    return guard([&] {
        require(s && q < s->clusters.size() && codes, "bad cluster index or null argument");
        DeviceGuard dg(s->device);
        d2h(codes, s->clusters[q].p, s->counts.nobs, s->st);
        stream_wait(s->st);
        if (g) *g = s->cluster_g[q];
    });
}

static int gather_impl(const fh_sample* s, const double* col, double* out, bool dev) {
    return guard([&] {
        require(s && col && out, "null argument");
        DeviceGuard dg(s->device);
        StreamScope ss(s->st);
        const size_t n = s->counts.n_input, nk = s->counts.nobs;
        std::vector<std::shared_ptr<DBuf<unsigned char>>> hold;
        const double* c = on_dev(col, n, dev, hold, s->st);
        DBuf<double> o;
        double* op = out;
        if (!dev) {
            o.alloc(nk);
            op = o.p;
        }
        hipLaunchKernelGGL(k_gather_f64, dim3(fgrid(nk)), dim3(FT), 0, s->st, c, s->kept.p, nk, op);
        FH_CHECK_LAUNCH();
        if (!dev) d2h(out, o.p, nk, s->st);
        stream_wait(s->st);
    });
}

int fh_sample_gather(const fh_sample* s, const double* col, double* out) { return gather_impl(s, col, out, false); }
int fh_sample_gather_dev(const fh_sample* s, const double* col, double* out) { return gather_impl(s, col, out, true); }

}  // extern "C"
