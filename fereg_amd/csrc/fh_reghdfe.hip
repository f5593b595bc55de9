// OLS with reghdfe 6.13.1's conventions on the GPU (ols.rs, dof.rs, stats.rs and vcov.rs of the
// reference). reghdfe standardizes every variable by its unweighted sd and takes every
// scale-sensitive decision (the collinearity sweep, the variance matrix, its PSD repair, the F
// test) on that scale. Residualization is linear, so the raw columns are residualized once and the
// Gram and meat matrices are rescaled on the host; the device passes are
//   column statistics (sums, then centered sums about the means: two passes per column block),
//   the residualization, the weighted and unweighted Grams, the residual pass, the design block
//   a_i = (x~_i + xbar, 1) and its robust or multiway-cluster meat,
// and the degrees of freedom come from device passes over the term codes: nesting in a cluster
// (first cluster value per level by compare-and-swap, then a check pass) and the mobility groups
// of each pair of terms (connected components by a lock-free union-find over the rows' edges).
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "fh_fit.hpp"

namespace fh {

namespace {

constexpr int RT = 256;
constexpr unsigned RGRID = 1024;
constexpr int NS = 5;  // sums of the first pass: x, x^2, w, w x, w x^2

__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < RT / WAVE; q++) t += red[q];
    __syncthreads();
    return t;  // valid in thread 0
}

// part[(block * R + j) * NS + s]: the first-pass sums of column j over the block's rows.
__global__ __launch_bounds__(RT) void k_col_sums(const double* __restrict__ z, size_t n, int ld, int R,
                                                 const double* __restrict__ w, double* __restrict__ part) {
    __shared__ double red[RT / WAVE];
    const int j = blockIdx.y;
    double s[NS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (size_t i = (size_t)blockIdx.x * RT + threadIdx.x; i < n; i += (size_t)gridDim.x * RT) {
        const double x = z[i * ld + j];
        const double wi = w ? w[i] : 1.0;
        s[0] += x;
        s[1] += x * x;
        s[2] += wi;
        s[3] += wi * x;
        s[4] += wi * x * x;
    }
    for (int q = 0; q < NS; q++) {
        const double t = block_sum(s[q], red);
        if (threadIdx.x == 0) part[((size_t)blockIdx.x * R + j) * NS + q] = t;
    }
}

// part[(block * R + j) * 2 + s]: sum (x - mean)^2 and sum w (x - wmean)^2 of column j.
__global__ __launch_bounds__(RT) void k_col_centered(const double* __restrict__ z, size_t n, int ld, int R,
                                                     const double* __restrict__ w, const double* __restrict__ mean,
                                                     const double* __restrict__ wmean, double* __restrict__ part) {
    __shared__ double red[RT / WAVE];
    const int j = blockIdx.y;
    const double m = mean[j], mw = wmean[j];
    double c = 0.0, cw = 0.0;
    for (size_t i = (size_t)blockIdx.x * RT + threadIdx.x; i < n; i += (size_t)gridDim.x * RT) {
        const double x = z[i * ld + j];
        const double wi = w ? w[i] : 1.0;
        c += (x - m) * (x - m);
        cw += wi * (x - mw) * (x - mw);
    }
    const double t0 = block_sum(c, red);
    const double t1 = block_sum(cw, red);
    if (threadIdx.x == 0) {
        part[((size_t)blockIdx.x * R + j) * 2] = t0;
        part[((size_t)blockIdx.x * R + j) * 2 + 1] = t1;
    }
}

// a_i = (z_i[kept] + xbar, 1): the design rows of the variance matrix (ols.rs:425-433).
__global__ void k_design(const double* __restrict__ z, size_t n, int ld, const int* __restrict__ kept, int k,
                         const double* __restrict__ xbar, int constant, double* __restrict__ A) {
    const int m = k + constant;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * m; t += (size_t)gridDim.x * blockDim.x) {
        const size_t i = t / m;
        const int p = (int)(t % m);
        A[t] = p < k ? z[i * ld + kept[p]] + (constant ? xbar[p] : 0.0) : 1.0;
    }
}

__global__ void k_sqrt(const double* __restrict__ w, size_t n, double* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = std::sqrt(w[i]);
}

// ------------------------------------------------------------------ degrees of freedom

__global__ void k_nest_first(const uint32_t* __restrict__ codes, const uint32_t* __restrict__ cl, size_t n,
                             uint32_t* map) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        atomicCAS(&map[codes[i]], 0xFFFFFFFFu, cl[i]);
}

__global__ void k_nest_check(const uint32_t* __restrict__ codes, const uint32_t* __restrict__ cl, size_t n,
                             const uint32_t* __restrict__ map, uint32_t* flag) {
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bad |= map[codes[i]] != cl[i];
    if (__ballot(bad) != 0 && lane_id() == 0) atomicOr(flag, 1u);
}

__device__ __forceinline__ uint32_t uf_load(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving. Parents only ever point to smaller vertices, so the walk ends.
__device__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    while (true) {
        const uint32_t p = uf_load(&parent[x]);
        if (p == x) return x;
        const uint32_t gp = uf_load(&parent[p]);
        if (gp == p) return p;
        atomicCAS(&parent[x], p, gp);
        x = gp;
    }
}

// Links the components of the two levels of every row (the larger root under the smaller); a
// failed link means another row changed that root first, and the roots are found again.
__global__ void k_uf_link(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint32_t ga, size_t n,
                          uint32_t* parent) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t u = a[i], v = ga + b[i];
        while (true) {
            u = uf_find(parent, u);
            v = uf_find(parent, v);
            if (u == v) break;
            if (u > v) {
                const uint32_t t = u;
                u = v;
                v = t;
            }
            if (atomicCAS(&parent[v], v, u) == v) break;
        }
    }
}

__global__ void k_iota_u32(uint32_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

__global__ void k_count_roots(uint32_t* parent, size_t nv, unsigned long long* cnt) {
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (size_t)gridDim.x * blockDim.x)
        c += uf_load(&parent[i]) == (uint32_t)i;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, WAVE);
    if (lane_id() == 0 && c) atomicAdd(cnt, c);
}

inline unsigned rgrid(size_t n) { return grid_for(n, RT * 4, 8192); }

bool nested_within(Absorber& a, const uint32_t* codes, uint64_t g, const uint32_t* cl) {
    hipStream_t st = a.st;
    DBuf<uint32_t> map(g), flag(1);
    FH_HIP(hipMemsetAsync(map.p, 0xFF, g * sizeof(uint32_t), st));
    FH_HIP(hipMemsetAsync(flag.p, 0, sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_nest_first, dim3(rgrid(a.n)), dim3(RT), 0, st, codes, cl, a.n, map.p);
    hipLaunchKernelGGL(k_nest_check, dim3(rgrid(a.n)), dim3(RT), 0, st, codes, cl, a.n, map.p, flag.p);
    FH_CHECK_LAUNCH();
    uint32_t h = 0;
    download(&h, flag.p, sizeof(h), st);
    return h == 0;
}

uint64_t mobility_groups(Absorber& a, const uint32_t* ca, uint64_t ga, const uint32_t* cb, uint64_t gb) {
    hipStream_t st = a.st;
    const size_t nv = ga + gb;
    require(nv < 0xFFFFFFFFull, "too many levels for the mobility groups");
    DBuf<uint32_t> parent(nv);
    DBuf<unsigned long long> cnt(1);
    hipLaunchKernelGGL(k_iota_u32, dim3(rgrid(nv)), dim3(RT), 0, st, parent.p, nv);
    hipLaunchKernelGGL(k_uf_link, dim3(rgrid(a.n)), dim3(RT), 0, st, ca, cb, (uint32_t)ga, a.n, parent.p);
    FH_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_count_roots, dim3(rgrid(nv)), dim3(RT), 0, st, parent.p, nv, cnt.p);
    FH_CHECK_LAUNCH();
    unsigned long long h = 0;
    download(&h, cnt.p, sizeof(h), st);
    return h;
}

struct Slot {
    uint64_t k = 0, m = 0;
    bool exact = false, nested = false;
};
struct Dof {
    std::vector<Slot> slots;
    uint64_t df_a = 0, initial = 0, redundant = 0, nested = 0;
};

// reghdfe estimate_dof (dof.rs:64-182) for intercept terms; no terms = one constant slot.
Dof estimate_dof(Absorber& a, const std::vector<const uint32_t*>& codes, const std::vector<uint64_t>& g,
                 const std::vector<const uint32_t*>& clusters, const fh_reghdfe_spec& sp) {
    Dof d;
    const size_t nt = codes.size();
    if (nt == 0) {
        d.slots.push_back(Slot{1, 0, true, false});
    } else {
        for (size_t q = 0; q < nt; q++) d.slots.push_back(Slot{g[q], 0, false, false});
    }
    std::vector<int> active(d.slots.size(), 1);
    if (!clusters.empty() && sp.dof_clusters && nt > 0) {
        for (const uint32_t* c : clusters)
            for (size_t q = 0; q < nt; q++) {
                if (!active[q] || d.slots[q].exact) continue;
                if (nested_within(a, codes[q], g[q], c)) {
                    d.slots[q].m = d.slots[q].k;
                    d.slots[q].exact = d.slots[q].nested = true;
                    d.nested += d.slots[q].k;
                    active[q] = 0;
                }
            }
    }
    size_t remaining = 0;
    for (int x : active) remaining += x;
    if (remaining > 1 || (remaining > 0 && d.nested > 0)) {
        bool skip = d.nested == 0;
        for (size_t q = 0; q < d.slots.size(); q++) {
            if (!active[q]) continue;
            if (skip) {
                skip = false;
                continue;
            }
            d.slots[q].m = 1;
        }
    }
    for (size_t q = 0; q < d.slots.size(); q++)
        if (active[q]) {
            d.slots[q].exact = true;
            break;
        }
    if ((sp.dof_pairwise || sp.dof_firstpair) && nt > 0) {
        std::vector<size_t> idx;
        for (size_t q = 0; q < nt; q++)
            if (active[q]) idx.push_back(q);
        int pairs = 0;
        bool done = false;
        for (size_t i = 0; i < idx.size() && !done; i++)
            for (size_t j = i + 1; j < idx.size(); j++) {
                const size_t qa = idx[i], qb = idx[j];
                pairs++;
                const uint64_t m = mobility_groups(a, codes[qa], g[qa], codes[qb], g[qb]);
                d.slots[qb].m = std::max(d.slots[qb].m, m);
                if (pairs == 1) d.slots[qb].exact = true;
                if (sp.dof_firstpair) {
                    done = true;
                    break;
                }
            }
    }
    for (const Slot& s : d.slots) {
        d.initial += s.k;
        d.redundant += s.m;
    }
    d.df_a = d.initial - d.redundant;
    return d;
}

// D M D scale, symmetrized before scaling (vcov.rs:254-266).
std::vector<double> sandwich(const std::vector<double>& D, const std::vector<double>& M, int m, double scale) {
    std::vector<double> t((size_t)m * m, 0.0), v((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++)
        for (int l = 0; l < m; l++) {
            const double x = D[(size_t)i * m + l];
            if (x == 0.0) continue;
            for (int j = 0; j < m; j++) t[(size_t)i * m + j] += x * M[(size_t)l * m + j];
        }
    for (int i = 0; i < m; i++)
        for (int l = 0; l < m; l++) {
            const double x = t[(size_t)i * m + l];
            if (x == 0.0) continue;
            for (int j = 0; j < m; j++) v[(size_t)i * m + j] += x * D[(size_t)l * m + j];
        }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < i; j++) {
            const double s = 0.5 * (v[(size_t)i * m + j] + v[(size_t)j * m + i]);
            v[(size_t)i * m + j] = v[(size_t)j * m + i] = s;
        }
    for (double& x : v) x *= scale;
    return v;
}

double missing_if_nonfinite(double x) { return std::isfinite(x) ? x : std::numeric_limits<double>::quiet_NaN(); }

}  // namespace

void reghdfe_fit(Absorber& a, const double* y, const double* X, size_t K, const fh_reghdfe_spec& sp,
                 const std::vector<const uint32_t*>& tcodes, const std::vector<uint64_t>& tg, const ClusterSet& cs,
                 double* b_out, double* V_out, int32_t* status_out, double* dof_out, fh_reghdfe_stats* st_out,
                 const fh_ctx* ctx) {
    hipStream_t st = a.st;
    const size_t n = a.n;
    const int R = (int)K + 1;
    require(K >= 1 && K <= 20, "1..20 regressors are supported");
    require(sp.vce >= FH_VCE_UNADJUSTED && sp.vce <= FH_VCE_CLUSTER, "unknown vce");
    require(sp.weight_kind >= FH_WEIGHT_NONE && sp.weight_kind <= FH_WEIGHT_PWEIGHT, "unknown weight kind");
    require((sp.weight_kind != FH_WEIGHT_NONE) == a.weighted,
            "the weight kind must be given exactly when the absorber holds weights");
    require(sp.vce != FH_VCE_CLUSTER || !cs.dims.empty(), "a cluster vce needs cluster codes");
    require(sp.tol > 0.0, "the tolerance must be positive");
    const int kind = sp.weight_kind;
    const bool awpw = kind == FH_WEIGHT_AWEIGHT || kind == FH_WEIGHT_PWEIGHT;
    const double* w = a.weighted ? a.w_rows.p : nullptr;  // raw weights, row order
    const bool report_constant = !sp.noconstant;         // every term has an intercept
    int vce = sp.vce;
    if (vce == FH_VCE_UNADJUSTED && kind == FH_WEIGHT_PWEIGHT) vce = FH_VCE_ROBUST;  // ols.rs:150-156

    // rows [y, X] in solver order
    const size_t ld = R <= 16 ? (size_t)padded_cols(R) : (size_t)R;
    a.rows_tmp.reserve(n * ld);
    double* rows = a.rows_tmp.p;
    gather_block(y, X, K, a.pi.p, n, ld, rows, st);

    // degrees of freedom (independent of the data columns)
    const Dof dof = estimate_dof(a, tcodes, tg, vce == FH_VCE_CLUSTER ? cs.dims : std::vector<const uint32_t*>{}, sp);
    check_interrupt(ctx);

    // pass A: raw column statistics (stats.rs:40-44, 154-184)
    std::vector<double> sumsq(R), sd(R), mean(R), wmean(R), wss(R), wss_unc(R);
    double sumw = 0.0;
    {
        const unsigned gx = std::max(1u, std::min<unsigned>(RGRID, (unsigned)((n + RT - 1) / RT)));
        DBuf<double> part((size_t)gx * R * NS), m2(2 * R);
        hipLaunchKernelGGL(k_col_sums, dim3(gx, R), dim3(RT), 0, st, rows, n, (int)ld, R, w, part.p);
        FH_CHECK_LAUNCH();
        const std::vector<double> s = sum_parts_dev(part.p, gx, (size_t)R * NS, st);
        const double dn = (double)n;
        for (int j = 0; j < R; j++) {
            mean[j] = s[(size_t)j * NS] / dn;
            sumsq[j] = s[(size_t)j * NS + 1];
            sumw = s[(size_t)j * NS + 2];
            wmean[j] = s[(size_t)j * NS + 3] / s[(size_t)j * NS + 2];
            wss_unc[j] = s[(size_t)j * NS + 4];
        }
        upload(m2.p, mean.data(), R * sizeof(double), st);
        upload(m2.p + R, wmean.data(), R * sizeof(double), st);
        DBuf<double> part2((size_t)gx * R * 2);
        hipLaunchKernelGGL(k_col_centered, dim3(gx, R), dim3(RT), 0, st, rows, n, (int)ld, R, w, m2.p, m2.p + R,
                           part2.p);
        FH_CHECK_LAUNCH();
        const std::vector<double> c = sum_parts_dev(part2.p, gx, (size_t)R * 2, st);
        for (int j = 0; j < R; j++) {
            sd[j] = std::max(std::sqrt(c[(size_t)j * 2] / (dn - 1.0)), 1e-3);
            wss[j] = c[(size_t)j * 2 + 1];
        }
    }
    const double n_obs = (double)n;
    if (!w) sumw = n_obs;
    const double big_n = kind == FH_WEIGHT_FWEIGHT ? sumw : n_obs;
    const double wscale = awpw ? n_obs / sumw : 1.0;  // wv = w * wscale
    const double tss_scale = awpw ? n_obs / sumw : 1.0;

    // residualize every column
    const int prec = sp.precision < 0 ? a.precision : sp.precision;
    const ResidOut ro = residualize_any(a, rows, ld, sp.tol, sp.maxiter ? sp.maxiter : 100000, nullptr, nullptr,
                                        prec, ctx);

    // pass B: Gram with the regression weights and the unweighted sums of squares
    std::vector<double> G((size_t)R * R), Gu;
    gram_cols(a, rows, n, (int)ld, 0, R, w, nullptr, G.data());
    if (w) {
        Gu.resize((size_t)R * R);
        gram_cols(a, rows, n, (int)ld, 0, R, nullptr, nullptr, Gu.data());
        for (double& x : G) x *= wscale;
    } else {
        Gu = G;
    }

    // collinearity, stage 1 (ols.rs:35-40) and stage 2 (ols.rs:46-56)
    const double ctol = std::min(1e-6, sp.tol / 10.0);
    std::vector<int32_t> status(K, 0);
    std::vector<bool> absorbed(R);
    for (int j = 0; j < R; j++) absorbed[j] = Gu[(size_t)j * R + j] / sumsq[j] <= ctol;
    const bool y_absorbed = absorbed[0];
    if (y_absorbed)
        for (int j = 0; j < R; j++) G[j] = G[(size_t)j * R] = 0.0;
    for (int j = 1; j < R; j++)
        if (absorbed[j]) status[j - 1] = 2;
    std::vector<int> cand;
    for (int j = 1; j < R; j++)
        if (status[j - 1] == 0) cand.push_back(j);
    const int nc = (int)cand.size();
    std::vector<double> gs((size_t)nc * nc);
    for (int p = 0; p < nc; p++)
        for (int q = 0; q < nc; q++)
            gs[(size_t)p * nc + q] = G[(size_t)cand[p] * R + cand[q]] / (sd[cand[p]] * sd[cand[q]]);
    std::vector<int> dropped;
    const std::vector<double> inv_c = nc ? invsym(gs, nc, INVSYM_TOL, &dropped) : std::vector<double>{};
    std::vector<int> keep_pos, kept;
    for (int c = 0; c < nc; c++) {
        if (dropped[c]) status[cand[c] - 1] = 3;
        else {
            keep_pos.push_back(c);
            kept.push_back(cand[c]);
        }
    }
    const int k = (int)kept.size();
    std::vector<double> inv_s((size_t)k * k);
    for (int p = 0; p < k; p++)
        for (int q = 0; q < k; q++) inv_s[(size_t)p * k + q] = inv_c[(size_t)keep_pos[p] * nc + keep_pos[q]];
    const double sd_y = sd[0];
    std::vector<double> b_s(k, 0.0), beta(k);
    for (int p = 0; p < k; p++) {
        double s = 0.0;
        for (int q = 0; q < k; q++) s += inv_s[(size_t)p * k + q] * (G[(size_t)kept[q] * R] / (sd[kept[q]] * sd_y));
        b_s[p] = s;
        beta[p] = s * sd_y / sd[kept[p]];
    }

    // pass C: residuals and rss = sum wv e^2
    std::vector<double> c(R, 0.0);
    c[0] = y_absorbed ? 0.0 : 1.0;
    for (int p = 0; p < k; p++) c[kept[p]] = -beta[p];
    DBuf<double> e(n);
    const double rss = residual_pass(a, rows, n, (int)ld, R, c.data(), w, e.p, -1) * wscale;
    const double rss_s = rss / (sd_y * sd_y);

    // constant and the extended bread (vcov.rs:215-236)
    std::vector<double> means_s(k);
    for (int p = 0; p < k; p++) means_s[p] = wmean[kept[p]] / sd[kept[p]];
    double b0_s = wmean[0] / sd_y;
    for (int p = 0; p < k; p++) b0_s -= means_s[p] * b_s[p];
    const int msize = k + (report_constant ? 1 : 0);
    const double n_c = awpw ? n_obs : sumw;
    std::vector<double> bread;
    if (report_constant) {
        const int k1 = k + 1;
        bread.assign((size_t)k1 * k1, 0.0);
        std::vector<double> side(k, 0.0);
        for (int j = 0; j < k; j++) {
            double s = 0.0;
            for (int i = 0; i < k; i++) s += means_s[i] * inv_s[(size_t)i * k + j];
            side[j] = -s;
        }
        double corner = 1.0 / n_c;
        for (int j = 0; j < k; j++) corner -= side[j] * means_s[j];
        for (int i = 0; i < k; i++) {
            for (int j = 0; j < k; j++) bread[(size_t)i * k1 + j] = inv_s[(size_t)i * k + j];
            bread[(size_t)i * k1 + k] = bread[(size_t)k * k1 + i] = side[i];
        }
        bread[(size_t)k * k1 + k] = corner;
    } else {
        bread = inv_s;
    }
    std::vector<double> b_full_s(b_s);
    if (report_constant) b_full_s.push_back(b0_s);

    // variance on the standardized scale (ols.rs:404-500)
    const double df_a = (double)dof.df_a, df_m = (double)k;
    double df_r = big_n - df_a - df_m;
    std::vector<double> sc(msize);
    for (int p = 0; p < k; p++) sc[p] = sd[kept[p]];
    if (report_constant) sc[k] = 1.0;
    std::vector<double> v_s((size_t)msize * msize, 0.0);
    double gmin = 0.0;
    std::vector<double> glist;
    if (vce == FH_VCE_CLUSTER) {
        gmin = std::numeric_limits<double>::infinity();
        for (size_t d = 0; d < cs.dims.size(); d++) {
            const double gd = (double)count_nonempty_clusters(a, cs.dims[d], cs.G[d]);
            glist.push_back(gd);
            gmin = std::min(gmin, gd);
        }
    }
    bool psd_fixed = false;
    if (msize > 0) {
        if (vce == FH_VCE_UNADJUSTED) {
            const double f = (rss_s / big_n) * (big_n / df_r);
            for (size_t i = 0; i < v_s.size(); i++) v_s[i] = bread[i] * f;
        } else {
            // the design rows a_i = (x~_i + xbar, 1)
            std::vector<int> kcols(kept.begin(), kept.end());
            std::vector<double> xbar(k);
            for (int p = 0; p < k; p++) xbar[p] = wmean[kept[p]];
            DBuf<int> kd(std::max(k, 1));
            DBuf<double> xd(std::max(k, 1)), A(n * msize);
            if (k) {
                upload(kd.p, kcols.data(), k * sizeof(int), st);
                upload(xd.p, xbar.data(), k * sizeof(double), st);
            }
            hipLaunchKernelGGL(k_design, dim3(grid_for(n * msize, RT, 1u << 16)), dim3(RT), 0, st, rows, n, (int)ld,
                               kd.p, k, xd.p, report_constant ? 1 : 0, A.p);
            FH_CHECK_LAUNCH();
            std::vector<double> meat((size_t)msize * msize);
            if (vce == FH_VCE_ROBUST) {
                // multiplier: e^2 (none), e^2 w (fw), (e wv)^2 (aw, pw): ols.rs:124-133
                DBuf<double> sw;
                const double* wp = nullptr;
                if (kind == FH_WEIGHT_FWEIGHT) {
                    sw.alloc(n);
                    hipLaunchKernelGGL(k_sqrt, dim3(grid_for(n, RT, 4096)), dim3(RT), 0, st, w, n, sw.p);
                    FH_CHECK_LAUNCH();
                    wp = sw.p;
                } else if (awpw) {
                    wp = w;
                }
                hetero_meat(a, A.p, n, msize, 0, msize, e.p, wp, meat.data());
                if (awpw)
                    for (double& x : meat) x *= wscale * wscale;
            } else {
                multiway_meat(a, A.p, msize, 0, msize, e.p, 1, w, cs, false, meat.data());
                if (awpw)
                    for (double& x : meat) x *= wscale * wscale;
            }
            for (int p = 0; p < msize; p++)
                for (int q = 0; q < msize; q++) meat[(size_t)p * msize + q] /= sc[p] * sc[q] * sd_y * sd_y;
            if (vce == FH_VCE_ROBUST) {
                v_s = sandwich(bread, meat, msize, big_n / (big_n - df_a - df_m));
            } else {
                const double nested_adj = dof.nested > 0 ? 1.0 : 0.0;
                const double q_cluster = (big_n - 1.0) / (big_n - nested_adj - df_m - df_a) * gmin / (gmin - 1.0);
                v_s = sandwich(bread, meat, msize, q_cluster);
                if (cs.dims.size() > 1) psd_fixed = fix_psd(v_s, msize, report_constant);
            }
        }
    }
    if (vce == FH_VCE_CLUSTER) df_r = std::min(df_r, gmin - 1.0);
    for (int i = 0; i < msize; i++)  // Mata _makesymmetric: the lower triangle wins
        for (int j = 0; j < i; j++) v_s[(size_t)j * msize + i] = v_s[(size_t)i * msize + j];

    // Wald F on the slopes (ols.rs:603-620)
    double f_stat = std::numeric_limits<double>::quiet_NaN();
    if (k > 0) {
        std::vector<double> vkk((size_t)k * k);
        for (int p = 0; p < k; p++)
            for (int q = 0; q < k; q++) vkk[(size_t)p * k + q] = v_s[(size_t)p * msize + q];
        std::vector<int> dv;
        const std::vector<double> iv = invsym_greedy(vkk, k, 1e-9, &dv);
        bool any = false;
        for (int x : dv) any |= x != 0;
        if (!any) {
            double qf = 0.0;
            for (int p = 0; p < k; p++)
                for (int q = 0; q < k; q++) qf += b_s[p] * iv[(size_t)p * k + q] * b_s[q];
            f_stat = missing_if_nonfinite(qf / df_m);
        }
    }
    bool any_nan = false;
    for (double x : v_s) any_nan |= std::isnan(x);
    if (any_nan) std::fill(v_s.begin(), v_s.end(), 0.0);

    // undo the standardization and expand to every regressor (+ _cons)
    const size_t k_all = K + (report_constant ? 1 : 0);
    std::vector<size_t> pos;
    for (int p = 0; p < k; p++) pos.push_back((size_t)kept[p] - 1);
    if (report_constant) pos.push_back(K);
    std::fill(b_out, b_out + k_all, 0.0);
    std::fill(V_out, V_out + k_all * k_all, 0.0);
    for (int p = 0; p < msize; p++) {
        const double sxp = sc[p] / sd_y;
        b_out[pos[p]] = b_full_s[p] / sxp;
        for (int q = 0; q < msize; q++)
            V_out[pos[p] * k_all + pos[q]] = v_s[(size_t)p * msize + q] / (sxp * (sc[q] / sd_y));
    }
    std::copy(status.begin(), status.end(), status_out);
    for (size_t s = 0; s < dof.slots.size(); s++) {
        const Slot& sl = dof.slots[s];
        dof_out[s * 5 + 0] = (double)sl.k;
        dof_out[s * 5 + 1] = (double)sl.m;
        dof_out[s * 5 + 2] = (double)(sl.k - sl.m);
        dof_out[s * 5 + 3] = sl.exact ? 0.0 : 1.0;
        dof_out[s * 5 + 4] = sl.nested ? 1.0 : 0.0;
    }

    // fit statistics (stats.rs:267-280)
    if (st_out) {
        fh_reghdfe_stats& o = *st_out;
        o = fh_reghdfe_stats{};
        const double tss = wss[0] * tss_scale;  // every model here has an intercept
        const double tss_within = G[0];
        const double used_df_r = big_n - df_a - df_m - (double)dof.nested;
        const double ln2pi = std::log(2.0 * M_PI);
        o.N = big_n;
        o.rank = df_m;
        o.df_r = missing_if_nonfinite(df_r);
        o.df_m = df_m;
        o.df_a = df_a;
        o.df_a_initial = (double)dof.initial;
        o.df_a_redundant = (double)dof.redundant;
        o.df_a_nested = (double)dof.nested;
        o.tss = missing_if_nonfinite(tss);
        o.tss_within = missing_if_nonfinite(tss_within);
        o.rss = missing_if_nonfinite(rss);
        o.mss = missing_if_nonfinite(tss - rss);
        o.rmse = missing_if_nonfinite(used_df_r != 0.0 ? std::sqrt(rss / used_df_r) : std::sqrt(rss));
        o.F = f_stat;
        o.ll = missing_if_nonfinite(-0.5 * big_n * (1.0 + ln2pi + std::log(rss / big_n)));
        o.ll_0 = missing_if_nonfinite(-0.5 * big_n * (1.0 + ln2pi + std::log(tss_within / big_n)));
        o.r2 = missing_if_nonfinite(1.0 - rss / tss);
        o.r2_a = missing_if_nonfinite(1.0 - (rss / used_df_r) / (tss / (big_n - 1.0)));
        o.r2_within = missing_if_nonfinite(1.0 - rss / tss_within);
        o.r2_a_within = missing_if_nonfinite(1.0 - (rss / used_df_r) / (tss_within / (used_df_r + df_m)));
        o.sumweights = sumw;
        o.num_singletons = (double)sp.num_singletons;
        o.N_full = big_n + (double)sp.num_singletons;
        o.report_constant = report_constant ? 1.0 : 0.0;
        o.converged = 1.0;
        o.y_absorbed = y_absorbed ? 1.0 : 0.0;
        o.psd_fixed = psd_fixed ? 1 : 0;
        o.iterations = ro.iterations;
        o.nslots = dof.slots.size();
        if (vce == FH_VCE_CLUSTER) {
            o.N_clust = gmin;
            o.N_clustervars = (double)glist.size();
            for (size_t d = 0; d < glist.size(); d++) o.N_clust_dim[d] = glist[d];
        } else {
            o.N_clust = std::numeric_limits<double>::quiet_NaN();
        }
    }
}

}  // namespace fh

using namespace fh;

extern "C" {

static int reghdfe_impl(fh_absorber* ab, const double* y, const double* X, size_t K, const fh_reghdfe_spec* sp,
                        double* b, double* V, int32_t* status, double* dof_table, fh_reghdfe_stats* stats,
                        const fh_ctx* ctx, bool dev) {
    return guard([&] {
        require(ab && y && X && sp && b && V && status && dof_table, "null argument");
        Absorber& a = ab->a;
        AbsScope g(a);
        require(!a.slope, "the reghdfe convention with varying slopes is not supported");
        restore_user_weights(a);
        const size_t n = a.n;
        require(sp->nterms == 0 || sp->terms, "null terms");
        require(sp->nclusters <= 4, "at most four cluster dimensions are supported");
        require(sp->nclusters == 0 || sp->clusters, "null cluster dimensions");
        std::vector<DBuf<uint32_t>> up;
        up.reserve(sp->nterms + sp->nclusters);
        auto dev_codes = [&](const uint32_t* p) -> const uint32_t* {
            require(p != nullptr, "null codes");
            if (dev) return p;
            up.emplace_back(n);
            h2d(up.back().p, p, n, a.st);
            return up.back().p;
        };
        std::vector<const uint32_t*> tc;
        std::vector<uint64_t> tgs;
        for (size_t q = 0; q < sp->nterms; q++) {
            tc.push_back(dev_codes(sp->terms[q].codes));
            tgs.push_back(sp->terms[q].g);
        }
        std::vector<const uint32_t*> cd;
        for (size_t q = 0; q < sp->nclusters; q++) cd.push_back(dev_codes(sp->clusters[q]));
        const ClusterSet cs = cluster_set(a, cd.data(), cd.size(), FH_GDF_MIN);
        if (dev) {
            reghdfe_fit(a, y, X, K, *sp, tc, tgs, cs, b, V, status, dof_table, stats, ctx);
        } else {
            DBuf<double> dy(n), dx(n * K);
            h2d(dy.p, y, n, a.st);
            h2d(dx.p, X, n * K, a.st);
            reghdfe_fit(a, dy.p, dx.p, K, *sp, tc, tgs, cs, b, V, status, dof_table, stats, ctx);
        }
    });
}

void fh_reghdfe_spec_default(fh_reghdfe_spec* sp) {
    *sp = fh_reghdfe_spec{};
    sp->vce = FH_VCE_UNADJUSTED;
    sp->weight_kind = FH_WEIGHT_NONE;
    sp->dof_pairwise = 1;
    sp->dof_clusters = 1;
    sp->dof_continuous = 1;
    sp->tol = 1e-8;
    sp->maxiter = 16000;
    sp->precision = -1;
}

int fh_reghdfe_fit(fh_absorber* ab, const double* y, const double* X, size_t K, const fh_reghdfe_spec* spec,
                   double* b, double* V, int32_t* status, double* dof_table, fh_reghdfe_stats* stats,
                   const fh_ctx* ctx) {
    return reghdfe_impl(ab, y, X, K, spec, b, V, status, dof_table, stats, ctx, false);
}
int fh_reghdfe_fit_dev(fh_absorber* ab, const double* y, const double* X, size_t K, const fh_reghdfe_spec* spec,
                       double* b, double* V, int32_t* status, double* dof_table, fh_reghdfe_stats* stats,
                       const fh_ctx* ctx) {
    return reghdfe_impl(ab, y, X, K, spec, b, V, status, dof_table, stats, ctx, true);
}

}  // extern "C"
