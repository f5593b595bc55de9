// CRV3: the leave-one-cluster-out jackknife (crv3.rs:1-30, fit/vcov.rs:369-397). For every cluster g
// the model is re-estimated without g's rows and V = factor sum_S sign_S scale(G_S)
// sum_g (beta_(-g) - beta)(beta_(-g) - beta)' over the non-empty subsets S of the cluster dimensions
// (inclusion-exclusion as for CRV1, each subset clustered by the intersection of its dimensions).
//
// Leaving a cluster out is a weight of zero on its rows: the absorber keeps its structure, the
// emptied fixed-effect levels drop out of the reduced system as redundant coordinates (n_r = 0)
// and empty FE1 levels contribute nothing, so every refit is one weight refresh and one solve on
// the device. A dropped row's level that the reference's fresh fit would remove as a singleton or
// a perfect fit does not move beta, in either form.
#include <functional>

#include "fh_fit.hpp"
#include "fh_internal.hpp"

namespace fh {

namespace {

constexpr int TPB = 256;

// wm[i] = (w ? w[i] : 1) [cl(pi[i]) != g] for the rows in solver order; cl in kept order.
__global__ void k_leave_out(const double* __restrict__ w, const uint32_t* __restrict__ cl_kept,
                            const uint32_t* __restrict__ pi, size_t n, uint32_t g, double* __restrict__ wm) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        wm[i] = cl_kept[pi[i]] == g ? 0.0 : (w ? w[i] : 1.0);
}

__global__ void k_cluster_count(const uint32_t* __restrict__ cl, size_t n, uint32_t* __restrict__ cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[cl[i]], 1u);
}

// sum over the clusters with rows of (b_g - beta)(b_g - beta)' (crv3.rs:89-104), in cluster order.
void jackknife_meat(Absorber& a, const uint32_t* cl_kept, uint64_t G, const double* w_row,
                    const std::vector<double>& beta, const CrvRefit& refit, std::vector<double>& meat) {
    hipStream_t st = a.st;
    const size_t k = beta.size();
    meat.assign(k * k, 0.0);
    DBuf<uint32_t> cnt(std::max<uint64_t>(G, 1));
    FH_HIP(hipMemsetAsync(cnt.p, 0, G * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_cluster_count, dim3(grid_for(a.n, TPB, 4096)), dim3(TPB), 0, st, cl_kept, a.n, cnt.p);
    FH_CHECK_LAUNCH();
    std::vector<uint32_t> hc(G);
    d2h(hc.data(), cnt.p, G, st);
    DBuf<double> wm(a.n);
    for (uint64_t g = 0; g < G; g++) {
        if (hc[g] == 0) continue;
        hipLaunchKernelGGL(k_leave_out, dim3(grid_for(a.n, TPB, 4096)), dim3(TPB), 0, st, w_row, cl_kept, a.pi.p, a.n,
                           (uint32_t)g, wm.p);
        FH_CHECK_LAUNCH();
        std::vector<double> b = refit(wm.p, g);
        for (size_t i = 0; i < k; i++)
            for (size_t j = 0; j < k; j++) meat[i * k + j] += (b[i] - beta[i]) * (b[j] - beta[j]);
    }
}

}  // namespace

void crv3_vcov(Absorber& a, const ClusterSet& cs, const fh_ssc& ssc, double dfk, const double* w_row,
               const std::vector<double>& beta, const CrvRefit& refit, std::vector<double>& V, uint64_t* ncl) {
    const int q = (int)cs.dims.size();
    require(q >= 1 && q <= 4, "CRV3: 1..4 cluster dimensions are supported");
    const size_t k = beta.size(), kk = k * k;
    const double n = (double)a.n;
    const bool conventional = ssc.g_adj && cs.g_df == FH_GDF_CONVENTIONAL;
    std::vector<double> meat(kk, 0.0), part;
    DBuf<uint32_t> inter(q > 1 ? a.n : 0);
    uint64_t gmin = UINT64_MAX;
    for (int size = 1; size <= q; size++) {
        const double sign = (size % 2) ? 1.0 : -1.0;
        std::vector<std::vector<int>> subs;
        std::vector<int> cur;
        std::function<void(int)> rec = [&](int start) {
            if ((int)cur.size() == size) {
                subs.push_back(cur);
                return;
            }
            for (int d = start; d < q; d++) {
                cur.push_back(d);
                rec(d + 1);
                cur.pop_back();
            }
        };
        rec(0);
        for (const auto& sub : subs) {
            const uint32_t* codes;
            uint64_t G;
            if (size == 1) {
                codes = cs.dims[sub[0]];
                G = cs.G[sub[0]];
            } else {
                std::vector<const uint32_t*> parts;
                std::vector<uint64_t> pg;
                for (int d : sub) {
                    parts.push_back(cs.dims[d]);
                    pg.push_back(cs.G[d]);
                }
                G = combine_dev(parts, pg, a.n, inter.p, a.st);
                codes = inter.p;
            }
            jackknife_meat(a, codes, G, w_row, beta, refit, part);
            const uint64_t g = count_nonempty_clusters(a, codes, G);
            require(g >= 2, "CRV3: at least two clusters are required");
            if (size == 1) gmin = std::min(gmin, g);
            const double f = sign * (conventional ? (double)g / ((double)g - 1.0) : 1.0);
            for (size_t i = 0; i < kk; i++) meat[i] += f * part[i];
        }
    }
    double factor = 1.0;
    if (ssc.g_adj && cs.g_df == FH_GDF_MIN) factor *= (double)gmin / ((double)gmin - 1.0);
    if (ssc.k_adj) factor *= (n - 1.0) / (n - dfk);
    V.resize(kk);
    for (size_t i = 0; i < kk; i++) V[i] = meat[i] * factor;
    if (ncl) *ncl = gmin;
    bool finite = true;
    for (double v : V) finite &= std::isfinite(v);
    if (finite) posdef_fix(V, (int)k);
}

}  // namespace fh
