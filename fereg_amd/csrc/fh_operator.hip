// Reduced operator, FP64 instantiation, rhs pass and back-substitution (see fh_operator_impl.hpp).
#include <map>
#include <mutex>

#include "fh_operator_impl.hpp"

namespace fh {

template <class T>
OpArgs<T> op_args(const Absorber& a) {
    return op_args_impl<T>(a);
}
template OpArgs<double> op_args<double>(const Absorber&);
template OpArgs<float> op_args<float>(const Absorber&);

template <class T>
void spmv(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, const uint32_t* done, hipStream_t st,
          double* pq) {
    if (a.slope) {
        slope_spmv(a, p, y, r, done, st, pq);
        return;
    }
    spmv_planned<T>(a, p, y, r, s, done, st, pq);
}
template void spmv<double>(const Absorber&, const double*, double*, int, Scratch<double>&, const uint32_t*, hipStream_t,
                           double*);
template SpmvNeed spmv_scratch_need<double>(const Absorber&, const SpmvPlan&, int);
template void spmv_fixed<double>(const Absorber&, const double*, double*, int, Scratch<double>&, hipStream_t);

namespace {
std::mutex tune_mu;
std::map<std::string, SpmvPlan>& tune_cache() {
    static auto* c = new std::map<std::string, SpmvPlan>();
    return *c;
}
}  // namespace

std::string plan_key(const Absorber& a, int tbytes, int rclass) {
    std::string k = std::to_string(tbytes) + "/" + std::to_string(rclass) + "/" + std::to_string(a.M) + "/" +
                    std::to_string(a.g1) + "/" + std::to_string(a.m) + "/" + std::to_string(a.ms) + "/" +
                    std::to_string(a.unit() ? 0 : 1) + "/" + std::to_string(a.small_dense ? 1 : 0);
    for (size_t b = 0; b < a.nb; b++)
        k += "/" + std::to_string(a.blocks[b].g) + ":" + std::to_string(a.nbk[b]) + ":" +
             std::to_string(a.blocks[b].small ? 1 : 0);
    return k;
}

bool tune_cache_get(const std::string& key, SpmvPlan& plan) {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tune_cache().find(key);
    if (it == tune_cache().end()) return false;
    plan = it->second;
    return true;
}

void tune_cache_put(const std::string& key, const SpmvPlan& plan) {
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_cache()[key] = plan;
}

void tune_cache_clear() {
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_cache().clear();
}

namespace {

// rhs of the dense small block: part[wg][k][j] = sum of z_c[j] over the cells c of this
// workgroup's range whose small code is k (chunk of KC levels in blockIdx.y). RMP >= r threads
// share a cell (one column each), so every z_c element is read once; lanes with the same column
// are reduced by shuffles, then the waves in order.
template <int RMP, int KC>
__global__ __launch_bounds__(WG) void k_small_rhs(const uint32_t* __restrict__ cell_codes, int nb, int bs,
                                                  const double* __restrict__ zc, size_t M, int r, int ms,
                                                  double* __restrict__ part) {
    constexpr int CPB = WG / RMP;
    __shared__ double red[WAVES][KC][RMP];
    const int j = threadIdx.x % RMP, ci = threadIdx.x / RMP, k0 = blockIdx.y * KC;
    double acc[KC];
#pragma unroll
    for (int kk = 0; kk < KC; kk++) acc[kk] = 0.0;
    if (j < r)
        for (size_t c = (size_t)blockIdx.x * CPB + ci; c < M; c += (size_t)gridDim.x * CPB) {
            const int k = (int)cell_codes[c * nb + bs] - k0;
            const double z = zc[c * r + j];
#pragma unroll
            for (int kk = 0; kk < KC; kk++) acc[kk] += k == kk ? z : 0.0;
        }
    const int w = threadIdx.x / WAVE, l = lane_id();
#pragma unroll
    for (int kk = 0; kk < KC; kk++) {
        double v = acc[kk];
#pragma unroll
        for (int o = RMP; o < WAVE; o <<= 1) v += __shfl_xor(v, o, WAVE);
        if (l < RMP) red[w][kk][l] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < KC * r; e += WG) {
        const int kk = e / r, jj = e % r;
        if (k0 + kk >= ms) continue;
        double sm = 0.0;
#pragma unroll
        for (int ww = 0; ww < WAVES; ww++) sm += red[ww][kk][jj];
        part[(size_t)blockIdx.x * ms * r + (size_t)(k0 + kk) * r + jj] = sm;
    }
}

void small_rhs(const Absorber& a, int b, const double* zc, int r, double* rhs_b, hipStream_t st) {
    const int ms = (int)a.blocks[b].g, nb = (int)a.nb;
    auto go = [&](auto rmp, auto kc) {
        constexpr int RMP = decltype(rmp)::value, KC = decltype(kc)::value;
        const unsigned nwg = grid_for(a.M, WG / RMP, 1024), nch = (unsigned)((ms + KC - 1) / KC);
        DBuf<double> part((size_t)nwg * ms * r);
        hipLaunchKernelGGL((k_small_rhs<RMP, KC>), dim3(nwg, nch), dim3(WG), 0, st, a.cell_codes.p, nb, b, zc,
                           (size_t)a.M, r, ms, part.p);
        FH_CHECK_LAUNCH();
        sum_parts_to(part.p, nwg, (size_t)ms * r, rhs_b, st);
    };
    auto by_r = [&](auto kc) {
        if (r <= 1) go(std::integral_constant<int, 1>{}, kc);
        else if (r <= 2) go(std::integral_constant<int, 2>{}, kc);
        else if (r <= 4) go(std::integral_constant<int, 4>{}, kc);
        else if (r <= 8) go(std::integral_constant<int, 8>{}, kc);
        else go(std::integral_constant<int, 16>{}, kc);
    };
    if (ms <= 8) by_r(std::integral_constant<int, 8>{});
    else if (ms <= 16) by_r(std::integral_constant<int, 16>{});
    else by_r(std::integral_constant<int, 32>{});
}

}  // namespace

unsigned rhs_pass(Absorber& a, const double* rows, int r, int rp, double* rhs, double* b1, size_t extra) {
    require(r >= 1 && r <= rp && rp <= 16, "rhs: 1..16 columns");
    hipStream_t st = a.st;
    OpArgs<double> A = op_args_impl<double>(a);
    const unsigned grid = seg_grid(a);
    Scratch<double>& s = a.s64;
    s.dpart.reserve((size_t)grid * 2 * rp + extra);
    if (a.nb > 0) s.zc.reserve(a.M * rp);
    const int L = (int)a.seg_lanes;
    const double* wr = a.weighted ? a.w_rows.p : nullptr;
    FH_DISPATCH_R(rp, hipLaunchKernelGGL((k_rhs_seg_reg<RM>), dim3(grid), dim3(WG), 0, st, A, rows, wr, r, rp, L, b1,
                                         a.nb > 0 ? s.zc.p : nullptr, s.dpart.p));
    FH_CHECK_LAUNCH();
    for (size_t b = 0; b < a.nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (a.boff[b].n == 0) {  // the dense small block: binned by code, no level pass
            small_rhs(a, (int)b, s.zc.p, rp, rhs + bi.off * rp, st);
            continue;
        }
        const int Ll = (int)a.blanes[b];
        const uint32_t nbk = a.nbk[b];
        const size_t items = (size_t)nbk * bi.g;
        double* ypart = nullptr;
        if (nbk > 1) {
            s.ypart.reserve(items * rp);
            ypart = s.ypart.p;
        }
        FH_DISPATCH_R(rp, hipLaunchKernelGGL((k_lvl_pull<RM>), dim3(lvl_grid(Ll, items)), dim3(WG), 0, st,
                                             a.boff[b].p, a.bperm[b].p, items, (uint32_t)bi.g, s.zc.p, rp, Ll,
                                             rhs + bi.off * rp, ypart));
        FH_CHECK_LAUNCH();
        if (ypart) bucket_reduce<double>(ypart, nbk, (size_t)bi.g, rp, rhs + bi.off * rp, nullptr, st);
    }
    return grid;
}

void residualize_rows(Absorber& a, double* rows, int r, int rp, const double* x, const double* b1, double* alpha) {
    require(r >= 1 && r <= rp && rp <= 16, "residualize: 1..16 columns");
    OpArgs<double> A = op_args_impl<double>(a);
    const unsigned grid = seg_grid(a);
    // rows of one or two columns: half the segment pass's lanes, so a lane's eight trips cover
    // more of its segment (C2, 10 cells per segment: 2 lanes 142 us, 4 lanes 157 us; wider rows
    // keep the segment pass's count, 12 columns: 1298 against 1178 us)
    const int L = rp <= 2 && a.seg_lanes > 1 ? (int)a.seg_lanes / 2 : (int)a.seg_lanes;
    if (a.nb <= 4) {
        FH_DISPATCH_R(rp, with_nb0((int)a.nb, [&](auto nbc) {
            constexpr int NB = decltype(nbc)::value;
            hipLaunchKernelGGL((k_resid_seg<RM, NB>), dim3(grid), dim3(WG), 0, a.st, A, rows, r, rp, L, x, b1, alpha);
        }));
    } else {
        FH_DISPATCH_R(rp, hipLaunchKernelGGL((k_resid_seg_any<RM>), dim3(grid), dim3(WG), 0, a.st, A, rows, r, rp, L, x,
                                             b1, alpha));
    }
    FH_CHECK_LAUNCH();
}

}  // namespace fh
