// Exact leverage of the fixed effects and the HC2 / HC3 scores (leverage.rs:1-130,
// fit/vcov.rs:230-254 and 333-364).
//
// With FE1 eliminated (its levels are the segments s), the FE indicators d of a row in cell c
// give d'(D'WD)^+ d = 1/n1_s + u_c' S^+ u_c, where S is the reduced matrix and u_c = e_c - v_s the
// cell's reduced indicators minus the segment's weighted mean v_s = (1/n1_s) sum_{c' in s} W_c' e_c'.
// Every pair of coordinates that meets in a segment is a structural entry of S, hence inside the
// pattern of its factor, so S^+ is only needed there: the joint factor (assembled on the GPU,
// factored on the host) is inverted on its own pattern by selected inversion, and one wave per
// segment forms u_c' S^+ u_c for its cells from the coordinates of its items:
//   u_c' S^+ u_c = sum_{t,u in c} P_tu - (2/n1) sum_{t in c} A_t + (1/n1^2) sum_t W_t A_t,
//   P_tu = S^+[k_t, k_u], A_t = sum_{u in s} W_u P_tu.
// The leverage of a row is then c (hcell + x~' B x~) with c = w_i, or 1 for fixest's OLS without
// absorbed terms (fit/vcov.rs:235).
#include <cmath>

#include "fh_fit.hpp"
#include "fh_internal.hpp"
#include "fh_ldl.hpp"

namespace fh {

namespace {

constexpr int TPB = 256;

struct ZDev {
    const uint32_t* pos;  // m: factor position of each reduced coordinate (UINT32_MAX: not kept)
    const uint32_t* zp;   // n + 1: columns of the selected inverse (the factor's pattern)
    const uint32_t* zi;   // rows, ascending within a column
    const double* zx;
    const double* zd;     // diagonal
    const double* sc;     // scaling of each position
};

// S^+[kp, kq] for two reduced coordinates.
__device__ double zlook(const ZDev& Z, uint64_t kp, uint64_t kq) {
    uint32_t p = Z.pos[kp], q = Z.pos[kq];
    if (p == UINT32_MAX || q == UINT32_MAX) return 0.0;
    const double s = Z.sc[p] * Z.sc[q];
    if (p == q) return Z.zd[p] * s;
    if (p < q) {
        const uint32_t t = p;
        p = q;
        q = t;
    }
    uint32_t lo = Z.zp[q], hi = Z.zp[q + 1];  // row p in column q
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (Z.zi[mid] < p) lo = mid + 1;
        else hi = mid;
    }
    return (lo < Z.zp[q + 1] && Z.zi[lo] == p) ? Z.zx[lo] * s : 0.0;
}

struct LevArgs {
    const uint32_t* seg_off;
    const uint32_t* codes;  // M x nb
    const double* cw;       // null = unit
    const double* inv_n1;
    uint64_t g1;
    int nb;
    uint64_t boff[NB_MAX];
};

// One wave per segment: hcell[c] = 1/n1_s + u_c' S^+ u_c for its cells (has_z = false: no reduced
// coordinates, 1/n1_s).
__global__ __launch_bounds__(TPB) void k_fe_leverage(LevArgs A, ZDev Z, int has_z, double* __restrict__ hcell) {
    const int lane = lane_id();
    for (uint64_t s = ((uint64_t)blockIdx.x * TPB + threadIdx.x) / WAVE; s < A.g1;
         s += (uint64_t)gridDim.x * (TPB / WAVE)) {
        const uint64_t c0 = A.seg_off[s], c1 = A.seg_off[s + 1];
        const double in1 = A.inv_n1[s];
        if (!has_z) {
            for (uint64_t c = c0 + lane; c < c1; c += WAVE) hcell[c] = in1;
            continue;
        }
        double bpart = 0.0;
        for (uint64_t c = c0 + lane; c < c1; c += WAVE) {
            const double wc = A.cw ? A.cw[c] : 1.0;
            double inner = 0.0, ac = 0.0;
            for (int b = 0; b < A.nb; b++) {
                const uint64_t kt = A.boff[b] + A.codes[c * (uint64_t)A.nb + b];
                for (int b2 = 0; b2 < A.nb; b2++)
                    inner += zlook(Z, kt, A.boff[b2] + A.codes[c * (uint64_t)A.nb + b2]);
                double at = 0.0;
                for (uint64_t c2 = c0; c2 < c1; c2++) {
                    const double w2 = A.cw ? A.cw[c2] : 1.0;
                    for (int b2 = 0; b2 < A.nb; b2++)
                        at += w2 * zlook(Z, kt, A.boff[b2] + A.codes[c2 * (uint64_t)A.nb + b2]);
                }
                ac += at;
            }
            bpart += wc * ac;
            hcell[c] = inner - 2.0 * in1 * ac;
        }
        const double beta = wave_sum(bpart) * in1 * in1;
        for (uint64_t c = c0 + lane; c < c1; c += WAVE) hcell[c] += beta + in1;
    }
}

constexpr int HC_KMAX = 32;

// Per cell, its rows: h = c (hcell + x~' B x~), the check h <= 1 - sqrt(eps), the scaled residual.
__global__ void k_hc_scores(const uint32_t* __restrict__ cell_obs_off, uint64_t M, const double* __restrict__ hcell,
                            const double* __restrict__ rows, int ld, const int32_t* __restrict__ kcol, int nk,
                            const double* __restrict__ B, const double* __restrict__ w, int row_weight,
                            const double* __restrict__ e, uint64_t es, int hc3, double* __restrict__ e_out,
                            uint32_t* __restrict__ bad) {
    const double thr = 1.0 - sqrt(2.220446049250313e-16);
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r0 = cell_obs_off ? cell_obs_off[c] : c, r1 = cell_obs_off ? cell_obs_off[c + 1] : c + 1;
        const double hc = hcell[c];
        for (uint64_t i = r0; i < r1; i++) {
            double x[HC_KMAX];
            for (int j = 0; j < nk; j++) x[j] = rows[i * (uint64_t)ld + 1 + kcol[j]];
            double q = 0.0;
            for (int j = 0; j < nk; j++) {
                double t = 0.0;
                for (int l = 0; l < nk; l++) t += B[j * nk + l] * x[l];
                q += x[j] * t;
            }
            const double h = (hc + q) * (row_weight && w ? w[i] : 1.0);
            if (!(h <= thr)) atomicOr(bad, 1u);
            const double d = 1.0 - h;
            e_out[i] = e[i * es] / (hc3 ? d : sqrt(d));
        }
    }
}

}  // namespace

void fe_leverage_cells(Absorber& a, double* hcell, const fh_ctx* ctx) {
    hipStream_t st = a.st;
    LevArgs A{};
    A.seg_off = a.seg_off.p;
    A.codes = a.cell_codes.p;
    A.cw = a.unit() ? nullptr : a.w64.cw.p;
    A.inv_n1 = a.w64.inv_n1.p;
    A.g1 = a.g1;
    A.nb = (int)a.nb;
    for (size_t b = 0; b < a.nb; b++) A.boff[b] = a.blocks[b].off;
    ZDev Z{};
    LdlFactor f;
    std::vector<uint32_t> keep;
    DBuf<uint32_t> pos, zp, zi;
    DBuf<double> zx, zd, sc;
    const bool has_z = a.m > 0;
    if (has_z) {
        std::string why;
        require(factor_joint_host(a, f, keep, why), "leverage: " + why);
        check_interrupt(ctx);
        std::vector<double> hzd, hzx;
        ldl_selected_inverse(f, hzd, hzx);
        std::vector<uint32_t> hpos(a.m, UINT32_MAX);
        for (size_t k = 0; k < f.n; k++) hpos[keep[f.perm[k]]] = (uint32_t)k;
        require(f.len() < (1ull << 32), "leverage: factor too large for 32-bit offsets", FH_INTERNAL);
        std::vector<uint32_t> hzp(f.n + 1), hzi(f.li.begin(), f.li.end());
        for (size_t k = 0; k <= f.n; k++) hzp[k] = (uint32_t)f.lp[k];
        auto put32 = [&](DBuf<uint32_t>& d, const std::vector<uint32_t>& h) {
            d.alloc(std::max<size_t>(h.size(), 1));
            if (!h.empty()) h2d(d.p, h.data(), h.size(), st);
        };
        auto put64 = [&](DBuf<double>& d, const std::vector<double>& h) {
            d.alloc(std::max<size_t>(h.size(), 1));
            if (!h.empty()) h2d(d.p, h.data(), h.size(), st);
        };
        put32(pos, hpos);
        put32(zp, hzp);
        put32(zi, hzi);
        put64(zx, hzx);
        put64(zd, hzd);
        put64(sc, f.scale);
        Z = ZDev{pos.p, zp.p, zi.p, zx.p, zd.p, sc.p};
        log_line(ctx, 2, "fereg: leverage: joint factor n=" + std::to_string(f.n) + " L=" + std::to_string(f.len()));
    }
    const unsigned grid = grid_for(a.g1, TPB / WAVE, 1u << 16);
    hipLaunchKernelGGL(k_fe_leverage, dim3(grid), dim3(TPB), 0, st, A, Z, has_z ? 1 : 0, hcell);
    FH_CHECK_LAUNCH();
    stream_wait(st);  // the host vectors behind the uploads go out of scope
}

void hc_scores(Absorber& a, const double* rows, int ld, const std::vector<int>& kept, const std::vector<double>& bread,
               const double* e, size_t es, const double* w, bool row_weight, bool hc3, double* e_out) {
    hipStream_t st = a.st;
    require(!a.slope, "HC2/HC3 with varying slopes are not supported (the leverage of slope terms is not built)");
    const int nk = (int)kept.size();
    require(nk <= HC_KMAX, "HC2/HC3: at most 32 kept regressors");
    DBuf<double> hcell(std::max<uint64_t>(a.M, 1)), B(std::max(nk * nk, 1));
    DBuf<int32_t> kc(std::max(nk, 1));
    DBuf<uint32_t> bad(1);
    fe_leverage_cells(a, hcell.p, nullptr);
    std::vector<int32_t> hk(kept.begin(), kept.end());
    if (nk) {
        h2d(kc.p, hk.data(), nk, st);
        h2d(B.p, bread.data(), (size_t)nk * nk, st);
    }
    FH_HIP(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_hc_scores, dim3(grid_for(a.M, TPB, 8192)), dim3(TPB), 0, st,
                       a.trivial() ? nullptr : a.cell_obs_off.p, a.M, hcell.p, rows, ld, kc.p, nk, B.p, w,
                       row_weight ? 1 : 0, e, es, hc3 ? 1 : 0, e_out, bad.p);
    FH_CHECK_LAUNCH();
    uint32_t hb = 0;
    d2h(&hb, bad.p, 1, st);
    require(hb == 0,
            "When calculating the diagonals of the projection matrix, a leverage was found to equal 1 (a fixed "
            "effect or dummy with one observation): HC2/HC3 are not defined");
}

}  // namespace fh
