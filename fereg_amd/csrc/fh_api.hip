// C-ABI entry points of fereg_hip (include/fereg_hip.h) and the host orchestration of a
// residualization (solve.rs:1141-1311) and an OLS fit (fit/ols.rs:213-330, fit/vcov.rs:63-382);
// the Poisson IRLS is in fh_irls.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>

#include "fh_fit.hpp"
#include "fh_sort.hpp"

namespace fh {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

void ensure_w32(Absorber&) {}  // refresh_weights fills both precisions

namespace {

constexpr int TPB = 256;
inline unsigned vgrid(size_t n) { return grid_for(n, TPB, 8192); }

// rows[i][0] = y[pi[i]], rows[i][1 + k] = X[pi[i]][k], rows[i][K+1 .. ld) = 0: one thread per
// output element (index type I), so the writes (and, for near-identity pi, the reads of X) are
// coalesced.
template <class I>
__global__ void k_gather_rows(const double* __restrict__ y, const double* __restrict__ X, I K,
                              const uint32_t* __restrict__ pi, I n, I ld, double* __restrict__ rows) {
    const I total = n * ld, stride = (I)gridDim.x * blockDim.x;
    for (I e = (I)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const I i = e / ld, c = e - i * ld;
        const size_t o = pi[i];
        if (c == 0) {
            if (y) rows[e] = y[o];
        } else {
            rows[e] = c <= K ? X[o * K + c - 1] : 0.0;
        }
    }
}

// The same with one thread per row for ld in {2, 4, 8, 12, 16}: the row's loads are issued
// together (element threads each wait on pi, then on X) and the row is stored as 16-byte pairs.
// X2: K even and X 16-byte aligned, so the random X row is read as K/2 loads of 16 bytes (one
// request each) instead of K.
template <int LD, bool X2>
__global__ void k_gather_rows_row(const double* __restrict__ y, const double* __restrict__ X, int K,
                                  const uint32_t* __restrict__ pi, size_t n, double* __restrict__ rows) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t o = pi[i];
        double v[LD];
        v[0] = y[o];
        if constexpr (X2) {
            const double2* xr = reinterpret_cast<const double2*>(X + o * K);
#pragma unroll
            for (int q = 0; q < (LD - 1) / 2; q++) {
                const double2 x2 = 2 * q + 1 < K ? xr[q] : make_double2(0.0, 0.0);
                v[1 + 2 * q] = x2.x;
                v[2 + 2 * q] = x2.y;
            }
            if constexpr (LD % 2 == 0) v[LD - 1] = 0.0;  // the pad column (K <= LD - 2 here)
        } else {
#pragma unroll
            for (int c = 1; c < LD; c++) v[c] = c <= K ? X[o * K + c - 1] : 0.0;
        }
        double2* dst = reinterpret_cast<double2*>(rows + i * LD);
#pragma unroll
        for (int q = 0; q < LD / 2; q++) dst[q] = make_double2(v[2 * q], v[2 * q + 1]);
    }
}

void gather_rows(const double* y, const double* X, size_t K, const uint32_t* pi, size_t n, size_t ld, double* rows,
                 hipStream_t st) {
    if (y && (ld == 2 || ld == 4 || ld == 8 || ld == 12 || ld == 16) && K + 1 <= ld) {
        const unsigned g = grid_for(n, TPB, 1u << 16);
        const bool x2 = K % 2 == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0;
        auto go = [&](auto ldc) {
            constexpr int LDC = decltype(ldc)::value;
            if (x2)
                hipLaunchKernelGGL((k_gather_rows_row<LDC, true>), dim3(g), dim3(TPB), 0, st, y, X, (int)K, pi, n, rows);
            else
                hipLaunchKernelGGL((k_gather_rows_row<LDC, false>), dim3(g), dim3(TPB), 0, st, y, X, (int)K, pi, n, rows);
        };
        if (ld == 2) go(std::integral_constant<int, 2>{});
        else if (ld == 4) go(std::integral_constant<int, 4>{});
        else if (ld == 8) go(std::integral_constant<int, 8>{});
        else if (ld == 12) go(std::integral_constant<int, 12>{});
        else go(std::integral_constant<int, 16>{});
        FH_CHECK_LAUNCH();
        return;
    }
    const unsigned grid = grid_for(n * ld, TPB, 1u << 16);
    if (n * ld < (1ull << 32) - (size_t)grid * TPB)
        hipLaunchKernelGGL(k_gather_rows<uint32_t>, dim3(grid), dim3(TPB), 0, st, y, X, (uint32_t)K, pi, (uint32_t)n,
                           (uint32_t)ld, rows);
    else
        hipLaunchKernelGGL(k_gather_rows<size_t>, dim3(grid), dim3(TPB), 0, st, y, X, K, pi, n, ld, rows);
    FH_CHECK_LAUNCH();
}

// flag = 1 when some code >= g (structure.rs:483-489) or some weight is not positive and finite.
// The same for up to NB_MAX + 1 terms in one launch (blockIdx.y = term), 16-byte loads where the
// array is 16-byte aligned.
struct CodeCheck {
    const uint32_t* c[NB_MAX + 1];
    uint64_t g[NB_MAX + 1];
};
__global__ void k_check_codes_all(CodeCheck cc, size_t n, uint32_t* flag) {
    const uint32_t* c = cc.c[blockIdx.y];
    const uint64_t g = cc.g[blockIdx.y];
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    bool bad = false;
    size_t head = 0;
    if ((reinterpret_cast<uintptr_t>(c) & 15) == 0) {
        const size_t nq = n / 4;
        const uint4* c4 = reinterpret_cast<const uint4*>(c);
        for (size_t i = t0; i < nq; i += stride) {
            const uint4 v = c4[i];
            bad |= v.x >= g || v.y >= g || v.z >= g || v.w >= g;
        }
        head = nq * 4;
    }
    for (size_t i = head + t0; i < n; i += stride) bad |= c[i] >= g;
    if (__ballot(bad) && lane_id() == 0) atomicOr(flag, 1u);
}
__global__ void k_check_weights(const double* __restrict__ w, size_t n, uint32_t* flag) {
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bad |= !(w[i] > 0.0) || !isfinite(w[i]);
    if (__ballot(bad) && lane_id() == 0) atomicOr(flag, 1u);
}
__global__ void k_max_u32(const uint32_t* __restrict__ c, size_t n, uint32_t* out) {
    uint32_t m = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 7 * stride < n; i += 8 * stride) {  // 8 independent loads in flight per lane
        uint32_t v[8];
#pragma unroll
        for (int h = 0; h < 8; h++) v[h] = c[i + h * stride];
#pragma unroll
        for (int h = 0; h < 8; h++) m = max(m, v[h]);
    }
    for (; i < n; i += stride) m = max(m, c[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, WAVE));
    __shared__ uint32_t wm[TPB / WAVE];
    if (lane_id() == 0) wm[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {  // one atomic per workgroup
        uint32_t b = 0;
        for (int w = 0; w < TPB / WAVE; w++) b = max(b, wm[w]);
        atomicMax(out, b);
    }
}
__global__ void k_gather_vec(const double* __restrict__ v, const uint32_t* __restrict__ pi, size_t n,
                             double* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = v[pi[i]];
}

// Columns [c0, c0+rc) of an n x ld block <-> a contiguous n x rc block.
__global__ void k_cols_out(const double* __restrict__ z, size_t n, int ld, int c0, int rc, double* __restrict__ out) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * rc; e += (size_t)gridDim.x * blockDim.x)
        out[e] = z[(e / rc) * ld + c0 + e % rc];
}
__global__ void k_cols_in(double* __restrict__ z, size_t n, int ld, int c0, int rc, const double* __restrict__ in) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * rc; e += (size_t)gridDim.x * blockDim.x)
        z[(e / rc) * ld + c0 + e % rc] = in[e];
}

__global__ void k_zero_cols(double* __restrict__ x, size_t m, int r, const int32_t* __restrict__ zero) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m * r; e += (size_t)gridDim.x * blockDim.x)
        if (zero[e % r]) x[e] = 0.0;
}

__global__ void k_wsq_b1(const double* __restrict__ b1, const double* __restrict__ n1, size_t g1, int r,
                         double* __restrict__ part) {
    // part[blockIdx.x * r + j] = sum_s b1^2 / n1 over the block's segments (fixed grid)
    __shared__ double red[TPB / WAVE][16];
    double acc[16];
    for (int j = 0; j < 16; j++) acc[j] = 0.0;
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < g1; s += (size_t)gridDim.x * blockDim.x)
        for (int j = 0; j < r; j++) {
            const double v = b1[s * r + j];
            if (n1[s] > 0.0) acc[j] += v * v / n1[s];  // empty FE1 levels: b1 = 0, no term
        }
    for (int j = 0; j < 16; j++)
        if (j < r) {
            double v = wave_sum(acc[j]);
            if (lane_id() == 0) red[threadIdx.x / WAVE][j] = v;
        }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        double v = 0.0;
        for (int q = 0; q < TPB / WAVE; q++) v += red[q][threadIdx.x];
        part[(size_t)blockIdx.x * r + threadIdx.x] = v;
    }
}

// fit_rows (solve.rs:1313-1317): out[i] = alpha1[s(c)] + sum_b x[off_b + code_b(c)] for every row i
// of cell c, column col of the (m + g1) x r coefficients; one thread per cell.
struct BlockOffs {
    uint32_t off[NB_MAX];
};
__global__ void k_fit_rows(const double* __restrict__ coefs, int r, int col, const uint32_t* __restrict__ cell_codes,
                           const uint32_t* __restrict__ cell_seg, const uint32_t* __restrict__ cell_obs_off,
                           size_t M, int nb, BlockOffs bo, size_t m, double* __restrict__ out) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (size_t)gridDim.x * blockDim.x) {
        double v = coefs[(m + cell_seg[c]) * r + col];
        for (int b = 0; b < nb; b++) v += coefs[((size_t)bo.off[b] + cell_codes[c * nb + b]) * r + col];
        const size_t r0 = cell_obs_off ? cell_obs_off[c] : c, r1 = cell_obs_off ? cell_obs_off[c + 1] : c + 1;
        for (size_t i = r0; i < r1; i++) out[i] = v;
    }
}

}  // namespace

// Rows [y, X] of the kept-order inputs in solver row order, padded to ld columns with zeros.
void gather_block(const double* y, const double* X, size_t K, const uint32_t* pi, size_t n, size_t ld, double* rows,
                  hipStream_t st) {
    gather_rows(y, X, K, pi, n, ld, rows, st);
}

double sum_parts(const double* dev, size_t count, hipStream_t st) {
    std::vector<double> h(count);
    d2h(h.data(), dev, count, st);
    stream_wait(st);
    double s = 0.0;
    for (double v : h) s += v;
    return s;
}

int guard(const std::function<void()>& f) {
    try {
        f();
        return FH_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.rc;
    } catch (const std::bad_alloc&) {
        set_last_error("out of host memory");
        return FH_OOM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return FH_INTERNAL;
    }
}

void check_interrupt(const fh_ctx* ctx) {
    if (ctx && ctx->interrupt && ctx->interrupt(ctx->user)) throw Error(FH_INTERRUPTED, "interrupted");
}

void log_line(const fh_ctx* ctx, int level, const std::string& s) {
    if (ctx && ctx->log && ctx->verbose >= level) ctx->log(ctx->user, s.c_str());
}

bool weights_ok(Absorber& a, const double* w) {
    DBuf<uint32_t> bad(1);
    FH_HIP(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), a.st));
    hipLaunchKernelGGL(k_check_weights, dim3(1024), dim3(TPB), 0, a.st, w, a.n, bad.p);
    FH_CHECK_LAUNCH();
    uint32_t h = 0;
    d2h(&h, bad.p, 1, a.st);
    stream_wait(a.st);
    return h == 0;
}

void restore_user_weights(Absorber& a) {
    if (!a.irls_weights) return;
    a.irls_weights = false;
    if (a.user_w.n) {
        a.w_rows = std::move(a.user_w);
        refresh_weights(a, a.w_rows.p);
    } else {
        refresh_weights(a, nullptr);
    }
}

// ------------------------------------------------------------------ residualize


// One residualization of a device block of r <= 16 columns (solve.rs:1141-1311). The solver's
// vectors carry rp = padded_cols(r) columns (zero, inactive pads) so that their rows are whole
// vectors; the caller's block, warm start and coefficients keep r columns.
ResidOut residualize_block(Absorber& a, double* z, int r, double tol, uint64_t maxiter, const double* warm_dev,
                           double* coefs_dev, int precision, const fh_ctx* ctx) {
    if (a.slope) return slope_residualize(a, z, r, tol, maxiter, warm_dev, coefs_dev, ctx);
    hipStream_t st = a.st;
    const size_t m = a.m, g1 = a.g1;
    const int rp = padded_cols(r);
    ResidOut out;
    a.b1.reserve(g1 * rp);
    a.s64.rhs.reserve(std::max<size_t>(m * rp, 1));
    a.s64.x.reserve(std::max<size_t>(m * rp, 1));
    // ss0 = wsq - wsum^2/sw (0 if <= 1e-15 wsq), ss1 = wsq - sum_s b1^2/n1 (solve.rs:1210-1217);
    // the partials of wsq / wsum (rhs pass) and of sum b1^2/n1 are read back in one transfer
    const unsigned pg = 256;
    const unsigned grid = rhs_pass(a, z, r, rp, a.s64.rhs.p, a.b1.p, (size_t)pg * rp);
    const size_t nws = (size_t)grid * 2 * rp;
    hipLaunchKernelGGL(k_wsq_b1, dim3(pg), dim3(TPB), 0, st, a.b1.p, a.n1.p, g1, rp, a.s64.dpart.p + nws);
    FH_CHECK_LAUNCH();
    // one fixed-order device reduction for each, one readback
    DBuf<double> sums(3 * (size_t)rp + 1);
    sum_parts_to(a.s64.dpart.p, grid, (size_t)(2 * rp), sums.p, st);
    sum_parts_to(a.s64.dpart.p + nws, pg, (size_t)rp, sums.p + 2 * rp, st);
    const bool with_sw = !a.sw_ready;  // the weight sum rides along when it is still on the device
    if (with_sw)
        FH_HIP(hipMemcpyAsync(sums.p + 3 * rp, a.sw_dev.p, sizeof(double), hipMemcpyDeviceToDevice, st));
    std::vector<double> hs(3 * (size_t)rp + (with_sw ? 1 : 0));
    d2h(hs.data(), sums.p, hs.size(), st);
    if (with_sw) {
        a.sw = hs[3 * rp];
        a.sw_ready = true;
    }
    check_interrupt(ctx);
    const std::vector<double> ws(hs.begin(), hs.begin() + 2 * rp);
    const double* hp = hs.data() + 2 * rp;
    std::vector<double> ss0(rp), ss1(rp);
    std::vector<int32_t> zero(rp);
    for (int j = 0; j < rp; j++) {  // pad columns: all sums zero, so ss0 = 0 (inactive)
        const double b1sq = hp[j];
        const double wsq = ws[j], wsum = ws[rp + j];
        const double v = wsq - wsum * wsum / absorber_sw(a);
        ss0[j] = v > 1e-15 * wsq ? v : 0.0;
        ss1[j] = wsq - b1sq;
        zero[j] = ss0[j] == 0.0;
    }
    double* x = a.s64.x.p;
    if (m > 0) {
        if (warm_dev && rp == r) {
            FH_HIP(hipMemcpyAsync(x, warm_dev, m * r * sizeof(double), hipMemcpyDeviceToDevice, st));
        } else {
            FH_HIP(hipMemsetAsync(x, 0, m * rp * sizeof(double), st));
            if (warm_dev) {
                hipLaunchKernelGGL(k_cols_in, dim3(vgrid(m * r)), dim3(TPB), 0, st, x, m, rp, 0, r, warm_dev);
                FH_CHECK_LAUNCH();
            }
        }
        DBuf<int32_t> dz(rp);
        h2d(dz.p, zero.data(), rp, st);
        hipLaunchKernelGGL(k_zero_cols, dim3(vgrid(m * rp)), dim3(TPB), 0, st, x, m, rp, dz.p);
        FH_CHECK_LAUNCH();
        const bool warm = warm_dev != nullptr;
        auto solve = [&](bool from_x, bool can_esc, uint64_t mi, JointProbe* jp) {
            return precision == FH_MIXED
                       ? pcg_mixed(a, a.s64.rhs.p, x, from_x, rp, ss1.data(), ss0.data(), tol, mi, ctx, can_esc, jp)
                       : pcg_fp64(a, a.s64.rhs.p, x, from_x, rp, ss1.data(), ss0.data(), tol, mi, ctx, can_esc, jp);
        };
        // the early probe (iteration 6) may build the joint factor when the projected PCG time
        // exceeds its estimated cost; the escalation probe (iteration 12) factors the reduced
        // matrix when the rate is slow (cg.rs:254-274, solve.rs:932-975); either way the solve
        // restarts from the current iterate
        JointProbe jp;
        uint64_t used = 0;
        const bool joint_ok = escalation_possible(a) && a.joint_policy != 0 && !a.fac_failed_joint;
        PcgResult pr = solve(warm, escalation_possible(a), maxiter, joint_ok ? &jp : nullptr);
        if (jp.fired) {
            used += pr.iterations;
            out.escalated = escalate_joint(a, ctx, jp.without);
            pr = solve(true, escalation_possible(a), maxiter > used ? maxiter - used : 1, nullptr);
        }
        if (pr.escalate) {
            // with no factor within the budgets the solve restarts on Jacobi
            used += pr.iterations;
            out.escalated = escalate(a, ctx) || out.escalated;
            pr = solve(true, false, maxiter > used ? maxiter - used : 1, nullptr);
        }
        pr.iterations += used;
        if (pr.rc == FH_CONVERGENCE)
            throw Error(FH_CONVERGENCE, "the FE solver did not converge within " + std::to_string(maxiter) + " iterations");
        out.iterations = std::max<uint64_t>(pr.iterations, 1);
        bool allconst = true;
        for (int j = 0; j < r; j++) allconst &= ss0[j] == 0.0;
        out.exact = pr.exact || allconst;
        out.refinements = pr.refinements;
    }
    if (rp == r) {
        residualize_rows(a, z, r, rp, x, a.b1.p, coefs_dev ? coefs_dev + m * r : nullptr);
        if (coefs_dev && m > 0) FH_HIP(hipMemcpyAsync(coefs_dev, x, m * r * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
        DBuf<double> alpha;
        if (coefs_dev) alpha.alloc(std::max<size_t>(g1 * rp, 1));
        residualize_rows(a, z, r, rp, x, a.b1.p, coefs_dev ? alpha.p : nullptr);
        if (coefs_dev) {  // FeCoefs [x; alpha1] with the caller's r columns
            if (m > 0)
                hipLaunchKernelGGL(k_cols_out, dim3(vgrid(m * r)), dim3(TPB), 0, st, x, m, rp, 0, r, coefs_dev);
            hipLaunchKernelGGL(k_cols_out, dim3(vgrid(g1 * r)), dim3(TPB), 0, st, alpha.p, g1, rp, 0, r,
                               coefs_dev + m * r);
            FH_CHECK_LAUNCH();
        }
    }
    return out;
}

// Any r: columns in chunks of <= 16 (every column is an independent system).
ResidOut residualize_any(Absorber& a, double* z, size_t r, double tol, uint64_t maxiter, const double* warm_dev,
                         double* coefs_dev, int precision, const fh_ctx* ctx) {
    if (r <= 16) return residualize_block(a, z, (int)r, tol, maxiter, warm_dev, coefs_dev, precision, ctx);
    const size_t n = a.n, p = a.m + a.g1;
    ResidOut tot;
    tot.iterations = 0;
    DBuf<double> zc(n * 16), wc(p * 16), cc(p * 16);
    for (size_t c0 = 0; c0 < r; c0 += 16) {
        const int rc = (int)std::min<size_t>(16, r - c0);
        hipLaunchKernelGGL(k_cols_out, dim3(vgrid(n * rc)), dim3(TPB), 0, a.st, z, n, (int)r, (int)c0, rc, zc.p);
        if (warm_dev)
            hipLaunchKernelGGL(k_cols_out, dim3(vgrid(p * rc)), dim3(TPB), 0, a.st, warm_dev, p, (int)r, (int)c0, rc, wc.p);
        FH_CHECK_LAUNCH();
        ResidOut o = residualize_block(a, zc.p, rc, tol, maxiter, warm_dev ? wc.p : nullptr, coefs_dev ? cc.p : nullptr,
                                       precision, ctx);
        hipLaunchKernelGGL(k_cols_in, dim3(vgrid(n * rc)), dim3(TPB), 0, a.st, z, n, (int)r, (int)c0, rc, zc.p);
        if (coefs_dev)
            hipLaunchKernelGGL(k_cols_in, dim3(vgrid(p * rc)), dim3(TPB), 0, a.st, coefs_dev, p, (int)r, (int)c0, rc, cc.p);
        FH_CHECK_LAUNCH();
        tot.iterations = std::max(tot.iterations, o.iterations);
        tot.exact = tot.exact && o.exact;
        tot.refinements = std::max(tot.refinements, o.refinements);
        tot.escalated |= o.escalated;
    }
    return tot;
}

// ------------------------------------------------------------------ packed requests

namespace {

// rows[i][c] *= f[req(i)][c] for the solver rows of a packed solve (requests contiguous).
__global__ void k_scale_ranges(double* __restrict__ rows, size_t n, int ld, int R, const uint32_t* __restrict__ off,
                               int nreq, const double* __restrict__ f) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nreq;  // the request of row i: off[j] <= i < off[j + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (off[mid] <= i) lo = mid;
            else hi = mid;
        }
        for (int c = 0; c < R; c++) rows[i * ld + c] *= f[lo * R + c];
    }
}

}  // namespace

void ols_fit_packed(const std::vector<PackedOls*>& reqs, int precision) {
    const size_t nreq = reqs.size();
    size_t Q = 0, Kmax = 0, ntot = 0;
    double tol = INFINITY;
    uint64_t maxiter = 0;
    for (const PackedOls* r : reqs) {
        Q = std::max(Q, r->codes.size());
        Kmax = std::max(Kmax, r->K);
        ntot += r->n;
        tol = std::min(tol, r->tol);
        maxiter = std::max(maxiter, r->maxiter);
    }
    const size_t R = Kmax + 1;
    require(R <= 16 && Q >= 1 && Q <= NB_MAX + 1, "packed solve: 1..15 regressors and 1..9 terms");
    // packed codes (disjoint level ranges per request) and the scaled [y, X] block, kept order
    std::vector<std::vector<uint32_t>> pc(Q, std::vector<uint32_t>(ntot));
    std::vector<uint64_t> pg(Q, 0);
    std::vector<double> yp(ntot), xp(ntot * Kmax, 0.0), scale(nreq * R, 1.0);
    std::vector<uint32_t> off(nreq + 1, 0);
    size_t r0 = 0;
    for (size_t j = 0; j < nreq; j++) {
        const PackedOls& r = *reqs[j];
        off[j] = (uint32_t)r0;
        for (size_t q = 0; q < Q; q++) {
            if (q < r.codes.size()) {
                for (size_t i = 0; i < r.n; i++) pc[q][r0 + i] = (uint32_t)(r.codes[q][i] + pg[q]);
                pg[q] += r.g[q];
            } else {
                for (size_t i = 0; i < r.n; i++) pc[q][r0 + i] = (uint32_t)pg[q];
                pg[q] += 1;
            }
        }
        for (size_t c = 0; c <= r.K; c++) {
            double ss = 0.0;
            for (size_t i = 0; i < r.n; i++) {
                const double v = c == 0 ? r.y[i] : r.X[i * r.K + c - 1];
                ss += v * v;
            }
            const double sc = ss > 0.0 ? 1.0 / std::sqrt(ss / (double)r.n) : 1.0;
            scale[j * R + c] = sc;
            for (size_t i = 0; i < r.n; i++) {
                if (c == 0) yp[r0 + i] = r.y[i] * sc;
                else xp[(r0 + i) * Kmax + c - 1] = r.X[i * r.K + c - 1] * sc;
            }
        }
        r0 += r.n;
    }
    off[nreq] = (uint32_t)ntot;
    for (uint64_t gq : pg) require(gq < (1ull << 32), "packed solve: too many levels");
    std::vector<fh_term> terms(Q);
    for (size_t q = 0; q < Q; q++) terms[q] = fh_term{pc[q].data(), pg[q], 1};
    fh_solver_config cfg;
    fh_config_default(&cfg);
    cfg.precision = precision;
    fh_absorber* ab = nullptr;
    int rc = fh_absorber_create(terms.data(), Q, nullptr, ntot, &cfg, nullptr, &ab);
    if (rc != FH_OK) {
        for (PackedOls* r : reqs) r->rc = rc;
        return;
    }
    std::unique_ptr<fh_absorber, void (*)(fh_absorber*)> hold(ab, fh_absorber_destroy);
    Absorber& a = ab->a;
    AbsScope g(a);
    hipStream_t st = a.st;
    const size_t ld = padded_cols((int)R);
    DBuf<double> dy(ntot), dx(std::max<size_t>(ntot * Kmax, 1)), rows(ntot * ld), dscale(nreq * R), inv(nreq * R);
    DBuf<uint32_t> doff(nreq + 1);
    h2d(dy.p, yp.data(), ntot, st);
    h2d(dx.p, xp.data(), ntot * Kmax, st);
    gather_rows(dy.p, dx.p, Kmax, a.pi.p, ntot, ld, rows.p, st);
    // each request's FE1 levels come after the previous request's, so its rows stay contiguous
    // in the solver order (the structure sorts by FE1 first): the kept offsets are row offsets
    std::vector<double> unscale(nreq * R);
    for (size_t i = 0; i < nreq * R; i++) unscale[i] = 1.0 / scale[i];
    h2d(doff.p, off.data(), nreq + 1, st);
    h2d(inv.p, unscale.data(), nreq * R, st);
    const double tol_p = tol / std::sqrt((double)nreq);  // every request within its own tolerance
    ResidOut ro;
    try {
        ro = residualize_any(a, rows.p, ld, tol_p, maxiter, nullptr, nullptr, precision, nullptr);
    } catch (const Error& e) {
        for (PackedOls* r : reqs) r->rc = e.rc;
        return;
    }
    hipLaunchKernelGGL(k_scale_ranges, dim3(grid_for(ntot, TPB, 4096)), dim3(TPB), 0, st, rows.p, ntot, (int)ld, (int)R,
                       doff.p, (int)nreq, inv.p);
    FH_CHECK_LAUNCH();
    for (size_t j = 0; j < nreq; j++) {
        PackedOls& r = *reqs[j];
        r.iterations = ro.iterations;
        try {
            double* rj = rows.p + (size_t)off[j] * ld;
            const size_t K = r.K, Rj = K + 1, n = r.n;
            std::vector<double> G(Rj * Rj);
            gram_cols(a, rj, n, (int)ld, 0, (int)Rj, nullptr, nullptr, G.data());
            std::vector<double> Gxx(K * K);
            for (size_t i = 0; i < K; i++)
                for (size_t l = 0; l < K; l++) Gxx[i * K + l] = G[(i + 1) * Rj + l + 1];
            std::vector<int> excluded;
            std::vector<double> binv;
            chol_collin(Gxx, (int)K, 1e-9, excluded, binv);
            std::vector<int> kept;
            for (size_t i = 0; i < K; i++)
                if (!excluded[i]) kept.push_back((int)i);
            const int mk = (int)kept.size();
            require(mk > 0, "all the regressors are collinear (with the fixed effects)");
            std::vector<double> coef(mk, 0.0), c(Rj, 0.0);
            for (int i = 0; i < mk; i++)
                for (int l = 0; l < mk; l++) coef[i] += binv[(size_t)i * mk + l] * G[(size_t)(kept[l] + 1) * Rj];
            c[0] = 1.0;
            for (int i = 0; i < mk; i++) c[kept[i] + 1] = -coef[i];
            DBuf<double> e(n);
            const double ssr = residual_pass(a, rj, n, (int)ld, (int)Rj, c.data(), nullptr, e.p, -1);
            double fe_par = 1.0;  // the constant without absorbed terms
            if (!r.codes.empty()) {
                fe_par = 0.0;
                for (uint64_t gq : r.g) fe_par += (double)gq;
                fe_par -= (double)r.codes.size() - 1.0;
            }
            const double np = (double)mk + fe_par;
            const double dfk = r.ssc.k_fixef == FH_KFIXEF_NONE ? np - fe_par : np;
            std::vector<double> V;
            vcov_from(a, rj, (int)ld, (int)Rj, excluded, binv, e.p, -1, nullptr, r.kind, ClusterSet{}, r.ssc, dfk,
                      ssr / ((double)n - 1.0), V, nullptr, nullptr, false, n);
            for (size_t i = 0; i < K; i++) {
                r.beta[i] = 0.0;
                r.kept[i] = !excluded[i];
            }
            for (int i = 0; i < mk; i++) r.beta[kept[i]] = coef[i];
            std::memcpy(r.vcov, V.data(), K * K * sizeof(double));
            r.rc = FH_OK;
        } catch (const Error& e) {
            set_last_error(e.what());
            r.rc = e.rc;
        }
    }
    stream_wait(st);
}

// ------------------------------------------------------------------ estimation helpers (host, K x K)

static bool term_has_icpt(const Absorber& a, size_t q) { return q >= a.term_icpt.size() || a.term_icpt[q]; }
static double term_ns(const Absorber& a, size_t q) { return q < a.term_ns.size() ? (double)a.term_ns[q] : 0.0; }

// fixest nparams' FE part (fit/ols.rs:25-37): every term G (intercept + slopes), minus one per
// intercept term beyond the first.
double fe_params(const Absorber& a) {
    double s = 0.0;
    int nic = 0;
    for (size_t q = 0; q < a.nterms; q++) {
        const bool ic = term_has_icpt(a, q);
        s += (double)a.term_g[q] * ((ic ? 1.0 : 0.0) + term_ns(a, q));
        nic += ic ? 1 : 0;
    }
    return s - (double)std::max(nic - 1, 0);
}

// The FE parameters df.K removes (fit/vcov.rs:63-69): the intercept terms' levels (slope-only
// terms count 0) minus one per such term beyond the first.
static double df_size(const Absorber& a, size_t q) { return term_has_icpt(a, q) ? (double)a.term_g[q] : 0.0; }
static double fe_params_df(const Absorber& a, int* n_ok_out) {
    double s = 0.0;
    int n_ok = 0;
    for (size_t q = 0; q < a.nterms; q++) {
        s += df_size(a, q);
        n_ok += df_size(a, q) > 0.0 ? 1 : 0;
    }
    if (n_ok_out) *n_ok_out = n_ok;
    return n_ok > 0 ? s - ((double)n_ok - 1.0) : 0.0;
}

// fixest df.K (fit/vcov.rs:63-112, K.exact = FALSE).
double df_k(const Absorber& a, int kind, const fh_ssc& ssc, const int32_t* nested, double np, int kdesign) {
    const size_t nt = a.nterms;
    int n_ok = 0;
    const double fe_par = fe_params_df(a, &n_ok);
    int kf = ssc.k_fixef;
    if (kind != FH_VCOV_CLUSTER && kind != FH_VCOV_CRV3 && kf == FH_KFIXEF_NONNESTED) kf = FH_KFIXEF_FULL;
    if (kf == FH_KFIXEF_NONE) return np - fe_par;
    if (kf == FH_KFIXEF_FULL || nt == 0 || n_ok == 0) return np;
    int n_nested = 0;
    for (size_t q = 0; q < nt; q++) n_nested += nested && nested[q] ? 1 : 0;
    double k = np;
    if ((size_t)n_nested == nt) {
        k = k - fe_par + 1.0;
    } else {
        double s = 0.0, c = 0.0;
        for (size_t q = 0; q < nt; q++)
            if (nested && nested[q]) {
                s += df_size(a, q);
                c += df_size(a, q) > 0.0 ? 1.0 : 0.0;
            }
        k -= s - c;
    }
    return std::max(k, (double)kdesign + 1.0);
}

struct FitCommon {
    std::vector<double> beta, vcov;
    std::vector<int32_t> kept;
};

// V for a residualized block rows (n x R: [target, X...]) with e the residual combo applied
// (fit/vcov.rs:316-382). bread over kept columns; scores x_i e_i w_i.
// A fit that feeds a wild cluster bootstrap: the bootstrap's per-cluster sums are formed in the
// CRV1 step (their first block is the CRV1 score), so the rows are gathered by cluster once.
struct WbPre {
    std::vector<double> rmat;  // restriction over all K regressors
    DBuf<double> S;            // G x 3 x kept, set by vcov_from
    bool ready = false;
};

void vcov_from(Absorber& a, double* rows, int ld, int R, const std::vector<int>& excluded,
               const std::vector<double>& bread, const double* e, int ecol, const double* w, int kind,
               const ClusterSet& cs, const fh_ssc& ssc, double dfk, double iid_scale,
               std::vector<double>& V, uint64_t* ncl, WbPre* wb, bool glm, size_t nrows) {
    require(kind >= FH_VCOV_IID && kind <= FH_VCOV_HC3, "unknown variance kind " + std::to_string(kind));
    require(nrows == 0 || kind <= FH_VCOV_HETERO, "a row range supports iid and hetero variances");
    const size_t nr = nrows ? nrows : a.n;
    const int K = R - 1;
    std::vector<int> kept;
    for (int j = 0; j < K; j++)
        if (!excluded[j]) kept.push_back(j);
    const int mk = (int)kept.size();
    const double n = (double)nr;
    V.assign((size_t)K * K, 0.0);
    std::vector<double> vk((size_t)mk * mk, 0.0);
    if (mk == 0) return;
    if (kind == FH_VCOV_IID) {
        const double factor = ssc.k_adj ? (n - 1.0) / (n - dfk) : 1.0;
        for (size_t i = 0; i < vk.size(); i++) vk[i] = bread[i] * iid_scale * factor;
    } else {
        std::vector<double> meat((size_t)K * K);
        double factor = 1.0;
        if (kind == FH_VCOV_HETERO) {
            hetero_meat(a, rows, nr, ld, 1, K, e, w, meat.data());
            if (ssc.k_adj) factor = n / (n - dfk);
        } else if (kind == FH_VCOV_HC2 || kind == FH_VCOV_HC3) {
            // fit/vcov.rs:333-364, fixest flavour: the scores over sqrt(1 - h) or (1 - h) with the
            // full model's leverage, no small-sample factor; the regressor part of the leverage
            // carries the weight except in an OLS without absorbed terms (fit/vcov.rs:235)
            DBuf<double> es(a.n);
            hc_scores(a, rows, ld, kept, bread, ecol >= 0 ? rows + ecol : e, ecol >= 0 ? (size_t)ld : 1, w,
                      glm || a.nterms > 0, kind == FH_VCOV_HC3, es.p);
            hetero_meat(a, rows, a.n, ld, 1, K, es.p, w, meat.data());
        } else if (cs.dims.size() > 1) {
            // several dimensions (fit/vcov.rs:366-380): G.df "min" scales by the smallest G once,
            // "conventional" every subset meat by its own G/(G-1)
            const double* ep = ecol >= 0 ? rows + ecol : e;
            const size_t es = ecol >= 0 ? (size_t)ld : 1;
            const bool conventional = ssc.g_adj && cs.g_df == FH_GDF_CONVENTIONAL;
            const uint64_t gmin = multiway_meat(a, rows, ld, 1, K, ep, es, w, cs, conventional, meat.data());
            if (ncl) *ncl = gmin;
            if (ssc.g_adj && cs.g_df == FH_GDF_MIN) factor *= (double)gmin / ((double)gmin - 1.0);
            if (ssc.k_adj) factor *= (n - 1.0) / (n - dfk);
        } else {
            require(!cs.dims.empty() && cs.dims[0] != nullptr, "a clustered variance matrix needs cluster codes");
            const uint32_t* cl_dev = cs.dims[0];
            const uint64_t G = cs.G[0];
            // residuals from the rows' pad column when they are there: one row read per observation
            const double* ep = ecol >= 0 ? rows + ecol : e;
            const size_t es = ecol >= 0 ? (size_t)ld : 1;
            uint64_t g = 0;
            if (wb && mk + 2 <= 21) {
                // the bootstrap's cluster sums; their first block, sum w e x~ over the kept
                // regressors, is the CRV1 score (same runs, lanes and products: the same values)
                std::vector<double> rk(mk), C, MR;
                for (int i = 0; i < mk; i++) rk[i] = wb->rmat[kept[i]];
                wb_constraint(bread, mk, rk.data(), C, MR);
                wb_cluster_sums(a, rows, ld, kept, C, MR, ep, es, w, cl_dev, G, wb->S);
                wb->ready = true;
                std::vector<double> mkk0((size_t)mk * mk);
                gram_cols(a, wb->S.p, G, 3 * mk, 0, mk, nullptr, nullptr, mkk0.data());
                std::fill(meat.begin(), meat.end(), 0.0);
                for (int i = 0; i < mk; i++)
                    for (int j = 0; j < mk; j++) meat[(size_t)kept[i] * K + kept[j]] = mkk0[(size_t)i * mk + j];
                g = count_nonempty_clusters(a, cl_dev, G);
            } else {
                g = cluster_meat(a, rows, a.n, ld, 1, K, ep, es, w, cl_dev, G, meat.data());
            }
            if (ncl) *ncl = g;
            if (ssc.g_adj) factor *= (double)g / ((double)g - 1.0);
            if (ssc.k_adj) factor *= (n - 1.0) / (n - dfk);
        }
        std::vector<double> mkk((size_t)mk * mk), t((size_t)mk * mk, 0.0);
        for (int i = 0; i < mk; i++)
            for (int j = 0; j < mk; j++) mkk[(size_t)i * mk + j] = meat[(size_t)kept[i] * K + kept[j]];
        for (int i = 0; i < mk; i++)
            for (int j = 0; j < mk; j++) {
                double s = 0.0;
                for (int l = 0; l < mk; l++) s += bread[(size_t)i * mk + l] * mkk[(size_t)l * mk + j];
                t[(size_t)i * mk + j] = s;
            }
        for (int i = 0; i < mk; i++)
            for (int j = 0; j < mk; j++) {
                double s = 0.0;
                for (int l = 0; l < mk; l++) s += t[(size_t)i * mk + l] * bread[(size_t)l * mk + j];
                vk[(size_t)i * mk + j] = s * factor;
            }
    }
    // fixest repairs a matrix with an eigenvalue <= 0 (fit/vcov.rs:487-491, every kind)
    bool finite = true;
    for (double v : vk) finite &= std::isfinite(v);
    if (finite) posdef_fix(vk, mk);
    for (int i = 0; i < mk; i++)
        for (int j = 0; j < mk; j++) V[(size_t)kept[i] * K + kept[j]] = vk[(size_t)i * mk + j];
}

ClusterSet cluster_set(Absorber& a, const uint32_t* const* dims, size_t ndim, int g_df) {
    require(ndim <= 4, "at most four cluster dimensions are supported");
    require(g_df == FH_GDF_MIN || g_df == FH_GDF_CONVENTIONAL, "unknown G.df rule");
    ClusterSet cs;
    cs.g_df = g_df;
    for (size_t d = 0; d < ndim; d++) {
        require(dims[d] != nullptr, "null cluster codes");
        cs.dims.push_back(dims[d]);
        cs.G.push_back(cluster_levels(a, dims[d]));
    }
    return cs;
}

// Cluster levels: max code + 1.
uint64_t cluster_levels(Absorber& a, const uint32_t* cl_dev) {
    require(cl_dev != nullptr, "a clustered variance matrix needs cluster codes");
    if (a.cl_runs && a.cl_runs->cl == cl_dev) return a.cl_runs->G;
    DBuf<uint32_t> mx(1);
    FH_HIP(hipMemsetAsync(mx.p, 0, sizeof(uint32_t), a.st));
    hipLaunchKernelGGL(k_max_u32, dim3(grid_for(a.n, TPB * 8, 1024)), dim3(TPB), 0, a.st, cl_dev, a.n, mx.p);
    FH_CHECK_LAUNCH();
    uint32_t h = 0;
    d2h(&h, mx.p, 1, a.st);
    stream_wait(a.st);
    return (uint64_t)h + 1;
}

// What the bootstrap needs from the OLS fit.
struct OlsState {
    std::vector<int> kept;
    std::vector<double> inv, coef;  // bread and beta over the kept columns
    DBuf<double> e;                  // residuals, row order
    double dfk = 0.0;
    uint64_t ncl = 0;
    int ld = 0, ecol = -1;  // stride of the solver rows, their residual column (-1: none)
    WbPre* wb = nullptr;    // set by the bootstrap: form its cluster sums in the CRV1 step
};

// OLS on device inputs (kept order). rows scratch: n x (1+K) in row order.
void ols_fit(Absorber& a, const double* y, const double* X, size_t K, int kind, const ClusterSet& cs,
             const int32_t* nested, const fh_ssc& ssc, const fh_solve_opts& o, double* beta, double* vcov,
             int32_t* kept_out, fh_fit_stats* stats, const fh_ctx* ctx, OlsState* state = nullptr) {
    hipStream_t st = a.st;
    const size_t n = a.n, R = K + 1;
    require(K >= 1 && K <= 21, "1..21 regressors are supported");
    require((kind != FH_VCOV_CLUSTER && kind != FH_VCOV_CRV3) || !cs.dims.empty(),
            "a clustered variance matrix needs cluster codes");
    require(kind >= FH_VCOV_IID && kind <= FH_VCOV_CRV3, "unknown variance kind " + std::to_string(kind));
    // solver rows padded to the solver's column class (zero pads, residualized as inactive
    // columns): whole 16-byte vectors per row; a pad column takes the residuals, so the cluster
    // sums read one aligned row per observation
    const size_t ld = R <= 16 ? (size_t)padded_cols((int)R) : R;
    const int ecol = ld > R ? (int)R : -1;
    a.rows_tmp.reserve(n * ld);
    double* rows = a.rows_tmp.p;
    gather_rows(y, X, K, a.pi.p, n, ld, rows, st);
    const int prec = o.precision < 0 ? a.precision : o.precision;
    DBuf<double> raw;  // CRV3 refits start from the raw rows
    if (kind == FH_VCOV_CRV3) {
        raw.alloc(n * ld);
        FH_HIP(hipMemcpyAsync(raw.p, rows, n * ld * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    ResidOut ro = residualize_any(a, rows, ld, o.tol, o.maxiter, nullptr, nullptr, prec, ctx);
    const double* w = a.weighted ? a.w_rows.p : nullptr;
    std::vector<double> G(R * R);
    gram_cols(a, rows, n, (int)ld, 0, (int)R, w, nullptr, G.data());
    // regress column 0 on 1..K (fit/design.rs:64-95)
    std::vector<double> Gxx(K * K);
    for (size_t i = 0; i < K; i++)
        for (size_t j = 0; j < K; j++) Gxx[i * K + j] = G[(i + 1) * R + (j + 1)];
    std::vector<int> excluded;
    std::vector<double> inv;
    chol_collin(Gxx, (int)K, 1e-9, excluded, inv);
    std::vector<int> kept;
    for (size_t j = 0; j < K; j++)
        if (!excluded[j]) kept.push_back((int)j);
    const int mk = (int)kept.size();
    require(mk > 0, "all the regressors are collinear (with the fixed effects)");
    std::vector<double> coef(mk, 0.0);
    for (int i = 0; i < mk; i++) {
        double s = 0.0;
        for (int j = 0; j < mk; j++) s += inv[(size_t)i * mk + j] * G[(size_t)(kept[j] + 1) * R];
        coef[i] = s;
    }
    std::vector<double> c(R, 0.0);
    c[0] = 1.0;
    for (int i = 0; i < mk; i++) c[kept[i] + 1] = -coef[i];
    DBuf<double> e(n);
    const double ssr = residual_pass(a, rows, n, (int)ld, (int)R, c.data(), w, e.p, ecol);
    // fixed-effect parameters; without absorbed terms the absorber's constant is fixest's
    // intercept, one parameter
    const double np = (double)mk + (a.nterms > 0 ? fe_params(a) : 1.0);
    const double dfk = df_k(a, kind, ssc, nested, np, mk);
    std::vector<double> V;
    uint64_t ncl = 0;
    if (kind == FH_VCOV_CRV3) {
        // each refit: the cluster's rows get weight zero, the raw rows are residualized again and
        // the kept regressors' normal equations solved (crv3.rs: every FE re-estimated)
        DBuf<double> w0, tmp(n * ld);
        const bool had_w = a.weighted;
        if (had_w) {
            w0.alloc(n);
            FH_HIP(hipMemcpyAsync(w0.p, a.w_rows.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        CrvRefit refit = [&](const double* wm, uint64_t g) {
            refresh_weights(a, wm);
            FH_HIP(hipMemcpyAsync(tmp.p, raw.p, n * ld * sizeof(double), hipMemcpyDeviceToDevice, st));
            residualize_any(a, tmp.p, ld, o.tol, o.maxiter, nullptr, nullptr, prec, ctx);
            std::vector<double> Gg(R * R);
            gram_cols(a, tmp.p, n, (int)ld, 0, (int)R, a.w_rows.p, nullptr, Gg.data());
            std::vector<double> sub((size_t)mk * mk), ginv;
            for (int i = 0; i < mk; i++)
                for (int j = 0; j < mk; j++) sub[(size_t)i * mk + j] = Gg[(size_t)(kept[i] + 1) * R + kept[j] + 1];
            std::vector<int> ex;
            chol_collin(sub, mk, 1e-9, ex, ginv);
            for (int i = 0; i < mk; i++)
                require(!ex[i], "CRV3: coefficient " + std::to_string(kept[i]) +
                                    " is not estimable when cluster " + std::to_string(g) + " is left out");
            std::vector<double> b(mk, 0.0);
            for (int i = 0; i < mk; i++)
                for (int j = 0; j < mk; j++) b[i] += ginv[(size_t)i * mk + j] * Gg[(size_t)(kept[j] + 1) * R];
            return b;
        };
        std::vector<double> Vk;
        crv3_vcov(a, cs, ssc, dfk, had_w ? w0.p : nullptr, coef, refit, Vk, &ncl);
        refresh_weights(a, had_w ? w0.p : nullptr);
        V.assign(K * K, 0.0);
        for (int i = 0; i < mk; i++)
            for (int j = 0; j < mk; j++) V[(size_t)kept[i] * K + kept[j]] = Vk[(size_t)i * mk + j];
    } else {
        vcov_from(a, rows, (int)ld, (int)R, excluded, inv, e.p, ecol, w, kind, cs, ssc, dfk,
                  ssr / ((double)n - 1.0), V, &ncl, state ? state->wb : nullptr);
    }
    for (size_t j = 0; j < K; j++) {
        beta[j] = 0.0;
        kept_out[j] = !excluded[j];
    }
    for (int i = 0; i < mk; i++) beta[kept[i]] = coef[i];
    std::memcpy(vcov, V.data(), K * K * sizeof(double));
    if (stats) {
        stats->ssr = ssr;
        stats->df_k = dfk;
        stats->nparams = np;
        stats->deviance = 0.0;
        stats->iterations = ro.iterations;
        stats->pcg_iterations = ro.iterations;
        stats->n_clusters = ncl;
        stats->converged = 1;
    }
    if (state) {
        state->kept = kept;
        state->inv = inv;
        state->coef = coef;
        state->e = std::move(e);
        state->dfk = dfk;
        state->ncl = ncl;
        state->ld = (int)ld;
        state->ecol = ecol;
    }
}

}  // namespace fh

using namespace fh;

extern "C" {

const char* fh_last_error(void) { return g_last_error.c_str(); }

void fh_config_default(fh_solver_config* cfg) {
    cfg->small_limit = 4096;
    cfg->precision = FH_PLAIN_FP64;
    cfg->bucket_log2 = -1;
    cfg->autotune = 0;
    cfg->mixed_lead = -1;
    cfg->sparse_chol = 1;
    cfg->joint = 1;
    cfg->dense_limit = -1;
}

static int create_impl(const fh_term* terms, size_t nterms, const double* w_kept, size_t n, const fh_solver_config* cfg,
                       const fh_ctx* ctx, fh_absorber** out, bool dev) {
    return guard([&] {
        require(out != nullptr, "null output handle");
        *out = nullptr;
        require(nterms == 0 || terms != nullptr, "null terms");
        bool slopes = false;
        for (size_t q = 0; q < nterms; q++) {
            require(terms[q].intercept == 0 || terms[q].intercept == 1, "intercept must be 0 or 1");
            require(terms[q].nslopes >= 0 && (terms[q].nslopes == 0 || terms[q].slopes != nullptr),
                    "null slope columns");
            require(terms[q].intercept == 1 || terms[q].nslopes > 0, "absorb term without intercept or slopes");
            require(terms[q].codes != nullptr || n == 0, "null term codes");
            slopes |= terms[q].intercept != 1 || terms[q].nslopes > 0;
        }
        require(n >= 1, "no observations", FH_TOO_FEW);
        auto h = std::make_unique<fh_absorber>();
        Absorber& a = h->a;
        FH_HIP(hipGetDevice(&a.device));
        a.st = stream_acquire();
        // on a failure below the stream goes back to the pool after the absorber's buffers
        struct StreamBack {
            hipStream_t st;
            int dev;
            bool armed = true;
            ~StreamBack() {
                if (armed) stream_release(st, dev);
            }
        } back{a.st, a.device};
        auto hold = std::move(h);  // destroyed before `back`: buffers first, then the stream
        StreamScope ss(a.st);
        a.n = n;
        a.nterms = nterms;
        fh_solver_config c;
        fh_config_default(&c);
        if (cfg) c = *cfg;
        a.small_limit = c.small_limit;
        a.precision = c.precision;
        require(c.bucket_log2 >= -1 && c.bucket_log2 <= 31, "bucket_log2 must be in -1..31");
        a.bucket_log2 = c.bucket_log2;
        require(c.autotune == 0 || c.autotune == 1, "autotune must be 0 or 1");
        a.autotune = c.autotune == 1;
        require(c.mixed_lead >= -1 && c.mixed_lead <= 8, "mixed_lead must be in -1..8");
        a.mixed_lead = c.mixed_lead < 0 ? MIXED_LEAD_DEFAULT : (uint32_t)c.mixed_lead;
        require(c.sparse_chol == 0 || c.sparse_chol == 1, "sparse_chol must be 0 or 1");
        require(c.joint == 0 || c.joint == 1, "joint must be 0 or 1");
        a.sparse_chol = c.sparse_chol == 1;
        a.joint_policy = c.joint;
        require(c.dense_limit >= -1, "dense_limit must be -1 (the default) or at least 0");
        a.dense_limit = c.dense_limit < 0 ? DENSE_LIMIT_DEFAULT : (uint64_t)c.dense_limit;
        std::vector<DBuf<uint32_t>> up(nterms);
        std::vector<const uint32_t*> dc(nterms);
        for (size_t q = 0; q < nterms; q++) {
            a.term_g.push_back(terms[q].g);
            if (dev) {
                dc[q] = terms[q].codes;
            } else {
                up[q].alloc(n);
                h2d(up[q].p, terms[q].codes, n, a.st);
                dc[q] = up[q].p;
            }
        }
        // code range check (structure.rs:483-489): every code below its level count; its flag is
        // read while the codes are copied (no code indexes memory before the check has passed)
        DBuf<uint32_t> bad(1);
        AsyncRead bad_read(sizeof(uint32_t));
        {
            FH_HIP(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), a.st));
            for (size_t q0 = 0; q0 < nterms; q0 += NB_MAX + 1) {
                CodeCheck cc{};
                const size_t nq = std::min<size_t>(nterms - q0, NB_MAX + 1);
                for (size_t q = 0; q < nq; q++) {
                    cc.c[q] = dc[q0 + q];
                    cc.g[q] = terms[q0 + q].g;
                }
                hipLaunchKernelGGL(k_check_codes_all, dim3(grid_for(n / 4 + 1, TPB, 2048), (unsigned)nq), dim3(TPB), 0,
                                   a.st, cc, n, bad.p);
            }
            FH_CHECK_LAUNCH();
            bad_read.start(bad.p, a.st);
        }
        // the terms' codes stay with the absorber (the bootstrap's non-nested projection)
        a.codes_kept.resize(nterms);
        for (size_t q = 0; q < nterms; q++) {
            if (dev) {
                a.codes_kept[q].alloc(n);
                FH_HIP(hipMemcpyAsync(a.codes_kept[q].p, dc[q], n * sizeof(uint32_t), hipMemcpyDeviceToDevice, a.st));
            } else {
                a.codes_kept[q] = std::move(up[q]);
            }
            dc[q] = a.codes_kept[q].p;
        }
        require(*static_cast<const uint32_t*>(bad_read.wait()) == 0, "a term code is not below its level count");
        DBuf<double> wd;
        const double* wk = nullptr;
        if (w_kept) {
            wd.alloc(n);
            if (dev) FH_HIP(hipMemcpyAsync(wd.p, w_kept, n * sizeof(double), hipMemcpyDeviceToDevice, a.st));
            else h2d(wd.p, w_kept, n, a.st);
            require(weights_ok(a, wd.p), "weights must be positive and finite");
            wk = wd.p;
        }
        for (size_t q = 0; q < nterms; q++) {
            a.term_icpt.push_back(terms[q].intercept == 1 ? 1 : 0);
            a.term_ns.push_back((uint32_t)terms[q].nslopes);
        }
        if (slopes) {
            // the varying-slopes solver (solve.rs:338-349 chooses it the same way)
            std::vector<SlopeTerm> sts(nterms);
            std::vector<DBuf<double>> sup;
            for (size_t q = 0; q < nterms; q++) {
                sts[q].codes = dc[q];
                sts[q].g = terms[q].g;
                sts[q].icpt = terms[q].intercept == 1;
                for (int k = 0; k < terms[q].nslopes; k++) {
                    require(terms[q].slopes[k] != nullptr, "null slope column");
                    if (dev) {
                        sts[q].slopes.push_back(terms[q].slopes[k]);
                    } else {
                        sup.emplace_back(n);
                        h2d(sup.back().p, terms[q].slopes[k], n, a.st);
                        sts[q].slopes.push_back(sup.back().p);
                    }
                }
            }
            slope_create(a, sts, wk);
            stream_wait(a.st);
            back.armed = false;
            *out = hold.release();
            return;
        }
        auto t0 = std::chrono::steady_clock::now();
        build_structure(a, dc);
        log_line(ctx, 2, "fereg_hip: structure " +
                             std::to_string(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count()) + "s");
        a.ms_struct = a.ms;
        a.small_struct.clear();
        for (const BlockInfo& bi : a.blocks) a.small_struct.push_back(bi.small ? 1 : 0);
        if (a.ms > a.dense_limit) {
            // The structure keeps the reference's small/large order (the row order and cell codes
            // depend on it); the small blocks are preconditioned like large ones (Jacobi, level
            // passes) when their dense factor would cost more than the iterations it saves
            for (size_t b = 0; b < a.nb; b++)
                if (a.blocks[b].small) {
                    a.blocks[b].small = false;
                    ensure_block_transpose(a, (int)b);
                }
            a.ms = 0;
        }
        DBuf<double> wrow;
        if (wk) {
            wrow.alloc(n);
            hipLaunchKernelGGL(k_gather_vec, dim3(vgrid(n)), dim3(TPB), 0, a.st, wk, a.pi.p, n, wrow.p);
            FH_CHECK_LAUNCH();
        }
        refresh_weights(a, wk ? wrow.p : nullptr);
        back.armed = false;
        *out = hold.release();
    });
}

int fh_absorber_create(const fh_term* terms, size_t nterms, const double* w_kept, size_t n, const fh_solver_config* cfg,
                       const fh_ctx* ctx, fh_absorber** out) {
    return create_impl(terms, nterms, w_kept, n, cfg, ctx, out, false);
}
int fh_absorber_create_dev(const fh_term* terms, size_t nterms, const double* w_kept, size_t n,
                           const fh_solver_config* cfg, const fh_ctx* ctx, fh_absorber** out) {
    return create_impl(terms, nterms, w_kept, n, cfg, ctx, out, true);
}

void fh_absorber_destroy(fh_absorber* ab) {
    if (!ab) return;
    hipStream_t st = ab->a.st;
    {
        const int dev = ab->a.device;
        DeviceGuard g(dev);
        (void)hipStreamSynchronize(st);
        delete ab;
        stream_release(st, dev);
    }
}

size_t fh_absorber_n(const fh_absorber* ab) { return ab ? ab->a.n : 0; }

const uint32_t* fh_absorber_row_order(fh_absorber* ab) {
    if (!ab) return nullptr;
    Absorber& a = ab->a;
    if (!a.host_pi_valid) {
        int rc = guard([&] {
            AbsScope g(a);
            a.host_pi.resize(a.n);
            d2h(a.host_pi.data(), a.pi.p, a.n, a.st);
            stream_wait(a.st);
            a.host_pi_valid = true;
        });
        if (rc) return nullptr;
    }
    return a.host_pi.data();
}
const uint32_t* fh_absorber_row_order_dev(const fh_absorber* ab) { return ab ? ab->a.pi.p : nullptr; }

static int set_weights_impl(fh_absorber* ab, const double* w, const fh_ctx* ctx, bool dev) {
    return guard([&] {
        require(ab != nullptr, "null absorber");
        Absorber& a = ab->a;
        AbsScope g(a);
        check_interrupt(ctx);
        if (!w) {
            a.irls_weights = false;
            a.user_w.release();
            refresh_weights(a, nullptr);
            return;
        }
        DBuf<double> d(a.n);
        if (dev) FH_HIP(hipMemcpyAsync(d.p, w, a.n * sizeof(double), hipMemcpyDeviceToDevice, a.st));
        else h2d(d.p, w, a.n, a.st);
        require(weights_ok(a, d.p), "weights must be positive and finite");
        a.irls_weights = false;
        a.user_w.release();
        refresh_weights(a, d.p);
    });
}
int fh_absorber_set_weights(fh_absorber* ab, const double* w, const fh_ctx* ctx) { return set_weights_impl(ab, w, ctx, false); }
int fh_absorber_set_weights_dev(fh_absorber* ab, const double* w, const fh_ctx* ctx) { return set_weights_impl(ab, w, ctx, true); }

static int residualize_impl(fh_absorber* ab, double* z, size_t r, const fh_solve_opts* o, const double* warm,
                            double* coefs_out, fh_solve_info* info, const fh_ctx* ctx, bool dev) {
    return guard([&] {
        require(ab != nullptr && o != nullptr, "null argument");
        Absorber& a = ab->a;
        AbsScope g(a);
        restore_user_weights(a);
        auto t0 = std::chrono::steady_clock::now();
        const size_t n = a.n, p = a.m + a.g1;
        if (info) *info = fh_solve_info{1, 0, 1, 0, 0.0};
        if (r == 0) return;
        require(z != nullptr, "null block");
        check_interrupt(ctx);
        const int prec = o->precision < 0 ? a.precision : o->precision;
        ResidOut ro;
        if (dev) {
            ro = residualize_any(a, z, r, o->tol, o->maxiter, warm, coefs_out, prec, ctx);
            stream_wait(a.st);  // the caller's device block is final on return
        } else {
            DBuf<double> dz(n * r), dw, dcf;
            h2d(dz.p, z, n * r, a.st);
            if (warm) {
                dw.alloc(p * r);
                h2d(dw.p, warm, p * r, a.st);
            }
            if (coefs_out) dcf.alloc(p * r);
            ro = residualize_any(a, dz.p, r, o->tol, o->maxiter, warm ? dw.p : nullptr, coefs_out ? dcf.p : nullptr, prec, ctx);
            d2h(z, dz.p, n * r, a.st);
            if (coefs_out) d2h(coefs_out, dcf.p, p * r, a.st);
            stream_wait(a.st);
        }
        if (info) {
            info->iterations = ro.iterations;
            info->exact = ro.exact;
            info->escalated = ro.escalated ? 1 : 0;
            info->refinements = ro.refinements;
            info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    });
}
int fh_absorber_residualize(fh_absorber* ab, double* z, size_t r, const fh_solve_opts* o, const double* warm,
                            double* coefs_out, fh_solve_info* info, const fh_ctx* ctx) {
    return residualize_impl(ab, z, r, o, warm, coefs_out, info, ctx, false);
}
int fh_absorber_residualize_dev(fh_absorber* ab, double* z, size_t r, const fh_solve_opts* o, const double* warm,
                                double* coefs_out, fh_solve_info* info, const fh_ctx* ctx) {
    return residualize_impl(ab, z, r, o, warm, coefs_out, info, ctx, true);
}

size_t fh_absorber_coef_rows(const fh_absorber* ab) { return ab ? ab->a.m + ab->a.g1 : 0; }

size_t fh_absorber_term_slots(const fh_absorber* ab) {
    if (!ab) return 0;
    const Absorber& a = ab->a;
    size_t s = 0;
    for (size_t q = 0; q < a.nterms; q++)
        s += (q >= a.term_icpt.size() || a.term_icpt[q] ? 1 : 0) + (q < a.term_ns.size() ? a.term_ns[q] : 0);
    return s;
}

int fh_absorber_info(const fh_absorber* ab, fh_structure_info* o) {
    return guard([&] {
        require(ab && o, "null argument");
        const Absorber& a = ab->a;
        if (a.slope) {
            // E's levels as the segments, one cell per row, the non-E blocks
            int64_t e = 0;
            uint64_t ge = 0, be = 0, msm = 0, nb = 0;
            slope_summary(a, &e, &ge, &be, &msm, &nb);
            *o = fh_structure_info{a.n, a.nterms, ge, a.n, nb, a.m, msm, 0, e, a.norm_fe1, 0, (int32_t)(a.ms > 0)};
            return;
        }
        *o = fh_structure_info{a.n,        a.nterms,  a.g1,
                               a.M,        a.nb,      a.m,
                               a.ms_struct, a.small_work, a.fe1,
                               a.norm_fe1, (int32_t)!a.trivial(), (int32_t)(a.ms > 0)};
    });
}

int fh_absorber_copy_structure(fh_absorber* ab, uint32_t* pi, uint32_t* seg_off, uint32_t* cc, uint32_t* coo) {
    return guard([&] {
        require(ab != nullptr, "null absorber");
        Absorber& a = ab->a;
        AbsScope g(a);
        require(!a.slope, "not available for an absorber with varying slopes (no cell structure)");
        if (pi) d2h(pi, a.pi.p, a.n, a.st);
        if (seg_off) d2h(seg_off, a.seg_off.p, a.g1 + 1, a.st);
        if (cc) d2h(cc, a.cell_codes.p, a.M * a.nb, a.st);
        if (coo && !a.trivial()) d2h(coo, a.cell_obs_off.p, a.M + 1, a.st);
        stream_wait(a.st);
    });
}

int fh_absorber_copy_blocks(const fh_absorber* ab, int64_t* term, uint64_t* g, uint64_t* off, int32_t* small) {
    return guard([&] {
        require(ab != nullptr, "null absorber");
        const Absorber& a = ab->a;
        if (a.slope) {
            slope_blocks(a, term, g, off, small);
            return;
        }
        for (size_t b = 0; b < a.nb; b++) {
            term[b] = a.blocks[b].term;
            g[b] = a.blocks[b].g;
            off[b] = a.blocks[b].off;
            small[b] = a.small_struct[b];
        }
    });
}

int fh_absorber_copy_level_cells(fh_absorber* ab, size_t b, uint32_t* perm, uint32_t* off) {
    return guard([&] {
        require(ab != nullptr, "null absorber");
        Absorber& a = ab->a;
        require(!a.slope, "not available for an absorber with varying slopes (no cell structure)");
        require(b < a.nb, "block index out of range");
        AbsScope g(a);
        ensure_level_cells(a, b);
        d2h(perm, a.lperm[b].p, a.M, a.st);
        d2h(off, a.loff[b].p, a.blocks[b].g + 1, a.st);
        stream_wait(a.st);
    });
}

int fh_absorber_copy_weight_state(fh_absorber* ab, double* n1, double* nr, double* diag, uint8_t* red, double* sw) {
    return guard([&] {
        require(ab != nullptr, "null absorber");
        Absorber& a = ab->a;
        AbsScope g(a);
        require(!a.slope, "not available for an absorber with varying slopes (no cell structure)");
        restore_user_weights(a);
        if (n1) d2h(n1, a.n1.p, a.g1, a.st);
        if (nr) d2h(nr, a.nr.p, a.m, a.st);
        if (diag) d2h(diag, a.diag.p, a.m, a.st);
        if (red) d2h(red, a.red.p, a.m, a.st);
        if (sw) *sw = absorber_sw(a);
        stream_wait(a.st);
    });
}

int fh_absorber_apply(fh_absorber* ab, const double* p, size_t r, double* y, int32_t precision) {
    return guard([&] {
        require(ab != nullptr && p && y, "null argument");
        require(r >= 1 && r <= 16, "1..16 columns");
        Absorber& a = ab->a;
        AbsScope g(a);
        restore_user_weights(a);
        const size_t m = a.m;
        if (m == 0) return;
        require(!a.slope || precision != FH_MIXED, "an absorber with varying slopes applies S in FP64");
        if (precision == FH_MIXED) {
            std::vector<float> hp(p, p + m * r), hy(m * r);
            DBuf<float> dp(m * r), dy(m * r);
            h2d(dp.p, hp.data(), m * r, a.st);
            spmv<float>(a, dp.p, dy.p, (int)r, a.s32, nullptr, a.st);
            d2h(hy.data(), dy.p, m * r, a.st);
            stream_wait(a.st);
            for (size_t i = 0; i < m * r; i++) y[i] = hy[i];
        } else {
            DBuf<double> dp(m * r), dy(m * r);
            h2d(dp.p, p, m * r, a.st);
            spmv<double>(a, dp.p, dy.p, (int)r, a.s64, nullptr, a.st);
            d2h(y, dy.p, m * r, a.st);
            stream_wait(a.st);
        }
    });
}

int fh_absorber_get_plan(const fh_absorber* ab, int32_t precision, size_t r, fh_spmv_plan* out) {
    return guard([&] {
        require(ab && out && r >= 1 && r <= 16, "bad argument");
        require(precision == FH_PLAIN_FP64 || precision == FH_MIXED, "unknown precision");
        const Absorber& a = ab->a;
        require(!a.slope, "not available for an absorber with varying slopes (no cell structure)");
        const SpmvPlan& pl = a.plan_for(precision == FH_PLAIN_FP64, (int)r);
        *out = fh_spmv_plan{};
        out->dense = pl.dense;
        out->seg_lanes = pl.seg_lanes ? pl.seg_lanes
                                      : spmv_seg_lanes(a, padded_cols((int)r), precision == FH_PLAIN_FP64 ? 8 : 4);
        for (size_t b = 0; b < a.nb && b < 9; b++) {
            const LvlMode md = lvl_mode(a, pl, b, (int)r, precision == FH_PLAIN_FP64 ? 8 : 4);
            out->block_lanes[b] = (uint32_t)md.lanes;
            out->block_mode[b] = a.nbk[b] > 1 ? (md.merge ? 0 : 1) : -1;
        }
        out->tuned = pl.tuned ? 1 : 0;
    });
}

int fh_absorber_set_plan(fh_absorber* ab, int32_t precision, size_t r, const fh_spmv_plan* in) {
    return guard([&] {
        require(ab && in && r >= 1 && r <= 16, "bad argument");
        require(precision == FH_PLAIN_FP64 || precision == FH_MIXED, "unknown precision");
        const Absorber& a = ab->a;
        require(!a.slope, "not available for an absorber with varying slopes (no cell structure)");
        auto pow2 = [](uint32_t v) { return v == 0 || (v <= 64 && (v & (v - 1)) == 0); };
        require(in->dense == 0 || in->dense == 1, "plan: dense must be 0 or 1");
        require(pow2(in->seg_lanes), "plan: lanes must be 0 or a power of two up to 64");
        SpmvPlan pl;
        pl.dense = in->dense;
        pl.seg_lanes = in->seg_lanes;
        pl.blanes.assign(a.nb, 0);
        pl.xcd.assign(a.nb, -1);
        for (size_t b = 0; b < a.nb && b < 9; b++) {
            require(pow2(in->block_lanes[b]), "plan: lanes must be 0 or a power of two up to 64");
            require(in->block_mode[b] >= -1 && in->block_mode[b] <= 1, "plan: block mode must be -1, 0 or 1");
            pl.blanes[b] = in->block_lanes[b];
            pl.xcd[b] = in->block_mode[b];
        }
        pl.tuned = true;
        a.plan_for(precision == FH_PLAIN_FP64, (int)r) = pl;
    });
}

void fh_tune_cache_clear(void) { tune_cache_clear(); }

int fh_absorber_time_spmv(fh_absorber* ab, size_t r, int32_t precision, uint32_t reps, double* ms, double* bytes) {
    return guard([&] {
        require(ab && ms && bytes && r >= 1 && r <= 16 && reps >= 1, "bad argument");
        Absorber& a = ab->a;
        require(!a.slope, "not available for an absorber with varying slopes (no cell structure)");
        AbsScope g(a);
        restore_user_weights(a);
        const size_t m = a.m;
        const double s = precision == FH_MIXED ? 4.0 : 8.0;
        *bytes = 4.0 * a.nb * a.M + 12.0 * a.g1 + 2.0 * s * m * r + (a.unit() ? 0.0 : s * a.M);
        *ms = 0.0;
        if (m == 0) return;
        hipEvent_t e0, e1;
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        // the solvers run S*p on columns padded to the kernels' class (zero pads): time that width
        const size_t rq = padded_cols(r);
        auto run = [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            DBuf<T> dp(m * rq), dy(m * rq);
            std::vector<T> hp(m * rq, T(0));
            for (size_t i = 0; i < m; i++)
                for (size_t j = 0; j < r; j++) hp[i * rq + j] = T(0.001 * (double)((i * r + j) % 997));
            h2d(dp.p, hp.data(), m * rq, a.st);
            Scratch<T>& sc = precision == FH_MIXED ? reinterpret_cast<Scratch<T>&>(a.s32) : reinterpret_cast<Scratch<T>&>(a.s64);
            spmv<T>(a, dp.p, dy.p, (int)rq, sc, nullptr, a.st);  // warm-up (scratch allocation)
            FH_HIP(hipEventRecord(e0, a.st));
            for (uint32_t k = 0; k < reps; k++) spmv<T>(a, dp.p, dy.p, (int)rq, sc, nullptr, a.st);
            FH_HIP(hipEventRecord(e1, a.st));
            FH_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            FH_HIP(hipEventElapsedTime(&t, e0, e1));
            *ms = (double)t / reps;
        };
        if (precision == FH_MIXED) run((float*)nullptr);
        else run((double*)nullptr);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

// fit_rows (solve.rs:1313-1317; operator.rs fit_rows): alpha1[s(i)] + sum_b x[off_b + code_b(c(i))].
int fh_absorber_fit_rows(fh_absorber* ab, const double* coefs, size_t r, size_t col, double* out) {
    return guard([&] {
        require(ab && coefs && out && col < r, "bad argument");
        Absorber& a = ab->a;
        AbsScope g(a);
        if (a.slope) {
            const size_t p = a.m + a.g1;
            DBuf<double> dc(p * r), dout(a.n);
            h2d(dc.p, coefs, p * r, a.st);
            slope_fit_rows(a, dc.p, (int)r, (int)col, dout.p);
            d2h(out, dout.p, a.n, a.st);
            stream_wait(a.st);
            return;
        }
        require(a.nb <= (size_t)NB_MAX, "fit_rows: too many absorbed terms");
        const size_t p = a.m + a.g1;
        DBuf<double> dc(p * r), dout(a.n);
        h2d(dc.p, coefs, p * r, a.st);
        BlockOffs bo{};
        for (size_t b = 0; b < a.nb; b++) bo.off[b] = (uint32_t)a.blocks[b].off;
        hipLaunchKernelGGL(k_fit_rows, dim3(grid_for(a.M, TPB, 8192)), dim3(TPB), 0, a.st, dc.p, (int)r, (int)col,
                           a.cell_codes.p, a.cell_seg.p, a.trivial() ? nullptr : a.cell_obs_off.p, (size_t)a.M,
                           (int)a.nb, bo, (size_t)a.m, dout.p);
        FH_CHECK_LAUNCH();
        d2h(out, dout.p, a.n, a.st);
        stream_wait(a.st);
    });
}

// term_values (solve.rs:1319-1343): per term levels, normalized so every term except FE1
// (fe1_index) has weighted mean 0 and FE1 absorbs the shifts.
int fh_absorber_term_values(fh_absorber* ab, const double* coefs, size_t r, size_t col, double* out, uint64_t* lens) {
    return guard([&] {
        require(ab && coefs && out && lens && col < r, "bad argument");
        Absorber& a = ab->a;
        AbsScope g(a);
        restore_user_weights(a);
        const size_t nt = a.nterms;
        if (a.slope) {
            size_t pos = 0, k = 0;
            for (auto& v : slope_term_values(a, coefs, r, col)) {
                lens[k++] = v.size();
                std::memcpy(out + pos, v.data(), v.size() * sizeof(double));
                pos += v.size();
            }
            return;
        }
        if (a.fe1 < 0 || a.norm_fe1 < 0) {
            for (size_t q = 0; q < nt; q++) lens[q] = 0;
            return;
        }
        std::vector<double> n1(a.g1), nr(a.m);
        d2h(n1.data(), a.n1.p, a.g1, a.st);
        if (a.m) d2h(nr.data(), a.nr.p, a.m, a.st);
        stream_wait(a.st);
        std::vector<std::vector<double>> vals(nt), lw(nt);
        vals[a.fe1].resize(a.g1);
        lw[a.fe1] = n1;
        for (size_t s = 0; s < a.g1; s++) vals[a.fe1][s] = coefs[(a.m + s) * r + col];
        for (auto& b : a.blocks) {
            vals[b.term].resize(b.g);
            lw[b.term].assign(nr.begin() + b.off, nr.begin() + b.off + b.g);
            for (size_t l = 0; l < b.g; l++) vals[b.term][l] = coefs[(b.off + l) * r + col];
        }
        double shift = 0.0;
        for (size_t t = 0; t < nt; t++) {
            if ((int64_t)t == a.norm_fe1) continue;
            double mean = 0.0;
            for (size_t l = 0; l < vals[t].size(); l++) mean += vals[t][l] * lw[t][l];
            mean /= absorber_sw(a);
            for (double& v : vals[t]) v -= mean;
            shift += mean;
        }
        for (double& v : vals[a.norm_fe1]) v += shift;
        size_t pos = 0;
        for (size_t t = 0; t < nt; t++) {
            lens[t] = vals[t].size();
            std::memcpy(out + pos, vals[t].data(), vals[t].size() * sizeof(double));
            pos += vals[t].size();
        }
    });
}

// Device codes of the cluster dimensions (host codes are uploaded into hold) and their set.
struct ClusterArgs {
    std::vector<DBuf<uint32_t>> up;
    std::vector<const uint32_t*> ptr;
};
static ClusterSet clusters_in(Absorber& a, const uint32_t* const* dims, size_t ndim, int g_df, bool dev,
                              ClusterArgs& hold) {
    require(ndim == 0 || dims != nullptr, "null cluster dimensions");
    hold.ptr.assign(ndim, nullptr);
    for (size_t d = 0; d < ndim; d++) {
        require(dims[d] != nullptr, "null cluster codes");
        if (dev) {
            hold.ptr[d] = dims[d];
        } else {
            hold.up.emplace_back(a.n);
            h2d(hold.up.back().p, dims[d], a.n, a.st);
            hold.ptr[d] = hold.up.back().p;
        }
    }
    return cluster_set(a, hold.ptr.data(), ndim, g_df);
}

static int ols_impl(fh_absorber* ab, const double* y, const double* X, size_t K, int32_t kind,
                    const uint32_t* const* dims, size_t ndim, int32_t g_df, const int32_t* nested, const fh_ssc* ssc,
                    const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats, bool dev) {
    return guard([&] {
        require(ab && y && X && ssc && o && beta && vcov && kept, "null argument");
        Absorber& a = ab->a;
        AbsScope g(a);
        restore_user_weights(a);
        const size_t n = a.n;
        ClusterArgs hold;
        const ClusterSet cs = clusters_in(a, dims, ndim, g_df, dev, hold);
        if (dev) {
            ols_fit(a, y, X, K, kind, cs, nested, *ssc, *o, beta, vcov, kept, stats, nullptr);
        } else {
            DBuf<double> dy(n), dx(n * K);
            h2d(dy.p, y, n, a.st);
            h2d(dx.p, X, n * K, a.st);
            ols_fit(a, dy.p, dx.p, K, kind, cs, nested, *ssc, *o, beta, vcov, kept, stats, nullptr);
        }
    });
}
static int wildboot_impl(fh_absorber* ab, const double* y, const double* X, size_t K, const uint32_t* cluster,
                         const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o, const fh_wboot_spec* spec,
                         double* beta, double* vcov, int32_t* kept, fh_wboot_result* res, double* t_boot, bool dev) {
    return guard([&] {
        require(ab && y && X && cluster && ssc && o && spec && spec->r_mat && beta && vcov && kept && res,
                "null argument");
        Absorber& a = ab->a;
        AbsScope g(a);
        require(!a.slope, "the wild bootstrap with varying slopes is not supported");
        restore_user_weights(a);
        const size_t n = a.n;
        DBuf<double> dy, dx;
        DBuf<uint32_t> dc;
        const double *py = y, *px = X;
        const uint32_t* pc = cluster;
        if (!dev) {
            dy.alloc(n);
            dx.alloc(n * K);
            dc.alloc(n);
            h2d(dy.p, y, n, a.st);
            h2d(dx.p, X, n * K, a.st);
            h2d(dc.p, cluster, n, a.st);
            py = dy.p;
            px = dx.p;
            pc = dc.p;
        }
        // the fit's CRV1 scores and the bootstrap's cluster sums share one set of cluster runs
        const uint64_t G = cluster_levels(a, pc);
        ClusterRuns runs;
        cluster_runs(a, pc, G, runs);
        struct RunsScope {
            Absorber& a;
            ~RunsScope() { a.cl_runs = nullptr; }
        } runs_scope{a};
        a.cl_runs = &runs;
        OlsState os;
        WbPre pre;
        pre.rmat.assign(spec->r_mat, spec->r_mat + K);
        os.wb = &pre;
        fh_fit_stats stats{};
        ClusterSet cs;
        cs.dims = {pc};
        cs.G = {G};
        ols_fit(a, py, px, K, FH_VCOV_CLUSTER, cs, nested, *ssc, *o, beta, vcov, kept, &stats, nullptr, &os);
        // R over the kept columns; a restriction on a dropped column is not testable
        std::vector<double> rk;
        for (size_t j = 0; j < K; j++) {
            if (kept[j]) rk.push_back(spec->r_mat[j]);
            else require(spec->r_mat[j] == 0.0, "the hypothesis involves a regressor dropped as collinear");
        }
        // fixed effects not nested in the clusters (wildboot/mod.rs:959-997) are projected out of
        // the draws exactly. At most min(G, 512) levels (Path 1): as appended pseudo-regressors
        // when they fit beside the kept regressors, else through Path 1's algebra with the draws as
        // FP64 products (m_nn / 16 solves). More levels with G <= 512 (Path 2): one cluster at a
        // time (G / 16 solves). More than 512 of both: skipped, as the reference does.
        std::vector<int> nest = terms_nested_in(a, pc);
        uint64_t m_nn = 0;
        for (size_t q = 0; q < a.nterms; q++)
            if (!nest[q]) m_nn += a.term_g[q];
        const bool ref_path1 = m_nn > 0 && m_nn <= std::min<uint64_t>(G, 512);
        const bool ref_path2 = m_nn > 0 && !ref_path1 && G <= 512;
        const bool path1 = ref_path1 && os.kept.size() + m_nn + 2 <= 21;
        const bool cells = ref_path2;
        const bool wide = ref_path1 && !path1;
        const int bt = spec->boot_type == 0 ? 11 : spec->boot_type;
        require(bt == 11 || bt == 13 || bt == 31 || bt == 33, "wildboot: boot_type must be 11, 13, 31 or 33");
        WbootInput in;
        in.rows = a.rows_tmp.p;
        in.R = os.ld;
        in.kept = os.kept;
        in.bread = os.inv;
        in.beta = os.coef;
        in.e = os.ecol >= 0 ? a.rows_tmp.p + os.ecol : os.e.p;
        in.es = os.ecol >= 0 ? (size_t)os.ld : 1;
        in.w = a.weighted ? a.w_rows.p : nullptr;
        in.cl = pc;
        in.G = G;
        const double nn = (double)n, gg = (double)os.ncl;
        in.ssc = (ssc->k_adj ? (nn - 1.0) / (nn - os.dfk) : 1.0) * (ssc->g_adj && gg > 1.0 ? gg / (gg - 1.0) : 1.0);
        in.S = pre.ready ? pre.S.p : nullptr;
        if (bt != 11) {
            // MNW: the non-nested terms' orthonormal basis joins the regressors when it fits; else
            // their leave-cluster-out refitting is left out, as the reference does with a warning
            DBuf<double> Qb;
            int r_nn = 0;
            const bool use_q = path1 && in.kept.size() + m_nn + 2 <= 21;
            if (use_q) mnw_basis(a, nest, m_nn, in, Qb, r_nn);
            in.S = nullptr;
            ols_wildboot_mnw(a, in, *spec, bt, rk.data(), Qb.p, r_nn, *res, t_boot);
            res->fe_not_projected = m_nn > 0 && !use_q;
            return;
        }
        if (cells) {
            ols_wildboot_cells(a, in, *spec, rk.data(), o->tol, o->maxiter, *res, t_boot);
            res->fe_not_projected = false;
            return;
        }
        if (wide) {
            ols_wildboot_wide(a, nest, m_nn, in, *spec, rk.data(), *res, t_boot);
            res->fe_not_projected = false;
            return;
        }
        WbExt ext;
        if (path1) wb_path1(a, nest, m_nn, in, rk, ext);
        ols_wildboot(a, in, *spec, rk.data(), *res, t_boot);
        res->fe_not_projected = m_nn > 0 && !path1;
    });
}
int fh_ols_wildboot(fh_absorber* ab, const double* y, const double* X, size_t K, const uint32_t* cluster,
                    const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o, const fh_wboot_spec* spec,
                    double* beta, double* vcov, int32_t* kept, fh_wboot_result* res, double* t_boot) {
    return wildboot_impl(ab, y, X, K, cluster, nested, ssc, o, spec, beta, vcov, kept, res, t_boot, false);
}
int fh_ols_wildboot_dev(fh_absorber* ab, const double* y, const double* X, size_t K, const uint32_t* cluster,
                        const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o, const fh_wboot_spec* spec,
                        double* beta, double* vcov, int32_t* kept, fh_wboot_result* res, double* t_boot) {
    return wildboot_impl(ab, y, X, K, cluster, nested, ssc, o, spec, beta, vcov, kept, res, t_boot, true);
}

int fh_ols_fit(fh_absorber* ab, const double* y, const double* X, size_t K, int32_t kind, const uint32_t* cluster,
               const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o, double* beta, double* vcov,
               int32_t* kept, fh_fit_stats* stats) {
    return ols_impl(ab, y, X, K, kind, &cluster, cluster ? 1 : 0, FH_GDF_MIN, nested, ssc, o, beta, vcov, kept, stats,
                    false);
}
int fh_ols_fit_dev(fh_absorber* ab, const double* y, const double* X, size_t K, int32_t kind, const uint32_t* cluster,
                   const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o, double* beta, double* vcov,
                   int32_t* kept, fh_fit_stats* stats) {
    return ols_impl(ab, y, X, K, kind, &cluster, cluster ? 1 : 0, FH_GDF_MIN, nested, ssc, o, beta, vcov, kept, stats,
                    true);
}
int fh_ols_fit_multi(fh_absorber* ab, const double* y, const double* X, size_t K, int32_t kind, const fh_clusters* cl,
                     const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o, double* beta, double* vcov,
                     int32_t* kept, fh_fit_stats* stats) {
    return ols_impl(ab, y, X, K, kind, cl ? cl->codes : nullptr, cl ? cl->ndim : 0, cl ? cl->g_df : FH_GDF_MIN, nested,
                    ssc, o, beta, vcov, kept, stats, false);
}
int fh_ols_fit_multi_dev(fh_absorber* ab, const double* y, const double* X, size_t K, int32_t kind,
                         const fh_clusters* cl, const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o,
                         double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats) {
    return ols_impl(ab, y, X, K, kind, cl ? cl->codes : nullptr, cl ? cl->ndim : 0, cl ? cl->g_df : FH_GDF_MIN, nested,
                    ssc, o, beta, vcov, kept, stats, true);
}

static int ppml_impl(fh_absorber* ab, const double* y, const double* X, size_t K, const double* offset, int32_t kind,
                     const uint32_t* const* dims, size_t ndim, int32_t g_df, const int32_t* nested, const fh_ssc* ssc,
                     double tol, uint64_t maxiter, const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept,
                     fh_fit_stats* stats, const fh_ctx* ctx, bool dev) {
    return guard([&] {
        require(ab && y && X && ssc && o && beta && vcov && kept, "null argument");
        Absorber& a = ab->a;
        AbsScope g(a);
        const size_t n = a.n;
        ClusterArgs hold;
        const ClusterSet cs = clusters_in(a, dims, ndim, g_df, dev, hold);
        if (dev) {
            ppml_fit(a, y, X, K, offset, kind, cs, nested, *ssc, tol, maxiter, *o, beta, vcov, kept, stats, ctx);
        } else {
            DBuf<double> dy(n), dx(n * K), doff;
            h2d(dy.p, y, n, a.st);
            h2d(dx.p, X, n * K, a.st);
            if (offset) {
                doff.alloc(n);
                h2d(doff.p, offset, n, a.st);
            }
            ppml_fit(a, dy.p, dx.p, K, offset ? doff.p : nullptr, kind, cs, nested, *ssc, tol, maxiter, *o, beta, vcov,
                     kept, stats, ctx);
        }
    });
}
int fh_ppml_fit(fh_absorber* ab, const double* y, const double* X, size_t K, const double* offset, int32_t kind,
                const uint32_t* cluster, const int32_t* nested, const fh_ssc* ssc, double tol, uint64_t maxiter,
                const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats,
                const fh_ctx* ctx) {
    return ppml_impl(ab, y, X, K, offset, kind, &cluster, cluster ? 1 : 0, FH_GDF_MIN, nested, ssc, tol, maxiter, o,
                     beta, vcov, kept, stats, ctx, false);
}
int fh_ppml_fit_dev(fh_absorber* ab, const double* y, const double* X, size_t K, const double* offset, int32_t kind,
                    const uint32_t* cluster, const int32_t* nested, const fh_ssc* ssc, double tol, uint64_t maxiter,
                    const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats,
                    const fh_ctx* ctx) {
    return ppml_impl(ab, y, X, K, offset, kind, &cluster, cluster ? 1 : 0, FH_GDF_MIN, nested, ssc, tol, maxiter, o,
                     beta, vcov, kept, stats, ctx, true);
}
int fh_ppml_fit_multi(fh_absorber* ab, const double* y, const double* X, size_t K, const double* offset, int32_t kind,
                      const fh_clusters* cl, const int32_t* nested, const fh_ssc* ssc, double tol, uint64_t maxiter,
                      const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats,
                      const fh_ctx* ctx) {
    return ppml_impl(ab, y, X, K, offset, kind, cl ? cl->codes : nullptr, cl ? cl->ndim : 0,
                     cl ? cl->g_df : FH_GDF_MIN, nested, ssc, tol, maxiter, o, beta, vcov, kept, stats, ctx, false);
}
int fh_ppml_fit_multi_dev(fh_absorber* ab, const double* y, const double* X, size_t K, const double* offset,
                          int32_t kind, const fh_clusters* cl, const int32_t* nested, const fh_ssc* ssc, double tol,
                          uint64_t maxiter, const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept,
                          fh_fit_stats* stats, const fh_ctx* ctx) {
    return ppml_impl(ab, y, X, K, offset, kind, cl ? cl->codes : nullptr, cl ? cl->ndim : 0,
                     cl ? cl->g_df : FH_GDF_MIN, nested, ssc, tol, maxiter, o, beta, vcov, kept, stats, ctx, true);
}

int fh_device_alloc(void** p, size_t bytes) {
    return guard([&] { FH_HIP(hipMalloc(p, bytes)); });
}
int fh_device_free(void* p) {
    return guard([&] { FH_HIP(hipFree(p)); });
}
int fh_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    return guard([&] { FH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int fh_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    return guard([&] { FH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int fh_device_synchronize(void) {
    return guard([&] { FH_HIP(hipDeviceSynchronize()); });
}
int fh_release_cached_memory(void) {
    return guard([&] { pool_release_cached(); });
}
int fh_memory_stats(size_t* in_use, size_t* cached) {
    return guard([&] { pool_stats(in_use, cached); });
}

}  // extern "C"
