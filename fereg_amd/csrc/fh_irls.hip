// Poisson pseudo-maximum likelihood by IRLS with absorbed fixed effects (glm/irls.rs:367-913),
// in fixest's flavour (glm/mod.rs:798-835, fit/glm.rs):
//   - initial values mu = y + 0.1, the deviance anchor at eta = 1e-5 (offset left out) and the
//     first step halved toward it at most twice (irls.rs:432-447, 693-710);
//   - per iteration: weights w = tw mu, working variable z = (eta - off) - 1 + y/mu, the block
//     [z, X] residualized at the inner tolerance (warm-started), tiny entries edited to 0;
//   - the collinearity of X decided once, at the first iteration (irls.rs:505-545);
//   - beta by normal equations while far from convergence, by a QR factor after that
//     (irls.rs:459, 548-569);
//   - fixest's convergence test and the inexact-Newton inner tolerance eps^(3/2)
//     (irls.rs:824-827, 893-901).
// Every per-row pass runs on the device; the host keeps the K x K algebra and the scalars.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "fh_fit.hpp"

namespace fh {

namespace {

constexpr int TPB = 256;
constexpr int GRID = 1024;  // fixed grid of the partial sums: a fixed summation order
constexpr int KC = 24;      // coefficient slots of a step (K <= 21 regressors + the target)
constexpr double MU_MIN = 100.0 * DBL_EPSILON;  // family.rs:13
constexpr double ETA_ANCHOR = 1e-5;             // irls.rs:432-447
constexpr double START_INNER_TOL = 1e-4;        // irls.rs:26
constexpr double COLLIN_TOL = 1e-9;             // fit/mod.rs:325
constexpr double COLLIN_MARGIN = 100.0;         // irls.rs:35

inline unsigned vgrid(size_t n) { return grid_for(n, TPB, 8192); }

struct Coefs {
    double c[KC];
};

__global__ void k_gather_vec(const double* __restrict__ v, const uint32_t* __restrict__ pi, size_t n,
                             double* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = v[pi[i]];
}

__global__ void k_gather_x(const double* __restrict__ X, size_t K, const uint32_t* __restrict__ pi, size_t n,
                           double* __restrict__ out) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * K; e += (size_t)gridDim.x * blockDim.x)
        out[e] = X[(size_t)pi[e / K] * K + e % K];
}

// flag bit 0: some y is negative or not finite; bit 1: some x lies in (0, ztol] in magnitude
// (the fill then edits X too, irls.rs:425-430). lny = ln y (0 where y = 0, irls.rs:225-243).
__global__ void k_irls_check(const double* __restrict__ y, const double* __restrict__ x, size_t n, size_t K,
                             double ztol, uint32_t* __restrict__ flag, double* __restrict__ lny) {
    bool bad = false, tiny = false;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = y[i];
        bad |= !(v >= 0.0) || !isfinite(v);
        lny[i] = v > 0.0 ? log(v) : 0.0;
    }
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * K; e += stride) {
        const double v = x[e];
        tiny |= v != 0.0 && fabs(v) <= ztol;
    }
    if (__ballot(bad) && lane_id() == 0) atomicOr(flag, 1u);
    if (__ballot(tiny) && lane_id() == 0) atomicOr(flag, 2u);
}

__device__ __forceinline__ double pois_dev(double mu, double eta, double y, double w, double lny) {
    return w * (mu - y) + (y > 0.0 ? w * y * (lny - eta) : 0.0);  // irls.rs:268-270
}

__device__ __forceinline__ double edit0(double v, double ztol) { return fabs(v) <= ztol ? 0.0 : v; }

// The IRLS rows of the next solve from eta and mu (irls.rs:467-493): w = tw mu, the working
// variable z = (eta - off) - 1 + y/mu (family.rs working) kept raw in zraw, and the block row
// [z, x...] with entries of magnitude <= ztol edited to 0 (x only when some x is that small).
__device__ __forceinline__ void fill_row(size_t i, double et, double m, double yi, double oi, double twi,
                                         const double* __restrict__ xr, size_t K, double ztol, bool xtiny,
                                         double* __restrict__ rows, double* __restrict__ zraw,
                                         double* __restrict__ w) {
    const size_t R = K + 1;
    const double z = (et - oi) - 1.0 + yi / m;
    zraw[i] = z;
    w[i] = twi * m;
    rows[i * R] = edit0(z, ztol);
    for (size_t k = 0; k < K; k++) {
        const double v = xr[i * K + k];
        rows[i * R + 1 + k] = xtiny ? edit0(v, ztol) : v;
    }
}

__device__ void block_sum(double acc, double* __restrict__ part) {
    __shared__ double red[TPB / WAVE];
    acc = wave_sum(acc);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < TPB / WAVE; q++) t += red[q];
        part[blockIdx.x] = t;
    }
}

// fixest's initial values mu = y + 0.1, eta = ln mu (the offset is inside eta), the first
// solve's rows, and the partial deviance of the anchor eta = 1e-5 without the offset.
__global__ __launch_bounds__(TPB) void k_irls_init(const double* __restrict__ y, const double* __restrict__ lny,
                                                   const double* __restrict__ tw, const double* __restrict__ off,
                                                   const double* __restrict__ xr, size_t n, size_t K, double ztol,
                                                   int xtiny, double* __restrict__ mu, double* __restrict__ eta,
                                                   double* __restrict__ rows, double* __restrict__ zraw,
                                                   double* __restrict__ w, double* __restrict__ part) {
    const double mu0 = exp(ETA_ANCHOR);
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double yi = y[i], twi = tw ? tw[i] : 1.0;
        const double m = yi + 0.1, et = log(m);
        mu[i] = m;
        eta[i] = et;
        fill_row(i, et, m, yi, off ? off[i] : 0.0, twi, xr, K, ztol, xtiny != 0, rows, zraw, w);
        acc += pois_dev(mu0, ETA_ANCHOR, yi, twi, lny[i]);
    }
    block_sum(acc, part);
}

// One IRLS step (irls.rs:575-650): e = z~ - x~'beta on the residualized rows (coefficients
// c = [1, -beta]), eta_old = the anchor 1e-5 + off at the first iteration, else eta;
// eta = zraw - e + off; mu = max(exp eta, MU_MIN); the deviance partials; and the next
// iteration's rows, weights and working variable into spare buffers.
__global__ __launch_bounds__(TPB) void k_irls_step(const double* __restrict__ rows, size_t K, Coefs cf,
                                                   const double* __restrict__ zraw, const double* __restrict__ off,
                                                   const double* __restrict__ y, const double* __restrict__ lny,
                                                   const double* __restrict__ tw, const double* __restrict__ xr,
                                                   size_t n, double ztol, int xtiny, int anchor,
                                                   double* __restrict__ eta, double* __restrict__ eta_old,
                                                   double* __restrict__ mu, double* __restrict__ e,
                                                   double* __restrict__ part, double* __restrict__ rows_n,
                                                   double* __restrict__ zraw_n, double* __restrict__ w_n) {
    const size_t R = K + 1;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double* row = rows + i * R;
        double ei = 0.0;
        for (size_t k = 0; k < R; k++) ei += cf.c[k] * row[k];
        e[i] = ei;
        const double oi = off ? off[i] : 0.0;
        eta_old[i] = anchor ? ETA_ANCHOR + oi : eta[i];
        const double et = zraw[i] - ei + oi;
        double m = exp(et);
        if (!(m >= MU_MIN)) m = MU_MIN;
        eta[i] = et;
        mu[i] = m;
        const double yi = y[i], twi = tw ? tw[i] : 1.0;
        acc += pois_dev(m, et, yi, twi, lny[i]);
        fill_row(i, et, m, yi, oi, twi, xr, K, ztol, xtiny != 0, rows_n, zraw_n, w_n);
    }
    block_sum(acc, part);
}

// fixest step halving (irls.rs:693-710): eta = (eta + eta_old) / 2, mu, the deviance partials and
// the next rows again.
__global__ __launch_bounds__(TPB) void k_irls_halve(double* __restrict__ eta, const double* __restrict__ eta_old,
                                                    const double* __restrict__ off, const double* __restrict__ y,
                                                    const double* __restrict__ lny, const double* __restrict__ tw,
                                                    const double* __restrict__ xr, size_t n, size_t K, double ztol,
                                                    int xtiny, double* __restrict__ mu, double* __restrict__ part,
                                                    double* __restrict__ rows_n, double* __restrict__ zraw_n,
                                                    double* __restrict__ w_n) {
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double et = 0.5 * (eta[i] + eta_old[i]);
        double m = exp(et);
        if (!(m >= MU_MIN)) m = MU_MIN;
        eta[i] = et;
        mu[i] = m;
        const double yi = y[i], twi = tw ? tw[i] : 1.0;
        acc += pois_dev(m, et, yi, twi, lny[i]);
        fill_row(i, et, m, yi, off ? off[i] : 0.0, twi, xr, K, ztol, xtiny != 0, rows_n, zraw_n, w_n);
    }
    block_sum(acc, part);
}

// irls.rs:212-222: a raw deviance below rounding is 0; deviance = 2 x the sum.
double finish_deviance(double raw, size_t n) {
    double d = raw;
    if (2.0 * d / (double)n < DBL_EPSILON) d = 0.0;
    if (std::fabs(d) <= DBL_EPSILON) d = 0.0;
    return 2.0 * d;
}

// ppmlhdfe predict_eps (irls.rs:105-141): ln(eps) history of three; once full, the cubic
// through (i, ln eps_i) without a constant is extrapolated to i = 4. NaN until then.
double predict_eps(double (&hist)[3], double eps) {
    int pos = 0;
    for (double v : hist) pos += std::isnan(v) ? 0 : 1;
    const bool abort = pos + 1 < 3;
    const double le = eps > 0.0 ? std::log(eps) : NAN;
    if (pos < 3) {
        hist[pos] = le;
    } else {
        hist[0] = hist[1];
        hist[1] = hist[2];
        hist[2] = le;
    }
    if (abort || std::isnan(hist[0]) || std::isnan(hist[1]) || std::isnan(hist[2])) return NAN;
    // [i, i^2, i^3] c = h at i = 1, 2, 3 by Gaussian elimination without pivoting
    double A[3][4] = {{1, 1, 1, hist[0]}, {2, 4, 8, hist[1]}, {3, 9, 27, hist[2]}};
    for (int p = 0; p < 3; p++)
        for (int row = p + 1; row < 3; row++) {
            const double f = A[row][p] / A[p][p];
            for (int j = p; j < 4; j++) A[row][j] -= f * A[p][j];
        }
    double c[3];
    for (int p = 2; p >= 0; p--) {
        double s = A[p][3];
        for (int j = p + 1; j < 3; j++) s -= A[p][j] * c[j];
        c[p] = s / A[p][p];
    }
    return std::exp(4.0 * c[0] + 16.0 * c[1] + 64.0 * c[2]);
}

// Coefficients of column 0 on the kept columns of the weighted Gram G (R x R, [z, X]) by the
// normal equations with the sweep inverse (wls.rs:160-169, normal_coefs).
std::vector<double> normal_coefs(const std::vector<double>& G, int R, const std::vector<int>& kept) {
    const int mk = (int)kept.size();
    std::vector<double> gxx((size_t)mk * mk);
    for (int a = 0; a < mk; a++)
        for (int b = 0; b < mk; b++) gxx[(size_t)a * mk + b] = G[(size_t)(1 + kept[a]) * R + 1 + kept[b]];
    const std::vector<double> inv = invsym(gxx, mk, INVSYM_TOL);
    std::vector<double> b(mk, 0.0);
    for (int a = 0; a < mk; a++) {
        double s = 0.0;
        for (int c = 0; c < mk; c++) s += inv[(size_t)a * mk + c] * G[(size_t)(1 + kept[c]) * R];
        b[a] = s;
    }
    return b;
}

}  // namespace

void ppml_fit(Absorber& a, const double* y, const double* X, size_t K, const double* offset, int kind,
              const ClusterSet& cs, const int32_t* nested, const fh_ssc& ssc, double tol, uint64_t maxiter,
              const fh_solve_opts& o, double* beta, double* vcov, int32_t* kept_out, fh_fit_stats* stats,
              const fh_ctx* ctx) {
    hipStream_t st = a.st;
    const size_t n = a.n, R = K + 1;
    require(K >= 1 && K <= 21, "1..21 regressors are supported");
    require(maxiter >= 1, "the IRLS needs at least one iteration");
    require((kind != FH_VCOV_CLUSTER && kind != FH_VCOV_CRV3) || !cs.dims.empty(),
            "a clustered variance matrix needs cluster codes");
    // The caller's weights are the prior weights; the absorber holds the IRLS weights from here on
    // and gets the caller's back before its next non-Poisson call.
    if (!a.irls_weights) {
        if (a.weighted) a.user_w = std::move(a.w_rows);
        else a.user_w.release();
        a.irls_weights = true;
    }
    const double* tw = a.user_w.n ? a.user_w.p : nullptr;
    const double ztol = std::min(tol, 1e-12);  // irls.rs:418
    // row-order copies of y, X and the offset
    DBuf<double> yr(n), xr(n * K), offr, lny(n), mu(n), eta(n), eta_old(n), zraw(n), w(n), e(n), part(GRID);
    hipLaunchKernelGGL(k_gather_vec, dim3(vgrid(n)), dim3(TPB), 0, st, y, a.pi.p, n, yr.p);
    hipLaunchKernelGGL(k_gather_x, dim3(vgrid(n * K)), dim3(TPB), 0, st, X, K, a.pi.p, n, xr.p);
    if (offset) {
        offr.alloc(n);
        hipLaunchKernelGGL(k_gather_vec, dim3(vgrid(n)), dim3(TPB), 0, st, offset, a.pi.p, n, offr.p);
    }
    FH_CHECK_LAUNCH();
    const double* off = offset ? offr.p : nullptr;
    uint32_t flags = 0;
    {
        DBuf<uint32_t> flag(1);
        FH_HIP(hipMemsetAsync(flag.p, 0, sizeof(uint32_t), st));
        hipLaunchKernelGGL(k_irls_check, dim3(GRID), dim3(TPB), 0, st, yr.p, xr.p, n, K, ztol, flag.p, lny.p);
        FH_CHECK_LAUNCH();
        d2h(&flags, flag.p, 1, st);
        stream_wait(st);
    }
    require(!(flags & 1u), "Poisson outcomes must be non-negative and finite");
    const int xtiny = (flags & 2u) ? 1 : 0;
    a.rows_tmp.reserve(n * R);
    double* rows = a.rows_tmp.p;
    DBuf<double> rows_n(n * R), zraw_n(n), w_n(n);
    hipLaunchKernelGGL(k_irls_init, dim3(GRID), dim3(TPB), 0, st, yr.p, lny.p, tw, off, xr.p, n, K, ztol, xtiny, mu.p,
                       eta.p, rows, zraw.p, w.p, part.p);
    FH_CHECK_LAUNCH();
    const double dev0 = finish_deviance(sum_parts(part.p, GRID, st), n);
    const size_t p = a.m + a.g1;
    DBuf<double> warm(p * R), coefs(p * R);
    bool have_warm = false, have_next = false;
    const int prec = o.precision < 0 ? a.precision : o.precision;
    // inner tolerances (glm/mod.rs:826-831): start at max(1e-4, tol), never below highest
    const double highest = std::max(1e-12, std::min(o.tol, 0.1 * tol));
    double tol_in = std::max(START_INNER_TOL, tol);
    double hist[3] = {NAN, NAN, NAN};
    double eps = NAN, dev = NAN, dev_prev = NAN;
    std::vector<int> excluded(K, 0), kept;
    std::vector<double> G(R * R), coef;
    uint64_t it = 0, pcg_total = 0;
    bool converged = false;
    for (it = 1; it <= maxiter; it++) {
        check_interrupt(ctx);
        const double predicted = predict_eps(hist, eps);
        // ppmlhdfe iter_fast_solver (irls.rs:459): normal equations while far from convergence
        const bool fast = tol_in > tol * 11.0 && (std::isnan(predicted) || predicted > tol);
        if (have_next) {
            std::swap(a.rows_tmp, rows_n);
            rows = a.rows_tmp.p;
            std::swap(zraw, zraw_n);
            std::swap(w, w_n);
        }
        // the weights move into the absorber; the buffer they replace becomes the next spare
        a.w_rows.reserve(n);
        std::swap(a.w_rows, w);
        refresh_weights(a, a.w_rows.p);
        ResidOut ro = residualize_any(a, rows, R, tol_in, o.maxiter, have_warm ? warm.p : nullptr, coefs.p, prec, ctx);
        pcg_total += ro.iterations;
        std::swap(warm, coefs);  // this solve's coefficients start the next one
        have_warm = true;
        // the weighted Gram of [z~, X~]; entries of magnitude <= ztol are edited to 0 on the way
        gram_cols(a, rows, n, (int)R, 0, (int)R, a.w_rows.p, nullptr, G.data(), ztol);
        if (it == 1) {
            // fixest's collinearity rule on the weighted Gram of X~ (irls.rs:505-545): the first
            // solve decides unless its error could change the outcome; then the block is
            // residualized again at the final tolerance and decides
            std::vector<double> raw(K * K);
            for (size_t i = 0; i < K; i++)
                for (size_t j = 0; j < K; j++) raw[i * K + j] = G[(i + 1) * R + j + 1];
            bool certain = false;
            if (ro.exact) {
                std::vector<double> inv;
                chol_collin(raw, (int)K, COLLIN_TOL, excluded, inv);
                certain = true;
            } else {
                std::vector<double> err(K);
                for (size_t j = 0; j < K; j++) err[j] = COLLIN_MARGIN * tol_in * std::sqrt(std::max(raw[j * K + j], 0.0));
                certain = chol_collin_certain(raw, (int)K, COLLIN_TOL, err, excluded);
            }
            if (!certain) {
                const ResidOut r2 =
                    residualize_any(a, rows, R, std::min(tol_in, highest), o.maxiter, nullptr, nullptr, prec, ctx);
                pcg_total += r2.iterations;
                gram_cols(a, rows, n, (int)R, 0, (int)R, a.w_rows.p, nullptr, G.data(), ztol);
                for (size_t i = 0; i < K; i++)
                    for (size_t j = 0; j < K; j++) raw[i * K + j] = G[(i + 1) * R + j + 1];
                std::vector<double> inv;
                chol_collin(raw, (int)K, COLLIN_TOL, excluded, inv);
            }
            kept.clear();
            for (size_t j = 0; j < K; j++)
                if (!excluded[j]) kept.push_back((int)j);
            require(!kept.empty(), "all the regressors are collinear (with the fixed effects)");
        }
        const int mk = (int)kept.size();
        bool have_coef = false;
        if (!fast) {
            // QR of sqrt(w) [X~_kept, z~] (wls.rs:103-176)
            const int m = mk + 1;
            std::vector<double> T0(R * (size_t)m, 0.0), G0((size_t)m * m);
            for (int j = 0; j < mk; j++) T0[(size_t)(1 + kept[j]) * m + j] = 1.0;
            T0[(size_t)m - 1] = 1.0;  // z (row 0) last
            std::vector<int> col(m);
            for (int j = 0; j < mk; j++) col[j] = 1 + kept[j];
            col[mk] = 0;
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++) G0[(size_t)i * m + j] = G[(size_t)col[i] * R + col[j]];
            std::vector<double> Rf;
            if (qr_factor(a, rows, n, (int)R, (int)R, T0, m, a.w_rows.p, G0, Rf)) {
                coef = qr_solve_upper(Rf, m);
                have_coef = true;
            }
        }
        if (!have_coef) coef = normal_coefs(G, (int)R, kept);
        Coefs cf{};
        cf.c[0] = 1.0;
        for (int i = 0; i < mk; i++) cf.c[1 + kept[i]] = -coef[i];
        hipLaunchKernelGGL(k_irls_step, dim3(GRID), dim3(TPB), 0, st, rows, K, cf, zraw.p, off, yr.p, lny.p, tw, xr.p,
                           n, ztol, xtiny, it == 1 ? 1 : 0, eta.p, eta_old.p, mu.p, e.p, part.p, rows_n.p, zraw_n.p,
                           w_n.p);
        FH_CHECK_LAUNCH();
        have_next = true;
        dev = finish_deviance(sum_parts(part.p, GRID, st), n);
        const double old_dev = it == 1 ? dev0 : dev_prev;
        // fixest step halving: unless the change is already below tol, a non-finite or higher
        // deviance halves the step toward the previous eta until it is finite and not higher;
        // at the first iteration at most twice when finite
        bool halved = false;
        {
            auto evol = [&](double d) { return d - old_dev; };
            const bool small = std::isfinite(dev) && (std::fabs(evol(dev)) < tol ||
                                                      std::fabs(evol(dev)) / (0.1 + std::fabs(dev)) < tol);
            if (!small && (!std::isfinite(dev) || evol(dev) > 0.0)) {
                uint64_t h = 0;
                while ((!std::isfinite(dev) || evol(dev) > 0.0) && !(it == 1 && std::isfinite(dev) && h >= 2) &&
                       h < maxiter) {
                    hipLaunchKernelGGL(k_irls_halve, dim3(GRID), dim3(TPB), 0, st, eta.p, eta_old.p, off, yr.p, lny.p,
                                       tw, xr.p, n, K, ztol, xtiny, mu.p, part.p, rows_n.p, zraw_n.p, w_n.p);
                    FH_CHECK_LAUNCH();
                    dev = finish_deviance(sum_parts(part.p, GRID, st), n);
                    h++;
                }
                halved = true;
            }
        }
        double delta = old_dev - dev;
        if (!std::isnan(delta) && dev < 0.1 * delta) delta = dev;
        if (it > 1) eps = std::fabs(delta) / std::max(std::min(dev, old_dev), 0.1);
        const double crit = halved ? INFINITY : std::fabs(old_dev - dev) / (0.1 + std::fabs(dev));
        log_line(ctx, 2, "fereg_hip: irls iteration " + std::to_string(it) + " deviance " + std::to_string(dev) +
                             " inner tol " + std::to_string(tol_in) + (fast ? "" : " (qr)"));
        if (crit < tol && (ro.exact || tol_in <= 1.1 * highest)) {
            converged = true;
            break;
        }
        if (it > 1) tol_in = std::max(std::min(eps * std::sqrt(eps), tol_in), highest);
        dev_prev = dev;
    }
    if (it > maxiter) it = maxiter;  // fixest returns the last iterate, flagged as not converged
    // variance at the last IRLS weights (fit/glm.rs:521-525): bread = (X~'W X~)^-1 over the kept
    // columns, scores x~ e w
    const int mk = (int)kept.size();
    std::vector<double> sub((size_t)mk * mk), inv;
    for (int i = 0; i < mk; i++)
        for (int j = 0; j < mk; j++) sub[(size_t)i * mk + j] = G[(size_t)(1 + kept[i]) * R + 1 + kept[j]];
    std::vector<int> singular;
    chol_collin(sub, mk, DBL_MIN, singular, inv);
    for (int s : singular) require(!s, "the Poisson information matrix is singular", FH_INTERNAL);
    const double np = (double)mk + (a.nterms > 0 ? fe_params(a) : 1.0);  // the constant without absorbed terms
    const double dfk = df_k(a, kind, ssc, nested, np, mk);
    uint64_t ncl = 0;
    std::vector<double> V;
    if (kind == FH_VCOV_CRV3) {
        // each refit: the IRLS again with the left-out cluster's prior weights at zero
        // (crv3.rs: GLMs always refit); the prior weights come back afterwards
        DBuf<double> prior;
        const bool had_tw = a.user_w.n > 0;
        if (had_tw) {
            prior.alloc(n);
            FH_HIP(hipMemcpyAsync(prior.p, a.user_w.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        CrvRefit refit = [&](const double* wm, uint64_t g) {
            DBuf<double> saved = std::move(a.user_w);
            a.user_w.alloc(n);
            FH_HIP(hipMemcpyAsync(a.user_w.p, wm, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            std::vector<double> b(K), Vt(K * K);
            std::vector<int32_t> kp(K);
            ppml_fit(a, y, X, K, offset, FH_VCOV_IID, ClusterSet{}, nested, ssc, tol, maxiter, o, b.data(), Vt.data(),
                     kp.data(), nullptr, ctx);
            a.user_w = std::move(saved);
            std::vector<double> bk(mk);
            for (int i = 0; i < mk; i++) {
                require(kp[kept[i]], "CRV3: coefficient " + std::to_string(kept[i]) +
                                         " is not estimable when cluster " + std::to_string(g) + " is left out");
                bk[i] = b[kept[i]];
            }
            return bk;
        };
        std::vector<double> Vk;
        crv3_vcov(a, cs, ssc, dfk, had_tw ? prior.p : nullptr, coef, refit, Vk, &ncl);
        V.assign(K * K, 0.0);
        for (int i = 0; i < mk; i++)
            for (int j = 0; j < mk; j++) V[(size_t)kept[i] * K + kept[j]] = Vk[(size_t)i * mk + j];
    } else {
        vcov_from(a, rows, (int)R, (int)R, excluded, inv, e.p, -1, a.w_rows.p, kind, cs, ssc, dfk, 1.0, V, &ncl,
                  nullptr, true);
    }
    for (size_t j = 0; j < K; j++) {
        beta[j] = 0.0;
        kept_out[j] = !excluded[j];
    }
    for (int i = 0; i < mk; i++) beta[kept[i]] = coef[i];
    std::memcpy(vcov, V.data(), K * K * sizeof(double));
    if (stats) {
        stats->ssr = 0.0;
        stats->df_k = dfk;
        stats->nparams = np;
        stats->deviance = dev;
        stats->iterations = it;
        stats->pcg_iterations = pcg_total;
        stats->n_clusters = ncl;
        stats->converged = converged ? 1 : 0;
    }
}

}  // namespace fh
