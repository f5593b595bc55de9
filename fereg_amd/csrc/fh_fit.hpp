// Host-side pieces shared by the estimation entry points: the residualization driver, the
// variance assembly, the small dense algebra of the K x K systems, and the transfer helpers.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

// ------------------------------------------------------------------ transfers and control

template <class T>
inline void h2d(T* dst, const T* src, size_t count, hipStream_t st) {
    if (count) upload(dst, src, count * sizeof(T), st);
}
template <class T>
inline void d2h(T* dst, const T* src, size_t count, hipStream_t st) {
    if (count) download(dst, src, count * sizeof(T), st);
}
// Runs f and maps an exception to the reference's error code (fh_last_error holds the message).
int guard(const std::function<void()>& f);
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        FH_HIP(hipGetDevice(&prev));
        if (prev != dev) FH_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

// The absorber's device is current and pool buffers are ordered on its stream.
struct AbsScope {
    DeviceGuard dg;
    StreamScope ss;
    explicit AbsScope(const Absorber& a) : dg(a.device), ss(a.st) {}
};

// Sum of `count` device partials, read back (waits for the stream).
double sum_parts(const double* dev, size_t count, hipStream_t st);
// Throws FH_INTERRUPTED when the caller's interrupt callback asks for it.
void check_interrupt(const fh_ctx* ctx);
void log_line(const fh_ctx* ctx, int level, const std::string& s);

// ------------------------------------------------------------------ small dense algebra (fh_linalg.hip)

struct CholPivots {
    std::vector<double> R, pivot;  // k x k upper factor over the kept columns, every column's pivot
    std::vector<int> excluded;
};
CholPivots chol_pivots(const std::vector<double>& A, int k, double tol);
// Column-order Cholesky excluding pivots below tol; inv = inverse of the kept block.
void chol_collin(const std::vector<double>& A, int k, double tol, std::vector<int>& excluded, std::vector<double>& inv);
// The exclusions of chol_collin, when columns carrying errors up to err could not change them.
bool chol_collin_certain(const std::vector<double>& A, int k, double tol, const std::vector<double>& err,
                         std::vector<int>& excluded);
std::vector<double> upper_inverse(const std::vector<double>& R, int k);
// Sweep generalized inverse (Mata invsym in column order), relative pivot tolerance.
std::vector<double> invsym(const std::vector<double>& A, int k, double tol, std::vector<int>* dropped = nullptr);
constexpr double INVSYM_TOL = 1e-13;
// Stata invsym(A) with the pivot order of the largest remaining diagonal.
std::vector<double> invsym_greedy(const std::vector<double>& A, int k, double tol, std::vector<int>* dropped = nullptr);
// Symmetric eigendecomposition by cyclic Jacobi rotations: U's columns are the eigenvectors.
void sym_eigen(const std::vector<double>& A, int k, std::vector<double>& lambda, std::vector<double>& U);
// fixest's PSD repair (eigenvalues clipped at 1e-16); true when an entry moved by more than 1e-8.
bool posdef_fix(std::vector<double>& V, int k);
// reghdfe's PSD repair (Cameron-Gelbach-Miller); true when applied.
bool fix_psd(std::vector<double>& V, int k, bool report_constant);
bool chol_upper(const std::vector<double>& A, int k, std::vector<double>& R);
std::vector<double> upper_mul(const std::vector<double>& A, const std::vector<double>& B, int k);
// b solving R_xx b = R_xy for the triangular factor of [X, y] (y last).
std::vector<double> qr_solve_upper(const std::vector<double>& R, int m);

// ------------------------------------------------------------------ device estimation passes (fh_estimate.hip)

// G = sum_i w_i (z_i T)'(z_i T) over the first R columns of the rows z (stride ld), T R x m
// (row-major, host): the Gram of transformed rows.
void gram_transformed(Absorber& a, const double* z, size_t n, int ld, int R, const std::vector<double>& T, int m,
                      const double* w, double* G_host);
// e = z[:, :k] c (n), also written to column ecol of z when ecol >= 0; returns sum w e^2.
double residual_pass(Absorber& a, double* z, size_t n, int ld, int k, const double* c_host, const double* w, double* e,
                     int ecol);
void hetero_meat(Absorber& a, const double* z, size_t n, int ld, int c0, int k, const double* e, const double* w,
                 double* meat);
// e_i at e[i * es]; returns the non-empty clusters.
uint64_t cluster_meat(Absorber& a, const double* z, size_t n, int ld, int c0, int k, const double* e, size_t es,
                      const double* w, const uint32_t* cl_kept_dev, uint64_t G, double* meat);

// R factor (m x m, upper) of the QR decomposition of sqrt(w) [z_i T0] by shifted Cholesky QR
// with two refinement passes; T0 (R x m) selects and orders the columns. False when a
// refinement pass meets a non-positive pivot (numerically rank-deficient columns).
bool qr_factor(Absorber& a, const double* z, size_t n, int ld, int R, const std::vector<double>& T0, int m,
               const double* w, const std::vector<double>& G0, std::vector<double>& Rf);

// ------------------------------------------------------------------ residualization (fh_api.hip)

struct ResidOut {
    uint64_t iterations = 1;
    bool exact = true;
    uint64_t refinements = 0;
    bool escalated = false;
};
// Residualizes the n x r device block z (row order) in place; warm / coefs: (m + g1) x r.
ResidOut residualize_any(Absorber& a, double* z, size_t r, double tol, uint64_t maxiter, const double* warm_dev,
                         double* coefs_dev, int precision, const fh_ctx* ctx);
// Rows [y, X] of kept-order device inputs in solver row order (pi), padded to ld columns.
void gather_block(const double* y, const double* X, size_t K, const uint32_t* pi, size_t n, size_t ld, double* rows,
                  hipStream_t st);
// Puts the caller's own weights back when a Poisson fit left its IRLS weights in the absorber.
void restore_user_weights(Absorber& a);

// ------------------------------------------------------------------ variance (fh_api.hip)

// fixest df.K (fit/vcov.rs:63-112).
double df_k(const Absorber& a, int kind, const fh_ssc& ssc, const int32_t* nested, double np, int kdesign);
// Number of fixed-effect parameters: sum of the term levels minus (terms - 1).
double fe_params(const Absorber& a);
struct WbPre;
// Cluster dimensions of a variance request: kept-order device codes with their level counts
// (max code + 1), and fixest's ssc(G.df) rule for several dimensions.
struct ClusterSet {
    std::vector<const uint32_t*> dims;
    std::vector<uint64_t> G;
    int g_df = FH_GDF_MIN;
};
uint64_t cluster_levels(Absorber& a, const uint32_t* cl_dev);
ClusterSet cluster_set(Absorber& a, const uint32_t* const* dims, size_t ndim, int g_df);
// Multiway meat (vcov.rs:480-558): sum over the non-empty subsets S of the dimensions of
// (-1)^(|S|+1) scale(G_S) meat_S, meat_S clustered by the intersection of S; scale(G) = G/(G-1)
// when `conventional`, else 1. Scores z[i][c0..c0+k) e_i w_i. Returns the smallest G of a single
// dimension; G_S counts non-empty clusters.
uint64_t multiway_meat(Absorber& a, const double* z, int ld, int c0, int k, const double* e, size_t es,
                       const double* w, const ClusterSet& cs, bool conventional, double* meat);
void vcov_from(Absorber& a, double* rows, int ld, int R, const std::vector<int>& excluded,
               const std::vector<double>& bread, const double* e, int ecol, const double* w, int kind,
               const ClusterSet& cs, const fh_ssc& ssc, double dfk, double iid_scale,
               std::vector<double>& V, uint64_t* ncl, WbPre* wb = nullptr, bool glm = false, size_t nrows = 0);

// ------------------------------------------------------------------ packed requests (fh_api.hip)

// One request of a packed solve: an unweighted OLS with iid or hetero variance (host inputs in
// kept order), its outputs and return code.
struct PackedOls {
    size_t n = 0, K = 0;
    std::vector<const uint32_t*> codes;
    std::vector<uint64_t> g;
    const double* y = nullptr;
    const double* X = nullptr;
    int kind = FH_VCOV_IID;
    fh_ssc ssc{};
    double tol = 1e-8;
    uint64_t maxiter = 16000;
    double* beta = nullptr;
    double* vcov = nullptr;
    int32_t* kept = nullptr;
    int rc = FH_OK;
    uint64_t iterations = 0;
};
// Several independent requests in one absorber: their fixed effects take disjoint level ranges
// (a request with fewer terms gets one constant level in the missing ones, absorbed by its own
// first term), so the packed projection is block diagonal and each request's residuals are its
// own; every column is scaled to unit mean square per request for the solve (the tolerance is
// relative per request) and unscaled before the fits. One structure build, one multi-RHS PCG.
void ols_fit_packed(const std::vector<PackedOls*>& reqs, int precision);

// ------------------------------------------------------------------ CRV3 (fh_crv3.hip)

// A refit with row weights wm (solver order, zero on the rows of the left-out cluster g): the
// coefficients over the full fit's kept regressors.
using CrvRefit = std::function<std::vector<double>(const double* wm_row, uint64_t g)>;
// V = factor sum_S sign scale(G_S) jackknife meat_S over the subsets of the cluster dimensions
// (crv3.rs, fit/vcov.rs:369-397), factor = G_min/(G_min - 1) (G.df min) x (n - 1)/(n - df_k), then
// fixest's PSD repair; beta over the kept regressors, w_row the fit's own row weights (or null).
void crv3_vcov(Absorber& a, const ClusterSet& cs, const fh_ssc& ssc, double dfk, const double* w_row,
               const std::vector<double>& beta, const CrvRefit& refit, std::vector<double>& V, uint64_t* ncl);

// ------------------------------------------------------------------ factors (fh_factor.hip)

// Dense codes of the interaction of device code arrays (n rows) into out; returns the levels.
uint64_t combine_dev(const std::vector<const uint32_t*>& parts, const std::vector<uint64_t>& g, size_t n,
                     uint32_t* out, hipStream_t st);

// ------------------------------------------------------------------ Poisson PML (fh_irls.hip)

void ppml_fit(Absorber& a, const double* y, const double* X, size_t K, const double* offset, int kind,
              const ClusterSet& cs, const int32_t* nested, const fh_ssc& ssc, double tol, uint64_t maxiter,
              const fh_solve_opts& o, double* beta, double* vcov, int32_t* kept_out, fh_fit_stats* stats,
              const fh_ctx* ctx);

}  // namespace fh

// The handle behind the C ABI's opaque fh_absorber.
struct fh_absorber {
    fh::Absorber a;
};
