// Host linear algebra on the small K x K systems of an estimation (K <= 22, row-major, FP64):
// the collinearity-aware Cholesky of fixest, the Goodnight sweep inverse of Stata, plain
// Cholesky and triangular helpers for the QR factor, and the triangular least-squares solve.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "fh_fit.hpp"

namespace fh {

CholPivots chol_pivots(const std::vector<double>& A, int k, double tol) {
    // vcov.rs:163-188: column order; column j is excluded when its pivot
    // A_jj - sum_{kept i<j} R_ij^2 is below tol (absolute).
    CholPivots c;
    c.R.assign((size_t)k * k, 0.0);
    c.excluded.assign(k, 0);
    c.pivot.assign(k, 0.0);
    for (int j = 0; j < k; j++) {
        double rjj = A[(size_t)j * k + j];
        for (int i = 0; i < j; i++)
            if (!c.excluded[i]) rjj -= c.R[(size_t)i * k + j] * c.R[(size_t)i * k + j];
        c.pivot[j] = rjj;
        if (!(rjj >= tol)) {
            c.excluded[j] = 1;
            continue;
        }
        const double d = std::sqrt(rjj);
        c.R[(size_t)j * k + j] = d;
        for (int col = j + 1; col < k; col++) {
            double v = A[(size_t)col * k + j];
            for (int i = 0; i < j; i++)
                if (!c.excluded[i]) v -= c.R[(size_t)i * k + col] * c.R[(size_t)i * k + j];
            c.R[(size_t)j * k + col] = v / d;
        }
    }
    return c;
}

void chol_collin(const std::vector<double>& A, int k, double tol, std::vector<int>& excluded, std::vector<double>& inv) {
    // vcov.rs:124-128: the exclusions and (R'R)^-1 over the kept columns (kept order), from the
    // inverse of the kept block of R.
    CholPivots c = chol_pivots(A, k, tol);
    excluded = c.excluded;
    std::vector<int> kept;
    for (int j = 0; j < k; j++)
        if (!excluded[j]) kept.push_back(j);
    const int mk = (int)kept.size();
    std::vector<double> Rk((size_t)mk * mk, 0.0);
    for (int a = 0; a < mk; a++)
        for (int b = a; b < mk; b++) Rk[(size_t)a * mk + b] = c.R[(size_t)kept[a] * k + kept[b]];
    const std::vector<double> Ri = upper_inverse(Rk, mk);
    inv.assign((size_t)mk * mk, 0.0);
    for (int i = 0; i < mk; i++)
        for (int j = 0; j < mk; j++) {
            double s = 0.0;
            for (int l = std::max(i, j); l < mk; l++) s += Ri[(size_t)i * mk + l] * Ri[(size_t)j * mk + l];
            inv[(size_t)i * mk + j] = s;
        }
}

bool chol_collin_certain(const std::vector<double>& A, int k, double tol, const std::vector<double>& err_in,
                         std::vector<int>& excluded) {
    // vcov.rs:141-159: the exclusions of chol_collin hold whatever errors of size err_j the
    // columns carry. The pivot of column j is the squared distance of column j from the span of
    // the kept earlier columns; a column error moves that distance by at most err_j plus the kept
    // columns' errors weighted by |b|, the coefficients of j on them (b = R_kk^-1 R_kj). The
    // decision stands when the distance stays on its side of sqrt(tol) by that margin. The
    // pivot rounding of the Gram's sums enters as 2 sqrt(1e-11 A_jj).
    CholPivots c = chol_pivots(A, k, tol);
    excluded = c.excluded;
    std::vector<double> err(k), b(k, 0.0);
    for (int j = 0; j < k; j++) err[j] = err_in[j] + 2.0 * std::sqrt(1e-11 * A[(size_t)j * k + j]);
    for (int j = 0; j < k; j++) {
        std::vector<int> kept;
        for (int i = 0; i < j; i++)
            if (!excluded[i]) kept.push_back(i);
        for (int q = (int)kept.size() - 1; q >= 0; q--) {
            const int i = kept[q];
            double s = 0.0;
            for (size_t l = q + 1; l < kept.size(); l++) s += c.R[(size_t)i * k + kept[l]] * b[kept[l]];
            b[i] = (c.R[(size_t)i * k + j] - s) / c.R[(size_t)i * k + i];
        }
        double margin = err[j];
        for (int i : kept) margin += std::fabs(b[i]) * err[i];
        const double dist = std::sqrt(std::max(c.pivot[j], 0.0));
        const bool stays = excluded[j] ? dist + margin < std::sqrt(tol) : dist - margin >= std::sqrt(tol);
        if (!stays) return false;
    }
    return true;
}

std::vector<double> upper_inverse(const std::vector<double>& R, int k) {
    // inverse of an upper-triangular matrix by column back-substitution
    std::vector<double> Ri((size_t)k * k, 0.0);
    for (int c = 0; c < k; c++) {
        Ri[(size_t)c * k + c] = 1.0 / R[(size_t)c * k + c];
        for (int row = c - 1; row >= 0; row--) {
            double s = 0.0;
            for (int l = row + 1; l <= c; l++) s += R[(size_t)row * k + l] * Ri[(size_t)l * k + c];
            Ri[(size_t)row * k + c] = -s / R[(size_t)row * k + row];
        }
    }
    return Ri;
}

std::vector<double> invsym(const std::vector<double>& A, int k, double tol, std::vector<int>* dropped_out) {
    // Stata/Mata invsym(A, 1..k) (vcov.rs:22-65): Goodnight sweeps in column order. A pivot
    // <= tol times its original diagonal (or <= 0, NaN) drops the column: its row and column of
    // the result are zero. Sweeping every pivot of a positive definite A leaves +A^-1.
    std::vector<double> m(A);
    std::vector<int> dropped(k, 0);
    for (int p = 0; p < k; p++) {
        const double d = m[(size_t)p * k + p];
        if (std::isnan(d) || d <= tol * A[(size_t)p * k + p] || d <= 0.0) {
            dropped[p] = 1;
            for (int j = 0; j < k; j++) m[(size_t)p * k + j] = m[(size_t)j * k + p] = 0.0;
            continue;
        }
        double* row = &m[(size_t)p * k];
        for (int j = 0; j < k; j++) row[j] /= d;
        for (int i = 0; i < k; i++) {
            if (i == p) continue;
            const double f = m[(size_t)i * k + p];
            if (f == 0.0) continue;
            for (int j = 0; j < k; j++) m[(size_t)i * k + j] -= f * row[j];
            m[(size_t)i * k + p] = -f / d;
        }
        m[(size_t)p * k + p] = 1.0 / d;
    }
    for (int i = 0; i < k; i++)
        for (int j = 0; j < i; j++) {
            const double v = 0.5 * (m[(size_t)i * k + j] + m[(size_t)j * k + i]);
            m[(size_t)i * k + j] = m[(size_t)j * k + i] = v;
        }
    if (dropped_out) *dropped_out = dropped;
    return m;
}

bool chol_upper(const std::vector<double>& A, int k, std::vector<double>& R) {
    // A = R'R with R upper triangular; false when a pivot is not positive
    R.assign((size_t)k * k, 0.0);
    for (int j = 0; j < k; j++) {
        double d = A[(size_t)j * k + j];
        for (int i = 0; i < j; i++) d -= R[(size_t)i * k + j] * R[(size_t)i * k + j];
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        d = std::sqrt(d);
        R[(size_t)j * k + j] = d;
        for (int c = j + 1; c < k; c++) {
            double v = A[(size_t)c * k + j];
            for (int i = 0; i < j; i++) v -= R[(size_t)i * k + c] * R[(size_t)i * k + j];
            R[(size_t)j * k + c] = v / d;
        }
    }
    return true;
}

std::vector<double> upper_mul(const std::vector<double>& A, const std::vector<double>& B, int k) {
    std::vector<double> C((size_t)k * k, 0.0);
    for (int i = 0; i < k; i++)
        for (int j = i; j < k; j++) {
            double s = 0.0;
            for (int l = i; l <= j; l++) s += A[(size_t)i * k + l] * B[(size_t)l * k + j];
            C[(size_t)i * k + j] = s;
        }
    return C;
}

std::vector<double> qr_solve_upper(const std::vector<double>& R, int m) {
    // wls.rs:138-158 (Mata _solveupper): R is the triangular factor of [X, y] (m = k + 1
    // columns, y last); solves R_xx b = R_xy by back substitution, with diagonal entries
    // <= 1e-13 max|R_ii| taken as zero (coefficient 0).
    const int k = m - 1;
    double dmax = 0.0;
    for (int j = 0; j < k; j++) dmax = std::max(dmax, std::fabs(R[(size_t)j * m + j]));
    std::vector<double> b(k, 0.0);
    for (int j = k - 1; j >= 0; j--) {
        const double d = R[(size_t)j * m + j];
        if (std::fabs(d) <= 1e-13 * dmax || d == 0.0) continue;
        double s = R[(size_t)j * m + k];
        for (int l = j + 1; l < k; l++) s -= R[(size_t)j * m + l] * b[l];
        b[j] = s / d;
    }
    return b;
}

std::vector<double> invsym_greedy(const std::vector<double>& A, int k, double tol, std::vector<int>* dropped_out) {
    // Stata invsym(A) without an order (vcov.rs:70-116): the next pivot is the unvisited one with
    // the largest remaining diagonal (lowest index on ties); a pivot <= tol times its original
    // diagonal (or <= 0, NaN) is dropped with its row and column.
    std::vector<double> m(A);
    std::vector<int> visited(k, 0), dropped(k, 0);
    for (int it = 0; it < k; it++) {
        int p = -1;
        for (int i = 0; i < k; i++)
            if (!visited[i] && (p < 0 || m[(size_t)i * k + i] > m[(size_t)p * k + p])) p = i;
        visited[p] = 1;
        const double d = m[(size_t)p * k + p];
        if (std::isnan(d) || d <= tol * A[(size_t)p * k + p] || d <= 0.0) {
            dropped[p] = 1;
            for (int j = 0; j < k; j++) m[(size_t)p * k + j] = m[(size_t)j * k + p] = 0.0;
            continue;
        }
        double* row = &m[(size_t)p * k];
        for (int j = 0; j < k; j++) row[j] /= d;
        for (int i = 0; i < k; i++) {
            if (i == p) continue;
            const double f = m[(size_t)i * k + p];
            if (f == 0.0) continue;
            for (int j = 0; j < k; j++) m[(size_t)i * k + j] -= f * row[j];
            m[(size_t)i * k + p] = -f / d;
        }
        m[(size_t)p * k + p] = 1.0 / d;
    }
    for (int i = 0; i < k; i++)
        for (int j = 0; j < i; j++) {
            const double v = 0.5 * (m[(size_t)i * k + j] + m[(size_t)j * k + i]);
            m[(size_t)i * k + j] = m[(size_t)j * k + i] = v;
        }
    if (dropped_out) *dropped_out = dropped;
    return m;
}

void sym_eigen(const std::vector<double>& A, int k, std::vector<double>& lambda, std::vector<double>& U) {
    // Cyclic Jacobi rotations (Golub and Van Loan, Matrix Computations, 4th ed., Alg. 8.5.3):
    // each rotation zeroes one off-diagonal pair with the smaller of the two angles that do; the
    // sweeps end when the off-diagonal mass is below (1e-2 eps)^2 of the total. U holds the
    // eigenvectors as columns (row-major), lambda the eigenvalues (unsorted).
    std::vector<double> m(A);
    U.assign((size_t)k * k, 0.0);
    for (int i = 0; i < k; i++) U[(size_t)i * k + i] = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < k; i++)
            for (int j = 0; j < k; j++) {
                const double v = m[(size_t)i * k + j] * m[(size_t)i * k + j];
                tot += v;
                if (i != j) off += v;
            }
        if (off == 0.0 || off <= (1e-2 * eps) * (1e-2 * eps) * tot) break;
        for (int p = 0; p < k; p++)
            for (int q = p + 1; q < k; q++) {
                const double apq = m[(size_t)p * k + q];
                if (apq == 0.0) continue;
                const double tau = (m[(size_t)q * k + q] - m[(size_t)p * k + p]) / (2.0 * apq);
                const double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau))
                                            : -1.0 / (-tau + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int r = 0; r < k; r++) {  // columns p, q of m
                    const double xp = m[(size_t)r * k + p], xq = m[(size_t)r * k + q];
                    m[(size_t)r * k + p] = c * xp - s * xq;
                    m[(size_t)r * k + q] = s * xp + c * xq;
                }
                for (int r = 0; r < k; r++) {  // rows p, q of m
                    const double xp = m[(size_t)p * k + r], xq = m[(size_t)q * k + r];
                    m[(size_t)p * k + r] = c * xp - s * xq;
                    m[(size_t)q * k + r] = s * xp + c * xq;
                }
                for (int r = 0; r < k; r++) {
                    const double up = U[(size_t)r * k + p], uq = U[(size_t)r * k + q];
                    U[(size_t)r * k + p] = c * up - s * uq;
                    U[(size_t)r * k + q] = s * up + c * uq;
                }
            }
    }
    lambda.resize(k);
    for (int i = 0; i < k; i++) lambda[i] = m[(size_t)i * k + i];
}

bool posdef_fix(std::vector<double>& V, int k) {
    // fixest mat_posdef_fix (fit/vcov.rs:514-534): when an eigenvalue is <= 0, V is rebuilt with
    // the eigenvalues clipped at 1e-16; true when that moved an entry by more than 1e-8.
    if (k == 0) return false;
    std::vector<double> lam, U;
    sym_eigen(V, k, lam, U);
    bool any = false;
    for (double l : lam) any |= !(l > 0.0);
    if (!any) return false;
    std::vector<double> f((size_t)k * k, 0.0);
    for (int e = 0; e < k; e++) {
        const double l = std::max(lam[e], 1e-16);
        for (int i = 0; i < k; i++) {
            const double ui = U[(size_t)i * k + e] * l;
            for (int j = 0; j < k; j++) f[(size_t)i * k + j] += ui * U[(size_t)j * k + e];
        }
    }
    bool changed = false;
    for (size_t i = 0; i < f.size(); i++) changed |= std::fabs(V[i] - f[i]) > 1e-8;
    V = f;
    return changed;
}

namespace {
// U diag(max(lambda, 0)) U' from the lower triangle, mirrored; false when no eigenvalue is negative.
bool clip_psd(const std::vector<double>& V, int k, std::vector<double>& out) {
    std::vector<double> lam, U;
    sym_eigen(V, k, lam, U);
    bool neg = false;
    for (double l : lam) neg |= l < 0.0;
    if (!neg) return false;
    out.assign((size_t)k * k, 0.0);
    for (int e = 0; e < k; e++) {
        if (lam[e] <= 0.0) continue;
        for (int i = 0; i < k; i++) {
            const double ui = U[(size_t)i * k + e] * lam[e];
            for (int j = 0; j <= i; j++) out[(size_t)i * k + j] += ui * U[(size_t)j * k + e];
        }
    }
    for (int i = 0; i < k; i++)
        for (int j = 0; j < i; j++) out[(size_t)j * k + i] = out[(size_t)i * k + j];
    return true;
}
}  // namespace

bool fix_psd(std::vector<double>& V, int k, bool report_constant) {
    // reghdfe_fix_psd (Cameron, Gelbach and Miller 2011; vcov.rs:689-712): symmetrize; with a
    // negative eigenvalue, replace V by its clipped form, and with a reported constant (last
    // row and column) put back the separately clipped regressor block.
    for (int i = 0; i < k; i++)
        for (int j = 0; j < i; j++) {
            const double s = 0.5 * (V[(size_t)i * k + j] + V[(size_t)j * k + i]);
            V[(size_t)i * k + j] = V[(size_t)j * k + i] = s;
        }
    const int kb = report_constant ? k - 1 : k;
    std::vector<double> backup((size_t)kb * kb);
    for (int i = 0; i < kb; i++)
        for (int j = 0; j < kb; j++) backup[(size_t)i * kb + j] = V[(size_t)i * k + j];
    std::vector<double> fixed;
    if (!clip_psd(V, k, fixed)) return false;
    V = fixed;
    if (report_constant) {
        std::vector<double> sub;
        if (!clip_psd(backup, kb, sub)) sub = backup;
        for (int i = 0; i < kb; i++)
            for (int j = 0; j < kb; j++) V[(size_t)i * k + j] = sub[(size_t)i * kb + j];
    }
    return true;
}

}  // namespace fh
