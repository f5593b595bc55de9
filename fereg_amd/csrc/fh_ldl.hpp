// Exact sparse LDL^T of a reduced-system matrix (a large block S_bb, or the joint S over every
// coordinate), for the escalated preconditioner of the FE solver (precond/sparse_chol.rs,
// solve.rs:685-891). The factor is built on the host, as the reference builds its own: the matrix
// is scaled to unit diagonal, ordered by an approximate minimum-degree ordering (a chosen set of
// coordinates, the small blocks' dense-ish rows, last), factored by the up-looking LDL^T with
// dynamic regularization (a pivot <= PIVOT_TOL grounds its coordinate with GROUND, as the dense
// small factor does: precond/dense.rs:26-40), and applied on the GPU by level-scheduled
// triangular solves (fh_ldl.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "fh_common.hpp"

namespace fh {

constexpr double LDL_PIVOT_TOL = 1e-10;  // precond/dense.rs:27
constexpr double LDL_GROUND = 1e30;      // precond/dense.rs:29

// Symmetric matrix by its lower triangle, row-major: row i holds columns j <= i, ascending.
struct SymLower {
    size_t n = 0;
    std::vector<uint64_t> ptr;  // n + 1
    std::vector<uint32_t> col;
    std::vector<double> val;
};

// LDL^T of P (s A s) P^T with s = diag(A)^-1/2 (1 where A_ii <= 0).
struct LdlFactor {
    size_t n = 0;
    std::vector<uint32_t> perm;   // perm[k]: matrix index eliminated k-th
    std::vector<double> scale;    // per elimination position k: s of perm[k]
    std::vector<int64_t> parent;  // elimination tree over positions (-1: root)
    std::vector<uint64_t> lp;     // strict lower L by columns (positions), n + 1
    std::vector<uint32_t> li;     // row positions, ascending within a column
    std::vector<double> lx;
    std::vector<double> d;        // pivots (LDL_GROUND where grounded)
    // the permuted upper triangle of A by columns, kept for a numeric refactor on new values of
    // the same pattern: entry e of column k reads A.val[upos[e]] at row position uri[e] <= k
    std::vector<uint64_t> up;
    std::vector<uint32_t> uri;
    std::vector<uint64_t> upos;
    std::vector<uint64_t> diag_pos;  // A.val index of each matrix index's diagonal (UINT64_MAX: none)
    uint64_t grounded = 0;
    uint64_t height = 0;            // elimination-tree height: the levels of each triangular solve
    double flops = 0.0;             // multiply-adds of the numeric factorization (sum of squared column counts)
    uint64_t len() const { return lp.empty() ? 0 : lp[n]; }
};

// Ordering and symbolic analysis. Matrix indices [0, nlast) are eliminated last, in their order
// (the reference orders the small coordinates last: solve.rs:776-789). The ordering is minimum
// degree, or nested dissection when that gives a much shorter elimination tree (a cheaper
// level-scheduled solve). Returns false when the factor would exceed max_len entries or
// max_flops multiply-adds (f then holds the counts only).
bool ldl_analyze(const SymLower& a, size_t nlast, uint64_t max_len, double max_flops, LdlFactor& f);
// Numeric factorization of a (same pattern as at analysis). Returns false on a non-finite pivot.
bool ldl_numeric(const SymLower& a, LdlFactor& f);
// x = A^+ b in place on the host (matrix index space), nrhs columns with row stride ld.
void ldl_solve_host(const LdlFactor& f, double* b, size_t nrhs, size_t ld);

// Selected inverse (Takahashi, Yamamoto and Kanenaga's recurrences, the form leverage.rs:22-26
// uses): Z = L^-T D^-1 L^-1 of the scaled permuted matrix on the pattern of L, zd its diagonal
// and zx its strictly lower entries at L's positions (column j, rows li[lp[j]..lp[j+1])). Every
// pair of indices that share a structural entry of A is in that pattern.
void ldl_selected_inverse(const LdlFactor& f, std::vector<double>& zd, std::vector<double>& zx);
// Z at positions (i, j) (any order) of the selected inverse; 0 outside the pattern.
double selinv_at(const LdlFactor& f, const std::vector<double>& zd, const std::vector<double>& zx, uint32_t i,
                 uint32_t j);

}  // namespace fh
