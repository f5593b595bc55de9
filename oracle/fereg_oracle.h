/* CPU oracle of the fixed-effect demeaning path (FP64, OpenMP threads; OMP_NUM_THREADS sets them).
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, the smoke check and bench.py's cpu_baseline leg may
 * load this library, and only as the checker: the product path never calls it.
 *
 * A plain-C restatement of fereg-core's cell structure, weight state, reduced operator, PCG
 * and back-substitution (crates/fereg-core/src/{structure,operator,cg,solve}.rs and
 * precond/{dense,jacobi}.rs). Every function cites the reference lines it follows.
 */
#ifndef FEREG_ORACLE_H
#define FEREG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_solver or_solver;

typedef struct {
    uint64_t n, nterms, g1, ncells, nb, m, m_small, small_work;
    int64_t fe1;          /* eliminated term, -1 = none (no absorb terms) */
    int64_t norm_fe1;     /* term absorbing the normalization shift (fe1_index) */
    int has_cell_obs_off; /* 0 when every cell holds one row */
} or_info;

/* codes[q] (n entries, kept order) with g[q] levels for nterms terms. */
int or_build(const uint32_t* const* codes, const uint64_t* g, uint64_t nterms, uint64_t n,
             uint64_t small_limit, const double* w_kept, or_solver** out);
void or_free(or_solver* s);
/* Threads the oracle's parallel loops use (omp_get_max_threads). */
int or_threads(void);
void or_get_info(const or_solver* s, or_info* out);
/* Block table: term, levels, offset, small flag for each of the nb blocks. */
void or_get_blocks(const or_solver* s, int64_t* term, uint64_t* g, uint64_t* off, int32_t* small);
/* Copies pi (n), seg_off (g1+1), cell_codes (M*nb), cell_obs_off (M+1 when present). */
void or_get_structure(const or_solver* s, uint32_t* pi, uint32_t* seg_off, uint32_t* cell_codes,
                      uint32_t* cell_obs_off);
/* Level-sorted cells of block b: perm (M), off (g_b + 1). */
void or_get_level_cells(const or_solver* s, uint64_t b, uint32_t* perm, uint32_t* off);
/* Weight state: n1 (g1), nr (m), diag (m), redundant (m), and sum of weights. */
void or_get_weight_state(const or_solver* s, double* n1, double* nr, double* diag, uint8_t* redundant,
                         double* sw);
/* Replaces the row weights (row order, NULL = unit). Returns 198 on non-positive weights. */
int or_set_weights(or_solver* s, const double* w_rows);

/* Reduced operator pieces (row-major, r columns). */
void or_apply(const or_solver* s, const double* p, uint64_t r, double* y);
void or_rhs(const or_solver* s, const double* rows, uint64_t r, double* rhs, double* b1, double* wsq,
            double* wsum);
void or_residualize_rows(const or_solver* s, double* rows, uint64_t r, const double* x,
                         const double* b1, double* alpha);
/* Dense grounded pseudo-inverse application on the small coordinates (all columns). */
void or_small_apply(const or_solver* s, const double* rin, uint64_t r, double* z);

/* Absorber::residualize: z (n x r, row order) in place. warm and coefs_out are (m+g1) x r or
 * NULL. Returns 0, 198 (shape) or 430 (maxiter). */
int or_residualize(or_solver* s, double* z, uint64_t r, double tol, uint64_t maxiter, const double* warm,
                   double* coefs_out, uint64_t* iterations, int32_t* exact);

#ifdef __cplusplus
}
#endif
#endif
