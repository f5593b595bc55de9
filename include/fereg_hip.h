/* fereg_hip: MI355X (gfx950) fixed-effect demeaning and estimation, C ABI.
 *
 * Drop-in for fereg-core's `Absorber` boundary (crates/fereg-core/src/absorber.rs:62-97) and the
 * estimation steps its callers run on the residualized block (ols.rs:234-449, fit/ols.rs:213-330,
 * fit/vcov.rs:278-382, glm/irls.rs:367-913). Plain pointers and sizes only. Every function that
 * can fail returns one of the reference's error codes (error.rs:6-37):
 *   FH_OK 0, FH_INTERRUPTED 1, FH_INVALID 198, FH_CONVERGENCE 430, FH_OOM 909,
 *   FH_TOO_FEW 2001, FH_INTERNAL 9999 (HIP errors are reported as FH_INTERNAL),
 * and fh_last_error() returns the message of the last failure on the calling thread.
 *
 * Threading: an absorber is Send-not-Sync (absorber.rs:62): one caller at a time, movable
 * between threads. Each absorber owns one HIP stream of default priority on the device that was
 * current when it was created; every call is synchronous with respect to the host. That stream
 * does not wait for other streams: device inputs (and outputs) of a _dev call must have no writes
 * still queued on another stream when the call is made (synchronize the producing stream first);
 * device outputs are complete when the call returns.
 *
 * Memory: functions ending in _dev take device pointers (HBM-resident inputs/outputs on the
 * absorber's device); the others take host pointers.
 */
#ifndef FEREG_HIP_H
#define FEREG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FH_OK 0
#define FH_INTERRUPTED 1
#define FH_INVALID 198
#define FH_CONVERGENCE 430
#define FH_OOM 909
#define FH_TOO_FEW 2001
#define FH_INTERNAL 9999

/* Precision of the reduced-system solve. */
#define FH_PLAIN_FP64 0 /* the reference's method in FP64 */
#define FH_MIXED 1      /* FP32 PCG with FP64 iterative refinement */

typedef struct fh_absorber fh_absorber;

/* One absorbed term: dense codes 0..g-1 in kept order, with level intercepts (intercept = 1) and/or
 * varying slopes (fe#c.x: nslopes kept-order columns, the term's slope variables). Replaces
 * `Term { codes, g, intercept, slopes }` (crates/fereg-core/src/session.rs). An absorber with a
 * slope term or a term without intercept is the varying-slopes solver (slopes.rs: SlopeSolver,
 * solve.rs:338-349 picks it the same way); at most 3 slopes per term (4 with the intercept).
 * Zero-initialized trailing fields give an intercept-only term. */
typedef struct {
    const uint32_t* codes;
    uint64_t g;
    int32_t intercept;            /* 1: level intercepts; 0: slopes only (nslopes >= 1) */
    int32_t nslopes;              /* slope columns */
    const double* const* slopes;  /* nslopes arrays of n doubles (host, or device for _dev) */
} fh_term;

/* SolverConfig (solve.rs:53-80): the knobs this solver honours. */
typedef struct {
    uint64_t small_limit; /* levels factored jointly and exactly (default 4096) */
    int32_t precision;    /* FH_PLAIN_FP64 or FH_MIXED */
    int32_t bucket_log2;  /* level records in buckets of 2^k FE1 levels; 0 = one bucket;
                             -1 (default) = one bucket per XCD from 2^19 FE1 levels on */
    int32_t autotune;     /* 1: the first S*p of each precision and column count times the
                             candidate launch plans and keeps the fastest, cached per structure
                             shape for the process (results may then differ in the last bits
                             between processes); 0: fixed default plans (bit-reproducible) */
    int32_t mixed_lead;   /* FH_MIXED: most FP64 PCG iterations before the FP32 correction
                             sweeps (0..8; -1 = the default, 8); the lead hands over earlier
                             once one FP32 sweep can reach the tolerance */
    int32_t sparse_chol;  /* 1 (default): a slowly converging solve escalates to exact sparse
                             factors of the reduced matrix; 0: Jacobi throughout */
    int32_t joint;        /* 1 (default): the escalation factors the joint matrix over every
                             coordinate when its budgets allow (the restart is then exact),
                             else each large block; 0: each large block only */
    int32_t dense_limit;  /* largest small super-block (levels) preconditioned by its dense
                             inverse (-1: the default, 64); a larger one keeps small_limit's
                             structure (row order) but is preconditioned like large blocks */
} fh_solver_config;

/* SolveOpts (absorber.rs:18-24) plus the precision of this solve. */
typedef struct {
    double tol;
    uint64_t maxiter;
    int32_t want_coefs;
    int32_t precision; /* -1 = the absorber's configured precision */
} fh_solve_opts;

/* SolveInfo (absorber.rs:27-37); refinements = FP64 outer sweeps of the mixed form; escalated = 1
 * when the solve's rate probe (iteration 12) replaced Jacobi by exact sparse factors (the joint
 * reduced matrix, or each large block) and restarted; the absorber keeps them for later solves. */
typedef struct {
    uint64_t iterations;
    int32_t escalated;
    int32_t exact;
    uint64_t refinements;
    double seconds;
} fh_solve_info;

/* Ctx (api.rs:198-229): cooperative interrupt polled between kernels, log sink. */
typedef struct {
    int (*interrupt)(void* user);
    void (*log)(void* user, const char* line);
    void* user;
    int32_t verbose;
} fh_ctx;

/* Structure summary (structure.rs:55-92). */
typedef struct {
    uint64_t n, nterms, g1, ncells, nb, m, m_small, small_work;
    int64_t fe1, norm_fe1;
    int32_t has_cell_obs_off;
    int32_t dense_small; /* 1: the small super-block is preconditioned by its dense inverse (m_small
                            <= dense_limit); 0: its blocks take Jacobi like the large ones */
} fh_structure_info;

const char* fh_last_error(void);
void fh_config_default(fh_solver_config* cfg);

/* solve::build / FeSolver::new (solve.rs:338-349, 396-465). w_kept: kept-order weights or NULL. */
int fh_absorber_create(const fh_term* terms, size_t nterms, const double* w_kept, size_t n,
                       const fh_solver_config* cfg, const fh_ctx* ctx, fh_absorber** out);
int fh_absorber_create_dev(const fh_term* terms, size_t nterms, const double* w_kept, size_t n,
                           const fh_solver_config* cfg, const fh_ctx* ctx, fh_absorber** out);
void fh_absorber_destroy(fh_absorber* ab);

/* Absorber::n, row_order (absorber.rs:64-69). row_order is borrowed, n entries, host memory. */
size_t fh_absorber_n(const fh_absorber* ab);
const uint32_t* fh_absorber_row_order(fh_absorber* ab);
const uint32_t* fh_absorber_row_order_dev(const fh_absorber* ab);

/* Absorber::set_weights (absorber.rs:74-78; solve.rs:1010-1139): row order, NULL = unit. */
int fh_absorber_set_weights(fh_absorber* ab, const double* w_rows, const fh_ctx* ctx);
int fh_absorber_set_weights_dev(fh_absorber* ab, const double* w_rows, const fh_ctx* ctx);

/* Absorber::residualize (absorber.rs:80-90; solve.rs:1141-1311): z is n x r row-major in row
 * order, residualized in place. warm / coefs_out: (m + g1) x r row-major or NULL. */
int fh_absorber_residualize(fh_absorber* ab, double* z, size_t r, const fh_solve_opts* o, const double* warm,
                            double* coefs_out, fh_solve_info* info, const fh_ctx* ctx);
int fh_absorber_residualize_dev(fh_absorber* ab, double* z, size_t r, const fh_solve_opts* o,
                                const double* warm, double* coefs_out, fh_solve_info* info,
                                const fh_ctx* ctx);

/* FeCoefs rows p = m + g1 (solve.rs:7-9); with varying slopes p = m + G_E b_E
 * ([x (m x r); alpha_E], E's levels level-major: intercept, then slopes; slopes.rs:30-31). */
size_t fh_absorber_coef_rows(const fh_absorber* ab);
/* Value vectors term_values returns: one per term for intercept-only absorbers; with varying
 * slopes one per slot (each term's intercept, then each slope; slopes.rs:2435-2464). */
size_t fh_absorber_term_slots(const fh_absorber* ab);
/* Absorber::fit_rows (absorber.rs:92; solve.rs:1313-1317): FE fit of column col per row. */
int fh_absorber_fit_rows(fh_absorber* ab, const double* coefs, size_t r, size_t col, double* out_rows);
/* Absorber::term_values (absorber.rs:94-96; solve.rs:1319-1343): out holds sum_q g_q values in
 * term order; lens (nterms) receives g_q. With varying slopes out holds one vector per slot
 * (fh_absorber_term_slots entries in lens): every intercept-only term but FE1 has weighted mean 0
 * and FE1's intercept takes the shifts; slope values are reported as solved. */
int fh_absorber_term_values(fh_absorber* ab, const double* coefs, size_t r, size_t col, double* out,
                            uint64_t* lens);

/* Structure and weight-state introspection (the bit-exact and parity checks). */
int fh_absorber_info(const fh_absorber* ab, fh_structure_info* out);
int fh_absorber_copy_structure(fh_absorber* ab, uint32_t* pi, uint32_t* seg_off, uint32_t* cell_codes,
                               uint32_t* cell_obs_off);
int fh_absorber_copy_blocks(const fh_absorber* ab, int64_t* term, uint64_t* g, uint64_t* off, int32_t* small);
int fh_absorber_copy_level_cells(fh_absorber* ab, size_t b, uint32_t* perm, uint32_t* off);
int fh_absorber_copy_weight_state(fh_absorber* ab, double* n1, double* nr, double* diag, uint8_t* redundant,
                                  double* sw);
/* Reduced operator y = S p (operator.rs:1375), host buffers m x r: the SpMV parity check. */
int fh_absorber_apply(fh_absorber* ab, const double* p, size_t r, double* y, int32_t precision);

/* y = S p as fh_absorber_apply, on DEVICE buffers p and y (m x r on the absorber's device), with
 * no allocation of any kind: no counterpart in the reference; for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Every
 * scratch array comes from one static device array of fh_noalloc_scratch_bytes() (64 MiB) per
 * GPU, held from the library's load to the process's end; a product that needs more is refused
 * with FH_INVALID (n = 1e7 with 1e6 / 1e5 / 10 levels fits in FP64 up to 2 columns, in FP32 up to
 * 4). Calls on one GPU share that array: calls that overlap on one GPU race. The absorber must
 * hold the caller's weights (not a Poisson fit's IRLS weights, whose restoring allocates).
 * Same kernels and launch plan as fh_absorber_apply: the same values, bit for bit. */
int fh_absorber_apply_noalloc(fh_absorber* ab, const double* p, size_t r, double* y, int32_t precision);
size_t fh_noalloc_scratch_bytes(void);

/* SpMV timing probe (no counterpart in the reference; measurement only): runs `reps` products
 * y = S p on the absorber's stream and returns the mean milliseconds per product (HIP events on
 * that stream) and the compulsory HBM bytes of one product,
 * 4 nb M + 12 g1 + 2 s m r + s M [weighted] (s = 8 FP64, 4 FP32). The product runs on rows
 * padded as the solvers pad them (r = 3 -> 4, 5..8 -> 8, 9..12 -> 12, 13..16 -> 16 columns, the
 * pads zero); the bytes count r columns. */
int fh_absorber_time_spmv(fh_absorber* ab, size_t r, int32_t precision, uint32_t reps, double* ms_per_spmv,
                          double* bytes_per_spmv);

/* Calibration of the S*p passes' ceiling (no counterpart in the reference): the time of one pass
 * that streams `count` random 32-bit codes below `levels` and gathers a table entry of entry_bytes
 * (4, 8 or 16) per code from an L2-resident table -- the access pattern of the segment and level
 * passes, whose limit is the L2's request rate, not HBM bandwidth. */
int fh_time_gather(size_t levels, size_t count, int32_t entry_bytes, uint32_t reps, double* ms_per_pass);

/* Launch plan of S*p for one precision and column count (no counterpart in the reference, whose
 * CPU operator has no launch shapes): the single small block as dense products inside the
 * segment and level passes (dense = 1) or gathered there (0), lanes per segment and lanes per
 * item of each block's level pass (powers of two 1..64, 0 = the default), and for a block whose
 * level records are cut into FE1 buckets the item shape (block_mode 1: one item per (bucket,
 * level), buckets spread over the XCDs; 0: one item per level over its buckets; -1: the default
 * rule, XCD buckets while a bucket's share of the segment sums fits in half an XCD's L2). get
 * returns the plan in use (the tuned one after an autotuned S*p; block_mode -1 for unbucketed
 * blocks); set pins one (e.g. a plan tuned in another process, for bit-identical results) and
 * marks it tuned. */
typedef struct {
    int32_t dense;
    uint32_t seg_lanes;
    uint32_t block_lanes[9];
    int32_t block_mode[9];
    int32_t tuned;
} fh_spmv_plan;
int fh_absorber_get_plan(const fh_absorber* ab, int32_t precision, size_t r, fh_spmv_plan* plan);
int fh_absorber_set_plan(fh_absorber* ab, int32_t precision, size_t r, const fh_spmv_plan* plan);
/* Forgets every tuned plan cached in this process. */
void fh_tune_cache_clear(void);

/* ------------------------------------------------------------------ estimation */

#define FH_VCOV_IID 0
#define FH_VCOV_HETERO 1
#define FH_VCOV_CLUSTER 2
#define FH_VCOV_HC2 3 /* fit/mod.rs:108-118: scores / sqrt(1 - h), the full model's exact leverage */
#define FH_VCOV_HC3 4 /* scores / (1 - h) */
#define FH_VCOV_CRV3 5 /* crv3.rs: leave-one-cluster-out jackknife (every FE re-estimated), 1-4 dimensions */

/* fixest small-sample rules (fit/vcov.rs:63-112, ssc of fit/mod.rs). */
#define FH_KFIXEF_NONE 0
#define FH_KFIXEF_FULL 1
#define FH_KFIXEF_NONNESTED 2
typedef struct {
    int32_t k_adj;
    int32_t k_fixef;
    int32_t g_adj;
    int32_t reserved;
} fh_ssc;

/* Result of an estimation on the absorber's sample. beta (K), vcov (K x K), kept (K: 1 = kept,
 * 0 = dropped as collinear; dropped coefficients are 0 with zero variance). */
typedef struct {
    double ssr;
    double df_k;
    double nparams;
    double deviance;    /* Poisson */
    uint64_t iterations; /* solver iterations (OLS) or IRLS iterations (Poisson) */
    uint64_t pcg_iterations;
    uint64_t n_clusters;
    int32_t converged;
    int32_t reserved;
} fh_fit_stats;

/* OLS of y on X (K columns) with the absorber's FEs (fit/ols.rs:213-330; fit/vcov.rs:316-382).
 * y (n), X (n x K row-major) in KEPT order; cluster: kept-order codes (n) or NULL; nested: per
 * absorbed term, whether it is nested in the cluster (fixest K.fixef = "nonnested"). */
int fh_ols_fit(fh_absorber* ab, const double* y, const double* X, size_t K, int32_t vcov_kind,
               const uint32_t* cluster, const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o,
               double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats);
int fh_ols_fit_dev(fh_absorber* ab, const double* y, const double* X, size_t K, int32_t vcov_kind,
                   const uint32_t* cluster, const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o,
                   double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats);

/* Several cluster dimensions (fit/vcov.rs:366-380; vcov.rs:480-558): the meat is the
 * inclusion-exclusion sum over the non-empty subsets of the dimensions of meats clustered by each
 * subset's intersection. g_df: FH_GDF_MIN (fixest's default ssc(G.df = "min"): the matrix is scaled
 * once by G/(G-1) with the smallest G) or FH_GDF_CONVENTIONAL (every subset meat by its own G/(G-1)).
 * As fixest does, a variance matrix with an eigenvalue <= 0 is rebuilt with the eigenvalues clipped
 * at 1e-16 (fit/vcov.rs:487-534; every fit entry point). stats->n_clusters = the smallest G. */
#define FH_GDF_MIN 0
#define FH_GDF_CONVENTIONAL 1
typedef struct {
    const uint32_t* const* codes; /* ndim arrays of n kept-order codes (host, or device for _dev) */
    size_t ndim;                  /* 1..4 */
    int32_t g_df;
    int32_t reserved;
} fh_clusters;
int fh_ols_fit_multi(fh_absorber* ab, const double* y, const double* X, size_t K, int32_t vcov_kind,
                     const fh_clusters* clusters, const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o,
                     double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats);
int fh_ols_fit_multi_dev(fh_absorber* ab, const double* y, const double* X, size_t K, int32_t vcov_kind,
                         const fh_clusters* clusters, const int32_t* nested, const fh_ssc* ssc,
                         const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats);

/* Poisson PML by IRLS with warm-started weighted residualization (glm/irls.rs:367-913, fixest
 * flavour: fit/glm.rs; family.rs:103-145), robust (hetero), iid or cluster variance. y, X and
 * offset (NULL = none) in KEPT order. The absorber's weights (fh_absorber_create / set_weights)
 * are the prior weights. The solver's weights are set every IRLS iteration; the absorber puts
 * the prior weights back before its next call that is not a Poisson fit. ctx may be NULL; its
 * interrupt callback is polled between IRLS iterations and between PCG iterations (rc 1). */
int fh_ppml_fit(fh_absorber* ab, const double* y, const double* X, size_t K, const double* offset, int32_t vcov_kind,
                const uint32_t* cluster, const int32_t* nested, const fh_ssc* ssc, double tol, uint64_t maxiter,
                const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept, fh_fit_stats* stats,
                const fh_ctx* ctx);
int fh_ppml_fit_dev(fh_absorber* ab, const double* y, const double* X, size_t K, const double* offset,
                    int32_t vcov_kind, const uint32_t* cluster, const int32_t* nested, const fh_ssc* ssc, double tol,
                    uint64_t maxiter, const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept,
                    fh_fit_stats* stats, const fh_ctx* ctx);

/* fh_ppml_fit with several cluster dimensions (see fh_clusters). */
int fh_ppml_fit_multi(fh_absorber* ab, const double* y, const double* X, size_t K, const double* offset,
                      int32_t vcov_kind, const fh_clusters* clusters, const int32_t* nested, const fh_ssc* ssc,
                      double tol, uint64_t maxiter, const fh_solve_opts* o, double* beta, double* vcov, int32_t* kept,
                      fh_fit_stats* stats, const fh_ctx* ctx);
int fh_ppml_fit_multi_dev(fh_absorber* ab, const double* y, const double* X, size_t K, const double* offset,
                          int32_t vcov_kind, const fh_clusters* clusters, const int32_t* nested, const fh_ssc* ssc,
                          double tol, uint64_t maxiter, const fh_solve_opts* o, double* beta, double* vcov,
                          int32_t* kept, fh_fit_stats* stats, const fh_ctx* ctx);

/* WCR11 wild cluster bootstrap of one linear hypothesis R beta = r on the OLS fit, Rademacher
 * weights, clusters = bootstrap clusters (wildboot/ols.rs:243-649 run_cluster_11; weights.rs;
 * rng.rs SplitMix64; mod.rs:529-749 p-value and test-inverted interval). The B x G sign by
 * cluster-score contraction runs on int8 MFMA with the scores split into 7 signed 8-bit limbs
 * (exact int32 accumulation, FP64-equivalent). Fixed effects not nested in the clusters
 * (wildboot/mod.rs:959-997) are projected out of the draws exactly. Path 1 (m_nn <= min(G, 512)
 * levels): as appended pseudo-regressors when kept + m_nn <= 19, else by Path 1's algebra with
 * the draws as FP64 matrix products (rocBLAS, atomics off). Path 2 (G <= 512 and more levels):
 * one cluster at a time, one FP64 residualization per cluster at tolerance min(o->tol, 1e-10).
 * More than 512 levels and clusters: not projected (the reference's "TooLarge" plan), result
 * flag set. */
#define FH_PVAL_TWO_TAILED 0
#define FH_PVAL_EQUAL_TAILED 1
#define FH_PVAL_GREATER 2
#define FH_PVAL_LESS 3
typedef struct {
    uint64_t reps;        /* replications (full enumeration when G <= 20 and 2^G <= reps) */
    uint64_t seed;        /* SplitMix64 seed; draw b uses stream(seed, b) */
    const double* r_mat;  /* R, K entries (host memory) */
    double r;             /* tested value */
    double level;         /* confidence level of the interval, e.g. 0.95 */
    double tol;           /* interval search tolerance (1e-6) */
    uint64_t maxiter;     /* interval search iterations (50) */
    int32_t pval_type;    /* FH_PVAL_* */
    int32_t conf_int;     /* 1: invert the test for an interval */
    int32_t boot_type;    /* 11 (or 0): WCR11; 13, 31, 33: MacKinnon-Nielsen-Webb (2023) with the
                             CRV3 jackknife in the variance (13), the data generating process (31)
                             or both (33) (wildboot/ols.rs:651-1082) */
    int32_t reserved;
} fh_wboot_spec;
typedef struct {
    double stat, p_val, se, point, ci_lo, ci_hi, ssc;
    uint64_t reps, reps_feasible, n_clusters;
    int32_t full_enumeration;
    int32_t fe_not_projected; /* absorbed terms not nested in the clusters were left in the scores */
    double seconds_bootstrap;
    double contraction_ms;    /* the MFMA contraction kernel (HIP events) */
    double contraction_ops;   /* its int8 operations (2 per multiply-add, padding included) */
} fh_wboot_result;
/* OLS (as fh_ols_fit, one-way clustered variance by `cluster`) followed by the bootstrap;
 * t_boot (reps entries, host) may be NULL. Codes must be the dense cluster factor 0..G-1. */
int fh_ols_wildboot(fh_absorber* ab, const double* y, const double* X, size_t K, const uint32_t* cluster,
                    const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o, const fh_wboot_spec* spec,
                    double* beta, double* vcov, int32_t* kept, fh_wboot_result* res, double* t_boot);
int fh_ols_wildboot_dev(fh_absorber* ab, const double* y, const double* X, size_t K, const uint32_t* cluster,
                        const int32_t* nested, const fh_ssc* ssc, const fh_solve_opts* o, const fh_wboot_spec* spec,
                        double* beta, double* vcov, int32_t* kept, fh_wboot_result* res, double* t_boot);

/* OLS with reghdfe 6.13.1's conventions (Session::ols, ols.rs:234-706; dof.rs:64-182): the
 * variables are standardized by their unweighted sd; stage 1 drops regressors explained by the
 * fixed effects (sum x~^2 / sum x^2 <= min(1e-6, tol / 10)), stage 2 sweeps the rest in order
 * (Mata invsym on the standardized Gram); the constant is the regression through the weighted
 * means; unadjusted, robust (pweight + unadjusted = robust) or multiway cluster variance with
 * reghdfe's degrees of freedom (nested terms, redundant levels, pairwise mobility groups) and
 * small-sample factors, the CGM PSD repair for several cluster dimensions; the Wald F test.
 * The absorber holds the model's raw weights (kept order at creation) and weight_kind says what
 * they are; a model without absorbed terms uses an absorber with one one-level term and nterms =
 * 0 here. Outputs: b (K + 1: the regressors, then _cons when reported), V ((K + 1) x (K + 1)
 * row-major over the same k_all = K + report_constant entries), status (K: 0 kept, 2 absorbed by
 * the fixed effects, 3 collinear), dof_table (5 per slot, max(nterms, 1) slots: K, M, K - M,
 * inexact, nested). Omitted coefficients are 0 with zero variance. */
#define FH_VCE_UNADJUSTED 0
#define FH_VCE_ROBUST 1
#define FH_VCE_CLUSTER 2
typedef struct {
    int32_t vce;                     /* FH_VCE_* */
    int32_t weight_kind;             /* FH_WEIGHT_* of the absorber's weights */
    const fh_term* terms;            /* kept-order codes of the absorbed terms (degrees of freedom) */
    size_t nterms;
    const uint32_t* const* clusters; /* FH_VCE_CLUSTER: kept-order codes of 1..4 dimensions */
    size_t nclusters;
    int32_t noconstant;
    int32_t dof_firstpair;           /* DofOptions (api.rs): dofadjustments() */
    int32_t dof_pairwise;
    int32_t dof_clusters;
    int32_t dof_continuous;
    int32_t precision;               /* -1 = the absorber's */
    double tol;                      /* solver tolerance (reghdfe tolerance()) */
    uint64_t maxiter;
    uint64_t num_singletons;         /* reported in N_full */
} fh_reghdfe_spec;
/* e() scalars (NaN = missing). */
typedef struct {
    double N, rank, df_r, df_m, df_a, df_a_initial, df_a_redundant, df_a_nested;
    double N_clust, N_clustervars, N_clust_dim[4];
    double tss, tss_within, rss, mss, rmse, F, ll, ll_0, r2, r2_within, r2_a, r2_a_within, sumweights;
    double num_singletons, N_full, report_constant, converged, y_absorbed;
    uint64_t iterations, nslots;
    int32_t psd_fixed;
    int32_t reserved;
} fh_reghdfe_stats;
void fh_reghdfe_spec_default(fh_reghdfe_spec* spec);
int fh_reghdfe_fit(fh_absorber* ab, const double* y, const double* X, size_t K, const fh_reghdfe_spec* spec,
                   double* b, double* V, int32_t* status, double* dof_table, fh_reghdfe_stats* stats,
                   const fh_ctx* ctx);
int fh_reghdfe_fit_dev(fh_absorber* ab, const double* y, const double* X, size_t K, const fh_reghdfe_spec* spec,
                       double* b, double* V, int32_t* status, double* dof_table, fh_reghdfe_stats* stats,
                       const fh_ctx* ctx);

/* ------------------------------------------------------------------ factors and the sample */

/* A factor column (fit::FactorCol, fit/mod.rs; factor::FactorValues::Num, api.rs). Missing
 * entries: non-finite doubles, INT32_MIN / INT64_MIN, codes outside base .. base + nlevels - 1. */
#define FH_COL_F64 0
#define FH_COL_I32 1
#define FH_COL_I64 2
#define FH_COL_CODES 3 /* uint32 codes */
typedef struct {
    int32_t kind;
    int32_t reserved;
    const void* data; /* n values of the kind */
    int64_t base;     /* FH_COL_CODES */
    uint64_t nlevels; /* FH_COL_CODES */
} fh_factor_col;

/* factor::factorize_col (factor.rs:307-345): dense codes over the n rows in ascending value
 * order (numeric order; -0.0 and 0.0 are one level); missing entries share one extra level, code
 * 0, present only when some row is missing (*n_missing rows; may be NULL). strict = 1 is
 * factor::factorize (factor.rs:44-76) of a double column: infinities are levels and a NaN is an
 * error (FH_INVALID "missing values in factor"). Codes are n uint32, host or device (_dev). */
int fh_factorize(const fh_factor_col* col, size_t n, int32_t strict, uint32_t* codes, uint64_t* nlevels,
                 uint64_t* n_missing);
int fh_factorize_dev(const fh_factor_col* col, size_t n, int32_t strict, uint32_t* codes, uint64_t* nlevels,
                     uint64_t* n_missing);
/* factor::combine (factor.rs:259-281): the interaction a#b#..., codes in lexicographic order of
 * (a, b, ...) over the combinations that occur. FH_INVALID when the product of the level counts
 * reaches 2^63. parts: nparts arrays of n codes, each below its nlevels. */
int fh_combine(const uint32_t* const* parts, const uint64_t* nlevels, size_t nparts, size_t n, uint32_t* codes,
               uint64_t* out_nlevels);
int fh_combine_dev(const uint32_t* const* parts, const uint64_t* nlevels, size_t nparts, size_t n,
                   uint32_t* codes, uint64_t* out_nlevels);
/* factor::restrict (factor.rs:285-305): the codes of the rows kept (ascending row indices),
 * renumbered densely in their order. */
int fh_restrict(const uint32_t* codes, uint64_t nlevels, size_t n, const uint32_t* kept, size_t nkept, uint32_t* out,
                uint64_t* out_nlevels);
int fh_restrict_dev(const uint32_t* codes, uint64_t nlevels, size_t n, const uint32_t* kept, size_t nkept,
                    uint32_t* out, uint64_t* out_nlevels);

/* sample::drop_singletons (sample.rs:44-74): cycles over the terms in order, dropping the live
 * rows (alive[i] = 1) of levels with one live row (singletons = 1) and, with y_sep, of levels whose
 * outcome sum is 0 (Poisson) or 0 or the count (logit: nonzero outcomes count 1), until no term
 * changes. fw: frequency weights (counts are weight sums) or NULL. Terms' intercept is ignored. */
#define FH_FAMILY_GAUSSIAN 0
#define FH_FAMILY_POISSON 1
#define FH_FAMILY_LOGIT 2
int fh_drop_singletons(const fh_term* terms, size_t nterms, size_t n, uint8_t* alive, const double* fw,
                       const double* y_sep, int32_t family, int32_t singletons, uint64_t* n_singletons,
                       uint64_t* n_separated);
int fh_drop_singletons_dev(const fh_term* terms, size_t nterms, size_t n, uint8_t* alive, const double* fw,
                           const double* y_sep, int32_t family, int32_t singletons, uint64_t* n_singletons,
                           uint64_t* n_separated);

/* The estimation sample (fit/input.rs:193-407 scan + build; session.rs:458-490 prepare): factor
 * codes, the missing-value and zero-weight scan, interactions, singleton and fe-separation
 * removal, then the kept rows and the term and cluster codes over them. */
#define FH_WEIGHT_NONE 0
#define FH_WEIGHT_AWEIGHT 1
#define FH_WEIGHT_FWEIGHT 2
#define FH_WEIGHT_PWEIGHT 3
typedef struct {
    size_t n;
    const double* y;              /* n, or NULL */
    const double* const* x;       /* nx columns of n values */
    size_t nx;
    const double* weights;        /* n, or NULL */
    const double* offset;         /* n, or NULL */
    const uint8_t* drop;          /* n, or NULL: 1 = a row the caller excludes (FitInput::drop) */
    const fh_factor_col* factors; /* nfactors columns of n values */
    size_t nfactors;
    const uint32_t* term_factors; /* factor indices of every absorbed term, concatenated; term q */
    const uint32_t* term_off;     /*   is the interaction of term_factors[term_off[q] .. term_off[q+1]) */
    size_t nterms;
    const uint32_t* cluster_factors; /* the same for cluster dimensions (at most 4) */
    const uint32_t* cluster_off;
    size_t nclusters;
    int32_t weight_kind;     /* FH_WEIGHT_*: frequency weights are summed in the singleton counts */
    int32_t family;          /* FH_FAMILY_* */
    int32_t missing;         /* 1: fit layer: rows with a missing value (y, x, offset, weights, a
                                used factor, drop) or a zero weight are dropped and counted; a
                                negative weight is an error. 0: session: every row is used and a
                                missing factor value is an error */
    int32_t drop_singletons; /* 1: cyclic singleton removal */
    int32_t separation;      /* 1: fixest's fe separation rule for Poisson and logit */
    int32_t reserved;
} fh_sample_spec;
/* fit::Dropped (fit/mod.rs:473-500). */
typedef struct {
    uint64_t n_input, nobs;
    uint64_t na, zero_weight, singletons, perfect_fit;
    uint64_t na_y, na_x, na_fe, na_cluster, na_weights, na_offset, na_extra;
} fh_sample_counts;
typedef struct fh_sample fh_sample;
/* Builds on the current device; spec arrays on the host (fh_sample_build) or the device (_dev).
 * FH_TOO_FEW when fewer than 2 rows remain. */
int fh_sample_build(const fh_sample_spec* spec, const fh_ctx* ctx, fh_sample** out);
int fh_sample_build_dev(const fh_sample_spec* spec, const fh_ctx* ctx, fh_sample** out);
void fh_sample_destroy(fh_sample* s);
int fh_sample_get_counts(const fh_sample* s, fh_sample_counts* out);
/* Kept input rows (nobs, ascending); term / cluster codes over them (nobs, dense) with their
 * level counts. Device pointers stay valid until fh_sample_destroy. */
const uint32_t* fh_sample_kept_dev(const fh_sample* s);
int fh_sample_kept(const fh_sample* s, uint32_t* rows);
const uint32_t* fh_sample_term_dev(const fh_sample* s, size_t q, uint64_t* g);
int fh_sample_term(const fh_sample* s, size_t q, uint32_t* codes, uint64_t* g);
const uint32_t* fh_sample_cluster_dev(const fh_sample* s, size_t q, uint64_t* g);
int fh_sample_cluster(const fh_sample* s, size_t q, uint32_t* codes, uint64_t* g);
/* out[j] = col[kept[j]]: an input column (n) on the kept rows (nobs). */
int fh_sample_gather(const fh_sample* s, const double* col, double* out);
int fh_sample_gather_dev(const fh_sample* s, const double* col, double* out);

/* Serving engine (no reference counterpart: fereg-core is a library called once per fit).
 * Continuous batching of independent regressions: a FIFO of requests, each admitted when one of
 * `slots` workers is free and its HBM footprint fits `hbm_budget` beside the requests already
 * on the device; every worker runs its request on its own stream, so requests overlap on the
 * GPU. A result does not depend on what else is running (fixed-order reductions). */
#define FH_MODEL_OLS 0
#define FH_MODEL_PPML 1
typedef struct fh_server fh_server;
typedef struct {
    uint32_t slots;      /* requests in flight at once (1..64) */
    uint64_t hbm_budget; /* bytes admitted at once; 0 = three quarters of the free memory */
    int32_t precision;   /* FH_PLAIN_FP64 or FH_MIXED */
    int32_t pack;        /* most requests packed into one solve (default 32; 0 or 1: each alone):
                            consecutive small unweighted OLS requests with iid or hetero variance
                            share one absorber on disjoint level ranges and one PCG launch
                            sequence, each result within its own tolerance of its solve alone */
} fh_server_config;
typedef struct {
    int32_t model;       /* FH_MODEL_OLS or FH_MODEL_PPML */
    int32_t vcov_kind;   /* FH_VCOV_* */
    size_t n, K, nterms;
    const fh_term* terms; /* host codes, kept order */
    const double* y;      /* host, n */
    const double* X;      /* host, n x K */
    const uint32_t* cluster;
    const int32_t* nested;
    fh_ssc ssc;
    double tol;
    uint64_t maxiter;
    double* beta;         /* outputs, host: K, K x K, K */
    double* vcov;
    int32_t* kept;
} fh_request;
typedef struct {
    int32_t rc;
    int32_t packed;                      /* requests in the solve this one ran in (1: alone) */
    double queued_s, service_s, total_s; /* wait for admission, on the device, submit to done */
    uint64_t iterations, pcg_iterations;
    double start_s, end_s;               /* admission and completion, seconds since fh_server_create */
} fh_response;
void fh_server_config_default(fh_server_config* cfg);
int fh_server_create(const fh_server_config* cfg, fh_server** out);
/* Every pointer of the request must stay valid until fh_server_wait returns for its ticket. */
int fh_server_submit(fh_server* s, const fh_request* req, uint64_t* ticket);
int fh_server_wait(fh_server* s, uint64_t ticket, fh_response* resp);
void fh_server_destroy(fh_server* s);

/* SparseChol::try_factor / solve (crates/fereg-core/src/precond/sparse_chol.rs:216-449): the exact
 * LDL^T of a symmetric positive semidefinite matrix given by its lower triangle (CSR: row i holds
 * columns <= i, ascending), scaled to unit diagonal, with every pivot <= 1e-10 grounded (its
 * coordinate's solution component becomes ~0: a reflexive generalized inverse, as the solver's
 * escalated preconditioner uses it). Matrix indices [0, nlast) are eliminated last. Host memory
 * and host compute; solve works in place on an n x nrhs row-major block. */
typedef struct fh_sparse_chol fh_sparse_chol;
int fh_sparse_chol_factor(size_t n, const uint64_t* ptr, const uint32_t* col, const double* val, size_t nlast,
                          fh_sparse_chol** out);
/* len: factor entries; flops: multiply-adds; grounded: grounded pivots; height: the elimination
 * tree's height (levels of a triangular solve; nested dissection replaces minimum degree when it
 * shortens the tree enough). */
int fh_sparse_chol_info(const fh_sparse_chol* f, uint64_t* len, double* flops, uint64_t* grounded, uint64_t* height);
int fh_sparse_chol_solve(const fh_sparse_chol* f, double* b, size_t nrhs);
/* Entries (rows[q], cols[q]) of the same generalized inverse by selected inversion on the factor's
 * pattern (leverage.rs:22-26: Takahashi recurrences; every pair sharing a structural entry of the
 * matrix is covered, other pairs read 0). */
int fh_sparse_chol_inverse_entries(const fh_sparse_chol* f, size_t npairs, const uint32_t* rows,
                                   const uint32_t* cols, double* out);
void fh_sparse_chol_destroy(fh_sparse_chol* f);

/* Device memory helpers for callers without another allocator. */
int fh_device_alloc(void** p, size_t bytes);
int fh_device_free(void* p);
int fh_memcpy_h2d(void* dst, const void* src, size_t bytes);
int fh_memcpy_d2h(void* dst, const void* src, size_t bytes);
int fh_device_synchronize(void);
/* The library keeps freed scratch blocks in a stream-ordered cache (no reference counterpart:
 * the CPU solver allocates with the system allocator). Returns every cached block whose last
 * use has completed to the device; in_use/cached report the pool's bytes (either may be null). */
int fh_release_cached_memory(void);
int fh_memory_stats(size_t* in_use, size_t* cached);

#ifdef __cplusplus
}
#endif
#endif
