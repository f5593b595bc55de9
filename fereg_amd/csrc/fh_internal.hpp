// Internal state of an absorber and the host launchers of its kernels.
//
// Data layout in HBM (one absorber, n rows, M cells, g1 segments, m reduced coordinates):
//   pi            u32[n]      kept-observation index of every row (row order = sorted order)
//   seg_off       u32[g1+1]   first cell of every FE1 level
//   cell_codes    u32[M*nb]   block codes per cell, row-major (cell, block)
//   large_codes   u32[M*(nb-1)] the same without block 0, when block 0 is the dense small block
//   cell_obs_off  u32[M+1]    first row of every cell (absent when every cell is one row)
//   cell_seg      u32[M]      FE1 level of every cell
//   per block b, the level-major transpose of the cells in segment buckets (built once):
//     the FE1 levels are cut into nbk[b] buckets of bwidth[b] consecutive segments (one per XCD
//     once g1 is large, more when the block's levels are long: ~2048 cells per item); cells are
//     ordered by (bucket, level, cell). Item i = bucket * g_b + level:
//     boff        u32[nbk*g_b+1] item offsets;  bperm u32[M] cells in (bucket, level, cell) order
//     lseg        u32[M]      FE1 level of the cell at each position
//     lother      u32[M*(nb-1)] codes of the other blocks at each position
//     lw          T[M]        cell weight at each position (weighted designs)
//   lperm/loff: the plain (level, cell) lists of structure.rs:144-153, built on demand.
//   weight state: cw T[M] (null = unit), inv_n1 T[g1], inv_diag T[m], minv T[ms*ms] (dense
//   grounded pseudo-inverse of the small super-block); with one small block of <= 32 levels,
//   the dense small-row operator E T[ms*sm_ld] (D and C side by side), nrs T[ms]; that block
//   then has no level-major transpose unless one is requested.
#pragma once

#include <algorithm>
#include <cmath>
#include <memory>

#include "fh_common.hpp"

namespace fh {

constexpr int NB_MAX = 8;  // non-FE1 blocks supported by the kernels
// Largest small super-block (levels) preconditioned by its dense grounded inverse; above it the
// small blocks take Jacobi and level passes like large ones. Measured on MI355X (three-term panels,
// n = 1.2e5 - 4e5): a dense block of 1203 / 2753 / 4010 levels cost 9.9 / 78 / 194 ms per fit
// against 1.6 / 2.8 / 2.5 ms with Jacobi (8 PCG iterations); up to 64 levels the inverse is a host
// computation of microseconds.
constexpr uint64_t DENSE_LIMIT_DEFAULT = 64;
// Most FP64 iterations a mixed-precision solve runs before its FP32 correction sweeps (default).
constexpr uint32_t MIXED_LEAD_DEFAULT = 8;
// MI355X: 8 accelerator complex dies, each with its own L2; the hardware hands consecutive
// workgroups of a launch to the XCDs in turn.
constexpr int XCD_COUNT = 8;
// Working set of gathered rows that a level pass keeps per XCD (half of its 4 MB L2).
constexpr size_t L2_TARGET = 2u << 20;
// Automatic FE1 buckets (one per XCD) from this many FE1 levels on.
constexpr uint64_t AUTO_BUCKET_G1 = 1u << 19;

struct BlockInfo {
    int64_t term;
    uint64_t g, off;
    bool small;
};

// Weight-dependent arrays in one precision.
template <class T>
struct WeightArrays {
    DBuf<T> cw;        // M (empty = every cell weight is 1)
    DBuf<T> inv_n1;    // g1
    DBuf<T> inv_diag;  // m (0 on redundant coordinates)
    std::vector<DBuf<T>> lw;  // per block, M (empty when unit)
    DBuf<T> minv;      // ms*ms
    // dense small-block operator (Absorber::small_dense): y_S = nr_S p_S + C p_L - D t
    // E = ms x sm_ld: D (ms x g1, D[k][s] = sum of W_c over the cells of segment s with small
    // level k) in columns [0, g1), C (ms x (m - ms), C[k][l] = sum of W_c over the cells with small
    // level k and large coordinate l) in columns [sm_coff, sm_coff + m - ms), zero pads
    DBuf<T> esm;
    DBuf<T> nrs;       // ms, n_r of the small coordinates
};

// Per-call scratch of the PCG and operator kernels (grown on demand).
template <class T>
struct Scratch {
    DBuf<T> r, z, p, q, x, rhs, t, zc, part, ypart, ds, dl;
    DBuf<double> dpart;  // FP64 partial sums (dots)
};

// Launch choices of S*p (route and lane widths); zero widths mean the absorber's defaults.
struct SpmvPlan {
    int dense = 1;           // the single small block as dense products inside the passes too
    uint32_t seg_lanes = 0;  // lanes per segment of the segment pass
    std::vector<uint32_t> blanes;  // lanes per item of each block's level pass
    std::vector<int> xcd;          // per block: 1 (bucket, level) items spread over the XCDs, 0 one
                                   // item per level over its buckets, -1 the default rule
    bool tuned = false;
};

struct ClusterRuns;
struct FacGpu;      // fh_escalate.hip
struct SlopeState;  // fh_slopes.hip

// The PCG's status as the device publishes it to pinned host memory after every iteration
// (zero-copy: the host polls it without a copy kernel); iters_done (the iteration count, the done
// flag in bit 63) is written last.
struct PollMirror {
    uint32_t err, pad;
    uint64_t iters_done;
    double probe_rho, probe_rem;
};
// Its pinned host allocation (fh_pcg.hip), one per absorber.
struct HostPoll {
    PollMirror* h = nullptr;  // host view
    PollMirror* d = nullptr;  // device view of the same memory
    HostPoll() = default;
    HostPoll(const HostPoll&) = delete;
    HostPoll& operator=(const HostPoll&) = delete;
    ~HostPoll();
    void ensure();
};

struct Absorber {
    int device = 0;
    hipStream_t st = nullptr;
    // structure
    uint64_t n = 0, nterms = 0, g1 = 0, M = 0, nb = 0, m = 0, ms = 0, small_work = 0;
    int64_t fe1 = -1, norm_fe1 = -1;
    std::vector<uint64_t> term_g;
    std::vector<DBuf<uint32_t>> codes_kept;  // the terms' codes, kept order (n each)
    std::vector<BlockInfo> blocks;
    uint64_t small_limit = 4096;
    uint64_t dense_limit = DENSE_LIMIT_DEFAULT;  // largest small super-block factored densely
    // the structure's own small/large split (as the reference classifies): ms and the blocks'
    // small flags describe the solver's dense region, which is empty when it exceeds dense_limit
    uint64_t ms_struct = 0;
    std::vector<uint8_t> small_struct;
    int precision = FH_PLAIN_FP64;
    DBuf<uint32_t> pi, seg_off, cell_codes, cell_obs_off, cell_seg;
    // codes of the large blocks only (M x (nb - 1), row-major), for the segment pass when the one
    // small block (block 0) goes through the dense products: it then streams only these
    DBuf<uint32_t> large_codes;
    std::vector<DBuf<uint32_t>> lperm, loff;                // plain level lists, on demand
    std::vector<DBuf<uint32_t>> bperm, boff, lseg, lother;  // bucketed transpose (large blocks)
    std::vector<uint32_t> nbk;                              // buckets per block
    std::vector<uint32_t> bwidth;                           // FE1 levels per bucket
    std::vector<uint32_t> blanes;                           // lanes per item in the level passes
    std::vector<uint32_t> llanes;                           // lanes per level (one item per level)
    int bucket_log2 = -1;  // configured FE1 bucket size 2^k of the level records (0 one bucket, -1 auto)
    std::vector<uint32_t> host_pi;
    bool host_pi_valid = false;
    // weights
    bool weighted = false;
    DBuf<double> w_rows;  // n (row order) when weighted
    // A Poisson fit sets the solver's weights every IRLS iteration and leaves its last ones in
    // place (irls_weights); the caller's own weights (row order, empty = unit) wait in user_w and
    // come back before the next call that is not a Poisson fit (restore_user_weights).
    bool irls_weights = false;
    DBuf<double> user_w;
    // sum of the weights: reduced on the device at each weights refresh (sw_dev) and read with
    // the next readback that needs it (absorber_sw, or the rhs sums of a residualization)
    double sw = 0.0;
    bool sw_ready = false;
    DBuf<double> sw_dev;
    DBuf<double> n1, nr, diag;
    DBuf<uint8_t> red;
    uint64_t nkeep = 0;
    WeightArrays<double> w64;
    WeightArrays<float> w32;
    bool w32_valid = false;
    // scratch
    Scratch<double> s64;
    Scratch<float> s32;
    DBuf<double> b1, rows_tmp;
    DBuf<double> status;  // PCG scalars
    DBuf<uint32_t> u32tmp;

    // one small block with few levels: its part of S p is dense products with D and C (fewer
    // bytes than a level pass over the cells), set by refresh_weights
    bool small_dense = false;
    size_t sm_ld = 0, sm_coff = 0;  // pitch of E and its first C column (multiples of 4)
    // E as u8 counts (no prior weights), valid while *esm8_bad == 0 on the device
    DBuf<uint8_t> esm8;
    DBuf<uint32_t> esm8_bad;
    bool esm8_ok = false;  // host copy of *esm8_bad == 0 after the last weight refresh

    // S*p plans per precision (FP32, FP64) and column class (r <= 1, 2, 4, 8, 12, 16); with
    // autotune the first S*p of a class times the candidates (autotune_spmv)
    bool autotune = false;
    mutable SpmvPlan plans[2][6];
    SpmvPlan& plan_for(bool fp64, int r) const {
        const int c = r <= 1 ? 0 : r <= 2 ? 1 : r <= 4 ? 2 : r <= 8 ? 3 : r <= 12 ? 4 : 5;
        return plans[fp64 ? 1 : 0][c];
    }

    // runs of the clusters of the current call's codes (set for the duration of one API call
    // that uses one clustering more than once)
    const ClusterRuns* cl_runs = nullptr;

    // escalated preconditioner (fh_escalate.hip): exact sparse factors of the joint S or of the
    // large blocks, with the solver knobs of SolverConfig::{sparse_chol, joint}
    std::vector<std::shared_ptr<FacGpu>> facs;
    bool fac_joint = false;   // one factor over every coordinate
    bool fac_fresh = false;   // built for the current weights
    bool fac_failed = false;  // no factor fit the budgets: Jacobi for good
    bool fac_failed_joint = false;  // the joint factor exceeded a hard budget (or failed)
    bool sparse_chol = true;
    int joint_policy = 1;     // 0 off, 1 on escalation and at the early probe's cost rule
    // what is known of the joint factor's cost (structure only): the assembly's item pairs, then
    // the nonzeros, factor entries and multiply-adds once it was assembled and analyzed
    double joint_pairs = -1.0, joint_nnz = -1.0, joint_len = -1.0, joint_flops = -1.0;
    uint32_t mixed_lead = MIXED_LEAD_DEFAULT;  // most FP64 iterations of a mixed solve before its FP32 sweeps
    HostPoll poll;  // the PCG status mirror

    // Varying slopes (fh_slopes.hip): set when a term has slopes or no intercept. The absorber then
    // has an observation-level structure of its own; pi, n, m, ms, g1 (= E's levels x its width,
    // the FeCoefs tail), the weights, w64.inv_diag (zero: the block-Jacobi factors cover the large
    // coordinates) and w64.minv (the dense small inverse) keep their meaning for the PCG.
    std::shared_ptr<SlopeState> slope;
    std::vector<uint8_t> term_icpt;  // per term: level intercepts (1 for every intercept-only term)
    std::vector<uint32_t> term_ns;   // per term: slope columns

    bool trivial() const { return cell_obs_off.n == 0; }
    bool unit() const { return !weighted && trivial(); }
    uint32_t seg_lanes = 8;  // lanes per segment in the segment passes
};

// ---------------------------------------------------------------- structure (fh_structure.hip)
// Builds the cell structure from device term codes (kept order).
void build_structure(Absorber& a, const std::vector<const uint32_t*>& dcodes);
// Level-sorted cells of block b (built on demand for small blocks).
void ensure_level_cells(Absorber& a, size_t b);
// The bucketed level-major transpose of block b (every block but the dense small one is built
// with the structure; ensure builds a missing one).
void build_block_transpose(Absorber& a, int b);
void ensure_block_transpose(Absorber& a, int b);

// The dense small-row operator applies: a single small block of <= 32 levels whose D and C are
// not much larger than the cell records a level pass over it would stream. A function of the
// structure only.
inline bool small_dense_rule(const Absorber& a) {
    int nsb = 0;
    for (auto& bi : a.blocks) nsb += bi.small ? 1 : 0;
    const size_t ms = a.ms, m = a.m, g1 = a.g1, M = a.M;
    return nsb == 1 && ms >= 1 && ms <= 32 && ms * (g1 + (m - ms)) <= 4 * std::max<size_t>(M, 1);
}

// ---------------------------------------------------------------- weights (fh_weights.hip)
// Escalation (solve.rs:685-891, cg.rs:254-274; fh_escalate.hip): when the PCG converges slowly
// (rho_hat > 0.6 at its probe) the reduced matrix is factored exactly: the joint S over every
// coordinate within the budgets (M = S: the restart is exact), else each large block S_bb. The
// factors stay with the absorber; after a weight refresh they remain the preconditioner (stale,
// so not exact) while their kept coordinates are unchanged, and the next escalation rebuilds them.
bool escalation_possible(const Absorber& a);
bool escalate(Absorber& a, const fh_ctx* ctx);  // false: no factor within the budgets (Jacobi stays)
void facs_after_refresh(Absorber& a);
// The exact LDL^T of the joint reduced matrix for the current weights on the host (assembled on the
// GPU, small coordinates last, the escalation's budgets); keep: the factored (kept) coordinates.
// False with the reason when a budget is exceeded.
struct LdlFactor;
bool factor_joint_host(Absorber& a, LdlFactor& f, std::vector<uint32_t>& keep, std::string& why);
// Leverage of the fixed effects per cell: hcell[c] = d'(D'WD)^+ d / w for the FE indicators d of
// cell c's rows (leverage.rs:1-30), as 1/n1_s + u_c' S^+ u_c with u_c the FE1-demeaned reduced
// coordinates of cell c, S^+ by selected inversion of the joint factor (fh_leverage.hip).
void fe_leverage_cells(Absorber& a, double* hcell_dev, const fh_ctx* ctx);
// HC2 / HC3 scores: e_out = e / sqrt(1 - h) (HC2) or e / (1 - h) (HC3) with the full model's
// leverage h = c (hcell + x~' B x~), c = w (row_weight) or 1, over the kept regressors (columns
// 1 + kept[j] of the rows, stride ld; B = bread over them). Throws when a leverage is 1
// (fit/vcov.rs:337-351: fixest's check at 1 - sqrt(eps)).
void hc_scores(Absorber& a, const double* rows, int ld, const std::vector<int>& kept, const std::vector<double>& bread,
               const double* e, size_t es, const double* w, bool row_weight, bool hc3, double* e_out);
bool facs_exact(const Absorber& a);  // a joint factor for the current weights: M = S
// The joint factor's estimated cost in seconds from what is known of it so far (a lower bound
// before it was first assembled), or INFINITY when the joint factor is not available.
double joint_build_estimate(Absorber& a);
// Builds the joint factor when every stage's estimated remaining cost stays within the margin of
// `without` (the projected PCG seconds): true when built (M = S for the restart).
bool escalate_joint(Absorber& a, const fh_ctx* ctx, double without);
// z = M^-1 r on the factored coordinates (FP64, m x r rows) and the partial r'z per factor in
// rzf_part (SPMV_DOTG x 16 entries, each factor its own share of them; zeroed by the caller).
void facs_apply(Absorber& a, const double* rr, double* z, int r, double* rzf_part, const uint32_t* done,
                hipStream_t st);
// The sum of the weights on the host (one readback after a refresh, none when a residualization
// has already brought it back).
double absorber_sw(Absorber& a);
void refresh_weights(Absorber& a, const double* w_rows_dev /* null = unit */);
void ensure_w32(Absorber& a);

// ---------------------------------------------------------------- operator (fh_operator.hip)
template <class T>
struct OpArgs {
    const uint32_t* cell_codes;
    const uint32_t* large_codes;  // cell_codes without block 0 (dense small block), or null
    const uint32_t* seg_off;
    const uint32_t* cell_obs_off;
    const T* cw;
    const T* inv_n1;
    uint64_t g1, M, m, ms;
    int nb, nsb;  // blocks, small blocks (the first nsb blocks)
    uint32_t boff[NB_MAX];
};

template <class T>
OpArgs<T> op_args(const Absorber& a);

// y = S p (m x r). t_buf >= g1*r, part >= spmv_part_len(a, r).
// pq (optional): SPMV_DOTG x 16 per-workgroup partials of p'y per column (fixed order).
constexpr int SPMV_DOTG = 256;
template <class T>
void spmv(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, const uint32_t* done, hipStream_t st,
          double* pq = nullptr);
// Scratch elements of one S*p (t, ds, dl, part, ypart of Scratch<T>) under a plan, and an S*p with
// the plan in use that never tunes (for callers that must not allocate).
struct SpmvNeed {
    size_t t = 0, ds = 0, dl = 0, part = 0, ypart = 0;
};
template <class T>
SpmvNeed spmv_scratch_need(const Absorber& a, const SpmvPlan& plan, int r);
template <class T>
void spmv_fixed(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, hipStream_t st);
// Level pass of block b under a plan: XCD buckets ((bucket, level) items, partial rows reduced in
// bucket order) while a bucket's share of t fits comfortably in an XCD's L2, else one item per
// level over its buckets (merge); and the lanes per item.
struct LvlMode {
    bool merge;
    int lanes;
};
// Default lanes per segment of the S*p segment pass: the absorber's (~4 cells per lane), halved
// for FP32 rows of more than 8 columns (12 columns at C2: 2 lanes 388 us, 4 lanes 412 us; FP64
// and narrower rows are fastest at the absorber's count).
inline uint32_t spmv_seg_lanes(const Absorber& a, int r, int tbytes) {
    return tbytes == 4 && r > 8 && a.seg_lanes > 1 ? a.seg_lanes / 2 : a.seg_lanes;
}

inline LvlMode lvl_mode(const Absorber& a, const SpmvPlan& plan, size_t b, int r, int tbytes) {
    const uint32_t nbk = a.nbk[b];
    const int x = b < plan.xcd.size() && plan.xcd[b] >= 0 ? plan.xcd[b]
                  : nbk < (uint32_t)XCD_COUNT                               ? 1
                  : (a.g1 / nbk + 1) * (size_t)r * (size_t)tbytes <= L2_TARGET ? 1
                                                                             : 0;
    LvlMode m;
    m.merge = nbk > 1 && x == 0;
    // unweighted FP64 rows of one or two columns: half the item lanes (C2: 196 us against 202 us
    // at 4 lanes); FP32, and weighted passes that also stream the level weights (Poisson panel:
    // 134 against 127 us), are fastest at the default
    const uint32_t def = m.merge ? a.llanes[b]
                         : tbytes == 8 && r <= 2 && !a.weighted && a.blanes[b] > 1 ? a.blanes[b] / 2
                                                                                   : a.blanes[b];
    m.lanes = (int)(b < plan.blanes.size() && plan.blanes[b] ? plan.blanes[b] : def);
    return m;
}

// Process-wide cache of tuned S*p plans keyed by the structure signature (plan_key).
std::string plan_key(const Absorber& a, int tbytes, int rclass);
bool tune_cache_get(const std::string& key, SpmvPlan& plan);
void tune_cache_put(const std::string& key, const SpmvPlan& plan);
void tune_cache_clear();

// rhs pass over rows (n x r, FP64): rhs (m x r, T), b1 (g1 x r), wsq/wsum partials -> out2 (2r).
// Leaves the per-workgroup partials of wsq / wsum (grid x 2r, returns grid) at the start of
// a.s64.dpart, which holds room for `extra` more doubles after them.
// rows: n x r; rhs (m x rp) and b1 (g1 x rp) with zero pad columns r..rp-1.
unsigned rhs_pass(Absorber& a, const double* rows, int r, int rp, double* rhs, double* b1, size_t extra);
// rows (n x r) -= alpha_s + u_c, with x (m x rp) and b1, alpha (g1 x rp; alpha optional).
void residualize_rows(Absorber& a, double* rows, int r, int rp, const double* x, const double* b1, double* alpha);

// Fixed-order device reduction of per-workgroup partials part[q * len + e] (q < nparts) read
// back as len values (fh_estimate.hip).
std::vector<double> sum_parts_dev(const double* part, size_t nparts, size_t len, hipStream_t st);
void sum_parts_to(const double* part, size_t nparts, size_t len, double* out_dev, hipStream_t st);

// Column count the solver's m x r vectors are padded to: the register classes of the kernels,
// so every row of p, t and x is whole 16-byte vectors (r <= 2 rows are 8 or 16 bytes already).
// The pad columns are zero and inactive.
inline int padded_cols(int r) { return r <= 2 ? r : r <= 4 ? 4 : r <= 8 ? 8 : r <= 12 ? 12 : 16; }

// ---------------------------------------------------------------- varying slopes (fh_slopes.hip)
// One absorbed term of a slope absorber: device codes and slope columns, kept order.
struct SlopeTerm {
    const uint32_t* codes = nullptr;
    uint64_t g = 0;
    bool icpt = true;
    std::vector<const double*> slopes;
};
constexpr int SL_BMAX = 8;  // widest term: intercept + slopes
// Builds the observation-level structure (slopes.rs:354-527) and the weight state for kept-order
// weights w_kept (device, null = unit).
void slope_create(Absorber& a, const std::vector<SlopeTerm>& terms, const double* w_kept);
// Weight state for row-order weights (device, null = unit; zeros allowed: rows left out).
void slope_refresh(Absorber& a, const double* w_rows);
// y = S p (m x r, FP64) with optional p'y partials in the PCG's slot layout (SPMV_DOTG x 16).
void slope_spmv(const Absorber& a, const double* p, double* y, int r, const uint32_t* done, hipStream_t st,
                double* pq);
// z = M^-1 r on the block-Jacobi coordinates [ms, m) and their r'z partials (SPMV_DOTG x 16).
void slope_precond(Absorber& a, const double* rr, double* z, int r, double* rzf_part, const uint32_t* done,
                   hipStream_t st);
struct ResidOut;
// One residualization of the n x r device block z (row order, r <= 16), solve.rs:1141-1311 with the
// slope operator; warm / coefs (m + g1) x r.
ResidOut slope_residualize(Absorber& a, double* z, int r, double tol, uint64_t maxiter, const double* warm,
                           double* coefs, const fh_ctx* ctx);
// Fit of column col of coefs ((m + g1) x r, device) per row (row order).
void slope_fit_rows(Absorber& a, const double* coefs, int r, int col, double* out);
// term_values (slopes.rs:2435-2464): one vector per slot (each term's intercept, then its slopes),
// from host coefs ((m + g1) x r).
std::vector<std::vector<double>> slope_term_values(Absorber& a, const double* coefs, size_t r, size_t col);
// Block layout (term, levels, first coordinate, small flag) of a slope absorber.
void slope_blocks(const Absorber& a, int64_t* term, uint64_t* g, uint64_t* off, int32_t* small);
// The E term, its levels and width, and the small coordinates.
void slope_summary(const Absorber& a, int64_t* e, uint64_t* ge, uint64_t* be, uint64_t* m_small, uint64_t* nb);
// The grounded explicit inverse of a dense symmetric block (precond/dense.rs), redundant
// coordinates left out; nkeep receives the coordinates kept (fh_weights.hip).
std::vector<double> grounded_inverse(const std::vector<double>& a, size_t ms, const std::vector<uint8_t>& red,
                                     uint64_t& nkeep);

// ---------------------------------------------------------------- PCG (fh_pcg.hip)
struct PcgResult {
    uint64_t iterations = 0;
    bool exact = false;
    int rc = FH_OK;
    uint64_t refinements = 0;
    bool escalate = false;       // stopped at the escalation probe (x holds the progress so far)
};
// The early probe's joint-factor rule (solve.rs:932-975): a solve that may escalate stops after its
// iteration-6 probe when the joint factor's estimated cost (joint_build_estimate, seconds) is
// within the margin of the projected remaining PCG time (fired; x holds the progress so far).
struct JointProbe {
    bool fired = false;
    double without = 0.0, rho = 0.0, remaining = 0.0;
};
// Solves S x = rhs for r columns; x holds the start (x_maybe_nonzero false: x is zero, which
// spares the device read that tests it). FP64 throughout.
PcgResult pcg_fp64(Absorber& a, const double* rhs, double* x, bool x_maybe_nonzero, int r, const double* ss1,
                   const double* ss0, double tol, uint64_t maxiter, const fh_ctx* ctx, bool can_escalate = false,
                   JointProbe* jp = nullptr);
// Mixed form: FP32 inner PCG, FP64 outer residual and update.
PcgResult pcg_mixed(Absorber& a, const double* rhs, double* x, bool x_maybe_nonzero, int r, const double* ss1,
                    const double* ss0, double tol, uint64_t maxiter, const fh_ctx* ctx, bool can_escalate = false,
                    JointProbe* jp = nullptr);

// ---------------------------------------------------------------- estimation (fh_estimate.hip)
// G = sum_i w_i z_i z_i' (r x r, FP64 on the host), z n x r row-major; w may be null.
void gram(Absorber& a, const double* z, const double* w, size_t n, int r, double* G_host);
// Columns c0..c0+k of z (ld), optional row weights w and v: G = sum_i w_i v_i z_i z_i' (host).
// ztol >= 0: entries of magnitude <= ztol are set to 0 in z as they are read (Mata
// _edittozerotol), and the Gram is of the edited block.
void gram_cols(Absorber& a, const double* z, size_t n, int ld, int c0, int k, const double* w, const double* v,
               double* G_host, double ztol = -1.0);
// Clusters as runs of adjacent solver rows with one cluster code: the rows of cluster g are the
// runs order[off[g] .. off[g+1]) (in row order), run q spanning rows run_off[q] .. run_off[q+1].
// Rows in solver order follow the absorbed terms, so a cluster that is (or nests) a term comes in
// runs of many rows and the runs, not the rows, are sorted.
struct ClusterRuns {
    const uint32_t* cl = nullptr;  // codes described (kept order)
    uint64_t G = 0;
    size_t nrun = 0;
    DBuf<uint32_t> run_off;  // nrun + 1
    DBuf<uint32_t> order;    // nrun, stable by cluster
    DBuf<uint32_t> off;      // G + 1
};
void cluster_runs(Absorber& a, const uint32_t* cl_kept_dev, uint64_t G, ClusterRuns& cr);
// The runs of these codes: the ones set on the absorber for the current call when they match,
// else built into `local`.
const ClusterRuns& runs_for(Absorber& a, const uint32_t* cl_kept_dev, uint64_t G, ClusterRuns& local);
// L lanes per cluster for ~2 runs per lane.
uint32_t lanes_per_cluster(const ClusterRuns& cr);

// ---------------------------------------------------------------- wild bootstrap (fh_wboot.hip)
// WCR11 on the residualized OLS block (rows n x K+1 with stride R: column 0 = y~, 1..K = X~), the kept
// regressors `kept`, bread (mk x mk), residuals e (n), clusters cl (kept order, G levels).
struct WbootInput {
    const double* rows;
    int R;
    std::vector<int> kept;
    std::vector<double> bread, beta;  // kept columns
    const double* e;  // e_i at e[i * es]
    size_t es = 1;
    const double* w;
    const uint32_t* cl;
    uint64_t G;
    double ssc;
    const double* S = nullptr;  // per-cluster sums already formed by the fit (G x 3 x kept), or null
};
// ConstraintProj for one restriction rk over the kept regressors: C = bread rk, M_R = C / (rk'C).
void wb_constraint(const std::vector<double>& bread, int mk, const double* rk, std::vector<double>& C,
                   std::vector<double>& MR);
// S[g][0..2][a] = sums over the rows of cluster g of (w e, w u1, w z) x~_a over the kept
// regressors (z = x~'C, u1 = -x~'M_R), on the call's cluster runs; S[g][0] is the CRV1 score.
void wb_cluster_sums(Absorber& a, const double* rows, int R, const std::vector<int>& kept, const std::vector<double>& C,
                     const std::vector<double>& MR, const double* e, size_t es, const double* w, const uint32_t* cl,
                     uint64_t G, DBuf<double>& S);
// Buffers of the bootstrap's non-nested projection (kept alive for the contraction).
struct WbExt {
    DBuf<double> rows, S;
};
// Fixed effects not nested in the clusters (m_nn levels): extends in (and rk) to the nested form
// with the projection's pseudo-regressors appended.
void wb_path1(Absorber& a, const std::vector<int>& nested, uint64_t m_nn, WbootInput& in, std::vector<double>& rk,
              WbExt& ext);
// Non-empty clusters of the call's runs.
uint64_t count_nonempty_clusters(Absorber& a, const uint32_t* cl, uint64_t G);
void ols_wildboot(Absorber& a, const WbootInput& in, const fh_wboot_spec& spec, const double* r_mat_kept,
                  fh_wboot_result& res, double* t_boot_host);
// WCR11 with the non-nested terms (m_nn <= 512 levels) projected out through Path 1's algebra with
// the draws as two FP64 matrix products (any number of clusters).
void ols_wildboot_wide(Absorber& a, const std::vector<int>& nested, uint64_t m_nn, const WbootInput& in,
                       const fh_wboot_spec& spec, const double* r_mat_kept, fh_wboot_result& res,
                       double* t_boot_host);
// WCR11 with every fixed effect projected out of the draws one cluster at a time (2 .. 512
// clusters): one residualization (tolerance min(tol, 1e-10), at most maxiter iterations) per
// cluster, then the same contraction over G x (2 + 2G) per-cluster columns.
void ols_wildboot_cells(Absorber& a, const WbootInput& in, const fh_wboot_spec& spec, const double* r_mat_kept,
                        double tol, uint64_t maxiter, fh_wboot_result& res, double* t_boot_host);
// MNW 13 / 31 / 33 (wildboot/ols.rs:651-1082) on the same input: Q (n x r_nn, row order, or null)
// the orthonormal basis of the non-nested terms' dummies residualized on the nested ones.
void ols_wildboot_mnw(Absorber& a, const WbootInput& in, const fh_wboot_spec& spec, int boot_type,
                      const double* r_mat_kept, const double* Q, int r_nn, fh_wboot_result& res, double* t_boot_host);
// That basis for the non-nested terms (m_nn levels): Q'WQ = I over the eigenvalues of the
// residualized dummies' Gram above 1e-10 of the largest (r_nn of them).
void mnw_basis(Absorber& a, const std::vector<int>& nested, uint64_t m_nn, const WbootInput& in, DBuf<double>& Q,
               int& r_nn);
// Per absorbed term: 1 when each of its levels lies in a single cluster (codes in kept order).
std::vector<int> terms_nested_in(Absorber& a, const uint32_t* cl);

}  // namespace fh
