/* CPU oracle of the fixed-effect demeaning path: see fereg_oracle.h.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product path). Plain C, FP64, one thread,
 * written from the definitions in the reference rather than its parallel task layout:
 *   structure   crates/fereg-core/src/structure.rs:247-527 (choice of FE1, block order, sort
 *               key, cells) and its brute-force definition `check` at structure.rs:1093-1171;
 *   weights     operator.rs:211-338 (cell weights, n1, sw, diag, redundant) and :344-420
 *               (n_r and sum_s c_{s,k}^2/n1_s);
 *   operator    operator.rs:894-948 (S*P), :969-1066 (rhs), :1107-1179 (back-substitution);
 *   small block operator.rs:1457-1560 (dense S_SS), precond/dense.rs:25-37,79-160 (grounded LDL');
 *   Jacobi      precond/jacobi.rs:7-41; block preconditioner precond/mod.rs:74-100;
 *   PCG         cg.rs:132-292 (lockstep columns, windowed stopping rule, exact one-shot path);
 *   residualize solve.rs:1141-1311 (ss0, ss1, warm start, FeCoefs = [x; alpha1]).
 * Not restated (the checks pin results at solution level instead): the cost model, probes,
 * escalation to sparse or approximate Cholesky factors, and the thread-chunk summation order.
 */
#include "fereg_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define OR_REDUNDANT 1e-12  /* operator.rs:21 */
#define OR_PIVOT_TOL 1e-10  /* precond/dense.rs:26 */
#define OR_GROUND 1e30      /* precond/dense.rs:28 */
#define OR_WINDOW 5         /* cg.rs:23 */

struct or_solver {
    uint64_t n, nterms, g1, M, nb, m, ms;
    int64_t fe1, norm_fe1;
    uint64_t small_work;
    uint32_t* pi;
    uint32_t* seg_off;
    uint32_t* cell_obs_off; /* NULL when trivial */
    uint32_t* cell_codes;
    int64_t* bterm;
    uint64_t* bg;
    uint64_t* boff;
    int32_t* bsmall;
    uint32_t** lperm;
    uint32_t** loff;
    /* weight state */
    double* w_rows; /* NULL = unit */
    double* cw;     /* cell weights (always materialized, 1.0 for unit) */
    double* n1;
    double sw;
    double* nr;
    double* diag;
    uint8_t* red;
    /* dense grounded factor of S_SS over the kept small coordinates */
    uint64_t nkeep;
    uint32_t* keep;
    double* scale;
    double* L; /* nkeep x nkeep, unit lower + D on the diagonal */
};

/* ---------------------------------------------------------------- structure */

/* fe1_index: the intercept-only term with most levels, first on ties (absorber.rs:99-108). */
static int64_t fe1_index(const uint64_t* g, uint64_t nt) {
    int64_t best = -1;
    for (uint64_t q = 0; q < nt; q++)
        if (best < 0 || g[q] > g[best]) best = (int64_t)q;
    return best;
}

/* E of eliminating term f: sum over other terms o and levels s of f of d_{s,o}^2, d the
 * distinct levels of o among the observations of level s (structure.rs:325-354). */
static uint64_t distinct_work(const uint32_t* const* codes, const uint64_t* g, uint64_t nt, uint64_t n,
                              uint64_t f) {
    /* observations sorted by f: counting sort */
    uint64_t gf = g[f];
    uint32_t* off = calloc(gf + 1, sizeof(uint32_t));
    uint32_t* ids = malloc((n ? n : 1) * sizeof(uint32_t));
    for (uint64_t i = 0; i < n; i++) off[codes[f][i] + 1]++;
    for (uint64_t k = 0; k < gf; k++) off[k + 1] += off[k];
    uint32_t* cur = malloc((gf + 1) * sizeof(uint32_t));
    memcpy(cur, off, (gf + 1) * sizeof(uint32_t));
    for (uint64_t i = 0; i < n; i++) ids[cur[codes[f][i]]++] = (uint32_t)i;
    uint64_t e = 0;
    for (uint64_t o = 0; o < nt; o++) {
        if (o == f) continue;
        uint32_t* stamp = calloc(g[o] ? g[o] : 1, sizeof(uint32_t));
        for (uint64_t s = 0; s < gf; s++) {
            uint64_t d = 0;
            uint32_t tag = (uint32_t)s + 1;
            for (uint32_t k = off[s]; k < off[s + 1]; k++) {
                uint32_t c = codes[o][ids[k]];
                if (stamp[c] != tag) { stamp[c] = tag; d++; }
            }
            e += d * d;
        }
        free(stamp);
    }
    free(off); free(ids); free(cur);
    return e;
}

/* choose_elimination (structure.rs:256-313). */
static int64_t choose_elimination(const uint32_t* const* codes, const uint64_t* g, uint64_t nt, uint64_t n) {
    int64_t f0 = fe1_index(g, nt);
    if (f0 < 0 || nt < 3) return f0;
    uint64_t ncand = 0;
    for (uint64_t q = 0; q < nt; q++)
        if (g[q] * 4 >= g[f0]) ncand++;
    if (ncand < 2) return f0;
    uint64_t* size = calloc(g[f0] ? g[f0] : 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) size[codes[f0][i]]++;
    uint64_t ssq = 0;
    for (uint64_t s = 0; s < g[f0]; s++) ssq += size[s] * size[s];
    free(size);
    if ((nt - 1) * ssq <= 64 * n) return f0;
    int64_t best = -1;
    uint64_t bestE = 0;
    for (uint64_t f = 0; f < nt; f++) {
        if (!(g[f] * 4 >= g[f0])) continue;
        uint64_t e = distinct_work(codes, g, nt, n, f);
        /* key (E, Reverse(g), f): smaller E, then more levels, then lower index */
        if (best < 0 || e < bestE || (e == bestE && g[f] > g[best])) { best = (int64_t)f; bestE = e; }
    }
    return best;
}

/* sort context for the row order (structure.rs:1099-1102: sort by (tuple, obs)) */
static const uint32_t* const* g_sort_cols;
static uint64_t g_sort_ncols;
static int cmp_rows(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    for (uint64_t k = 0; k < g_sort_ncols; k++) {
        uint32_t cx = g_sort_cols[k][x], cy = g_sort_cols[k][y];
        if (cx != cy) return cx < cy ? -1 : 1;
    }
    return x < y ? -1 : (x > y);
}

static void weight_refresh(or_solver* s);

int or_build(const uint32_t* const* codes, const uint64_t* g, uint64_t nt, uint64_t n, uint64_t small_limit,
             const double* w_kept, or_solver** out) {
    *out = NULL;
    for (uint64_t q = 0; q < nt; q++)
        for (uint64_t i = 0; i < n; i++)
            if (codes[q][i] >= g[q]) return 198;
    if (w_kept)
        for (uint64_t i = 0; i < n; i++)
            if (!(w_kept[i] > 0.0) || !isfinite(w_kept[i])) return 198;
    or_solver* s = calloc(1, sizeof(or_solver));
    s->n = n;
    s->nterms = nt;
    s->fe1 = choose_elimination(codes, g, nt, n);
    s->norm_fe1 = fe1_index(g, nt);
    s->g1 = s->fe1 >= 0 ? g[s->fe1] : 1;

    /* small/large classification: ascending (G, q) while the running total fits
     * (structure.rs:366-385) */
    uint64_t nb = nt > 0 ? (s->fe1 >= 0 ? nt - 1 : nt) : 0;
    s->nb = nb;
    int64_t* others = malloc((nb + 1) * sizeof(int64_t));
    uint64_t no = 0;
    for (uint64_t q = 0; q < nt; q++)
        if ((int64_t)q != s->fe1) others[no++] = (int64_t)q;
    int64_t* by_g = malloc((nb + 1) * sizeof(int64_t));
    memcpy(by_g, others, nb * sizeof(int64_t));
    for (uint64_t i = 1; i < nb; i++) { /* insertion sort by (g, q) */
        int64_t v = by_g[i];
        uint64_t j = i;
        while (j > 0 && (g[by_g[j - 1]] > g[v] || (g[by_g[j - 1]] == g[v] && by_g[j - 1] > v))) { by_g[j] = by_g[j - 1]; j--; }
        by_g[j] = v;
    }
    int* small_set = calloc(nt + 1, sizeof(int));
    uint64_t total = 0;
    for (uint64_t i = 0; i < nb; i++) {
        int64_t q = by_g[i];
        if (total + g[q] > small_limit) break;
        total += g[q];
        small_set[q] = 1;
    }
    s->bterm = malloc((nb + 1) * sizeof(int64_t));
    s->bg = malloc((nb + 1) * sizeof(uint64_t));
    s->boff = malloc((nb + 1) * sizeof(uint64_t));
    s->bsmall = malloc((nb + 1) * sizeof(int32_t));
    uint64_t nbk = 0, off = 0;
    for (int pass = 0; pass < 2; pass++)
        for (uint64_t i = 0; i < nb; i++) {
            int64_t q = others[i];
            if ((pass == 0) != (small_set[q] != 0)) continue;
            s->bterm[nbk] = q; s->bg[nbk] = g[q]; s->boff[nbk] = off; s->bsmall[nbk] = small_set[q];
            off += g[q];
            nbk++;
        }
    s->m = off;
    s->ms = 0;
    for (uint64_t b = 0; b < nb; b++)
        if (s->bsmall[b]) s->ms += s->bg[b];

    /* rows sorted by (FE1 code, block codes in block order, obs) */
    const uint32_t** cols = malloc((nb + 2) * sizeof(uint32_t*));
    uint64_t nc = 0;
    uint32_t* zeros = NULL;
    if (s->fe1 >= 0) cols[nc++] = codes[s->fe1];
    else { zeros = calloc(n ? n : 1, sizeof(uint32_t)); cols[nc++] = zeros; }
    for (uint64_t b = 0; b < nb; b++) cols[nc++] = codes[s->bterm[b]];
    s->pi = malloc((n ? n : 1) * sizeof(uint32_t));
    g_sort_cols = cols; g_sort_ncols = nc;
    {
        /* The sort by (FE1 code, tuple, obs) as a stable counting sort by the FE1 code (the
         * primary key; obs ascending within each code) followed by a sort of each FE1 group by
         * (tuple, obs): the same order as one sort, the groups sorted in parallel. */
        uint64_t g1c = s->fe1 >= 0 ? g[s->fe1] : 1;
        uint64_t* goff = calloc(g1c + 1, sizeof(uint64_t));
        for (uint64_t i = 0; i < n; i++) goff[cols[0][i] + 1]++;
        for (uint64_t k = 0; k < g1c; k++) goff[k + 1] += goff[k];
        uint64_t* cur = malloc((g1c + 1) * sizeof(uint64_t));
        memcpy(cur, goff, (g1c + 1) * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; i++) s->pi[cur[cols[0][i]]++] = (uint32_t)i;
        free(cur);
#pragma omp parallel for schedule(dynamic, 4096)
        for (uint64_t k = 0; k < g1c; k++)
            if (goff[k + 1] - goff[k] > 1) qsort(s->pi + goff[k], goff[k + 1] - goff[k], sizeof(uint32_t), cmp_rows);
        free(goff);
    }

    /* cells = runs of equal tuples; seg_off by FE1 level (structure.rs:1103-1120) */
    uint32_t* starts = malloc((n + 1) * sizeof(uint32_t));
    s->cell_codes = malloc((n * nb + 1) * sizeof(uint32_t));
    s->seg_off = malloc((s->g1 + 1) * sizeof(uint32_t));
    uint64_t M = 0, nseg = 0;
    for (uint64_t row = 0; row < n; row++) {
        uint32_t o = s->pi[row];
        int newcell = row == 0;
        if (!newcell) {
            uint32_t p = s->pi[row - 1];
            for (uint64_t k = 0; k < nc; k++)
                if (cols[k][p] != cols[k][o]) { newcell = 1; break; }
        }
        if (newcell) {
            uint32_t c1 = cols[0][o];
            while (nseg <= c1) s->seg_off[nseg++] = (uint32_t)M;
            starts[M] = (uint32_t)row;
            for (uint64_t b = 0; b < nb; b++) s->cell_codes[M * nb + b] = cols[1 + b][o];
            M++;
        }
    }
    while (nseg <= s->g1) s->seg_off[nseg++] = (uint32_t)M;
    s->M = M;
    starts[M] = (uint32_t)n;
    if (M != n) {
        s->cell_obs_off = malloc((M + 1) * sizeof(uint32_t));
        memcpy(s->cell_obs_off, starts, (M + 1) * sizeof(uint32_t));
    }
    free(starts);

    /* small_work = sum_seg (distinct small coordinates)^2 (structure.rs:1122-1146) */
    s->small_work = 0;
    if (s->ms > 0) {
        uint32_t* stamp = calloc(s->ms, sizeof(uint32_t));
        for (uint64_t sg = 0; sg < s->g1; sg++) {
            uint64_t d = 0;
            for (uint32_t c = s->seg_off[sg]; c < s->seg_off[sg + 1]; c++)
                for (uint64_t b = 0; b < nb; b++) {
                    if (!s->bsmall[b]) continue;
                    uint64_t k = s->boff[b] + s->cell_codes[(uint64_t)c * nb + b];
                    if (stamp[k] != (uint32_t)sg + 1) { stamp[k] = (uint32_t)sg + 1; d++; }
                }
            s->small_work += d * d;
        }
        free(stamp);
        /* demotion (structure.rs:521-525) */
        if (s->small_work > 64 * M) {
            for (uint64_t b = 0; b < nb; b++) s->bsmall[b] = 0;
            s->ms = 0;
        }
    }

    /* level-sorted cell lists of every block (structure.rs:144-153, 1152-1167) */
    s->lperm = calloc(nb + 1, sizeof(uint32_t*));
    s->loff = calloc(nb + 1, sizeof(uint32_t*));
    for (uint64_t b = 0; b < nb; b++) {
        /* cells by (level, cell): a stable counting sort by the level */
        uint32_t* perm = malloc((M ? M : 1) * sizeof(uint32_t));
        uint32_t* lo = calloc(s->bg[b] + 1, sizeof(uint32_t));
        for (uint64_t c = 0; c < M; c++) lo[s->cell_codes[c * nb + b] + 1]++;
        for (uint64_t k = 0; k < s->bg[b]; k++) lo[k + 1] += lo[k];
        uint32_t* cur = malloc((s->bg[b] + 1) * sizeof(uint32_t));
        memcpy(cur, lo, (s->bg[b] + 1) * sizeof(uint32_t));
        for (uint64_t c = 0; c < M; c++) perm[cur[s->cell_codes[c * nb + b]]++] = (uint32_t)c;
        free(cur);
        s->lperm[b] = perm;
        s->loff[b] = lo;
    }
    free(others); free(by_g); free(small_set); free(cols); free(zeros);

    /* weights in row order (solve.rs:408-429) */
    if (w_kept) {
        s->w_rows = malloc((n ? n : 1) * sizeof(double));
        for (uint64_t i = 0; i < n; i++) s->w_rows[i] = w_kept[s->pi[i]];
    }
    s->cw = malloc((M ? M : 1) * sizeof(double));
    s->n1 = malloc(s->g1 * sizeof(double));
    s->nr = malloc((s->m ? s->m : 1) * sizeof(double));
    s->diag = malloc((s->m ? s->m : 1) * sizeof(double));
    s->red = malloc((s->m ? s->m : 1));
    weight_refresh(s);
    *out = s;
    return 0;
}

void or_free(or_solver* s) {
    if (!s) return;
    for (uint64_t b = 0; b < s->nb; b++) { free(s->lperm[b]); free(s->loff[b]); }
    free(s->lperm); free(s->loff);
    free(s->pi); free(s->seg_off); free(s->cell_obs_off); free(s->cell_codes);
    free(s->bterm); free(s->bg); free(s->boff); free(s->bsmall);
    free(s->w_rows); free(s->cw); free(s->n1); free(s->nr); free(s->diag); free(s->red);
    free(s->keep); free(s->scale); free(s->L);
    free(s);
}

void or_get_info(const or_solver* s, or_info* o) {
    o->n = s->n; o->nterms = s->nterms; o->g1 = s->g1; o->ncells = s->M; o->nb = s->nb;
    o->m = s->m; o->m_small = s->ms; o->small_work = s->small_work; o->fe1 = s->fe1;
    o->norm_fe1 = s->norm_fe1; o->has_cell_obs_off = s->cell_obs_off != NULL;
}

void or_get_blocks(const or_solver* s, int64_t* term, uint64_t* g, uint64_t* off, int32_t* small) {
    for (uint64_t b = 0; b < s->nb; b++) { term[b] = s->bterm[b]; g[b] = s->bg[b]; off[b] = s->boff[b]; small[b] = s->bsmall[b]; }
}

void or_get_structure(const or_solver* s, uint32_t* pi, uint32_t* seg_off, uint32_t* cc, uint32_t* coo) {
    if (pi) memcpy(pi, s->pi, s->n * sizeof(uint32_t));
    if (seg_off) memcpy(seg_off, s->seg_off, (s->g1 + 1) * sizeof(uint32_t));
    if (cc) memcpy(cc, s->cell_codes, s->M * s->nb * sizeof(uint32_t));
    if (coo && s->cell_obs_off) memcpy(coo, s->cell_obs_off, (s->M + 1) * sizeof(uint32_t));
}

void or_get_level_cells(const or_solver* s, uint64_t b, uint32_t* perm, uint32_t* off) {
    memcpy(perm, s->lperm[b], s->M * sizeof(uint32_t));
    memcpy(off, s->loff[b], (s->bg[b] + 1) * sizeof(uint32_t));
}

void or_get_weight_state(const or_solver* s, double* n1, double* nr, double* diag, uint8_t* red, double* sw) {
    if (n1) memcpy(n1, s->n1, s->g1 * sizeof(double));
    if (nr) memcpy(nr, s->nr, s->m * sizeof(double));
    if (diag) memcpy(diag, s->diag, s->m * sizeof(double));
    if (red) memcpy(red, s->red, s->m);
    if (sw) *sw = s->sw;
}

/* ---------------------------------------------------------------- weights */

static inline uint64_t cell_row0(const or_solver* s, uint64_t c) { return s->cell_obs_off ? s->cell_obs_off[c] : c; }
static inline uint64_t cell_row1(const or_solver* s, uint64_t c) { return s->cell_obs_off ? s->cell_obs_off[c + 1] : c + 1; }

static void small_factor(or_solver* s);

/* refresh_weights / level_weights (operator.rs:211-338, 344-420). */
static void weight_refresh(or_solver* s) {
    for (uint64_t c = 0; c < s->M; c++) {
        double v = 0.0;
        for (uint64_t i = cell_row0(s, c); i < cell_row1(s, c); i++) v += s->w_rows ? s->w_rows[i] : 1.0;
        s->cw[c] = v;
    }
    s->sw = 0.0;
    for (uint64_t sg = 0; sg < s->g1; sg++) {
        double v = 0.0;
        for (uint32_t c = s->seg_off[sg]; c < s->seg_off[sg + 1]; c++) v += s->cw[c];
        s->n1[sg] = v;
        s->sw += v;
    }
    uint64_t m = s->m, nb = s->nb;
    double* dq = calloc(m ? m : 1, sizeof(double));
    double* cs = calloc(m ? m : 1, sizeof(double));
    for (uint64_t k = 0; k < m; k++) s->nr[k] = 0.0;
    for (uint64_t sg = 0; sg < s->g1; sg++) {
        uint32_t c0 = s->seg_off[sg], c1 = s->seg_off[sg + 1];
        for (uint32_t c = c0; c < c1; c++)
            for (uint64_t b = 0; b < nb; b++) cs[s->boff[b] + s->cell_codes[(uint64_t)c * nb + b]] += s->cw[c];
        for (uint32_t c = c0; c < c1; c++)
            for (uint64_t b = 0; b < nb; b++) {
                uint64_t k = s->boff[b] + s->cell_codes[(uint64_t)c * nb + b];
                s->nr[k] += s->cw[c];
                dq[k] += s->cw[c] * cs[k] / s->n1[sg];
            }
        for (uint32_t c = c0; c < c1; c++)
            for (uint64_t b = 0; b < nb; b++) cs[s->boff[b] + s->cell_codes[(uint64_t)c * nb + b]] = 0.0;
    }
    for (uint64_t k = 0; k < m; k++) {
        s->diag[k] = s->nr[k] - dq[k];
        s->red[k] = s->diag[k] <= OR_REDUNDANT * s->nr[k];
    }
    free(dq); free(cs);
    small_factor(s);
}

int or_set_weights(or_solver* s, const double* w_rows) {
    if (w_rows) {
        for (uint64_t i = 0; i < s->n; i++)
            if (!(w_rows[i] > 0.0) || !isfinite(w_rows[i])) return 198;
        if (!s->w_rows) s->w_rows = malloc((s->n ? s->n : 1) * sizeof(double));
        memcpy(s->w_rows, w_rows, s->n * sizeof(double));
    } else {
        free(s->w_rows);
        s->w_rows = NULL;
    }
    weight_refresh(s);
    return 0;
}

/* ---------------------------------------------------------------- small dense factor */

/* Dense S_SS (operator.rs:1457-1560): A_SS minus sum_s c_s c_s' / n1_s over small coordinates. */
static double* assemble_small(const or_solver* s) {
    uint64_t ms = s->ms, nb = s->nb;
    double* a = calloc(ms * ms, sizeof(double));
    double* cs = calloc(ms, sizeof(double));
    uint64_t* touched = malloc(ms * sizeof(uint64_t));
    for (uint64_t sg = 0; sg < s->g1; sg++) {
        uint64_t nt = 0;
        for (uint32_t c = s->seg_off[sg]; c < s->seg_off[sg + 1]; c++) {
            double w = s->cw[c];
            for (uint64_t b = 0; b < nb; b++) {
                if (!s->bsmall[b]) continue;
                uint64_t k = s->boff[b] + s->cell_codes[(uint64_t)c * nb + b];
                if (cs[k] == 0.0) touched[nt++] = k;
                cs[k] += w;
                for (uint64_t b2 = 0; b2 < nb; b2++) {
                    if (!s->bsmall[b2]) continue;
                    uint64_t l = s->boff[b2] + s->cell_codes[(uint64_t)c * nb + b2];
                    a[k * ms + l] += w;
                }
            }
        }
        for (uint64_t i = 0; i < nt; i++)
            for (uint64_t j = 0; j < nt; j++) a[touched[i] * ms + touched[j]] -= cs[touched[i]] * cs[touched[j]] / s->n1[sg];
        for (uint64_t i = 0; i < nt; i++) cs[touched[i]] = 0.0;
    }
    free(cs); free(touched);
    return a;
}

/* Dense::factor (precond/dense.rs:79-115): unit-diagonal scaling and LDL' in which a pivot
 * below PIVOT_TOL is replaced by GROUND (dense.rs:25-37). */
static void small_factor(or_solver* s) {
    free(s->keep); free(s->scale); free(s->L);
    s->keep = NULL; s->scale = NULL; s->L = NULL; s->nkeep = 0;
    uint64_t ms = s->ms;
    if (ms == 0) return;
    double* a = assemble_small(s);
    s->keep = malloc(ms * sizeof(uint32_t));
    uint64_t k = 0;
    for (uint64_t i = 0; i < ms; i++)
        if (!s->red[i] && a[i * ms + i] > 0.0) s->keep[k++] = (uint32_t)i;
    s->nkeep = k;
    s->scale = malloc((k ? k : 1) * sizeof(double));
    s->L = calloc(k * k + 1, sizeof(double));
    for (uint64_t i = 0; i < k; i++) s->scale[i] = 1.0 / sqrt(a[(uint64_t)s->keep[i] * (ms + 1)]);
    double* L = s->L;
    for (uint64_t i = 0; i < k; i++)
        for (uint64_t j = 0; j <= i; j++) L[i * k + j] = a[(uint64_t)s->keep[i] * ms + s->keep[j]] * s->scale[i] * s->scale[j];
    for (uint64_t j = 0; j < k; j++) {
        double d = L[j * k + j];
        for (uint64_t p = 0; p < j; p++) d -= L[j * k + p] * L[j * k + p] * L[p * k + p];
        if (d <= OR_PIVOT_TOL) d = OR_GROUND;
        L[j * k + j] = d;
        for (uint64_t i = j + 1; i < k; i++) {
            double v = L[i * k + j];
            for (uint64_t p = 0; p < j; p++) v -= L[i * k + p] * L[j * k + p] * L[p * k + p];
            L[i * k + j] = v / d;
        }
    }
    free(a);
}

/* Dense::apply (precond/dense.rs:117-160): z_S = S_SS^+ r_S, redundant rows 0. */
void or_small_apply(const or_solver* s, const double* rin, uint64_t r, double* z) {
    uint64_t k = s->nkeep;
    for (uint64_t i = 0; i < s->ms * r; i++) z[i] = 0.0;
    if (k == 0) return;
    double* t = malloc(k * sizeof(double));
    const double* L = s->L;
    for (uint64_t j = 0; j < r; j++) {
        for (uint64_t i = 0; i < k; i++) t[i] = rin[(uint64_t)s->keep[i] * r + j] * s->scale[i];
        for (uint64_t i = 0; i < k; i++)
            for (uint64_t p = 0; p < i; p++) t[i] -= L[i * k + p] * t[p];
        for (uint64_t i = 0; i < k; i++) t[i] /= L[i * k + i];
        for (uint64_t ii = k; ii-- > 0;)
            for (uint64_t p = ii + 1; p < k; p++) t[ii] -= L[p * k + ii] * t[p];
        for (uint64_t i = 0; i < k; i++) z[(uint64_t)s->keep[i] * r + j] = t[i] * s->scale[i];
    }
    free(t);
}

/* Precond::apply (precond/mod.rs:74-100) with Jacobi on the large blocks (jacobi.rs:7-41). */
static void precond_apply(const or_solver* s, const double* rin, uint64_t r, double* z) {
    for (uint64_t k = s->ms; k < s->m; k++) {
        double inv = (s->red[k] || !(s->diag[k] > 0.0)) ? 0.0 : 1.0 / s->diag[k];
        for (uint64_t j = 0; j < r; j++) z[k * r + j] = inv * rin[k * r + j];
    }
    if (s->ms > 0) or_small_apply(s, rin, r, z);
}

/* out[e] = sum over threads t = 0..nt-1 (in order) of part[t * len + e] */
static void sum_threads(const double* part, int nt, uint64_t len, double* out) {
#pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < len; e++) {
        double v = 0.0;
        for (int t = 0; t < nt; t++) v += part[(size_t)t * len + e];
        out[e] = v;
    }
}

int or_threads(void) { return omp_get_max_threads(); }

/* ---------------------------------------------------------------- operator */

/* S*P (operator.rs:894-948): u_c = sum_b p[off_b + code_b(c)], t_s = sum_c W_c u_c / n1_s,
 * y[off_b + code_b(c)] += W_c (u_c - t_s). */
void or_apply(const or_solver* s, const double* p, uint64_t r, double* y) {
    uint64_t nb = s->nb, my = s->m * r;
    int nt = omp_get_max_threads();
    /* per-thread partial y over a static split of the segments, summed in thread order */
    double* ys = calloc((size_t)nt * my + 1, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        double* yt = ys + (size_t)omp_get_thread_num() * my;
        double* u = malloc((r ? r : 1) * sizeof(double));
        double* t = malloc((r ? r : 1) * sizeof(double));
#pragma omp for schedule(static)
        for (uint64_t sg = 0; sg < s->g1; sg++) {
            uint32_t c0 = s->seg_off[sg], c1 = s->seg_off[sg + 1];
            for (uint64_t j = 0; j < r; j++) t[j] = 0.0;
            for (uint32_t c = c0; c < c1; c++)
                for (uint64_t j = 0; j < r; j++) {
                    double v = 0.0;
                    for (uint64_t b = 0; b < nb; b++) v += p[(s->boff[b] + s->cell_codes[(uint64_t)c * nb + b]) * r + j];
                    t[j] += s->cw[c] * v;
                }
            for (uint64_t j = 0; j < r; j++) t[j] = s->n1[sg] > 0.0 ? t[j] / s->n1[sg] : 0.0;
            for (uint32_t c = c0; c < c1; c++) {
                for (uint64_t j = 0; j < r; j++) {
                    double v = 0.0;
                    for (uint64_t b = 0; b < nb; b++) v += p[(s->boff[b] + s->cell_codes[(uint64_t)c * nb + b]) * r + j];
                    u[j] = s->cw[c] * (v - t[j]);
                }
                for (uint64_t b = 0; b < nb; b++) {
                    uint64_t k = s->boff[b] + s->cell_codes[(uint64_t)c * nb + b];
                    for (uint64_t j = 0; j < r; j++) yt[k * r + j] += u[j];
                }
            }
        }
        free(u); free(t);
    }
    sum_threads(ys, nt, my, y);
    free(ys);
}

/* Ops::rhs (operator.rs:969-1066): s_c = sum w v, b1_s, z_c = s_c - W_c b1_s / n1_s,
 * rhs[k] = sum z_c, wsq = sum w v^2, wsum = sum w v. */
void or_rhs(const or_solver* s, const double* rows, uint64_t r, double* rhs, double* b1, double* wsq, double* wsum) {
    uint64_t nb = s->nb, mr = s->m * r;
    uint64_t maxc = 0;
    for (uint64_t sg = 0; sg < s->g1; sg++) {
        uint64_t k = s->seg_off[sg + 1] - s->seg_off[sg];
        if (k > maxc) maxc = k;
    }
    int nt = omp_get_max_threads();
    /* per thread: rhs (m r), wsq (r), wsum (r); summed in thread order */
    uint64_t stride = mr + 2 * r;
    double* acc = calloc((size_t)nt * stride + 1, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        double* at = acc + (size_t)omp_get_thread_num() * stride;
        double *rt = at, *sq = at + mr, *sm = at + mr + r;
        double* sc = malloc((maxc * r + 1) * sizeof(double));
#pragma omp for schedule(static)
        for (uint64_t sg = 0; sg < s->g1; sg++) {
            uint32_t c0 = s->seg_off[sg], c1 = s->seg_off[sg + 1];
            for (uint64_t j = 0; j < r; j++) b1[sg * r + j] = 0.0;
            for (uint32_t c = c0; c < c1; c++) {
                double* v = sc + (c - c0) * r;
                for (uint64_t j = 0; j < r; j++) v[j] = 0.0;
                for (uint64_t i = cell_row0(s, c); i < cell_row1(s, c); i++) {
                    double w = s->w_rows ? s->w_rows[i] : 1.0;
                    for (uint64_t j = 0; j < r; j++) {
                        double wv = w * rows[i * r + j];
                        v[j] += wv;
                        sq[j] += wv * rows[i * r + j];
                        sm[j] += wv;
                    }
                }
                for (uint64_t j = 0; j < r; j++) b1[sg * r + j] += v[j];
            }
            for (uint32_t c = c0; c < c1; c++) {
                double* v = sc + (c - c0) * r;
                double f = s->cw[c] / s->n1[sg];
                for (uint64_t j = 0; j < r; j++) v[j] -= f * b1[sg * r + j];
                for (uint64_t b = 0; b < nb; b++) {
                    uint64_t k = s->boff[b] + s->cell_codes[(uint64_t)c * nb + b];
                    for (uint64_t j = 0; j < r; j++) rt[k * r + j] += v[j];
                }
            }
        }
        free(sc);
    }
    double* tot = malloc((stride + 1) * sizeof(double));
    sum_threads(acc, nt, stride, tot);
    memcpy(rhs, tot, mr * sizeof(double));
    memcpy(wsq, tot + mr, r * sizeof(double));
    memcpy(wsum, tot + mr + r, r * sizeof(double));
    free(tot);
    free(acc);
}

/* Ops::residualize_rows (operator.rs:1107-1179): alpha_s = (b1_s - sum W_c u_c) / n1_s,
 * v_i -= alpha_s + u_c. */
void or_residualize_rows(const or_solver* s, double* rows, uint64_t r, const double* x, const double* b1, double* alpha) {
    uint64_t nb = s->nb;
#pragma omp parallel
    {
    double* a = malloc((r ? r : 1) * sizeof(double));
    double* u = malloc((r ? r : 1) * sizeof(double));
#pragma omp for schedule(static)
    for (uint64_t sg = 0; sg < s->g1; sg++) {
        uint32_t c0 = s->seg_off[sg], c1 = s->seg_off[sg + 1];
        for (uint64_t j = 0; j < r; j++) a[j] = b1[sg * r + j];
        for (uint32_t c = c0; c < c1; c++)
            for (uint64_t j = 0; j < r; j++) {
                double v = 0.0;
                for (uint64_t b = 0; b < nb; b++) v += x[(s->boff[b] + s->cell_codes[(uint64_t)c * nb + b]) * r + j];
                a[j] -= s->cw[c] * v;
            }
        for (uint64_t j = 0; j < r; j++) a[j] = s->n1[sg] > 0.0 ? a[j] / s->n1[sg] : 0.0;
        if (alpha)
            for (uint64_t j = 0; j < r; j++) alpha[sg * r + j] = a[j];
        for (uint32_t c = c0; c < c1; c++) {
            for (uint64_t j = 0; j < r; j++) {
                double v = 0.0;
                for (uint64_t b = 0; b < nb; b++) v += x[(s->boff[b] + s->cell_codes[(uint64_t)c * nb + b]) * r + j];
                u[j] = v + a[j];
            }
            for (uint64_t i = cell_row0(s, c); i < cell_row1(s, c); i++)
                for (uint64_t j = 0; j < r; j++) rows[i * r + j] -= u[j];
        }
    }
    free(a); free(u);
    }
}

/* ---------------------------------------------------------------- PCG */

static void dots(const double* a, const double* b, uint64_t m, uint64_t w, double* out) {
    for (uint64_t j = 0; j < w; j++) out[j] = 0.0;
    for (uint64_t i = 0; i < m; i++)
        for (uint64_t j = 0; j < w; j++) out[j] += a[i * w + j] * b[i * w + j];
}

/* cg::pcg (cg.rs:132-292) without probes/escalation: the oracle keeps its block-Jacobi
 * preconditioner (dense small block + Jacobi) for the whole solve. */
static int pcg(const or_solver* s, const double* rhs, double* x, uint64_t w, const double* ss1, const double* ss0,
               double tol, uint64_t maxiter, uint64_t* iters_out, int* exact_out) {
    uint64_t m = s->m, n = m * w;
    double* r = calloc(n + 1, sizeof(double));
    double* z = calloc(n + 1, sizeof(double));
    double* p = calloc(n + 1, sizeof(double));
    double* q = calloc(n + 1, sizeof(double));
    int* active = malloc((w + 1) * sizeof(int));
    double* floor_ = malloc((w + 1) * sizeof(double));
    double* rho = malloc((w + 1) * sizeof(double));
    double* rho_new = malloc((w + 1) * sizeof(double));
    double* pq = malloc((w + 1) * sizeof(double));
    double* gamma = malloc((w + 1) * sizeof(double));
    double* xb = malloc((w + 1) * sizeof(double));
    double* xr = malloc((w + 1) * sizeof(double));
    double* ring = calloc(w * OR_WINDOW + 1, sizeof(double));
    int rc = 0;
    uint64_t iterations = 0;
    *exact_out = 0;
    for (uint64_t j = 0; j < w; j++) { active[j] = ss0[j] > 0.0; floor_[j] = 1e-26 * ss0[j]; }
    int xnz = 0;
    for (uint64_t i = 0; i < n; i++)
        if (x[i] != 0.0) { xnz = 1; break; }
    if (xnz) {
        or_apply(s, x, w, q);
        for (uint64_t i = 0; i < n; i++) r[i] = rhs[i] - q[i];
    } else {
        memcpy(r, rhs, n * sizeof(double));
    }
    precond_apply(s, r, w, z);
    if (s->ms == s->m && s->nkeep > 0) { /* exact: M = S (solve.rs:923-930) */
        for (uint64_t i = 0; i < m; i++)
            for (uint64_t j = 0; j < w; j++)
                if (active[j]) x[i * w + j] += z[i * w + j];
        iterations = 1;
        *exact_out = 1;
        goto done;
    }
    memcpy(p, z, n * sizeof(double));
    dots(r, z, m, w, rho);
    for (uint64_t j = 0; j < w; j++)
        if (!(rho[j] > 0.0)) active[j] = 0;
    for (uint64_t i = 0; i < m; i++)
        for (uint64_t j = 0; j < w; j++)
            if (!active[j]) p[i * w + j] = 0.0;
    uint64_t k = 0;
    for (;;) {
        int any = 0;
        for (uint64_t j = 0; j < w; j++) any |= active[j];
        if (!any) break;
        if (iterations >= maxiter) { rc = 430; goto done; }
        or_apply(s, p, w, q);
        dots(p, q, m, w, pq);
        for (uint64_t j = 0; j < w; j++) {
            gamma[j] = 0.0;
            if (active[j]) {
                if (pq[j] > 0.0 && isfinite(pq[j])) gamma[j] = rho[j] / pq[j];
                else active[j] = 0;
            }
        }
        for (uint64_t j = 0; j < w; j++) { xb[j] = 0.0; xr[j] = 0.0; }
        for (uint64_t i = 0; i < m; i++)
            for (uint64_t j = 0; j < w; j++) {
                double xv = x[i * w + j] + gamma[j] * p[i * w + j];
                double rv = r[i * w + j] - gamma[j] * q[i * w + j];
                x[i * w + j] = xv;
                r[i * w + j] = rv;
                xb[j] += xv * rhs[i * w + j];
                xr[j] += xv * rv;
            }
        precond_apply(s, r, w, z);
        dots(r, z, m, w, rho_new);
        iterations++;
        k++;
        for (uint64_t j = 0; j < w; j++) {
            if (!active[j]) continue;
            ring[j * OR_WINDOW + (k - 1) % OR_WINDOW] = gamma[j] * rho[j];
            double ssk = ss1[j] - (xb[j] + xr[j]);
            double thr = tol * tol * ssk;
            if (floor_[j] > thr) thr = floor_[j];
            double sum = 0.0;
            for (int t = 0; t < OR_WINDOW; t++) sum += ring[j * OR_WINDOW + t];
            int done_ = k >= OR_WINDOW && sum <= thr;
            int at_roundoff = rho_new[j] <= floor_[j];
            if (done_ || at_roundoff || !(rho_new[j] > 0.0)) active[j] = 0;
        }
        for (uint64_t i = 0; i < m; i++)
            for (uint64_t j = 0; j < w; j++)
                p[i * w + j] = active[j] ? z[i * w + j] + (rho_new[j] / rho[j]) * p[i * w + j] : 0.0;
        memcpy(rho, rho_new, w * sizeof(double));
    }
done:
    *iters_out = iterations;
    free(r); free(z); free(p); free(q); free(active); free(floor_); free(rho); free(rho_new);
    free(pq); free(gamma); free(xb); free(xr); free(ring);
    return rc;
}

/* FeSolver::residualize (solve.rs:1141-1311). */
int or_residualize(or_solver* s, double* zrows, uint64_t r, double tol, uint64_t maxiter, const double* warm,
                   double* coefs_out, uint64_t* iterations, int32_t* exact) {
    uint64_t m = s->m, g1 = s->g1;
    *iterations = 1;
    *exact = 1;
    if (r == 0) return 0;
    double* rhs = calloc(m * r + 1, sizeof(double));
    double* b1 = calloc(g1 * r + 1, sizeof(double));
    double* wsq = calloc(r, sizeof(double));
    double* wsum = calloc(r, sizeof(double));
    double* ss0 = calloc(r, sizeof(double));
    double* ss1 = calloc(r, sizeof(double));
    double* x = calloc(m * r + 1, sizeof(double));
    or_rhs(s, zrows, r, rhs, b1, wsq, wsum);
    for (uint64_t j = 0; j < r; j++) {
        double v = wsq[j] - wsum[j] * wsum[j] / s->sw;
        ss0[j] = v > 1e-15 * wsq[j] ? v : 0.0;
        double b1sq = 0.0;
        for (uint64_t sg = 0; sg < g1; sg++)
            if (s->n1[sg] > 0.0) b1sq += b1[sg * r + j] * b1[sg * r + j] / s->n1[sg]; /* empty levels: no term */
        ss1[j] = wsq[j] - b1sq;
    }
    if (warm) memcpy(x, warm, m * r * sizeof(double));
    for (uint64_t j = 0; j < r; j++)
        if (ss0[j] == 0.0)
            for (uint64_t i = 0; i < m; i++) x[i * r + j] = 0.0;
    int rc = 0;
    if (m > 0) {
        int ex = 0;
        rc = pcg(s, rhs, x, r, ss1, ss0, tol, maxiter, iterations, &ex);
        if (*iterations < 1) *iterations = 1;
        int allconst = 1;
        for (uint64_t j = 0; j < r; j++) allconst &= ss0[j] == 0.0;
        *exact = ex || allconst;
    }
    if (rc == 0) {
        double* alpha = coefs_out ? coefs_out + m * r : NULL;
        or_residualize_rows(s, zrows, r, x, b1, alpha);
        if (coefs_out) memcpy(coefs_out, x, m * r * sizeof(double));
    }
    free(rhs); free(b1); free(wsq); free(wsum); free(ss0); free(ss1); free(x);
    return rc;
}
