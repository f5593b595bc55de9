// Varying slopes: absorbed terms fe#c.x, fe##c.x and slope-only terms (slopes.rs:1-33), behind the
// same absorber entry points as the intercept-only solver.
//
// Structure (slopes.rs:354-527). Rows are sorted by the level of the eliminated term E (the
// intercept-only term with the most levels, else the term with the most levels) and every row is
// its own cell. A term of width b = [intercept] + slopes owns b consecutive reduced coordinates per
// level (intercept first); the non-E terms form blocks, the small set first (the reference's size
// and cost rule, which fixes the FeCoefs layout), then the others, each group in term order.
// Per row the device holds the first coordinate of its level in every block (coord, n x nb), the
// blocks' slope values (sv, n x nsv) and E's slopes (he, n x e_ns); per block the rows sorted by
// (level, row) (perm, poff), which within a level are grouped by E's level because rows are in
// E order.
//
// Operator (slopes.rs:13-17). With h_q,i the design row of term q (1?, x...) and M_s^+ the
// generalized inverse of M_s = sum_{i in s} w_i h_E h_E' per level s of E:
//   u_i = sum_q h_q,i' P[q, l_q(i)];  t_s = M_s^+ sum_{i in s} w_i h_E,i u_i';
//   z_i = w_i (u_i - h_E,i' t_s);     Y[q, l] = sum_{i in l} h_q,i z_i'.
// The first line is one pass per segment of E (one wave per segment, lanes over its rows: the
// wave reduction gives t_s in fixed order); the last is a pull per level of each block (one wave
// per level over its row list, fixed order). No atomics: results are bitwise reproducible.
// The right-hand side is the same with u_i = v_i; the residualization replaces v by
// d - h_E' t(d) with d = v - u(x).
//
// Preconditioner (slopes.rs:16-21): the dense grounded inverse of the small coordinates when they
// are few (the intercept-only solver's dense_limit; their block of S is assembled by products with
// unit vectors), and b x b block Jacobi on the Schur complement's diagonal blocks of every other
// level (generalized inverse by the symmetric sweep with pivots <= 1e-10 of their diagonal dropped,
// redundant coordinates (diagonal <= 1e-12 of the level's Gram) zero). The PCG is the
// intercept-only solver's own, in FP64. Escalation to sparse factors is not built for slopes: the
// block Jacobi stays.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fh_fit.hpp"
#include "fh_internal.hpp"
#include "fh_sort.hpp"

namespace fh {

namespace {

constexpr int TPB = 256;
constexpr int WPB = TPB / WAVE;
constexpr int RC = 4;                   // columns per sweep of the row kernels
// A block with a level longer than this many rows runs its level passes over items of at most CH
// rows (fixed boundaries) whose partial sums are added per level in item order.
constexpr uint32_t CH = 1024;
constexpr int SL_SMAX = NB_MAX * SL_BMAX;  // slope columns of the non-E blocks
constexpr int DOTG = SPMV_DOTG;
constexpr double LEVEL_TOL = 1e-12;     // slopes.rs:58: per-level inverses of E
constexpr double JACOBI_TOL = 1e-10;    // slopes.rs:60: block-Jacobi inverses
constexpr double REDUNDANT_REL = 1e-12; // operator.rs:21
// The small set's size and cost rule (slopes.rs:63-71, 444-460). It decides which blocks come
// first, hence the FeCoefs layout, so it is kept as the reference states it (seconds per unit).
constexpr double C_ROW = 1e-9, C_SMALL_ASM = 1e-9, C_FLOP = 1.2e-11, SMALL_FLOOR = 1e-3, SMALL_ITERS = 8.0;

}  // namespace

struct SlBlock {
    uint32_t term;
    uint64_t g;
    int icpt, ns, b;
    uint64_t off;  // first reduced coordinate
    int s0;        // first slope column in sv
    bool small;
    uint64_t joff;  // first entry of its block-Jacobi inverses (jacobi blocks)
    DBuf<uint32_t> perm, poff;  // rows by (level, row); offsets per level (g + 1)
    // long levels: items of <= CH rows of the level lists (first / end position, level) and the
    // items of each level (g + 1)
    uint64_t poff_n = 0;  // rows over its levels (n)
    DBuf<double> svp;     // its slope values in level-list order (n x ns)
    bool chunked = false;
    uint64_t nitems = 0;
    DBuf<uint32_t> it_first, it_end, it_lvl, lvl_it;
};

struct SlopeState {
    uint32_t e = 0;
    uint64_t ge = 0;
    int e_icpt = 1, e_ns = 0, be = 1;
    std::vector<SlBlock> blocks;
    int nsmall = 0, nsv = 0;
    uint64_t m_small = 0;
    bool dense = false;  // the small coordinates take the dense grounded inverse
    std::vector<DBuf<double>> slope_cols;  // kept-order slope columns (owned copies)
    DBuf<uint32_t> seg_off, row_seg, coord;
    DBuf<double> sv, he;
    // weights
    DBuf<double> minv;   // ge x be^2
    DBuf<double> jinv;   // block-Jacobi inverses
    DBuf<double> gdiag;  // m: diagonal of sum w h h' per coordinate
    DBuf<uint8_t> red;   // m: redundant coordinates
    DBuf<double> zbuf;   // n x r scratch of the row kernels
    DBuf<double> ipart;  // item partials of the chunked level passes
    // E's segments in items of <= CH rows when one is longer (first / end row, segment; items of
    // each segment, ge + 1), their partials and t per segment (ge x be x r)
    bool e_chunked = false;
    uint64_t e_nitems = 0;
    DBuf<uint32_t> e_first, e_end, e_lvl, seg_it;
    DBuf<double> epart, tbuf, segpart;
};

namespace {

// ------------------------------------------------------------------ device helpers

// Generalized inverse of the symmetric K x K row-major a in place by the symmetric sweep
// (Goodnight); pivot p is dropped (row and column p zero) unless its Schur complement given the
// earlier kept pivots exceeds tol * refd[p] (slopes.rs:99-132). K is a compile-time constant so
// that a stays in registers.
template <int K>
__device__ void sweep_ginv(double (&a)[K * K], const double (&refd)[K], double tol) {
#pragma unroll
    for (int p = 0; p < K; p++) {
        const double d = a[p * K + p];
        if (!(d > tol * refd[p] && d > 0.0 && isfinite(d))) {
#pragma unroll
            for (int j = 0; j < K; j++) {
                a[p * K + j] = 0.0;
                a[j * K + p] = 0.0;
            }
            continue;
        }
        const double inv = 1.0 / d;
#pragma unroll
        for (int j = 0; j < K; j++) a[p * K + j] *= inv;
#pragma unroll
        for (int i = 0; i < K; i++) {
            if (i == p) continue;
            const double c = a[i * K + p];
#pragma unroll
            for (int j = 0; j < K; j++) a[i * K + j] -= c * a[p * K + j];
            a[i * K + p] = -c * inv;
        }
        a[p * K + p] = inv;
    }
}

// Design row (1?, x...) of E for row i.
__device__ __forceinline__ void e_row(const double* he, int e_ns, int e_icpt, size_t i, double* h) {
    int k = 0;
    if (e_icpt) h[k++] = 1.0;
    for (int s = 0; s < e_ns; s++) h[k++] = he[i * e_ns + s];
}

struct SegArgs {
    const uint32_t* seg_off;
    uint64_t ge;
    const uint32_t* coord;  // n x nb
    int nb;
    int bw[NB_MAX], icpt[NB_MAX], s0[NB_MAX];
    const double* sv;
    int nsv;
    const double* he;
    int e_ns, e_icpt, be;
    const double* minv;
    const double* w;  // null = unit
};

// Value of slot d of block b in row i (intercept 1, then the slopes).
__device__ __forceinline__ double slot_val(const SegArgs& A, int b, int d, size_t i) {
    if (A.icpt[b]) {
        if (d == 0) return 1.0;
        d -= 1;
    }
    return A.sv[i * A.nsv + A.s0[b] + d];
}

enum { M_APPLY = 0, M_RHS = 1, M_RESID = 2 };

// Groups of L lanes per segment of E (BE = E's width, a compile-time constant so that the
// accumulators stay in registers; a group's lanes stride the segment's rows and its sums are
// reduced by xor shuffles within the group, in fixed order).
//   M_APPLY: src = P (m x r); u = D_R P; dst = Z = w (u - h_E' t) (n x r).
//   M_RHS:   src = V (n x ld); u = v; dst = Z; part: per-wave [wsq | wsum | ss1] (3 r).
//   M_RESID: src = dst = V (n x ld); d = v - D_R x; v = d - h_E' t; alpha (ge be x r) = t.
template <int MODE, int L, int BE>
__global__ __launch_bounds__(TPB) void k_sl_seg(SegArgs A, const double* src, int ld, const double* __restrict__ x,
                                                int r, double* dst, double* __restrict__ alpha,
                                                double* __restrict__ part, const uint32_t* done) {
    if (done && *done) return;
    __shared__ double wsums[WPB][3][16];
    const int lane = lane_id(), wv = threadIdx.x / WAVE, gl = lane % L;
    if (MODE == M_RHS) {
        for (int e = lane; e < 48; e += WAVE) wsums[wv][e / 16][e % 16] = 0.0;
        __syncthreads();
    }
    constexpr int GPW = WAVE / L;  // groups per wave
    const uint64_t ngroups = (uint64_t)gridDim.x * WPB * GPW;
    const uint64_t g0 = ((uint64_t)blockIdx.x * WPB + wv) * GPW + lane / L;
    // every lane of a wave runs the same number of segment rounds (the shuffles of the RHS sums
    // and the group sums need their whole group)
    const uint64_t rounds = (A.ge + ngroups - 1) / ngroups;
    for (uint64_t rd = 0; rd < rounds; rd++) {
        const uint64_t s = g0 + rd * ngroups;
        const bool live = s < A.ge;
        const size_t i0 = live ? A.seg_off[s] : 0, i1 = live ? A.seg_off[s + 1] : 0;
        const double* mv = A.minv + (live ? s : 0) * BE * BE;
        for (int c0 = 0; c0 < r; c0 += RC) {
            const int nc = min(RC, r - c0);
            double acc[BE][RC], wsq[RC], wsm[RC];
#pragma unroll
            for (int c = 0; c < RC; c++) {
                wsq[c] = 0.0;
                wsm[c] = 0.0;
#pragma unroll
                for (int e = 0; e < BE; e++) acc[e][c] = 0.0;
            }
            for (size_t i = i0 + gl; i < i1; i += L) {
                const double wi = A.w ? A.w[i] : 1.0;
                double u[RC];
#pragma unroll
                for (int c = 0; c < RC; c++) u[c] = 0.0;
                if (MODE == M_APPLY || MODE == M_RESID) {
                    const double* P = MODE == M_APPLY ? src : x;
                    for (int b = 0; b < A.nb; b++) {
                        const size_t k0 = A.coord[i * A.nb + b];
                        for (int d = 0; d < A.bw[b]; d++) {
                            const double h = slot_val(A, b, d, i);
                            const double* pr = P + (k0 + d) * r + c0;
#pragma unroll
                            for (int c = 0; c < RC; c++)
                                if (c < nc) u[c] += h * pr[c];
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    if (c >= nc) continue;
                    if (MODE == M_RESID) {
                        u[c] = src[i * ld + c0 + c] - u[c];
                        dst[i * ld + c0 + c] = u[c];  // d, read back by this lane below
                    } else if (MODE == M_RHS) {
                        u[c] = src[i * ld + c0 + c];
                        wsq[c] += wi * u[c] * u[c];
                        wsm[c] += wi * u[c];
                    } else {
                        dst[i * r + c0 + c] = u[c];  // u, read back by this lane below
                    }
                }
                double h[BE];
#pragma unroll
                for (int e = 0; e < BE; e++)
                    h[e] = (A.e_icpt && e == 0) ? 1.0 : A.he[i * A.e_ns + e - A.e_icpt];
#pragma unroll
                for (int e = 0; e < BE; e++)
#pragma unroll
                    for (int c = 0; c < RC; c++) acc[e][c] += wi * h[e] * u[c];
            }
            // t = M_s^+ acc (every lane of the group holds the group's sums)
#pragma unroll
            for (int e = 0; e < BE; e++)
#pragma unroll
                for (int c = 0; c < RC; c++) acc[e][c] = group_sum<L>(acc[e][c]);
            double t[BE][RC];
#pragma unroll
            for (int e = 0; e < BE; e++) {
                double me[BE];
#pragma unroll
                for (int f = 0; f < BE; f++) me[f] = live ? mv[e * BE + f] : 0.0;
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    double v = 0.0;
#pragma unroll
                    for (int f = 0; f < BE; f++) v += me[f] * acc[f][c];
                    t[e][c] = v;
                }
            }
            if (MODE == M_RHS) {
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    const double q = group_sum<L>(wsq[c]), sm = group_sum<L>(wsm[c]);
                    double at = 0.0;
#pragma unroll
                    for (int e = 0; e < BE; e++) at += acc[e][c] * t[e][c];
                    // the wave's groups of this round, added by the wave reduction (fixed order)
                    const bool own = live && gl == 0;
                    const double wq = wave_sum(own ? q : 0.0), wsm2 = wave_sum(own ? sm : 0.0),
                                 ws1 = wave_sum(own ? q - at : 0.0);
                    if (lane == 0 && c < nc) {
                        wsums[wv][0][c0 + c] += wq;
                        wsums[wv][1][c0 + c] += wsm2;
                        wsums[wv][2][c0 + c] += ws1;
                    }
                }
            }
            if (MODE == M_RESID && alpha && live && gl == 0) {
#pragma unroll
                for (int e = 0; e < BE; e++)
#pragma unroll
                    for (int c = 0; c < RC; c++)
                        if (c < nc) alpha[(s * BE + e) * r + c0 + c] = t[e][c];
            }
            for (size_t i = i0 + gl; i < i1; i += L) {
                const double wi = A.w ? A.w[i] : 1.0;
                double h[BE];
#pragma unroll
                for (int e = 0; e < BE; e++)
                    h[e] = (A.e_icpt && e == 0) ? 1.0 : A.he[i * A.e_ns + e - A.e_icpt];
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    if (c >= nc) continue;
                    double ht = 0.0;
#pragma unroll
                    for (int e = 0; e < BE; e++) ht += h[e] * t[e][c];
                    if (MODE == M_APPLY) {
                        const size_t o = i * r + c0 + c;
                        dst[o] = wi * (dst[o] - ht);
                    } else if (MODE == M_RHS) {
                        dst[i * r + c0 + c] = wi * (src[i * ld + c0 + c] - ht);
                    } else {
                        const size_t o = i * ld + c0 + c;
                        dst[o] = dst[o] - ht;
                    }
                }
            }
        }
    }
    if (MODE == M_RHS) {
        __syncthreads();
        const size_t gw = (size_t)blockIdx.x * WPB + wv;
        for (int e = lane; e < 3 * r; e += WAVE) part[gw * 3 * r + e] = wsums[wv][e / r][e % r];
    }
}

struct ItemArgs {
    const uint32_t* first;
    const uint32_t* end;
    const uint32_t* lvl;
    const uint32_t* lvl_it;
    uint64_t nitems;
};

// E's long segments, phase A: one wave per item of a segment. u as in k_sl_seg (stashed the same
// way); part[it][e][c] = sum over the item's rows of w h_E,e u_c, and for M_RHS also sum w v^2 and
// sum w v (rows be and be + 1 of the item's (be + 2) x r partials).
template <int MODE>
__global__ __launch_bounds__(TPB) void k_sl_segA(SegArgs A, ItemArgs I, const double* src, int ld,
                                                 const double* __restrict__ x, int r, double* dst,
                                                 double* __restrict__ part, const uint32_t* done) {
    if (done && *done) return;
    const int lane = lane_id(), be = A.be, pw = be + 2;
    for (uint64_t it = ((uint64_t)blockIdx.x * TPB + threadIdx.x) / WAVE; it < I.nitems;
         it += (uint64_t)gridDim.x * WPB) {
        const size_t i0 = I.first[it], i1 = I.end[it];
        for (int c0 = 0; c0 < r; c0 += RC) {
            const int nc = min(RC, r - c0);
            double acc[SL_BMAX][RC], wsq[RC], wsm[RC];
#pragma unroll
            for (int c = 0; c < RC; c++) {
                wsq[c] = 0.0;
                wsm[c] = 0.0;
#pragma unroll
                for (int e = 0; e < SL_BMAX; e++) acc[e][c] = 0.0;
            }
            for (size_t i = i0 + lane; i < i1; i += WAVE) {
                const double wi = A.w ? A.w[i] : 1.0;
                double u[RC];
#pragma unroll
                for (int c = 0; c < RC; c++) u[c] = 0.0;
                if (MODE == M_APPLY || MODE == M_RESID) {
                    const double* P = MODE == M_APPLY ? src : x;
                    for (int b = 0; b < A.nb; b++) {
                        const size_t k0 = A.coord[i * A.nb + b];
                        for (int d = 0; d < A.bw[b]; d++) {
                            const double h = slot_val(A, b, d, i);
#pragma unroll
                            for (int c = 0; c < RC; c++)
                                if (c < nc) u[c] += h * P[(k0 + d) * r + c0 + c];
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    if (c >= nc) continue;
                    if (MODE == M_RESID) {
                        u[c] = src[i * ld + c0 + c] - u[c];
                        dst[i * ld + c0 + c] = u[c];
                    } else if (MODE == M_RHS) {
                        u[c] = src[i * ld + c0 + c];
                        wsq[c] += wi * u[c] * u[c];
                        wsm[c] += wi * u[c];
                    } else {
                        dst[i * r + c0 + c] = u[c];
                    }
                }
                double h[SL_BMAX];
                e_row(A.he, A.e_ns, A.e_icpt, i, h);
#pragma unroll
                for (int e = 0; e < SL_BMAX; e++)
                    if (e < be)
#pragma unroll
                        for (int c = 0; c < RC; c++) acc[e][c] += wi * h[e] * u[c];
            }
#pragma unroll
            for (int e = 0; e < SL_BMAX + 2; e++)
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    if (e >= be && !(MODE == M_RHS && e < be + 2)) continue;
                    double v = 0.0;
                    if (e < be) {
#pragma unroll
                        for (int ee = 0; ee < SL_BMAX; ee++)
                            if (ee == e) v = acc[ee][c];
                    } else {
                        v = e == be ? wsq[c] : wsm[c];
                    }
                    v = wave_sum(v);
                    if (lane == 0 && c < nc) part[(it * pw + e) * r + c0 + c] = v;
                }
        }
    }
}

// Phase B: per (segment, column) the items' partials in order, t = M_s^+ acc into T (and alpha),
// and for M_RHS the segment's [wsq | wsum | ss1] (3 r per segment).
template <int MODE>
__global__ void k_sl_segB(SegArgs A, ItemArgs I, const double* __restrict__ part, int r, double* __restrict__ T,
                          double* __restrict__ alpha, double* __restrict__ segpart, const uint32_t* done) {
    if (done && *done) return;
    const int be = A.be, pw = be + 2;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < A.ge * (uint64_t)r;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = e / r;
        const int c = (int)(e % r);
        double acc[SL_BMAX], q = 0.0, sm = 0.0;
        for (int f = 0; f < SL_BMAX; f++) acc[f] = 0.0;
        for (uint32_t it = I.lvl_it[s]; it < I.lvl_it[s + 1]; it++) {
            for (int f = 0; f < be; f++) acc[f] += part[((uint64_t)it * pw + f) * r + c];
            if (MODE == M_RHS) {
                q += part[((uint64_t)it * pw + be) * r + c];
                sm += part[((uint64_t)it * pw + be + 1) * r + c];
            }
        }
        const double* mv = A.minv + s * be * be;
        double at = 0.0;
        for (int f = 0; f < be; f++) {
            double t = 0.0;
            for (int g2 = 0; g2 < be; g2++) t += mv[f * be + g2] * acc[g2];
            T[(s * be + f) * r + c] = t;
            if (MODE == M_RESID && alpha) alpha[(s * be + f) * r + c] = t;
            at += acc[f] * t;
        }
        if (MODE == M_RHS) {
            segpart[s * 3 * r + c] = q;
            segpart[s * 3 * r + r + c] = sm;
            segpart[s * 3 * r + 2 * r + c] = q - at;
        }
    }
}

// Phase C, one thread per row: z = w (u - h_E' t_s) (M_APPLY, M_RHS into dst), v = d - h_E' t_s
// (M_RESID in place).
template <int MODE>
__global__ void k_sl_segC(SegArgs A, const uint32_t* __restrict__ row_seg, size_t n, const double* src, int ld,
                          int r, const double* __restrict__ T, double* dst, const uint32_t* done) {
    if (done && *done) return;
    const int be = A.be;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t s = row_seg[i];
        const double wi = A.w ? A.w[i] : 1.0;
        double h[SL_BMAX];
        e_row(A.he, A.e_ns, A.e_icpt, i, h);
        for (int c = 0; c < r; c++) {
            double ht = 0.0;
            for (int f = 0; f < be; f++) ht += h[f] * T[(s * be + f) * r + c];
            if (MODE == M_APPLY) dst[i * r + c] = wi * (dst[i * r + c] - ht);
            else if (MODE == M_RHS) dst[i * r + c] = wi * (src[i * ld + c] - ht);
            else dst[i * ld + c] = dst[i * ld + c] - ht;
        }
    }
}

struct PullArgs {
    const uint32_t* perm;
    const uint32_t* poff;
    uint64_t g;
    uint64_t off;
    int b, icpt, ns;
    const double* svp;  // the block's slope values in level-list order (n x ns)
};

// Y[off + l b + d][c] = sum over the rows i of level l of h_d(i) Z[i][c]: groups of L lanes per
// level, B = the block's width (compile time), group sums by xor shuffles (fixed order).
template <int L, int B>
__global__ __launch_bounds__(TPB) void k_sl_pull(PullArgs P, const double* __restrict__ Z, int r,
                                                 double* __restrict__ Y, const uint32_t* done) {
    if (done && *done) return;
    constexpr int GPW = WAVE / L;
    const int lane = lane_id(), gl = lane % L;
    const uint64_t ngroups = (uint64_t)gridDim.x * WPB * GPW;
    const uint64_t g0 = ((uint64_t)blockIdx.x * WPB + threadIdx.x / WAVE) * GPW + lane / L;
    for (uint64_t l = g0; l < P.g; l += ngroups) {
        const size_t k0 = P.poff[l], k1 = P.poff[l + 1];
        for (int c0 = 0; c0 < r; c0 += RC) {
            const int nc = min(RC, r - c0);
            double acc[B][RC];
#pragma unroll
            for (int d = 0; d < B; d++)
#pragma unroll
                for (int c = 0; c < RC; c++) acc[d][c] = 0.0;
            for (size_t k = k0 + gl; k < k1; k += L) {
                const size_t i = P.perm[k];
                double z[RC];
#pragma unroll
                for (int c = 0; c < RC; c++) z[c] = c < nc ? Z[i * r + c0 + c] : 0.0;
#pragma unroll
                for (int d = 0; d < B; d++) {
                    const double h = P.icpt ? (d == 0 ? 1.0 : P.svp[k * P.ns + d - 1]) : P.svp[k * P.ns + d];
#pragma unroll
                    for (int c = 0; c < RC; c++) acc[d][c] += h * z[c];
                }
            }
#pragma unroll
            for (int d = 0; d < B; d++)
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    const double v = group_sum<L>(acc[d][c]);
                    if (gl == 0 && c < nc) Y[(P.off + l * B + d) * r + c0 + c] = v;
                }
        }
    }
}

// The pull over items: part[it][d][c] = sum over the item's rows of h_d(i) Z[i][c] (b x r per item).
__global__ __launch_bounds__(TPB) void k_sl_pull_items(PullArgs P, ItemArgs I, const double* __restrict__ Z, int r,
                                                       double* __restrict__ part, const uint32_t* done) {
    if (done && *done) return;
    const int lane = lane_id();
    for (uint64_t it = ((uint64_t)blockIdx.x * TPB + threadIdx.x) / WAVE; it < I.nitems;
         it += (uint64_t)gridDim.x * WPB) {
        const size_t k0 = I.first[it], k1 = I.end[it];
        for (int c0 = 0; c0 < r; c0 += RC) {
            const int nc = min(RC, r - c0);
            double acc[SL_BMAX][RC];
#pragma unroll
            for (int d = 0; d < SL_BMAX; d++)
#pragma unroll
                for (int c = 0; c < RC; c++) acc[d][c] = 0.0;
            for (size_t k = k0 + lane; k < k1; k += WAVE) {
                const size_t i = P.perm[k];
                double z[RC];
#pragma unroll
                for (int c = 0; c < RC; c++) z[c] = c < nc ? Z[i * r + c0 + c] : 0.0;
#pragma unroll
                for (int d = 0; d < SL_BMAX; d++) {
                    if (d >= P.b) break;
                    const double h = P.icpt ? (d == 0 ? 1.0 : P.svp[k * P.ns + d - 1]) : P.svp[k * P.ns + d];
#pragma unroll
                    for (int c = 0; c < RC; c++) acc[d][c] += h * z[c];
                }
            }
#pragma unroll
            for (int d = 0; d < SL_BMAX; d++)
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    if (d >= P.b) continue;
                    const double v = wave_sum(acc[d][c]);
                    if (lane == 0 && c < nc) part[(it * P.b + d) * r + c0 + c] = v;
                }
        }
    }
}

// out[l * len + e] = sum of part[it * len + e] over the items of level l: one workgroup per level,
// its threads over the items (it = first + t, t + TPB, ...), then the waves in order (fixed).
__global__ __launch_bounds__(TPB) void k_sl_item_sum(const uint32_t* __restrict__ lvl_it, uint64_t g,
                                                     const double* __restrict__ part, int len,
                                                     double* __restrict__ out, const uint32_t* done) {
    if (done && *done) return;
    __shared__ double red[WPB];
    for (uint64_t l = blockIdx.x; l < g; l += gridDim.x) {
        const uint32_t a0 = lvl_it[l], a1 = lvl_it[l + 1];
        for (int q = 0; q < len; q++) {
            double s = 0.0;
            for (uint32_t it = a0 + threadIdx.x; it < a1; it += TPB) s += part[(uint64_t)it * len + q];
            s = wave_sum(s);
            if (lane_id() == 0) red[threadIdx.x / WAVE] = s;
            __syncthreads();
            if (threadIdx.x == 0) {
                double t = 0.0;
                for (int w = 0; w < WPB; w++) t += red[w];
                out[l * len + q] = t;
            }
            __syncthreads();
        }
    }
}

// Fixed-order partial dot products per column: part[wg * 16 + j] = sum over this workgroup's
// elements e (e % r == j) of a[e] b[e], over m x r.
__global__ __launch_bounds__(TPB) void k_sl_dot(const double* __restrict__ a, const double* __restrict__ b, size_t len,
                                                int r, double* __restrict__ part, const uint32_t* done) {
    if (done && *done) return;
    __shared__ double red[WPB][16];
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = 0.0;
    for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < len; e += (size_t)gridDim.x * TPB) {
        const int j = (int)(e % r);
        const double v = a[e] * b[e];
#pragma unroll
        for (int jj = 0; jj < 16; jj++)
            if (jj == j) acc[jj] += v;
    }
    const int wv = threadIdx.x / WAVE;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const double v = wave_sum(acc[j]);
        if (lane_id() == 0) red[wv][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        double s = 0.0;
        for (int w = 0; w < WPB; w++) s += red[w][threadIdx.x];
        part[(size_t)blockIdx.x * 16 + threadIdx.x] = s;
    }
}

struct JacArgs {
    int nblk;
    uint64_t lcum[NB_MAX + 1];  // cumulative levels over the Jacobi blocks
    uint64_t off[NB_MAX], joff[NB_MAX];
    int b[NB_MAX];
    const double* jinv;
};

// z = J_l r on every level of the Jacobi blocks (b x b inverses) and the partial r'z per column.
__global__ __launch_bounds__(TPB) void k_sl_jacobi(JacArgs J, const double* __restrict__ rr, double* __restrict__ z,
                                                   int r, double* __restrict__ part, const uint32_t* done) {
    if (done && *done) return;
    __shared__ double red[WPB][16];
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = 0.0;
    const uint64_t L = J.lcum[J.nblk];
    for (uint64_t gl = (uint64_t)blockIdx.x * TPB + threadIdx.x; gl < L; gl += (uint64_t)gridDim.x * TPB) {
        int q = 0;
        while (gl >= J.lcum[q + 1]) q++;
        const uint64_t l = gl - J.lcum[q];
        const int b = J.b[q];
        const size_t k0 = J.off[q] + l * b;
        const double* M = J.jinv + J.joff[q] + l * b * b;
        for (int c = 0; c < r; c++) {
            double rv[SL_BMAX];
#pragma unroll
            for (int d = 0; d < SL_BMAX; d++) rv[d] = d < b ? rr[(k0 + d) * r + c] : 0.0;
            double s = 0.0;
#pragma unroll
            for (int d = 0; d < SL_BMAX; d++) {
                if (d >= b) break;
                double v = 0.0;
#pragma unroll
                for (int f = 0; f < SL_BMAX; f++)
                    if (f < b) v += M[d * b + f] * rv[f];
                z[(k0 + d) * r + c] = v;
                s += rv[d] * v;
            }
#pragma unroll
            for (int jj = 0; jj < 16; jj++)
                if (jj == c) acc[jj] += s;
        }
    }
    const int wv = threadIdx.x / WAVE;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const double v = wave_sum(acc[j]);
        if (lane_id() == 0) red[wv][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        double s = 0.0;
        for (int w = 0; w < WPB; w++) s += red[w][threadIdx.x];
        part[(size_t)blockIdx.x * 16 + threadIdx.x] = s;
    }
}

// ------------------------------------------------------------------ structure kernels

// keys[i] = code(i) << 32 | i, vals[i] = i; code(i) = codes[idx ? idx[i] : i].
__global__ void k_sl_keys(const uint32_t* __restrict__ codes, const uint32_t* __restrict__ idx, size_t n,
                          uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t c = codes[idx ? idx[i] : i];
        keys[i] = ((uint64_t)c << 32) | (uint64_t)i;
        vals[i] = (uint32_t)i;
    }
}

// off[l] = first position of code l in the sorted keys (n + 1 threads; empty levels share the
// next present one's position, off[g] = n).
__global__ void k_sl_offsets(const uint64_t* __restrict__ keys, size_t n, uint64_t g, uint32_t* __restrict__ off) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (size_t)gridDim.x * blockDim.x) {
        const int64_t prev = i == 0 ? -1 : (int64_t)(keys[i - 1] >> 32);
        const int64_t cur = i == n ? (int64_t)g : (int64_t)(keys[i] >> 32);
        for (int64_t l = prev + 1; l <= cur; l++) off[l] = (uint32_t)i;
    }
}

__global__ void k_sl_row_seg(const uint32_t* __restrict__ seg_off, uint64_t ge, uint32_t* __restrict__ row_seg) {
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ge; s += (uint64_t)gridDim.x * blockDim.x)
        for (uint32_t i = seg_off[s]; i < seg_off[s + 1]; i++) row_seg[i] = (uint32_t)s;
}

struct RowSrc {
    const uint32_t* codes[NB_MAX];
    uint64_t off[NB_MAX];
    int b[NB_MAX];
    const double* sl[SL_SMAX];
    const double* hel[SL_BMAX];
    int nb, nsv, e_ns;
};

// Row-order copies of the blocks' first coordinates and slope values and of E's slopes.
__global__ void k_sl_rows(RowSrc S, const uint32_t* __restrict__ pi, size_t n, uint32_t* __restrict__ coord,
                          double* __restrict__ sv, double* __restrict__ he) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t o = pi[i];
        for (int b = 0; b < S.nb; b++) coord[i * S.nb + b] = (uint32_t)(S.off[b] + (uint64_t)S.codes[b][o] * S.b[b]);
        for (int s = 0; s < S.nsv; s++) sv[i * S.nsv + s] = S.sl[s][o];
        for (int s = 0; s < S.e_ns; s++) he[i * S.e_ns + s] = S.hel[s][o];
    }
}

// svp[k * ns + j] = sv[perm[k] * nsv + s0 + j]: a block's slope values in its level-list order.
__global__ void k_sl_svp(const uint32_t* __restrict__ perm, size_t n, const double* __restrict__ sv, int nsv, int s0,
                         int ns, double* __restrict__ svp) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * ns; e += (size_t)gridDim.x * blockDim.x) {
        const size_t k = e / ns;
        svp[e] = sv[(size_t)perm[k] * nsv + s0 + e % ns];
    }
}

// ------------------------------------------------------------------ weight kernels

// M_s^+ of every level of E (slopes.rs:551-622): with an intercept, the centred slope block is
// swept (pivots <= 1e-12 of sum w x^2 dropped, so a level without within variation gets slope 0
// and the mean as intercept) and bordered; without, M_s itself is swept. Levels without weight
// get 0 (their rows contribute nothing). NS = E's slopes.
template <int NS>
__global__ void k_sl_minv(const uint32_t* __restrict__ seg_off, uint64_t ge, const double* __restrict__ he,
                          int e_icpt, const double* __restrict__ w, double* __restrict__ minv) {
    constexpr int G2 = NS > 0 ? NS : 1;
    const int be = e_icpt + NS;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ge; s += (uint64_t)gridDim.x * blockDim.x) {
        const size_t i0 = seg_off[s], i1 = seg_off[s + 1];
        double* out = minv + s * be * be;
        double G[G2 * G2], xbar[G2], refd[G2];
#pragma unroll
        for (int k = 0; k < G2 * G2; k++) G[k] = 0.0;
#pragma unroll
        for (int k = 0; k < G2; k++) {
            xbar[k] = 0.0;
            refd[k] = 0.0;
        }
        for (int k = 0; k < be * be; k++) out[k] = 0.0;
        if (e_icpt) {
            double sw = 0.0;
            for (size_t i = i0; i < i1; i++) {
                const double wi = w ? w[i] : 1.0;
                sw += wi;
#pragma unroll
                for (int k = 0; k < NS; k++) {
                    const double x = he[i * NS + k];
                    xbar[k] += wi * x;
                    refd[k] += wi * x * x;
                }
            }
            if (!(sw > 0.0)) continue;
#pragma unroll
            for (int k = 0; k < NS; k++) xbar[k] /= sw;
            for (size_t i = i0; i < i1; i++) {
                const double wi = w ? w[i] : 1.0;
#pragma unroll
                for (int k = 0; k < NS; k++) {
                    const double dk = wi * (he[i * NS + k] - xbar[k]);
#pragma unroll
                    for (int l = 0; l < NS; l++) G[k * G2 + l] += dk * (he[i * NS + l] - xbar[l]);
                }
            }
            if (NS > 0) sweep_ginv<G2>(G, refd, LEVEL_TOL);
            double q = 0.0;
#pragma unroll
            for (int k = 0; k < NS; k++) {
                double gx = 0.0;
#pragma unroll
                for (int l = 0; l < NS; l++) gx += G[k * G2 + l] * xbar[l];
                q += xbar[k] * gx;
                out[1 + k] = -gx;
                out[(1 + k) * be] = -gx;
#pragma unroll
                for (int l = 0; l < NS; l++) out[(1 + k) * be + 1 + l] = G[k * G2 + l];
            }
            out[0] = 1.0 / sw + q;
        } else {
            for (size_t i = i0; i < i1; i++) {
                const double wi = w ? w[i] : 1.0;
#pragma unroll
                for (int k = 0; k < NS; k++)
#pragma unroll
                    for (int l = 0; l < NS; l++) G[k * G2 + l] += wi * he[i * NS + k] * he[i * NS + l];
            }
#pragma unroll
            for (int k = 0; k < NS; k++) refd[k] = G[k * G2 + k];
            if (NS > 0) sweep_ginv<G2>(G, refd, LEVEL_TOL);
#pragma unroll
            for (int k = 0; k < NS; k++)
#pragma unroll
                for (int l = 0; l < NS; l++) out[k * be + l] = G[k * G2 + l];
        }
    }
}

struct DiagArgs {
    const uint32_t* perm;
    const uint32_t* poff;
    uint64_t g, off;
    int b, icpt, ns;
    const double* svp;  // the block's slope values in level-list order
    const double* he;
    int e_ns, e_icpt, be;
    const uint32_t* row_seg;
    const double* minv;
    const double* w;
    int jacobi;    // write the block-Jacobi inverses
    double* jinv;  // this block's inverses (g x b^2)
};

// Over the level-list positions [k, k1): gl += sum w h h' and cmc += sum_s C_s' M_s^+ C_s over the
// runs of equal E level (C_s = sum w h_E h'), runs taken whole (the caller's range holds whole runs).
// B, BE: the block's and E's widths.
template <int B, int BE>
__device__ void diag_accumulate(const DiagArgs& A, size_t k, size_t k1, double (&gl)[B * B], double (&dl)[B * B]) {
    while (k < k1) {
        const uint32_t s = A.row_seg[A.perm[k]];
        double C[BE][B];
#pragma unroll
        for (int e = 0; e < BE; e++)
#pragma unroll
            for (int p = 0; p < B; p++) C[e][p] = 0.0;
        for (; k < k1 && A.row_seg[A.perm[k]] == s; k++) {
            const size_t i = A.perm[k];
            const double wi = A.w ? A.w[i] : 1.0;
            double hq[B], he[BE];
#pragma unroll
            for (int d = 0; d < B; d++)
                hq[d] = A.icpt ? (d == 0 ? 1.0 : A.svp[k * A.ns + d - 1]) : A.svp[k * A.ns + d];
#pragma unroll
            for (int e = 0; e < BE; e++) he[e] = (A.e_icpt && e == 0) ? 1.0 : A.he[i * A.e_ns + e - A.e_icpt];
#pragma unroll
            for (int p = 0; p < B; p++)
#pragma unroll
                for (int q = 0; q < B; q++) gl[p * B + q] += wi * hq[p] * hq[q];
#pragma unroll
            for (int e = 0; e < BE; e++)
#pragma unroll
                for (int p = 0; p < B; p++) C[e][p] += wi * he[e] * hq[p];
        }
        const double* mv = A.minv + (size_t)s * BE * BE;
        double M[BE][BE];
#pragma unroll
        for (int e = 0; e < BE; e++)
#pragma unroll
            for (int f = 0; f < BE; f++) M[e][f] = mv[e * BE + f];
#pragma unroll
        for (int p = 0; p < B; p++)
#pragma unroll
            for (int q = 0; q < B; q++) {
                double acc = 0.0;
#pragma unroll
                for (int e = 0; e < BE; e++) {
                    double t = 0.0;
#pragma unroll
                    for (int f = 0; f < BE; f++) t += M[e][f] * C[f][q];
                    acc += C[e][p] * t;
                }
                dl[p * B + q] += acc;
            }
    }
}

// The diagonal block of level l from its Gram gl and its sum_s C_s' M_s^+ C_s (in dl): D_l =
// gl - dl, the redundancy flags, the Gram diagonal and (jacobi) the generalized inverse of D_l
// with the redundant coordinates left out.
template <int B>
__device__ void diag_finish(const DiagArgs& A, uint64_t l, const double (&gl)[B * B], double (&dl)[B * B],
                            double* __restrict__ gdiag, uint8_t* __restrict__ red) {
    bool rd[B];
    double refd[B];
#pragma unroll
    for (int q = 0; q < B * B; q++) dl[q] = gl[q] - dl[q];
#pragma unroll
    for (int p = 0; p < B; p++) {
        rd[p] = !(dl[p * B + p] > REDUNDANT_REL * gl[p * B + p]);
        gdiag[A.off + l * B + p] = gl[p * B + p];
        red[A.off + l * B + p] = rd[p] ? 1 : 0;
        refd[p] = dl[p * B + p];
    }
    if (!A.jacobi) return;
#pragma unroll
    for (int p = 0; p < B; p++)
        if (rd[p])
#pragma unroll
            for (int q = 0; q < B; q++) {
                dl[p * B + q] = 0.0;
                dl[q * B + p] = 0.0;
            }
    sweep_ginv<B>(dl, refd, JACOBI_TOL);
#pragma unroll
    for (int q = 0; q < B * B; q++) A.jinv[l * B * B + q] = dl[q];
}

// Per level of one block with short levels (slopes.rs:624-686, 787-834): one thread per level.
template <int B, int BE>
__global__ void k_sl_diag(DiagArgs A, double* __restrict__ gdiag, uint8_t* __restrict__ red) {
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l < A.g; l += (uint64_t)gridDim.x * blockDim.x) {
        double gl[B * B], dl[B * B];
#pragma unroll
        for (int q = 0; q < B * B; q++) {
            gl[q] = 0.0;
            dl[q] = 0.0;
        }
        diag_accumulate<B, BE>(A, A.poff[l], A.poff[l + 1], gl, dl);
        diag_finish<B>(A, l, gl, dl, gdiag, red);
    }
}

// Long levels: one thread per item. An item owns the runs of equal E level that start inside its
// nominal range: it skips the run continuing from the previous item and finishes its last run past
// its nominal end. part[it] = [gl | cmc] (2 B^2).
template <int B, int BE>
__global__ void k_sl_diag_items(DiagArgs A, ItemArgs I, double* __restrict__ part) {
    constexpr int BB = B * B;
    for (uint64_t it = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; it < I.nitems;
         it += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t l = I.lvl[it];
        const size_t lo = A.poff[l], hi = A.poff[l + 1];
        size_t k0 = I.first[it], k1 = I.end[it];
        if (k0 > lo) {
            const uint32_t sp = A.row_seg[A.perm[k0 - 1]];
            while (k0 < hi && A.row_seg[A.perm[k0]] == sp) k0++;
        }
        if (k1 < hi && k1 > lo) {
            const uint32_t sl = A.row_seg[A.perm[k1 - 1]];
            while (k1 < hi && A.row_seg[A.perm[k1]] == sl) k1++;
        }
        double gl[BB], dl[BB];
#pragma unroll
        for (int q = 0; q < BB; q++) {
            gl[q] = 0.0;
            dl[q] = 0.0;
        }
        if (k0 < k1) diag_accumulate<B, BE>(A, k0, k1, gl, dl);
#pragma unroll
        for (int q = 0; q < BB; q++) {
            part[it * 2 * BB + q] = gl[q];
            part[it * 2 * BB + BB + q] = dl[q];
        }
    }
}

template <int B>
__global__ void k_sl_diag_finish(DiagArgs A, ItemArgs I, const double* __restrict__ part, double* __restrict__ gdiag,
                                 uint8_t* __restrict__ red) {
    constexpr int BB = B * B;
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l < A.g; l += (uint64_t)gridDim.x * blockDim.x) {
        double gl[BB], dl[BB];
#pragma unroll
        for (int q = 0; q < BB; q++) {
            gl[q] = 0.0;
            dl[q] = 0.0;
        }
        for (uint32_t it = I.lvl_it[l]; it < I.lvl_it[l + 1]; it++)
#pragma unroll
            for (int q = 0; q < BB; q++) {
                gl[q] += part[(uint64_t)it * 2 * BB + q];
                dl[q] += part[(uint64_t)it * 2 * BB + BB + q];
            }
        diag_finish<B>(A, l, gl, dl, gdiag, red);
    }
}

__global__ void k_sl_units(double* __restrict__ P, size_t m, int r, uint64_t k0) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m * r; e += (size_t)gridDim.x * blockDim.x) {
        const size_t k = e / r;
        const int c = (int)(e % r);
        P[e] = k == k0 + (uint64_t)c ? 1.0 : 0.0;
    }
}

__global__ void k_sl_gather_w(const double* __restrict__ wk, const uint32_t* __restrict__ pi, size_t n,
                              double* __restrict__ wr) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        wr[i] = wk[pi[i]];
}

__global__ void k_sl_zero_cols(double* __restrict__ x, size_t m, int r, uint32_t mask) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m * r; e += (size_t)gridDim.x * blockDim.x)
        if ((mask >> (e % r)) & 1u) x[e] = 0.0;
}

// out[i] = D_R x + h_E' alpha_s for column col of coefs ((m + ge be) x r).
__global__ void k_sl_fit(SegArgs A, const uint32_t* __restrict__ row_seg, size_t n, const double* __restrict__ coefs,
                         int r, int col, uint64_t m, double* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int b = 0; b < A.nb; b++) {
            const size_t k0 = A.coord[i * A.nb + b];
            for (int d = 0; d < A.bw[b]; d++) v += slot_val(A, b, d, i) * coefs[(k0 + d) * r + col];
        }
        double h[SL_BMAX];
        e_row(A.he, A.e_ns, A.e_icpt, i, h);
        const size_t s = row_seg[i];
        for (int e = 0; e < A.be; e++) v += h[e] * coefs[(m + s * A.be + e) * r + col];
        out[i] = v;
    }
}

// Sum of the weights per level of one block (unit weights: the row counts).
__global__ void k_sl_level_w(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ poff, uint64_t g,
                             const double* __restrict__ w, double* __restrict__ out) {
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l < g; l += (uint64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (uint32_t k = poff[l]; k < poff[l + 1]; k++) s += w ? w[perm[k]] : 1.0;
        out[l] = s;
    }
}

__global__ void k_sl_sum(const double* __restrict__ v, size_t n, double* __restrict__ part) {
    __shared__ double red[WPB];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) s += v[i];
    s = wave_sum(s);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < WPB; w++) t += red[w];
        part[blockIdx.x] = t;
    }
}

// ------------------------------------------------------------------ host helpers

SegArgs seg_args(const Absorber& a) {
    const SlopeState& S = *a.slope;
    SegArgs A{};
    A.seg_off = S.seg_off.p;
    A.ge = S.ge;
    A.coord = S.coord.p;
    A.nb = (int)S.blocks.size();
    for (size_t b = 0; b < S.blocks.size(); b++) {
        A.bw[b] = S.blocks[b].b;
        A.icpt[b] = S.blocks[b].icpt;
        A.s0[b] = S.blocks[b].s0;
    }
    A.sv = S.sv.p;
    A.nsv = S.nsv;
    A.he = S.he.p;
    A.e_ns = S.e_ns;
    A.e_icpt = S.e_icpt;
    A.be = S.be;
    A.minv = S.minv.p;
    A.w = a.weighted ? a.w_rows.p : nullptr;
    return A;
}

PullArgs pull_args(const SlopeState& S, const SlBlock& B) {
    PullArgs P{};
    P.perm = B.perm.p;
    P.poff = B.poff.p;
    P.g = B.g;
    P.off = B.off;
    P.b = B.b;
    P.icpt = B.icpt;
    P.ns = B.ns;
    P.svp = B.svp.p;
    return P;
}

// Calls f(integral_constant L, integral_constant W) for lanes L in {4, 16, 64} and width W in 1..4.
template <class F>
void with_lanes_width(int L, int w, F&& f) {
    auto by_w = [&](auto lc) {
        switch (w) {
            case 1: f(lc, std::integral_constant<int, 1>{}); break;
            case 2: f(lc, std::integral_constant<int, 2>{}); break;
            case 3: f(lc, std::integral_constant<int, 3>{}); break;
            default: f(lc, std::integral_constant<int, 4>{}); break;
        }
    };
    if (L == 4) by_w(std::integral_constant<int, 4>{});
    else if (L == 16) by_w(std::integral_constant<int, 16>{});
    else by_w(std::integral_constant<int, 64>{});
}

ItemArgs item_args(const SlBlock& B) {
    return ItemArgs{B.it_first.p, B.it_end.p, B.it_lvl.p, B.lvl_it.p, B.nitems};
}

// The pass over E's segments in one mode: one wave per segment, or, when a segment is longer than
// CH rows, phases A-C over items. M_RHS leaves its per-column [wsq | wsum | ss1] partials in
// *part_out (count *nparts_out, 3 r each) for a fixed-order sum.
template <int MODE>
void seg_pass(const Absorber& a, SlopeState& S, const double* src, int ld, const double* x, int r, double* dst,
              double* alpha, const uint32_t* done, hipStream_t st, DBuf<double>* rhs_part = nullptr,
              size_t* nparts_out = nullptr) {
    const SegArgs A = seg_args(a);
    if (!S.e_chunked) {
        const double avg = (double)a.n / (double)std::max<uint64_t>(S.ge, 1);
        const int L = avg <= 8.0 ? 4 : avg <= 32.0 ? 16 : 64;
        // the right-hand side's per-wave partials are summed afterwards: fewer waves there
        const unsigned grid = grid_for(std::max<uint64_t>(S.ge, 1), WPB * (WAVE / L), MODE == M_RHS ? 4096u : 1u << 15);
        double* part = nullptr;
        if (MODE == M_RHS) {
            const size_t nw = (size_t)grid * WPB;
            rhs_part->reserve(nw * 3 * r);
            part = rhs_part->p;
            *nparts_out = nw;
        }
        with_lanes_width(L, S.be, [&](auto lc, auto bc) {
            constexpr int LL = decltype(lc)::value, BE = decltype(bc)::value;
            hipLaunchKernelGGL((k_sl_seg<MODE, LL, BE>), dim3(grid), dim3(TPB), 0, st, A, src, ld, x, r, dst, alpha,
                               part, done);
        });
        FH_CHECK_LAUNCH();
        return;
    }
    const ItemArgs I{S.e_first.p, S.e_end.p, S.e_lvl.p, S.seg_it.p, S.e_nitems};
    S.epart.reserve(S.e_nitems * (S.be + 2) * r);
    S.tbuf.reserve(S.ge * S.be * r);
    double* segpart = nullptr;
    if (MODE == M_RHS) {
        rhs_part->reserve(S.ge * 3 * r);
        segpart = rhs_part->p;
        *nparts_out = S.ge;
    }
    hipLaunchKernelGGL((k_sl_segA<MODE>), dim3(grid_for(S.e_nitems, WPB, 1u << 15)), dim3(TPB), 0, st, A, I, src, ld,
                       x, r, dst, S.epart.p, done);
    hipLaunchKernelGGL((k_sl_segB<MODE>), dim3(grid_for(S.ge * r, TPB, 8192)), dim3(TPB), 0, st, A, I, S.epart.p, r,
                       S.tbuf.p, alpha, segpart, done);
    hipLaunchKernelGGL((k_sl_segC<MODE>), dim3(grid_for(a.n, TPB, 8192)), dim3(TPB), 0, st, A, S.row_seg.p, a.n, src,
                       ld, r, S.tbuf.p, dst, done);
    FH_CHECK_LAUNCH();
}

// E's segments in items of <= CH rows when the longest exceeds CH.
void build_e_items(SlopeState& S, const std::vector<uint32_t>& hseg, hipStream_t st) {
    uint32_t maxlen = 0;
    for (uint64_t s = 0; s < S.ge; s++) maxlen = std::max(maxlen, hseg[s + 1] - hseg[s]);
    S.e_chunked = maxlen > CH;
    if (!S.e_chunked) return;
    std::vector<uint32_t> first, end, lvl, seg_it(S.ge + 1);
    for (uint64_t s = 0; s < S.ge; s++) {
        seg_it[s] = (uint32_t)first.size();
        const uint32_t len = hseg[s + 1] - hseg[s];
        const uint32_t nch = std::max<uint32_t>(1, (len + CH - 1) / CH);
        for (uint32_t c = 0; c < nch; c++) {
            first.push_back(hseg[s] + c * CH);
            end.push_back(std::min(hseg[s + 1], hseg[s] + (c + 1) * CH));
            lvl.push_back((uint32_t)s);
        }
    }
    seg_it[S.ge] = (uint32_t)first.size();
    S.e_nitems = first.size();
    S.e_first.alloc(S.e_nitems);
    S.e_end.alloc(S.e_nitems);
    S.e_lvl.alloc(S.e_nitems);
    S.seg_it.alloc(S.ge + 1);
    h2d(S.e_first.p, first.data(), S.e_nitems, st);
    h2d(S.e_end.p, end.data(), S.e_nitems, st);
    h2d(S.e_lvl.p, lvl.data(), S.e_nitems, st);
    h2d(S.seg_it.p, seg_it.data(), S.ge + 1, st);
    stream_wait(st);
}

// Y rows of block B (b x r per level) = the pull of Z over its levels.
void pull_block(SlopeState& S, const SlBlock& B, const double* Z, int r, double* Y, const uint32_t* done,
                hipStream_t st) {
    if (!B.chunked) {
        // lanes per level from the mean level length
        const double avg = (double)B.poff_n / (double)std::max<uint64_t>(B.g, 1);
        const int L = avg <= 8.0 ? 4 : avg <= 32.0 ? 16 : 64;
        const unsigned grid = grid_for(std::max<uint64_t>(B.g, 1), WPB * (WAVE / L), 1u << 15);
        with_lanes_width(L, B.b, [&](auto lc, auto bc) {
            constexpr int LL = decltype(lc)::value, BB = decltype(bc)::value;
            hipLaunchKernelGGL((k_sl_pull<LL, BB>), dim3(grid), dim3(TPB), 0, st, pull_args(S, B), Z, r, Y, done);
        });
        return;
    }
    S.ipart.reserve(B.nitems * B.b * r);
    hipLaunchKernelGGL(k_sl_pull_items, dim3(grid_for(B.nitems, WPB, 1u << 15)), dim3(TPB), 0, st, pull_args(S, B),
                       item_args(B), Z, r, S.ipart.p, done);
    hipLaunchKernelGGL(k_sl_item_sum, dim3((unsigned)std::min<uint64_t>(std::max<uint64_t>(B.g, 1), 8192)), dim3(TPB), 0,
                       st, B.lvl_it.p, B.g, S.ipart.p, B.b * r, Y + B.off * r, done);
}

// Items of <= CH rows over the level lists of a block whose longest level exceeds CH.
void build_items(SlBlock& B, hipStream_t st) {
    std::vector<uint32_t> off(B.g + 1);
    d2h(off.data(), B.poff.p, B.g + 1, st);
    stream_wait(st);
    B.poff_n = off[B.g];
    uint32_t maxlen = 0;
    for (uint64_t l = 0; l < B.g; l++) maxlen = std::max(maxlen, off[l + 1] - off[l]);
    B.chunked = maxlen > CH;
    if (!B.chunked) return;
    std::vector<uint32_t> first, end, lvl, lvl_it(B.g + 1);
    for (uint64_t l = 0; l < B.g; l++) {
        lvl_it[l] = (uint32_t)first.size();
        const uint32_t len = off[l + 1] - off[l];
        const uint32_t nch = std::max<uint32_t>(1, (len + CH - 1) / CH);
        for (uint32_t c = 0; c < nch; c++) {
            first.push_back(off[l] + c * CH);
            end.push_back(std::min(off[l + 1], off[l] + (c + 1) * CH));
            lvl.push_back((uint32_t)l);
        }
    }
    lvl_it[B.g] = (uint32_t)first.size();
    B.nitems = first.size();
    B.it_first.alloc(B.nitems);
    B.it_end.alloc(B.nitems);
    B.it_lvl.alloc(B.nitems);
    B.lvl_it.alloc(B.g + 1);
    h2d(B.it_first.p, first.data(), B.nitems, st);
    h2d(B.it_end.p, end.data(), B.nitems, st);
    h2d(B.it_lvl.p, lvl.data(), B.nitems, st);
    h2d(B.lvl_it.p, lvl_it.data(), B.g + 1, st);
}

// Rows (of the index space 0..n-1, codes[idx[i]] when idx) sorted by (code, index): the sorted
// indices and the offsets of every level (g + 1).
void sort_by_code(const uint32_t* codes, const uint32_t* idx, size_t n, uint64_t g, DBuf<uint32_t>& order,
                  DBuf<uint32_t>& off, hipStream_t st) {
    DBuf<uint64_t> keys(n), ka(n);
    DBuf<uint32_t> vals(n), va(n);
    hipLaunchKernelGGL(k_sl_keys, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, codes, idx, n, keys.p, vals.p);
    FH_CHECK_LAUNCH();
    int bits = 1;
    while (bits < 32 && (1ull << bits) < g) bits++;
    radix_sort_pairs(keys, vals, ka, va, n, 32 + bits, st);
    off.alloc(g + 1);
    hipLaunchKernelGGL(k_sl_offsets, dim3(grid_for(n + 1, TPB, 8192)), dim3(TPB), 0, st, keys.p, n, g, off.p);
    FH_CHECK_LAUNCH();
    order = std::move(vals);
}

}  // namespace

// ------------------------------------------------------------------ entry points

void slope_spmv(const Absorber& a, const double* p, double* y, int r, const uint32_t* done, hipStream_t st,
                double* pq) {
    SlopeState& S = *a.slope;
    require(r >= 1 && r <= 16, "slopes: 1..16 columns");
    S.zbuf.reserve(a.n * r);
    seg_pass<M_APPLY>(a, S, p, 0, nullptr, r, S.zbuf.p, nullptr, done, st);
    for (const SlBlock& B : S.blocks) pull_block(S, B, S.zbuf.p, r, y, done, st);
    if (pq) hipLaunchKernelGGL(k_sl_dot, dim3(DOTG), dim3(TPB), 0, st, p, y, a.m * r, r, pq, done);
    FH_CHECK_LAUNCH();
}

void slope_precond(Absorber& a, const double* rr, double* z, int r, double* rzf_part, const uint32_t* done,
                   hipStream_t st) {
    const SlopeState& S = *a.slope;
    JacArgs J{};
    J.jinv = S.jinv.p;
    for (const SlBlock& B : S.blocks) {
        if (S.dense && B.small) continue;
        J.off[J.nblk] = B.off;
        J.joff[J.nblk] = B.joff;
        J.b[J.nblk] = B.b;
        J.lcum[J.nblk + 1] = J.lcum[J.nblk] + B.g;
        J.nblk++;
    }
    hipLaunchKernelGGL(k_sl_jacobi, dim3(DOTG), dim3(TPB), 0, st, J, rr, z, r, rzf_part, done);
    FH_CHECK_LAUNCH();
}

void slope_refresh(Absorber& a, const double* w_rows) {
    hipStream_t st = a.st;
    SlopeState& S = *a.slope;
    const size_t n = a.n, m = a.m;
    a.weighted = w_rows != nullptr;
    if (a.weighted && w_rows != a.w_rows.p) {
        a.w_rows.reserve(n);
        FH_HIP(hipMemcpyAsync(a.w_rows.p, w_rows, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    const double* w = a.weighted ? a.w_rows.p : nullptr;
    // the weight sum, fixed order
    if (w) {
        DBuf<double> part(DOTG);
        hipLaunchKernelGGL(k_sl_sum, dim3(DOTG), dim3(TPB), 0, st, w, n, part.p);
        FH_CHECK_LAUNCH();
        a.sw_dev.reserve(1);
        sum_parts_to(part.p, DOTG, 1, a.sw_dev.p, st);
        download(&a.sw, a.sw_dev.p, sizeof(double), st);
    } else {
        a.sw = (double)n;
        a.sw_dev.reserve(1);
        upload(a.sw_dev.p, &a.sw, sizeof(double), st);
    }
    a.sw_ready = true;
    // M_s^+ per level of E
    S.minv.reserve(std::max<uint64_t>(S.ge * S.be * S.be, 1));
    {
        const unsigned grid = grid_for(std::max<uint64_t>(S.ge, 1), TPB, 8192);
        switch (S.e_ns) {
            case 0: hipLaunchKernelGGL(k_sl_minv<0>, dim3(grid), dim3(TPB), 0, st, S.seg_off.p, S.ge, S.he.p, S.e_icpt, w, S.minv.p); break;
            case 1: hipLaunchKernelGGL(k_sl_minv<1>, dim3(grid), dim3(TPB), 0, st, S.seg_off.p, S.ge, S.he.p, S.e_icpt, w, S.minv.p); break;
            case 2: hipLaunchKernelGGL(k_sl_minv<2>, dim3(grid), dim3(TPB), 0, st, S.seg_off.p, S.ge, S.he.p, S.e_icpt, w, S.minv.p); break;
            case 3: hipLaunchKernelGGL(k_sl_minv<3>, dim3(grid), dim3(TPB), 0, st, S.seg_off.p, S.ge, S.he.p, S.e_icpt, w, S.minv.p); break;
            default: hipLaunchKernelGGL(k_sl_minv<4>, dim3(grid), dim3(TPB), 0, st, S.seg_off.p, S.ge, S.he.p, S.e_icpt, w, S.minv.p); break;
        }
    }
    FH_CHECK_LAUNCH();
    // diagonal blocks, redundancy, block-Jacobi inverses
    S.gdiag.reserve(std::max<size_t>(m, 1));
    S.red.reserve(std::max<size_t>(m, 1));
    uint64_t jlen = 0;
    for (SlBlock& B : S.blocks) {
        if (S.dense && B.small) continue;
        B.joff = jlen;
        jlen += B.g * B.b * B.b;
    }
    S.jinv.reserve(std::max<uint64_t>(jlen, 1));
    for (SlBlock& B : S.blocks) {
        DiagArgs D{};
        D.perm = B.perm.p;
        D.poff = B.poff.p;
        D.g = B.g;
        D.off = B.off;
        D.b = B.b;
        D.icpt = B.icpt;
        D.ns = B.ns;
        D.svp = B.svp.p;
        D.he = S.he.p;
        D.e_ns = S.e_ns;
        D.e_icpt = S.e_icpt;
        D.be = S.be;
        D.row_seg = S.row_seg.p;
        D.minv = S.minv.p;
        D.w = w;
        D.jacobi = !(S.dense && B.small);
        D.jinv = D.jacobi ? S.jinv.p + B.joff : nullptr;
        // widths as compile-time constants (the lanes argument of the helper is unused here)
        with_lanes_width(4, B.b, [&](auto, auto bc) {
            constexpr int BW = decltype(bc)::value;
            auto by_be = [&](auto ec) {
                constexpr int BE = decltype(ec)::value;
                if (B.chunked) {
                    DBuf<double> part(B.nitems * 2 * BW * BW);
                    hipLaunchKernelGGL((k_sl_diag_items<BW, BE>), dim3(grid_for(B.nitems, TPB, 8192)), dim3(TPB), 0,
                                       st, D, item_args(B), part.p);
                    hipLaunchKernelGGL((k_sl_diag_finish<BW>), dim3(grid_for(std::max<uint64_t>(B.g, 1), TPB, 8192)),
                                       dim3(TPB), 0, st, D, item_args(B), part.p, S.gdiag.p, S.red.p);
                } else {
                    hipLaunchKernelGGL((k_sl_diag<BW, BE>), dim3(grid_for(std::max<uint64_t>(B.g, 1), TPB, 8192)),
                                       dim3(TPB), 0, st, D, S.gdiag.p, S.red.p);
                }
            };
            switch (S.be) {
                case 1: by_be(std::integral_constant<int, 1>{}); break;
                case 2: by_be(std::integral_constant<int, 2>{}); break;
                case 3: by_be(std::integral_constant<int, 3>{}); break;
                default: by_be(std::integral_constant<int, 4>{}); break;
            }
        });
        FH_CHECK_LAUNCH();
    }
    // Jacobi leaves every coordinate to the block inverses (zero inverse diagonal)
    a.w64.inv_diag.reserve(std::max<size_t>(m, 1));
    FH_HIP(hipMemsetAsync(a.w64.inv_diag.p, 0, std::max<size_t>(m, 1) * sizeof(double), st));
    // the dense small block: S_SS by products with unit vectors, its grounded inverse on the host
    const size_t ms = a.ms;
    a.nkeep = 0;
    if (ms > 0) {
        std::vector<double> Sss(ms * ms), gd(ms), col;
        std::vector<uint8_t> rd(ms);
        DBuf<double> P, Y;
        for (size_t k0 = 0; k0 < ms; k0 += 16) {
            const int rc = (int)std::min<size_t>(16, ms - k0);
            P.reserve(m * rc);
            Y.reserve(m * rc);
            hipLaunchKernelGGL(k_sl_units, dim3(grid_for(m * rc, TPB, 4096)), dim3(TPB), 0, st, P.p, m, rc,
                               (uint64_t)k0);
            FH_CHECK_LAUNCH();
            slope_spmv(a, P.p, Y.p, rc, nullptr, st, nullptr);
            col.resize(ms * rc);
            d2h(col.data(), Y.p, ms * rc, st);
            stream_wait(st);
            for (size_t k = 0; k < ms; k++)
                for (int c = 0; c < rc; c++) Sss[k * ms + k0 + c] = col[k * rc + c];
        }
        d2h(gd.data(), S.gdiag.p, ms, st);
        stream_wait(st);
        for (size_t k = 0; k < ms; k++) rd[k] = !(Sss[k * ms + k] > REDUNDANT_REL * gd[k]);
        std::vector<double> inv = grounded_inverse(Sss, ms, rd, a.nkeep);
        a.w64.minv.reserve(ms * ms);
        h2d(a.w64.minv.p, inv.data(), ms * ms, st);
        std::vector<uint8_t> rdu(rd.begin(), rd.end());
        h2d(S.red.p, rdu.data(), ms, st);
        stream_wait(st);
    }
}

void slope_create(Absorber& a, const std::vector<SlopeTerm>& terms, const double* w_kept) {
    hipStream_t st = a.st;
    const size_t n = a.n;
    const size_t nt = terms.size();
    require(nt >= 1, "slopes: no absorbed terms");
    size_t ntot_sl = 0;
    for (const SlopeTerm& t : terms) {
        require(t.icpt || !t.slopes.empty(), "absorb term without intercept or slopes");
        require((t.icpt ? 1 : 0) + t.slopes.size() <= (size_t)SL_BMAX,
                "slopes: at most " + std::to_string(SL_BMAX) + " coordinates per level (intercept + slopes)");
        ntot_sl += t.slopes.size();
        for (const double* x : t.slopes) require(x != nullptr, "null slope column");
    }
    require(nt - 1 <= (size_t)NB_MAX, "slopes: at most " + std::to_string(NB_MAX + 1) + " absorbed terms");
    auto S = std::make_shared<SlopeState>();
    a.slope = S;
    // the slope columns are copied (the caller's arrays need not outlive the absorber)
    std::vector<std::vector<const double*>> own(nt);
    for (size_t q = 0; q < nt; q++)
        for (const double* x : terms[q].slopes) {
            S->slope_cols.emplace_back(n);
            FH_HIP(hipMemcpyAsync(S->slope_cols.back().p, x, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            own[q].push_back(S->slope_cols.back().p);
        }
    // E (absorber.rs:100-108; slopes.rs:364-372)
    int64_t fe1 = -1;
    for (size_t q = 0; q < nt; q++)
        if (terms[q].icpt && terms[q].slopes.empty() && (fe1 < 0 || terms[q].g > terms[fe1].g)) fe1 = (int64_t)q;
    size_t e = 0;
    if (fe1 >= 0) {
        e = (size_t)fe1;
    } else {
        for (size_t q = 1; q < nt; q++)
            if (terms[q].g > terms[e].g) e = q;
    }
    const SlopeTerm& te = terms[e];
    S->e = (uint32_t)e;
    S->ge = te.g;
    S->e_icpt = te.icpt ? 1 : 0;
    S->e_ns = (int)te.slopes.size();
    S->be = S->e_icpt + S->e_ns;
    // rows sorted by (level of E, observation)
    {
        DBuf<uint32_t> order;
        sort_by_code(te.codes, nullptr, n, te.g, order, S->seg_off, st);
        a.pi = std::move(order);
    }
    S->row_seg.alloc(n);
    hipLaunchKernelGGL(k_sl_row_seg, dim3(grid_for(std::max<uint64_t>(S->ge, 1), TPB, 8192)), dim3(TPB), 0, st,
                       S->seg_off.p, S->ge, S->row_seg.p);
    FH_CHECK_LAUNCH();
    // the small set (slopes.rs:444-460): by coordinates, while within small_limit and the dense
    // factor's modelled cost
    std::vector<uint32_t> hseg(S->ge + 1);
    d2h(hseg.data(), S->seg_off.p, S->ge + 1, st);
    stream_wait(st);
    build_e_items(*S, hseg, st);
    auto width = [&](size_t q) { return (size_t)(terms[q].icpt ? 1 : 0) + terms[q].slopes.size(); };
    std::vector<size_t> others;
    for (size_t q = 0; q < nt; q++)
        if (q != e) others.push_back(q);
    size_t slots = (size_t)S->be;
    for (size_t q : others) slots += width(q);
    const double matvec = C_ROW * (double)n * (double)(slots + 2);
    auto small_cost = [&](size_t msz, size_t bs) {
        double work = (double)n * (double)(bs * bs);
        for (uint64_t s = 0; s < S->ge; s++) {
            const double d = std::min<double>((double)(hseg[s + 1] - hseg[s]) * (double)bs, (double)msz);
            work += d * d;
        }
        return C_SMALL_ASM * work + C_FLOP * std::pow((double)msz, 3) / 3.0;
    };
    std::vector<size_t> by_size = others;
    std::stable_sort(by_size.begin(), by_size.end(), [&](size_t x, size_t y) {
        const uint64_t cx = terms[x].g * width(x), cy = terms[y].g * width(y);
        return cx != cy ? cx < cy : x < y;
    });
    std::vector<char> small(nt, 0);
    size_t total = 0, bs = 0;
    for (size_t q : by_size) {
        const size_t c = terms[q].g * width(q);
        if (total + c > a.small_limit || small_cost(total + c, bs + width(q)) > std::max(SMALL_FLOOR, SMALL_ITERS * matvec))
            break;
        total += c;
        bs += width(q);
        small[q] = 1;
    }
    std::vector<size_t> order;
    for (size_t q : others)
        if (small[q]) order.push_back(q);
    S->nsmall = (int)order.size();
    for (size_t q : others)
        if (!small[q]) order.push_back(q);
    uint64_t off = 0;
    int s0 = 0;
    for (size_t q : order) {
        SlBlock B{};
        B.term = (uint32_t)q;
        B.g = terms[q].g;
        B.icpt = terms[q].icpt ? 1 : 0;
        B.ns = (int)terms[q].slopes.size();
        B.b = B.icpt + B.ns;
        B.off = off;
        B.s0 = s0;
        B.small = small[q] != 0;
        off += B.g * B.b;
        s0 += B.ns;
        S->blocks.push_back(std::move(B));
    }
    require(off < (1ull << 32), "too many fixed-effect coefficients");
    S->nsv = s0;
    a.m = off;
    S->m_small = 0;
    for (const SlBlock& B : S->blocks)
        if (B.small) S->m_small += B.g * B.b;
    S->dense = S->m_small > 0 && S->m_small <= a.dense_limit;
    a.ms = S->dense ? S->m_small : 0;
    a.g1 = S->ge * S->be;
    a.nb = 0;  // the cell-structure kernels have nothing to do on a slope absorber
    a.M = n;
    // per-row arrays
    RowSrc R{};
    R.nb = (int)S->blocks.size();
    R.nsv = S->nsv;
    R.e_ns = S->e_ns;
    for (size_t b = 0; b < S->blocks.size(); b++) {
        const SlBlock& B = S->blocks[b];
        R.codes[b] = terms[B.term].codes;
        R.off[b] = B.off;
        R.b[b] = B.b;
        for (int k = 0; k < B.ns; k++) R.sl[B.s0 + k] = own[B.term][k];
    }
    for (int k = 0; k < S->e_ns; k++) R.hel[k] = own[e][k];
    S->coord.alloc(std::max<size_t>(n * S->blocks.size(), 1));
    S->sv.alloc(std::max<size_t>(n * (size_t)S->nsv, 1));
    S->he.alloc(std::max<size_t>(n * (size_t)S->e_ns, 1));
    hipLaunchKernelGGL(k_sl_rows, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, R, a.pi.p, n, S->coord.p, S->sv.p,
                       S->he.p);
    FH_CHECK_LAUNCH();
    // per block: rows by (level, row)
    for (SlBlock& B : S->blocks) {
        sort_by_code(terms[B.term].codes, a.pi.p, n, B.g, B.perm, B.poff, st);
        build_items(B, st);
        B.svp.alloc(std::max<size_t>(n * B.ns, 1));
        if (B.ns > 0) {
            hipLaunchKernelGGL(k_sl_svp, dim3(grid_for(n * B.ns, TPB, 8192)), dim3(TPB), 0, st, B.perm.p, n, S->sv.p,
                               S->nsv, B.s0, B.ns, B.svp.p);
            FH_CHECK_LAUNCH();
        }
    }
    // the kept-order slope columns are no longer needed once in row order
    S->slope_cols.clear();
    a.fe1 = fe1;
    a.norm_fe1 = fe1;
    a.precision = FH_PLAIN_FP64;
    DBuf<double> wr;
    if (w_kept) {
        wr.alloc(n);
        hipLaunchKernelGGL(k_sl_gather_w, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, w_kept, a.pi.p, n, wr.p);
        FH_CHECK_LAUNCH();
    }
    slope_refresh(a, w_kept ? wr.p : nullptr);
}

ResidOut slope_residualize(Absorber& a, double* z, int r, double tol, uint64_t maxiter, const double* warm,
                           double* coefs, const fh_ctx* ctx) {
    hipStream_t st = a.st;
    SlopeState& S = *a.slope;
    require(r >= 1 && r <= 16, "slopes: 1..16 columns per solve");
    const size_t n = a.n, m = a.m;
    ResidOut out;
    out.exact = false;  // the slope solver reports no exact solves (slopes.rs:2419)
    S.zbuf.reserve(n * r);
    a.s64.rhs.reserve(std::max<size_t>(m * r, 1));
    a.s64.x.reserve(std::max<size_t>(m * r, 1));
    DBuf<double> part, sums(3 * (size_t)r);
    size_t nparts = 0;
    seg_pass<M_RHS>(a, S, z, r, nullptr, r, S.zbuf.p, nullptr, nullptr, st, &part, &nparts);
    for (const SlBlock& B : S.blocks) pull_block(S, B, S.zbuf.p, r, a.s64.rhs.p, nullptr, st);
    FH_CHECK_LAUNCH();
    sum_parts_to(part.p, nparts, 3 * (size_t)r, sums.p, st);
    std::vector<double> hs(3 * (size_t)r);
    d2h(hs.data(), sums.p, hs.size(), st);
    stream_wait(st);
    check_interrupt(ctx);
    // ss0 = wsq - wsum^2/sw (0 if <= 1e-15 wsq), ss1 = sum w (v - P_E v)^2 (solve.rs:1210-1217)
    std::vector<double> ss0(r), ss1(r);
    uint32_t zmask = 0;
    bool allconst = true;
    for (int j = 0; j < r; j++) {
        const double wsq = hs[j], wsum = hs[r + j];
        const double v = wsq - wsum * wsum / a.sw;
        ss0[j] = v > 1e-15 * wsq ? v : 0.0;
        ss1[j] = hs[2 * r + j];
        if (ss0[j] == 0.0) zmask |= 1u << j;
        allconst &= ss0[j] == 0.0;
    }
    double* x = a.s64.x.p;
    if (m > 0) {
        if (warm) FH_HIP(hipMemcpyAsync(x, warm, m * r * sizeof(double), hipMemcpyDeviceToDevice, st));
        else FH_HIP(hipMemsetAsync(x, 0, m * r * sizeof(double), st));
        if (zmask) {
            hipLaunchKernelGGL(k_sl_zero_cols, dim3(grid_for(m * r, TPB, 4096)), dim3(TPB), 0, st, x, m, r, zmask);
            FH_CHECK_LAUNCH();
        }
        PcgResult pr = pcg_fp64(a, a.s64.rhs.p, x, warm != nullptr, r, ss1.data(), ss0.data(), tol, maxiter, ctx);
        if (pr.rc == FH_CONVERGENCE)
            throw Error(FH_CONVERGENCE, "the FE solver did not converge within " + std::to_string(maxiter) + " iterations");
        out.iterations = std::max<uint64_t>(pr.iterations, 1);
    }
    (void)allconst;
    seg_pass<M_RESID>(a, S, z, r, m > 0 ? x : nullptr, r, z, coefs ? coefs + m * r : nullptr, nullptr, st);
    if (coefs && m > 0) FH_HIP(hipMemcpyAsync(coefs, x, m * r * sizeof(double), hipMemcpyDeviceToDevice, st));
    return out;
}

void slope_fit_rows(Absorber& a, const double* coefs, int r, int col, double* out) {
    const SlopeState& S = *a.slope;
    hipLaunchKernelGGL(k_sl_fit, dim3(grid_for(a.n, TPB, 8192)), dim3(TPB), 0, a.st, seg_args(a), S.row_seg.p, a.n,
                       coefs, r, col, (uint64_t)a.m, out);
    FH_CHECK_LAUNCH();
}

std::vector<std::vector<double>> slope_term_values(Absorber& a, const double* coefs, size_t r, size_t col) {
    const SlopeState& S = *a.slope;
    const size_t nt = a.nterms, m = a.m;
    std::vector<std::vector<std::vector<double>>> vals(nt);
    vals[S.e].assign(S.be, std::vector<double>(S.ge));
    for (int k = 0; k < S.be; k++)
        for (uint64_t g = 0; g < S.ge; g++) vals[S.e][k][g] = coefs[(m + g * S.be + k) * r + col];
    for (const SlBlock& B : S.blocks) {
        vals[B.term].assign(B.b, std::vector<double>(B.g));
        for (int k = 0; k < B.b; k++)
            for (uint64_t l = 0; l < B.g; l++) vals[B.term][k][l] = coefs[(B.off + l * B.b + k) * r + col];
    }
    // normalization (slopes.rs:2447-2462): every intercept-only term but FE1 gets weighted mean 0,
    // FE1's intercept takes the shifts
    if (a.fe1 >= 0) {
        const double* w = a.weighted ? a.w_rows.p : nullptr;
        for (const SlBlock& B : S.blocks) {
            if (!B.icpt || B.ns > 0) continue;
            DBuf<double> lw(std::max<uint64_t>(B.g, 1));
            hipLaunchKernelGGL(k_sl_level_w, dim3(grid_for(std::max<uint64_t>(B.g, 1), TPB, 8192)), dim3(TPB), 0, a.st,
                               B.perm.p, B.poff.p, B.g, w, lw.p);
            FH_CHECK_LAUNCH();
            std::vector<double> h(B.g);
            d2h(h.data(), lw.p, B.g, a.st);
            stream_wait(a.st);
            double mean = 0.0;
            for (uint64_t l = 0; l < B.g; l++) mean += vals[B.term][0][l] * h[l];
            mean /= a.sw;
            for (double& v : vals[B.term][0]) v -= mean;
            for (double& v : vals[a.fe1][0]) v += mean;
        }
    }
    std::vector<std::vector<double>> flat;
    for (auto& t : vals)
        for (auto& s : t) flat.push_back(std::move(s));
    return flat;
}

void slope_blocks(const Absorber& a, int64_t* term, uint64_t* g, uint64_t* off, int32_t* small) {
    const SlopeState& S = *a.slope;
    for (size_t b = 0; b < S.blocks.size(); b++) {
        if (term) term[b] = S.blocks[b].term;
        if (g) g[b] = S.blocks[b].g;
        if (off) off[b] = S.blocks[b].off;
        if (small) small[b] = S.blocks[b].small ? 1 : 0;
    }
}

void slope_summary(const Absorber& a, int64_t* e, uint64_t* ge, uint64_t* be, uint64_t* m_small, uint64_t* nb) {
    const SlopeState& S = *a.slope;
    if (e) *e = S.e;
    if (ge) *ge = S.ge;
    if (be) *be = (uint64_t)S.be;
    if (m_small) *m_small = S.m_small;
    if (nb) *nb = S.blocks.size();
}

}  // namespace fh
