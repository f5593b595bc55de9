// The escalated preconditioner of the FE solver on the GPU (solve.rs:685-891, precond/mod.rs:74-185):
// exact sparse factors of the joint reduced matrix S (every coordinate; M = S, so a restarted
// solve is exact) or, when the joint factor exceeds its budgets, of each large block S_bb
// (block-Jacobi with exact blocks beside the dense small factor).
//
//  - Assembly on the GPU: S = sum_s S_s with, per FE1 level s, the ordered item pairs (t, u) of
//    its cells' coordinates, entry (k_t, k_u) += [cell t == cell u] W_c - W_t W_u / n1_s
//    (operator.rs:1829-2054 assembles the same matrix from segment aggregates on the CPU). One
//    wave per segment emits its lower-triangle triples; a radix sort by (row, column) and a
//    fixed-order sum of each run (emission order is stable) give the CSR, deterministically.
//  - The factor is built on the host (fh_ldl.hip), as the reference builds its own.
//  - The apply: z = s P^T L^-T D^-1 L^-1 P s r by level-scheduled triangular solves. A factor
//    whose vector fits in LDS is solved by one workgroup with the vector staged in LDS (forward
//    levels, then backward levels, a barrier between levels, the r'z partial at the end: one
//    launch). Larger factors run one launch per wide level and one single-workgroup launch per run
//    of narrow levels on a vector in HBM. Rows are processed one wave each, lanes over the row's
//    entries, R columns per lane.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "fh_fit.hpp"
#include "fh_internal.hpp"
#include "fh_ldl.hpp"
#include "fh_sort.hpp"

namespace fh {

namespace {

constexpr int TPB = 256;
constexpr int FAC_RMAX = 16;
constexpr int FAC_LDS_THREADS = 1024;
constexpr size_t FAC_LDS_BYTES = 60 * 1024;  // the LDS vector of the one-workgroup solve (+ 2 KB of partials)
constexpr uint32_t FAC_WIDE = 128;           // rows from which a level gets its own launch

// Budgets of an escalated factor (solve.rs:31-36 sets the reference's; these bound GPU memory
// for the assembly triples and the host factorization's time).
constexpr uint64_t ESC_MAX_PAIRS = 400'000'000;  // lower-triangle item pairs emitted by the assembly
constexpr uint64_t ESC_MAX_NNZ = 30'000'000;     // structural nonzeros of the assembled matrix
constexpr uint64_t ESC_MAX_LEN = 100'000'000;    // factor entries
constexpr double ESC_MAX_FLOPS = 2e10;           // multiply-adds of the host factorization

// Cost model of a factor build here (seconds; the reference's is solve.rs:82-121 on its CPU):
// GPU assembly per emitted item pair (count, emit, radix sort, run sums), per structural nonzero
// (download, the host CSR, the ordering and the symbolic pass), per multiply-add of the host
// numeric factorization and per factor entry (transposes, level schedules, upload).
// Measured on MI355X with its host (rings and coupled rings of 1.5e3 to 1.5e5 coordinates): the
// assembly 27 ms for 1e6 pairs, the analysis 12-30 ms for 0.44-0.93 M nonzeros, the numeric pass
// and the upload 3-10 ms each for 0.5-1.1 M entries.
constexpr double C_PAIR = 3e-8;
constexpr double C_NNZ = 4e-8;
constexpr double C_FLOP = 1.5e-9;
constexpr double C_ENTRY = 2e-8;

struct AsmArgs {
    const uint32_t* seg_off;
    const uint32_t* codes;  // M x nb
    const double* cw;       // null = unit
    const double* inv_n1;
    const uint8_t* red;
    uint64_t g1, lo, range;
    int nb, nbs;
    uint32_t bsel[NB_MAX];   // blocks whose coordinates intersect [lo, lo + range)
    uint64_t boff[NB_MAX];   // their coordinate offsets
};

// Item t of segment s (cell c0 + t / nbs, block bsel[t % nbs]): its coordinate relative to lo and
// the cell weight; false when the coordinate is outside the range or redundant.
__device__ __forceinline__ bool asm_item(const AsmArgs& A, uint64_t c0, uint32_t t, uint32_t& k, double& w,
                                         uint64_t& cell) {
    cell = c0 + t / (uint32_t)A.nbs;
    const int q = (int)(t % (uint32_t)A.nbs);
    const uint64_t g = A.boff[q] + A.codes[cell * (uint64_t)A.nb + A.bsel[q]];
    if (g < A.lo || g >= A.lo + A.range || A.red[g]) return false;
    k = (uint32_t)(g - A.lo);
    w = A.cw ? A.cw[cell] : 1.0;
    return true;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Lower-triangle pairs of segment s counted by this lane (outer items lane, lane + 64, ...).
__device__ uint32_t asm_lane_count(const AsmArgs& A, uint64_t c0, uint32_t items, int lane) {
    uint32_t cnt = 0;
    for (uint32_t t = lane; t < items; t += WAVE) {
        uint32_t kt, ku;
        double wt, wu;
        uint64_t ct, cu;
        if (!asm_item(A, c0, t, kt, wt, ct)) continue;
        for (uint32_t u = 0; u < items; u++)
            if (asm_item(A, c0, u, ku, wu, cu) && ku <= kt) cnt++;
    }
    return cnt;
}

// One wave per segment: the number of lower-triangle pairs.
__global__ __launch_bounds__(TPB) void k_asm_count(AsmArgs A, uint32_t* __restrict__ cnt) {
    const int lane = lane_id();
    for (uint64_t s = ((uint64_t)blockIdx.x * TPB + threadIdx.x) / WAVE; s < A.g1;
         s += (uint64_t)gridDim.x * (TPB / WAVE)) {
        const uint64_t c0 = A.seg_off[s];
        const uint32_t items = (uint32_t)((A.seg_off[s + 1] - c0) * (uint64_t)A.nbs);
        const uint32_t c = wave_sum_u32(asm_lane_count(A, c0, items, lane));
        if (lane == 0) cnt[s] = c;
    }
}

// One wave per segment: its triples at off[s] (lanes in order, each lane's pairs in its loop order).
__global__ __launch_bounds__(TPB) void k_asm_emit(AsmArgs A, const uint32_t* __restrict__ off,
                                                  uint64_t* __restrict__ keys, uint32_t* __restrict__ idx,
                                                  double* __restrict__ vals) {
    const int lane = lane_id();
    for (uint64_t s = ((uint64_t)blockIdx.x * TPB + threadIdx.x) / WAVE; s < A.g1;
         s += (uint64_t)gridDim.x * (TPB / WAVE)) {
        const uint64_t c0 = A.seg_off[s];
        const uint32_t items = (uint32_t)((A.seg_off[s + 1] - c0) * (uint64_t)A.nbs);
        const uint32_t mine = asm_lane_count(A, c0, items, lane);
        uint32_t before = mine;  // inclusive scan over the lanes
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const uint32_t v = __shfl_up(before, o, WAVE);
            if (lane >= o) before += v;
        }
        uint64_t pos = (uint64_t)off[s] + before - mine;
        const double in1 = A.inv_n1[s];
        for (uint32_t t = lane; t < items; t += WAVE) {
            uint32_t kt, ku;
            double wt, wu;
            uint64_t ct, cu;
            if (!asm_item(A, c0, t, kt, wt, ct)) continue;
            for (uint32_t u = 0; u < items; u++) {
                if (!asm_item(A, c0, u, ku, wu, cu) || ku > kt) continue;
                keys[pos] = (uint64_t)kt * A.range + ku;
                idx[pos] = (uint32_t)pos;
                vals[pos] = (ct == cu ? wt : 0.0) - wt * wu * in1;
                pos++;
            }
        }
    }
}

__global__ void k_run_flags(const uint64_t* __restrict__ keys, size_t n, uint32_t* __restrict__ flag) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_run_starts(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ uid, size_t n,
                             uint32_t* __restrict__ start) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (flag[i]) start[uid[i]] = (uint32_t)i;
}

// Entry u of the CSR: the sum of its run in sorted (= emission) order, its column and a row count.
__global__ void k_run_reduce(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                             const double* __restrict__ vals, const uint32_t* __restrict__ start, size_t nu,
                             size_t n, uint64_t range, uint32_t* __restrict__ col, double* __restrict__ val,
                             uint32_t* __restrict__ rowcnt) {
    for (size_t u = (size_t)blockIdx.x * blockDim.x + threadIdx.x; u < nu; u += (size_t)gridDim.x * blockDim.x) {
        const size_t i0 = start[u], i1 = u + 1 < nu ? start[u + 1] : n;
        double s = 0.0;
        for (size_t i = i0; i < i1; i++) s += vals[idx[i]];
        const uint64_t key = keys[i0];
        col[u] = (uint32_t)(key % range);
        val[u] = s;
        atomicAdd(&rowcnt[key / range], 1u);
    }
}

// ------------------------------------------------------------------ triangular solves

struct FacDev {
    const uint32_t *gmap, *fptr, *fidx, *bptr, *bidx, *fsched, *fofs, *bsched, *bofs;
    const double *sc, *dinv, *fval, *bval;
};

// Row k of the forward (FWD) or backward solve, by one wave: y_k = s_k r_g(k) - sum_j L_kj y_j, or
// x_k = y_k / d_k - sum_i L_ik x_i with z_g(k) = s_k x_k.
template <bool FWD>
__device__ __forceinline__ void fac_row(uint32_t k, const FacDev& F, const double* __restrict__ rr, double* y,
                                        double* __restrict__ z, int r, int lane) {
    const uint32_t* ptr = FWD ? F.fptr : F.bptr;
    const uint32_t* idx = FWD ? F.fidx : F.bidx;
    const double* val = FWD ? F.fval : F.bval;
    double acc[FAC_RMAX];
#pragma unroll
    for (int c = 0; c < FAC_RMAX; c++) acc[c] = 0.0;
    const uint32_t e1 = ptr[k + 1];
    for (uint32_t e = ptr[k] + lane; e < e1; e += WAVE) {
        const double v = val[e];
        const double* yj = y + (size_t)idx[e] * r;
#pragma unroll
        for (int c = 0; c < FAC_RMAX; c++)
            if (c < r) acc[c] += v * yj[c];
    }
    double mine = 0.0;
#pragma unroll
    for (int c = 0; c < FAC_RMAX; c++)
        if (c < r) {
            const double s = wave_sum(acc[c]);
            if (lane == c) mine = s;
        }
    if (lane < r) {
        const size_t g = (size_t)F.gmap[k] * r + lane;
        if (FWD) {
            y[(size_t)k * r + lane] = F.sc[k] * rr[g] - mine;
        } else {
            const double x = F.dinv[k] * y[(size_t)k * r + lane] - mine;
            y[(size_t)k * r + lane] = x;
            z[g] = F.sc[k] * x;
        }
    }
}

// Levels [l0, l1) of one direction by the waves of one workgroup, a barrier after each level.
template <bool FWD>
__device__ void fac_levels_wg(const FacDev& F, uint32_t l0, uint32_t l1, const double* __restrict__ rr, double* y,
                              double* __restrict__ z, int r) {
    const uint32_t* sched = FWD ? F.fsched : F.bsched;
    const uint32_t* ofs = FWD ? F.fofs : F.bofs;
    const int lane = lane_id();
    const uint32_t wv = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    for (uint32_t L = l0; L < l1; L++) {
        const uint32_t t1 = ofs[L + 1];
        for (uint32_t t = ofs[L] + wv; t < t1; t += nw) fac_row<FWD>(sched[t], F, rr, y, z, r, lane);
        __syncthreads();
    }
}

// sum over the factor's positions of r_g(k) z_g(k) per column, by this workgroup's threads over
// positions blockIdx.x*blockDim.x + threadIdx.x (stride gridDim.x * blockDim.x); written to
// part[(base + blockIdx.x) * 16 + j] in a fixed order.
__device__ void fac_rz(const FacDev& F, size_t nf, const double* __restrict__ rr, const double* __restrict__ z, int r,
                       double* __restrict__ part, int base) {
    __shared__ double red[FAC_LDS_THREADS / WAVE][FAC_RMAX];
    double acc[FAC_RMAX];
#pragma unroll
    for (int c = 0; c < FAC_RMAX; c++) acc[c] = 0.0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nf; k += (size_t)gridDim.x * blockDim.x) {
        const size_t g = (size_t)F.gmap[k] * r;
#pragma unroll
        for (int c = 0; c < FAC_RMAX; c++)
            if (c < r) acc[c] += rr[g + c] * z[g + c];
    }
    const int w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
    for (int c = 0; c < FAC_RMAX; c++)
        if (c < r) {
            const double s = wave_sum(acc[c]);
            if (lane_id() == 0) red[w][c] = s;
        }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        double s = 0.0;
        for (int q = 0; q < nw; q++) s += red[q][threadIdx.x];
        part[((size_t)base + blockIdx.x) * FAC_RMAX + threadIdx.x] = s;
    }
}

// The whole solve in one workgroup with the vector in LDS.
__global__ __launch_bounds__(FAC_LDS_THREADS) void k_fac_lds(FacDev F, size_t nf, uint32_t nlf, uint32_t nlb,
                                                             const double* __restrict__ rr, double* __restrict__ z,
                                                             int r, double* __restrict__ part, int base,
                                                             const uint32_t* __restrict__ done) {
    if (done && *done) return;
    extern __shared__ double ys[];
    fac_levels_wg<true>(F, 0, nlf, rr, ys, z, r);
    fac_levels_wg<false>(F, 0, nlb, rr, ys, z, r);
    fac_rz(F, nf, rr, z, r, part, base);
}

// A run of levels on a vector in HBM: one wide level over the grid, or narrow levels in one
// workgroup (gridDim.x == 1).
template <bool FWD>
__global__ __launch_bounds__(FAC_LDS_THREADS) void k_fac_level(FacDev F, uint32_t l0, uint32_t l1,
                                                               const double* __restrict__ rr, double* y,
                                                               double* __restrict__ z, int r,
                                                               const uint32_t* __restrict__ done) {
    if (done && *done) return;
    if (gridDim.x == 1) {
        fac_levels_wg<FWD>(F, l0, l1, rr, y, z, r);
        return;
    }
    const uint32_t* sched = FWD ? F.fsched : F.bsched;
    const uint32_t* ofs = FWD ? F.fofs : F.bofs;
    const uint32_t nw = blockDim.x / WAVE;
    const uint32_t t1 = ofs[l0 + 1];
    for (uint32_t t = ofs[l0] + blockIdx.x * nw + threadIdx.x / WAVE; t < t1; t += gridDim.x * nw)
        fac_row<FWD>(sched[t], F, rr, y, z, r, lane_id());
}

__global__ __launch_bounds__(TPB) void k_fac_dot(FacDev F, size_t nf, const double* __restrict__ rr,
                                                 const double* __restrict__ z, int r, double* __restrict__ part,
                                                 int base, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    fac_rz(F, nf, rr, z, r, part, base);
}

template <class T>
__global__ void k_zero_range(T* __restrict__ v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = T(0);
}

}  // namespace

// ------------------------------------------------------------------ factor on the device

struct FacGpu {
    uint64_t lo = 0, range = 0;   // coordinates [lo, lo + range) it covers
    std::vector<uint32_t> keep;   // kept coordinates (global, ascending)
    LdlFactor f;
    size_t nf = 0;
    uint32_t nlf = 0, nlb = 0;    // forward / backward levels
    uint32_t maxw = 1;            // most rows in one level
    std::vector<uint32_t> hfofs, hbofs;
    DBuf<uint32_t> gmap, fptr, fidx, bptr, bidx, fsched, fofs, bsched, bofs;
    DBuf<double> sc, dinv, fval, bval, y;
    FacDev dev() const {
        return FacDev{gmap.p, fptr.p, fidx.p, bptr.p, bidx.p, fsched.p, fofs.p, bsched.p, bofs.p,
                      sc.p,   dinv.p, fval.p, bval.p};
    }
};

namespace {

// Level schedule: positions grouped by level (stable), and the level offsets.
void schedule(const std::vector<uint32_t>& lev, uint32_t nl, std::vector<uint32_t>& sched, std::vector<uint32_t>& ofs) {
    ofs.assign(nl + 1, 0);
    for (uint32_t l : lev) ofs[l + 1]++;
    for (uint32_t l = 0; l < nl; l++) ofs[l + 1] += ofs[l];
    sched.resize(lev.size());
    std::vector<uint32_t> fill(ofs.begin(), ofs.end() - 1);
    for (size_t k = 0; k < lev.size(); k++) sched[fill[lev[k]]++] = (uint32_t)k;
}

template <class T>
void put(DBuf<T>& d, const std::vector<T>& h, hipStream_t st) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty()) FH_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

// Moves the host factor to the device with its level schedules.
void upload_factor(FacGpu& F, hipStream_t st) {
    const LdlFactor& f = F.f;
    const size_t n = f.n;
    F.nf = n;
    std::vector<uint32_t> gmap(n);
    std::vector<double> dinv(n);
    for (size_t k = 0; k < n; k++) {
        gmap[k] = F.keep[f.perm[k]];
        dinv[k] = 1.0 / f.d[k];
    }
    const uint64_t len = f.len();
    require(len < (1ull << 32), "escalation: factor too large for 32-bit offsets", FH_INTERNAL);
    std::vector<uint32_t> bptr(n + 1), fptr(n + 1, 0), fidx(len);
    std::vector<double> fval(len);
    for (size_t k = 0; k <= n; k++) bptr[k] = (uint32_t)f.lp[k];
    // rows of L (forward solve) from its columns: column j's rows ascend, so each row's entries
    // come out in ascending column order
    for (uint64_t p = 0; p < len; p++) fptr[f.li[p] + 1]++;
    for (size_t k = 0; k < n; k++) fptr[k + 1] += fptr[k];
    {
        std::vector<uint32_t> fill(fptr.begin(), fptr.end() - 1);
        for (size_t j = 0; j < n; j++)
            for (uint64_t p = f.lp[j]; p < f.lp[j + 1]; p++) {
                const uint32_t i = f.li[p];
                fidx[fill[i]] = (uint32_t)j;
                fval[fill[i]++] = f.lx[p];
            }
    }
    // levels: forward, a row after every row it reads; backward, after every row below it
    std::vector<uint32_t> levf(n, 0), levb(n, 0);
    uint32_t nlf = n ? 1 : 0, nlb = n ? 1 : 0;
    for (size_t j = 0; j < n; j++)
        for (uint64_t p = f.lp[j]; p < f.lp[j + 1]; p++) levf[f.li[p]] = std::max(levf[f.li[p]], levf[j] + 1);
    for (size_t j = n; j-- > 0;)
        for (uint64_t p = f.lp[j]; p < f.lp[j + 1]; p++) levb[j] = std::max(levb[j], levb[f.li[p]] + 1);
    for (size_t k = 0; k < n; k++) {
        nlf = std::max(nlf, levf[k] + 1);
        nlb = std::max(nlb, levb[k] + 1);
    }
    std::vector<uint32_t> fs, bs;
    schedule(levf, nlf, fs, F.hfofs);
    schedule(levb, nlb, bs, F.hbofs);
    F.nlf = nlf;
    F.nlb = nlb;
    F.maxw = 1;
    for (uint32_t l = 0; l < nlf; l++) F.maxw = std::max(F.maxw, F.hfofs[l + 1] - F.hfofs[l]);
    for (uint32_t l = 0; l < nlb; l++) F.maxw = std::max(F.maxw, F.hbofs[l + 1] - F.hbofs[l]);
    std::vector<uint32_t> li32(f.li.begin(), f.li.end());
    put(F.gmap, gmap, st);
    put(F.sc, f.scale, st);
    put(F.dinv, dinv, st);
    put(F.bptr, bptr, st);
    put(F.bidx, li32, st);
    put(F.bval, f.lx, st);
    put(F.fptr, fptr, st);
    put(F.fidx, fidx, st);
    put(F.fval, fval, st);
    put(F.fsched, fs, st);
    put(F.fofs, F.hfofs, st);
    put(F.bsched, bs, st);
    put(F.bofs, F.hbofs, st);
    stream_wait(st);  // the host vectors above go out of scope
}

// S over the coordinates [lo, lo + range) (redundant ones excluded), assembled on the GPU and
// brought to the host as a lower CSR over the kept coordinates. Returns false over a budget.
bool assemble(Absorber& a, uint64_t lo, uint64_t range, SymLower& out, std::vector<uint32_t>& keep) {
    hipStream_t st = a.st;
    AsmArgs A{};
    A.seg_off = a.seg_off.p;
    A.codes = a.cell_codes.p;
    A.cw = a.unit() ? nullptr : a.w64.cw.p;
    A.inv_n1 = a.w64.inv_n1.p;
    A.red = a.red.p;
    A.g1 = a.g1;
    A.lo = lo;
    A.range = range;
    A.nb = (int)a.nb;
    A.nbs = 0;
    for (size_t b = 0; b < a.nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (bi.off + bi.g <= lo || bi.off >= lo + range) continue;
        A.bsel[A.nbs] = (uint32_t)b;
        A.boff[A.nbs++] = bi.off;
    }
    require(A.nbs > 0, "escalation: no block in the range", FH_INTERNAL);
    // the work bound from the segment sizes (host): sum_s (d_s nbs)^2 / 2 pairs at most
    std::vector<uint32_t> so(a.g1 + 1);
    download(so.data(), a.seg_off.p, so.size() * sizeof(uint32_t), st);
    double work = 0.0;
    for (size_t s = 0; s < a.g1; s++) {
        const double it = (double)(so[s + 1] - so[s]) * A.nbs;
        work += it * (it + 1) / 2;
    }
    if (work > (double)ESC_MAX_PAIRS) return false;
    const unsigned grid = grid_for(a.g1, TPB / WAVE, 1u << 16);
    DBuf<uint32_t> cnt(a.g1 + 1), total(1);
    hipLaunchKernelGGL(k_asm_count, dim3(grid), dim3(TPB), 0, st, A, cnt.p);
    FH_CHECK_LAUNCH();
    exclusive_scan_u32(cnt.p, cnt.p, a.g1, total.p, st);
    uint32_t np = 0;
    download(&np, total.p, sizeof(np), st);
    if (np == 0) return false;
    DBuf<uint64_t> keys(np), keys2(np);
    DBuf<uint32_t> idx(np), idx2(np);
    DBuf<double> vals(np);
    hipLaunchKernelGGL(k_asm_emit, dim3(grid), dim3(TPB), 0, st, A, cnt.p, keys.p, idx.p, vals.p);
    FH_CHECK_LAUNCH();
    radix_sort_pairs(keys, idx, keys2, idx2, np, (int)bits_for(range * range), st);
    DBuf<uint32_t> flag(np), uid(np);
    const unsigned vg = grid_for(np, TPB, 8192);
    hipLaunchKernelGGL(k_run_flags, dim3(vg), dim3(TPB), 0, st, keys.p, (size_t)np, flag.p);
    FH_CHECK_LAUNCH();
    exclusive_scan_u32(flag.p, uid.p, np, total.p, st);
    uint32_t nu = 0;
    download(&nu, total.p, sizeof(nu), st);
    if (nu > ESC_MAX_NNZ) return false;
    DBuf<uint32_t> start(nu), col(nu), rowcnt(range + 1);
    DBuf<double> val(nu);
    FH_HIP(hipMemsetAsync(rowcnt.p, 0, (range + 1) * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_run_starts, dim3(vg), dim3(TPB), 0, st, flag.p, uid.p, (size_t)np, start.p);
    hipLaunchKernelGGL(k_run_reduce, dim3(grid_for(nu, TPB, 8192)), dim3(TPB), 0, st, keys.p, idx.p, vals.p, start.p,
                       (size_t)nu, (size_t)np, range, col.p, val.p, rowcnt.p);
    FH_CHECK_LAUNCH();
    std::vector<uint32_t> hcnt(range), hcol(nu);
    std::vector<double> hval(nu);
    std::vector<uint8_t> hred(range);
    download(hcnt.data(), rowcnt.p, range * sizeof(uint32_t), st);
    download(hcol.data(), col.p, nu * sizeof(uint32_t), st);
    download(hval.data(), val.p, nu * sizeof(double), st);
    download(hred.data(), a.red.p + lo, range, st);
    // kept coordinates: not redundant, with a positive diagonal; rows of the others are empty
    std::vector<int64_t> kidx(range, -1);
    std::vector<uint64_t> rp(range + 1, 0);
    for (uint64_t i = 0; i < range; i++) rp[i + 1] = rp[i] + hcnt[i];
    keep.clear();
    for (uint64_t i = 0; i < range; i++) {
        if (hred[i] || hcnt[i] == 0) continue;
        const uint64_t e = rp[i + 1] - 1;  // the diagonal is the row's last entry
        if (hcol[e] != i || !(hval[e] > 0.0)) continue;
        kidx[i] = (int64_t)keep.size();
        keep.push_back((uint32_t)(lo + i));
    }
    out = SymLower();
    out.n = keep.size();
    out.ptr.assign(out.n + 1, 0);
    out.col.reserve(nu);
    out.val.reserve(nu);
    for (uint64_t i = 0; i < range; i++) {
        if (kidx[i] < 0) continue;
        for (uint64_t e = rp[i]; e < rp[i + 1]; e++) {
            const int64_t j = kidx[hcol[e]];
            if (j < 0) continue;
            out.col.push_back((uint32_t)j);
            out.val.push_back(hval[e]);
        }
        out.ptr[kidx[i] + 1] = out.col.size();
    }
    return out.n > 0;
}

double secs_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

// Builds the factor of [lo, lo + range) (small_last: kept coordinates below a.ms ordered last).
// budget (seconds): the build stops before a stage whose estimated remaining cost exceeds it
// (costs already spent are sunk). joint: the joint factor's statistics are kept on the absorber.
std::shared_ptr<FacGpu> build_factor(Absorber& a, uint64_t lo, uint64_t range, bool small_last, const fh_ctx* ctx,
                                     const char* what, double budget, bool joint) {
    const auto t0 = std::chrono::steady_clock::now();
    auto F = std::make_shared<FacGpu>();
    F->lo = lo;
    F->range = range;
    SymLower S;
    const std::string tag = std::string("fereg: escalation: ") + what;
    if (!assemble(a, lo, range, S, F->keep)) {
        log_line(ctx, 2, tag + " over the assembly budgets");
        if (joint) a.fac_failed_joint = true;
        return nullptr;
    }
    const double t_asm = secs_since(t0);
    const double nnz = (double)S.ptr[S.n];
    if (joint) a.joint_nnz = nnz;
    if (C_NNZ * nnz > budget) {
        log_line(ctx, 2, tag + " rejected (cost) after assembly: nnz=" + std::to_string((uint64_t)nnz));
        return nullptr;
    }
    size_t nlast = 0;
    if (small_last)
        while (nlast < F->keep.size() && F->keep[nlast] < a.ms) nlast++;
    const bool fits = ldl_analyze(S, nlast, ESC_MAX_LEN, ESC_MAX_FLOPS, F->f);
    const double t_an = secs_since(t0) - t_asm;
    if (joint) {
        a.joint_len = (double)F->f.len();
        a.joint_flops = F->f.flops;
    }
    if (!fits) {
        log_line(ctx, 2, tag + " over the fill budget (L=" + std::to_string(F->f.len()) + ")");
        if (joint) a.fac_failed_joint = true;
        return nullptr;
    }
    if (C_FLOP * F->f.flops + C_ENTRY * (double)F->f.len() > budget) {
        log_line(ctx, 2, tag + " rejected (cost) after analysis: L=" + std::to_string(F->f.len()));
        return nullptr;
    }
    check_interrupt(ctx);
    if (!ldl_numeric(S, F->f)) {
        if (joint) a.fac_failed_joint = true;
        return nullptr;
    }
    const double t_num = secs_since(t0) - t_asm - t_an;
    upload_factor(*F, a.st);
    const double secs = secs_since(t0);
    log_line(ctx, 2, tag + " factor n=" + std::to_string(F->nf) + " nnz=" + std::to_string((uint64_t)nnz) +
                         " L=" + std::to_string(F->f.len()) + " flops=" + std::to_string(F->f.flops) +
                         " grounded=" + std::to_string(F->f.grounded) + " levels=" + std::to_string(F->nlf) + "/" +
                         std::to_string(F->nlb) + " in " + std::to_string(secs) + "s (assembly " +
                         std::to_string(t_asm) + ", analysis " + std::to_string(t_an) + ", numeric " +
                         std::to_string(t_num) + ", upload " + std::to_string(secs - t_asm - t_an - t_num) + ")");
    return F;
}

// Item pairs the joint assembly emits at most (structure only; one download of seg_off).
double joint_pairs(Absorber& a) {
    if (a.joint_pairs < 0.0) {
        std::vector<uint32_t> so(a.g1 + 1);
        download(so.data(), a.seg_off.p, so.size() * sizeof(uint32_t), a.st);
        double work = 0.0;
        for (size_t s = 0; s < a.g1; s++) {
            const double it = (double)(so[s + 1] - so[s]) * (double)a.nb;
            work += it * (it + 1) / 2;
        }
        a.joint_pairs = work;
    }
    return a.joint_pairs;
}

// Jacobi and the dense small factor leave the factored coordinates to the factors: their inverse
// diagonal (and, under a joint factor, the small block's inverse) are zero, so the fused update
// kernel writes z = 0 and r'z = 0 there, and the factor solves write z and their own r'z.
void mask_factored(Absorber& a) {
    hipStream_t st = a.st;
    for (auto& F : a.facs) {
        const size_t lo = F->lo, n = F->range;
        hipLaunchKernelGGL(k_zero_range<double>, dim3(grid_for(n, TPB, 4096)), dim3(TPB), 0, st, a.w64.inv_diag.p + lo, n);
        FH_CHECK_LAUNCH();
    }
    if (a.fac_joint && a.ms > 0 && a.w64.minv.p) {
        hipLaunchKernelGGL(k_zero_range<double>, dim3(grid_for(a.ms * a.ms, TPB, 4096)), dim3(TPB), 0, st,
                           a.w64.minv.p, a.ms * a.ms);
        FH_CHECK_LAUNCH();
    }
}

}  // namespace

bool factor_joint_host(Absorber& a, LdlFactor& f, std::vector<uint32_t>& keep, std::string& why) {
    SymLower S;
    if (!assemble(a, 0, a.m, S, keep)) {
        why = "the reduced matrix exceeds the assembly budgets";
        return false;
    }
    size_t nlast = 0;
    while (nlast < keep.size() && keep[nlast] < a.ms) nlast++;
    if (!ldl_analyze(S, nlast, ESC_MAX_LEN, ESC_MAX_FLOPS, f)) {
        why = "the factor exceeds its fill budget (" + std::to_string(f.len()) + " entries)";
        return false;
    }
    if (!ldl_numeric(S, f)) {
        why = "a non-finite pivot";
        return false;
    }
    return true;
}

bool escalation_possible(const Absorber& a) {
    if (!a.sparse_chol || a.fac_failed || a.m <= a.ms) return false;
    return a.facs.empty() || !a.fac_fresh;
}

bool escalate(Absorber& a, const fh_ctx* ctx) {
    require(escalation_possible(a), "escalation: nothing to escalate");
    if (!a.facs.empty()) {
        // stale factors: drop them and recompute the Jacobi and small-block state they masked
        a.facs.clear();
        a.fac_joint = false;
        refresh_weights(a, a.weighted ? a.w_rows.p : nullptr);
    }
    if (a.joint_policy != 0 && !a.fac_failed_joint) {
        auto J = build_factor(a, 0, a.m, true, ctx, "joint", INFINITY, true);
        if (J) {
            a.facs.push_back(J);
            a.fac_joint = true;
        }
    }
    if (a.facs.empty()) {
        for (const BlockInfo& bi : a.blocks) {
            if (bi.small) continue;
            auto B = build_factor(a, bi.off, bi.g, false, ctx, "block", INFINITY, false);
            if (B) a.facs.push_back(B);
        }
    }
    if (a.facs.empty()) {
        a.fac_failed = true;
        return false;
    }
    a.fac_fresh = true;
    mask_factored(a);
    return true;
}

void facs_after_refresh(Absorber& a) {
    if (a.facs.empty()) return;
    // a factor stays as the preconditioner for new weights while its kept coordinates are the
    // same (it is still SPD there, so the PCG converges to the same answer); otherwise Jacobi
    // returns. A stale factor is rebuilt when the PCG escalates again.
    std::vector<uint8_t> red(a.m);
    download(red.data(), a.red.p, a.m, a.st);
    bool same = true;
    for (auto& F : a.facs) {
        size_t q = 0;
        for (uint64_t i = 0; i < F->range && same; i++) {
            const bool kept = !red[F->lo + i];
            const bool was = q < F->keep.size() && F->keep[q] == F->lo + i;
            if (was) q++;
            same = kept == was;
        }
    }
    a.fac_fresh = false;
    if (!same) {
        a.facs.clear();
        a.fac_joint = false;
        return;
    }
    mask_factored(a);
}

bool facs_exact(const Absorber& a) { return a.fac_joint && a.fac_fresh && !a.facs.empty(); }

double joint_build_estimate(Absorber& a) {
    if (a.joint_policy == 0 || !a.sparse_chol || a.fac_failed || a.fac_failed_joint || !a.facs.empty() ||
        a.m <= a.ms)
        return INFINITY;
    double c = C_PAIR * joint_pairs(a);
    if (a.joint_nnz >= 0.0) c += C_NNZ * a.joint_nnz;
    if (a.joint_flops >= 0.0) c += C_FLOP * a.joint_flops + C_ENTRY * a.joint_len;
    return c;
}

bool escalate_joint(Absorber& a, const fh_ctx* ctx, double without) {
    if (joint_build_estimate(a) == INFINITY) return false;
    log_line(ctx, 2, "fereg: early probe: joint factor est " + std::to_string(joint_build_estimate(a)) +
                         "s vs pcg est " + std::to_string(without) + "s");
    auto J = build_factor(a, 0, a.m, true, ctx, "joint (probe)", 0.8 * without, true);
    if (!J) return false;
    a.facs.push_back(J);
    a.fac_joint = true;
    a.fac_fresh = true;
    mask_factored(a);
    return true;
}

void facs_apply(Absorber& a, const double* rr, double* z, int r, double* rzf_part, const uint32_t* done,
                hipStream_t st) {
    const int nfac = (int)a.facs.size();
    if (nfac == 0) return;
    const int span = SPMV_DOTG / nfac;
    for (int q = 0; q < nfac; q++) {
        FacGpu& F = *a.facs[q];
        if (F.nf == 0) continue;
        const FacDev D = F.dev();
        const size_t vbytes = F.nf * (size_t)r * sizeof(double);
        if (vbytes <= FAC_LDS_BYTES) {
            // one wave per row of the widest level, up to 16 (a barrier costs every wave of the group)
            const unsigned threads = WAVE * std::min<unsigned>(FAC_LDS_THREADS / WAVE, F.maxw);
            hipLaunchKernelGGL(k_fac_lds, dim3(1), dim3(threads), vbytes, st, D, F.nf, F.nlf, F.nlb, rr, z, r,
                               rzf_part, q * span, done);
            FH_CHECK_LAUNCH();
            continue;
        }
        F.y.reserve(F.nf * (size_t)r);
        for (int dir = 0; dir < 2; dir++) {
            const std::vector<uint32_t>& ofs = dir == 0 ? F.hfofs : F.hbofs;
            const uint32_t nl = dir == 0 ? F.nlf : F.nlb;
            uint32_t l = 0;
            while (l < nl) {
                const uint32_t rows = ofs[l + 1] - ofs[l];
                uint32_t l1 = l + 1;
                unsigned grid = grid_for(rows, TPB / WAVE, 1u << 14);
                unsigned threads = TPB;
                if (rows < FAC_WIDE) {  // a run of narrow levels in one workgroup, a wave per row
                    uint32_t w = rows;
                    while (l1 < nl && ofs[l1 + 1] - ofs[l1] < FAC_WIDE) {
                        w = std::max(w, ofs[l1 + 1] - ofs[l1]);
                        l1++;
                    }
                    grid = 1;
                    threads = WAVE * std::min<unsigned>(FAC_LDS_THREADS / WAVE, w);
                }
                if (dir == 0)
                    hipLaunchKernelGGL(k_fac_level<true>, dim3(grid), dim3(threads), 0, st, D, l, l1, rr, F.y.p, z, r, done);
                else
                    hipLaunchKernelGGL(k_fac_level<false>, dim3(grid), dim3(threads), 0, st, D, l, l1, rr, F.y.p, z, r,
                                       done);
                FH_CHECK_LAUNCH();
                l = l1;
            }
        }
        hipLaunchKernelGGL(k_fac_dot, dim3(span), dim3(TPB), 0, st, D, F.nf, rr, z, r, rzf_part, q * span, done);
        FH_CHECK_LAUNCH();
    }
}

}  // namespace fh
