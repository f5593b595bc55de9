// Lockstep PCG over r columns on the GPU (cg.rs:132-292), with every per-column scalar kept in
// device memory: a one-workgroup scalar kernel applies the reference's rules (gamma, the
// Hestenes-Stiefel window of 5, thr = max(tol^2 (ss1 - x'rhs - x'r), 1e-26 ss0), the roundoff
// stop, beta, freezing of converged columns), so the host only polls a status word every few
// iterations; every kernel of an iteration returns at once after the solve has finished.
// The mixed form runs the first iterations in FP64, then the same PCG in FP32 on the correction
// system, refining in FP64 (residual rhs - S x recomputed with the FP64 operator after every
// inner solve; a column is accepted only on that FP64 residual).
#include <algorithm>
#include <atomic>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <initializer_list>

#include "fh_internal.hpp"

namespace fh {

namespace {

constexpr int RMAX = 16;
constexpr int WINDOW = 5;  // cg.rs:23
constexpr int DOTG = 256;  // workgroups of the dot kernels (fixed: fixed summation order)
constexpr int TPB = 256;

// Largest drop of r'M^-1 r one FP32 correction sweep of the mixed form is asked for (rel_one^2).
constexpr double HANDOVER_GAIN = 1e10;

// Escalation probe (cg.rs:29-37, 254-274): r'z after PROBE_FROM iterations of a solve and the
// largest rate rho_hat = (r'z_k / r'z_from)^(1 / (k - from)) at k = PROBE_AT over the active
// columns; rho_hat > PROBE_RATIO asks for a stronger preconditioner (err = 2).
constexpr int PROBE_FROM = 7, PROBE_AT = 12;
constexpr double PROBE_RATIO = 0.6;
// Early probe (cg.rs:38-42): rho_hat = (r'z_6 / r'z_3)^(1/3) and the projected remaining
// iterations max_j ln(r'z_j / thr_j) / ln(1 / rho_hat_j), for the joint factor's cost rule.
constexpr int EARLY_FROM = 3, EARLY_AT = 6;

struct PcgState {
    double rho[RMAX], rho_new[RMAX], gamma[RMAX], beta[RMAX], ss1[RMAX], ss0[RMAX], floor_[RMAX], thr[RMAX];
    double ring[RMAX][WINDOW];
    double xb[RMAX], xr[RMAX], rho0[RMAX], absthr[RMAX], rho_from[RMAX], rho_early[RMAX];
    int32_t active[RMAX];
    uint32_t done, err;  // err: 1 = maxiter reached, 2 = escalation requested, 3 = hand-over
    uint64_t iters;
    double probe_rho, probe_rem;  // the early probe's rate and remaining iterations (-1: not yet)
    uint64_t k, maxiter;
    int32_t r;
    int32_t can_escalate;
    int32_t fac;  // escalated factors write their r'z partials to slot S_RZF
    double tol, rel2;
    uint64_t handover_at;  // > 0: stop with err = 3 after this iteration while columns remain
};

enum Slot { S_PQ = 0, S_XB = 1, S_XR = 2, S_RZL = 3, S_RZS = 4, S_RZF = 5, NSLOT = 6 };
constexpr int NSLOT_MAX = 5;

// Per-workgroup partial dots: dpart[(slot*DOTG + wg)*RMAX + j] over rows [row0, row1).
template <class T>
__global__ __launch_bounds__(TPB) void k_dot(const T* __restrict__ a, const T* __restrict__ b, size_t row0,
                                             size_t row1, int r, double* __restrict__ dpart, int slot,
                                             const uint32_t* __restrict__ done) {
    if (done && *done) return;
    __shared__ double red[TPB / WAVE][RMAX];
    double acc[RMAX];
#pragma unroll
    for (int j = 0; j < RMAX; j++) acc[j] = 0.0;
    const size_t n = (row1 - row0) * r;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
        const size_t e = row0 * r + i;
        const int j = (int)(i % r);
        const double v = (double)a[e] * (double)b[e];
#pragma unroll
        for (int jj = 0; jj < RMAX; jj++)
            if (jj == j) acc[jj] += v;
    }
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int j = 0; j < RMAX; j++) {
        if (j < r) {
            double s = wave_sum(acc[j]);
            if (lane_id() == 0) red[w][j] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        double s = 0.0;
        for (int ww = 0; ww < TPB / WAVE; ww++) s += red[ww][threadIdx.x];
        dpart[((size_t)slot * DOTG + blockIdx.x) * RMAX + threadIdx.x] = s;
    }
}

// sums[slot][j] = sum over the DOTG workgroup partials of `nslot` slots (first, ...), computed
// by the whole 256-thread block: thread t holds partial t, a wave sum, then 4 wave sums (fixed
// order). Every thread must call it.
__device__ void slot_sums(const double* __restrict__ dpart, const int* slots, int nslot, int r,
                          double (*sums)[RMAX]) {
    __shared__ double red[TPB / WAVE][NSLOT_MAX][RMAX];
    const int t = threadIdx.x, w = t / WAVE;
    for (int a = 0; a < nslot; a++)
        for (int j = 0; j < r; j++) {
            double v = dpart[((size_t)slots[a] * DOTG + t) * RMAX + j];
            v = wave_sum(v);
            if (lane_id() == 0) red[w][a][j] = v;
        }
    __syncthreads();
    if (t < nslot * r) {
        const int a = t / r, j = t % r;
        double v = 0.0;
        for (int ww = 0; ww < TPB / WAVE; ww++) v += red[ww][a][j];
        sums[a][j] = v;
    }
    __syncthreads();
}

__global__ void k_zero_slots(double* dpart, int slot, int nslots) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)nslots * DOTG * RMAX;
         i += (size_t)gridDim.x * blockDim.x)
        dpart[(size_t)slot * DOTG * RMAX + i] = 0.0;
}

// z = inv_diag .* r on rows [ms, m) (jacobi.rs:12-41).
template <class T>
__global__ void k_jacobi(const T* __restrict__ inv_diag, const T* __restrict__ rr, T* __restrict__ z, size_t ms,
                         size_t m, int r, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const size_t n = (m - ms) * r;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t e = ms * r + i;
        z[e] = inv_diag[e / r] * rr[e];
    }
}

// z_S = M_S r_S with the explicit grounded pseudo-inverse (precond/dense.rs:117-160).
template <class T>
__global__ void k_small_apply(const T* __restrict__ minv, const T* __restrict__ rr, T* __restrict__ z, size_t ms,
                              int r, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < ms * r; i += (size_t)gridDim.x * blockDim.x) {
        const size_t k = i / r;
        const int j = (int)(i % r);
        T s = T(0);
        for (size_t l = 0; l < ms; l++) s += minv[k * ms + l] * rr[l * r + j];
        z[i] = s;
    }
}

template <class T>
__global__ void k_sub(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = a[i] - (b ? b[i] : T(0));
}

template <class TO, class TI>
__global__ void k_convert(const TI* __restrict__ a, TO* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (TO)a[i];
}

template <class TI>
__global__ void k_add_to64(double* __restrict__ x, const TI* __restrict__ d, size_t n, int r,
                           const int32_t* __restrict__ mask) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (!mask || mask[i % r]) x[i] += (double)d[i];
}

// Start of a (re)start: rho = r'z, inactive when rho <= 0 (cg.rs:199-205).
// The poll fields to the host's pinned mirror: the data, then after a system-scope fence the word
// the host waits on (the iteration count with the done flag in its top bit), so a host that sees
// that word sees the data.
__device__ void publish(const PcgState* S, PollMirror* m) {
    m->err = S->err;
    m->probe_rho = S->probe_rho;
    m->probe_rem = S->probe_rem;
    __threadfence_system();
    *(volatile uint64_t*)&m->iters_done = S->iters | ((uint64_t)(S->done != 0) << 63);
}

__global__ __launch_bounds__(TPB) void k_start(PcgState* S, const double* __restrict__ dpart, PollMirror* mirror) {
    __shared__ double sums[NSLOT_MAX][RMAX];
    const int sl[3] = {S_RZL, S_RZS, S_RZF};
    slot_sums(dpart, sl, S->fac ? 3 : 2, S->r, sums);
    const int j = threadIdx.x;
    if (j < S->r) {
        double rho = sums[0][j] + sums[1][j] + (S->fac ? sums[2][j] : 0.0);
        S->rho[j] = rho;
        S->rho0[j] = rho;
        if (!(rho > 0.0)) S->active[j] = 0;
        for (int q = 0; q < WINDOW; q++) S->ring[j][q] = 0.0;
    }
    __syncthreads();
    if (j == 0) {
        int any = 0;
        for (int c = 0; c < S->r; c++) any |= S->active[c];
        S->done = any ? 0u : 1u;
        S->k = 0;
        if (!any && mirror) publish(S, mirror);
    }
}

template <class T>
__global__ void k_mask_copy(const T* __restrict__ z, T* __restrict__ p, size_t n, int r, const PcgState* __restrict__ S) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = S->active[i % r] ? z[i] : T(0);
}

// gamma = rho / p'q (cg.rs:224-234) from the p'q partials, in every workgroup (same fixed order,
// so every workgroup holds the same values; workgroup 0 records them), then x += gamma p,
// r -= gamma q, partial x'rhs and x'r (cg.rs:349-454), Jacobi z on the large rows and its
// partial r'z. With SMALL (ms r <= TPB, ms <= SMALL_FUSE_MAX) workgroup 0 owns every small-row
// element and also applies the small block's explicit inverse, z_S = M_S r_S
// (precond/dense.rs:117-160), writing the partial r_S'z_S; the other workgroups write zeros there.
constexpr int SMALL_FUSE_MAX = 32;

template <class T, bool SMALL>
__global__ __launch_bounds__(TPB) void k_update(T* __restrict__ x, T* __restrict__ rr, const T* __restrict__ p,
                                                const T* __restrict__ q, const T* __restrict__ rhs,
                                                const T* __restrict__ inv_diag, const T* __restrict__ minv,
                                                T* __restrict__ z, size_t ms, size_t m, int r, PcgState* __restrict__ S,
                                                double* __restrict__ dpart) {
    if (S->done) return;
    __shared__ double sums[NSLOT_MAX][RMAX];
    __shared__ double red[TPB / WAVE][4][RMAX];
    __shared__ T rs[SMALL ? TPB : 1];
    {
        const int sl[1] = {S_PQ};
        slot_sums(dpart, sl, 1, r, sums);
    }
    double g[RMAX];
#pragma unroll
    for (int j = 0; j < RMAX; j++) {
        g[j] = 0.0;
        if (j < r && S->active[j]) {
            const double pq = sums[0][j];
            if (pq > 0.0 && isfinite(pq)) g[j] = S->rho[j] / pq;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)r) {
        const int j = threadIdx.x;
        const double pq = sums[0][j];
        double gj = 0.0;
#pragma unroll
        for (int jj = 0; jj < RMAX; jj++)
            if (jj == j) gj = g[jj];
        if (S->active[j] && !(pq > 0.0 && isfinite(pq))) S->active[j] = 0;
        S->gamma[j] = gj;
    }
    double xb[RMAX], xr[RMAX], rz[RMAX], rzs[RMAX];
#pragma unroll
    for (int j = 0; j < RMAX; j++) { xb[j] = 0.0; xr[j] = 0.0; rz[j] = 0.0; rzs[j] = 0.0; }
    const size_t n = m * r;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
        const int j = (int)(i % r);
        const size_t k = i / r;
        double gj = 0.0;
#pragma unroll
        for (int jj = 0; jj < RMAX; jj++)
            if (jj == j) gj = g[jj];
        const T xv = x[i] + (T)gj * p[i];
        const T rv = rr[i] - (T)gj * q[i];
        x[i] = xv;
        rr[i] = rv;
        const double a = (double)xv * (double)rhs[i], b = (double)xv * (double)rv;
        double c = 0.0;
        if (k >= ms) {
            const T zv = inv_diag[k] * rv;
            z[i] = zv;
            c = (double)rv * (double)zv;
        } else if (SMALL) {
            rs[i] = rv;  // workgroup 0, first step: i < ms r <= TPB
        }
#pragma unroll
        for (int jj = 0; jj < RMAX; jj++)
            if (jj == j) { xb[jj] += a; xr[jj] += b; rz[jj] += c; }
    }
    if (SMALL && blockIdx.x == 0) {
        __syncthreads();
        const size_t e = threadIdx.x;
        if (e < ms * r) {
            const size_t k = e / r;
            const int j = (int)(e % r);
            T sm = T(0);
            for (size_t l = 0; l < ms; l++) sm += minv[k * ms + l] * rs[l * r + j];
            z[e] = sm;
            const double c = (double)rs[e] * (double)sm;
#pragma unroll
            for (int jj = 0; jj < RMAX; jj++)
                if (jj == j) rzs[jj] += c;
        }
    }
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int j = 0; j < RMAX; j++) {
        if (j < r) {
            double a = wave_sum(xb[j]), b = wave_sum(xr[j]), c = wave_sum(rz[j]), d = SMALL ? wave_sum(rzs[j]) : 0.0;
            if (lane_id() == 0) { red[w][0][j] = a; red[w][1][j] = b; red[w][2][j] = c; red[w][3][j] = d; }
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        const int j = threadIdx.x;
        double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
        for (int ww = 0; ww < TPB / WAVE; ww++) { a += red[ww][0][j]; b += red[ww][1][j]; c += red[ww][2][j]; d += red[ww][3][j]; }
        dpart[((size_t)S_XB * DOTG + blockIdx.x) * RMAX + j] = a;
        dpart[((size_t)S_XR * DOTG + blockIdx.x) * RMAX + j] = b;
        dpart[((size_t)S_RZL * DOTG + blockIdx.x) * RMAX + j] = c;
        if (SMALL) dpart[((size_t)S_RZS * DOTG + blockIdx.x) * RMAX + j] = d;
    }
}

// Stopping rule and beta (cg.rs:237-289).
// Called by every thread of one 256-thread workgroup.
__device__ void stop_rule(PcgState* S, const double* __restrict__ dpart, PollMirror* mirror) {
    __shared__ double sums[NSLOT_MAX][RMAX];
    const int sl[5] = {S_RZL, S_RZS, S_XB, S_XR, S_RZF};
    slot_sums(dpart, sl, S->fac ? 5 : 4, S->r, sums);
    const int j = threadIdx.x;
    __shared__ int any, not_ready, soon;
    if (j == 0) {
        any = 0;
        not_ready = 0;
        soon = 1;
    }
    __syncthreads();
    const uint64_t k = S->k + 1;
    if (j < S->r) {
        const double rn = sums[0][j] + sums[1][j] + (S->fac ? sums[4][j] : 0.0);
        const double rho_prev = S->rho[j];
        double beta = 0.0;
        if (S->active[j]) {
            const double xb = sums[2][j], xr = sums[3][j];
            S->xb[j] = xb;
            S->xr[j] = xr;
            S->ring[j][(k - 1) % WINDOW] = S->gamma[j] * S->rho[j];
            const double ssk = S->ss1[j] - (xb + xr);
            double thr = S->tol * S->tol * ssk;
            if (S->absthr[j] > 0.0) thr = S->absthr[j];
            thr = fmax(thr, S->floor_[j]);
            S->thr[j] = thr;
            double sum = 0.0;
            for (int q = 0; q < WINDOW; q++) sum += S->ring[j][q];
            const bool done_w = k >= WINDOW && sum <= thr;
            const bool at_roundoff = rn <= S->floor_[j];
            const bool rel = S->rel2 > 0.0 && rn <= S->rel2 * S->rho0[j];
            // a correction solve of the mixed form also stops once its r'M^-1 r meets the outer
            // test (absthr is that test's threshold; the caller decides on the FP64 residual)
            const bool at_abs = S->absthr[j] > 0.0 && rn <= S->absthr[j];
            if (done_w || at_roundoff || rel || at_abs || !(rn > 0.0)) S->active[j] = 0;
            else beta = rn / S->rho[j];
        }
        S->beta[j] = beta;
        S->rho_new[j] = rn;
        S->rho[j] = rn;
        if (S->active[j]) {
            atomicOr(&any, 1);
            if (S->handover_at) {
                // hand-over test of a mixed solve's FP64 lead: one FP32 sweep (a drop of r'M^-1 r by
                // at most HANDOVER_GAIN) reaches the sweep target 0.01 thr; the next FP64 step is
                // predicted to reach the roundoff floor at the current rate
                const double target = fmax(0.01 * S->thr[j], S->floor_[j]);
                if (!(rn <= HANDOVER_GAIN * target)) atomicOr(&not_ready, 1);
                const double next = rho_prev > 0.0 ? rn * (rn / rho_prev) : INFINITY;
                if (!(next <= S->floor_[j])) atomicAnd(&soon, 0);
            }
        }
        if (k == PROBE_FROM) S->rho_from[j] = rn;
        if (k == EARLY_FROM) S->rho_early[j] = rn;
    }
    __shared__ double rh[RMAX], rem[RMAX];
    if (j < RMAX) {
        rh[j] = 0.0;
        rem[j] = 0.0;
    }
    __syncthreads();
    if (k == PROBE_AT && S->can_escalate && j < S->r && S->active[j] && S->rho_from[j] > 0.0)
        rh[j] = pow(S->rho_new[j] / S->rho_from[j], 1.0 / (double)(PROBE_AT - PROBE_FROM));
    if (k == EARLY_AT && S->can_escalate && j < S->r && S->active[j] && S->rho_early[j] > 0.0) {
        const double rr = pow(S->rho_new[j] / S->rho_early[j], 1.0 / (double)(EARLY_AT - EARLY_FROM));
        const double left = S->maxiter > S->iters + 1 ? (double)(S->maxiter - S->iters - 1) : 0.0;
        const double rm = rr < 1.0 ? fmax(log(S->rho_new[j] / S->thr[j]), 0.0) / log(1.0 / rr) : left;
        rh[j] = rr;
        rem[j] = fmin(rm, left);
    }
    __syncthreads();
    if (j == 0) {
        S->k = k;
        S->iters += 1;
        double rho_hat = 0.0, remaining = 0.0;
        for (int q = 0; q < RMAX; q++) {
            rho_hat = fmax(rho_hat, rh[q]);
            remaining = fmax(remaining, rem[q]);
        }
        if (k == EARLY_AT && S->can_escalate) {
            S->probe_rho = rho_hat;
            S->probe_rem = remaining;
        }
        if (!any) S->done = 1;
        else if (S->iters >= S->maxiter) { S->err = 1; S->done = 1; }
        else if (k == PROBE_AT && S->can_escalate && rho_hat > PROBE_RATIO) { S->err = 2; S->done = 1; }
        else if (S->handover_at && (k >= S->handover_at || (!not_ready && !soon))) { S->err = 3; S->done = 1; }
        if (mirror) publish(S, mirror);
    }
}

__global__ __launch_bounds__(TPB) void k_stop(PcgState* S, const double* __restrict__ dpart, PollMirror* mirror) {
    if (S->done) return;
    stop_rule(S, dpart, mirror);
}

template <class T>
__global__ void k_update_p(T* __restrict__ p, const T* __restrict__ z, size_t n, int r, const PcgState* __restrict__ S) {
    if (S->done) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(i % r);
        p[i] = S->active[j] ? z[i] + (T)S->beta[j] * p[i] : T(0);
    }
}

inline unsigned vgrid(size_t n) { return grid_for(n, TPB, 4096); }

template <class T>
WeightArrays<T>& warr(Absorber& a) {
    if constexpr (std::is_same<T, double>::value) return a.w64; else return a.w32;
}
template <class T>
Scratch<T>& scr(Absorber& a) {
    if constexpr (std::is_same<T, double>::value) return a.s64; else return a.s32;
}

// z = M^-1 r in one launch: Jacobi on the large rows with the partial r'z (slot S_RZL), and in
// workgroup 0 the small block's explicit inverse with the partial r_S'z_S (slot S_RZS; zeros from
// the other workgroups). Used when ms r <= TPB and ms <= SMALL_FUSE_MAX.
template <class T>
__global__ __launch_bounds__(TPB) void k_precond(const T* __restrict__ inv_diag, const T* __restrict__ minv,
                                                 const T* __restrict__ rr, T* __restrict__ z, size_t ms, size_t m,
                                                 int r, double* __restrict__ dpart) {
    __shared__ double red[TPB / WAVE][2][RMAX];
    double rzl[RMAX], rzs[RMAX];
#pragma unroll
    for (int j = 0; j < RMAX; j++) { rzl[j] = 0.0; rzs[j] = 0.0; }
    const size_t n = m * r;
    for (size_t i = ms * r + (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
        const int j = (int)(i % r);
        const T zv = inv_diag[i / r] * rr[i];
        z[i] = zv;
        const double c = (double)rr[i] * (double)zv;
#pragma unroll
        for (int jj = 0; jj < RMAX; jj++)
            if (jj == j) rzl[jj] += c;
    }
    if (blockIdx.x == 0 && threadIdx.x < ms * r) {
        const size_t e = threadIdx.x, k = e / r;
        const int j = (int)(e % r);
        T sm = T(0);
        for (size_t l = 0; l < ms; l++) sm += minv[k * ms + l] * rr[l * r + j];
        z[e] = sm;
        const double c = (double)rr[e] * (double)sm;
#pragma unroll
        for (int jj = 0; jj < RMAX; jj++)
            if (jj == j) rzs[jj] += c;
    }
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int j = 0; j < RMAX; j++)
        if (j < r) {
            const double a = wave_sum(rzl[j]), b = wave_sum(rzs[j]);
            if (lane_id() == 0) { red[w][0][j] = a; red[w][1][j] = b; }
        }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        const int j = threadIdx.x;
        double a = 0.0, b = 0.0;
        for (int ww = 0; ww < TPB / WAVE; ww++) { a += red[ww][0][j]; b += red[ww][1][j]; }
        dpart[((size_t)S_RZL * DOTG + blockIdx.x) * RMAX + j] = a;
        dpart[((size_t)S_RZS * DOTG + blockIdx.x) * RMAX + j] = b;
    }
}

// The escalated factors' part of z = M^-1 r (FP64 only: an absorber with factors solves in FP64).
template <class T>
void precond_factors(Absorber& a, const T* rr, T* z, int r, double* dpart, const uint32_t* done, hipStream_t st) {
    if (a.slope) {
        if constexpr (std::is_same<T, double>::value)
            slope_precond(a, rr, z, r, dpart + (size_t)S_RZF * DOTG * RMAX, done, st);
        else
            throw Error(FH_INTERNAL, "slope absorbers solve in FP64");
        return;
    }
    if (a.facs.empty()) return;
    if constexpr (std::is_same<T, double>::value)
        facs_apply(a, rr, z, r, dpart + (size_t)S_RZF * DOTG * RMAX, done, st);
    else
        throw Error(FH_INTERNAL, "escalated factors are applied in FP64");
}

// z = M^-1 r and the partial r'z of both parts.
template <class T>
void precond(Absorber& a, const T* rr, T* z, int r, double* dpart, const uint32_t* done, hipStream_t st) {
    WeightArrays<T>& W = warr<T>(a);
    const size_t m = a.m, ms = a.ms;
    if (!done && ms <= (size_t)SMALL_FUSE_MAX && ms * (size_t)r <= (size_t)TPB) {
        hipLaunchKernelGGL(k_precond<T>, dim3(DOTG), dim3(TPB), 0, st, W.inv_diag.p, W.minv.p, rr, z, ms, m, r, dpart);
        FH_CHECK_LAUNCH();
        precond_factors(a, rr, z, r, dpart, done, st);
        return;
    }
    if (m > ms) {
        hipLaunchKernelGGL(k_jacobi<T>, dim3(vgrid((m - ms) * r)), dim3(TPB), 0, st, W.inv_diag.p, rr, z, ms, m, r, done);
        hipLaunchKernelGGL(k_dot<T>, dim3(DOTG), dim3(TPB), 0, st, rr, z, ms, m, r, dpart, (int)S_RZL, done);
    }
    if (ms > 0) {
        hipLaunchKernelGGL(k_small_apply<T>, dim3(vgrid(ms * r)), dim3(TPB), 0, st, W.minv.p, rr, z, ms, r, done);
        hipLaunchKernelGGL(k_dot<T>, dim3(DOTG), dim3(TPB), 0, st, rr, z, (size_t)0, ms, r, dpart, (int)S_RZS, done);
    }
    FH_CHECK_LAUNCH();
    precond_factors(a, rr, z, r, dpart, done, st);
}

// Accept the joint factor when its estimated cost is below this share of the projected PCG cost
// (solve.rs:115-117: the cost model's margin).
constexpr double JOINT_MARGIN = 0.8;

struct PcgRun {
    uint64_t iterations = 0;
    bool exact = false;
    bool maxiter_hit = false;
    bool escalate = false;  // stopped at the probe: the caller strengthens the preconditioner
    bool handover = false;  // stopped at handover_at with columns still active (active[] below)
    int32_t active[RMAX] = {};
};

// One PCG solve in precision T from x (x_nonzero: start from x, else from 0).
template <class T>
PcgRun run_pcg(Absorber& a, const T* rhs, T* x, bool x_nonzero, int r, const double* ss1, const double* ss0,
               double tol, uint64_t maxiter, double rel2, const double* absthr, const fh_ctx* ctx,
               bool can_escalate = false, uint64_t handover_at = 0, bool r_ready = false, JointProbe* jp = nullptr) {
    hipStream_t st = a.st;
    Scratch<T>& s = scr<T>(a);
    const size_t m = a.m, ms = a.ms, nv = m * r;
    s.r.reserve(nv); s.z.reserve(nv); s.p.reserve(nv); s.q.reserve(nv);
    s.dpart.reserve((size_t)NSLOT * DOTG * RMAX);
    a.status.reserve((sizeof(PcgState) + 7) / 8);
    PcgState* S = reinterpret_cast<PcgState*>(a.status.p);
    const uint32_t* done = &S->done;
    PcgState h{};
    h.r = r;
    h.tol = tol;
    h.rel2 = rel2;
    h.maxiter = maxiter;
    h.can_escalate = can_escalate ? 1 : 0;
    h.fac = (a.facs.empty() && !a.slope) ? 0 : 1;  // the slope block Jacobi writes the factor slot
    h.handover_at = handover_at;
    h.probe_rho = -1.0;
    h.probe_rem = -1.0;
    for (int j = 0; j < r; j++) {
        h.ss1[j] = ss1[j];
        h.ss0[j] = ss0[j];
        h.floor_[j] = 1e-26 * ss0[j];
        h.active[j] = ss0[j] > 0.0;
        h.absthr[j] = absthr ? absthr[j] : 0.0;
    }
    upload(S, &h, sizeof(PcgState), st);
    hipLaunchKernelGGL(k_zero_slots, dim3(8), dim3(TPB), 0, st, s.dpart.p, 0, (int)NSLOT);
    // r = rhs - S x (already in s.r when r_ready)
    if (r_ready) {
    } else if (x_nonzero) {
        spmv<T>(a, x, s.q.p, r, s, nullptr, st);
        hipLaunchKernelGGL(k_sub<T>, dim3(vgrid(nv)), dim3(TPB), 0, st, rhs, s.q.p, s.r.p, nv);
    } else {
        hipLaunchKernelGGL(k_sub<T>, dim3(vgrid(nv)), dim3(TPB), 0, st, rhs, (const T*)nullptr, s.r.p, nv);
    }
    FH_CHECK_LAUNCH();
    precond<T>(a, s.r.p, s.z.p, r, s.dpart.p, nullptr, st);
    PcgRun out;
    if ((ms == m && a.nkeep > 0) || facs_exact(a)) {
        // M = S exactly (the dense small factor covers S, or a joint factor for the current
        // weights): x += M^-1 (rhs - S x) on the active columns (cg.rs:171-198)
        DBuf<int32_t> mask(r);
        std::vector<int32_t> hm(h.active, h.active + r);
        upload(mask.p, hm.data(), r * sizeof(int32_t), st);
        if constexpr (std::is_same<T, double>::value)
            hipLaunchKernelGGL(k_add_to64<double>, dim3(vgrid(nv)), dim3(TPB), 0, st, x, s.z.p, nv, r, mask.p);
        else
            throw Error(FH_INTERNAL, "exact path runs in FP64");
        FH_CHECK_LAUNCH();
        stream_wait(st);
        out.iterations = 1;
        out.exact = true;
        return out;
    }
    a.poll.ensure();
    volatile PollMirror* mh = a.poll.h;
    mh->err = 0;
    mh->probe_rho = -1.0;
    mh->probe_rem = -1.0;
    mh->iters_done = 0;
    std::atomic_thread_fence(std::memory_order_seq_cst);
    PollMirror* md = a.poll.d;
    hipLaunchKernelGGL(k_start, dim3(1), dim3(TPB), 0, st, S, s.dpart.p, md);
    hipLaunchKernelGGL(k_mask_copy<T>, dim3(vgrid(nv)), dim3(TPB), 0, st, s.z.p, s.p.p, nv, r, S);
    FH_CHECK_LAUNCH();
    WeightArrays<T>& W = warr<T>(a);
    static_assert(SPMV_DOTG == DOTG && RMAX == 16, "the S*p p'q partials use the PCG slot layout");
    const bool small_fused = ms > 0 && ms <= (size_t)SMALL_FUSE_MAX && ms * (size_t)r <= (size_t)TPB;
    auto enqueue = [&](int batch) {
        for (int b = 0; b < batch; b++) {
            spmv<T>(a, s.p.p, s.q.p, r, s, done, st, s.dpart.p + (size_t)S_PQ * DOTG * RMAX);
            if (small_fused) {
                hipLaunchKernelGGL((k_update<T, true>), dim3(DOTG), dim3(TPB), 0, st, x, s.r.p, s.p.p, s.q.p, rhs,
                                   W.inv_diag.p, W.minv.p, s.z.p, ms, m, r, S, s.dpart.p);
            } else {
                hipLaunchKernelGGL((k_update<T, false>), dim3(DOTG), dim3(TPB), 0, st, x, s.r.p, s.p.p, s.q.p, rhs,
                                   W.inv_diag.p, W.minv.p, s.z.p, ms, m, r, S, s.dpart.p);
                if (ms > 0) {
                    hipLaunchKernelGGL(k_small_apply<T>, dim3(vgrid(ms * r)), dim3(TPB), 0, st, W.minv.p, s.r.p, s.z.p,
                                       ms, r, done);
                    hipLaunchKernelGGL(k_dot<T>, dim3(DOTG), dim3(TPB), 0, st, s.r.p, s.z.p, (size_t)0, ms, r,
                                       s.dpart.p, (int)S_RZS, done);
                }
            }
            precond_factors(a, (const T*)s.r.p, s.z.p, r, s.dpart.p, done, st);
            hipLaunchKernelGGL(k_stop, dim3(1), dim3(TPB), 0, st, S, s.dpart.p, md);
            hipLaunchKernelGGL(k_update_p<T>, dim3(vgrid(nv)), dim3(TPB), 0, st, s.p.p, s.z.p, nv, r, S);
            FH_CHECK_LAUNCH();
        }
    };
    // Pipelined polling: the status after iteration k is read while iteration k+1 is already
    // queued, so the device never waits for the host's decision. The device publishes it to pinned
    // host memory (no copy kernel between iterations). Once the stopping rule has fired every
    // kernel of an iteration returns at once, so the one iteration queued past the end costs only
    // its launches.
    struct Poll {
        uint32_t done, err;
        uint64_t iters;
        double probe_rho, probe_rem;
    };
    static_assert(offsetof(PcgState, err) == offsetof(PcgState, done) + 4 &&
                      offsetof(PcgState, iters) == offsetof(PcgState, done) + 8 &&
                      offsetof(PcgState, probe_rem) == offsetof(PcgState, done) + 24,
                  "poll fields are contiguous");
    static_assert(sizeof(Poll) == sizeof(PollMirror), "the mirror holds the poll fields");
    // Waits until the device has published iteration `target` (or the end). A status that does not
    // come while the stream has drained (nothing left that could publish it) is read from the
    // device state instead.
    auto wait_for = [&](uint64_t target) {
        Poll p{};
        auto t_spin = std::chrono::steady_clock::now();
        constexpr uint64_t DONE = 1ull << 63;
        for (uint64_t spins = 0;; spins++) {
            const uint64_t w = mh->iters_done;
            if ((w & ~DONE) >= target || (w & DONE)) {
                std::atomic_thread_fence(std::memory_order_acquire);
                p.done = (w & DONE) ? 1u : 0u;
                p.err = mh->err;
                p.iters = w & ~DONE;
                p.probe_rho = mh->probe_rho;
                p.probe_rem = mh->probe_rem;
                return p;
            }
            if ((spins & 1023) == 1023 &&
                std::chrono::steady_clock::now() - t_spin > std::chrono::milliseconds(20) &&
                hipStreamQuery(st) == hipSuccess) {
                download(&p, &S->done, sizeof(Poll), st);
                return p;
            }
        }
    };
    const auto t0 = std::chrono::steady_clock::now();
    bool probed = jp == nullptr;
    enqueue(1);
    Poll pl{};
    for (uint64_t k = 1;; k++) {
        enqueue(1);
        pl = wait_for(k);
        if (pl.done) break;
        if (ctx && ctx->interrupt && ctx->interrupt(ctx->user)) {
            stream_wait(st);
            throw Error(FH_INTERRUPTED, "interrupted");
        }
        if (!probed && pl.probe_rem >= 0.0) {
            // the early probe: the projected remaining iterations at the measured time per
            // iteration against the joint factor's estimated cost (solve.rs:932-975); the
            // fast-converging rule of the reference first (rho < 0.5 and < 20 iterations)
            probed = true;
            if (!(pl.probe_rho < 0.5 && pl.probe_rem < 20.0)) {
                const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                const double without = pl.probe_rem * el / (double)std::max<uint64_t>(pl.iters, 1);
                if (joint_build_estimate(a) <= JOINT_MARGIN * without) {
                    jp->fired = true;
                    jp->without = without;
                    jp->rho = pl.probe_rho;
                    jp->remaining = pl.probe_rem;
                    // the iteration queued after this one completes too; x holds every finished step
                    download(&pl, &S->done, sizeof(Poll), st);
                    if (pl.done) jp->fired = false;
                    break;
                }
            }
        }
    }
    out.iterations = pl.iters;
    out.maxiter_hit = pl.err == 1;
    out.escalate = pl.err == 2;
    out.handover = pl.err == 3;
    if (out.handover) download(out.active, S->active, sizeof(out.active), st);
    return out;
}

template <class T>
__global__ void k_dot_pair(const T* __restrict__ a, const T* __restrict__ b, size_t n, int r, double* __restrict__ out) {
    // out[blockIdx.x * r + j] partial sums (fixed grid)
    __shared__ double red[TPB / WAVE][RMAX];
    double acc[RMAX];
#pragma unroll
    for (int j = 0; j < RMAX; j++) acc[j] = 0.0;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n * r; i += (size_t)gridDim.x * TPB) {
        const int j = (int)(i % r);
        const double v = (double)a[i] * (double)b[i];
#pragma unroll
        for (int jj = 0; jj < RMAX; jj++)
            if (jj == j) acc[jj] += v;
    }
#pragma unroll
    for (int j = 0; j < RMAX; j++)
        if (j < r) {
            double s = wave_sum(acc[j]);
            if (lane_id() == 0) red[threadIdx.x / WAVE][j] = s;
        }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        double s = 0.0;
        for (int ww = 0; ww < TPB / WAVE; ww++) s += red[ww][threadIdx.x];
        out[(size_t)blockIdx.x * r + threadIdx.x] = s;
    }
}

// Per-column dot products x_i'y_i of pairs of m x r FP64 blocks (fixed order), read back in one
// transfer: out[i][j].
std::vector<std::vector<double>> dots64(Absorber& a, std::initializer_list<std::pair<const double*, const double*>> xy,
                                        int r) {
    const size_t np = xy.size();
    DBuf<double> part(np * DOTG * r);
    size_t i = 0;
    for (auto& pr : xy) {
        hipLaunchKernelGGL(k_dot_pair<double>, dim3(DOTG), dim3(TPB), 0, a.st, pr.first, pr.second, (size_t)a.m, r,
                           part.p + i * DOTG * r);
        FH_CHECK_LAUNCH();
        i++;
    }
    std::vector<double> h(np * DOTG * r);
    download(h.data(), part.p, h.size() * sizeof(double), a.st);
    std::vector<std::vector<double>> out(np, std::vector<double>(r, 0.0));
    for (size_t k = 0; k < np; k++)
        for (int q = 0; q < DOTG; q++)
            for (int j = 0; j < r; j++) out[k][j] += h[(k * DOTG + q) * r + j];
    return out;
}
std::vector<double> dots64(Absorber& a, const double* x, const double* y, int r) { return dots64(a, {{x, y}}, r)[0]; }

}  // namespace

// Mirrors come from pinned pages shared by the process (a pinned allocation costs far more than a
// PCG iteration, and absorbers are often created per call): a slot is taken on an absorber's first
// solve and returned when it is destroyed.
namespace {
std::mutex g_poll_mu;
std::vector<std::pair<PollMirror*, PollMirror*>> g_poll_free;  // (host, device) views
}  // namespace

HostPoll::~HostPoll() {
    if (!h) return;
    std::lock_guard<std::mutex> lk(g_poll_mu);
    g_poll_free.emplace_back(h, d);
}

void HostPoll::ensure() {
    if (h) return;
    std::lock_guard<std::mutex> lk(g_poll_mu);
    if (g_poll_free.empty()) {
        constexpr size_t SLOTS = 64;
        void* p = nullptr;
        FH_HIP(hipHostMalloc(&p, SLOTS * sizeof(PollMirror), hipHostMallocCoherent | hipHostMallocMapped));
        void* dp = nullptr;
        FH_HIP(hipHostGetDevicePointer(&dp, p, 0));
        for (size_t i = SLOTS; i-- > 0;)
            g_poll_free.emplace_back(static_cast<PollMirror*>(p) + i, static_cast<PollMirror*>(dp) + i);
    }
    h = g_poll_free.back().first;
    d = g_poll_free.back().second;
    g_poll_free.pop_back();
}

PcgResult pcg_fp64(Absorber& a, const double* rhs, double* x, bool x_maybe_nonzero, int r, const double* ss1,
                   const double* ss0, double tol, uint64_t maxiter, const fh_ctx* ctx, bool can_escalate,
                   JointProbe* jp) {
    // a warm start may still be all zero: test it on the device
    bool nz = false;
    if (x_maybe_nonzero)
        for (double v : dots64(a, x, x, r)) nz |= v != 0.0;
    PcgRun run = run_pcg<double>(a, rhs, x, nz, r, ss1, ss0, tol, maxiter, 0.0, nullptr, ctx, can_escalate, 0, false,
                                 can_escalate ? jp : nullptr);
    PcgResult res;
    res.iterations = run.iterations;
    res.exact = run.exact;
    res.rc = run.maxiter_hit ? FH_CONVERGENCE : FH_OK;
    res.escalate = run.escalate;
    return res;
}

PcgResult pcg_mixed(Absorber& a, const double* rhs, double* x, bool x_maybe_nonzero, int r, const double* ss1,
                    const double* ss0, double tol, uint64_t maxiter, const fh_ctx* ctx, bool can_escalate,
                    JointProbe* jp) {
    PcgResult res;
    // exact dense path (nothing to iterate), or escalated sparse factors, whose grounded inverse of
    // a slowly mixing system is too ill-conditioned for FP32 sweeps (and is applied in FP64 only)
    if (a.ms == a.m || !a.facs.empty())
        return pcg_fp64(a, rhs, x, x_maybe_nonzero, r, ss1, ss0, tol, maxiter, ctx, can_escalate && a.ms < a.m, jp);
    ensure_w32(a);
    hipStream_t st = a.st;
    const size_t m = a.m, nv = m * r;
    Scratch<double>& s64 = a.s64;
    Scratch<float>& s32 = a.s32;
    s64.r.reserve(nv); s64.z.reserve(nv); s64.q.reserve(nv);
    s64.dpart.reserve((size_t)NSLOT * DOTG * RMAX);
    s32.x.reserve(nv); s32.rhs.reserve(nv);
    bool nz = false;
    if (x_maybe_nonzero)
        for (double v : dots64(a, x, x, r)) nz |= v != 0.0;
    std::vector<int32_t> act(r);
    for (int j = 0; j < r; j++) act[j] = ss0[j] > 0.0;
    uint64_t total_iters = 0;
    // The FP64 outer residual of the start and its preconditioned norm decide whether an FP64 lead
    // runs. A lead keeps the reference's FP64 iteration, with its stopping rule, while the next
    // step is predicted to reach the roundoff floor (the difficult panel: the preconditioner nearly
    // solves it, two FP64 steps against two FP32 sweeps), and hands over to the FP32 sweeps once
    // one sweep can reach the target, at the latest after mixed_lead iterations. A warm start that
    // is already that close (IRLS) goes straight to the sweeps.
    auto outer_residual = [&] {
        if (nz) spmv<double>(a, x, s64.q.p, r, s64, nullptr, st);
        hipLaunchKernelGGL(k_sub<double>, dim3(vgrid(nv)), dim3(TPB), 0, st, rhs, nz ? s64.q.p : nullptr, s64.r.p, nv);
        FH_CHECK_LAUNCH();
        precond<double>(a, s64.r.p, s64.z.p, r, s64.dpart.p, nullptr, st);
        return dots64(a, {{s64.r.p, s64.z.p}, {x, rhs}, {x, s64.r.p}}, r);
    };
    auto thr_of = [&](int j, const std::vector<std::vector<double>>& d) {
        return std::max(tol * tol * (ss1[j] - (d[1][j] + d[2][j])), 1e-26 * ss0[j]);
    };
    bool have_r = false;
    std::vector<std::vector<double>> dd;
    if (a.mixed_lead > 0) {
        // from x = 0 the residual is the right-hand side, whose r'M^-1 r is of the order of the
        // column's sum of squares: never within one sweep of the target at a tolerance the
        // reference would use, so a cold start goes to the lead without the readback
        bool ready = false;
        if (nz) {
            dd = outer_residual();
            ready = true;
            for (int j = 0; j < r; j++)
                if (act[j] && dd[0][j] > 0.0) ready &= dd[0][j] <= HANDOVER_GAIN * 0.01 * thr_of(j, dd);
        }
        if (ready) {
            have_r = true;
        } else {
            PcgRun lead = run_pcg<double>(a, rhs, x, nz, r, ss1, ss0, tol, maxiter, 0.0, nullptr, ctx, false,
                                          a.mixed_lead, nz);
            total_iters += lead.iterations;
            if (!lead.handover) {
                res.iterations = std::max<uint64_t>(total_iters, 1);
                res.exact = lead.exact;
                res.rc = lead.maxiter_hit ? FH_CONVERGENCE : FH_OK;
                return res;
            }
            for (int j = 0; j < r; j++) act[j] = act[j] && lead.active[j];
            nz = true;
            dd.clear();  // s64.r holds the lead's FP64 residual (the PCG's own recursion, as in cg.rs)
            have_r = true;
        }
    }
    DBuf<int32_t> mask(r);
    bool finished = false;
    // Inner FP32 solve: stops at its column thresholds (the window estimate, or r'M^-1 r below
    // the outer test's threshold) or when r'M^-1 r has dropped by rel^2. When every column's
    // threshold lies within 1e-10 of its r'M^-1 r, one sweep can finish (rel 1e-5); beyond that
    // two sweeps are needed anyway, and a first stop at 1e-8 costs fewer iterations in total than
    // one at 1e-10.
    const double rel_one = 1e-5, rel_split = 1e-4;
    for (int sweep = 0; sweep < 40; sweep++) {
        // FP64 outer residual r = rhs - S x (from the FP64 operator; the lead's residual at the
        // first sweep) and its preconditioned norm
        std::vector<std::vector<double>> d;
        if (!have_r) {
            d = outer_residual();
        } else if (!dd.empty()) {
            d = std::move(dd);  // the start's residual, already reduced
            dd.clear();
        } else {
            precond<double>(a, s64.r.p, s64.z.p, r, s64.dpart.p, nullptr, st);
            d = dots64(a, {{s64.r.p, s64.z.p}, {x, rhs}, {x, s64.r.p}}, r);
        }
        have_r = false;
        const std::vector<double>&rz = d[0], &xb = d[1], &xr = d[2];
        std::vector<double> ss1c(r), ss0c(r), absthr(r);
        bool any = false;
        for (int j = 0; j < r; j++) {
            if (!act[j]) { ss0c[j] = 0.0; ss1c[j] = 0.0; continue; }
            const double ssk = ss1[j] - (xb[j] + xr[j]);
            const double thr = std::max(tol * tol * ssk, 1e-26 * ss0[j]);
            // the error ||x* - x||_S^2 is estimated by r'M^-1 r on the FP64 residual; a margin
            // of 100 covers the preconditioner's misestimate
            if (sweep > 0 && rz[j] <= 0.01 * thr) { act[j] = 0; continue; }
            if (!(rz[j] > 0.0)) { act[j] = 0; continue; }
            any = true;
            ss0c[j] = ss0[j];
            ss1c[j] = ss1[j];
            absthr[j] = 0.01 * thr;
        }
        if (!any) {
            finished = true;
            break;
        }
        double need = 1.0;
        for (int j = 0; j < r; j++)
            if (act[j]) need = std::min(need, absthr[j] / rz[j]);
        const double rel = need >= rel_one * rel_one ? rel_one : rel_split;
        res.refinements++;
        // FP32 correction solve S d = r from d = 0
        hipLaunchKernelGGL((k_convert<float, double>), dim3(vgrid(nv)), dim3(TPB), 0, st, s64.r.p, s32.rhs.p, nv);
        FH_HIP(hipMemsetAsync(s32.x.p, 0, nv * sizeof(float), st));
        FH_CHECK_LAUNCH();
        PcgRun run = run_pcg<float>(a, s32.rhs.p, s32.x.p, false, r, ss1c.data(), ss0c.data(), tol,
                                    maxiter > total_iters ? maxiter - total_iters : 1, rel * rel, absthr.data(), ctx,
                                    can_escalate && sweep == 0, 0, false, can_escalate && sweep == 0 ? jp : nullptr);
        total_iters += run.iterations;
        upload(mask.p, act.data(), r * sizeof(int32_t), st);
        hipLaunchKernelGGL(k_add_to64<float>, dim3(vgrid(nv)), dim3(TPB), 0, st, x, s32.x.p, nv, r, mask.p);
        FH_CHECK_LAUNCH();
        nz = true;
        if (run.escalate || (jp && jp->fired)) {  // x holds the progress so far; the caller restarts
            res.escalate = run.escalate;
            finished = true;
            break;
        }
        if (run.maxiter_hit || total_iters >= maxiter) {
            res.rc = FH_CONVERGENCE;
            finished = true;
            break;
        }
    }
    if (!finished) {
        // the refinement sweeps stopped gaining (an FP32-limited preconditioner): finish in FP64
        // from the current iterate
        PcgResult f = pcg_fp64(a, rhs, x, true, r, ss1, ss0, tol, maxiter > total_iters ? maxiter - total_iters : 1,
                               ctx, false);
        total_iters += f.iterations;
        res.rc = f.rc;
        res.exact = f.exact;
    }
    stream_wait(st);
    res.iterations = std::max<uint64_t>(total_iters, 1);
    return res;
}

}  // namespace fh
