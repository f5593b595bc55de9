// The reduced operator and the row passes on the GPU (operator.rs:894-1179):
//   S*p   = segment pass (t_s = sum_c W_c u_c / n1_s) followed by one level pass per block
//           over its level-major transpose (y_l = sum_{c in l} W_c (u_c - t_{s(c)})); blocks
//           with long levels are cut into FE1 buckets, so a level's sum is a fixed-order sum of
//           per-bucket partials (no float atomics anywhere);
//   rhs   = segment pass over the rows (cell sums, b1, sum w v^2, sum w v; z_c to a scratch)
//           + one level pull of z_c per block;
//   back-substitution: alpha_s = (b1_s - sum_c W_c u_c)/n1_s, v_i -= alpha_s + u_c in place.
// Every pass assigns L lanes to a segment (or level item), L = next power of two of the mean
// length over the unroll depth, so cell reads are coalesced; all sums are in a fixed order
// (run-to-run reproducible).
#pragma once
#include <algorithm>
#include <type_traits>

#include "fh_internal.hpp"

namespace fh {

namespace {

constexpr int WG = 256;
constexpr int WAVES = WG / WAVE;

template <class T>
__device__ __forceinline__ T shfl_xor_t(T v, int o) { return __shfl_xor(v, o, WAVE); }

// Streamed (read-once) loads: non-temporal, so the index streams do not evict the gathered
// vectors (p, t) from L2.
template <class T>
__device__ __forceinline__ T ldnt(const T* p) { return __builtin_nontemporal_load(p); }

// row[j] = src[j] (j < r) for a valid lane, else 0. When r == RM the row is read with 16-byte
// (or 8-byte) vector loads: one gather instruction per row instead of one per column, which
// matters because random gathers are bounded by instructions x distinct lines, not by bytes.
template <class T, int RM>
__device__ __forceinline__ void ld_row(const T* __restrict__ src, int r, bool v, T (&row)[RM]) {
#pragma unroll
    for (int j = 0; j < RM; j++) row[j] = T(0);
    if (!v) return;
    if constexpr (std::is_same<T, float>::value && RM % 4 == 0) {
        if (r == RM) {
#pragma unroll
            for (int q = 0; q < RM / 4; q++) {
                const float4 x = reinterpret_cast<const float4*>(src)[q];
                row[4 * q] = x.x; row[4 * q + 1] = x.y; row[4 * q + 2] = x.z; row[4 * q + 3] = x.w;
            }
            return;
        }
    } else if constexpr (std::is_same<T, float>::value && RM == 2) {
        if (r == 2) {
            const float2 x = *reinterpret_cast<const float2*>(src);
            row[0] = x.x; row[1] = x.y;
            return;
        }
    } else if constexpr (std::is_same<T, double>::value && RM % 2 == 0) {
        if (r == RM) {
#pragma unroll
            for (int q = 0; q < RM / 2; q++) {
                const double2 x = reinterpret_cast<const double2*>(src)[q];
                row[2 * q] = x.x; row[2 * q + 1] = x.y;
            }
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < RM; j++)
        if (j < r) row[j] = src[j];
}

// Unroll depth of the hoisted passes: positions per lane whose gathers are in flight at once,
// sized so the gathered rows stay within ~64 VGPRs (no spills at 16 FP64 columns).
template <class T, int RM>
constexpr int gather_trips() {
    constexpr int w = RM * (int)sizeof(T) / 4;  // VGPRs per row
    return w <= 4 ? 8 : (w <= 8 ? 4 : (w <= 16 ? 2 : 1));
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, WAVE));
    return v;
}

template <class T>
__device__ __forceinline__ T gsum(T v, int L) {
    for (int o = L >> 1; o > 0; o >>= 1) v += shfl_xor_t(v, o);
    return v;
}

template <class T, int RM>
__device__ __forceinline__ void gather_u(const OpArgs<T>& A, const T* __restrict__ p, uint32_t c, int r, T (&u)[RM]) {
#pragma unroll
    for (int j = 0; j < RM; j++) u[j] = T(0);
    for (int b = 0; b < A.nb; b++) {
        const size_t k = (size_t)(A.boff[b] + A.cell_codes[(size_t)c * A.nb + b]) * r;
#pragma unroll
        for (int j = 0; j < RM; j++)
            if (j < r) u[j] += p[k + j];
    }
}

// ---------------------------------------------------------------- S*p, segment pass
// t_s = sum_{c in s} W_c u_c / n1_s for any block count (runtime loop over the blocks).
template <class T, int RM>
__global__ __launch_bounds__(WG) void k_seg_t_any(OpArgs<T> A, const T* __restrict__ p, int r, int L,
                                                  T* __restrict__ t, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int spw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * spw; base < A.g1; base += nwt * spw) {
        const size_t s = base + gl;
        const bool sv = s < A.g1;
        const uint32_t c0 = sv ? A.seg_off[s] : 0, c1 = sv ? A.seg_off[s + 1] : 0;
        T tv[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) tv[j] = T(0);
        for (uint32_t c = c0 + li; c < c1; c += L) {
            T u[RM];
            gather_u<T, RM>(A, p, c, r, u);
            const T wc = A.cw ? A.cw[c] : T(1);
#pragma unroll
            for (int j = 0; j < RM; j++) tv[j] += wc * u[j];
        }
        const T inv = sv ? A.inv_n1[s] : T(0);
#pragma unroll
        for (int j = 0; j < RM; j++) tv[j] = gsum(tv[j], L) * inv;
        if (sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) t[s * r + j] = tv[j];
    }
}

// t_s = sum_{c in s} W_c u_c / n1_s with NB (compile-time block count) and U = 8 trips per lane with every index load issued before the
// gathers, so a lane keeps ~8 x NB independent loads in flight (the passes are latency-bound
// with fewer: 32 resident waves per CU at most).
template <class T, int RM, int NB>
__global__ __launch_bounds__(WG) void k_seg_t(OpArgs<T> A, const T* __restrict__ p, int r, int L, T* __restrict__ t,
                                              const uint32_t* __restrict__ done) {
    if (done && *done) return;
    constexpr int U = gather_trips<T, RM>();
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int spw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * spw; base < A.g1; base += nwt * spw) {
        const size_t s = base + gl;
        const bool sv = s < A.g1;
        const uint32_t c0 = sv ? A.seg_off[s] : 0, c1 = sv ? A.seg_off[s + 1] : 0;
        const uint32_t tmax = wave_max_u32((c1 - c0 + L - 1) / L);
        const T inv = sv ? A.inv_n1[s] : T(0);
        T tv[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) tv[j] = T(0);
        for (uint32_t it = 0; it < tmax; it += U) {
            uint32_t code[U][NB];
            T wc[U];
            bool v[U];
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t c = c0 + (it + h) * L + li;
                v[h] = c < c1;
#pragma unroll
                for (int b = 0; b < NB; b++) code[h][b] = v[h] ? ldnt(&A.cell_codes[(size_t)c * NB + b]) : 0u;
                wc[h] = v[h] ? (A.cw ? ldnt(&A.cw[c]) : T(1)) : T(0);
            }
#pragma unroll
            for (int h = 0; h < U; h++) {
                T u[RM];
#pragma unroll
                for (int j = 0; j < RM; j++) u[j] = T(0);
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    T row[RM];
                    ld_row<T, RM>(p + (size_t)(A.boff[b] + code[h][b]) * r, r, v[h], row);
#pragma unroll
                    for (int j = 0; j < RM; j++) u[j] += row[j];
                }
#pragma unroll
                for (int j = 0; j < RM; j++) tv[j] += wc[h] * u[j];
            }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) tv[j] = gsum(tv[j], L) * inv;
        if (sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) t[s * r + j] = tv[j];
    }
}

// ---------------------------------------------------------------- S*p, dense small rows
// Partial sums over a column range of y_S[k] = sum_l C[k][l] p_L[l] - sum_s D[k][s] t_s for the
// small levels k in chunk blockIdx.y (KC levels): columns are the g1 segments (D, t) followed by
// the m - ms large coordinates (C, p_L). Each workgroup writes part[blockIdx.x][k][j].
template <class T, int RM>
constexpr int small_chunk() { return (sizeof(T) == 8 ? 32 : 64) / RM < 32 ? (sizeof(T) == 8 ? 32 : 64) / RM : 32; }

template <class T, int RM>
__global__ __launch_bounds__(WG) void k_small_dense(const T* __restrict__ D, const T* __restrict__ t,
                                                    const T* __restrict__ C, const T* __restrict__ pL, size_t g1,
                                                    size_t mL, int ms, int r, T* __restrict__ part,
                                                    const uint32_t* __restrict__ done) {
    if (done && *done) return;
    constexpr int KC = small_chunk<T, RM>();
    __shared__ T red[WAVES][KC][RM];
    const int k0 = blockIdx.y * KC;
    T acc[KC][RM];
#pragma unroll
    for (int kk = 0; kk < KC; kk++)
#pragma unroll
        for (int j = 0; j < RM; j++) acc[kk][j] = T(0);
    const size_t nc = g1 + mL;
    for (size_t col = (size_t)blockIdx.x * WG + threadIdx.x; col < nc; col += (size_t)gridDim.x * WG) {
        const bool seg = col < g1;
        const size_t c = seg ? col : col - g1;
        const T* M = seg ? D : C;
        const size_t ld = seg ? g1 : mL;
        T v[RM];
        ld_row<T, RM>((seg ? t : pL) + c * r, r, true, v);
        if (seg)
#pragma unroll
            for (int j = 0; j < RM; j++) v[j] = -v[j];
#pragma unroll
        for (int kk = 0; kk < KC; kk++) {
            const int k = k0 + kk;
            if (k < ms) {
                const T mk = M[(size_t)k * ld + c];
#pragma unroll
                for (int j = 0; j < RM; j++) acc[kk][j] += mk * v[j];
            }
        }
    }
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int kk = 0; kk < KC; kk++)
#pragma unroll
        for (int j = 0; j < RM; j++) {
            const T sm = wave_sum(acc[kk][j]);
            if (lane_id() == 0) red[w][kk][j] = sm;
        }
    __syncthreads();
    for (int e = threadIdx.x; e < KC * r; e += WG) {
        const int kk = e / r, j = e % r;
        if (k0 + kk >= ms) continue;
        T sm = T(0);
#pragma unroll
        for (int ww = 0; ww < WAVES; ww++) sm += red[ww][kk][j];
        part[(size_t)blockIdx.x * ms * r + (size_t)(k0 + kk) * r + j] = sm;
    }
}

// y_S[k][j] = nr_k p_S[k][j] + sum over the workgroup partials (one wave per entry, fixed order).
template <class T>
__global__ __launch_bounds__(WG) void k_small_dense_final(const T* __restrict__ part, unsigned nparts, int ms, int r,
                                                          const T* __restrict__ nrs, const T* __restrict__ p,
                                                          T* __restrict__ y, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const size_t len = (size_t)ms * r;
    for (size_t e = (size_t)blockIdx.x * WAVES + threadIdx.x / WAVE; e < len; e += (size_t)gridDim.x * WAVES) {
        T sm = T(0);
        for (unsigned q = lane_id(); q < nparts; q += WAVE) sm += part[(size_t)q * len + e];
        sm = wave_sum(sm);
        if (lane_id() == 0) y[e] = nrs[e / r] * p[e] + sm;
    }
}

struct OtherOffs {
    uint32_t off[NB_MAX];
};

// ---------------------------------------------------------------- S*p, level pass of one block
// y_l = sum_{c in l} W_c (u_c - t_{s(c)}) over the level-major records of item
// (bucket, level); with several buckets the item sums go to ypart and k_bucket_reduce adds
// them in bucket order.
template <class T, int RM, int NO>
__global__ __launch_bounds__(WG) void k_lvl_spmv(const uint32_t* __restrict__ loff, const uint32_t* __restrict__ lseg,
                                                 const uint32_t* __restrict__ lother, const T* __restrict__ lw,
                                                 size_t g, uint32_t gb, uint32_t off_b, OtherOffs oo, int nother,
                                                 const T* __restrict__ p, const T* __restrict__ t, int r, int L,
                                                 T* __restrict__ y, T* __restrict__ ypart,
                                                 const uint32_t* __restrict__ done) {
    if (done && *done) return;
    constexpr int U = gather_trips<T, RM>();
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * gpw; base < g; base += nwt * gpw) {
        const size_t it_ = base + gl;  // item = bucket * gb + level
        const bool valid = it_ < g;
        const size_t lv = valid ? it_ % gb : 0;
        const uint32_t j0 = valid ? loff[it_] : 0, j1 = valid ? loff[it_ + 1] : 0;
        const uint32_t tmax = wave_max_u32((j1 - j0 + L - 1) / L);
        T pb[RM], acc[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) {
            pb[j] = (valid && j < r) ? p[(size_t)(off_b + lv) * r + j] : T(0);
            acc[j] = T(0);
        }
        for (uint32_t it = 0; it < tmax; it += U) {
            // indices of U positions per lane first, then the gathers
            uint32_t sg[U], oc[U][NO > 0 ? NO : 1];
            bool pv[U];
            T wc[U];
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t pos = j0 + (it + h) * L + li;
                pv[h] = pos < j1;
                sg[h] = pv[h] ? ldnt(&lseg[pos]) : 0u;
#pragma unroll
                for (int o = 0; o < NO; o++)
                    oc[h][o] = (pv[h] && o < nother) ? ldnt(&lother[(size_t)pos * nother + o]) : 0u;
                wc[h] = pv[h] ? (lw ? ldnt(&lw[pos]) : T(1)) : T(0);
            }
#pragma unroll
            for (int h = 0; h < U; h++) {
                T u[RM], z[RM];
#pragma unroll
                for (int j = 0; j < RM; j++) u[j] = pb[j];
#pragma unroll
                for (int o = 0; o < NO; o++) {
                    T row[RM];
                    ld_row<T, RM>(p + (size_t)(oo.off[o] + oc[h][o]) * r, r, pv[h] && o < nother, row);
#pragma unroll
                    for (int j = 0; j < RM; j++) u[j] += row[j];
                }
                T tr[RM];
                ld_row<T, RM>(t + (size_t)sg[h] * r, r, pv[h], tr);
#pragma unroll
                for (int j = 0; j < RM; j++) {
                    z[j] = (pv[h] && j < r) ? wc[h] * (u[j] - tr[j]) : T(0);
                    acc[j] += z[j];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) acc[j] = gsum(acc[j], L);
        if (valid && li == 0) {
            T* dst = ypart ? ypart + it_ * r : y + (size_t)(off_b + lv) * r;
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) dst[j] = acc[j];
        }
    }
}

// y[l] = sum over the buckets of ypart[k*gb + l] (fixed order): one thread per output for a few
// buckets; one wave per output (lane-strided sums, fixed shuffle tree) for many.
template <class T>
__global__ __launch_bounds__(WG) void k_bucket_reduce_wave(const T* __restrict__ ypart, uint32_t nbk, size_t gb, int r,
                                                           T* __restrict__ y, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const size_t nout = gb * r;
    for (size_t e = (size_t)blockIdx.x * WAVES + threadIdx.x / WAVE; e < nout; e += (size_t)gridDim.x * WAVES) {
        T s = T(0);
        for (uint32_t k = lane_id(); k < nbk; k += WAVE) s += ypart[(size_t)k * nout + e];
        s = wave_sum(s);
        if (lane_id() == 0) y[e] = s;
    }
}
template <class T>
__global__ void k_bucket_reduce(const T* __restrict__ ypart, uint32_t nbk, size_t gb, int r, T* __restrict__ y,
                                const uint32_t* __restrict__ done) {
    if (done && *done) return;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < gb * r; e += (size_t)gridDim.x * blockDim.x) {
        T s = T(0);
        for (uint32_t k = 0; k < nbk; k++) s += ypart[(size_t)k * gb * r + e];
        y[e] = s;
    }
}

// ---------------------------------------------------------------- rhs pass (FP64)
template <int RM>
__device__ __forceinline__ void cell_sums(const OpArgs<double>& A, const double* __restrict__ rows,
                                          const double* __restrict__ wr, uint32_t c, int r, double (&sc)[RM],
                                          double (&wsq)[RM], double (&wsum)[RM], bool stats) {
#pragma unroll
    for (int j = 0; j < RM; j++) sc[j] = 0.0;
    const size_t r0 = A.cell_obs_off ? A.cell_obs_off[c] : c, r1 = A.cell_obs_off ? A.cell_obs_off[c + 1] : c + 1;
    for (size_t i = r0; i < r1; i++) {
        const double w = wr ? wr[i] : 1.0;
#pragma unroll
        for (int j = 0; j < RM; j++)
            if (j < r) {
                const double v = rows[i * r + j];
                const double wv = w * v;
                sc[j] += wv;
                if (stats) {
                    wsq[j] += wv * v;
                    wsum[j] += wv;
                }
            }
    }
}

// rhs of a block: pull of z_c along the level-major lists of item (bucket, level).
template <int RM>
__global__ __launch_bounds__(WG) void k_lvl_pull(const uint32_t* __restrict__ loff, const uint32_t* __restrict__ lperm,
                                                 size_t g, uint32_t gb, const double* __restrict__ zc, int r, int L,
                                                 double* __restrict__ out, double* __restrict__ ypart) {
    constexpr int U = gather_trips<double, RM>();
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * gpw; base < g; base += nwt * gpw) {
        const size_t it_ = base + gl;  // item = bucket * gb + level
        const bool valid = it_ < g;
        const uint32_t j0 = valid ? loff[it_] : 0, j1 = valid ? loff[it_ + 1] : 0;
        const uint32_t tmax = wave_max_u32((j1 - j0 + L - 1) / L);
        double acc[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) acc[j] = 0.0;
        for (uint32_t it = 0; it < tmax; it += U) {
            uint32_t c[U];
            bool pv[U];
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t pos = j0 + (it + h) * L + li;
                pv[h] = pos < j1;
                c[h] = pv[h] ? ldnt(&lperm[pos]) : 0u;
            }
#pragma unroll
            for (int h = 0; h < U; h++) {
                double z[RM];
                ld_row<double, RM>(zc + (size_t)c[h] * r, r, pv[h], z);
#pragma unroll
                for (int j = 0; j < RM; j++) acc[j] += z[j];
            }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) acc[j] = gsum(acc[j], L);
        if (valid && li == 0) {
            double* dst = ypart ? ypart + it_ * r : out + (it_ % gb) * r;
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) dst[j] = acc[j];
        }
    }
}

// ---------------------------------------------------------------- back-substitution (FP64)
// Any block count (runtime loop over the blocks, u gathered twice).
template <int RM>
__global__ __launch_bounds__(WG) void k_resid_seg_any(OpArgs<double> A, double* __restrict__ rows, int r, int L,
                                                  const double* __restrict__ x, const double* __restrict__ b1,
                                                  double* __restrict__ alpha) {
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int spw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * spw; base < A.g1; base += nwt * spw) {
        const size_t s = base + gl;
        const bool sv = s < A.g1;
        const uint32_t c0 = sv ? A.seg_off[s] : 0, c1 = sv ? A.seg_off[s + 1] : 0;
        const uint32_t tmax = wave_max_u32((c1 - c0 + L - 1) / L);
        double av[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = 0.0;
        for (uint32_t it = 0; it < tmax; it++) {
            const uint32_t c = c0 + it * L + li;
            if (c < c1 && A.nb > 0) {
                double u[RM];
                gather_u<double, RM>(A, x, c, r, u);
                const double wc = A.cw ? A.cw[c] : 1.0;
#pragma unroll
                for (int j = 0; j < RM; j++) av[j] += wc * u[j];
            }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = gsum(av[j], L);
        const double inv = sv ? A.inv_n1[s] : 0.0;
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = sv && j < r ? (b1[s * r + j] - av[j]) * inv : 0.0;
        if (alpha && sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) alpha[s * r + j] = av[j];
        for (uint32_t it = 0; it < tmax; it++) {
            const uint32_t c = c0 + it * L + li;
            if (c < c1) {
                double u[RM];
                if (A.nb > 0) gather_u<double, RM>(A, x, c, r, u);
                else
#pragma unroll
                    for (int j = 0; j < RM; j++) u[j] = 0.0;
#pragma unroll
                for (int j = 0; j < RM; j++) u[j] += av[j];
                const size_t r0 = A.cell_obs_off ? A.cell_obs_off[c] : c;
                const size_t r1 = A.cell_obs_off ? A.cell_obs_off[c + 1] : c + 1;
                for (size_t i = r0; i < r1; i++)
#pragma unroll
                    for (int j = 0; j < RM; j++)
                        if (j < r) rows[i * r + j] -= u[j];
            }
        }
    }
}

// Back-substitution with a compile-time block count: U trips per lane with the index loads
// issued before the gathers; u_c stays in registers between the alpha sum and the row update
// when a segment fits one trip (the common case), so x is gathered once per cell.

template <int RM, int NB>
__global__ __launch_bounds__(WG) void k_resid_seg(OpArgs<double> A, double* __restrict__ rows, int r, int L,
                                                  const double* __restrict__ x, const double* __restrict__ b1,
                                                  double* __restrict__ alpha) {
    constexpr int U = gather_trips<double, RM>();
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int spw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * spw; base < A.g1; base += nwt * spw) {
        const size_t s = base + gl;
        const bool sv = s < A.g1;
        const uint32_t c0 = sv ? A.seg_off[s] : 0, c1 = sv ? A.seg_off[s + 1] : 0;
        const uint32_t tmax = wave_max_u32((c1 - c0 + L - 1) / L);
        double av[RM], uk[U][RM];
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = 0.0;
        auto gather = [&](uint32_t it, double (&u)[U][RM], double (&wc)[U]) {
            uint32_t code[U][NB > 0 ? NB : 1];
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t c = c0 + (it + h) * L + li;
                const bool v = c < c1;
#pragma unroll
                for (int b = 0; b < NB; b++) code[h][b] = v ? ldnt(&A.cell_codes[(size_t)c * NB + b]) : 0u;
                wc[h] = v ? (A.cw ? A.cw[c] : 1.0) : 0.0;
            }
#pragma unroll
            for (int h = 0; h < U; h++) {
                const bool v = c0 + (it + h) * L + li < c1;
#pragma unroll
                for (int j = 0; j < RM; j++) u[h][j] = 0.0;
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    double row[RM];
                    ld_row<double, RM>(x + (size_t)(A.boff[b] + code[h][b]) * r, r, v, row);
#pragma unroll
                    for (int j = 0; j < RM; j++) u[h][j] += row[j];
                }
            }
        };
        for (uint32_t it = 0; it < tmax; it += U) {
            double wc[U];
            gather(it, uk, wc);
#pragma unroll
            for (int h = 0; h < U; h++)
#pragma unroll
                for (int j = 0; j < RM; j++) av[j] += wc[h] * uk[h][j];
        }
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = gsum(av[j], L);
        const double inv = sv ? A.inv_n1[s] : 0.0;
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = sv && j < r ? (b1[s * r + j] - av[j]) * inv : 0.0;
        if (alpha && sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) alpha[s * r + j] = av[j];
        for (uint32_t it = 0; it < tmax; it += U) {
            if (tmax > U) {
                double wc[U];
                gather(it, uk, wc);
            }
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t c = c0 + (it + h) * L + li;
                if (c < c1) {
                    const size_t r0 = A.cell_obs_off ? A.cell_obs_off[c] : c;
                    const size_t r1 = A.cell_obs_off ? A.cell_obs_off[c + 1] : c + 1;
                    for (size_t i = r0; i < r1; i++)
#pragma unroll
                        for (int j = 0; j < RM; j++)
                            if (j < r) rows[i * r + j] -= uk[h][j] + av[j];
                }
            }
        }
    }
}

// rhs segment pass without the small scatter (the scatter runs in the level pull, or there are
// no small blocks): U trips per lane, cell sums kept in registers for z_c when a segment fits
// one trip, so every row is read once.
template <int RM>
__global__ __launch_bounds__(WG) void k_rhs_seg_reg(OpArgs<double> A, const double* __restrict__ rows,
                                                    const double* __restrict__ wr, int r, int L, double* __restrict__ b1,
                                                    double* __restrict__ zc, double* __restrict__ spart) {
    constexpr int U = gather_trips<double, RM>();
    __shared__ double red[WAVES][2 * 16];
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int spw = WAVE / L, gl = l / L, li = l % L;
    double wsq[RM], wsum[RM];
#pragma unroll
    for (int j = 0; j < RM; j++) { wsq[j] = 0.0; wsum[j] = 0.0; }
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * spw; base < A.g1; base += nwt * spw) {
        const size_t s = base + gl;
        const bool sv = s < A.g1;
        const uint32_t c0 = sv ? A.seg_off[s] : 0, c1 = sv ? A.seg_off[s + 1] : 0;
        const uint32_t tmax = wave_max_u32((c1 - c0 + L - 1) / L);
        double bv[RM], sk[U][RM];
#pragma unroll
        for (int j = 0; j < RM; j++) bv[j] = 0.0;
        for (uint32_t it = 0; it < tmax; it += U) {
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t c = c0 + (it + h) * L + li;
#pragma unroll
                for (int j = 0; j < RM; j++) sk[h][j] = 0.0;
                if (c < c1) cell_sums<RM>(A, rows, wr, c, r, sk[h], wsq, wsum, true);
#pragma unroll
                for (int j = 0; j < RM; j++) bv[j] += sk[h][j];
            }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) bv[j] = gsum(bv[j], L);
        if (sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) b1[s * r + j] = bv[j];
        if (!zc) continue;
        const double inv = sv ? A.inv_n1[s] : 0.0;
        for (uint32_t it = 0; it < tmax; it += U) {
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t c = c0 + (it + h) * L + li;
                if (c < c1) {
                    if (tmax > U) {
                        double d0[RM], d1[RM];
                        cell_sums<RM>(A, rows, wr, c, r, sk[h], d0, d1, false);
                    }
                    const double f = (A.cw ? A.cw[c] : 1.0) * inv;
#pragma unroll
                    for (int j = 0; j < RM; j++)
                        if (j < r) zc[(size_t)c * r + j] = sk[h][j] - f * bv[j];
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < RM; j++) {
        double a = wave_sum(wsq[j]), b = wave_sum(wsum[j]);
        if (l == 0 && j < r) { red[w][j] = a; red[w][16 + j] = b; }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        double a = 0.0, b = 0.0;
        for (int ww = 0; ww < WAVES; ww++) { a += red[ww][threadIdx.x]; b += red[ww][16 + threadIdx.x]; }
        spart[(size_t)blockIdx.x * 2 * r + threadIdx.x] = a;
        spart[(size_t)blockIdx.x * 2 * r + r + threadIdx.x] = b;
    }
}

inline unsigned seg_grid(const Absorber& a) {
    const size_t spw = WAVE / a.seg_lanes;
    return grid_for(a.g1, (unsigned)(spw * WAVES), 1u << 16);
}

}  // namespace

template <class T>
inline OpArgs<T> op_args_impl(const Absorber& a) {
    OpArgs<T> A{};
    A.cell_codes = a.cell_codes.p;
    A.seg_off = a.seg_off.p;
    A.cell_obs_off = a.trivial() ? nullptr : a.cell_obs_off.p;
    const WeightArrays<T>* W;
    if constexpr (std::is_same<T, double>::value) W = &a.w64; else W = &a.w32;
    A.cw = a.unit() ? nullptr : W->cw.p;
    A.inv_n1 = W->inv_n1.p;
    A.g1 = a.g1;
    A.M = a.M;
    A.m = a.m;
    A.ms = a.ms;
    A.nb = (int)a.nb;
    A.nsb = 0;
    for (size_t b = 0; b < a.nb; b++) {
        A.boff[b] = (uint32_t)a.blocks[b].off;
        if (a.blocks[b].small) A.nsb++;
    }
    return A;
}

#define FH_DISPATCH_R(r, ...)                                      \
    do {                                                           \
        if ((r) <= 1) { constexpr int RM = 1; __VA_ARGS__; }       \
        else if ((r) <= 2) { constexpr int RM = 2; __VA_ARGS__; }   \
        else if ((r) <= 4) { constexpr int RM = 4; __VA_ARGS__; }   \
        else if ((r) <= 8) { constexpr int RM = 8; __VA_ARGS__; }   \
        else { constexpr int RM = 16; __VA_ARGS__; }               \
    } while (0)

template <class F>
void with_nb(int nb, F&& f) {
    switch (nb) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        default: f(std::integral_constant<int, 4>{}); break;
    }
}
template <class F>
void with_nb0(int nb, F&& f) {
    if (nb == 0) f(std::integral_constant<int, 0>{});
    else with_nb(nb, f);
}
template <class F>
void with_no(int no, F&& f) {
    switch (no) {
        case 0: f(std::integral_constant<int, 0>{}); break;
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        default: f(std::integral_constant<int, NB_MAX - 1>{}); break;
    }
}


inline OtherOffs other_offs(const Absorber& a, int b) {
    OtherOffs oo{};
    int k = 0;
    for (int bb = 0; bb < (int)a.nb; bb++)
        if (bb != b) oo.off[k++] = (uint32_t)a.blocks[bb].off;
    return oo;
}

template <class T>
void bucket_reduce(const T* ypart, uint32_t nbk, size_t gb, int r, T* y, const uint32_t* done, hipStream_t st) {
    if (nbk <= 32)
        hipLaunchKernelGGL(k_bucket_reduce<T>, dim3(grid_for(gb * r, WG, 8192)), dim3(WG), 0, st, ypart, nbk, gb, r, y,
                           done);
    else
        hipLaunchKernelGGL(k_bucket_reduce_wave<T>, dim3(grid_for(gb * r, WAVES, 8192)), dim3(WG), 0, st, ypart, nbk,
                           gb, r, y, done);
    FH_CHECK_LAUNCH();
}

inline unsigned lvl_grid(int L, size_t items) {
    return grid_for(items, (unsigned)((WAVE / L) * WAVES), 1u << 20);
}

template <class T>
void spmv_impl(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, const uint32_t* done, hipStream_t st) {
    require(r >= 1 && r <= 16, "spmv: 1..16 columns");
    OpArgs<T> A = op_args_impl<T>(a);
    const int nb = (int)a.nb;
    if (nb == 0) return;
    const WeightArrays<T>* W;
    if constexpr (std::is_same<T, double>::value) W = &a.w64; else W = &a.w32;
    s.t.reserve(std::max<size_t>(a.g1 * r, 1));
    const int L = (int)a.seg_lanes;
    const unsigned grid = seg_grid(a);
    if (nb <= 4) {
        FH_DISPATCH_R(r, with_nb(nb, [&](auto nbc) {
            constexpr int NB = decltype(nbc)::value;
            hipLaunchKernelGGL((k_seg_t<T, RM, NB>), dim3(grid), dim3(WG), 0, st, A, p, r, L, s.t.p, done);
        }));
    } else {
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_seg_t_any<T, RM>), dim3(grid), dim3(WG), 0, st, A, p, r, L, s.t.p, done));
    }
    FH_CHECK_LAUNCH();
    if (a.small_dense) {
        const size_t ms = a.ms, mL = a.m - a.ms;
        unsigned nwg = 0, nch = 0;
        FH_DISPATCH_R(r, {
            constexpr int KC = small_chunk<T, RM>();
            nch = (unsigned)((ms + KC - 1) / KC);
            nwg = grid_for(a.g1 + mL, WG * 4, 512);
            s.part.reserve((size_t)nwg * ms * r);
            hipLaunchKernelGGL((k_small_dense<T, RM>), dim3(nwg, nch), dim3(WG), 0, st, W->dsm.p, s.t.p, W->csm.p,
                               p + ms * r, a.g1, mL, (int)ms, r, s.part.p, done);
        });
        FH_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_small_dense_final<T>, dim3(grid_for(ms * r, WAVES, 1024)), dim3(WG), 0, st, s.part.p, nwg,
                           (int)ms, r, W->nrs.p, p, y, done);
        FH_CHECK_LAUNCH();
    }
    for (int b = 0; b < nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (bi.small && a.small_dense) continue;
        const OtherOffs oo = other_offs(a, b);
        const int Ll = (int)a.blanes[b];
        const uint32_t nbk = a.nbk[b];
        const size_t items = (size_t)nbk * bi.g;
        const T* lw = a.unit() ? nullptr : W->lw[b].p;
        T* ypart = nullptr;
        if (nbk > 1) {
            s.ypart.reserve(items * r);
            ypart = s.ypart.p;
        }
        FH_DISPATCH_R(r, with_no(nb - 1, [&](auto noc) {
            constexpr int NO = decltype(noc)::value;
            hipLaunchKernelGGL((k_lvl_spmv<T, RM, NO>), dim3(lvl_grid(Ll, items)), dim3(WG), 0, st, a.boff[b].p,
                               a.lseg[b].p, a.lother[b].p, lw, items, (uint32_t)bi.g, (uint32_t)bi.off, oo, nb - 1, p,
                               s.t.p, r, Ll, y, ypart, done);
        }));
        FH_CHECK_LAUNCH();
        if (ypart) bucket_reduce<T>(ypart, nbk, (size_t)bi.g, r, y + bi.off * r, done, st);
    }
}
}  // namespace fh
