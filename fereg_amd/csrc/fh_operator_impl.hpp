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

// Index-stream loads (codes, level records, cell weights): plain cached loads. A lane's
// consecutive positions are L entries apart and share cache lines, which a non-temporal load
// does not keep: on the segment pass of a 1e7-cell design the non-temporal form took 66-95 us
// against 52-57 us for plain loads.
template <class T>
__device__ __forceinline__ T ld_idx(const T* p) { return *p; }

// row[j] = src[j] (j < r) for a valid lane, else 0. When r == RM the row is read with 16-byte
// (or 8-byte) vector loads: one gather instruction per row instead of one per column, which
// matters because random gathers are bounded by instructions x distinct lines, not by bytes.
// UA: rows with r < RM are read in element-aligned vector chunks. That pays for rows read
// in sequence (each line is used whole) and costs for random row gathers (a chunk that crosses a
// line boundary turns into two requests), so the gathers keep per-element loads.
template <class T, int RM, bool UA = false>
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
    // r < RM: the row start is only element-aligned, so the vector loads are 4-byte (FP32) or
    // 8-byte (FP64) aligned; the hardware serves dword-aligned vector loads from global memory,
    // so a row of 11 FP32 values is 4 loads (4 + 4 + 2 + 1) rather than 11
    if constexpr (UA && std::is_same<T, float>::value && RM % 4 == 0) {
#pragma unroll
        for (int q = 0; q < RM / 4; q++) {
            const int j = 4 * q;
            if (j + 4 <= r) {
                float4 x;
                __builtin_memcpy(&x, src + j, sizeof(float4));
                row[j] = x.x; row[j + 1] = x.y; row[j + 2] = x.z; row[j + 3] = x.w;
            } else if (j + 2 <= r) {
                float2 x;
                __builtin_memcpy(&x, src + j, sizeof(float2));
                row[j] = x.x; row[j + 1] = x.y;
                if (j + 3 <= r) row[j + 2] = src[j + 2];
            } else if (j + 1 <= r) {
                row[j] = src[j];
            }
        }
        return;
    } else if constexpr (UA && std::is_same<T, double>::value && RM % 2 == 0) {
#pragma unroll
        for (int q = 0; q < RM / 2; q++) {
            const int j = 2 * q;
            if (j + 2 <= r) {
                double2 x;
                __builtin_memcpy(&x, src + j, sizeof(double2));
                row[j] = x.x; row[j + 1] = x.y;
            } else if (j + 1 <= r) {
                row[j] = src[j];
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < RM; j++)
        if (j < r) row[j] = src[j];
}

// dst[j] = row[j] (j < r): the store counterpart of ld_row (16/8-byte vector stores when r == RM,
// element-aligned vector chunks otherwise).
template <class T, int RM>
__device__ __forceinline__ void st_row(T* __restrict__ dst, int r, const T (&row)[RM]) {
    if constexpr (std::is_same<T, double>::value && RM % 2 == 0) {
#pragma unroll
        for (int q = 0; q < RM / 2; q++) {
            const int j = 2 * q;
            if (j + 2 <= r) {
                const double2 x = make_double2(row[j], row[j + 1]);
                __builtin_memcpy(dst + j, &x, sizeof(double2));
            } else if (j + 1 <= r) {
                dst[j] = row[j];
            }
        }
    } else if constexpr (std::is_same<T, float>::value && RM % 4 == 0) {
#pragma unroll
        for (int q = 0; q < RM / 4; q++) {
            const int j = 4 * q;
            if (j + 4 <= r) {
                const float4 x = make_float4(row[j], row[j + 1], row[j + 2], row[j + 3]);
                __builtin_memcpy(dst + j, &x, sizeof(float4));
            } else {
#pragma unroll
                for (int i = 0; i < 3; i++)
                    if (j + i < r) dst[j + i] = row[j + i];
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < RM; j++)
            if (j < r) dst[j] = row[j];
    }
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
// With SKIP = 1 the first block (the single small block) is not gathered: its part of the
// segment sum, sum_c W_c p_S[k(c)] = (D' p_S)[s], comes in ds (dense, k_small_dense_t).
template <class T, int RM, int NB, int SKIP = 0>
__global__ __launch_bounds__(WG) void k_seg_t(OpArgs<T> A, const T* __restrict__ p, int r, int L, T* __restrict__ t,
                                              const T* __restrict__ ds, const uint32_t* __restrict__ done) {
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
                // with the small block skipped, the large blocks' own code array (no block-0 column)
                for (int b = SKIP; b < NB; b++)
                    code[h][b] = !v[h] ? 0u
                                 : SKIP ? ld_idx(&A.large_codes[(size_t)c * (NB - 1) + b - 1])
                                        : ld_idx(&A.cell_codes[(size_t)c * NB + b]);
                wc[h] = v[h] ? (A.cw ? ld_idx(&A.cw[c]) : T(1)) : T(0);
            }
#pragma unroll
            for (int h = 0; h < U; h++) {
                T u[RM];
#pragma unroll
                for (int j = 0; j < RM; j++) u[j] = T(0);
#pragma unroll
                for (int b = SKIP; b < NB; b++) {
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
        for (int j = 0; j < RM; j++) tv[j] = gsum(tv[j], L);
        if (SKIP) {
            T dv[RM];
            ld_row<T, RM>(ds + s * r, r, sv, dv);
#pragma unroll
            for (int j = 0; j < RM; j++) tv[j] += dv[j];
        }
#pragma unroll
        for (int j = 0; j < RM; j++) tv[j] *= inv;
        if (sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) t[s * r + j] = tv[j];
    }
}

// ---------------------------------------------------------------- S*p, dense small rows
// The dense small-block matrix E (ms rows, pitch ld, a multiple of 4) holds D in columns
// [0, g1) and C in columns [coff, coff + m - ms) (coff a multiple of 4, pads zero):
//   D[k][s] = sum of W_c over the cells of segment s with small level k,
//   C[k][l] = sum of W_c over the cells with small level k and large coordinate l.
// Threads own quads of 4 columns, so every E access is one 16-byte aligned load (two for FP64)
// and a quad never straddles the D/C boundary.
template <class T>
__device__ __forceinline__ void ld_quad(const T* __restrict__ src, T (&e)[4]) {
    if constexpr (sizeof(T) == 4) {
        const float4 x = *reinterpret_cast<const float4*>(src);
        e[0] = x.x; e[1] = x.y; e[2] = x.z; e[3] = x.w;
    } else {
        const double2 x = reinterpret_cast<const double2*>(src)[0], y = reinterpret_cast<const double2*>(src)[1];
        e[0] = x.x; e[1] = x.y; e[2] = y.x; e[3] = y.y;
    }
}

// The same from the u8 copy of E (cell counts without prior weights) when use8 (a device flag:
// the standalone D'p_S / C'p_S product); the fused partials pick their load path at compile time
// (ld_quad_sel) from the host copy of that flag.
template <class T>
__device__ __forceinline__ void ld_quad_e(const T* __restrict__ E, const uint8_t* __restrict__ E8, bool use8,
                                          size_t idx, T (&e)[4]) {
    if (use8) {
        const uint32_t v = *reinterpret_cast<const uint32_t*>(E8 + idx);
        e[0] = T(v & 0xFFu);
        e[1] = T((v >> 8) & 0xFFu);
        e[2] = T((v >> 16) & 0xFFu);
        e[3] = T(v >> 24);
    } else {
        ld_quad(E + idx, e);
    }
}

// A quad of E from the u8 counts (U8) or from E itself, chosen at compile time so that the fused
// level pass carries one load path only.
template <class T, bool U8>
__device__ __forceinline__ void ld_quad_sel(const T* __restrict__ E, const uint8_t* __restrict__ E8, size_t idx,
                                            T (&e)[4]) {
    if constexpr (U8) {
        const uint32_t v = *reinterpret_cast<const uint32_t*>(E8 + idx);
        e[0] = T(v & 0xFFu);
        e[1] = T((v >> 8) & 0xFFu);
        e[2] = T((v >> 16) & 0xFFu);
        e[3] = T(v >> 24);
    } else {
        ld_quad(E + idx, e);
    }
}

// Levels per chunk (blockIdx.y) of k_small_dense: KC x RM accumulators per thread.
template <class T, int RM>
constexpr int small_chunk() { return (sizeof(T) == 8 ? 32 : 64) / RM < 32 ? (sizeof(T) == 8 ? 32 : 64) / RM : 32; }

// ds[s] = sum_k D[k][s] p_S[k] (g1 rows) and dl[l] = sum_k C[k][l] p_S[k] (m - ms rows): the
// small block's contributions to the segment sums and to the large rows, one column quad per
// thread (E read once, coalesced; p_S staged in LDS).
template <class T, int RM>
__global__ __launch_bounds__(WG) void k_small_dense_t(const T* __restrict__ E, const uint8_t* __restrict__ E8,
                                                      const uint32_t* __restrict__ bad8, size_t ld, size_t coff,
                                                      size_t g1, size_t mL, const T* __restrict__ pS, int ms, int r,
                                                      T* __restrict__ ds, T* __restrict__ dl,
                                                      const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const bool use8 = E8 && *bad8 == 0;
    __shared__ T ps[32 * RM];
    for (int e = threadIdx.x; e < ms * RM; e += WG) {
        const int k = e / RM, j = e % RM;
        ps[e] = j < r ? pS[k * r + j] : T(0);
    }
    __syncthreads();
    const size_t nq = ld / 4;
    for (size_t q = (size_t)blockIdx.x * WG + threadIdx.x; q < nq; q += (size_t)gridDim.x * WG) {
        const size_t c4 = 4 * q;
        T acc[4][RM];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < RM; j++) acc[i][j] = T(0);
#pragma unroll 8
        for (int k = 0; k < ms; k++) {
            T e[4];
            ld_quad_e(E, E8, use8, (size_t)k * ld + c4, e);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < RM; j++) acc[i][j] += e[i] * ps[k * RM + j];
        }
        const bool seg = c4 < coff;
        T* o = seg ? ds : dl;
        const size_t base = seg ? c4 : c4 - coff, lim = seg ? g1 : mL;
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (base + i < lim)
#pragma unroll
                for (int j = 0; j < RM; j++)
                    if (j < r) o[(base + i) * r + j] = acc[i][j];
    }
}

// Partial sums over a column range of y_S[k] = sum_l C[k][l] p_L[l] - sum_s D[k][s] t_s for the
// small levels k in chunk ch (KC levels), one column quad per thread and step. Workgroup bx of nwg
// writes part[(k r + j) nwg + bx].
template <class T, int RM, int KC, bool U8>
__device__ __forceinline__ void small_dense_body(const T* __restrict__ E, const uint8_t* __restrict__ E8, size_t ld,
                                                 size_t coff, size_t g1, size_t mL,
                                                 const T* __restrict__ t, const T* __restrict__ pL, int ms, int r,
                                                 T* __restrict__ part, unsigned bx, unsigned nwg, unsigned ch) {
    __shared__ T red[WAVES][KC][RM];
    const int k0 = (int)ch * KC;
    T acc[KC][RM];
#pragma unroll
    for (int kk = 0; kk < KC; kk++)
#pragma unroll
        for (int j = 0; j < RM; j++) acc[kk][j] = T(0);
    const size_t nq = ld / 4;
    for (size_t q = (size_t)bx * WG + threadIdx.x; q < nq; q += (size_t)nwg * WG) {
        const size_t c4 = 4 * q;
        const bool seg = c4 < coff;
        const T* src = seg ? t : pL;
        const size_t base = seg ? c4 : c4 - coff, lim = seg ? g1 : mL;
        T v[4][RM];
#pragma unroll
        for (int i = 0; i < 4; i++) ld_row<T, RM>(src + (base + i) * r, r, base + i < lim, v[i]);
        if (seg)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < RM; j++) v[i][j] = -v[i][j];
#pragma unroll
        for (int kk = 0; kk < KC; kk++) {
            const int k = k0 + kk;
            if (k < ms) {
                T e[4];
                ld_quad_sel<T, U8>(E, E8, (size_t)k * ld + c4, e);
#pragma unroll
                for (int j = 0; j < RM; j++) acc[kk][j] += e[0] * v[0][j] + e[1] * v[1][j] + e[2] * v[2][j] + e[3] * v[3][j];
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
        part[((size_t)(k0 + kk) * r + j) * nwg + bx] = sm;
    }
}

// The same partials for wide blocks (RM >= 8): four threads share a column quad and own RM/4
// columns each, so one chunk covers up to 12..32 small levels and t, p_L are read once (KC x RM
// accumulators no longer fit one thread); the quad's E loads coincide across the four threads.
// Lanes with the same column slice are reduced by shuffles, then the waves in order.
template <class T, int RM, int KC, bool U8>
__device__ __forceinline__ void small_dense_cols(const T* __restrict__ E, const uint8_t* __restrict__ E8, size_t ld,
                                                 size_t coff, size_t g1, size_t mL,
                                                 const T* __restrict__ t, const T* __restrict__ pL, int ms, int r,
                                                 T* __restrict__ part, unsigned bx, unsigned nwg, unsigned ch) {
    constexpr int TPQ = 4, CW = RM / TPQ, QPB = WG / TPQ;
    __shared__ T red[WAVES][KC][RM];
    const int sl = threadIdx.x % TPQ, qi = threadIdx.x / TPQ, c0 = sl * CW, k0 = (int)ch * KC;
    T acc[KC][CW];
#pragma unroll
    for (int kk = 0; kk < KC; kk++)
#pragma unroll
        for (int c = 0; c < CW; c++) acc[kk][c] = T(0);
    const size_t nq = ld / 4;
    for (size_t q = (size_t)bx * QPB + qi; q < nq; q += (size_t)nwg * QPB) {
        const size_t c4 = 4 * q;
        const bool seg = c4 < coff;
        const T* src = seg ? t : pL;
        const size_t base = seg ? c4 : c4 - coff, lim = seg ? g1 : mL;
        T v[4][CW];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int c = 0; c < CW; c++) {
                const T x = (c0 + c < r && base + i < lim) ? src[(base + i) * r + c0 + c] : T(0);
                v[i][c] = seg ? -x : x;
            }
#pragma unroll
        for (int kk = 0; kk < KC; kk++) {
            const int k = k0 + kk;
            if (k < ms) {
                T e[4];
                ld_quad_sel<T, U8>(E, E8, (size_t)k * ld + c4, e);
#pragma unroll
                for (int c = 0; c < CW; c++) acc[kk][c] += e[0] * v[0][c] + e[1] * v[1][c] + e[2] * v[2][c] + e[3] * v[3][c];
            }
        }
    }
    const int w = threadIdx.x / WAVE, l = lane_id();
#pragma unroll
    for (int kk = 0; kk < KC; kk++)
#pragma unroll
        for (int c = 0; c < CW; c++) {
            T x = acc[kk][c];
#pragma unroll
            for (int o = TPQ; o < WAVE; o <<= 1) x += __shfl_xor(x, o, WAVE);
            if (l < TPQ) red[w][kk][l * CW + c] = x;
        }
    __syncthreads();
    for (int e = threadIdx.x; e < KC * r; e += WG) {
        const int kk = e / r, jj = e % r;
        if (k0 + kk >= ms) continue;
        T sm = T(0);
#pragma unroll
        for (int ww = 0; ww < WAVES; ww++) sm += red[ww][kk][jj];
        part[((size_t)(k0 + kk) * r + jj) * nwg + bx] = sm;
    }
}

// Dense small-row partials, picking the layout by the column class.
template <class T, int RM, int KC, bool U8>
__device__ __forceinline__ void small_dense_any(const T* __restrict__ E, const uint8_t* __restrict__ E8, size_t ld,
                                                size_t coff, size_t g1, size_t mL, const T* __restrict__ t,
                                                const T* __restrict__ pL, int ms, int r, T* __restrict__ part,
                                                unsigned bx, unsigned nwg, unsigned ch) {
    if constexpr (RM >= 8)
        small_dense_cols<T, RM, KC, U8>(E, E8, ld, coff, g1, mL, t, pL, ms, r, part, bx, nwg, ch);
    else
        small_dense_body<T, RM, KC, U8>(E, E8, ld, coff, g1, mL, t, pL, ms, r, part, bx, nwg, ch);
}

// Levels per chunk of the dense small-row partials: KC x RM accumulators per thread (rows layout)
// or KC x RM/4 (column layout of the wide blocks), at most 32.
template <class T, int RM>
constexpr int small_chunk_any() {
    if constexpr (RM >= 8) {
        constexpr int k = (sizeof(T) == 8 ? 48 : 64) / (RM / 4);
        return k < 32 ? k : 32;
    } else {
        return small_chunk<T, RM>();
    }
}

template <class T, int RM, int KC, bool U8>
__global__ __launch_bounds__(WG) void k_small_dense(const T* __restrict__ E, const uint8_t* __restrict__ E8, size_t ld,
                                                    size_t coff, size_t g1, size_t mL, const T* __restrict__ t,
                                                    const T* __restrict__ pL, int ms, int r, T* __restrict__ part,
                                                    const uint32_t* __restrict__ done) {
    if (done && *done) return;
    small_dense_any<T, RM, KC, U8>(E, E8, ld, coff, g1, mL, t, pL, ms, r, part, blockIdx.x, gridDim.x, blockIdx.y);
}

// Last step of S*p: the small rows y_S[k][j] = nr_k p_S[k][j] + the sum of the partials of entry
// e = k r + j (when part is set; one workgroup per entry, fixed order), then, when pq is set, the
// per-workgroup partials pq[wg * 16 + j] of p'y over all m rows (the PCG's p'q, fixed order).
template <class T>
__global__ __launch_bounds__(WG) void k_spmv_finish(const T* __restrict__ part, unsigned nparts, size_t ms, int r,
                                                    const T* __restrict__ nrs, const T* __restrict__ p,
                                                    T* __restrict__ y, size_t m, double* __restrict__ pq,
                                                    const uint32_t* __restrict__ done) {
    if (done && *done) return;
    constexpr int RQ = 16;
    __shared__ T red[WAVES];
    __shared__ double dred[WAVES][RQ];
    double acc[RQ];
#pragma unroll
    for (int j = 0; j < RQ; j++) acc[j] = 0.0;
    if (part) {
        for (size_t e = blockIdx.x; e < ms * r; e += gridDim.x) {
            T sm = T(0);
            for (unsigned q = threadIdx.x; q < nparts; q += WG) sm += part[e * nparts + q];
            sm = wave_sum(sm);
            if (lane_id() == 0) red[threadIdx.x / WAVE] = sm;
            __syncthreads();
            if (threadIdx.x == 0) {
                T tot = T(0);
#pragma unroll
                for (int w = 0; w < WAVES; w++) tot += red[w];
                const T ye = nrs[e / r] * p[e] + tot;
                y[e] = ye;
                const int j = (int)(e % r);
#pragma unroll
                for (int jj = 0; jj < RQ; jj++)
                    if (jj == j) acc[jj] += (double)p[e] * (double)ye;
            }
            __syncthreads();
        }
    }
    if (!pq) return;
    const size_t e0 = part ? ms * r : 0;
    for (size_t i = e0 + (size_t)blockIdx.x * WG + threadIdx.x; i < m * r; i += (size_t)gridDim.x * WG) {
        const int j = (int)(i % r);
        const double v = (double)p[i] * (double)y[i];
#pragma unroll
        for (int jj = 0; jj < RQ; jj++)
            if (jj == j) acc[jj] += v;
    }
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int j = 0; j < RQ; j++)
        if (j < r) {
            const double sm = wave_sum(acc[j]);
            if (lane_id() == 0) dred[w][j] = sm;
        }
    __syncthreads();
    if (threadIdx.x < (unsigned)r) {
        double sm = 0.0;
#pragma unroll
        for (int ww = 0; ww < WAVES; ww++) sm += dred[ww][threadIdx.x];
        pq[(size_t)blockIdx.x * RQ + threadIdx.x] = sm;
    }
}

struct OtherOffs {
    uint32_t off[NB_MAX];
};

// ---------------------------------------------------------------- S*p, level pass of one block
// y_l = sum_{c in l} W_c (u_c - t_{s(c)}) over the level-major records of item
// (bucket, level); with several buckets the item sums go to ypart and k_bucket_reduce adds
// them in bucket order.
// With oskip > 0 the first oskip other blocks (the single small block) are not gathered: their
// part of the row, sum_c W_c p_S[k(c)] = (C' p_S)[l], comes in dl (this block's rows, dense,
// k_small_dense_t).
template <class T>
struct LvlArgs {
    const uint32_t* loff;
    const uint32_t* lseg;
    const uint32_t* lother;
    const T* lw;
    size_t g;
    uint32_t gb, off_b;
    OtherOffs oo;
    int nother, oskip;
    const T* dl;
    const T* p;
    const T* t;
    int r, L;
    T* y;
    T* ypart;
    uint32_t nbk;    // buckets (items = nbk x gb, bucket-major)
    uint32_t X;      // 1, or XCD_COUNT: bucket q runs only on the workgroups of XCD q mod X
    uint32_t merge;  // 1: one item per level over all its buckets' ranges, y written directly (X = 1)
};

template <class T, int RM, int NO>
__device__ __forceinline__ void lvl_spmv_body(const LvlArgs<T>& A, unsigned bid, unsigned nblk) {
    const uint32_t* __restrict__ loff = A.loff;
    const uint32_t* __restrict__ lseg = A.lseg;
    const uint32_t* __restrict__ lother = A.lother;
    const T* __restrict__ lw = A.lw;
    const T* __restrict__ dl = A.dl;
    const T* __restrict__ p = A.p;
    const T* __restrict__ t = A.t;
    const size_t g = A.g;
    const uint32_t gb = A.gb, off_b = A.off_b;
    const int nother = A.nother, oskip = A.oskip, r = A.r, L = A.L;
    T* __restrict__ y = A.y;
    T* __restrict__ ypart = A.ypart;
    const OtherOffs& oo = A.oo;
    constexpr int U = gather_trips<T, RM>();
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    // Workgroups are dispatched round-robin over the XCDs (workgroup b runs on XCD b mod X), so
    // with X = XCD_COUNT every bucket's items, and with them the t rows of its FE1 range, stay in
    // one XCD's L2. The XCD's items are its buckets' items in order.
    // With merge, an item is a level and its positions are the ranges of that level in every
    // bucket, summed in bucket order.
    const uint32_t X = A.X, xq = bid % X, merge = A.merge;
    const size_t nl = nblk / X, lb = bid / X;
    const size_t nbx = A.nbk > xq ? (A.nbk - xq + X - 1) / X : 0, nxi = merge ? gb : nbx * gb;
    (void)g;
    const size_t nwt = nl * WAVES;
    for (size_t base = (lb * WAVES + w) * gpw; base < nxi; base += nwt * gpw) {
        const size_t kx = base + gl;
        const bool valid = kx < nxi;
        // item = bucket * gb + level
        const size_t it_ = !valid ? 0 : merge ? kx : (xq + (size_t)X * (kx / gb)) * gb + kx % gb;
        const size_t lv = valid ? it_ % gb : 0;
        T pb[RM], acc[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) {
            pb[j] = (valid && j < r) ? p[(size_t)(off_b + lv) * r + j] : T(0);
            acc[j] = T(0);
        }
        const uint32_t nq = merge ? A.nbk : 1;
        for (uint32_t bq = 0; bq < nq; bq++) {
        const size_t itq = merge ? (size_t)bq * gb + lv : it_;
        const uint32_t j0 = valid ? loff[itq] : 0, j1 = valid ? loff[itq + 1] : 0;
        const uint32_t tmax = wave_max_u32((j1 - j0 + L - 1) / L);
        for (uint32_t it = 0; it < tmax; it += U) {
            // indices of U positions per lane first, then the gathers
            uint32_t sg[U], oc[U][NO > 0 ? NO : 1];
            bool pv[U];
            T wc[U];
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t pos = j0 + (it + h) * L + li;
                pv[h] = pos < j1;
                sg[h] = pv[h] ? ld_idx(&lseg[pos]) : 0u;
#pragma unroll
                for (int o = 0; o < NO; o++)
                    oc[h][o] = (pv[h] && o < nother && o >= oskip) ? ld_idx(&lother[(size_t)pos * nother + o]) : 0u;
                wc[h] = pv[h] ? (lw ? ld_idx(&lw[pos]) : T(1)) : T(0);
            }
#pragma unroll
            for (int h = 0; h < U; h++) {
                T u[RM], z[RM];
#pragma unroll
                for (int j = 0; j < RM; j++) u[j] = pb[j];
#pragma unroll
                for (int o = 0; o < NO; o++) {
                    T row[RM];
                    ld_row<T, RM>(p + (size_t)(oo.off[o] + oc[h][o]) * r, r, pv[h] && o < nother && o >= oskip, row);
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
        }
#pragma unroll
        for (int j = 0; j < RM; j++) acc[j] = gsum(acc[j], L);
        if (dl && valid && it_ < gb) {  // once per level: the first bucket's item
            T dv[RM];
            ld_row<T, RM>(dl + (size_t)lv * r, r, true, dv);
#pragma unroll
            for (int j = 0; j < RM; j++) acc[j] += dv[j];
        }
        if (valid && li == 0) {
            T* dst = (ypart && !merge) ? ypart + it_ * r : y + (size_t)(off_b + lv) * r;
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) dst[j] = acc[j];
        }
    }
}

template <class T, int RM, int NO>
__global__ __launch_bounds__(WG) void k_lvl_spmv(LvlArgs<T> A, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    lvl_spmv_body<T, RM, NO>(A, blockIdx.x, gridDim.x);
}

// One launch for a level pass and the dense small-row partials: both read only p and t, so the
// bandwidth-bound dense product runs beside the latency-bound gathers. Workgroups [0, nlvl) run
// the level pass, the rest the dense chunks (nwg workgroups per chunk of KC levels).
template <class T, int RM, int NO, int KC, bool U8>
__global__ __launch_bounds__(WG) void k_lvl_small(LvlArgs<T> A, unsigned nlvl, const T* __restrict__ E,
                                                  const uint8_t* __restrict__ E8, size_t ld, size_t coff, size_t g1,
                                                  size_t mL, const T* __restrict__ pL, int ms, T* __restrict__ part,
                                                  unsigned nwg, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    if (blockIdx.x < nlvl) {
        lvl_spmv_body<T, RM, NO>(A, blockIdx.x, nlvl);
    } else {
        const unsigned b = blockIdx.x - nlvl;
        small_dense_any<T, RM, KC, U8>(E, E8, ld, coff, g1, mL, A.t, pL, ms, A.r, part, b % nwg, nwg, b / nwg);
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
        double v[RM];
        ld_row<double, RM, true>(rows + i * r, r, true, v);
#pragma unroll
        for (int j = 0; j < RM; j++)
            if (j < r) {
                const double wv = w * v[j];
                sc[j] += wv;
                if (stats) {
                    wsq[j] += wv * v[j];
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
                c[h] = pv[h] ? ld_idx(&lperm[pos]) : 0u;
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
__global__ __launch_bounds__(WG) void k_resid_seg_any(OpArgs<double> A, double* __restrict__ rows, int r, int rp,
                                                  int L, const double* __restrict__ x, const double* __restrict__ b1,
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
                gather_u<double, RM>(A, x, c, rp, u);
                const double wc = A.cw ? A.cw[c] : 1.0;
#pragma unroll
                for (int j = 0; j < RM; j++) av[j] += wc * u[j];
            }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = gsum(av[j], L);
        const double inv = sv ? A.inv_n1[s] : 0.0;
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = sv && j < rp ? (b1[s * rp + j] - av[j]) * inv : 0.0;
        if (alpha && sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < rp) alpha[s * rp + j] = av[j];
        for (uint32_t it = 0; it < tmax; it++) {
            const uint32_t c = c0 + it * L + li;
            if (c < c1) {
                double u[RM];
                if (A.nb > 0) gather_u<double, RM>(A, x, c, rp, u);
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
// when a segment fits one trip (the common case), so x is gathered once per cell. When every
// cell is one row, the segment fits one trip and a lane has four or more trips (rows of up to
// four columns), the rows, b1 and 1/n are loaded at the start
// with the cell codes, so their latency overlaps the gathers instead of following the alpha
// store (which the compiler cannot move them past).

template <int RM, int NB>
__global__ __launch_bounds__(WG) void k_resid_seg(OpArgs<double> A, double* __restrict__ rows, int r, int rp, int L,
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
        // wave-uniform; four or more trips per lane only (8 columns, U = 2: 688 -> 714 us)
        const bool pre = U >= 4 && !A.cell_obs_off && tmax <= U;
        double av[RM], uk[U][RM], rv[U][RM], bs[RM];
        if (pre) {
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t c = c0 + h * L + li;
                ld_row<double, RM, true>(rows + (size_t)c * r, r, c < c1, rv[h]);
            }
        }
        if (U >= 4)
#pragma unroll
            for (int j = 0; j < RM; j++) bs[j] = sv && j < rp ? b1[s * rp + j] : 0.0;
        const double inv_early = U >= 4 && sv ? A.inv_n1[s] : 0.0;
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = 0.0;
        auto gather = [&](uint32_t it, double (&u)[U][RM], double (&wc)[U]) {
            uint32_t code[U][NB > 0 ? NB : 1];
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t c = c0 + (it + h) * L + li;
                const bool v = c < c1;
#pragma unroll
                for (int b = 0; b < NB; b++) code[h][b] = v ? ld_idx(&A.cell_codes[(size_t)c * NB + b]) : 0u;
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
                    ld_row<double, RM>(x + (size_t)(A.boff[b] + code[h][b]) * rp, rp, v, row);
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
        const double inv = U >= 4 ? inv_early : (sv ? A.inv_n1[s] : 0.0);
        if (U < 4)
#pragma unroll
            for (int j = 0; j < RM; j++) bs[j] = sv && j < rp ? b1[s * rp + j] : 0.0;
#pragma unroll
        for (int j = 0; j < RM; j++) av[j] = sv && j < rp ? (bs[j] - av[j]) * inv : 0.0;
        if (alpha && sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < rp) alpha[s * rp + j] = av[j];
        if (pre) {
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t c = c0 + h * L + li;
                if (c < c1) {
#pragma unroll
                    for (int j = 0; j < RM; j++) rv[h][j] -= uk[h][j] + av[j];
                    st_row<double, RM>(rows + (size_t)c * r, r, rv[h]);
                }
            }
            continue;
        }
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
                    for (size_t i = r0; i < r1; i++) {
                        double v[RM];
                        ld_row<double, RM, true>(rows + i * r, r, true, v);
#pragma unroll
                        for (int j = 0; j < RM; j++) v[j] -= uk[h][j] + av[j];
                        st_row<double, RM>(rows + i * r, r, v);
                    }
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
                                                    const double* __restrict__ wr, int r, int rp, int L,
                                                    double* __restrict__ b1, double* __restrict__ zc,
                                                    double* __restrict__ spart) {
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
        double bv[RM], sk[U][RM], cwk[U];
#pragma unroll
        for (int j = 0; j < RM; j++) bv[j] = 0.0;
        // 1/n and (one row per cell) the cell weights are loaded with the rows, ahead of the b1
        // store that the compiler cannot move them past. Only with four or more trips per lane:
        // with fewer the extra code path costs occupancy (8 columns: 503 -> 530 us, 12 columns:
        // 709 -> 1130 us).
        const double inv_early = U >= 4 && zc && sv ? A.inv_n1[s] : 0.0;
        for (uint32_t it = 0; it < tmax; it += U) {
            if (U >= 4 && !A.cell_obs_off) {
                // one row per cell: the U rows (and weights) are loaded together
                double wk[U];
#pragma unroll
                for (int h = 0; h < U; h++) {
                    const uint32_t c = c0 + (it + h) * L + li;
                    ld_row<double, RM, true>(rows + (size_t)c * r, r, c < c1, sk[h]);
                    wk[h] = c < c1 ? (wr ? wr[c] : 1.0) : 0.0;
                    cwk[h] = zc && c < c1 && A.cw ? A.cw[c] : 1.0;
                }
#pragma unroll
                for (int h = 0; h < U; h++)
#pragma unroll
                    for (int j = 0; j < RM; j++)
                        if (j < r) {
                            const double wv = wk[h] * sk[h][j];
                            wsq[j] += wv * sk[h][j];
                            wsum[j] += wv;
                            sk[h][j] = wv;
                            bv[j] += wv;
                        }
                continue;
            }
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
                if (j < rp) b1[s * rp + j] = bv[j];
        if (!zc) continue;
        const double inv = U >= 4 ? inv_early : (sv ? A.inv_n1[s] : 0.0);
        const bool one = U >= 4 && !A.cell_obs_off && tmax <= U;  // cwk holds this segment's cell weights
        for (uint32_t it = 0; it < tmax; it += U) {
#pragma unroll
            for (int h = 0; h < U; h++) {
                const uint32_t c = c0 + (it + h) * L + li;
                if (c < c1) {
                    if (tmax > U) {
                        double d0[RM], d1[RM];
                        cell_sums<RM>(A, rows, wr, c, r, sk[h], d0, d1, false);
                    }
                    const double f = (one ? cwk[h] : (A.cw ? A.cw[c] : 1.0)) * inv;
                    double zv[RM];
#pragma unroll
                    for (int j = 0; j < RM; j++) zv[j] = sk[h][j] - f * bv[j];
                    st_row<double, RM>(zc + (size_t)c * rp, rp, zv);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < RM; j++) {
        double a = wave_sum(wsq[j]), b = wave_sum(wsum[j]);
        if (l == 0 && j < rp) { red[w][j] = a; red[w][16 + j] = b; }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)rp) {
        double a = 0.0, b = 0.0;
        for (int ww = 0; ww < WAVES; ww++) { a += red[ww][threadIdx.x]; b += red[ww][16 + threadIdx.x]; }
        spart[(size_t)blockIdx.x * 2 * rp + threadIdx.x] = a;
        spart[(size_t)blockIdx.x * 2 * rp + rp + threadIdx.x] = b;
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
    A.large_codes = a.small_dense && a.nb > 1 ? a.large_codes.p : nullptr;
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
        else if ((r) <= 12) { constexpr int RM = 12; __VA_ARGS__; } \
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
// The level-pass grid and XCD spread for nbk buckets: with at least XCD_COUNT buckets each XCD
// takes the buckets q = x mod XCD_COUNT (the grid a multiple of XCD_COUNT).
inline unsigned lvl_grid_x(int L, size_t items, uint32_t nbk, uint32_t* X) {
    const unsigned g = lvl_grid(L, items);
    if (nbk < (uint32_t)XCD_COUNT) {
        *X = 1;
        return g;
    }
    *X = XCD_COUNT;
    return (g + XCD_COUNT - 1) / XCD_COUNT * XCD_COUNT;
}

// The route and lane widths of one S*p: plan.dense selects the dense products for the single
// small block also inside the segment and level passes (its codes are then never read there).
// pq (optional, SPMV_DOTG x 16 doubles): per-workgroup partials of p'y, for the PCG.
template <class T>
void spmv_impl(const Absorber& a, const SpmvPlan& plan, const T* p, T* y, int r, Scratch<T>& s,
               const uint32_t* done, hipStream_t st, double* pq = nullptr) {
    require(r >= 1 && r <= 16, "spmv: 1..16 columns");
    OpArgs<T> A = op_args_impl<T>(a);
    const int nb = (int)a.nb;
    if (nb == 0) return;
    const WeightArrays<T>* W;
    if constexpr (std::is_same<T, double>::value) W = &a.w64; else W = &a.w32;
    s.t.reserve(std::max<size_t>(a.g1 * r, 1));
    const int L = (int)(plan.seg_lanes ? plan.seg_lanes : spmv_seg_lanes(a, r, (int)sizeof(T)));
    const unsigned grid = grid_for(a.g1, (unsigned)((WAVE / L) * WAVES), 1u << 16);
    const size_t ms = a.ms, mL = a.m - a.ms;
    const T* E = a.small_dense ? W->esm.p : nullptr;
    const uint8_t* E8 = a.small_dense && !a.weighted ? a.esm8.p : nullptr;  // u8 counts when they fit
    const uint32_t* bad8 = a.small_dense ? a.esm8_bad.p : nullptr;
    const bool u8 = E8 && a.esm8_ok;  // the fused partials read the u8 counts
    const bool dense = a.small_dense && plan.dense && nb <= 4 && r <= 4;
    if (dense) {
        s.ds.reserve(a.g1 * r);
        s.dl.reserve(std::max<size_t>(mL * r, 1));
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_small_dense_t<T, RM>), dim3(grid_for(a.sm_ld / 4, WG, 4096)), dim3(WG),
                                            0, st, E, E8, bad8, a.sm_ld, a.sm_coff, a.g1, mL, p, (int)ms, r, s.ds.p,
                                            s.dl.p, done));
        FH_CHECK_LAUNCH();
    }
    if (nb <= 4) {
        FH_DISPATCH_R(r, with_nb(nb, [&](auto nbc) {
            constexpr int NB = decltype(nbc)::value;
            if (dense)
                hipLaunchKernelGGL((k_seg_t<T, RM, NB, 1>), dim3(grid), dim3(WG), 0, st, A, p, r, L, s.t.p, s.ds.p,
                                   done);
            else
                hipLaunchKernelGGL((k_seg_t<T, RM, NB, 0>), dim3(grid), dim3(WG), 0, st, A, p, r, L, s.t.p,
                                   (const T*)nullptr, done);
        }));
    } else {
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_seg_t_any<T, RM>), dim3(grid), dim3(WG), 0, st, A, p, r, L, s.t.p, done));
    }
    FH_CHECK_LAUNCH();
    // dense small-row partials: one launch with the first large block's level pass (or alone)
    const unsigned nwg = a.small_dense ? grid_for(a.sm_ld / 4, WG, 2048) : 0;  // ~one column quad per thread
    if (a.small_dense) s.part.reserve((size_t)nwg * ms * r);
    int first_large = -1;
    for (int b = 0; b < nb; b++)
        if (!(a.blocks[b].small && a.small_dense)) {
            first_large = b;
            break;
        }
    auto with_kc = [&](auto rmc, auto&& f) {
        constexpr int KM = small_chunk_any<T, decltype(rmc)::value>();
        if (ms <= 4 || KM <= 4) f(std::integral_constant<int, (KM < 4 ? KM : 4)>{});
        else if (ms <= 8 || KM <= 8) f(std::integral_constant<int, (KM < 8 ? KM : 8)>{});
        else if (ms <= 16 || KM <= 16) f(std::integral_constant<int, (KM < 16 ? KM : 16)>{});
        else f(std::integral_constant<int, KM>{});
    };
    if (a.small_dense && first_large < 0) {
        FH_DISPATCH_R(r, with_kc(std::integral_constant<int, RM>{}, [&](auto kc) {
            constexpr int KC = decltype(kc)::value;
            const unsigned nch = (unsigned)((ms + KC - 1) / KC);
            if (u8)
                hipLaunchKernelGGL((k_small_dense<T, RM, KC, true>), dim3(nwg, nch), dim3(WG), 0, st, E, E8, a.sm_ld,
                                   a.sm_coff, a.g1, mL, s.t.p, p + ms * r, (int)ms, r, s.part.p, done);
            else
                hipLaunchKernelGGL((k_small_dense<T, RM, KC, false>), dim3(nwg, nch), dim3(WG), 0, st, E, E8, a.sm_ld,
                                   a.sm_coff, a.g1, mL, s.t.p, p + ms * r, (int)ms, r, s.part.p, done);
        }));
        FH_CHECK_LAUNCH();
    }
    int nsb = 0;
    for (auto& bi : a.blocks) nsb += bi.small ? 1 : 0;
    for (int b = 0; b < nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (bi.small && a.small_dense) continue;
        require(a.boff[b].n > 0 && a.lseg[b].n > 0, "S*p: a block without level records");
        const uint32_t nbk = a.nbk[b];
        const LvlMode md = lvl_mode(a, plan, b, r, (int)sizeof(T));
        const bool merge = md.merge;
        const int Ll = md.lanes;
        const size_t items = merge ? (size_t)bi.g : (size_t)nbk * bi.g;
        T* ypart = nullptr;
        if (nbk > 1 && !merge) {
            s.ypart.reserve(items * r);
            ypart = s.ypart.p;
        }
        LvlArgs<T> LA;
        LA.loff = a.boff[b].p;
        LA.lseg = a.lseg[b].p;
        LA.lother = a.lother[b].p;
        LA.lw = a.unit() ? nullptr : W->lw[b].p;
        LA.g = items;
        LA.gb = (uint32_t)bi.g;
        LA.off_b = (uint32_t)bi.off;
        LA.oo = other_offs(a, b);
        LA.nother = nb - 1;
        LA.oskip = dense ? nsb : 0;  // small blocks precede the large ones among the others
        LA.dl = dense ? s.dl.p + (bi.off - ms) * r : nullptr;
        LA.p = p;
        LA.t = s.t.p;
        LA.r = r;
        LA.L = Ll;
        LA.y = y;
        LA.ypart = ypart;
        LA.nbk = nbk;
        LA.merge = merge ? 1u : 0u;
        const unsigned nlvl = lvl_grid_x(Ll, items, merge ? 1u : nbk, &LA.X);
        FH_DISPATCH_R(r, with_no(nb - 1, [&](auto noc) {
            constexpr int NO = decltype(noc)::value;
            if (a.small_dense && b == first_large) {
                with_kc(std::integral_constant<int, RM>{}, [&](auto kc) {
                    constexpr int KC = decltype(kc)::value;
                    const unsigned nch = (unsigned)((ms + KC - 1) / KC);
                    if (u8)
                        hipLaunchKernelGGL((k_lvl_small<T, RM, NO, KC, true>), dim3(nlvl + nwg * nch), dim3(WG), 0, st,
                                           LA, nlvl, E, E8, a.sm_ld, a.sm_coff, a.g1, mL, p + ms * r, (int)ms,
                                           s.part.p, nwg, done);
                    else
                        hipLaunchKernelGGL((k_lvl_small<T, RM, NO, KC, false>), dim3(nlvl + nwg * nch), dim3(WG), 0,
                                           st, LA, nlvl, E, E8, a.sm_ld, a.sm_coff, a.g1, mL, p + ms * r, (int)ms,
                                           s.part.p, nwg, done);
                });
            } else {
                hipLaunchKernelGGL((k_lvl_spmv<T, RM, NO>), dim3(nlvl), dim3(WG), 0, st, LA, done);
            }
        }));
        FH_CHECK_LAUNCH();
        if (ypart) bucket_reduce<T>(ypart, nbk, (size_t)bi.g, r, y + bi.off * r, done, st);
    }
    if (a.small_dense || pq) {
        const unsigned fg = pq ? (unsigned)SPMV_DOTG : (unsigned)(ms * r);
        hipLaunchKernelGGL(k_spmv_finish<T>, dim3(fg), dim3(WG), 0, st, a.small_dense ? s.part.p : (const T*)nullptr,
                           nwg, ms, r, a.small_dense ? W->nrs.p : (const T*)nullptr, p, y, a.m, pq, done);
        FH_CHECK_LAUNCH();
    }
}

// Element counts of the scratch arrays spmv_impl reserves for this plan and column count (the
// same rules as its reserve calls).
template <class T>
SpmvNeed spmv_scratch_need(const Absorber& a, const SpmvPlan& plan, int r) {
    SpmvNeed n;
    const int nb = (int)a.nb;
    if (nb == 0) return n;
    const size_t ms = a.ms, mL = a.m - a.ms;
    n.t = std::max<size_t>(a.g1 * r, 1);
    if (a.small_dense && plan.dense && nb <= 4 && r <= 4) {
        n.ds = a.g1 * r;
        n.dl = std::max<size_t>(mL * r, 1);
    }
    if (a.small_dense) n.part = (size_t)grid_for(a.sm_ld / 4, WG, 2048) * ms * r;
    for (int b = 0; b < nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (bi.small && a.small_dense) continue;
        const LvlMode md = lvl_mode(a, plan, b, r, (int)sizeof(T));
        if (a.nbk[b] > 1 && !md.merge) n.ypart = std::max(n.ypart, (size_t)a.nbk[b] * bi.g * r);
    }
    return n;
}

// Times the candidate plans of S*p for this absorber, precision and column class on the caller's
// (p, y) (y is overwritten, p only read) and keeps the fastest: coordinate descent over the
// route, then the segment lanes, then each large block's level lanes, halving and doubling the
// default widths. Each candidate runs once untimed and three times timed (minimum kept). The
// choice is cached per structure signature, so later absorbers of the same shape skip the timing.
template <class T>
void autotune_spmv(const Absorber& a, SpmvPlan& plan, const T* p, T* y, int r, Scratch<T>& s, hipStream_t st) {
    const int rc = r <= 1 ? 1 : r <= 2 ? 2 : r <= 4 ? 4 : r <= 8 ? 8 : r <= 12 ? 12 : 16;
    const std::string key = plan_key(a, (int)sizeof(T), rc);
    if (tune_cache_get(key, plan)) {
        plan.tuned = true;
        return;
    }
    SpmvPlan best;
    best.dense = 1;
    best.seg_lanes = spmv_seg_lanes(a, r, (int)sizeof(T));
    best.blanes.assign(a.nb, 0);
    best.xcd.assign(a.nb, -1);
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    auto time_plan = [&](const SpmvPlan& c) {
        spmv_impl<T>(a, c, p, y, r, s, nullptr, st);
        float bestms = 1e30f;
        for (int k = 0; k < 3; k++) {
            FH_HIP(hipEventRecord(e0, st));
            spmv_impl<T>(a, c, p, y, r, s, nullptr, st);
            FH_HIP(hipEventRecord(e1, st));
            FH_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FH_HIP(hipEventElapsedTime(&ms, e0, e1));
            bestms = std::min(bestms, ms);
        }
        return bestms;
    };
    float tb = time_plan(best);
    auto consider = [&](const SpmvPlan& c) {
        const float tc = time_plan(c);
        if (tc < tb * 0.98f) {
            tb = tc;
            best = c;
        }
    };
    if (a.small_dense && a.nb <= 4 && r <= 4) {
        SpmvPlan c = best;
        c.dense = 0;
        consider(c);
    }
    for (uint32_t f : {2u, 0u}) {  // double, then halve the segment lanes
        SpmvPlan c = best;
        c.seg_lanes = f ? best.seg_lanes * 2 : best.seg_lanes / 2;
        if (c.seg_lanes >= 1 && c.seg_lanes <= WAVE && c.seg_lanes != best.seg_lanes) consider(c);
    }
    for (size_t b = 0; b < a.nb; b++) {
        if (a.blocks[b].small && a.small_dense) continue;
        if (a.nbk[b] >= (uint32_t)XCD_COUNT) {  // the other bucket mode, at its own default lanes
            const bool merged = lvl_mode(a, best, b, r, (int)sizeof(T)).merge;
            SpmvPlan c = best;
            c.xcd[b] = merged ? 1 : 0;
            c.blanes[b] = 0;
            consider(c);
        }
        const uint32_t cur = (uint32_t)lvl_mode(a, best, b, r, (int)sizeof(T)).lanes;
        for (uint32_t f : {2u, 0u}) {
            SpmvPlan c = best;
            c.blanes[b] = f ? cur * 2 : cur / 2;
            if (c.blanes[b] >= 1 && c.blanes[b] <= WAVE && c.blanes[b] != cur) consider(c);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    best.tuned = true;
    plan = best;
    tune_cache_put(key, plan);
}

// S*p with the plan in use (the default one when none has been tuned), never tuning.
template <class T>
void spmv_fixed(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, hipStream_t st) {
    spmv_impl<T>(a, a.plan_for(sizeof(T) == 8, r), p, y, r, s, nullptr, st);
}

template <class T>
void spmv_planned(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, const uint32_t* done, hipStream_t st,
                  double* pq) {
    SpmvPlan& plan = a.plan_for(sizeof(T) == 8, r);
    if (a.autotune && !plan.tuned) autotune_spmv<T>(a, plan, p, y, r, s, st);
    spmv_impl<T>(a, plan, p, y, r, s, done, st, pq);
}
}  // namespace fh
