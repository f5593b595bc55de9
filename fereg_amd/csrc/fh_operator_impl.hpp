// The reduced operator and the row passes on the GPU (operator.rs:894-1179):
//   S*p   = segment pass (t_s = sum_c W_c u_c / n1_s, small-coordinate scatter of
//           W_c (u_c - t_s) by wave multisplit into wave-private LDS accumulators) followed by
//           one level pass per large block over its level-major transpose
//           (y_l = sum_{c in l} W_c (u_c - t_{s(c)}), no float atomics);
//   rhs   = segment pass over the rows (cell sums, b1, sum w v^2, sum w v; z_c to a scratch)
//           + small scatter + level pull of z_c for the large blocks;
//   back-substitution: alpha_s = (b1_s - sum_c W_c u_c)/n1_s, v_i -= alpha_s + u_c in place.
// Every pass assigns L lanes to a segment (or level), L = next power of two of the mean length,
// so cell reads are coalesced; all sums are in a fixed order (run-to-run reproducible).
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

// Adds z (r columns) of every valid lane to acc[key*r + j], one wave reduction per distinct key
// (in lane order: deterministic).
template <class T, int RM>
__device__ __forceinline__ void multisplit_add(uint32_t key, bool valid, const T (&z)[RM], int r, T* acc,
                                               bool global_atomic) {
    uint64_t pending = __ballot(valid);
    const int l = lane_id();
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lk = __shfl((int)key, leader, WAVE);
        const uint64_t m = __ballot(valid && key == lk);
        const bool mine = (m >> l) & 1ull;
#pragma unroll
        for (int j = 0; j < RM; j++) {
            if (j < r) {
                T v = mine ? z[j] : T(0);
                v = wave_sum(v);
                if (l == leader) {
                    if (global_atomic) atomicAdd(&acc[(size_t)lk * r + j], v);
                    else acc[(size_t)lk * r + j] += v;
                }
            }
        }
        pending &= ~m;
    }
}

// Per-lane register accumulators of the small-coordinate scatter: a[k][j] += z[j] for key k,
// keys < NK. One predicated add per (key, column) and cell; the wave reduction happens once,
// at the end of the kernel (flush), not per cell.
template <class T, int RM, int NK>
struct SmallRegs {
    T a[NK > 0 ? NK : 1][RM];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int k = 0; k < (NK > 0 ? NK : 1); k++)
#pragma unroll
            for (int j = 0; j < RM; j++) a[k][j] = T(0);
    }
    __device__ __forceinline__ void add(uint32_t key, bool valid, const T (&z)[RM]) {
#pragma unroll
        for (int k = 0; k < NK; k++) {
            const bool hit = valid && key == (uint32_t)k;
#pragma unroll
            for (int j = 0; j < RM; j++) a[k][j] += hit ? z[j] : T(0);
        }
    }
    // part[k*r + j] (k < ms, j < r) = sum over the workgroup (fixed order). All threads call it.
    __device__ void flush(T* __restrict__ part, size_t ms, int r) {
        __shared__ T red[WAVES][NK > 0 ? NK : 1][RM];
        const int w = threadIdx.x / WAVE;
#pragma unroll
        for (int k = 0; k < NK; k++)
#pragma unroll
            for (int j = 0; j < RM; j++) {
                const T v = wave_sum(a[k][j]);
                if (lane_id() == 0) red[w][k][j] = v;
            }
        __syncthreads();
        for (int e = threadIdx.x; e < (int)ms * r; e += WG) {
            const int k = e / r, j = e % r;
            T v = T(0);
#pragma unroll
            for (int ww = 0; ww < WAVES; ww++) v += red[ww][k][j];
            part[e] = v;
        }
    }
};

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
template <class T, int RM>
__global__ __launch_bounds__(WG) void k_seg_spmv(OpArgs<T> A, const T* __restrict__ p, int r, int L,
                                                 T* __restrict__ t, T* __restrict__ part, int lds_acc,
                                                 const uint32_t* __restrict__ done) {
    if (done && *done) return;
    extern __shared__ unsigned char smem_raw[];
    T* acc = reinterpret_cast<T*>(smem_raw);
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int spw = WAVE / L, gl = l / L, li = l % L;
    const size_t accl = A.ms * (size_t)r;
    T* my = lds_acc ? acc + (size_t)w * accl : part;
    if (lds_acc)
        for (size_t i = l; i < accl; i += WAVE) my[i] = T(0);
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * spw; base < A.g1; base += nwt * spw) {
        const size_t s = base + gl;
        const bool sv = s < A.g1;
        const uint32_t c0 = sv ? A.seg_off[s] : 0, c1 = sv ? A.seg_off[s + 1] : 0;
        const uint32_t tmax = wave_max_u32((c1 - c0 + L - 1) / L);
        T tv[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) tv[j] = T(0);
        for (uint32_t it = 0; it < tmax; it += 2) {
            const uint32_t ca = c0 + it * L + li, cb = ca + L;
            const bool va = ca < c1, vb = cb < c1;
            T ua[RM], ub[RM];
#pragma unroll
            for (int j = 0; j < RM; j++) { ua[j] = T(0); ub[j] = T(0); }
            for (int b = 0; b < A.nb; b++) {
                const size_t ka = va ? (size_t)(A.boff[b] + A.cell_codes[(size_t)ca * A.nb + b]) * r : 0;
                const size_t kb = vb ? (size_t)(A.boff[b] + A.cell_codes[(size_t)cb * A.nb + b]) * r : 0;
#pragma unroll
                for (int j = 0; j < RM; j++)
                    if (j < r) {
                        if (va) ua[j] += p[ka + j];
                        if (vb) ub[j] += p[kb + j];
                    }
            }
            const T wa = va ? (A.cw ? A.cw[ca] : T(1)) : T(0);
            const T wb = vb ? (A.cw ? A.cw[cb] : T(1)) : T(0);
#pragma unroll
            for (int j = 0; j < RM; j++) tv[j] += wa * ua[j] + wb * ub[j];
        }
        const T inv = sv ? A.inv_n1[s] : T(0);
#pragma unroll
        for (int j = 0; j < RM; j++) tv[j] = gsum(tv[j], L) * inv;
        if (sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) t[s * r + j] = tv[j];
        if (A.nsb > 0) {
            for (uint32_t it = 0; it < tmax; it++) {
                const uint32_t c = c0 + it * L + li;
                const bool cv = c < c1;
                T z[RM];
#pragma unroll
                for (int j = 0; j < RM; j++) z[j] = T(0);
                if (cv) {
                    T u[RM];
                    gather_u<T, RM>(A, p, c, r, u);
                    const T wc = A.cw ? A.cw[c] : T(1);
#pragma unroll
                    for (int j = 0; j < RM; j++) z[j] = wc * (u[j] - tv[j]);
                }
                for (int b = 0; b < A.nsb; b++) {
                    const uint32_t key = cv ? A.boff[b] + A.cell_codes[(size_t)c * A.nb + b] : 0u;
                    multisplit_add<T, RM>(key, cv, z, r, my, !lds_acc);
                }
            }
        }
    }
    if (lds_acc) {
        __syncthreads();
        for (size_t i = threadIdx.x; i < accl; i += WG) {
            T v = T(0);
#pragma unroll
            for (int ww = 0; ww < WAVES; ww++) v += acc[(size_t)ww * accl + i];
            part[(size_t)blockIdx.x * accl + i] = v;
        }
    }
}

// t_s = sum_{c in s} W_c u_c / n1_s only (the small scatter runs in the level pass): NB
// (compile-time block count) and U = 8 trips per lane with every index load issued before the
// gathers, so a lane keeps ~8 x NB independent loads in flight (the passes are latency-bound
// with fewer: 32 resident waves per CU at most).
template <class T, int RM, int NB>
__global__ __launch_bounds__(WG) void k_seg_t(OpArgs<T> A, const T* __restrict__ p, int r, int L, T* __restrict__ t,
                                              const uint32_t* __restrict__ done) {
    if (done && *done) return;
    constexpr int U = 8;
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

// y[i] = sum over the workgroup partials part[q*len + i]: one workgroup per entry, strided
// per-thread sums and a fixed tree (deterministic, no serial chain of loads).
template <class T>
__global__ __launch_bounds__(WG) void k_reduce_small(const T* __restrict__ part, size_t nparts, size_t len,
                                                     T* __restrict__ y, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    __shared__ T red[WAVES];
    for (size_t i = blockIdx.x; i < len; i += gridDim.x) {
        T s = T(0);
        for (size_t q = threadIdx.x; q < nparts; q += WG) s += part[q * len + i];
        s = wave_sum(s);
        if (lane_id() == 0) red[threadIdx.x / WAVE] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            T t = T(0);
#pragma unroll
            for (int w = 0; w < WAVES; w++) t += red[w];
            y[i] = t;
        }
        __syncthreads();
    }
}

struct OtherOffs {
    uint32_t off[NB_MAX];
};

// ---------------------------------------------------------------- S*p, level pass of one large block
// y_l = sum_{c in l} W_c (u_c - t_{s(c)}) over the level-major records; with NK > 0 the same
// W_c (u_c - t_s) is also scattered into the nsb small blocks (keys < NK) through per-lane
// registers, and each workgroup writes its partial small vector to part[blockIdx.x][ms*r].
template <class T, int RM, int NK, int NO>
__global__ __launch_bounds__(WG) void k_lvl_spmv(const uint32_t* __restrict__ loff, const uint32_t* __restrict__ lseg,
                                                 const uint32_t* __restrict__ lother, const T* __restrict__ lw,
                                                 size_t g, uint32_t gb, uint32_t off_b, OtherOffs oo, int nother, int nsb,
                                                 size_t ms, const T* __restrict__ p, const T* __restrict__ t, int r,
                                                 int L, T* __restrict__ y, T* __restrict__ ypart, T* __restrict__ part,
                                                 const uint32_t* __restrict__ done) {
    if (done && *done) return;
    constexpr int U = 8;
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    SmallRegs<T, RM, NK> sm;
    if (NK > 0) sm.zero();
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
                if (NK > 0) {
                    // small blocks come first among the other blocks
#pragma unroll
                    for (int sb = 0; sb < NO; sb++)
                        if (sb < nsb) sm.add(oo.off[sb] + oc[h][sb], pv[h], z);
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
    if (NK > 0) sm.flush(part + (size_t)blockIdx.x * ms * r, ms, r);
}

// y[l] = sum over the buckets of ypart[k*gb + l] (fixed order).
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

template <int RM>
__global__ __launch_bounds__(WG) void k_rhs_seg(OpArgs<double> A, const double* __restrict__ rows,
                                                const double* __restrict__ wr, int r, int L, double* __restrict__ b1,
                                                double* __restrict__ zc, double* __restrict__ part,
                                                double* __restrict__ spart, int lds_acc) {
    extern __shared__ unsigned char smem_raw[];
    double* acc = reinterpret_cast<double*>(smem_raw);
    __shared__ double red[WAVES][2 * 16];
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int spw = WAVE / L, gl = l / L, li = l % L;
    const size_t accl = A.ms * (size_t)r;
    double* my = lds_acc ? acc + (size_t)w * accl : part;
    if (lds_acc)
        for (size_t i = l; i < accl; i += WAVE) my[i] = 0.0;
    double wsq[RM], wsum[RM];
#pragma unroll
    for (int j = 0; j < RM; j++) { wsq[j] = 0.0; wsum[j] = 0.0; }
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * spw; base < A.g1; base += nwt * spw) {
        const size_t s = base + gl;
        const bool sv = s < A.g1;
        const uint32_t c0 = sv ? A.seg_off[s] : 0, c1 = sv ? A.seg_off[s + 1] : 0;
        const uint32_t tmax = wave_max_u32((c1 - c0 + L - 1) / L);
        double bv[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) bv[j] = 0.0;
        for (uint32_t it = 0; it < tmax; it++) {
            const uint32_t c = c0 + it * L + li;
            if (c < c1) {
                double sc[RM];
                cell_sums<RM>(A, rows, wr, c, r, sc, wsq, wsum, true);
#pragma unroll
                for (int j = 0; j < RM; j++) bv[j] += sc[j];
            }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) bv[j] = gsum(bv[j], L);
        if (sv && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) b1[s * r + j] = bv[j];
        const double inv = sv ? A.inv_n1[s] : 0.0;
        for (uint32_t it = 0; it < tmax; it++) {
            const uint32_t c = c0 + it * L + li;
            const bool cv = c < c1;
            double z[RM];
#pragma unroll
            for (int j = 0; j < RM; j++) z[j] = 0.0;
            if (cv) {
                double sc[RM], d0[RM], d1[RM];
                cell_sums<RM>(A, rows, wr, c, r, sc, d0, d1, false);
                const double f = (A.cw ? A.cw[c] : 1.0) * inv;
#pragma unroll
                for (int j = 0; j < RM; j++) z[j] = sc[j] - f * bv[j];
                if (zc)
#pragma unroll
                    for (int j = 0; j < RM; j++)
                        if (j < r) zc[(size_t)c * r + j] = z[j];
            }
            for (int b = 0; b < A.nsb; b++) {
                const uint32_t key = cv ? A.boff[b] + A.cell_codes[(size_t)c * A.nb + b] : 0u;
                multisplit_add<double, RM>(key, cv, z, r, my, !lds_acc);
            }
        }
    }
    // per-workgroup sums of w v^2 and w v
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
    if (lds_acc) {
        for (size_t i = threadIdx.x; i < accl; i += WG) {
            double v = 0.0;
#pragma unroll
            for (int ww = 0; ww < WAVES; ww++) v += acc[(size_t)ww * accl + i];
            part[(size_t)blockIdx.x * accl + i] = v;
        }
    }
}

// rhs of a large block: pull of z_c along the level lists; with NK > 0 z_c is also scattered
// into the small blocks (keys from the level-major records) as in k_lvl_spmv.
template <int RM, int NK>
__global__ __launch_bounds__(WG) void k_lvl_pull(const uint32_t* __restrict__ loff, const uint32_t* __restrict__ lperm,
                                                 const uint32_t* __restrict__ lother, int nother, OtherOffs oo, int nsb,
                                                 size_t ms, size_t g, uint32_t gb, const double* __restrict__ zc, int r,
                                                 int L, double* __restrict__ out, double* __restrict__ ypart,
                                                 double* __restrict__ part) {
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    SmallRegs<double, RM, NK> sm;
    if (NK > 0) sm.zero();
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * gpw; base < g; base += nwt * gpw) {
        const size_t lv = base + gl;  // item = bucket * gb + level
        const bool valid = lv < g;
        const uint32_t j0 = valid ? loff[lv] : 0, j1 = valid ? loff[lv + 1] : 0;
        const uint32_t tmax = wave_max_u32((j1 - j0 + L - 1) / L);
        double acc[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) acc[j] = 0.0;
        for (uint32_t it = 0; it < tmax; it++) {
            const uint32_t pos = j0 + it * L + li;
            const bool pv = pos < j1;
            double z[RM];
            const size_t c = pv ? lperm[pos] : 0;
#pragma unroll
            for (int j = 0; j < RM; j++) {
                z[j] = (pv && j < r) ? zc[c * r + j] : 0.0;
                acc[j] += z[j];
            }
            if (NK > 0)
                for (int sb = 0; sb < nsb; sb++) {
                    const uint32_t code = pv ? lother[(size_t)pos * nother + sb] : 0u;
                    sm.add(oo.off[sb] + code, pv, z);
                }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) acc[j] = gsum(acc[j], L);
        if (valid && li == 0) {
            double* dst = ypart ? ypart + lv * r : out + (lv % gb) * r;
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) dst[j] = acc[j];
        }
    }
    if (NK > 0) sm.flush(part + (size_t)blockIdx.x * ms * r, ms, r);
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
template <int RM>
constexpr int resid_trips() { return RM <= 2 ? 8 : (RM <= 4 ? 4 : 2); }

template <int RM, int NB>
__global__ __launch_bounds__(WG) void k_resid_seg(OpArgs<double> A, double* __restrict__ rows, int r, int L,
                                                  const double* __restrict__ x, const double* __restrict__ b1,
                                                  double* __restrict__ alpha) {
    constexpr int U = resid_trips<RM>();
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
    constexpr int U = resid_trips<RM>();
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

template <class T>
size_t lds_bytes(const Absorber& a, int r) {
    return (size_t)WAVES * a.ms * r * sizeof(T);
}
constexpr size_t LDS_ACC_MAX = 96 * 1024;

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

// Register budget of the small-coordinate accumulators (bytes per lane).
constexpr int SMALL_REG_BYTES = 576;
template <class T, int RM, int NK>
constexpr bool small_fits() { return NK * RM * (int)sizeof(T) <= SMALL_REG_BYTES; }

// Calls f(std::integral_constant<int, NK>) with the smallest NK in {16, 32} >= ms whose
// accumulators fit the register budget; returns false when none does.
template <class T, int RM, class F>
bool with_small_nk(size_t ms, F&& f) {
    if constexpr (small_fits<T, RM, 16>()) if (ms <= 16) { f(std::integral_constant<int, 16>{}); return true; }
    if constexpr (small_fits<T, RM, 32>()) if (ms <= 32) { f(std::integral_constant<int, 32>{}); return true; }
    return false;
}

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

template <class T>
DBuf<T>& ypart_buf(Scratch<T>& s) { return s.ypart; }

// Index of the first large block (-1 when every block is small).
inline int first_large(const Absorber& a) {
    for (size_t b = 0; b < a.nb; b++)
        if (!a.blocks[b].small) return (int)b;
    return -1;
}

inline OtherOffs other_offs(const Absorber& a, int b) {
    OtherOffs oo{};
    int k = 0;
    for (int bb = 0; bb < (int)a.nb; bb++)
        if (bb != b) oo.off[k++] = (uint32_t)a.blocks[bb].off;
    return oo;
}

inline unsigned lvl_grid(const Absorber& a, size_t g) {
    const int Ll = (int)a.lvl_lanes;
    return grid_for(g, (unsigned)((WAVE / Ll) * WAVES), 1u << 20);
}

template <class T>
void spmv_impl(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, const uint32_t* done, hipStream_t st) {
    require(r >= 1 && r <= 16, "spmv: 1..16 columns");
    OpArgs<T> A = op_args_impl<T>(a);
    const size_t accl = a.ms * (size_t)r;
    const int fl = first_large(a);
    const int nb = (int)a.nb;
    const WeightArrays<T>* W;
    if constexpr (std::is_same<T, double>::value) W = &a.w64; else W = &a.w32;
    s.t.reserve(std::max<size_t>(a.g1 * r, 1));
    const int L = (int)a.seg_lanes;
    const unsigned grid = seg_grid(a);
    // Small scatter inside the first large block's level pass (per-lane registers) when the
    // accumulators fit; otherwise in the segment pass (wave multisplit).
    bool small_in_lvl = false;
    if (accl && fl >= 0) {
        FH_DISPATCH_R(r, small_in_lvl = with_small_nk<T, RM>(a.ms, [](auto) {}));
    }
    if (accl && !small_in_lvl) {
        const bool lds = lds_bytes<T>(a, r) <= LDS_ACC_MAX;
        s.part.reserve(std::max<size_t>(lds ? (size_t)grid * accl : accl, 1));
        if (!lds) FH_HIP(hipMemsetAsync(s.part.p, 0, accl * sizeof(T), st));
        const size_t shm = lds ? lds_bytes<T>(a, r) : 0;
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_seg_spmv<T, RM>), dim3(grid), dim3(WG), shm, st, A, p, r, L, s.t.p,
                                            s.part.p, (int)lds, done));
        FH_CHECK_LAUNCH();
        if (lds)
            hipLaunchKernelGGL(k_reduce_small<T>, dim3((unsigned)std::min<size_t>(accl, 8192)), dim3(WG), 0, st,
                               s.part.p, (size_t)grid, accl, y, done);
        else
            FH_HIP(hipMemcpyAsync(y, s.part.p, accl * sizeof(T), hipMemcpyDeviceToDevice, st));
        FH_CHECK_LAUNCH();
    } else if (nb >= 1 && nb <= 4) {
        FH_DISPATCH_R(r, with_nb(nb, [&](auto nbc) {
            constexpr int NB = decltype(nbc)::value;
            hipLaunchKernelGGL((k_seg_t<T, RM, NB>), dim3(grid), dim3(WG), 0, st, A, p, r, L, s.t.p, done);
        }));
        FH_CHECK_LAUNCH();
    } else {
        OpArgs<T> At = A;
        At.nsb = 0;  // t only
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_seg_spmv<T, RM>), dim3(grid), dim3(WG), 0, st, At, p, r, L, s.t.p,
                                            (T*)nullptr, 0, done));
        FH_CHECK_LAUNCH();
    }
    int nsb = 0;
    for (auto& bi : a.blocks) nsb += bi.small ? 1 : 0;
    for (int b = 0; b < nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (bi.small) continue;
        const OtherOffs oo = other_offs(a, b);
        const int Ll = (int)a.lvl_lanes;
        const uint32_t nbk = a.nbk[b];
        const size_t items = (size_t)nbk * bi.g;
        const unsigned lg = lvl_grid(a, items);
        const T* lw = a.unit() ? nullptr : W->lw[b].p;
        T* ypart = nullptr;
        if (nbk > 1) {
            ypart_buf<T>(s).reserve(items * r);
            ypart = ypart_buf<T>(s).p;
        }
        if (b == fl && small_in_lvl) {
            s.part.reserve((size_t)lg * accl);
            FH_DISPATCH_R(r, with_small_nk<T, RM>(a.ms, [&](auto nk) {
                constexpr int NK = decltype(nk)::value;
                with_no(nb - 1, [&](auto noc) {
                    constexpr int NO = decltype(noc)::value;
                    hipLaunchKernelGGL((k_lvl_spmv<T, RM, NK, NO>), dim3(lg), dim3(WG), 0, st, a.boff[b].p,
                                       a.lseg[b].p, a.lother[b].p, lw, items, (uint32_t)bi.g, (uint32_t)bi.off, oo,
                                       nb - 1, nsb, (size_t)a.ms, p, s.t.p, r, Ll, y, ypart, s.part.p, done);
                });
            }));
            FH_CHECK_LAUNCH();
            hipLaunchKernelGGL(k_reduce_small<T>, dim3((unsigned)std::min<size_t>(accl, 8192)), dim3(WG), 0, st,
                               s.part.p, (size_t)lg, accl, y, done);
        } else {
            FH_DISPATCH_R(r, with_no(nb - 1, [&](auto noc) {
                constexpr int NO = decltype(noc)::value;
                hipLaunchKernelGGL((k_lvl_spmv<T, RM, 0, NO>), dim3(lg), dim3(WG), 0, st, a.boff[b].p, a.lseg[b].p,
                                   a.lother[b].p, lw, items, (uint32_t)bi.g, (uint32_t)bi.off, oo, nb - 1, 0, (size_t)0,
                                   p, s.t.p, r, Ll, y, ypart, (T*)nullptr, done);
            }));
        }
        FH_CHECK_LAUNCH();
        if (ypart) {
            hipLaunchKernelGGL(k_bucket_reduce<T>, dim3(grid_for(bi.g * r, WG, 8192)), dim3(WG), 0, st, ypart, nbk,
                               (size_t)bi.g, r, y + bi.off * r, done);
            FH_CHECK_LAUNCH();
        }
    }
}
}  // namespace fh
