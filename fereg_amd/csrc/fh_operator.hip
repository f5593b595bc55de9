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
#include <algorithm>

#include "fh_internal.hpp"

namespace fh {

namespace {

constexpr int WG = 256;
constexpr int WAVES = WG / WAVE;

template <class T>
__device__ __forceinline__ T shfl_xor_t(T v, int o) { return __shfl_xor(v, o, WAVE); }

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
        for (uint32_t it = 0; it < tmax; it++) {
            const uint32_t c = c0 + it * L + li;
            if (c < c1) {
                T u[RM];
                gather_u<T, RM>(A, p, c, r, u);
                const T wc = A.cw ? A.cw[c] : T(1);
#pragma unroll
                for (int j = 0; j < RM; j++) tv[j] += wc * u[j];
            }
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
template <class T, int RM>
__global__ __launch_bounds__(WG) void k_lvl_spmv(const uint32_t* __restrict__ loff, const uint32_t* __restrict__ lseg,
                                                 const uint32_t* __restrict__ lother, const T* __restrict__ lw,
                                                 size_t g, uint32_t off_b, int nother, OtherOffs oo,
                                                 const T* __restrict__ p, const T* __restrict__ t, int r, int L,
                                                 T* __restrict__ y, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * gpw; base < g; base += nwt * gpw) {
        const size_t lv = base + gl;
        const bool valid = lv < g;
        const uint32_t j0 = valid ? loff[lv] : 0, j1 = valid ? loff[lv + 1] : 0;
        const uint32_t tmax = wave_max_u32((j1 - j0 + L - 1) / L);
        T pb[RM], acc[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) {
            pb[j] = (valid && j < r) ? p[(size_t)(off_b + lv) * r + j] : T(0);
            acc[j] = T(0);
        }
        for (uint32_t it = 0; it < tmax; it++) {
            const uint32_t pos = j0 + it * L + li;
            if (pos < j1) {
                const T wc = lw ? lw[pos] : T(1);
                const size_t s = lseg[pos];
                T u[RM];
#pragma unroll
                for (int j = 0; j < RM; j++) u[j] = pb[j];
                for (int o = 0; o < nother; o++) {
                    const size_t k = (size_t)(oo.off[o] + lother[(size_t)pos * nother + o]) * r;
#pragma unroll
                    for (int j = 0; j < RM; j++)
                        if (j < r) u[j] += p[k + j];
                }
#pragma unroll
                for (int j = 0; j < RM; j++)
                    if (j < r) acc[j] += wc * (u[j] - t[s * r + j]);
            }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) acc[j] = gsum(acc[j], L);
        if (valid && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) y[(size_t)(off_b + lv) * r + j] = acc[j];
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

// rhs of a large block: pull of z_c along the level lists.
template <int RM>
__global__ __launch_bounds__(WG) void k_lvl_pull(const uint32_t* __restrict__ loff, const uint32_t* __restrict__ lperm,
                                                 size_t g, const double* __restrict__ zc, int r, int L,
                                                 double* __restrict__ out) {
    const int w = threadIdx.x / WAVE, l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwt = (size_t)gridDim.x * WAVES;
    for (size_t base = ((size_t)blockIdx.x * WAVES + w) * gpw; base < g; base += nwt * gpw) {
        const size_t lv = base + gl;
        const bool valid = lv < g;
        const uint32_t j0 = valid ? loff[lv] : 0, j1 = valid ? loff[lv + 1] : 0;
        const uint32_t tmax = wave_max_u32((j1 - j0 + L - 1) / L);
        double acc[RM];
#pragma unroll
        for (int j = 0; j < RM; j++) acc[j] = 0.0;
        for (uint32_t it = 0; it < tmax; it++) {
            const uint32_t pos = j0 + it * L + li;
            if (pos < j1) {
                const size_t c = lperm[pos];
#pragma unroll
                for (int j = 0; j < RM; j++)
                    if (j < r) acc[j] += zc[c * r + j];
            }
        }
#pragma unroll
        for (int j = 0; j < RM; j++) acc[j] = gsum(acc[j], L);
        if (valid && li == 0)
#pragma unroll
            for (int j = 0; j < RM; j++)
                if (j < r) out[lv * r + j] = acc[j];
    }
}

// ---------------------------------------------------------------- back-substitution (FP64)
template <int RM>
__global__ __launch_bounds__(WG) void k_resid_seg(OpArgs<double> A, double* __restrict__ rows, int r, int L,
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

template <class T>
size_t lds_bytes(const Absorber& a, int r) {
    return (size_t)WAVES * a.ms * r * sizeof(T);
}
constexpr size_t LDS_ACC_MAX = 96 * 1024;

inline unsigned seg_grid(const Absorber& a) {
    const size_t spw = WAVE / a.seg_lanes;
    return grid_for(a.g1, (unsigned)(spw * WAVES), 2048);
}

}  // namespace

template <class T>
OpArgs<T> op_args(const Absorber& a) {
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
template OpArgs<double> op_args<double>(const Absorber&);
template OpArgs<float> op_args<float>(const Absorber&);

#define FH_DISPATCH_R(r, KERNEL_CALL)                            \
    do {                                                         \
        if ((r) <= 1) { constexpr int RM = 1; KERNEL_CALL; }     \
        else if ((r) <= 2) { constexpr int RM = 2; KERNEL_CALL; } \
        else if ((r) <= 4) { constexpr int RM = 4; KERNEL_CALL; } \
        else if ((r) <= 8) { constexpr int RM = 8; KERNEL_CALL; } \
        else { constexpr int RM = 16; KERNEL_CALL; }             \
    } while (0)

template <class T>
void spmv(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, const uint32_t* done, hipStream_t st) {
    require(r >= 1 && r <= 16, "spmv: 1..16 columns");
    OpArgs<T> A = op_args<T>(a);
    const unsigned grid = seg_grid(a);
    const size_t accl = a.ms * (size_t)r;
    const bool lds = lds_bytes<T>(a, r) <= LDS_ACC_MAX;
    s.t.reserve(std::max<size_t>(a.g1 * r, 1));
    s.part.reserve(std::max<size_t>(lds ? (size_t)grid * accl : accl, 1));
    if (!lds && accl) FH_HIP(hipMemsetAsync(s.part.p, 0, accl * sizeof(T), st));
    const int L = (int)a.seg_lanes;
    const size_t shm = lds ? lds_bytes<T>(a, r) : 0;
    FH_DISPATCH_R(r, hipLaunchKernelGGL((k_seg_spmv<T, RM>), dim3(grid), dim3(WG), shm, st, A, p, r, L, s.t.p, s.part.p,
                                        (int)lds, done));
    FH_CHECK_LAUNCH();
    if (accl) {
        if (lds) {
            hipLaunchKernelGGL(k_reduce_small<T>, dim3((unsigned)std::min<size_t>(accl, 8192)), dim3(WG), 0, st,
                               s.part.p, (size_t)grid, accl, y, done);
        } else {
            FH_HIP(hipMemcpyAsync(y, s.part.p, accl * sizeof(T), hipMemcpyDeviceToDevice, st));
        }
        FH_CHECK_LAUNCH();
    }
    const WeightArrays<T>* W;
    if constexpr (std::is_same<T, double>::value) W = &a.w64; else W = &a.w32;
    const int nb = (int)a.nb;
    for (int b = 0; b < nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (bi.small) continue;
        OtherOffs oo{};
        int k = 0;
        for (int bb = 0; bb < nb; bb++)
            if (bb != b) oo.off[k++] = (uint32_t)a.blocks[bb].off;
        const int Ll = (int)a.lvl_lanes;
        const unsigned lg = grid_for(bi.g, (unsigned)((WAVE / Ll) * WAVES), 4096);
        const T* lw = a.unit() ? nullptr : W->lw[b].p;
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_lvl_spmv<T, RM>), dim3(lg), dim3(WG), 0, st, a.loff[b].p, a.lseg[b].p,
                                            a.lother[b].p, lw, (size_t)bi.g, (uint32_t)bi.off, nb - 1, oo, p, s.t.p, r,
                                            Ll, y, done));
        FH_CHECK_LAUNCH();
    }
}
template void spmv<double>(const Absorber&, const double*, double*, int, Scratch<double>&, const uint32_t*, hipStream_t);
template void spmv<float>(const Absorber&, const float*, float*, int, Scratch<float>&, const uint32_t*, hipStream_t);

void rhs_pass(Absorber& a, const double* rows, int r, double* rhs, double* b1, double* wsq_wsum) {
    require(r >= 1 && r <= 16, "rhs: 1..16 columns");
    hipStream_t st = a.st;
    OpArgs<double> A = op_args<double>(a);
    const unsigned grid = seg_grid(a);
    const size_t accl = a.ms * (size_t)r;
    const bool lds = lds_bytes<double>(a, r) <= LDS_ACC_MAX;
    Scratch<double>& s = a.s64;
    s.part.reserve(std::max<size_t>(lds ? (size_t)grid * accl : accl, 1));
    if (!lds && accl) FH_HIP(hipMemsetAsync(s.part.p, 0, accl * sizeof(double), st));
    s.dpart.reserve((size_t)grid * 2 * r);
    bool has_large = false;
    for (auto& bi : a.blocks) has_large |= !bi.small;
    if (has_large) s.zc.reserve(a.M * r);
    const int L = (int)a.seg_lanes;
    const size_t shm = lds ? lds_bytes<double>(a, r) : 0;
    const double* wr = a.weighted ? a.w_rows.p : nullptr;
    FH_DISPATCH_R(r, hipLaunchKernelGGL((k_rhs_seg<RM>), dim3(grid), dim3(WG), shm, st, A, rows, wr, r, L, b1,
                                        has_large ? s.zc.p : nullptr, s.part.p, s.dpart.p, (int)lds));
    FH_CHECK_LAUNCH();
    if (accl) {
        if (lds)
            hipLaunchKernelGGL(k_reduce_small<double>, dim3((unsigned)std::min<size_t>(accl, 8192)), dim3(WG), 0, st,
                               s.part.p, (size_t)grid, accl, rhs, nullptr);
        else
            FH_HIP(hipMemcpyAsync(rhs, s.part.p, accl * sizeof(double), hipMemcpyDeviceToDevice, st));
        FH_CHECK_LAUNCH();
    }
    for (size_t b = 0; b < a.nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (bi.small) continue;
        const int Ll = (int)a.lvl_lanes;
        const unsigned lg = grid_for(bi.g, (unsigned)((WAVE / Ll) * WAVES), 4096);
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_lvl_pull<RM>), dim3(lg), dim3(WG), 0, st, a.loff[b].p, a.lperm[b].p,
                                            (size_t)bi.g, s.zc.p, r, Ll, rhs + bi.off * r));
        FH_CHECK_LAUNCH();
    }
    // wsq / wsum: fixed-order sum of the workgroup partials on the host
    std::vector<double> h((size_t)grid * 2 * r);
    FH_HIP(hipMemcpyAsync(h.data(), s.dpart.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    for (int j = 0; j < 2 * r; j++) wsq_wsum[j] = 0.0;
    for (unsigned q = 0; q < grid; q++)
        for (int j = 0; j < 2 * r; j++) wsq_wsum[j] += h[(size_t)q * 2 * r + j];
}

void residualize_rows(Absorber& a, double* rows, int r, const double* x, const double* b1, double* alpha) {
    require(r >= 1 && r <= 16, "residualize: 1..16 columns");
    OpArgs<double> A = op_args<double>(a);
    const unsigned grid = seg_grid(a);
    const int L = (int)a.seg_lanes;
    FH_DISPATCH_R(r, hipLaunchKernelGGL((k_resid_seg<RM>), dim3(grid), dim3(WG), 0, a.st, A, rows, r, L, x, b1, alpha));
    FH_CHECK_LAUNCH();
}

}  // namespace fh
