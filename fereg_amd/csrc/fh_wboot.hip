// WCR11 wild cluster bootstrap of one linear hypothesis on the residualized OLS block
// (crates/fereg-core/src/wildboot/ols.rs:243-649, clusters = bootstrap clusters, Rademacher
// weights weights.rs:65-139 from SplitMix64 rng.rs:4-31, p-value and test inversion
// mod.rs:529-749).
//
// Per cluster g (targets t = 0: u_hat, t = 1: -X~ M_R of the affine interval, ols.rs:274-312):
//   S_t[g] = sum_{i in g} w_i u_t,i x~_i, B[g] = sum w_i z_i x~_i (z = x~'C, C = bread R'),
//   Phi_t = bread S_t, A_t = C'S_t (= R Phi_t, the cell score of the cluster).
// Per draw b with signs V_b (exact +-1):
//   num_t = V_b.A_t, db_t = V_b.Phi_t, q[c,t] = V_bc A_t[c] - B_c.db_t, v = ssc sum_c q q'.
// Since V^2 = 1, sum_c q_t1 q_t2 = AA[t1,t2] - db_t2.EA_t1 - db_t1.EA_t2 + db_t1' BB db_t2 with
// EA_t = V_b.(A_t B), AA = sum A_t1 A_t2, BB = sum B B'. Every V-dependent quantity is one
// product V (B x G) . P (G x (2 + 4k)), P = [A_0, A_1, Phi_0, Phi_1, A_0 B, A_1 B].
//
// That product runs on int8 MFMA: V is exact in int8; each column of P is scaled into
// [-1/4, 1/4] and split into 7 signed 8-bit limbs (56 fixed-point bits); int32 accumulation
// of +-limb is exact over any G < 2^24, so the contraction is exact for the 56-bit P and
// bit-reproducible; the limbs are recombined in FP64 per draw.
#include <chrono>
#include <cmath>
#include <cstring>

#include "fh_fit.hpp"

#include <rocblas/rocblas.h>

namespace fh {

namespace {

constexpr unsigned TPB = 256;
constexpr int LIMBS = 7;
// 16-column tiles per GEMM workgroup: 8 keeps the accumulators small enough for several
// workgroups per CU, whose waves hide each step's HBM fetch (C3, 294 limb columns: 19 tiles per
// workgroup 0.97 ms, 8 tiles 0.60 ms; 4, 6, 10 and 12 tiles 0.64-0.68 ms)
constexpr int WB_NTC = 8;
constexpr int WB_ROWS = 128;  // draws per GEMM workgroup (4 waves x 2 tiles of 16)
constexpr int KMAX = 21;
constexpr size_t PC_GMAX = 512;  // clusters of the per-cluster projection (the reference's Path 2 limit)
constexpr uint64_t SM_GAMMA = 0x9E3779B97F4A7C15ull;
constexpr uint64_t SM_STREAM = 0xD1B54A32D192ED03ull;

typedef int v4i __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ uint64_t sm_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// row[q] = src[q], q < LD (LD even, src 16-byte aligned): LD / 2 loads of 16 bytes.
template <int LD>
__device__ __forceinline__ void ld_row16(const double* __restrict__ src, double (&row)[LD]) {
#pragma unroll
    for (int q = 0; q < LD / 2; q++) {
        const double2 v = reinterpret_cast<const double2*>(src)[q];
        row[2 * q] = v.x;
        row[2 * q + 1] = v.y;
    }
}

// S[g][0..2][a]: sums over the rows i of cluster g of (w u0, w u1, w z)_i x~_ia, with z = x~_i'C,
// u0 = e_i and u1 = -x~_i'M_R (x~ the kept regressors): L lanes per cluster, each summing whole
// runs of adjacent rows (the cluster's runs strided over the lanes), then the lanes reduced by
// shuffles in a fixed order.
// LD > 0: every regressor is kept (x~_ia in column 1 + a) and the row stride is LD, so a row is
// read whole with 16-byte loads (LD / 2 requests for a random row instead of one per column);
// with ecol >= 0 the residual is column ecol of the same row.
template <int MK, int LD = 0>
__global__ __launch_bounds__(TPB) void k_wb_clsums(const uint32_t* __restrict__ off, const uint32_t* __restrict__ order,
                                                   const uint32_t* __restrict__ run_off, size_t G,
                                                   const double* __restrict__ rows, int R,
                                                   const int* __restrict__ kept, int mk, const double* __restrict__ C,
                                                   const double* __restrict__ MR, const double* __restrict__ e,
                                                   size_t es, const double* __restrict__ w, int L,
                                                   double* __restrict__ S, int ecol = -1) {
    const int l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwaves = (size_t)gridDim.x * (TPB / WAVE);
    for (size_t base = ((size_t)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE) * gpw; base < G; base += nwaves * gpw) {
        const size_t q = base + gl;
        const bool valid = q < G;
        double s0[MK], s1[MK], sb[MK];
#pragma unroll
        for (int a = 0; a < MK; a++) s0[a] = s1[a] = sb[a] = 0.0;
        const uint32_t j0 = valid ? off[q] : 0, j1 = valid ? off[q + 1] : 0;
        for (uint32_t j = j0 + li; j < j1; j += L) {
            const uint32_t run = order[j];
            const uint32_t i1 = run_off[run + 1];
            for (uint32_t i = run_off[run]; i < i1; i++) {
                double x[MK], z = 0.0, u1 = 0.0, ev;
                if constexpr (LD > 0) {
                    double row[LD];
                    ld_row16<LD>(rows + (size_t)i * LD, row);
#pragma unroll
                    for (int a = 0; a < MK; a++) x[a] = a + 1 < LD && a < mk ? row[a + 1 < LD ? a + 1 : 0] : 0.0;
                    if (ecol >= 0) {
                        ev = 0.0;
#pragma unroll
                        for (int q = 0; q < LD; q++) ev = q == ecol ? row[q] : ev;
                    } else {
                        ev = e[(size_t)i * es];
                    }
                } else {
#pragma unroll
                    for (int a = 0; a < MK; a++) x[a] = a < mk ? rows[(size_t)i * R + 1 + kept[a]] : 0.0;
                    ev = e[(size_t)i * es];
                }
#pragma unroll
                for (int a = 0; a < MK; a++)
                    if (a < mk) {
                        z += x[a] * C[a];
                        u1 -= x[a] * MR[a];
                    }
                const double wi = w ? w[i] : 1.0;
                const double wz = wi * z, w0 = wi * ev, w1 = wi * u1;
#pragma unroll
                for (int a = 0; a < MK; a++) {
                    s0[a] += w0 * x[a];
                    s1[a] += w1 * x[a];
                    sb[a] += wz * x[a];
                }
            }
        }
#pragma unroll
        for (int a = 0; a < MK; a++)
            for (int o = L >> 1; o > 0; o >>= 1) {
                s0[a] += __shfl_xor(s0[a], o, WAVE);
                s1[a] += __shfl_xor(s1[a], o, WAVE);
                sb[a] += __shfl_xor(sb[a], o, WAVE);
            }
        if (valid && li == 0)
#pragma unroll
            for (int a = 0; a < MK; a++)
                if (a < mk) {
                    S[(q * 3 + 0) * mk + a] = s0[a];
                    S[(q * 3 + 1) * mk + a] = s1[a];
                    S[(q * 3 + 2) * mk + a] = sb[a];
                }
    }
}

// P[g] = [A0, A1, Phi0, Phi1, A0 B, A1 B] and Q[g] = [A0, A1, B]; per-workgroup column maxima
// of |P| in cmax[blockIdx.x][c].
__global__ __launch_bounds__(TPB) void k_wb_prow(const double* __restrict__ S, size_t G, int mk,
                                                 const double* __restrict__ bread, const double* __restrict__ C,
                                                 double* __restrict__ P, double* __restrict__ Q,
                                                 double* __restrict__ cmax) {
    const int NC = 2 + 4 * mk;
    __shared__ double red[TPB / WAVE][2 + 4 * KMAX];
    double mx[2 + 4 * KMAX];
    for (int c = 0; c < NC; c++) mx[c] = 0.0;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < G; g += (size_t)gridDim.x * blockDim.x) {
        const double* S0 = S + (g * 3) * mk;
        const double* S1 = S0 + mk;
        const double* Bc = S1 + mk;
        double* p = P + g * NC;
        double A0 = 0.0, A1 = 0.0;
        for (int a = 0; a < mk; a++) {
            A0 += C[a] * S0[a];
            A1 += C[a] * S1[a];
        }
        p[0] = A0;
        p[1] = A1;
        for (int a = 0; a < mk; a++) {
            double f0 = 0.0, f1 = 0.0;
            for (int b = 0; b < mk; b++) {
                f0 += bread[a * mk + b] * S0[b];
                f1 += bread[a * mk + b] * S1[b];
            }
            p[2 + a] = f0;
            p[2 + mk + a] = f1;
            p[2 + 2 * mk + a] = A0 * Bc[a];
            p[2 + 3 * mk + a] = A1 * Bc[a];
        }
        double* q = Q + g * (2 + mk);
        q[0] = A0;
        q[1] = A1;
        for (int a = 0; a < mk; a++) q[2 + a] = Bc[a];
        for (int c = 0; c < NC; c++) mx[c] = fmax(mx[c], fabs(p[c]));
    }
    const int wv = threadIdx.x / WAVE;
    for (int c = 0; c < NC; c++) {
        double v = mx[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
        if (lane_id() == 0) red[wv][c] = v;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < NC; c += TPB) {
        double v = 0.0;
        for (int ww = 0; ww < (int)(TPB / WAVE); ww++) v = fmax(v, red[ww][c]);
        cmax[(size_t)blockIdx.x * NC + c] = v;
    }
}

// Pq[(l NC + c) Gpad + g] = limb l (most significant first) of round(P[g][c] * inv_scale[c]).
__global__ void k_wb_quant(const double* __restrict__ P, size_t G, int NC, const double* __restrict__ inv_scale,
                           size_t Gpad, int8_t* __restrict__ Pq) {
    const size_t total = G * NC;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t g = e / NC;
        const int c = (int)(e % NC);
        long long X = llrint(P[e] * inv_scale[c]);
        for (int l = LIMBS - 1; l >= 0; l--) {
            const long long d = ((X + 128) & 255) - 128;  // balanced base-256 digit
            Pq[((size_t)l * NC + c) * Gpad + g] = (int8_t)d;
            X = (X - d) >> 8;
        }
    }
}

// 16 signs from 16 bits (bit 0 -> +1, bit 1 -> -1), byte j of dword d = bit 4d + j.
__device__ __forceinline__ v4i sign_bytes(uint32_t bits16) {
    v4i a;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const uint32_t nib = (bits16 >> (4 * d)) & 0xFu;
        const uint32_t spread = (nib * 0x00204081u) & 0x01010101u;  // bit j -> byte j
        a[d] = (int)(0x01010101u | (spread * 0xFEu));
    }
    return a;
}

// part[split][b][col] = sum over this split's g of V[b][g] Pq[col][g] (int32, exact).
// Workgroup: 128 draws (4 waves x 2 row tiles) x up to WB_NTC column tiles; each 64-g step stages
// the column slice in LDS (80-byte rows: conflict-free 16-byte reads) and generates the
// signs of its draws from SplitMix64 (word s of draw b covers g in [64 s, 64 s + 64)).
__global__ __launch_bounds__(TPB) void k_wb_gemm(const int8_t* __restrict__ Pq, size_t Gpad, int ntiles,
                                                 uint64_t seed, int full, uint32_t steps_per_split, uint32_t nsteps,
                                                 int32_t* __restrict__ part, size_t Bpad, int ncolp) {
    __shared__ v4i lds[WB_NTC * 16][5];
    const int lane = lane_id(), w = threadIdx.x / WAVE;
    const int kg = lane >> 4, li = lane & 15;
    const int tile0 = blockIdx.z * WB_NTC;
    const int ntc = min(WB_NTC, ntiles - tile0);
    const uint64_t bbase = (uint64_t)blockIdx.x * WB_ROWS + (uint64_t)w * 32;
    uint64_t st0[2], bid[2];
#pragma unroll
    for (int rt = 0; rt < 2; rt++) {
        bid[rt] = bbase + 16 * rt + li;
        st0[rt] = seed ^ (bid[rt] * SM_STREAM);
    }
    v4i acc[2][WB_NTC];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int t = 0; t < WB_NTC; t++) acc[rt][t] = v4i{0, 0, 0, 0};
    const uint32_t s0 = blockIdx.y * steps_per_split;
    const uint32_t s1 = min(nsteps, s0 + steps_per_split);
    // the next step's slice is loaded into registers while the current one is multiplied
    constexpr int PER = (WB_NTC * 64 + TPB - 1) / TPB;
    v4i pre[PER];
    auto fetch = [&](uint32_t s) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int it = threadIdx.x + u * TPB;
            if (it < ntc * 64) {
                const int col = it >> 2, q = it & 3;
                pre[u] = *reinterpret_cast<const v4i*>(Pq + (size_t)(tile0 * 16 + col) * Gpad + (size_t)s * 64 + q * 16);
            }
        }
    };
    if (s0 < s1) fetch(s0);
    for (uint32_t s = s0; s < s1; s++) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int it = threadIdx.x + u * TPB;
            if (it < ntc * 64) lds[it >> 2][it & 3] = pre[u];
        }
        __syncthreads();
        if (s + 1 < s1) fetch(s + 1);
        v4i a[2];
#pragma unroll
        for (int rt = 0; rt < 2; rt++) {
            const uint64_t word = full ? (s == 0 ? bid[rt] : 0ull) : sm_mix(st0[rt] + (uint64_t)(s + 2) * SM_GAMMA);
            a[rt] = sign_bytes((uint32_t)(word >> (16 * kg)) & 0xFFFFu);
        }
#pragma unroll
        for (int t = 0; t < WB_NTC; t++) {
            if (t < ntc) {
                const v4i bf = lds[t * 16 + li][kg];
                acc[0][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[0], bf, acc[0][t], 0, 0, 0);
                acc[1][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[1], bf, acc[1][t], 0, 0, 0);
            }
        }
    }
    // D layout of the 16x16 tile: lane = column (lane & 15), rows 4 (lane >> 4) + r
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int t = 0; t < WB_NTC; t++) {
            if (t >= ntc) continue;
            const int col = (tile0 + t) * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint64_t b = bbase + 16 * rt + 4 * kg + r;
                part[((size_t)blockIdx.y * Bpad + b) * ncolp + col] = acc[rt][t][r];
            }
        }
}

// Per draw (one wave each): sum the splits (exact, int64) lane-strided over the limb columns,
// recombine the limbs in FP64, then lane 0 forms the affine coefficients
// (num0, num1, v00, v01, v11) of t(delta).
__global__ __launch_bounds__(TPB) void k_wb_post(const int32_t* __restrict__ part, int nsplit, size_t Bpad, int ncolp,
                                                 int NC, int mk, uint64_t reps, const double* __restrict__ scale,
                                                 const double* __restrict__ AA, const double* __restrict__ BB,
                                                 double ssc, double* __restrict__ coefs) {
    __shared__ long long S[TPB / WAVE][LIMBS * (2 + 4 * KMAX)];
    __shared__ double Y[TPB / WAVE][2 + 4 * KMAX];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    for (uint64_t b = (uint64_t)blockIdx.x * (TPB / WAVE) + w; b < reps; b += (uint64_t)gridDim.x * (TPB / WAVE)) {
        for (int col = lane; col < LIMBS * NC; col += WAVE) {
            long long s = 0;
            for (int k = 0; k < nsplit; k++) s += part[((size_t)k * Bpad + b) * ncolp + col];
            S[w][col] = s;
        }
        __builtin_amdgcn_wave_barrier();
        for (int c = lane; c < NC; c += WAVE) {
            double v = 0.0;
            for (int l = LIMBS - 1; l >= 0; l--) v += ldexp((double)S[w][l * NC + c], -8 * (l + 1));
            Y[w][c] = v * scale[c];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
            const double* y = Y[w];
            const double* db0 = y + 2;
            const double* db1 = y + 2 + mk;
            const double* ea0 = y + 2 + 2 * mk;
            const double* ea1 = y + 2 + 3 * mk;
            double e00 = 0.0, e01 = 0.0, e10 = 0.0, e11 = 0.0, q00 = 0.0, q01 = 0.0, q11 = 0.0;
            for (int a = 0; a < mk; a++) {
                e00 += db0[a] * ea0[a];
                e01 += db0[a] * ea1[a];
                e10 += db1[a] * ea0[a];
                e11 += db1[a] * ea1[a];
                double r0 = 0.0, r1 = 0.0;
                for (int c = 0; c < mk; c++) {
                    r0 += BB[a * mk + c] * db0[c];
                    r1 += BB[a * mk + c] * db1[c];
                }
                q00 += db0[a] * r0;
                q01 += db0[a] * r1;
                q11 += db1[a] * r1;
            }
            double* o = coefs + b * 5;
            o[0] = y[0];
            o[1] = y[1];
            o[2] = ssc * (AA[0] - 2.0 * e00 + q00);
            o[3] = ssc * (AA[1] - e10 - e01 + q01);
            o[4] = ssc * (AA[3] - 2.0 * e11 + q11);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// The per-cluster form (one wave per draw): column c of part holds limb l of V_b . P[:, c] at
// l NC + c; num_t = column t, q_t[c] = column 2 + t nc + c, v_t1t2 = ssc sum_c q_t1[c] q_t2[c],
// summed per lane over c = lane, lane + 64, ... and reduced over the wave in a fixed order.
__global__ __launch_bounds__(TPB) void k_wb_post_cells(const int32_t* __restrict__ part, int nsplit, size_t Bpad,
                                                       int ncolp, int NC, int nc, uint64_t reps,
                                                       const double* __restrict__ scale, double ssc,
                                                       double* __restrict__ coefs) {
    const int w = threadIdx.x / WAVE, lane = lane_id();
    for (uint64_t b = (uint64_t)blockIdx.x * (TPB / WAVE) + w; b < reps; b += (uint64_t)gridDim.x * (TPB / WAVE)) {
        const int32_t* pb = part + b * (size_t)ncolp;
        auto col = [&](int c) {
            double v = 0.0;
            for (int l = LIMBS - 1; l >= 0; l--) {
                long long s = 0;
                for (int k = 0; k < nsplit; k++) s += pb[(size_t)k * Bpad * ncolp + (size_t)l * NC + c];
                v += ldexp((double)s, -8 * (l + 1));
            }
            return v * scale[c];
        };
        double s00 = 0.0, s01 = 0.0, s11 = 0.0;
        for (int c = lane; c < nc; c += WAVE) {
            const double q0 = col(2 + c), q1 = col(2 + nc + c);
            s00 += q0 * q0;
            s01 += q0 * q1;
            s11 += q1 * q1;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            s00 += __shfl_xor(s00, o, WAVE);
            s01 += __shfl_xor(s01, o, WAVE);
            s11 += __shfl_xor(s11, o, WAVE);
        }
        const double n0 = col(0), n1 = col(1);
        if (lane == 0) {
            double* o = coefs + b * 5;
            o[0] = n0;
            o[1] = n1;
            o[2] = ssc * s00;
            o[3] = ssc * s01;
            o[4] = ssc * s11;
        }
    }
}

// Nesting of the absorbed terms in the clusters (dof.rs nested_within, by value), with no
// atomics on hot addresses: min/max cluster per cell, then per FE1 level over its cells
// (seg_off) and per block level over its level-major positions.
__global__ void k_wb_cell_span(const uint32_t* __restrict__ coo, const uint32_t* __restrict__ pi,
                               const uint32_t* __restrict__ cl, size_t M, uint2* __restrict__ span) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (size_t)gridDim.x * blockDim.x) {
        const size_t r0 = coo ? coo[c] : c, r1 = coo ? coo[c + 1] : c + 1;
        uint32_t mn = 0xFFFFFFFFu, mx = 0u;
        for (size_t i = r0; i < r1; i++) {
            const uint32_t k = cl[pi[i]];
            mn = min(mn, k);
            mx = max(mx, k);
        }
        span[c] = make_uint2(mn, mx);
    }
}

// *bad |= some group (cells off[q]..off[q+1], through perm when given) spans clusters; one
// flag write per workgroup. Groups: FE1 levels (off = seg_off) or (bucket, level) items.
__global__ __launch_bounds__(TPB) void k_wb_group_span(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ perm, const uint2* __restrict__ span,
                                                       size_t ngroups, uint2* __restrict__ gspan,
                                                       uint32_t* __restrict__ bad) {
    bool b = false;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < ngroups; q += (size_t)gridDim.x * blockDim.x) {
        uint32_t mn = 0xFFFFFFFFu, mx = 0u;
        for (uint32_t j = off[q]; j < off[q + 1]; j++) {
            const uint2 s = span[perm ? perm[j] : j];
            mn = min(mn, s.x);
            mx = max(mx, s.y);
        }
        if (gspan) gspan[q] = make_uint2(mn, mx);
        b |= mn < mx;
    }
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (__ballot(b) != 0 && lane_id() == 0) atomicOr(&any, 1);
    __syncthreads();
    if (threadIdx.x == 0 && any) atomicOr(bad, 1u);
}

// Levels cut into nbk buckets: combine the item spans of each level.
__global__ __launch_bounds__(TPB) void k_wb_level_span(const uint2* __restrict__ gspan, uint32_t nbk, size_t g,
                                                       uint32_t* __restrict__ bad) {
    bool b = false;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < g; l += (size_t)gridDim.x * blockDim.x) {
        uint32_t mn = 0xFFFFFFFFu, mx = 0u;
        for (uint32_t k = 0; k < nbk; k++) {
            const uint2 s = gspan[(size_t)k * g + l];
            mn = min(mn, s.x);
            mx = max(mx, s.y);
        }
        b |= mn < mx;
    }
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (__ballot(b) != 0 && lane_id() == 0) atomicOr(&any, 1);
    __syncthreads();
    if (threadIdx.x == 0 && any) atomicOr(bad, 1u);
}

// Per level of a block of at most MS levels without level-major records (the dense small block):
// the smallest and largest cluster over its cells, register bins per lane, reduced over the wave
// by shuffles and over the workgroup in LDS; part[wg] holds each workgroup's (min, max) per level.
template <int MS>
__global__ __launch_bounds__(TPB) void k_small_code_span(const uint32_t* __restrict__ cell_codes, int nb, int b,
                                                         const uint2* __restrict__ span, size_t M,
                                                         uint2* __restrict__ part) {
    __shared__ uint2 red[TPB / WAVE][MS];
    uint32_t mn[MS], mx[MS];
#pragma unroll
    for (int k = 0; k < MS; k++) { mn[k] = 0xFFFFFFFFu; mx[k] = 0u; }
    for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < M; c += (size_t)gridDim.x * TPB) {
        const uint32_t lv = cell_codes[c * nb + b];
        const uint2 sp = span[c];
#pragma unroll
        for (int k = 0; k < MS; k++)
            if (lv == (uint32_t)k) { mn[k] = min(mn[k], sp.x); mx[k] = max(mx[k], sp.y); }
    }
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < MS; k++) {
        for (int o = 32; o > 0; o >>= 1) {
            mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], o, WAVE));
            mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], o, WAVE));
        }
        if (lane_id() == 0) red[w][k] = make_uint2(mn[k], mx[k]);
    }
    __syncthreads();
    if (threadIdx.x < MS) {
        uint32_t a = 0xFFFFFFFFu, z = 0u;
        for (int ww = 0; ww < TPB / WAVE; ww++) { a = min(a, red[ww][threadIdx.x].x); z = max(z, red[ww][threadIdx.x].y); }
        part[(size_t)blockIdx.x * MS + threadIdx.x] = make_uint2(a, z);
    }
}

// *bad |= some level (columns of the nparts x MS partials) spans clusters: workgroup k reduces
// level k's column over the partials.
template <int MS>
__global__ __launch_bounds__(TPB) void k_small_span_check(const uint2* __restrict__ part, unsigned nparts,
                                                          uint32_t* __restrict__ bad) {
    __shared__ uint2 red[TPB / WAVE];
    const unsigned k = blockIdx.x;
    uint32_t a = 0xFFFFFFFFu, z = 0u;
    for (unsigned q = threadIdx.x; q < nparts; q += TPB) {
        const uint2 v = part[(size_t)q * MS + k];
        a = min(a, v.x);
        z = max(z, v.y);
    }
    for (int o = 32; o > 0; o >>= 1) {
        a = min(a, (uint32_t)__shfl_xor((int)a, o, WAVE));
        z = max(z, (uint32_t)__shfl_xor((int)z, o, WAVE));
    }
    if (lane_id() == 0) red[threadIdx.x / WAVE] = make_uint2(a, z);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < TPB / WAVE; w++) { a = min(a, red[w].x); z = max(z, red[w].y); }
        if (a != 0xFFFFFFFFu && a < z) atomicOr(bad, 1u);
    }
}

// ------------------------------------------------------------------ host: inference

// ols.rs:139-142
double eval_affine_t(double delta, const double* c) {
    const double v = c[2] + delta * (2.0 * c[3] + delta * c[4]);
    if (!std::isfinite(v) || v <= 1e-300) return NAN;
    return (c[0] + delta * c[1]) / std::sqrt(v);
}

// mod.rs:529-580 for one restriction
std::pair<double, uint64_t> compute_pvalue(double stat, const std::vector<double>& tb, int pval_type) {
    if (!std::isfinite(stat)) return {NAN, 0};
    uint64_t feas = 0, hi = 0, lo = 0;
    const double eps = 1e-12 * (1.0 + std::fabs(stat)), as = std::fabs(stat);
    for (double t : tb) {
        if (!std::isfinite(t)) continue;
        feas++;
        if (pval_type == FH_PVAL_TWO_TAILED) {
            if (std::fabs(t) > as + eps) hi++;
        } else {
            if (t > stat + eps) hi++;
            if (t < stat - eps) lo++;
        }
    }
    if (feas == 0) return {NAN, 0};
    const double f = (double)feas;
    switch (pval_type) {
        case FH_PVAL_TWO_TAILED:
        case FH_PVAL_GREATER: return {hi / f, feas};
        case FH_PVAL_LESS: return {lo / f, feas};
        default: return {std::min(2.0 * ((double)std::min(hi, lo) / f), 1.0), feas};
    }
}

// Standard normal quantile: Wichura (1988), algorithm AS 241 (PPND16).
double qnorm(double p) {
    if (std::isnan(p) || p < 0.0 || p > 1.0) return NAN;
    if (p == 0.0) return -INFINITY;
    if (p == 1.0) return INFINITY;
    const double q = p - 0.5;
    if (std::fabs(q) <= 0.425) {
        const double r = 0.180625 - q * q;
        const double num = (((((((2509.0809287301226727 * r + 33430.575583588128105) * r + 67265.770927008700853) * r +
                                45921.953931549871457) * r + 13731.693765509461125) * r + 1971.5909503065514427) * r +
                             133.14166789178437745) * r + 3.387132872796366608);
        const double den = (((((((5226.495278852545925 * r + 28729.085735721942674) * r + 39307.89580009271061) * r +
                                21213.794301586595867) * r + 5394.1960214247511077) * r + 687.1870074920579083) * r +
                             42.313330701600911252) * r + 1.0);
        return q * num / den;
    }
    double r = std::sqrt(-std::log(q < 0.0 ? p : 1.0 - p));
    double v;
    if (r <= 5.0) {
        r -= 1.6;
        const double num = (((((((7.7454501427834140764e-4 * r + 0.0227238449892691845833) * r + 0.24178072517745061177) * r +
                                1.27045825245236838258) * r + 3.64784832476320460504) * r + 5.7694972214606914055) * r +
                             4.6303378461565452959) * r + 1.42343711074968357734);
        const double den = (((((((1.05075007164441684324e-9 * r + 5.475938084995344946e-4) * r + 0.0151986665636164571966) * r +
                                0.14810397642748007459) * r + 0.68976733498510000455) * r + 1.6763848301838038494) * r +
                             2.05319162663775882187) * r + 1.0);
        v = num / den;
    } else {
        r -= 5.0;
        const double num = (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r + 0.0012426609473880784386) * r +
                                0.026532189526576123093) * r + 0.29656057182850489123) * r + 1.7848265399172913358) * r +
                             5.4637849111641143699) * r + 6.6579046435011037772);
        const double den = (((((((2.04426310338993978564e-15 * r + 1.4215117583164458887e-7) * r + 1.8463183175100546818e-5) * r +
                                7.868691311456132591e-4) * r + 0.0148753612908506148525) * r + 0.13692988092273580531) * r +
                             0.59983220655588793769) * r + 1.0);
        v = num / den;
    }
    return q < 0.0 ? -v : v;
}

// The test-inverted confidence interval (mod.rs:637-749): each endpoint is where the bootstrap
// p-value of H0: R beta = x falls below alpha moving away from the point estimate. The settings
// that decide the result follow the reference: the first step hw = max(1.5 z se, 1e-6 (1 + |point|))
// with z = max(|qnorm(1 - alpha/2)|, 1.5), up to 16 doublings of the step to find a value below
// alpha, and a search that ends once the bracket is x_tol = tol (1 + |point|) wide or a bracket end
// lies within p_tol = 1.0001 (1 or 2 for equal tails) / (feasible draws) of alpha, i.e. within one
// draw of it. p(x) is a step function, so every point meeting that rule is a valid endpoint.
//
// The root finder is Chandrupatla's (T. R. Chandrupatla, "A new hybrid quadratic/bisection
// algorithm for finding the zero of a nonlinear function without using derivatives", Advances in
// Engineering Software 28 (1997) 145-149): it keeps a bracket [a, b] with the newest point a,
// the previous point c, and takes inverse quadratic interpolation through (a, b, c) when the
// paper's validity test on xi = (a - b)/(c - b) and phi = (fa - fb)/(fc - fb) holds, else
// bisects.
struct CiRoot {
    double x_tol, f_tol;
    uint64_t maxiter;

    // f(x_in) >= 0 > f(x_out); returns the end of the final bracket with the smaller |f|.
    template <class F>
    double solve(F&& f, double x_in, double f_in, double x_out, double f_out) const {
        double a = x_in, fa = f_in;  // newest point
        double b = x_out, fb = f_out;  // other end of the bracket (opposite sign)
        double c = b, fc = fb;         // point dropped last
        double t = 0.5;
        auto done = [&] { return std::fabs(b - a) <= x_tol || std::min(std::fabs(fa), std::fabs(fb)) <= f_tol; };
        for (uint64_t it = 0; it < maxiter && !done(); it++) {
            const double width = std::fabs(b - a);
            const double tlim = std::min(0.5, 0.5 * x_tol / width);
            t = std::min(std::max(t, tlim), 1.0 - tlim);
            double xt = a + t * (b - a), ft = f(xt);
            if (!std::isfinite(ft)) {  // an infeasible trial: halve the bracket instead
                xt = 0.5 * (a + b);
                ft = f(xt);
                if (!std::isfinite(ft)) break;
            }
            if ((ft >= 0.0) == (fa >= 0.0)) {
                c = a;
                fc = fa;
            } else {
                c = b;
                fc = fb;
                b = a;
                fb = fa;
            }
            a = xt;
            fa = ft;
            t = 0.5;
            const double dcb = c - b, dfcb = fc - fb;
            if (dcb != 0.0 && dfcb != 0.0 && fb != fa && fc != fa && b != a) {
                const double xi = (a - b) / dcb, phi = (fa - fb) / dfcb;
                if (phi * phi < xi && (1.0 - phi) * (1.0 - phi) < 1.0 - xi)
                    t = fa / (fb - fa) * fc / (fb - fc) + (c - a) / (b - a) * fa / (fc - fa) * fb / (fc - fb);
            }
        }
        return std::fabs(fa) <= std::fabs(fb) ? a : b;
    }
};

template <class F>
std::pair<double, double> invert_ci_search(double point, double se, double alpha, int pval_type, double tol,
                                           uint64_t maxiter, uint64_t reps_feas, F&& eval_p) {
    if (!std::isfinite(point) || !std::isfinite(se) || se <= 0.0 || reps_feas == 0) return {NAN, NAN};
    const double z = std::max(std::fabs(qnorm(1.0 - alpha / 2.0)), 1.5);
    const double hw = std::max(1.5 * z * se, 1e-6 * (1.0 + std::fabs(point)));
    const double tails = pval_type == FH_PVAL_EQUAL_TAILED ? 2.0 : 1.0;
    const CiRoot root{tol * (1.0 + std::fabs(point)), 1.0001 * tails / (double)reps_feas, maxiter};
    auto f = [&](double x) { return eval_p(x) - alpha; };
    auto endpoint = [&](double dir) -> double {
        // the last point with p >= alpha (the estimate itself counts as one) and the first below
        double inside = point, f_inside = f(point);
        if (!std::isfinite(f_inside) || f_inside < 0.0) f_inside = 1.0 - alpha;
        double step = hw;
        for (int k = 0; k < 16; k++, step *= 2.0) {
            const double x = point + dir * step, fx = f(x);
            if (!std::isfinite(fx)) continue;
            if (fx < 0.0) return root.solve(f, inside, f_inside, x, fx);
            inside = x;
            f_inside = fx;
        }
        return dir * INFINITY;  // p stays above alpha: the interval is open on this side
    };
    const double lo = pval_type == FH_PVAL_LESS ? -INFINITY : endpoint(-1.0);
    const double hi = pval_type == FH_PVAL_GREATER ? INFINITY : endpoint(1.0);
    return {lo, hi};
}

// rows_ext[i] = [rows[i] (ld columns), D_nn[i] (m columns)]: the indicator of row i's level in
// each non-nested term q at column ld + off[q] + code (the level of kept row pi[i]).
struct NnTerms {
    const uint32_t* codes[8];
    uint32_t off[8];
    int nq;
};
__global__ void k_nn_rows(const double* __restrict__ rows, size_t n, int ld, const uint32_t* __restrict__ pi,
                          NnTerms t, int m, double* __restrict__ out) {
    const int le = ld + m;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double* o = out + i * le;
        for (int c = 0; c < ld; c++) o[c] = rows[i * ld + c];
        for (int c = 0; c < m; c++) o[ld + c] = 0.0;
        const uint32_t k = pi[i];
        for (int q = 0; q < t.nq; q++) o[ld + t.off[q] + t.codes[q][k]] = 1.0;
    }
}

// The same indicator block alone, in another absorber's row order pi2 (n x m).
__global__ void k_nn_dummies(size_t n, const uint32_t* __restrict__ pi2, NnTerms t, int m, double* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double* o = out + i * m;
        for (int c = 0; c < m; c++) o[c] = 0.0;
        const uint32_t k = pi2[i];
        for (int q = 0; q < t.nq; q++) o[t.off[q] + t.codes[q][k]] = 1.0;
    }
}

__global__ void k_to_kept(const double* __restrict__ v_rows, const uint32_t* __restrict__ pi, size_t n,
                          double* __restrict__ v_kept) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v_kept[pi[i]] = v_rows[i];
}

// Per row (row order): cluster code, z = x~'C and the weighted targets w u_0 = w e, w u_1 = -w x~'M_R.
__global__ void k_pc_rows(const double* __restrict__ rows, int R, const int* __restrict__ kept, int mk,
                          const double* __restrict__ C, const double* __restrict__ MR, const double* __restrict__ e,
                          size_t es, const double* __restrict__ w, const uint32_t* __restrict__ cl,
                          const uint32_t* __restrict__ pi, size_t n, uint32_t* __restrict__ clr,
                          double* __restrict__ z, double2* __restrict__ wu) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double zi = 0.0, u1 = 0.0;
        for (int a = 0; a < mk; a++) {
            const double x = rows[i * R + 1 + kept[a]];
            zi += x * C[a];
            u1 -= x * MR[a];
        }
        const double wi = w ? w[i] : 1.0;
        clr[i] = cl[pi[i]];
        z[i] = zi;
        wu[i] = make_double2(wi * e[i * es], wi * u1);
    }
}

// H[i][j] = z_i when row i lies in cluster c0 + j, else 0 (n x nb, j < nb).
__global__ void k_pc_cols(const uint32_t* __restrict__ clr, const double* __restrict__ z, size_t n, uint32_t c0,
                          int nb, double* __restrict__ H) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * nb; e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e / nb;
        const uint32_t j = (uint32_t)(e % nb);
        H[e] = clr[i] == c0 + j ? z[i] : 0.0;
    }
}

// P[g][2 + t nc + c0 + j] = sum over the rows i of cluster g of (w u_t)_i H[i][j]: one wave per
// cluster, lanes strided over its runs, each summing whole runs in row order, then a fixed-order
// reduction over the wave.
__global__ __launch_bounds__(TPB) void k_pc_sums(const uint32_t* __restrict__ off, const uint32_t* __restrict__ order,
                                                 const uint32_t* __restrict__ run_off, size_t G,
                                                 const double2* __restrict__ wu, const double* __restrict__ H, int nb,
                                                 uint32_t c0, size_t nc, int NC, double* __restrict__ P) {
    const int lane = lane_id();
    for (size_t g = (size_t)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE; g < G;
         g += (size_t)gridDim.x * (TPB / WAVE)) {
        double s0[16], s1[16];
#pragma unroll
        for (int j = 0; j < 16; j++) s0[j] = s1[j] = 0.0;
        for (uint32_t k = off[g] + lane; k < off[g + 1]; k += WAVE) {
            const uint32_t run = order[k];
            for (uint32_t i = run_off[run]; i < run_off[run + 1]; i++) {
                const double2 u = wu[i];
#pragma unroll
                for (int j = 0; j < 16; j++)
                    if (j < nb) {
                        const double h = H[(size_t)i * nb + j];
                        s0[j] += u.x * h;
                        s1[j] += u.y * h;
                    }
            }
        }
#pragma unroll
        for (int j = 0; j < 16; j++)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                s0[j] += __shfl_xor(s0[j], o, WAVE);
                s1[j] += __shfl_xor(s1[j], o, WAVE);
            }
        if (lane == 0) {
            double* p = P + g * NC + 2 + c0;
#pragma unroll
            for (int j = 0; j < 16; j++)
                if (j < nb) {
                    p[j] = s0[j];
                    p[nc + j] = s1[j];
                }
        }
    }
}

// Phi[g][t] = bread S_t[g] (mk each) and P[g][t] = C.S_t[g] (the draw's numerators).
__global__ void k_pc_phi(const double* __restrict__ S, size_t G, int mk, const double* __restrict__ bread,
                         const double* __restrict__ C, int NC, double* __restrict__ Phi, double* __restrict__ P) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < G * 2; e += (size_t)gridDim.x * blockDim.x) {
        const size_t g = e / 2;
        const int t = (int)(e % 2);
        const double* St = S + (g * 3 + t) * mk;
        double num = 0.0;
        for (int a = 0; a < mk; a++) {
            double f = 0.0;
            for (int b = 0; b < mk; b++) f += bread[a * mk + b] * St[b];
            Phi[e * mk + a] = f;
            num += C[a] * St[a];
        }
        P[g * NC + t] = num;
    }
}

// P[g][2 + t nc + c] -= B_c . Phi_t[g] (B_c = S[c][2], the cluster's sum of w z x~): the regressors'
// part of the projection.
__global__ void k_pc_fix(const double* __restrict__ S, const double* __restrict__ Phi, size_t G, int mk, int NC,
                         double* __restrict__ P) {
    const size_t nc = G;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < G * 2 * nc; e += (size_t)gridDim.x * blockDim.x) {
        const size_t g = e / (2 * nc), r = e % (2 * nc);
        const int t = (int)(r / nc);
        const size_t c = r % nc;
        const double* Bc = S + (c * 3 + 2) * mk;
        const double* ph = Phi + (g * 2 + t) * mk;
        double d = 0.0;
        for (int a = 0; a < mk; a++) d += Bc[a] * ph[a];
        P[g * NC + 2 + r] -= d;
    }
}

// cmax[c] = max over g of |P[g][c]| (one thread per column, G rows).
__global__ void k_pc_cmax(const double* __restrict__ P, size_t G, int NC, double* __restrict__ cmax) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < NC; c += gridDim.x * blockDim.x) {
        double m = 0.0;
        for (size_t g = 0; g < G; g++) m = fmax(m, fabs(P[g * NC + c]));
        cmax[c] = m;
    }
}

}  // namespace

// Per absorbed term (absorber term order): 1 when every level lies in one cluster
// (dof.rs nested_within, by value).
std::vector<int> terms_nested_in(Absorber& a, const uint32_t* cl) {
    hipStream_t st = a.st;
    std::vector<int> nested(a.nterms, 1);
    if (a.M == 0) return nested;
    const size_t nflag = a.nb + 1;
    DBuf<uint2> span(a.M);
    DBuf<uint32_t> bad(nflag);
    FH_HIP(hipMemsetAsync(bad.p, 0, bad.bytes(), st));
    hipLaunchKernelGGL(k_wb_cell_span, dim3(grid_for(a.M, TPB, 8192)), dim3(TPB), 0, st,
                       a.trivial() ? nullptr : a.cell_obs_off.p, a.pi.p, cl, (size_t)a.M, span.p);
    if (a.fe1 >= 0)
        hipLaunchKernelGGL(k_wb_group_span, dim3(grid_for(a.g1, TPB, 4096)), dim3(TPB), 0, st, a.seg_off.p,
                           (const uint32_t*)nullptr, span.p, (size_t)a.g1, (uint2*)nullptr, bad.p);
    DBuf<uint2> gspan;
    for (size_t b = 0; b < a.nb; b++) {
        if (a.boff[b].n == 0 && a.blocks[b].g <= 32) {  // the dense small block: no level records
            const unsigned np = grid_for(a.M, TPB * 16, 1024);
            DBuf<uint2> part((size_t)np * 32);
            hipLaunchKernelGGL(k_small_code_span<32>, dim3(np), dim3(TPB), 0, st, a.cell_codes.p, (int)a.nb, (int)b,
                               span.p, (size_t)a.M, part.p);
            hipLaunchKernelGGL(k_small_span_check<32>, dim3(a.blocks[b].g), dim3(TPB), 0, st, part.p, np,
                               bad.p + 1 + b);
            continue;
        }
        ensure_block_transpose(a, (int)b);
        const size_t items = (size_t)a.nbk[b] * a.blocks[b].g;
        if (a.nbk[b] > 1) gspan.reserve(items);
        hipLaunchKernelGGL(k_wb_group_span, dim3(grid_for(items, TPB, 4096)), dim3(TPB), 0, st, a.boff[b].p,
                           a.bperm[b].p, span.p, items, a.nbk[b] > 1 ? gspan.p : nullptr, bad.p + 1 + b);
        if (a.nbk[b] > 1)
            hipLaunchKernelGGL(k_wb_level_span, dim3(grid_for(a.blocks[b].g, TPB, 4096)), dim3(TPB), 0, st, gspan.p,
                               a.nbk[b], (size_t)a.blocks[b].g, bad.p + 1 + b);
    }
    FH_CHECK_LAUNCH();
    std::vector<uint32_t> hb(nflag);
    FH_HIP(hipMemcpyAsync(hb.data(), bad.p, nflag * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    if (a.fe1 >= 0 && hb[0]) nested[a.fe1] = 0;
    for (size_t b = 0; b < a.nb; b++)
        if (hb[1 + b]) nested[a.blocks[b].term] = 0;
    return nested;
}

void wb_constraint(const std::vector<double>& bread, int mk, const double* rk, std::vector<double>& C,
                   std::vector<double>& MR) {
    // ConstraintProj (ols.rs:76-121) for one restriction
    C.assign(mk, 0.0);
    MR.assign(mk, 0.0);
    for (int i = 0; i < mk; i++)
        for (int j = 0; j < mk; j++) C[i] += bread[(size_t)i * mk + j] * rk[j];
    double v0 = 0.0;
    for (int j = 0; j < mk; j++) v0 += rk[j] * C[j];
    const double v0i = std::fabs(v0) > 1e-12 * std::max(1.0, std::fabs(v0)) ? 1.0 / v0 : 0.0;
    for (int i = 0; i < mk; i++) MR[i] = C[i] * v0i;
}

void wb_cluster_sums(Absorber& a, const double* rows, int R, const std::vector<int>& kept, const std::vector<double>& C,
                     const std::vector<double>& MR, const double* e, size_t es, const double* w, const uint32_t* cl,
                     uint64_t G, DBuf<double>& S) {
    hipStream_t st = a.st;
    const int mk = (int)kept.size();
    require(mk >= 1 && mk + 2 <= KMAX, "the bootstrap supports 1..19 kept regressors");
    DBuf<int> dkept(mk);
    DBuf<double> dC(mk), dMR(mk);
    FH_HIP(hipMemcpyAsync(dkept.p, kept.data(), mk * sizeof(int), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(dC.p, C.data(), mk * sizeof(double), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(dMR.p, MR.data(), mk * sizeof(double), hipMemcpyHostToDevice, st));
    ClusterRuns local;
    const ClusterRuns& cr = runs_for(a, cl, G, local);
    S.alloc(G * 3 * mk);
    const uint32_t L = lanes_per_cluster(cr);
    const unsigned cg = grid_for(G, (TPB / WAVE) * (WAVE / L), 1u << 16);
    auto launch = [&](auto mkc) {
        hipLaunchKernelGGL((k_wb_clsums<decltype(mkc)::value>), dim3(cg), dim3(TPB), 0, st, cr.off.p, cr.order.p,
                           cr.run_off.p, (size_t)G, rows, R, dkept.p, mk, dC.p, dMR.p, e, es, w, (int)L, S.p, -1);
    };
    // whole-row loads when every regressor is kept and the stride is a padded width
    bool ident = 1 + mk <= R && (reinterpret_cast<uintptr_t>(rows) & 15) == 0;
    for (int q = 0; q < mk; q++) ident &= kept[q] == q;
    const int ecol = es == (size_t)R && e >= rows && e < rows + R ? (int)(e - rows) : -1;
    auto launch_rows = [&](auto mkc, auto ldc) {
        hipLaunchKernelGGL((k_wb_clsums<decltype(mkc)::value, decltype(ldc)::value>), dim3(cg), dim3(TPB), 0, st,
                           cr.off.p, cr.order.p, cr.run_off.p, (size_t)G, rows, R, dkept.p, mk, dC.p, dMR.p, e, es, w,
                           (int)L, S.p, ecol);
    };
    using I4 = std::integral_constant<int, 4>;
    using I8 = std::integral_constant<int, 8>;
    using I12 = std::integral_constant<int, 12>;
    using I16 = std::integral_constant<int, 16>;
    using I20 = std::integral_constant<int, 20>;
    if (ident && mk <= 4 && R == 4) launch_rows(I4{}, I4{});
    else if (ident && mk <= 4 && R == 8) launch_rows(I4{}, I8{});
    else if (ident && mk <= 12 && R == 8) launch_rows(I12{}, I8{});
    else if (ident && mk <= 12 && R == 12) launch_rows(I12{}, I12{});
    else if (ident && mk <= 12 && R == 16) launch_rows(I12{}, I16{});
    else if (ident && mk <= 20 && R == 16) launch_rows(I20{}, I16{});
    else if (mk <= 4) launch(I4{});
    else if (mk <= 12) launch(I12{});
    else launch(I20{});
    FH_CHECK_LAUNCH();
}

// The inference from the draws' affine coefficients (reps x 5, host): the observed statistic,
// the bootstrap t's, the p-value and the test-inverted interval (ols.rs:553-649).
void wb_infer(const std::vector<double>& hc, uint64_t reps, bool full, size_t G, double point, double v_obs,
              double ssc, const fh_wboot_spec& spec, fh_wboot_result& res, double* t_boot_host,
              std::chrono::steady_clock::time_point t0) {
    const double num_obs = point - spec.r;
    const double se = (std::isfinite(v_obs) && v_obs > 1e-300) ? std::sqrt(v_obs) : NAN;
    const double stat = std::isfinite(se) ? num_obs / se : NAN;
    const double delta = spec.r - point;
    std::vector<double> tb(reps);
    for (uint64_t b = 0; b < reps; b++) tb[b] = eval_affine_t(delta, &hc[b * 5]);
    auto pv = compute_pvalue(stat, tb, spec.pval_type);
    res.stat = stat;
    res.p_val = pv.first;
    res.reps_feasible = pv.second;
    res.se = se;
    res.point = point;
    res.ssc = ssc;
    res.reps = reps;
    res.n_clusters = G;
    res.full_enumeration = full;
    res.ci_lo = NAN;
    res.ci_hi = NAN;
    if (spec.conf_int) {
        std::vector<double> trial(reps);
        auto ci = invert_ci_search(point, se, 1.0 - spec.level, spec.pval_type, spec.tol, spec.maxiter, pv.second,
                                   [&](double r_trial) {
                                       const double d = r_trial - point, st_trial = (point - r_trial) / se;
                                       for (uint64_t b = 0; b < reps; b++) trial[b] = eval_affine_t(d, &hc[b * 5]);
                                       return compute_pvalue(st_trial, trial, spec.pval_type).first;
                                   });
        res.ci_lo = ci.first;
        res.ci_hi = ci.second;
    }
    if (t_boot_host) std::memcpy(t_boot_host, tb.data(), reps * sizeof(double));
    res.seconds_bootstrap = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The draws and the inference shared by WCR11 and the MNW variants: the per-cluster rows P
// (device, G x NC with NC = 2 + 4 mk: [num_0, num_1, Phi_0, Phi_1, A_0 B, A_1 B] and their column
// scales), AA (the per-cluster scalars' Gram, 2 x 2) and BB (sum B B', mk x mk) give every draw's
// affine coefficients (num0, num1, v00, v01, v11) of t(delta) by the int8 MFMA contraction;
// then the observed statistic (point, v_obs), the p-value and the test-inverted interval.
// cells > 0: the per-cluster form instead, P = [num_0, num_1, q_0 (cells), q_1 (cells)] with the cell
// scores of a draw V_b . q_t[c] (AA and BB unused).
void wb_draws(Absorber& a, const double* P, size_t G, int NC, int mk, const std::vector<double>& scale,
              const std::vector<double>& inv_scale, const std::vector<double>& AA, const std::vector<double>& BB,
              double ssc, double point, double v_obs, const fh_wboot_spec& spec, fh_wboot_result& res,
              double* t_boot_host, std::chrono::steady_clock::time_point t0, size_t cells = 0) {
    hipStream_t st = a.st;
    // draws
    uint64_t reps = spec.reps;
    bool full = false;
    if (G <= 20 && (1ull << G) <= reps) {  // weights.rs:39-61
        reps = 1ull << G;
        full = true;
    }
    const size_t Gpad = (G + 63) / 64 * 64;
    const int ncol = LIMBS * NC, ntiles = (ncol + 15) / 16, ncolp = ntiles * 16;
    DBuf<int8_t> Pq((size_t)ncolp * Gpad);
    FH_HIP(hipMemsetAsync(Pq.p, 0, Pq.bytes(), st));
    DBuf<double> dinv(NC), dscale(NC), dAA(4), dBB((size_t)mk * mk);
    FH_HIP(hipMemcpyAsync(dinv.p, inv_scale.data(), NC * sizeof(double), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(dscale.p, scale.data(), NC * sizeof(double), hipMemcpyHostToDevice, st));
    if (cells == 0) {
        FH_HIP(hipMemcpyAsync(dAA.p, AA.data(), 4 * sizeof(double), hipMemcpyHostToDevice, st));
        FH_HIP(hipMemcpyAsync(dBB.p, BB.data(), (size_t)mk * mk * sizeof(double), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_wb_quant, dim3(grid_for(G * NC, TPB, 8192)), dim3(TPB), 0, st, P, G, NC, dinv.p, Gpad, Pq.p);
    FH_CHECK_LAUNCH();
    const uint32_t nsteps = (uint32_t)(Gpad / 64);
    const size_t bblocks = (reps + WB_ROWS - 1) / WB_ROWS, Bpad = bblocks * WB_ROWS;
    const int cgroups = (ntiles + WB_NTC - 1) / WB_NTC;
    // split the G range so the grid fills the chip (~8 workgroups per CU; 4 per CU: 0.67 ms at C3)
    uint32_t nsplit = (uint32_t)std::max<size_t>(1, std::min<size_t>(nsteps, (2048 + bblocks * cgroups - 1) / (bblocks * cgroups)));
    const uint32_t per = (nsteps + nsplit - 1) / nsplit;
    nsplit = (nsteps + per - 1) / per;
    DBuf<int32_t> part((size_t)nsplit * Bpad * ncolp);
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    FH_HIP(hipEventRecord(e0, st));
    hipLaunchKernelGGL(k_wb_gemm, dim3((unsigned)bblocks, nsplit, cgroups), dim3(TPB), 0, st, Pq.p, Gpad, ntiles,
                       (uint64_t)spec.seed, (int)full, per, nsteps, part.p, Bpad, ncolp);
    FH_CHECK_LAUNCH();
    FH_HIP(hipEventRecord(e1, st));
    DBuf<double> coefs(reps * 5);
    if (cells > 0)
        hipLaunchKernelGGL(k_wb_post_cells, dim3(grid_for(reps, TPB / WAVE, 8192)), dim3(TPB), 0, st, part.p,
                           (int)nsplit, Bpad, ncolp, NC, (int)cells, reps, dscale.p, ssc, coefs.p);
    else
        hipLaunchKernelGGL(k_wb_post, dim3(grid_for(reps, TPB / WAVE, 8192)), dim3(TPB), 0, st, part.p, (int)nsplit,
                           Bpad, ncolp, NC, mk, reps, dscale.p, dAA.p, dBB.p, ssc, coefs.p);
    FH_CHECK_LAUNCH();
    std::vector<double> hc(reps * 5);
    FH_HIP(hipMemcpyAsync(hc.data(), coefs.p, hc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    float gms = 0.f;
    FH_HIP(hipEventElapsedTime(&gms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    res.contraction_ms = gms;
    // int8 multiply-adds (2 ops each) of the draws x clusters x limb-columns product, padding included
    res.contraction_ops = 2.0 * (double)Bpad * (double)Gpad * (double)ncolp;
    wb_infer(hc, reps, full, G, point, v_obs, ssc, spec, res, t_boot_host, t0);
}

void ols_wildboot(Absorber& a, const WbootInput& in, const fh_wboot_spec& spec, const double* rk, fh_wboot_result& res,
                  double* t_boot_host) {
    hipStream_t st = a.st;
    const size_t G = in.G;
    const int mk = (int)in.kept.size();
    require(mk >= 1 && mk + 2 <= KMAX, "the bootstrap supports 1..19 kept regressors");
    require(G >= 2 && G < (1ull << 24), "the bootstrap needs 2 .. 2^24 clusters");
    require(spec.reps >= 1, "wildboot: reps must be at least 1");
    require(spec.level > 0.0 && spec.level < 1.0, "wildboot: confidence level must be in (0, 1)");
    require(spec.tol > 0.0 && std::isfinite(spec.tol) && spec.maxiter > 0, "wildboot: tol and maxiter must be positive");
    require(spec.pval_type >= 0 && spec.pval_type <= 3, "wildboot: unknown p-value type");
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> C, MR;
    wb_constraint(in.bread, mk, rk, C, MR);
    DBuf<double> dC(mk), dbread((size_t)mk * mk);
    FH_HIP(hipMemcpyAsync(dC.p, C.data(), mk * sizeof(double), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(dbread.p, in.bread.data(), (size_t)mk * mk * sizeof(double), hipMemcpyHostToDevice, st));
    // per-cluster sums (formed by the fit's CRV1 step when it ran for this bootstrap)
    DBuf<double> Sown;
    const double* Sp = in.S;
    if (!Sp) {
        wb_cluster_sums(a, in.rows, in.R, in.kept, C, MR, in.e, in.es, in.w, in.cl, G, Sown);
        Sp = Sown.p;
    }
    const int NC = 2 + 4 * mk, QC = 2 + mk;
    DBuf<double> P(G * NC), Q(G * QC);
    const unsigned pg = grid_for(G, TPB, 512);
    DBuf<double> cmax((size_t)pg * NC);
    hipLaunchKernelGGL(k_wb_prow, dim3(pg), dim3(TPB), 0, st, Sp, G, mk, dbread.p, dC.p, P.p, Q.p, cmax.p);
    FH_CHECK_LAUNCH();
    std::vector<double> QG((size_t)QC * QC);
    gram_cols(a, Q.p, G, QC, 0, QC, nullptr, nullptr, QG.data());  // synchronizes
    std::vector<double> hmax((size_t)pg * NC);
    FH_HIP(hipMemcpyAsync(hmax.data(), cmax.p, hmax.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    std::vector<double> scale(NC), inv_scale(NC);
    for (int c = 0; c < NC; c++) {
        double m = 0.0;
        for (unsigned q = 0; q < pg; q++) m = std::max(m, hmax[(size_t)q * NC + c]);
        scale[c] = m > 0.0 ? 4.0 * m : 1.0;  // |P / scale| <= 1/4: 7 balanced limbs cover it
        inv_scale[c] = std::ldexp(1.0 / scale[c], 56);
    }
    std::vector<double> AA = {QG[0], QG[1], QG[QC], QG[QC + 1]}, BB((size_t)mk * mk);
    for (int i = 0; i < mk; i++)
        for (int j = 0; j < mk; j++) BB[(size_t)i * mk + j] = QG[(size_t)(2 + i) * QC + 2 + j];
    // observed statistic: q_obs = A_0 (the cluster scores of the unrestricted residuals)
    double point = 0.0;
    for (int j = 0; j < mk; j++) point += rk[j] * in.beta[j];
    const double v_obs = in.ssc * AA[0];
    wb_draws(a, P.p, G, NC, mk, scale, inv_scale, AA, BB, in.ssc, point, v_obs, spec, res, t_boot_host, t0);
}

void wb_path1(Absorber& a, const std::vector<int>& nested, uint64_t m_nn, WbootInput& in, std::vector<double>& rk,
              WbExt& ext) {
    // Fixed effects not nested in the clusters (Path 1 of the reference, wildboot/mod.rs:959-1180;
    // ols.rs:413-448). With D_nn the non-nested terms' dummies, D~ = M_nest D_nn their residuals
    // on the nested terms and G_nn = D~' W D~, the FE-projected bootstrap residual of a draw is
    // u* - D~ G_nn^+ D_nn' W u* (u* = V_g u~ by cluster), so the cell score of cluster c becomes
    //   q[c, t] = V_c A_t[c] - B_c.(V Phi_t) - f_c.(V psi_t),
    // f_c = sum_{i in c} w_i z_i D_nn[i] (= the same sum over D~: z is orthogonal to every FE and
    // the nested terms' levels lie within clusters), psi_t[g] = G_nn^+ F_t[g],
    // F_t[g] = sum_{i in g} w_i u_t,i D_nn[i]. That is the nested form with m_nn pseudo-regressors
    // appended: S_t gets F_t, B gets f, the bread becomes blockdiag(bread, G_nn^+) and C, M_R,
    // beta and R get zeros -- one more cluster-sum pass over [rows, D_nn] and the same contraction.
    hipStream_t st = a.st;
    const size_t n = a.n;
    const int mk = (int)in.kept.size(), m = (int)m_nn;
    require(mk + m + 2 <= KMAX,
            "the bootstrap's non-nested projection supports kept regressors + non-nested levels <= 19");
    NnTerms nt{};
    uint32_t off = 0;
    for (size_t q = 0; q < a.nterms; q++) {
        if (nested[q]) continue;
        require(nt.nq < 8, "at most 8 non-nested terms");
        require(a.codes_kept.size() == a.nterms && a.codes_kept[q].p, "the absorber has no term codes", FH_INTERNAL);
        nt.codes[nt.nq] = a.codes_kept[q].p;
        nt.off[nt.nq++] = off;
        off += (uint32_t)a.term_g[q];
    }
    // G_nn = D~' W D~ (weights: the absorber's, row order)
    std::vector<double> Gnn((size_t)m * m);
    bool any_nested = false;
    for (size_t q = 0; q < a.nterms; q++) any_nested |= nested[q] != 0;
    const int ld = in.R, le = ld + m;
    ext.rows.alloc(n * le);
    hipLaunchKernelGGL(k_nn_rows, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, in.rows, n, ld, a.pi.p, nt, m,
                       ext.rows.p);
    FH_CHECK_LAUNCH();
    if (!any_nested) {
        gram_cols(a, ext.rows.p, n, le, ld, m, in.w, nullptr, Gnn.data());
    } else {
        // residualize D_nn on the nested terms through an absorber of those terms alone
        std::vector<fh_term> tt;
        for (size_t q = 0; q < a.nterms; q++)
            if (nested[q]) tt.push_back(fh_term{a.codes_kept[q].p, a.term_g[q], 1});
        DBuf<double> wk;
        if (in.w) {
            wk.alloc(n);
            hipLaunchKernelGGL(k_to_kept, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, in.w, a.pi.p, n, wk.p);
            FH_CHECK_LAUNCH();
        }
        stream_wait(st);
        fh_solver_config cfg;
        fh_config_default(&cfg);
        fh_absorber* tmp = nullptr;
        if (fh_absorber_create_dev(tt.data(), tt.size(), in.w ? wk.p : nullptr, n, &cfg, nullptr, &tmp) != FH_OK)
            throw Error(FH_INTERNAL, std::string("non-nested projection: ") + fh_last_error());
        struct Tmp {
            fh_absorber* p;
            ~Tmp() { fh_absorber_destroy(p); }
        } hold{tmp};
        DBuf<double> dn(n * m);
        hipLaunchKernelGGL(k_nn_dummies, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, n, tmp->a.pi.p, nt, m, dn.p);
        FH_CHECK_LAUNCH();
        stream_wait(st);
        fh_solve_opts o{1e-12, 100000, 0, FH_PLAIN_FP64};
        fh_solve_info info{};
        if (fh_absorber_residualize_dev(tmp, dn.p, (size_t)m, &o, nullptr, nullptr, &info, nullptr) != FH_OK)
            throw Error(FH_INTERNAL, std::string("non-nested projection: ") + fh_last_error());
        {
            AbsScope sc(tmp->a);
            gram_cols(tmp->a, dn.p, n, m, 0, m, tmp->a.weighted ? tmp->a.w_rows.p : nullptr, nullptr, Gnn.data());
        }
    }
    // G_nn^+ (wildboot/mod.rs:434-470: eigenvalues <= 1e-10 of the largest dropped)
    for (int i = 0; i < m; i++)
        for (int j = 0; j < i; j++) Gnn[(size_t)i * m + j] = Gnn[(size_t)j * m + i] = 0.5 * (Gnn[(size_t)i * m + j] + Gnn[(size_t)j * m + i]);
    std::vector<double> lam, U, Gp((size_t)m * m, 0.0);
    sym_eigen(Gnn, m, lam, U);
    double lmax = 0.0;
    for (double l : lam) lmax = std::max(lmax, std::fabs(l));
    for (int e = 0; e < m; e++) {
        if (!(lam[e] > 1e-10 * lmax) || lmax == 0.0) continue;
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) Gp[(size_t)i * m + j] += U[(size_t)i * m + e] * U[(size_t)j * m + e] / lam[e];
    }
    // the extended problem
    const int mk2 = mk + m;
    std::vector<double> bread2((size_t)mk2 * mk2, 0.0);
    for (int i = 0; i < mk; i++)
        for (int j = 0; j < mk; j++) bread2[(size_t)i * mk2 + j] = in.bread[(size_t)i * mk + j];
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) bread2[(size_t)(mk + i) * mk2 + mk + j] = Gp[(size_t)i * m + j];
    std::vector<int> kept2(in.kept);
    for (int j = 0; j < m; j++) kept2.push_back(ld - 1 + j);
    rk.resize(mk2, 0.0);
    std::vector<double> C2, MR2;
    wb_constraint(bread2, mk2, rk.data(), C2, MR2);
    // the residual: its own array, or the same column of the extended rows
    const double* e2 = in.e;
    size_t es2 = in.es;
    if (in.es == (size_t)ld && in.e >= in.rows && in.e < in.rows + ld) {
        e2 = ext.rows.p + (in.e - in.rows);
        es2 = (size_t)le;
    }
    wb_cluster_sums(a, ext.rows.p, le, kept2, C2, MR2, e2, es2, in.w, in.cl, in.G, ext.S);
    in.rows = ext.rows.p;
    in.R = le;
    in.kept = kept2;
    in.bread = bread2;
    in.beta.resize(mk2, 0.0);
    in.e = e2;
    in.es = es2;
    in.S = ext.S.p;
}


void ols_wildboot_cells(Absorber& a, const WbootInput& in, const fh_wboot_spec& spec, const double* rk, double tol,
                        uint64_t maxiter, fh_wboot_result& res, double* t_boot_host) {
    // The exact projection of every fixed effect, one cluster at a time (Path 2 of the reference,
    // wildboot/mod.rs:1290-1363, ols.rs:456-548). With the FE projector P_D, X~ the residualized
    // kept regressors, z = X~ C and u_t the draws' targets, a draw's cell score of cluster c is
    //   q_t[c] = sum_g V_g M_t[c, g],
    //   M_t[c, g] = <1_c z, M_X~ M_D (1_g u_t)>_W = <M_D (1_c z), 1_g u_t>_W - B_c . Phi_t[g],
    // the second form by the W-symmetry of M_D and X~'W P_D = 0. So one residualization per
    // cluster (the column 1_c z) serves both targets, where the reference projects 1_g u_t for
    // every target and cluster. The G x (2 + 2G) matrix P = [C.S_0, C.S_1, M_0', M_1'] goes through
    // the same int8 contraction as the other forms; each draw then sums its cells' squares.
    hipStream_t st = a.st;
    const size_t G = in.G, n = a.n;
    const int mk = (int)in.kept.size();
    require(mk >= 1 && mk + 2 <= KMAX, "the bootstrap supports 1..19 kept regressors");
    require(G >= 2 && G <= PC_GMAX, "the per-cluster projection supports 2 .. 512 clusters");
    require(spec.reps >= 1, "wildboot: reps must be at least 1");
    require(spec.level > 0.0 && spec.level < 1.0, "wildboot: confidence level must be in (0, 1)");
    require(spec.tol > 0.0 && std::isfinite(spec.tol) && spec.maxiter > 0, "wildboot: tol and maxiter must be positive");
    require(spec.pval_type >= 0 && spec.pval_type <= 3, "wildboot: unknown p-value type");
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> C, MR;
    wb_constraint(in.bread, mk, rk, C, MR);
    DBuf<double> Sown;
    const double* Sp = in.S;
    if (!Sp) {
        wb_cluster_sums(a, in.rows, in.R, in.kept, C, MR, in.e, in.es, in.w, in.cl, G, Sown);
        Sp = Sown.p;
    }
    DBuf<int> dkept(mk);
    DBuf<double> dC(mk), dMR(mk), dbread((size_t)mk * mk);
    FH_HIP(hipMemcpyAsync(dkept.p, in.kept.data(), mk * sizeof(int), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(dC.p, C.data(), mk * sizeof(double), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(dMR.p, MR.data(), mk * sizeof(double), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(dbread.p, in.bread.data(), (size_t)mk * mk * sizeof(double), hipMemcpyHostToDevice, st));
    DBuf<uint32_t> clr(n);
    DBuf<double> z(n);
    DBuf<double2> wu(n);
    hipLaunchKernelGGL(k_pc_rows, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, in.rows, in.R, dkept.p, mk, dC.p,
                       dMR.p, in.e, in.es, in.w, in.cl, a.pi.p, n, clr.p, z.p, wu.p);
    FH_CHECK_LAUNCH();
    ClusterRuns local;
    const ClusterRuns& cr = runs_for(a, in.cl, G, local);
    const int NC = (int)(2 + 2 * G);
    DBuf<double> P(G * NC), H(n * 16);
    // the columns 1_c z, 16 clusters per residualization (FP64, the reference's tolerance cap)
    for (size_t c0 = 0; c0 < G; c0 += 16) {
        const int nb = (int)std::min<size_t>(16, G - c0);
        hipLaunchKernelGGL(k_pc_cols, dim3(grid_for(n * nb, TPB, 8192)), dim3(TPB), 0, st, clr.p, z.p, n,
                           (uint32_t)c0, nb, H.p);
        FH_CHECK_LAUNCH();
        residualize_any(a, H.p, (size_t)nb, std::min(tol, 1e-10), maxiter, nullptr, nullptr, FH_PLAIN_FP64, nullptr);
        hipLaunchKernelGGL(k_pc_sums, dim3(grid_for(G, TPB / WAVE, 4096)), dim3(TPB), 0, st, cr.off.p, cr.order.p,
                           cr.run_off.p, G, wu.p, H.p, nb, (uint32_t)c0, G, NC, P.p);
        FH_CHECK_LAUNCH();
    }
    DBuf<double> Phi(G * 2 * mk), cmax(NC);
    hipLaunchKernelGGL(k_pc_phi, dim3(grid_for(G * 2, TPB, 1024)), dim3(TPB), 0, st, Sp, G, mk, dbread.p, dC.p, NC,
                       Phi.p, P.p);
    hipLaunchKernelGGL(k_pc_fix, dim3(grid_for(G * 2 * G, TPB, 8192)), dim3(TPB), 0, st, Sp, Phi.p, G, mk, NC, P.p);
    hipLaunchKernelGGL(k_pc_cmax, dim3(grid_for(NC, TPB, 64)), dim3(TPB), 0, st, P.p, G, NC, cmax.p);
    FH_CHECK_LAUNCH();
    std::vector<double> hmax(NC), hS(G * 3 * mk);
    FH_HIP(hipMemcpyAsync(hmax.data(), cmax.p, NC * sizeof(double), hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(hS.data(), Sp, hS.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    std::vector<double> scale(NC), inv_scale(NC);
    for (int c = 0; c < NC; c++) {
        scale[c] = hmax[c] > 0.0 ? 4.0 * hmax[c] : 1.0;  // |P / scale| <= 1/4: 7 balanced limbs cover it
        inv_scale[c] = std::ldexp(1.0 / scale[c], 56);
    }
    // observed statistic: the cluster scores of the unrestricted residuals, A_0[g] = C.S_0[g]
    double aa = 0.0;
    for (size_t g = 0; g < G; g++) {
        double a0 = 0.0;
        for (int j = 0; j < mk; j++) a0 += C[j] * hS[(g * 3) * mk + j];
        aa += a0 * a0;
    }
    double point = 0.0;
    for (int j = 0; j < mk; j++) point += rk[j] * in.beta[j];
    const std::vector<double> none;
    wb_draws(a, P.p, G, NC, mk, scale, inv_scale, none, none, in.ssc, point, in.ssc * aa, spec, res, t_boot_host, t0,
             G);
}

// ------------------------------------------------------------------ Path 1 with many levels
// Path 1 (wildboot/mod.rs:959-1180, ols.rs:413-448) when the non-nested levels do not fit beside the
// regressors (any number of clusters). With p = kept + m_nn,
// Phi~_t[g] = [bread S_t[g], G_nn^- F_t[g]] and B~_g = [B_g, f_g], a draw's cell scores are
//   q_t[c] = V_c A_t[c] - B~_c . d_t,  d_t = sum_g V_g Phi~_t[g],
// so the draws are two plain FP64 products: Y = V [A_0, A_1, Phi~_0, Phi~_1] (reps x G times
// G x (2 + 2p)) and Z_t = d_t B~' (reps x p times p x G), by rocBLAS with atomics off (fixed
// order), the signs V materialized from the same SplitMix64 stream; a wave per draw then sums
// its cells' squares. Any generalized inverse G_nn^- gives the same projection (F_t and f lie in
// the range of G_nn); the grounded LDL' inverse is used (O(m^3), no eigen-decomposition).

namespace {

// WS[g][t][j], t = 0, 1: sum of w u_t, t = 2: sum of w z over the rows of cluster g whose level in
// a non-nested term is column j (= the term's offset + its code): one wave per (cluster, chunk of
// 64 columns), lanes over the columns, the cluster's runs in row order.
__global__ __launch_bounds__(TPB) void k_wide_sums(const uint32_t* __restrict__ off, const uint32_t* __restrict__ order,
                                                   const uint32_t* __restrict__ run_off, size_t G,
                                                   const double2* __restrict__ wu, const double* __restrict__ z,
                                                   const double* __restrict__ w, const uint32_t* __restrict__ pi,
                                                   NnTerms nt, int m, double* __restrict__ WS) {
    const int lane = lane_id();
    const size_t nch = (size_t)(m + WAVE - 1) / WAVE, tasks = G * nch;
    for (size_t task = (size_t)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE; task < tasks;
         task += (size_t)gridDim.x * (TPB / WAVE)) {
        const size_t g = task / nch;
        const int j = (int)(task % nch) * WAVE + lane;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (uint32_t k = off[g]; k < off[g + 1]; k++) {
            const uint32_t run = order[k];
            for (uint32_t i = run_off[run]; i < run_off[run + 1]; i++) {
                const uint32_t kr = pi[i];
                bool hit = false;
                for (int q = 0; q < nt.nq; q++) hit |= (int)(nt.off[q] + nt.codes[q][kr]) == j;
                if (hit) {
                    const double2 u = wu[i];
                    s0 += u.x;
                    s1 += u.y;
                    s2 += (w ? w[i] : 1.0) * z[i];
                }
            }
        }
        if (j < m) {
            WS[(g * 3 + 0) * m + j] = s0;
            WS[(g * 3 + 1) * m + j] = s1;
            WS[(g * 3 + 2) * m + j] = s2;
        }
    }
}

// Pm[g] = [C.S_0, C.S_1, bread S_0 (mk, then m left for G^- F_0), bread S_1 (mk, then m)] (row
// stride NC1 = 2 + 2p) and Bt[g] = [B_g, f_g] (p).
__global__ void k_wide_rows(const double* __restrict__ S, const double* __restrict__ WS, size_t G, int mk, int m,
                            const double* __restrict__ bread, const double* __restrict__ C, double* __restrict__ Pm,
                            double* __restrict__ Bt) {
    const int p = mk + m, NC1 = 2 + 2 * p;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < G; g += (size_t)gridDim.x * blockDim.x) {
        double* row = Pm + g * NC1;
        for (int t = 0; t < 2; t++) {
            const double* St = S + (g * 3 + t) * mk;
            double num = 0.0;
            for (int a = 0; a < mk; a++) {
                double f = 0.0;
                for (int b = 0; b < mk; b++) f += bread[a * mk + b] * St[b];
                row[2 + t * p + a] = f;
                num += C[a] * St[a];
            }
            row[t] = num;
        }
        double* bt = Bt + g * p;
        for (int a = 0; a < mk; a++) bt[a] = S[(g * 3 + 2) * mk + a];
        for (int j = 0; j < m; j++) bt[mk + j] = WS[(g * 3 + 2) * m + j];
    }
}

// V[b][g] = the Rademacher sign of cluster g in draw b (weights.rs:65-80): bit g mod 64 of word
// g / 64 of the draw's SplitMix64 stream, 0 -> +1 (full enumeration: bit g of b).
__global__ void k_wide_signs(uint64_t seed, int full, uint64_t reps, size_t G, double* __restrict__ V) {
    const size_t words = (G + 63) / 64;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < reps * words; e += (size_t)gridDim.x * blockDim.x) {
        const uint64_t b = e / words;
        const size_t s = e % words;
        const uint64_t word = full ? (s == 0 ? b : 0ull) : sm_mix((seed ^ (b * SM_STREAM)) + (uint64_t)(s + 2) * SM_GAMMA);
        const size_t g1 = min(G, s * 64 + 64);
        for (size_t g = s * 64; g < g1; g++) V[b * G + g] = ((word >> (g - s * 64)) & 1ull) ? -1.0 : 1.0;
    }
}

// Per draw (one wave): num_t = Y[b][t], q_t[c] = V[b][c] A_t[c] - Z_t[b][c], v = ssc sum_c q q'.
__global__ __launch_bounds__(TPB) void k_wide_post(const double* __restrict__ Y, int NC1, const double* __restrict__ V,
                                                   const double* __restrict__ Pm, const double* __restrict__ Z0,
                                                   const double* __restrict__ Z1, size_t G, uint64_t reps, double ssc,
                                                   double* __restrict__ coefs) {
    const int w = threadIdx.x / WAVE, lane = lane_id();
    for (uint64_t b = (uint64_t)blockIdx.x * (TPB / WAVE) + w; b < reps; b += (uint64_t)gridDim.x * (TPB / WAVE)) {
        double s00 = 0.0, s01 = 0.0, s11 = 0.0;
        for (size_t c = lane; c < G; c += WAVE) {
            const double v = V[b * G + c];
            const double q0 = v * Pm[c * NC1] - Z0[b * G + c];
            const double q1 = v * Pm[c * NC1 + 1] - Z1[b * G + c];
            s00 += q0 * q0;
            s01 += q0 * q1;
            s11 += q1 * q1;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            s00 += __shfl_xor(s00, o, WAVE);
            s01 += __shfl_xor(s01, o, WAVE);
            s11 += __shfl_xor(s11, o, WAVE);
        }
        if (lane == 0) {
            double* o = coefs + b * 5;
            o[0] = Y[b * NC1];
            o[1] = Y[b * NC1 + 1];
            o[2] = ssc * s00;
            o[3] = ssc * s01;
            o[4] = ssc * s11;
        }
    }
}

__global__ void k_scale_rows_sqrt(double* __restrict__ D, size_t n, int m, const double* __restrict__ w) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * m; e += (size_t)gridDim.x * blockDim.x)
        D[e] *= sqrt(w[e / m]);
}

void blas_ok(rocblas_status s, const char* what) {
    if (s != rocblas_status_success)
        throw Error(FH_INTERNAL, std::string("rocBLAS ") + what + " failed (status " + std::to_string((int)s) + ")");
}

// A rocBLAS handle on the absorber's stream, atomics off (results independent of scheduling).
struct Blas {
    rocblas_handle h = nullptr;
    explicit Blas(hipStream_t st) {
        blas_ok(rocblas_create_handle(&h), "create_handle");
        blas_ok(rocblas_set_stream(h, st), "set_stream");
        blas_ok(rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed), "set_atomics_mode");
    }
    ~Blas() {
        if (h) (void)rocblas_destroy_handle(h);
    }
    Blas(const Blas&) = delete;
    Blas& operator=(const Blas&) = delete;
    // row-major C (rows x cols, ldc) = op(A) op(B) through the column-major identity C' = B' A'
    void gemm(rocblas_operation ta, rocblas_operation tb, int m, int n, int k, const double* A, int lda,
              const double* B, int ldb, double* C, int ldc) {
        const double one = 1.0, zero = 0.0;
        blas_ok(rocblas_dgemm(h, ta, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc), "dgemm");
    }
};

}  // namespace

void ols_wildboot_wide(Absorber& a, const std::vector<int>& nested, uint64_t m_nn, const WbootInput& in,
                       const fh_wboot_spec& spec, const double* rk, fh_wboot_result& res, double* t_boot_host) {
    hipStream_t st = a.st;
    const size_t G = in.G, n = a.n;
    const int mk = (int)in.kept.size(), m = (int)m_nn, p = mk + m, NC1 = 2 + 2 * p;
    require(mk >= 1 && mk + 2 <= KMAX, "the bootstrap supports 1..19 kept regressors");
    require(m >= 1 && m <= 512, "the non-nested projection supports 1..512 levels");
    require(G >= 2 && G < (1ull << 24), "the bootstrap needs 2 .. 2^24 clusters");
    require(n < (1ull << 31), "the non-nested projection with many levels supports < 2^31 rows");
    require(spec.reps >= 1, "wildboot: reps must be at least 1");
    require(spec.level > 0.0 && spec.level < 1.0, "wildboot: confidence level must be in (0, 1)");
    require(spec.tol > 0.0 && std::isfinite(spec.tol) && spec.maxiter > 0, "wildboot: tol and maxiter must be positive");
    require(spec.pval_type >= 0 && spec.pval_type <= 3, "wildboot: unknown p-value type");
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> C, MR;
    wb_constraint(in.bread, mk, rk, C, MR);
    DBuf<double> Sown;
    const double* Sp = in.S;
    if (!Sp) {
        wb_cluster_sums(a, in.rows, in.R, in.kept, C, MR, in.e, in.es, in.w, in.cl, G, Sown);
        Sp = Sown.p;
    }
    NnTerms nt{};
    uint32_t noff = 0;
    bool any_nested = false;
    for (size_t q = 0; q < a.nterms; q++) {
        if (nested[q]) {
            any_nested = true;
            continue;
        }
        require(nt.nq < 8, "at most 8 non-nested terms");
        require(a.codes_kept.size() == a.nterms && a.codes_kept[q].p, "the absorber has no term codes", FH_INTERNAL);
        nt.codes[nt.nq] = a.codes_kept[q].p;
        nt.off[nt.nq++] = noff;
        noff += (uint32_t)a.term_g[q];
    }
    Blas blas(st);
    // G_nn = D~' W D~ over the dummies, residualized on the nested terms when there are any
    std::vector<double> Gnn((size_t)m * m);
    {
        DBuf<double> D(n * m), dG((size_t)m * m);
        const double* wg = in.w;
        fh_absorber* tmp = nullptr;
        struct Tmp {
            fh_absorber* p;
            ~Tmp() {
                if (p) fh_absorber_destroy(p);
            }
        } hold{nullptr};
        if (!any_nested) {
            hipLaunchKernelGGL(k_nn_dummies, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, n, a.pi.p, nt, m, D.p);
            FH_CHECK_LAUNCH();
        } else {
            std::vector<fh_term> tt;
            for (size_t q = 0; q < a.nterms; q++)
                if (nested[q]) tt.push_back(fh_term{a.codes_kept[q].p, a.term_g[q], 1});
            DBuf<double> wk;
            if (in.w) {
                wk.alloc(n);
                hipLaunchKernelGGL(k_to_kept, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, in.w, a.pi.p, n, wk.p);
                FH_CHECK_LAUNCH();
            }
            stream_wait(st);
            fh_solver_config cfg;
            fh_config_default(&cfg);
            if (fh_absorber_create_dev(tt.data(), tt.size(), in.w ? wk.p : nullptr, n, &cfg, nullptr, &tmp) != FH_OK)
                throw Error(FH_INTERNAL, std::string("non-nested projection: ") + fh_last_error());
            hold.p = tmp;
            hipLaunchKernelGGL(k_nn_dummies, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, n, tmp->a.pi.p, nt, m, D.p);
            FH_CHECK_LAUNCH();
            stream_wait(st);
            fh_solve_opts o{1e-12, 100000, 0, FH_PLAIN_FP64};
            fh_solve_info info{};
            if (fh_absorber_residualize_dev(tmp, D.p, (size_t)m, &o, nullptr, nullptr, &info, nullptr) != FH_OK)
                throw Error(FH_INTERNAL, std::string("non-nested projection: ") + fh_last_error());
            stream_wait(tmp->a.st);
            wg = tmp->a.weighted ? tmp->a.w_rows.p : nullptr;
        }
        if (wg) {
            hipLaunchKernelGGL(k_scale_rows_sqrt, dim3(grid_for(n * m, TPB, 8192)), dim3(TPB), 0, st, D.p, n, m, wg);
            FH_CHECK_LAUNCH();
        }
        // row-major D (n x m) is column-major D' (m x n): G = D' D
        blas.gemm(rocblas_operation_none, rocblas_operation_transpose, m, m, (int)n, D.p, m, D.p, m, dG.p, m);
        download(Gnn.data(), dG.p, Gnn.size() * sizeof(double), st);
    }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < i; j++)
            Gnn[(size_t)i * m + j] = Gnn[(size_t)j * m + i] = 0.5 * (Gnn[(size_t)i * m + j] + Gnn[(size_t)j * m + i]);
    uint64_t nkeep = 0;
    const std::vector<double> Gm = grounded_inverse(Gnn, (size_t)m, std::vector<uint8_t>(m, 0), nkeep);
    // per-cluster sums of the dummies, the rows of the two products
    DBuf<uint32_t> clr(n);
    DBuf<double> z(n), dC(mk), dMR(mk), dbread((size_t)mk * mk), dGm((size_t)m * m);
    DBuf<double2> wu(n);
    DBuf<int> dkept(mk);
    upload(dkept.p, in.kept.data(), mk * sizeof(int), st);
    upload(dC.p, C.data(), mk * sizeof(double), st);
    upload(dMR.p, MR.data(), mk * sizeof(double), st);
    upload(dbread.p, in.bread.data(), (size_t)mk * mk * sizeof(double), st);
    upload(dGm.p, Gm.data(), Gm.size() * sizeof(double), st);
    hipLaunchKernelGGL(k_pc_rows, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, in.rows, in.R, dkept.p, mk, dC.p,
                       dMR.p, in.e, in.es, in.w, in.cl, a.pi.p, n, clr.p, z.p, wu.p);
    FH_CHECK_LAUNCH();
    ClusterRuns local;
    const ClusterRuns& cr = runs_for(a, in.cl, G, local);
    DBuf<double> WS(G * 3 * m), Pm(G * NC1), Bt(G * p);
    const size_t nch = (size_t)(m + WAVE - 1) / WAVE;
    hipLaunchKernelGGL(k_wide_sums, dim3(grid_for(G * nch, TPB / WAVE, 1u << 16)), dim3(TPB), 0, st, cr.off.p,
                       cr.order.p, cr.run_off.p, G, wu.p, z.p, in.w, a.pi.p, nt, m, WS.p);
    hipLaunchKernelGGL(k_wide_rows, dim3(grid_for(G, TPB, 4096)), dim3(TPB), 0, st, Sp, WS.p, G, mk, m, dbread.p, dC.p,
                       Pm.p, Bt.p);
    FH_CHECK_LAUNCH();
    // G^- F_t into Pm: column-major (m x G) = G^- (m x m, symmetric) times F_t' (m x G, stride 3m)
    for (int t = 0; t < 2; t++)
        blas.gemm(rocblas_operation_none, rocblas_operation_none, m, (int)G, m, dGm.p, m, WS.p + (size_t)t * m, 3 * m,
                  Pm.p + 2 + (size_t)t * p + mk, NC1);
    // draws
    uint64_t reps = spec.reps;
    bool full = false;
    if (G <= 20 && (1ull << G) <= reps) {  // weights.rs:39-61
        reps = 1ull << G;
        full = true;
    }
    require(reps < (1ull << 31), "the non-nested projection with many levels supports < 2^31 draws");
    DBuf<double> V(reps * G), Y(reps * NC1), Z0(reps * G), Z1(reps * G), coefs(reps * 5);
    hipLaunchKernelGGL(k_wide_signs, dim3(grid_for(reps * ((G + 63) / 64), TPB, 8192)), dim3(TPB), 0, st,
                       (uint64_t)spec.seed, (int)full, reps, G, V.p);
    FH_CHECK_LAUNCH();
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    FH_HIP(hipEventRecord(e0, st));
    // Y' (NC1 x reps) = Pm' (NC1 x G) V' (G x reps)
    blas.gemm(rocblas_operation_none, rocblas_operation_none, NC1, (int)reps, (int)G, Pm.p, NC1, V.p, (int)G, Y.p, NC1);
    // Z_t' (G x reps) = Bt (G x p) d_t' (p x reps), d_t = columns 2 + t p .. of Y
    blas.gemm(rocblas_operation_transpose, rocblas_operation_none, (int)G, (int)reps, p, Bt.p, p, Y.p + 2, NC1, Z0.p,
              (int)G);
    blas.gemm(rocblas_operation_transpose, rocblas_operation_none, (int)G, (int)reps, p, Bt.p, p, Y.p + 2 + p, NC1,
              Z1.p, (int)G);
    FH_HIP(hipEventRecord(e1, st));
    hipLaunchKernelGGL(k_wide_post, dim3(grid_for(reps, TPB / WAVE, 8192)), dim3(TPB), 0, st, Y.p, NC1, V.p, Pm.p,
                       Z0.p, Z1.p, G, reps, in.ssc, coefs.p);
    FH_CHECK_LAUNCH();
    std::vector<double> hc(reps * 5), hS(G * 3 * mk);
    FH_HIP(hipMemcpyAsync(hc.data(), coefs.p, hc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(hS.data(), Sp, hS.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    float gms = 0.f;
    FH_HIP(hipEventElapsedTime(&gms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    res.contraction_ms = gms;
    res.contraction_ops = 2.0 * (double)reps * (double)G * (double)NC1 + 4.0 * (double)reps * (double)G * (double)p;
    double aa = 0.0;
    for (size_t g = 0; g < G; g++) {
        double a0 = 0.0;
        for (int j = 0; j < mk; j++) a0 += C[j] * hS[(g * 3) * mk + j];
        aa += a0 * a0;
    }
    double point = 0.0;
    for (int j = 0; j < mk; j++) point += rk[j] * in.beta[j];
    wb_infer(hc, reps, full, G, point, in.ssc * aa, in.ssc, spec, res, t_boot_host, t0);
}

// ------------------------------------------------------------------ MNW 13 / 31 / 33
// MacKinnon, Nielsen and Webb (2023): the WCR bootstrap with the CRV3 jackknife in the data
// generating process (31), in the variance of the statistic (13), or both (33), for one
// restriction with the bootstrap clusters as the error clusters (wildboot/ols.rs:651-1082).
// Per cluster g, with A_g = X_g'W X_g and s_t,g = X_g'W u_t (t = 0: u_hat, t = 1: -X~ M_R, the
// affine interval's two targets), G = sum A_g and G_(-g) = G - A_g:
//   D_t = s_t                                     (DGP 1)
//   D_t = s_t + A_g (y1 - G_(-g)^+ R'h^+ R y1),   y1 = G_(-g)^+ s_t, h = R G_(-g)^+ R'   (DGP 3:
//         the score of the restricted leave-cluster-out residuals)
//   J = R G^+ (CRV1 variance) or R G_(-g)^+ (CRV3), k_t = J D_t, JA = J A_g,
//   phi_t = G^+ D_t, num_t = R phi_t.
// A draw's variance is mult sum_g (V_bg k_t - JA db_t)^2 with db_t = V_b phi: the form of the
// WCR11 contraction with P = [num_0, num_1, phi_0, phi_1, k_0 JA, k_1 JA], AA = sum k k',
// BB = sum JA JA'. mult = (G - 1)/G for the CRV3 variance, the CRV1 factor otherwise. Fixed
// effects not nested in the clusters enter as the reference's Path-1 pseudo-regressors: an
// orthonormal basis Q of their dummies residualized on the nested terms (Q'WQ = I).

namespace {

// out[g] = [A_g (p x p), s_0 (p), s_1 (p)] with x = [x~ (kept regressors), Q]: one wave per cluster,
// each lane summing its entries over the cluster's runs in row order (fixed order).
__global__ __launch_bounds__(TPB) void k_mnw_sums(const uint32_t* __restrict__ off, const uint32_t* __restrict__ order,
                                                  const uint32_t* __restrict__ run_off, size_t G,
                                                  const double* __restrict__ rows, int R, const int* __restrict__ kept,
                                                  int k, const double* __restrict__ Q, int r,
                                                  const double* __restrict__ MR, const double* __restrict__ e,
                                                  size_t es, const double* __restrict__ w, double* __restrict__ out) {
    const int p = k + r, E = p * p + 2 * p;
    for (size_t g = (size_t)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE; g < G;
         g += (size_t)gridDim.x * (TPB / WAVE)) {
        for (int ent = lane_id(); ent < E; ent += WAVE) {
            int a, b = 0, kind;
            if (ent < p * p) {
                kind = 0;
                a = ent / p;
                b = ent % p;
            } else if (ent < p * p + p) {
                kind = 1;
                a = ent - p * p;
            } else {
                kind = 2;
                a = ent - p * p - p;
            }
            double s = 0.0;
            for (uint32_t j = off[g]; j < off[g + 1]; j++) {
                const uint32_t run = order[j];
                for (uint32_t i = run_off[run]; i < run_off[run + 1]; i++) {
                    const double wi = w ? w[i] : 1.0;
                    const double xa = a < k ? rows[(size_t)i * R + 1 + kept[a]] : Q[(size_t)i * r + a - k];
                    double f;
                    if (kind == 0) {
                        f = b < k ? rows[(size_t)i * R + 1 + kept[b]] : Q[(size_t)i * r + b - k];
                    } else if (kind == 1) {
                        f = e[(size_t)i * es];
                    } else {
                        double z = 0.0;
                        for (int c = 0; c < k; c++) z += rows[(size_t)i * R + 1 + kept[c]] * MR[c];
                        f = -z;
                    }
                    s += wi * xa * f;
                }
            }
            out[g * E + ent] = s;
        }
    }
}

// dst[i] = src[inv[pi[i]]] (rows of m values): another absorber's row order into this one's.
__global__ void k_mnw_perm(const double* __restrict__ src, int m, const uint32_t* __restrict__ inv,
                           const uint32_t* __restrict__ pi, size_t n, double* __restrict__ dst) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t j = inv[pi[i]];
        for (int c = 0; c < m; c++) dst[i * m + c] = src[j * m + c];
    }
}

__global__ void k_mnw_inv(const uint32_t* __restrict__ pi, size_t n, uint32_t* __restrict__ inv) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        inv[pi[i]] = (uint32_t)i;
}

// Q[i] = D[i] T (n x m times m x r).
__global__ void k_mnw_basis(const double* __restrict__ D, size_t n, int m, const double* __restrict__ T, int r,
                            double* __restrict__ Q) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        for (int c = 0; c < r; c++) {
            double s = 0.0;
            for (int l = 0; l < m; l++) s += D[i * m + l] * T[l * r + c];
            Q[i * r + c] = s;
        }
}

// Symmetric pseudo-inverse: eigenvalues <= tol * (largest |eigenvalue|) dropped (wildboot/mod.rs:434-470).
std::vector<double> sym_pinv(std::vector<double> A, int n, double tol) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++) A[(size_t)i * n + j] = A[(size_t)j * n + i] = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
    std::vector<double> lam, U, out((size_t)n * n, 0.0);
    if (n == 0) return out;
    sym_eigen(A, n, lam, U);
    double lmax = 0.0;
    for (double l : lam) lmax = std::max(lmax, std::fabs(l));
    if (lmax == 0.0) return out;
    for (int e = 0; e < n; e++) {
        if (!(lam[e] > tol * lmax)) continue;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) out[(size_t)i * n + j] += U[(size_t)i * n + e] * U[(size_t)j * n + e] / lam[e];
    }
    return out;
}

}  // namespace

void mnw_basis(Absorber& a, const std::vector<int>& nested, uint64_t m_nn, const WbootInput& in, DBuf<double>& Q,
               int& r_nn) {
    hipStream_t st = a.st;
    const size_t n = a.n;
    const int m = (int)m_nn;
    NnTerms nt{};
    uint32_t off = 0;
    bool any_nested = false;
    for (size_t q = 0; q < a.nterms; q++) {
        if (nested[q]) {
            any_nested = true;
            continue;
        }
        require(nt.nq < 8, "at most 8 non-nested terms");
        require(a.codes_kept.size() == a.nterms && a.codes_kept[q].p, "the absorber has no term codes", FH_INTERNAL);
        nt.codes[nt.nq] = a.codes_kept[q].p;
        nt.off[nt.nq++] = off;
        off += (uint32_t)a.term_g[q];
    }
    DBuf<double> D(n * m);
    std::vector<double> Gnn((size_t)m * m);
    if (!any_nested) {
        hipLaunchKernelGGL(k_nn_dummies, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, n, a.pi.p, nt, m, D.p);
        FH_CHECK_LAUNCH();
        gram_cols(a, D.p, n, m, 0, m, in.w, nullptr, Gnn.data());
    } else {
        // the dummies residualized on the nested terms, in an absorber of those terms alone
        std::vector<fh_term> tt;
        for (size_t q = 0; q < a.nterms; q++)
            if (nested[q]) tt.push_back(fh_term{a.codes_kept[q].p, a.term_g[q], 1});
        DBuf<double> wk;
        if (in.w) {
            wk.alloc(n);
            hipLaunchKernelGGL(k_to_kept, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, in.w, a.pi.p, n, wk.p);
            FH_CHECK_LAUNCH();
        }
        stream_wait(st);
        fh_solver_config cfg;
        fh_config_default(&cfg);
        fh_absorber* tmp = nullptr;
        if (fh_absorber_create_dev(tt.data(), tt.size(), in.w ? wk.p : nullptr, n, &cfg, nullptr, &tmp) != FH_OK)
            throw Error(FH_INTERNAL, std::string("non-nested projection: ") + fh_last_error());
        struct Tmp {
            fh_absorber* p;
            ~Tmp() { fh_absorber_destroy(p); }
        } hold{tmp};
        DBuf<double> dn(n * m);
        hipLaunchKernelGGL(k_nn_dummies, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, n, tmp->a.pi.p, nt, m, dn.p);
        FH_CHECK_LAUNCH();
        stream_wait(st);
        fh_solve_opts o{1e-12, 100000, 0, FH_PLAIN_FP64};
        fh_solve_info info{};
        if (fh_absorber_residualize_dev(tmp, dn.p, (size_t)m, &o, nullptr, nullptr, &info, nullptr) != FH_OK)
            throw Error(FH_INTERNAL, std::string("non-nested projection: ") + fh_last_error());
        {
            AbsScope sc(tmp->a);
            gram_cols(tmp->a, dn.p, n, m, 0, m, tmp->a.weighted ? tmp->a.w_rows.p : nullptr, nullptr, Gnn.data());
            stream_wait(tmp->a.st);
        }
        DBuf<uint32_t> inv(n);
        hipLaunchKernelGGL(k_mnw_inv, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, tmp->a.pi.p, n, inv.p);
        hipLaunchKernelGGL(k_mnw_perm, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, dn.p, m, inv.p, a.pi.p, n, D.p);
        FH_CHECK_LAUNCH();
        stream_wait(st);
    }
    // Q = D~ U Lambda^-1/2 over the eigenvalues above 1e-10 of the largest (wildboot/mod.rs:1036-1076)
    for (int i = 0; i < m; i++)
        for (int j = 0; j < i; j++) Gnn[(size_t)i * m + j] = Gnn[(size_t)j * m + i] = 0.5 * (Gnn[(size_t)i * m + j] + Gnn[(size_t)j * m + i]);
    std::vector<double> lam, U;
    sym_eigen(Gnn, m, lam, U);
    double lmax = 0.0;
    for (double l : lam) lmax = std::max(lmax, std::fabs(l));
    std::vector<int> keep;
    for (int e = 0; e < m; e++)
        if (lmax > 0.0 && lam[e] > 1e-10 * lmax) keep.push_back(e);
    r_nn = (int)keep.size();
    if (r_nn == 0) return;
    std::vector<double> T((size_t)m * r_nn);
    for (int l = 0; l < m; l++)
        for (int c = 0; c < r_nn; c++) T[(size_t)l * r_nn + c] = U[(size_t)l * m + keep[c]] / std::sqrt(lam[keep[c]]);
    DBuf<double> dT(T.size());
    h2d(dT.p, T.data(), T.size(), st);
    Q.alloc(n * r_nn);
    hipLaunchKernelGGL(k_mnw_basis, dim3(grid_for(n, TPB, 8192)), dim3(TPB), 0, st, D.p, n, m, dT.p, r_nn, Q.p);
    FH_CHECK_LAUNCH();
    stream_wait(st);
}

void ols_wildboot_mnw(Absorber& a, const WbootInput& in, const fh_wboot_spec& spec, int boot_type, const double* rk,
                      const double* Q, int r_nn, fh_wboot_result& res, double* t_boot_host) {
    hipStream_t st = a.st;
    const size_t G = in.G;
    const int k = (int)in.kept.size(), p = k + r_nn;
    require(k >= 1 && p + 2 <= KMAX, "the MNW bootstrap supports kept regressors + non-nested levels <= 19");
    require(G >= 2 && G < (1ull << 24), "the bootstrap needs 2 .. 2^24 clusters");
    require(spec.reps >= 1, "wildboot: reps must be at least 1");
    require(spec.level > 0.0 && spec.level < 1.0, "wildboot: confidence level must be in (0, 1)");
    require(spec.tol > 0.0 && std::isfinite(spec.tol) && spec.maxiter > 0, "wildboot: tol and maxiter must be positive");
    require(spec.pval_type >= 0 && spec.pval_type <= 3, "wildboot: unknown p-value type");
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> C, MR;
    wb_constraint(in.bread, k, rk, C, MR);
    DBuf<int> dkept(k);
    DBuf<double> dMR(k);
    h2d(dkept.p, in.kept.data(), k, st);
    h2d(dMR.p, MR.data(), k, st);
    ClusterRuns local;
    const ClusterRuns& cr = runs_for(a, in.cl, G, local);
    const int E = p * p + 2 * p;
    DBuf<double> S(G * E);
    hipLaunchKernelGGL(k_mnw_sums, dim3(grid_for(G, TPB / WAVE, 1u << 16)), dim3(TPB), 0, st, cr.off.p, cr.order.p,
                       cr.run_off.p, G, in.rows, in.R, dkept.p, k, Q, r_nn, dMR.p, in.e, in.es, in.w, S.p);
    FH_CHECK_LAUNCH();
    std::vector<double> hs(G * E);
    d2h(hs.data(), S.p, hs.size(), st);
    stream_wait(st);
    // G, its pseudo-inverse (eigenvalues <= 1e-12 of the largest dropped) and R over [x~, Q]
    std::vector<double> Gt((size_t)p * p, 0.0), Ra(p, 0.0);
    for (size_t g = 0; g < G; g++)
        for (int e = 0; e < p * p; e++) Gt[e] += hs[g * E + e];
    for (int j = 0; j < k; j++) Ra[j] = rk[j];
    const std::vector<double> Gi = sym_pinv(Gt, p, 1e-12);
    std::vector<double> cR(p, 0.0);  // G^+ R'
    for (int i = 0; i < p; i++)
        for (int j = 0; j < p; j++) cR[i] += Gi[(size_t)i * p + j] * Ra[j];
    const bool dgp3 = boot_type == 31 || boot_type == 33, den3 = boot_type == 13 || boot_type == 33;
    const double mult = den3 ? ((double)G - 1.0) / (double)G : in.ssc;
    const int NC = 2 + 4 * p;
    std::vector<double> P(G * NC), AA(4, 0.0), BB((size_t)p * p, 0.0), cmax(NC, 0.0);
    double v_obs = 0.0;
    std::vector<double> A((size_t)p * p), J(p), JA(p), D(p), y1(p), gnr(p), Gm((size_t)p * p), Gmi;
    for (size_t g = 0; g < G; g++) {
        const double* hg = hs.data() + g * E;
        for (int e = 0; e < p * p; e++) A[e] = hg[e];
        const double* s0 = hg + p * p;
        const double* s1 = s0 + p;
        if (dgp3 || den3) {
            for (int e = 0; e < p * p; e++) Gm[e] = Gt[e] - A[e];
            Gmi = sym_pinv(Gm, p, 1e-10);
        }
        // J = R G_(-g)^+ (CRV3) or R G^+ (CRV1)
        for (int c = 0; c < p; c++) {
            double s = 0.0;
            if (den3)
                for (int j = 0; j < p; j++) s += Ra[j] * Gmi[(size_t)j * p + c];
            else
                s = cR[c];
            J[c] = s;
        }
        for (int c = 0; c < p; c++) {
            double s = 0.0;
            for (int j = 0; j < p; j++) s += J[j] * A[(size_t)j * p + c];
            JA[c] = s;
        }
        double jo = 0.0;
        for (int c = 0; c < p; c++) jo += J[c] * s0[c];
        v_obs += jo * jo;
        double hinv = 0.0;
        if (dgp3) {
            double h = 0.0;
            for (int i = 0; i < p; i++) {
                double s = 0.0;
                for (int j = 0; j < p; j++) s += Gmi[(size_t)i * p + j] * Ra[j];
                gnr[i] = s;
                h += Ra[i] * s;
            }
            hinv = h > 0.0 ? 1.0 / h : 0.0;  // the 1 x 1 pseudo-inverse
        }
        double kt[2], rphi[2];
        double* row = P.data() + g * NC;
        for (int t = 0; t < 2; t++) {
            const double* sin = t == 0 ? s0 : s1;
            if (!dgp3) {
                for (int c = 0; c < p; c++) D[c] = sin[c];
            } else {
                for (int i = 0; i < p; i++) {
                    double s = 0.0;
                    for (int j = 0; j < p; j++) s += Gmi[(size_t)i * p + j] * sin[j];
                    y1[i] = s;
                }
                double ry = 0.0;
                for (int i = 0; i < p; i++) ry += Ra[i] * y1[i];
                for (int i = 0; i < p; i++) y1[i] -= gnr[i] * hinv * ry;  // the restricted step
                for (int i = 0; i < p; i++) {
                    double s = 0.0;
                    for (int j = 0; j < p; j++) s += A[(size_t)i * p + j] * y1[j];
                    D[i] = sin[i] + s;
                }
            }
            double kk = 0.0, rp = 0.0;
            for (int c = 0; c < p; c++) kk += J[c] * D[c];
            for (int i = 0; i < p; i++) {
                double ph = 0.0;
                for (int j = 0; j < p; j++) ph += Gi[(size_t)i * p + j] * D[j];
                row[2 + t * p + i] = ph;
                rp += Ra[i] * ph;
            }
            kt[t] = kk;
            rphi[t] = rp;
        }
        row[0] = rphi[0];
        row[1] = rphi[1];
        for (int c = 0; c < p; c++) {
            row[2 + 2 * p + c] = kt[0] * JA[c];
            row[2 + 3 * p + c] = kt[1] * JA[c];
        }
        AA[0] += kt[0] * kt[0];
        AA[1] += kt[0] * kt[1];
        AA[3] += kt[1] * kt[1];
        for (int i = 0; i < p; i++)
            for (int j = 0; j < p; j++) BB[(size_t)i * p + j] += JA[i] * JA[j];
        for (int c = 0; c < NC; c++) cmax[c] = std::max(cmax[c], std::fabs(row[c]));
    }
    AA[2] = AA[1];
    std::vector<double> scale(NC), inv_scale(NC);
    for (int c = 0; c < NC; c++) {
        scale[c] = cmax[c] > 0.0 ? 4.0 * cmax[c] : 1.0;
        inv_scale[c] = std::ldexp(1.0 / scale[c], 56);
    }
    DBuf<double> dP(P.size());
    h2d(dP.p, P.data(), P.size(), st);
    double point = 0.0;
    for (int j = 0; j < k; j++) point += rk[j] * in.beta[j];
    wb_draws(a, dP.p, G, NC, p, scale, inv_scale, AA, BB, mult, point, mult * v_obs, spec, res, t_boot_host, t0);
}

}  // namespace fh
