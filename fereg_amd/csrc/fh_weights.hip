// Weight state on the GPU (operator.rs:211-338, 344-570): cell weights W_c, FE1 level weights
// n1 (and sum w), level weights n_r, diag(S) = n_r - sum_s c_{s,k}^2/n1_s, redundant flags,
// the Jacobi inverse diagonal (precond/jacobi.rs:7-9), and the dense small super-block S_SS
// whose grounded LDL' factor (precond/dense.rs:25-37, 79-115) is turned into an explicit
// pseudo-inverse on the host (an m_s x m_s FP64 factorization) and applied on the GPU.
#include <algorithm>
#include <cmath>

#include "fh_internal.hpp"

namespace fh {

namespace {

constexpr unsigned TPB = 256;
inline unsigned blocks_for(size_t n) { return grid_for(n, TPB, 8192); }

// out[c][b - 1] = codes[c][b] for b = 1 .. nb-1.
__global__ void k_drop_block0(const uint32_t* __restrict__ codes, size_t M, int nb, uint32_t* __restrict__ out) {
    const size_t w = (size_t)(nb - 1), total = M * w;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t c = e / w, b = e - c * w;
        out[e] = codes[c * nb + 1 + b];
    }
}

__global__ void k_cell_weights(const uint32_t* __restrict__ coo, const double* __restrict__ w, size_t M,
                               double* __restrict__ cw, float* __restrict__ cw32) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (size_t)gridDim.x * blockDim.x) {
        size_t r0 = coo ? coo[c] : c, r1 = coo ? coo[c + 1] : c + 1;
        double s = 0.0;
        for (size_t i = r0; i < r1; i++) s += w ? w[i] : 1.0;
        cw[c] = s;
        cw32[c] = (float)s;
    }
}

__global__ void k_n1(const uint32_t* __restrict__ seg_off, const double* __restrict__ cw, size_t g1,
                     double* __restrict__ n1, double* __restrict__ inv64, float* __restrict__ inv32) {
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < g1; s += (size_t)gridDim.x * blockDim.x) {
        uint32_t c0 = seg_off[s], c1 = seg_off[s + 1];
        double v = 0.0;
        if (cw)
            for (uint32_t c = c0; c < c1; c++) v += cw[c];
        else
            v = double(c1 - c0);
        n1[s] = v;
        // an FE1 level without rows (codes need not use every level) contributes nothing: its
        // segment sums are 0, not 0/0, so the dense small-block products stay finite
        const double inv = v > 0.0 ? 1.0 / v : 0.0;
        inv64[s] = inv;
        inv32[s] = (float)inv;
    }
}

// Deterministic partial sums (fixed grid, fixed order): out[blockIdx.x].
__global__ __launch_bounds__(TPB) void k_partial_sum(const double* __restrict__ v, size_t n, double* __restrict__ out) {
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) s += v[i];
    __shared__ double ws[TPB / WAVE];
    s = wave_sum(s);
    if (lane_id() == 0) ws[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(TPB / WAVE); i++) t += ws[i];
        out[blockIdx.x] = t;
    }
}

// Large block: L lanes per level read its positions coalesced, bucket by bucket ((bucket,
// level, cell) order). The cells of one FE1 level are adjacent within a bucket, so c_{s,k} are
// run sums: the lane holding the last position of a run sums the run forward from its start
// (each run is summed once, in position order) and adds run^2 / n1_s.
__global__ __launch_bounds__(TPB) void k_level_weights(const uint32_t* __restrict__ boff,
                                                       const uint32_t* __restrict__ lseg, uint32_t nbk,
                                                       const double* __restrict__ lw, const double* __restrict__ n1,
                                                       size_t g, size_t off, int L, double* __restrict__ nr,
                                                       double* __restrict__ diag, uint8_t* __restrict__ red,
                                                       double* __restrict__ inv64, float* __restrict__ inv32) {
    const int l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwaves = (size_t)gridDim.x * (TPB / WAVE);
    for (size_t base = ((size_t)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE) * gpw; base < g; base += nwaves * gpw) {
        const size_t lv = base + gl;
        const bool valid = lv < g;
        double tot = 0.0, dq = 0.0;
        for (uint32_t k = 0; k < nbk; k++) {
            const uint32_t j0 = valid ? boff[(size_t)k * g + lv] : 0, j1 = valid ? boff[(size_t)k * g + lv + 1] : 0;
            for (uint32_t j = j0 + li; j < j1; j += L) {
                const double w = lw ? lw[j] : 1.0;
                tot += w;
                const uint32_t s = lseg[j];
                if (j + 1 < j1 && lseg[j + 1] == s) continue;  // not the run's last position
                uint32_t q = j;
                while (q > j0 && lseg[q - 1] == s) q--;
                double run = 0.0;
                for (uint32_t i = q; i <= j; i++) run += lw ? lw[i] : 1.0;
                dq += run * run / n1[s];
            }
        }
        for (int o = L >> 1; o > 0; o >>= 1) {
            tot += __shfl_xor(tot, o, WAVE);
            dq += __shfl_xor(dq, o, WAVE);
        }
        if (!valid || li != 0) continue;
        const double d = tot - dq;
        const bool r = d <= 1e-12 * tot;  // operator.rs:21, 319-333
        nr[off + lv] = tot;
        diag[off + lv] = d;
        red[off + lv] = r;
        const double inv = (r || !(d > 0.0)) ? 0.0 : 1.0 / d;  // jacobi.rs:7-9
        inv64[off + lv] = inv;
        inv32[off + lv] = (float)inv;
    }
}

// lw[j] = W of the cell at position j of a bucketed block (bucket-major order). Bucket q's
// positions only reference cells of its FE1 range; with X = XCD_COUNT they are gathered by the
// workgroups of XCD q mod X, whose L2 then holds that slice of the cell weights.
__global__ __launch_bounds__(TPB) void k_gather_lw_x(const uint32_t* __restrict__ boff, uint32_t nbk, size_t g,
                                                     const uint32_t* __restrict__ bperm, const double* __restrict__ cw,
                                                     uint32_t X, double* __restrict__ lw64, float* __restrict__ lw32) {
    const uint32_t xq = blockIdx.x % X;
    const size_t lb = blockIdx.x / X, nl = gridDim.x / X;
    for (uint32_t q = xq; q < nbk; q += X) {
        const size_t j0 = boff[(size_t)q * g], j1 = boff[(size_t)(q + 1) * g];
        for (size_t j = j0 + lb * TPB + threadIdx.x; j < j1; j += nl * TPB) {
            const double w = cw[bperm[j]];
            lw64[j] = w;
            lw32[j] = (float)w;
        }
    }
}

// lw[j] = W of the cell at position j (level-major order of a small block).
__global__ void k_gather_lw(const uint32_t* __restrict__ bperm, const double* __restrict__ cw, size_t M,
                            double* __restrict__ lw64, float* __restrict__ lw32) {
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += (size_t)gridDim.x * blockDim.x) {
        const double w = cw[bperm[j]];
        lw64[j] = w;
        lw32[j] = (float)w;
    }
}

struct SmallOffs {
    uint32_t off[NB_MAX];
    int nsb, nb;
};

// S_SS (+ n_r of the small coordinates in row ms): one thread per segment adds
// sum_c W_c [k][l] and subtracts sum_{c,c'} W_c W_c' / n1_s [k(c)][l(c')] (operator.rs:1457-1560).
// The terms are rounded to 64-bit fixed point (fscale = 2^(62 - ceil log2 sw), every partial sum
// stays below 2^62) and added with integer atomics: the result is exact for the rounded terms and
// independent of the order the atomics land in (deterministic under any scheduling). With
// ms <= SA_LDS_MAX each workgroup accumulates in LDS and writes its partial to out[blockIdx.x].
constexpr int SA_LDS_MAX = 64;
__device__ __forceinline__ unsigned long long fixed(double v, double fscale) {
    return (unsigned long long)llrint(v * fscale);  // two's complement for negative terms
}
__global__ __launch_bounds__(TPB) void k_small_assemble(const uint32_t* __restrict__ seg_off,
                                                        const uint32_t* __restrict__ cell_codes,
                                                        const double* __restrict__ cw, const double* __restrict__ n1,
                                                        size_t g1, size_t ms, SmallOffs so, double fscale,
                                                        unsigned long long* __restrict__ out) {
    extern __shared__ unsigned long long acc[];
    const size_t len = ms * ms + ms;
    const bool lds = ms <= SA_LDS_MAX;
    unsigned long long* dst = lds ? acc : out;
    if (lds) {
        for (size_t i = threadIdx.x; i < len; i += blockDim.x) acc[i] = 0ull;
        __syncthreads();
    }
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < g1; s += (size_t)gridDim.x * blockDim.x) {
        uint32_t c0 = seg_off[s], c1 = seg_off[s + 1];
        double inv = 1.0 / n1[s];
        for (uint32_t c = c0; c < c1; c++) {
            double w = cw ? cw[c] : 1.0;
            const unsigned long long fw = fixed(w, fscale);
            for (int b = 0; b < so.nsb; b++) {
                uint32_t k = so.off[b] + cell_codes[(size_t)c * so.nb + b];
                atomicAdd(&dst[ms * ms + k], fw);
                for (int b2 = 0; b2 < so.nsb; b2++) {
                    uint32_t l = so.off[b2] + cell_codes[(size_t)c * so.nb + b2];
                    atomicAdd(&dst[(size_t)k * ms + l], fw);
                }
            }
            for (uint32_t c2 = c0; c2 < c1; c2++) {
                const unsigned long long fw2 = fixed(-(w * (cw ? cw[c2] : 1.0) * inv), fscale);
                for (int b = 0; b < so.nsb; b++) {
                    uint32_t k = so.off[b] + cell_codes[(size_t)c * so.nb + b];
                    for (int b2 = 0; b2 < so.nsb; b2++) {
                        uint32_t l = so.off[b2] + cell_codes[(size_t)c2 * so.nb + b2];
                        atomicAdd(&dst[(size_t)k * ms + l], fw2);
                    }
                }
            }
        }
    }
    if (lds) {
        __syncthreads();
        unsigned long long* part = out + (size_t)blockIdx.x * len;
        for (size_t i = threadIdx.x; i < len; i += blockDim.x) part[i] = acc[i];
    }
}

// out[i] = (sum over the partials of the fixed-point parts[q][i]) / fscale (integer sums: exact).
__global__ void k_fixed_to_double(const unsigned long long* __restrict__ parts, size_t nparts, size_t len,
                                  double inv_fscale, double* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
        unsigned long long s = 0;
        for (size_t q = 0; q < nparts; q++) s += parts[q * len + i];
        out[i] = (double)(long long)s * inv_fscale;
    }
}

// E8[i] = v when v is an integer in [0, 255], else *bad = 1 (the u8 copy of E is then unused).
__device__ __forceinline__ void put_e8(uint8_t* __restrict__ E8, uint32_t* __restrict__ bad, size_t i, double v) {
    if (!E8) return;
    if (v >= 0.0 && v <= 255.0 && v == rint(v)) E8[i] = (uint8_t)v;
    else atomicOr(bad, 1u);
}

// Level sums c_{s,k} of the single small block (block 0, whose codes ascend within a segment),
// one thread per segment, the row in registers. Optional outputs: row s of
// C = [c_{s,0..ms-1} / sqrt(n1_s), sqrt(n1_s)] (ld ms + 1) and the dense D[k][s] = c_{s,k}
// (FP64 and FP32, pitch ldd).
template <int MS>
__global__ __launch_bounds__(TPB) void k_seg_level_sums(const uint32_t* __restrict__ seg_off,
                                                        const uint32_t* __restrict__ cell_codes,
                                                        const double* __restrict__ cw, const double* __restrict__ n1,
                                                        size_t g1, int nb, uint32_t off0, int ms,
                                                        double* __restrict__ C, size_t ldd, double* __restrict__ D64,
                                                        float* __restrict__ D32, uint8_t* __restrict__ D8,
                                                        uint32_t* __restrict__ bad8, int L) {
    // L lanes per segment (its cells strided over the lanes, register bins, shuffle reduction)
    const int l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwaves = (size_t)gridDim.x * (TPB / WAVE);
    for (size_t base = ((size_t)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE) * gpw; base < g1; base += nwaves * gpw) {
        const size_t s = base + gl;
        const bool valid = s < g1;
        double row[MS];
#pragma unroll
        for (int k = 0; k < MS; k++) row[k] = 0.0;
        const uint32_t c0 = valid ? seg_off[s] : 0, c1 = valid ? seg_off[s + 1] : 0;
        for (uint32_t c = c0 + li; c < c1; c += L) {
            const uint32_t key = off0 + cell_codes[(size_t)c * nb];
            const double w = cw ? cw[c] : 1.0;
#pragma unroll
            for (int k = 0; k < MS; k++) row[k] += key == (uint32_t)k ? w : 0.0;
        }
#pragma unroll
        for (int k = 0; k < MS; k++)
            for (int o = L >> 1; o > 0; o >>= 1) row[k] += __shfl_xor(row[k], o, WAVE);
        if (!valid || li != 0) continue;
        if (D64)
#pragma unroll
            for (int k = 0; k < MS; k++)
                if (k < ms) {
                    D64[(size_t)k * ldd + s] = row[k];
                    D32[(size_t)k * ldd + s] = (float)row[k];
                    put_e8(D8, bad8, (size_t)k * ldd + s, row[k]);
                }
        if (!C) continue;
        const double sq = sqrt(n1[s]);
        const double f = sq > 0.0 ? 1.0 / sq : 0.0;
        double* out = C + s * (ms + 1);
#pragma unroll
        for (int k = 0; k < MS; k++)
            if (k < ms) out[k] = row[k] * f;
        out[ms] = sq;
    }
}

// C[k][coff + l] = sum of W over the positions of level l of a large block whose small code
// (other block 0) is k: L lanes per level over its buckets' ranges in order, MS register bins
// per lane, then the lanes reduced by shuffles (fixed order).
template <int MS>
__global__ __launch_bounds__(TPB) void k_level_small_sums(const uint32_t* __restrict__ boff,
                                                          const uint32_t* __restrict__ lother, int nother,
                                                          const double* __restrict__ lw, uint32_t nbk, size_t g,
                                                          int ms, size_t ldc, size_t coff, int L,
                                                          double* __restrict__ C64, float* __restrict__ C32,
                                                          uint8_t* __restrict__ C8, uint32_t* __restrict__ bad8) {
    const int l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwaves = (size_t)gridDim.x * (TPB / WAVE);
    for (size_t base = ((size_t)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE) * gpw; base < g; base += nwaves * gpw) {
        const size_t lv = base + gl;
        const bool valid = lv < g;
        double row[MS];
#pragma unroll
        for (int k = 0; k < MS; k++) row[k] = 0.0;
        for (uint32_t q = 0; q < nbk; q++) {
            const uint32_t j0 = valid ? boff[(size_t)q * g + lv] : 0, j1 = valid ? boff[(size_t)q * g + lv + 1] : 0;
            for (uint32_t j = j0 + li; j < j1; j += L) {
                const uint32_t key = lother[(size_t)j * nother];
                const double w = lw ? lw[j] : 1.0;
#pragma unroll
                for (int k = 0; k < MS; k++) row[k] += key == (uint32_t)k ? w : 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k < MS; k++)
            for (int o = L >> 1; o > 0; o >>= 1) row[k] += __shfl_xor(row[k], o, WAVE);
        if (!valid || li != 0) continue;
#pragma unroll
        for (int k = 0; k < MS; k++)
            if (k < ms) {
                C64[(size_t)k * ldc + coff + lv] = row[k];
                C32[(size_t)k * ldc + coff + lv] = (float)row[k];
                put_e8(C8, bad8, (size_t)k * ldc + coff + lv, row[k]);
            }
    }
}

// Spreads the packed small-block results [n_r | diag | redundant | inverse] into their arrays
// (FP64 and FP32 copies; the small coordinates get no Jacobi scale).
__global__ void k_small_spread(const double* __restrict__ pk, size_t ms, int with_inv, double* __restrict__ nr,
                               double* __restrict__ nrs64, float* __restrict__ nrs32, double* __restrict__ diag,
                               uint8_t* __restrict__ red, double* __restrict__ inv_diag64, float* __restrict__ inv_diag32,
                               double* __restrict__ minv64, float* __restrict__ minv32) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < ms + (with_inv ? ms * ms : 0);
         i += (size_t)gridDim.x * blockDim.x) {
        if (i < ms) {
            nr[i] = pk[i];
            if (nrs64) { nrs64[i] = pk[i]; nrs32[i] = (float)pk[i]; }
            diag[i] = pk[ms + i];
            red[i] = pk[2 * ms + i] != 0.0 ? 1 : 0;
            inv_diag64[i] = 0.0;
            inv_diag32[i] = 0.0f;
        } else {
            const double v = pk[3 * ms + (i - ms)];
            minv64[i - ms] = v;
            minv32[i - ms] = (float)v;
        }
    }
}

// Lanes per segment of the segment level sums: half the segment pass's, since every lane's MS
// register bins are reduced across the segment's lanes (C2, 10 cells per segment: 1, 2, 4, 8
// lanes gave 106, 106, 127, 201 us).
inline uint32_t seg_sums_lanes(const Absorber& a) { return a.seg_lanes > 1 ? a.seg_lanes / 2 : 1; }

inline unsigned seg_sums_grid(const Absorber& a) {
    return grid_for(a.g1, (TPB / WAVE) * (WAVE / seg_sums_lanes(a)), 1u << 16);
}


// ---- grounded pseudo-inverse on the GPU (large small super-blocks)
// diag and n_r of the assembled S_SS (len ms*ms + ms) for the host's redundancy decisions.
__global__ void k_extract_diag(const double* __restrict__ S, size_t ms, double* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < ms; k += (size_t)gridDim.x * blockDim.x) {
        out[k] = S[k * ms + k];
        out[ms + k] = S[ms * ms + k];
    }
}
// A[i][j] = S[keep_i][keep_j] sc_i sc_j (unit-diagonal scaling); col = column 0 of A.
__global__ void k_sweep_init(const double* __restrict__ S, size_t ms, const uint32_t* __restrict__ keep,
                             const double* __restrict__ sc, int k, double* __restrict__ A, double* __restrict__ col) {
    const size_t kk = (size_t)k * k;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < kk; e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e / k, l = e % k;
        const double v = S[(size_t)keep[i] * ms + keep[l]] * sc[i] * sc[l];
        A[e] = v;
        if (l == 0) col[i] = v;
    }
}
// Gauss-Jordan sweep of pivot j (Goodnight's operator on a symmetric matrix): with d the current
// diagonal, A[i][l] -= A[i][j] A[j][l] / d, row/column j become A[.][j] / d and A[j][j] = -1/d.
// A pivot d <= 1e-10 is grounded: left unswept (precond/dense.rs:25-37 replaces it by 1e30,
// which removes the direction), so after all pivots -A is the inverse over the swept indices.
// col holds column j before the step; the step writes column j+1 after it into colnext.
__global__ void k_sweep_step(double* __restrict__ A, int k, int j, const double* __restrict__ col,
                             double* __restrict__ colnext, uint8_t* __restrict__ grounded) {
    const double d = col[j];
    const bool g = !(d > 1e-10);
    const double inv = g ? 0.0 : 1.0 / d;
    const size_t kk = (size_t)k * k;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < kk; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / k), l = (int)(e % k);
        double v = A[e];
        if (!g) {
            if (i == j && l == j) v = -inv;
            else if (i == j) v = col[l] * inv;
            else if (l == j) v = col[i] * inv;
            else v -= col[i] * col[l] * inv;
            A[e] = v;
        }
        if (l == j + 1) colnext[i] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) grounded[j] = g;
}
// minv[keep_i][keep_l] = -A[i][l] sc_i sc_l over the swept pivots (0 elsewhere).
__global__ void k_sweep_scatter(const double* __restrict__ A, int k, const uint32_t* __restrict__ keep,
                                const double* __restrict__ sc, const uint8_t* __restrict__ grounded, size_t ms,
                                double* __restrict__ minv, float* __restrict__ minv32) {
    const size_t kk = (size_t)k * k;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < kk; e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e / k, l = e % k;
        const double v = (grounded[i] || grounded[l]) ? 0.0 : -A[e] * sc[i] * sc[l];
        const size_t o = (size_t)keep[i] * ms + keep[l];
        minv[o] = v;
        minv32[o] = (float)v;
    }
}

}  // namespace

// Grounded LDL' of S_SS over the kept coordinates and its explicit inverse
// (precond/dense.rs:25-37, 79-160): unit-diagonal scaling, pivots <= 1e-10 replaced by 1e30.
std::vector<double> grounded_inverse(const std::vector<double>& a, size_t ms, const std::vector<uint8_t>& red,
                                     uint64_t& nkeep) {
    std::vector<uint32_t> keep;
    for (size_t i = 0; i < ms; i++)
        if (!red[i] && a[i * ms + i] > 0.0) keep.push_back((uint32_t)i);
    const size_t k = keep.size();
    nkeep = k;
    std::vector<double> inv(ms * ms, 0.0);
    if (k == 0) return inv;
    std::vector<double> sc(k), L(k * k, 0.0);
    for (size_t i = 0; i < k; i++) sc[i] = 1.0 / std::sqrt(a[(size_t)keep[i] * (ms + 1)]);
    for (size_t i = 0; i < k; i++)
        for (size_t j = 0; j <= i; j++) L[i * k + j] = a[(size_t)keep[i] * ms + keep[j]] * sc[i] * sc[j];
    for (size_t j = 0; j < k; j++) {
        double d = L[j * k + j];
        for (size_t p = 0; p < j; p++) d -= L[j * k + p] * L[j * k + p] * L[p * k + p];
        if (d <= 1e-10) d = 1e30;
        L[j * k + j] = d;
        for (size_t i = j + 1; i < k; i++) {
            double v = L[i * k + j];
            for (size_t p = 0; p < j; p++) v -= L[i * k + p] * L[j * k + p] * L[p * k + p];
            L[i * k + j] = v / d;
        }
    }
    std::vector<double> t(k);
    for (size_t col = 0; col < k; col++) {
        std::fill(t.begin(), t.end(), 0.0);
        t[col] = sc[col];
        for (size_t i = 0; i < k; i++)
            for (size_t p = 0; p < i; p++) t[i] -= L[i * k + p] * t[p];
        for (size_t i = 0; i < k; i++) t[i] /= L[i * k + i];
        for (size_t ii = k; ii-- > 0;)
            for (size_t p = ii + 1; p < k; p++) t[ii] -= L[p * k + ii] * t[p];
        for (size_t i = 0; i < k; i++) inv[(size_t)keep[i] * ms + keep[col]] = t[i] * sc[i];
    }
    return inv;
}


double absorber_sw(Absorber& a) {
    if (!a.sw_ready) {
        download(&a.sw, a.sw_dev.p, sizeof(double), a.st);
        a.sw_ready = true;
    }
    return a.sw;
}

void refresh_weights(Absorber& a, const double* w_rows_dev) {
    if (a.slope) {
        slope_refresh(a, w_rows_dev);
        return;
    }
    hipStream_t st = a.st;
    const size_t M = a.M, g1 = a.g1, m = a.m, ms = a.ms;
    a.weighted = w_rows_dev != nullptr;
    if (a.weighted && w_rows_dev != a.w_rows.p) {
        a.w_rows.reserve(a.n);
        FH_HIP(hipMemcpyAsync(a.w_rows.p, w_rows_dev, a.n * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    const bool unit = a.unit();
    // cell weights
    if (!unit) {
        a.w64.cw.reserve(M);
        a.w32.cw.reserve(M);
        hipLaunchKernelGGL(k_cell_weights, dim3(blocks_for(M)), dim3(TPB), 0, st, a.trivial() ? nullptr : a.cell_obs_off.p,
                           a.weighted ? a.w_rows.p : nullptr, M, a.w64.cw.p, a.w32.cw.p);
        FH_CHECK_LAUNCH();
    } else {
        a.w64.cw.release();
        a.w32.cw.release();
    }
    const double* cw = unit ? nullptr : a.w64.cw.p;
    a.n1.reserve(g1);
    a.w64.inv_n1.reserve(g1);
    a.w32.inv_n1.reserve(g1);
    hipLaunchKernelGGL(k_n1, dim3(blocks_for(g1)), dim3(TPB), 0, st, a.seg_off.p, cw, g1, a.n1.p, a.w64.inv_n1.p,
                       a.w32.inv_n1.p);
    FH_CHECK_LAUNCH();
    {  // the weight sum stays on the device until a readback needs it
        DBuf<double> part(256);
        a.sw_dev.reserve(1);
        hipLaunchKernelGGL(k_partial_sum, dim3(256), dim3(TPB), 0, st, a.n1.p, g1, part.p);
        FH_CHECK_LAUNCH();
        sum_parts_to(part.p, 256, 1, a.sw_dev.p, st);
        a.sw_ready = false;
    }
    a.nr.reserve(std::max<size_t>(m, 1));
    a.diag.reserve(std::max<size_t>(m, 1));
    a.red.reserve(std::max<size_t>(m, 1));
    a.w64.inv_diag.reserve(std::max<size_t>(m, 1));
    a.w32.inv_diag.reserve(std::max<size_t>(m, 1));
    const int nb = (int)a.nb;
    a.w64.lw.resize(nb);
    a.w32.lw.resize(nb);
    const bool dense_small = small_dense_rule(a);
    for (int b = 0; b < nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (bi.small && (unit || dense_small)) continue;  // no level pass over the dense small block
        if (!unit) {
            a.w64.lw[b].reserve(M);
            a.w32.lw[b].reserve(M);
        } else {
            a.w64.lw[b].release();
            a.w32.lw[b].release();
        }
        if (bi.small) {  // level-order weights only
            hipLaunchKernelGGL(k_gather_lw, dim3(blocks_for(M)), dim3(TPB), 0, st, a.bperm[b].p, cw, M, a.w64.lw[b].p,
                               a.w32.lw[b].p);
            FH_CHECK_LAUNCH();
            continue;
        }
        if (!unit) {  // level-order weights first: the level sums then read them in order
            const uint32_t X = a.nbk[b] >= (uint32_t)XCD_COUNT ? (uint32_t)XCD_COUNT : 1u;
            const unsigned gg = (unsigned)(grid_for(M, TPB, 4096) / X * X + X);
            hipLaunchKernelGGL(k_gather_lw_x, dim3(gg), dim3(TPB), 0, st, a.boff[b].p, a.nbk[b], (size_t)bi.g,
                               a.bperm[b].p, cw, X, a.w64.lw[b].p, a.w32.lw[b].p);
            FH_CHECK_LAUNCH();
        }
        // lanes per level: ~4 positions per lane of each (bucket, level) item, since the lanes
        // walk a level's buckets one after another
        uint32_t L = 1;
        while (L < 64 && L * 4.0 < double(a.M) / double(std::max<uint64_t>(bi.g, 1) * std::max<uint32_t>(a.nbk[b], 1)))
            L <<= 1;
        const unsigned grid = grid_for(bi.g, (TPB / WAVE) * (WAVE / L), 1u << 16);
        hipLaunchKernelGGL(k_level_weights, dim3(grid), dim3(TPB), 0, st, a.boff[b].p, a.lseg[b].p, a.nbk[b],
                           unit ? nullptr : a.w64.lw[b].p, a.n1.p, (size_t)bi.g, (size_t)bi.off, (int)L, a.nr.p,
                           a.diag.p, a.red.p, a.w64.inv_diag.p, a.w32.inv_diag.p);
        FH_CHECK_LAUNCH();
    }
    // dense small-row operator: one small block of <= 32 levels whose D and C are not much
    // larger than the cell records a level pass would stream
    a.small_dense = dense_small;
    if (a.small_dense && nb > 1 && a.large_codes.n == 0) {
        a.large_codes.alloc(M * (nb - 1));
        hipLaunchKernelGGL(k_drop_block0, dim3(blocks_for(M * (nb - 1))), dim3(TPB), 0, st, a.cell_codes.p, M, nb,
                           a.large_codes.p);
        FH_CHECK_LAUNCH();
    }
    if (a.small_dense) {
        // E = [D | pad | C | pad], ms rows of pitch sm_ld: every column quad is 16-byte aligned
        const size_t mL = m - ms;
        a.sm_coff = (g1 + 3) / 4 * 4;
        a.sm_ld = a.sm_coff + (mL + 3) / 4 * 4;
        // every refresh rewrites all entries of E but the pads, so a buffer is zeroed when allocated
        const size_t ne = ms * a.sm_ld;
        if (a.w64.esm.n < ne) {
            a.w64.esm.reserve(ne);
            a.w32.esm.reserve(ne);
            FH_HIP(hipMemsetAsync(a.w64.esm.p, 0, ne * sizeof(double), st));
            FH_HIP(hipMemsetAsync(a.w32.esm.p, 0, ne * sizeof(float), st));
        }
        // without prior weights E holds cell counts: a u8 copy when every count fits (the S*p
        // kernels test the device flag and fall back to E64/E32 otherwise)
        a.esm8_bad.reserve(1);
        FH_HIP(hipMemsetAsync(a.esm8_bad.p, 0, sizeof(uint32_t), st));
        if (!a.weighted) {
            if (a.esm8.n < ne) {
                a.esm8.reserve(ne);
                FH_HIP(hipMemsetAsync(a.esm8.p, 0, ne, st));
            }
        } else {
            a.esm8.release();
        }
        for (int b = 0; b < nb; b++) {
            const BlockInfo& bi = a.blocks[b];
            if (bi.small) continue;
            // lanes per level: ~4 positions per lane of each (bucket, level) item (the lanes walk
            // a level's buckets one after another; C2: 86 -> 63 us, a sweep of 1..16 lanes gave
            // 98, 74, 63, 65, 86 us)
            uint32_t L = 1;
            while (L < 64 &&
                   L * 4.0 < double(a.M) / double(std::max<uint64_t>(bi.g, 1) * std::max<uint32_t>(a.nbk[b], 1)))
                L <<= 1;
            const unsigned grid = grid_for(bi.g, (TPB / WAVE) * (WAVE / L), 1u << 16);
            if (ms <= 16)
                hipLaunchKernelGGL(k_level_small_sums<16>, dim3(grid), dim3(TPB), 0, st, a.boff[b].p, a.lother[b].p,
                                   nb - 1, unit ? nullptr : a.w64.lw[b].p, a.nbk[b], (size_t)bi.g, (int)ms, a.sm_ld,
                                   a.sm_coff + (size_t)(bi.off - ms), (int)L, a.w64.esm.p, a.w32.esm.p,
                                   a.weighted ? (uint8_t*)nullptr : a.esm8.p, a.esm8_bad.p);
            else
                hipLaunchKernelGGL(k_level_small_sums<32>, dim3(grid), dim3(TPB), 0, st, a.boff[b].p, a.lother[b].p,
                                   nb - 1, unit ? nullptr : a.w64.lw[b].p, a.nbk[b], (size_t)bi.g, (int)ms, a.sm_ld,
                                   a.sm_coff + (size_t)(bi.off - ms), (int)L, a.w64.esm.p, a.w32.esm.p,
                                   a.weighted ? (uint8_t*)nullptr : a.esm8.p, a.esm8_bad.p);
            FH_CHECK_LAUNCH();
        }
        if (ms > 21)  // the S_SS assembly below does not produce D for these
            hipLaunchKernelGGL(k_seg_level_sums<32>, dim3(seg_sums_grid(a)), dim3(TPB), 0, st, a.seg_off.p,
                               a.cell_codes.p, cw, a.n1.p, g1, nb, (uint32_t)a.blocks[0].off, (int)ms, (double*)nullptr,
                               a.sm_ld, a.w64.esm.p, a.w32.esm.p, a.weighted ? (uint8_t*)nullptr : a.esm8.p,
                               a.esm8_bad.p, (int)seg_sums_lanes(a));
        FH_CHECK_LAUNCH();
    }
    // small super-block
    a.nkeep = 0;
    if (ms > 0) {
        SmallOffs so{};
        so.nb = nb;
        so.nsb = 0;
        for (int b = 0; b < nb; b++)
            if (a.blocks[b].small) so.off[so.nsb++] = (uint32_t)a.blocks[b].off;
        const size_t len = ms * ms + ms;
        DBuf<double> sfull(len);
        std::vector<double> h(len);
        if (so.nsb == 1 && ms <= 21) {
            // S_SS = diag(n_r) - C'C over the scaled level sums; n_r from the extra column
            DBuf<double> C(g1 * (ms + 1));
            double* D64 = a.small_dense ? a.w64.esm.p : nullptr;
            float* D32 = a.small_dense ? a.w32.esm.p : nullptr;
            uint8_t* D8 = a.small_dense && !a.weighted ? a.esm8.p : nullptr;
            if (ms <= 8)
                hipLaunchKernelGGL(k_seg_level_sums<8>, dim3(seg_sums_grid(a)), dim3(TPB), 0, st, a.seg_off.p,
                                   a.cell_codes.p, cw, a.n1.p, g1, nb, so.off[0], (int)ms, C.p, a.sm_ld, D64, D32,
                                   D8, a.esm8_bad.p, (int)seg_sums_lanes(a));
            else if (ms <= 16)
                hipLaunchKernelGGL(k_seg_level_sums<16>, dim3(seg_sums_grid(a)), dim3(TPB), 0, st, a.seg_off.p,
                                   a.cell_codes.p, cw, a.n1.p, g1, nb, so.off[0], (int)ms, C.p, a.sm_ld, D64, D32,
                                   D8, a.esm8_bad.p, (int)seg_sums_lanes(a));
            else
                hipLaunchKernelGGL(k_seg_level_sums<21>, dim3(seg_sums_grid(a)), dim3(TPB), 0, st, a.seg_off.p,
                                   a.cell_codes.p, cw, a.n1.p, g1, nb, so.off[0], (int)ms, C.p, a.sm_ld, D64, D32,
                                   D8, a.esm8_bad.p, (int)seg_sums_lanes(a));
            FH_CHECK_LAUNCH();
            std::vector<double> G((ms + 1) * (ms + 1));
            gram_cols(a, C.p, g1, (int)ms + 1, 0, (int)ms + 1, nullptr, nullptr, G.data());
            for (size_t k = 0; k < ms; k++) {
                h[ms * ms + k] = G[k * (ms + 1) + ms];
                for (size_t l = 0; l < ms; l++) h[k * ms + l] = -G[k * (ms + 1) + l];
                h[k * ms + k] += h[ms * ms + k];
            }
        } else {
            // 64-bit fixed point with every partial sum below 2^62 (terms and sums are <= sw)
            const double fscale = std::ldexp(1.0, 62 - (int)std::ceil(std::log2(std::max(absorber_sw(a), 1.0)) + 1e-12));
            const bool lds = ms <= (size_t)SA_LDS_MAX;
            const unsigned grid = lds ? std::min<unsigned>(blocks_for(g1), 512) : blocks_for(g1);
            const size_t nparts = lds ? grid : 1;
            DBuf<unsigned long long> parts(nparts * len);
            if (!lds) FH_HIP(hipMemsetAsync(parts.p, 0, parts.bytes(), st));
            hipLaunchKernelGGL(k_small_assemble, dim3(grid), dim3(TPB), lds ? len * sizeof(unsigned long long) : 0, st,
                               a.seg_off.p, a.cell_codes.p, cw, a.n1.p, g1, ms, so, fscale, parts.p);
            hipLaunchKernelGGL(k_fixed_to_double, dim3(grid_for(len, TPB, 4096)), dim3(TPB), 0, st, parts.p, nparts, len,
                               1.0 / fscale, sfull.p);
            FH_CHECK_LAUNCH();
        }
        // large blocks stay on the device: diag and n_r come to the host, the grounded inverse is
        // formed by pivot sweeps on the GPU (O(ms^3) work; the host version is for small blocks)
        const bool dev_inverse = !(so.nsb == 1 && ms <= 21) && ms > 64;
        std::vector<double> nrs(ms), dg(ms);
        std::vector<uint8_t> rd(ms);
        std::vector<double> inv;
        if (dev_inverse) {
            DBuf<double> dd(2 * ms);
            hipLaunchKernelGGL(k_extract_diag, dim3(grid_for(ms, TPB, 1024)), dim3(TPB), 0, st, sfull.p, ms, dd.p);
            FH_CHECK_LAUNCH();
            download(h.data(), dd.p, 2 * ms * sizeof(double), st);
            stream_wait(st);
            for (size_t k = 0; k < ms; k++) {
                dg[k] = h[k];
                nrs[k] = h[ms + k];
                rd[k] = dg[k] <= 1e-12 * nrs[k];
            }
            std::vector<uint32_t> keep;
            std::vector<double> sc;
            for (size_t i = 0; i < ms; i++)
                if (!rd[i] && dg[i] > 0.0) {
                    keep.push_back((uint32_t)i);
                    sc.push_back(1.0 / std::sqrt(dg[i]));
                }
            const int k = (int)keep.size();
            a.nkeep = (uint64_t)k;
            a.w64.minv.reserve(ms * ms);
            a.w32.minv.reserve(ms * ms);
            FH_HIP(hipMemsetAsync(a.w64.minv.p, 0, ms * ms * sizeof(double), st));
            FH_HIP(hipMemsetAsync(a.w32.minv.p, 0, ms * ms * sizeof(float), st));
            if (k > 0) {
                const size_t kk = (size_t)k * k;
                DBuf<uint32_t> dkeep(k);
                DBuf<double> dsc(k), A(kk), col(2 * (size_t)k);
                DBuf<uint8_t> grounded(k);
                upload(dkeep.p, keep.data(), k * sizeof(uint32_t), st);
                upload(dsc.p, sc.data(), k * sizeof(double), st);
                const unsigned gs = grid_for(kk, TPB, 8192);
                hipLaunchKernelGGL(k_sweep_init, dim3(gs), dim3(TPB), 0, st, sfull.p, ms, dkeep.p, dsc.p, k, A.p, col.p);
                for (int j = 0; j < k; j++)
                    hipLaunchKernelGGL(k_sweep_step, dim3(gs), dim3(TPB), 0, st, A.p, k, j, col.p + (size_t)(j & 1) * k,
                                       col.p + (size_t)((j + 1) & 1) * k, grounded.p);
                hipLaunchKernelGGL(k_sweep_scatter, dim3(gs), dim3(TPB), 0, st, A.p, k, dkeep.p, dsc.p, grounded.p, ms,
                                   a.w64.minv.p, a.w32.minv.p);
                FH_CHECK_LAUNCH();
            }
        } else {
            if (!(so.nsb == 1 && ms <= 21)) {
                download(h.data(), sfull.p, len * sizeof(double), st);
                stream_wait(st);
            }
            std::vector<double> S(h.begin(), h.begin() + ms * ms);
            for (size_t k = 0; k < ms; k++) {
                nrs[k] = h[ms * ms + k];
                dg[k] = S[k * ms + k];
                rd[k] = dg[k] <= 1e-12 * nrs[k];
            }
            inv = grounded_inverse(S, ms, rd, a.nkeep);
        }
        // the host results in one transfer ([n_r | diag | redundant | inverse]), spread on the device
        const bool with_inv = !dev_inverse;
        std::vector<double> pk(3 * ms + (with_inv ? ms * ms : 0));
        for (size_t k = 0; k < ms; k++) {
            pk[k] = nrs[k];
            pk[ms + k] = dg[k];
            pk[2 * ms + k] = rd[k] ? 1.0 : 0.0;
        }
        if (with_inv) std::copy(inv.begin(), inv.end(), pk.begin() + 3 * ms);
        DBuf<double> dpk(pk.size());
        upload(dpk.p, pk.data(), pk.size() * sizeof(double), st);
        if (a.small_dense) {
            a.w64.nrs.reserve(ms);
            a.w32.nrs.reserve(ms);
        }
        if (with_inv) {
            a.w64.minv.reserve(ms * ms);
            a.w32.minv.reserve(ms * ms);
        }
        hipLaunchKernelGGL(k_small_spread, dim3(grid_for(ms * ms + ms, TPB, 1024)), dim3(TPB), 0, st, dpk.p, ms,
                           with_inv ? 1 : 0, a.nr.p, a.small_dense ? a.w64.nrs.p : nullptr,
                           a.small_dense ? a.w32.nrs.p : nullptr, a.diag.p, a.red.p, a.w64.inv_diag.p,
                           a.w32.inv_diag.p, with_inv ? a.w64.minv.p : nullptr, with_inv ? a.w32.minv.p : nullptr);
        FH_CHECK_LAUNCH();
    }
    // the u8 counts serve S*p when every count fit (one flag read per refresh, not per product)
    a.esm8_ok = false;
    if (a.small_dense && !a.weighted && a.esm8.n) {
        uint32_t bad = 1;
        download(&bad, a.esm8_bad.p, sizeof(bad), st);
        a.esm8_ok = bad == 0;
    }
    facs_after_refresh(a);
}

}  // namespace fh
