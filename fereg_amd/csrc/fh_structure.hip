// Cell structure of the intercept-only FE design on the GPU (bit-exact with
// crates/fereg-core/src/structure.rs:356-527):
//   rows sorted by (FE1 code, block codes in block order, obs) with a stable LSD radix sort of
//   packed keys (obs order is the stable tie-break), cells = runs of equal tuples found by a
//   flag + exclusive-scan pass, seg_off from the cells' FE1 codes, sum_seg s^2 of the small
//   coordinates with a per-wave LDS bitmap, and the level-major transpose of every large block
//   (a second radix sort of the cells by code, structure.rs:144-153).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "fh_internal.hpp"
#include "fh_sort.hpp"

namespace fh {

namespace {

struct CodeSet {
    const uint32_t* c[NB_MAX + 1];
    uint32_t shift[NB_MAX + 1];
    int nq;
};

__global__ void k_hist(const uint32_t* __restrict__ codes, size_t n, uint32_t* __restrict__ cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[codes[i]], 1u);
}

// part[blockIdx.x] = the workgroup's sum of v (no atomics: one atomic per workgroup on a single
// address serializes over thousands of workgroups). Every thread calls it; blockDim.x <= 1024.
__device__ void block_sum_u64(unsigned long long v, unsigned long long* part) {
    __shared__ unsigned long long wpart[1024 / WAVE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    if (lane_id() == 0) wpart[threadIdx.x / WAVE] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (unsigned w = 0; w < blockDim.x / WAVE; w++) t += wpart[w];
        part[blockIdx.x] = t;
    }
}

__global__ void k_sumsq(const uint32_t* __restrict__ v, size_t g, unsigned long long* part) {
    unsigned long long s = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < g; i += (size_t)gridDim.x * blockDim.x)
        s += (unsigned long long)v[i] * v[i];
    block_sum_u64(s, part);
}

// key[i] = packed fields of row vals[i] (fields of the word: most significant first).
__global__ void k_pack(CodeSet cs, const uint32_t* __restrict__ vals, size_t n, uint64_t* __restrict__ keys) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t o = vals ? vals[i] : (uint32_t)i;
        uint64_t k = 0;
        for (int q = 0; q < cs.nq; q++) k |= (uint64_t)cs.c[q][o] << cs.shift[q];
        keys[i] = k;
    }
}

// FE1 in kept order: st[0] |= some code below its predecessor's, or more than 64 above it (the
// fast path is off); otherwise off[l] = the first row of level l (n + 1 threads, empty levels
// take the next level's first row, off[g1] = n), each offset written by the row where the code
// steps past it -- at most 64 writes per row, whatever the input.
__global__ void k_fe1_offsets(const uint32_t* __restrict__ fe1, size_t n, uint64_t g1, uint32_t* __restrict__ off,
                              uint32_t* __restrict__ st) {
    uint32_t bad = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (size_t)gridDim.x * blockDim.x) {
        const int64_t prev = i == 0 ? -1 : (int64_t)fe1[i - 1];
        const int64_t cur = i == n ? (int64_t)g1 : (int64_t)fe1[i];
        if (cur < prev || cur - prev > 64) {
            bad = 1;
            continue;
        }
        for (int64_t l = prev + 1; l <= cur; l++) off[l] = (uint32_t)i;
    }
    if (__any(bad) && lane_id() == 0) atomicOr(&st[0], 1u);
}

// part[blockIdx.x] = the longest level, off[s + 1] - off[s], among the workgroup's levels (no
// atomics: one contended atomic per workgroup costs more than the whole read).
__global__ __launch_bounds__(256) void k_max_len(const uint32_t* __restrict__ off, size_t g1,
                                                 uint32_t* __restrict__ part) {
    __shared__ uint32_t red[256 / WAVE];
    uint32_t m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < g1; i += (size_t)gridDim.x * blockDim.x)
        m = max(m, off[i + 1] - off[i]);
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, WAVE));
    if (lane_id() == 0) red[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < blockDim.x / WAVE; w++) m = max(m, red[w]);
        part[blockIdx.x] = m;
    }
}

// Rows sorted by (FE1, block codes, row) when FE1 ascends in kept order and no level has more than
// L rows: L lanes per FE1 level, lane j holding the level's row off[s] + j with key = its packed
// block codes; its position is its rank among the level's rows by (key, row), the order the
// stable radix sort of the packed (FE1, blocks) word gives.
template <int L>
__global__ __launch_bounds__(256) void k_level_rank_sort(CodeSet cs, const uint32_t* __restrict__ off, uint64_t g1,
                                                         uint32_t* __restrict__ pi) {
    constexpr int GPW = WAVE / L;
    const int lane = lane_id(), j = lane % L, base = lane - j;
    const uint64_t ngroups = (uint64_t)gridDim.x * (blockDim.x / WAVE) * GPW;
    const uint64_t g0 = ((uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * GPW + lane / L;
    const uint64_t rounds = (g1 + ngroups - 1) / ngroups;
    for (uint64_t rd = 0; rd < rounds; rd++) {  // the same rounds in every lane (the shuffles)
        const uint64_t s = g0 + rd * ngroups;
        const uint32_t lo = s < g1 ? off[s] : 0, len = s < g1 ? off[s + 1] - lo : 0;
        uint64_t key = ~0ull;
        if ((uint32_t)j < len) {
            key = 0;
            for (int q = 0; q < cs.nq; q++) key |= (uint64_t)cs.c[q][lo + j] << cs.shift[q];
        }
        const uint32_t khi = (uint32_t)(key >> 32), klo = (uint32_t)key;
        uint32_t rank = 0;
#pragma unroll
        for (int l = 0; l < L; l++) {
            const uint32_t ohi = (uint32_t)__shfl((int)khi, base + l, WAVE);
            const uint32_t olo = (uint32_t)__shfl((int)klo, base + l, WAVE);
            const uint64_t ok = ((uint64_t)ohi << 32) | olo;
            rank += ((uint32_t)l < len) && (ok < key || (ok == key && l < j)) ? 1u : 0u;
        }
        if ((uint32_t)j < len) pi[lo + rank] = lo + j;
    }
}

__global__ void k_iota(uint32_t* __restrict__ v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}

// Row i starts a cell when its (FE1, block codes) tuple differs from row i-1's.
// flags[i] = 1 where sorted row i starts a cell; part[blockIdx.x] = the workgroup's count of them
// (their sum is the cell count M; M = n means every row is its own cell).
__global__ __launch_bounds__(256) void k_cell_flags(CodeSet cs, const uint32_t* __restrict__ pi, size_t n,
                                                    uint32_t* __restrict__ flags, uint32_t* __restrict__ part) {
    uint32_t cnt = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t f = 1;
        if (i > 0) {
            uint32_t o = pi[i], p = pi[i - 1];
            f = 0;
            for (int q = 0; q < cs.nq; q++) f |= cs.c[q][o] != cs.c[q][p];
        }
        flags[i] = f;
        cnt += f;
    }
    __shared__ uint32_t red[256 / WAVE];
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) cnt += (uint32_t)__shfl_xor((int)cnt, o, WAVE);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (unsigned w = 0; w < blockDim.x / WAVE; w++) t += red[w];
        part[blockIdx.x] = t;
    }
}

__global__ void k_cell_start(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ scan, size_t n,
                             uint32_t* __restrict__ start) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (flags[i]) start[scan[i]] = (uint32_t)i;
}

// Per cell: FE1 level and block codes of its first row (row c when start is null).
__global__ void k_cell_codes(CodeSet cs, int has_fe1, const uint32_t* __restrict__ start,
                             const uint32_t* __restrict__ pi, size_t M, uint32_t* __restrict__ cell_seg,
                             uint32_t* __restrict__ cell_codes) {
    const int nb = cs.nq - 1;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (size_t)gridDim.x * blockDim.x) {
        uint32_t o = pi[start ? start[c] : (uint32_t)c];  // no start: every row its own cell
        cell_seg[c] = has_fe1 ? cs.c[0][o] : 0u;
        for (int b = 0; b < nb; b++) cell_codes[c * nb + b] = cs.c[1 + b][o];
    }
}

// off[k] = first position whose (ascending) key is >= k, for k in 0..=g.
__global__ void k_offsets(const uint32_t* __restrict__ key, size_t M, size_t g, uint32_t* __restrict__ off) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (size_t)gridDim.x * blockDim.x) {
        uint32_t s = key[c];
        int64_t prev = c ? (int64_t)key[c - 1] : -1;
        for (int64_t k = prev + 1; k <= (int64_t)s; k++) off[k] = (uint32_t)c;
        if (c == M - 1)
            for (size_t k = (size_t)s + 1; k <= g; k++) off[k] = (uint32_t)M;
    }
}

__global__ void k_offsets_u64(const uint64_t* __restrict__ key, size_t M, size_t g, uint32_t* __restrict__ off,
                              uint64_t mask = ~0ull) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (size_t)gridDim.x * blockDim.x) {
        uint64_t s = key[c] & mask;
        int64_t prev = c ? (int64_t)(key[c - 1] & mask) : -1;
        for (int64_t k = prev + 1; k <= (int64_t)s; k++) off[k] = (uint32_t)c;
        if (c == M - 1)
            for (size_t k = (size_t)s + 1; k <= g; k++) off[k] = (uint32_t)M;
    }
}

// sum over segments of (distinct small coordinates)^2: one wave per segment, LDS bitmap.
constexpr int SW_WG = 256;
__global__ __launch_bounds__(SW_WG) void k_small_work(const uint32_t* __restrict__ seg_off,
                                                      const uint32_t* __restrict__ cell_codes, size_t g1, int nb,
                                                      int nsb, const uint32_t* __restrict__ boff_small, int words,
                                                      unsigned long long* out) {
    extern __shared__ uint32_t bm[];
    const int w = threadIdx.x / WAVE, l = lane_id();
    uint32_t* my = bm + (size_t)w * words;
    unsigned long long acc = 0;
    const size_t nwaves = (size_t)gridDim.x * (SW_WG / WAVE);
    for (size_t s = (size_t)blockIdx.x * (SW_WG / WAVE) + w; s < g1; s += nwaves) {
        for (int i = l; i < words; i += WAVE) my[i] = 0;
        __builtin_amdgcn_wave_barrier();
        uint32_t c0 = seg_off[s], c1 = seg_off[s + 1];
        for (uint32_t c = c0 + l; c < c1; c += WAVE)
            for (int b = 0; b < nsb; b++) {
                uint32_t k = boff_small[b] + cell_codes[(size_t)c * nb + b];
                atomicOr(&my[k >> 5], 1u << (k & 31));
            }
        __builtin_amdgcn_wave_barrier();
        uint32_t d = 0;
        for (int i = l; i < words; i += WAVE) d += __popc(my[i]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, WAVE);
        acc += (unsigned long long)d * d;
        __builtin_amdgcn_wave_barrier();
    }
    if (l != 0) acc = 0;  // every lane holds the wave's sum
    block_sum_u64(acc, out);
}

// One small block (block 0, codes ascending within a segment): distinct = runs of equal codes.
// L lanes per segment read its cells coalesced; a cell starts a run when it is the segment's
// first or its code differs from the previous cell's.
__global__ void k_small_work1(const uint32_t* __restrict__ seg_off, const uint32_t* __restrict__ cell_codes, size_t g1,
                              int nb, int L, unsigned long long* out) {
    const int l = lane_id();
    const int spw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwaves = (size_t)gridDim.x * (blockDim.x / WAVE);
    unsigned long long acc = 0;
    for (size_t base = ((size_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * spw; base < g1;
         base += nwaves * spw) {
        const size_t s = base + gl;
        const bool sv = s < g1;
        const uint32_t c0 = sv ? seg_off[s] : 0, c1 = sv ? seg_off[s + 1] : 0;
        uint32_t d = 0;
        for (uint32_t c = c0 + li; c < c1; c += L)
            d += (c == c0 || cell_codes[(size_t)c * nb] != cell_codes[(size_t)(c - 1) * nb]) ? 1u : 0u;
        for (int o = L >> 1; o > 0; o >>= 1) d += (uint32_t)__shfl_xor((int)d, o, WAVE);
        if (li == 0) acc += (unsigned long long)d * d;
    }
    block_sum_u64(acc, out);
}

__global__ void k_level_keys(const uint32_t* __restrict__ cell_codes, size_t M, int nb, int b,
                             uint64_t* __restrict__ keys) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (size_t)gridDim.x * blockDim.x)
        keys[c] = cell_codes[c * nb + b];
}

// keys[c] = bucket(s(c)) * g + code_b(c): the (bucket, level) item of every cell, in the low
// `ib` bits; with `pay` set, s(c) and the codes of the other blocks ride in the bits above
// (the sort only looks at the low ib bits), so the level records come out of the sorted keys
// without a gather. *unsorted |= the items are not ascending in cell order.
struct PayloadLayout {
    int ib, sb;                 // item bits, segment bits
    int ob[NB_MAX];             // bits of each other block's code (block order without b)
    int pay;                    // payload present
};

__device__ __forceinline__ uint64_t bucket_key(const uint32_t* __restrict__ cell_codes,
                                               const uint32_t* __restrict__ cell_seg, size_t c, int nb, int b,
                                               uint32_t bw, uint64_t g, const PayloadLayout& pl) {
    const uint32_t sg = cell_seg[c];
    uint64_t k = (uint64_t)(sg / bw) * g + cell_codes[c * nb + b];
    if (pl.pay) {
        int at = pl.ib;
        k |= (uint64_t)sg << at;
        at += pl.sb;
        int o = 0;
        for (int bb = 0; bb < nb; bb++) {
            if (bb == b) continue;
            k |= (uint64_t)cell_codes[c * nb + bb] << at;
            at += pl.ob[o++];
        }
    }
    return k;
}

// keys[c] = (bucket, level) item of cell c with its payload; *unsorted (when given) |= the items
// are not in cell order.
__global__ void k_bucket_keys(const uint32_t* __restrict__ cell_codes, const uint32_t* __restrict__ cell_seg, size_t M,
                              int nb, int b, uint32_t bw, uint64_t g, PayloadLayout pl, uint64_t* __restrict__ keys,
                              uint32_t* __restrict__ unsorted) {
    const uint64_t mask = pl.ib >= 64 ? ~0ull : ((1ull << pl.ib) - 1);
    bool bad = false;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (size_t)gridDim.x * blockDim.x) {
        const uint64_t k = bucket_key(cell_codes, cell_seg, c, nb, b, bw, g, pl);
        keys[c] = k;
        if (unsorted && c > 0) {
            const uint32_t sp = cell_seg[c - 1];
            const uint64_t kp = (uint64_t)(sp / bw) * g + cell_codes[(c - 1) * nb + b];
            bad |= (k & mask) < kp;
        }
    }
    if (!unsorted) return;
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (__ballot(bad) != 0 && lane_id() == 0) atomicOr(&any, 1);
    __syncthreads();
    if (threadIdx.x == 0 && any) atomicOr(unsorted, 1u);
}

// Level records from the sorted payload keys (coalesced, no gather).
__global__ void k_level_records_keys(const uint64_t* __restrict__ keys, size_t M, int nother, PayloadLayout pl,
                                     uint32_t* __restrict__ lseg, uint32_t* __restrict__ lother) {
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += (size_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[j];
        int at = pl.ib;
        lseg[j] = (uint32_t)((k >> at) & ((1ull << pl.sb) - 1));
        at += pl.sb;
        for (int o = 0; o < nother; o++) {
            lother[j * nother + o] = (uint32_t)((k >> at) & ((1ull << pl.ob[o]) - 1));
            at += pl.ob[o];
        }
    }
}

__global__ void k_level_records(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ cell_seg,
                                const uint32_t* __restrict__ cell_codes, size_t M, int nb, int b,
                                uint32_t* __restrict__ lseg, uint32_t* __restrict__ lother) {
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += (size_t)gridDim.x * blockDim.x) {
        uint32_t c = perm[j];
        lseg[j] = cell_seg[c];
        int k = 0;
        for (int bb = 0; bb < nb; bb++)
            if (bb != b) lother[j * (nb - 1) + k++] = cell_codes[(size_t)c * nb + bb];
    }
}

// E of eliminating term f against one other term o: sum_s d_{s,o}^2 (structure.rs:325-354),
// from the sorted (code_f, code_o) pairs: d_s = distinct pairs of level s.
__global__ void k_pair_keys(const uint32_t* __restrict__ cf, const uint32_t* __restrict__ co, int bits_o, size_t n,
                            uint64_t* __restrict__ keys) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        keys[i] = ((uint64_t)cf[i] << bits_o) | co[i];
}
__global__ void k_distinct(const uint64_t* __restrict__ keys, size_t n, int bits_o, uint32_t* __restrict__ d) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (i == 0 || keys[i] != keys[i - 1]) atomicAdd(&d[keys[i] >> bits_o], 1u);
}

constexpr unsigned TPB = 256;
inline unsigned blocks_for(size_t n) { return grid_for(n, TPB, 8192); }

// Lanes per group: the next power of two of the mean length, within 1..64.
inline uint32_t lanes(double mean) {
    uint32_t L = 1;
    while (L < 64 && L < mean) L <<= 1;
    return L;
}
// Lanes per segment of the segment passes: about 4 cells per lane (the gathers of a lane are
// dependent on its code loads, so shorter chains keep more of them in flight).
inline uint32_t seg_lanes_for(size_t M, size_t g1) { return lanes(g1 ? double(M) / double(g1) / 4.0 : 1.0); }

template <class T>
T read_scalar(const T* dev, hipStream_t st) {
    T v;
    download(&v, dev, sizeof(T), st);
    stream_wait(st);
    return v;
}

// The sum of n per-workgroup partials (exact: integers), read back.
static unsigned long long read_part_sum(const DBuf<unsigned long long>& part, size_t n, hipStream_t st) {
    std::vector<unsigned long long> h(n);
    download(h.data(), part.p, n * sizeof(unsigned long long), st);
    unsigned long long t = 0;
    for (unsigned long long v : h) t += v;
    return t;
}

}  // namespace

// choose_elimination (structure.rs:256-313), with the histogram and E sums on the GPU.
static int64_t choose_fe1(Absorber& a, const std::vector<const uint32_t*>& dc) {
    const auto& g = a.term_g;
    const size_t nt = g.size(), n = a.n;
    int64_t f0 = -1;
    for (size_t q = 0; q < nt; q++)
        if (f0 < 0 || g[q] > g[f0]) f0 = (int64_t)q;  // fe1_index, absorber.rs:99-108
    a.norm_fe1 = f0;
    if (f0 < 0 || nt < 3) return f0;
    std::vector<size_t> cands;
    for (size_t q = 0; q < nt; q++)
        if (g[q] * 4 >= g[f0]) cands.push_back(q);
    if (cands.size() < 2) return f0;
    hipStream_t st = a.st;
    DBuf<uint32_t> cnt(std::max<uint64_t>(g[f0], 1));
    uint64_t gmax = g[f0];
    for (size_t q : cands) gmax = std::max<uint64_t>(gmax, g[q]);
    DBuf<unsigned long long> acc(blocks_for(gmax));
    FH_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), st));
    hipLaunchKernelGGL(k_hist, dim3(blocks_for(n)), dim3(TPB), 0, st, dc[f0], n, cnt.p);
    hipLaunchKernelGGL(k_sumsq, dim3(blocks_for(g[f0])), dim3(TPB), 0, st, cnt.p, (size_t)g[f0], acc.p);
    FH_CHECK_LAUNCH();
    unsigned long long ssq = read_part_sum(acc, blocks_for(g[f0]), st);
    if ((nt - 1) * ssq <= 64ull * n) return f0;
    DBuf<uint64_t> keys(n), kalt(n);
    DBuf<uint32_t> vals(n), valt(n);
    int64_t best = -1;
    unsigned long long bestE = 0;
    for (size_t f : cands) {
        unsigned long long e = 0;
        DBuf<uint32_t> d(std::max<uint64_t>(g[f], 1));
        for (size_t o = 0; o < nt; o++) {
            if (o == f) continue;
            int bo = (int)bits_for(g[o]);
            hipLaunchKernelGGL(k_pair_keys, dim3(blocks_for(n)), dim3(TPB), 0, st, dc[f], dc[o], bo, n, keys.p);
            hipLaunchKernelGGL(k_iota, dim3(blocks_for(n)), dim3(TPB), 0, st, vals.p, n);
            FH_CHECK_LAUNCH();
            radix_sort_pairs(keys, vals, kalt, valt, n, (int)(bits_for(g[f]) + bo), st);
            FH_HIP(hipMemsetAsync(d.p, 0, d.bytes(), st));
            hipLaunchKernelGGL(k_distinct, dim3(blocks_for(n)), dim3(TPB), 0, st, keys.p, n, bo, d.p);
            hipLaunchKernelGGL(k_sumsq, dim3(blocks_for(g[f])), dim3(TPB), 0, st, d.p, (size_t)g[f], acc.p);
            FH_CHECK_LAUNCH();
            e += read_part_sum(acc, blocks_for(g[f]), st);
        }
        // key (E, Reverse(g), f): smaller E, then more levels, then lower index
        if (best < 0 || e < bestE || (e == bestE && g[f] > g[best])) {
            best = (int64_t)f;
            bestE = e;
        }
    }
    return best;
}

// Rows not yet in (FE1, blocks) order but with FE1 ascending in kept order and no FE1 level
// longer than 64 rows (panels stored unit by unit, periods unordered): each level's rows are
// The row order of a single-word key: true when the packed (FE1, blocks) keys are already sorted
// (one check and readback); else, when FE1 ascends in kept order with no level above 64 rows, the
// rows are sorted per level in registers by their block codes (k_level_rank_sort), which replaces
// the global radix sort (one more pass and readback). Returns true when pi holds the sorted order.
static bool order_single_word(Absorber& a, const std::vector<const uint32_t*>& dc, const uint64_t* keys,
                              const std::vector<const uint32_t*>& fields, const std::vector<unsigned>& fbits,
                              hipStream_t st) {
    const size_t n = a.n;
    constexpr unsigned ML_BLOCKS = 256;
    DBuf<uint32_t> info(2 + ML_BLOCKS);
    check_sorted_u64(keys, n, info.p, st);
    if (read_scalar(info.p, st) == 0) return true;  // already in order
    if (a.fe1 < 0 || fields.size() < 2) return false;
    DBuf<uint32_t> off(a.g1 + 1);
    FH_HIP(hipMemsetAsync(info.p + 1, 0, sizeof(uint32_t), st));
    const unsigned nml = grid_for(a.g1, TPB, ML_BLOCKS);
    hipLaunchKernelGGL(k_fe1_offsets, dim3(blocks_for(n + 1)), dim3(TPB), 0, st, dc[a.fe1], n, (uint64_t)a.g1, off.p,
                       info.p + 1);
    hipLaunchKernelGGL(k_max_len, dim3(nml), dim3(TPB), 0, st, off.p, (size_t)a.g1, info.p + 2);
    FH_CHECK_LAUNCH();
    std::vector<uint32_t> hs(1 + nml);
    download(hs.data(), info.p + 1, hs.size() * sizeof(uint32_t), st);
    if (hs[0] != 0) return false;
    uint32_t maxlen = 0;
    for (unsigned b = 0; b < nml; b++) maxlen = std::max(maxlen, hs[1 + b]);
    if (maxlen > 64) return false;
    CodeSet cb{};
    unsigned sh = 0;
    cb.nq = (int)fields.size() - 1;
    for (size_t k = fields.size(); k-- > 1;) {
        cb.c[k - 1] = fields[k];
        cb.shift[k - 1] = sh;
        sh += fbits[k];
    }
    auto go = [&](auto lc) {
        constexpr int L = decltype(lc)::value;
        hipLaunchKernelGGL((k_level_rank_sort<L>), dim3(grid_for(a.g1, (TPB / WAVE) * (WAVE / L), 1u << 15)), dim3(TPB),
                           0, st, cb, off.p, (uint64_t)a.g1, a.pi.p);
    };
    if (maxlen <= 16) go(std::integral_constant<int, 16>{});
    else if (maxlen <= 32) go(std::integral_constant<int, 32>{});
    else go(std::integral_constant<int, 64>{});
    FH_CHECK_LAUNCH();
    return true;
}

void build_structure(Absorber& a, const std::vector<const uint32_t*>& dc) {
    hipStream_t st = a.st;
    const size_t n = a.n, nt = a.nterms;
    a.fe1 = choose_fe1(a, dc);
    a.g1 = a.fe1 >= 0 ? a.term_g[a.fe1] : 1;

    // small/large classification: ascending (G, q) while the running total fits (structure.rs:366-385)
    std::vector<int64_t> others;
    for (size_t q = 0; q < nt; q++)
        if ((int64_t)q != a.fe1) others.push_back((int64_t)q);
    require(others.size() <= (size_t)NB_MAX, "at most 9 absorbed terms are supported");
    std::vector<int64_t> by_g = others;
    std::stable_sort(by_g.begin(), by_g.end(), [&](int64_t x, int64_t y) {
        return a.term_g[x] != a.term_g[y] ? a.term_g[x] < a.term_g[y] : x < y;
    });
    std::vector<bool> small_set(nt, false);
    uint64_t total = 0;
    for (int64_t q : by_g) {
        if (total + a.term_g[q] > a.small_limit) break;
        total += a.term_g[q];
        small_set[q] = true;
    }
    a.blocks.clear();
    uint64_t off = 0;
    for (int pass = 0; pass < 2; pass++)
        for (int64_t q : others) {
            if ((pass == 0) != small_set[q]) continue;
            a.blocks.push_back(BlockInfo{q, a.term_g[q], off, small_set[q]});
            off += a.term_g[q];
        }
    a.nb = a.blocks.size();
    a.m = off;
    a.ms = 0;
    for (auto& b : a.blocks)
        if (b.small) a.ms += b.g;
    const int nb = (int)a.nb;
    require(a.m < (1ull << 32) && n < (1ull << 32), "sizes beyond 2^32 rows or levels are not supported");

    // field list: FE1 then blocks (most significant first)
    std::vector<const uint32_t*> fields;
    std::vector<unsigned> fbits;
    DBuf<uint32_t> zeros;
    if (a.fe1 >= 0) {
        fields.push_back(dc[a.fe1]);
        fbits.push_back(bits_for(a.g1));
    }
    for (auto& b : a.blocks) {
        fields.push_back(dc[b.term]);
        fbits.push_back(bits_for(b.g));
    }

    // rows sorted by the fields, obs as the stable tie-break: words of <= 64 bits from the least
    // significant end, each sorted stably in turn (LSD over words).
    a.pi.alloc(std::max<size_t>(n, 1));
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(n)), dim3(TPB), 0, st, a.pi.p, n);
    FH_CHECK_LAUNCH();
    if (n > 1 && !fields.empty()) {
        // group fields into words
        std::vector<std::pair<size_t, size_t>> words;  // [begin, end) field ranges
        size_t e = fields.size();
        while (e > 0) {
            size_t b = e;
            unsigned w = 0;
            while (b > 0 && w + fbits[b - 1] <= 64) { w += fbits[b - 1]; b--; }
            require(b < e, "a code wider than 64 bits");
            words.push_back({b, e});
            e = b;
        }
        DBuf<uint64_t> keys(n), kalt(n);
        DBuf<uint32_t> valt(n);
        bool first = true;
        for (auto [b, e2] : words) {
            CodeSet cs{};
            unsigned sh = 0, width = 0;
            cs.nq = (int)(e2 - b);
            for (size_t k = e2; k-- > b;) {
                cs.c[k - b] = fields[k];
                cs.shift[k - b] = sh;
                sh += fbits[k];
            }
            width = sh;
            if (width == 0) continue;
            hipLaunchKernelGGL(k_pack, dim3(blocks_for(n)), dim3(TPB), 0, st, cs, first ? nullptr : a.pi.p, n, keys.p);
            FH_CHECK_LAUNCH();
            bool sorted = false;
            if (first && words.size() == 1) sorted = order_single_word(a, dc, keys.p, fields, fbits, st);
            if (!sorted) radix_sort_pairs(keys, a.pi, kalt, valt, n, (int)width, st);
            first = false;
        }
    }

    // cells: flags, exclusive scan, starts, codes, segments
    CodeSet all{};
    all.nq = 0;
    if (a.fe1 >= 0) all.c[all.nq++] = dc[a.fe1];
    else {
        zeros.alloc(std::max<size_t>(n, 1));
        FH_HIP(hipMemsetAsync(zeros.p, 0, zeros.bytes(), st));
        all.c[all.nq++] = zeros.p;
    }
    for (auto& b : a.blocks) all.c[all.nq++] = dc[b.term];
    // the cell count from per-workgroup flag counts; when every row is its own cell (distinct key
    // tuples, C2: M = n = 1e7) the scan and the start pass are skipped (cell c = sorted row c)
    DBuf<uint32_t> flags(std::max<size_t>(n, 1));
    const unsigned fgrid = n ? blocks_for(n) : 0;
    DBuf<uint32_t> fpart(std::max<unsigned>(fgrid, 1));
    if (n) {
        hipLaunchKernelGGL(k_cell_flags, dim3(fgrid), dim3(TPB), 0, st, all, a.pi.p, n, flags.p, fpart.p);
        FH_CHECK_LAUNCH();
        std::vector<uint32_t> hp(fgrid);
        download(hp.data(), fpart.p, fgrid * sizeof(uint32_t), st);
        uint64_t tot = 0;
        for (uint32_t v : hp) tot += v;
        a.M = tot;
    } else {
        a.M = 0;
    }
    const size_t M = a.M;
    DBuf<uint32_t> start;
    if (M != n) {
        DBuf<uint32_t> scan(n);
        start.alloc(M + 1);
        exclusive_scan_u32(flags.p, scan.p, n, nullptr, st);
        hipLaunchKernelGGL(k_cell_start, dim3(blocks_for(n)), dim3(TPB), 0, st, flags.p, scan.p, n, start.p);
        FH_CHECK_LAUNCH();
    }
    a.cell_seg.alloc(std::max<size_t>(M, 1));
    a.cell_codes.alloc(std::max<size_t>(M * nb, 1));
    hipLaunchKernelGGL(k_cell_codes, dim3(blocks_for(M)), dim3(TPB), 0, st, all, (int)(a.fe1 >= 0),
                       M != n ? (const uint32_t*)start.p : nullptr, a.pi.p, M, a.cell_seg.p, a.cell_codes.p);
    FH_CHECK_LAUNCH();
    a.seg_off.alloc(a.g1 + 1);
    FH_HIP(hipMemsetAsync(a.seg_off.p, 0, a.seg_off.bytes(), st));
    if (M > 0) {
        hipLaunchKernelGGL(k_offsets, dim3(blocks_for(M)), dim3(TPB), 0, st, a.cell_seg.p, M, (size_t)a.g1, a.seg_off.p);
        FH_CHECK_LAUNCH();
    }
    if (M != n) {
        uint32_t nn = (uint32_t)n;
        upload(start.p + M, &nn, sizeof(uint32_t), st);
        a.cell_obs_off = std::move(start);
    } else {
        a.cell_obs_off.release();
    }

    // sum_seg s^2 of the small coordinates, then the demotion rule (structure.rs:521-525); the
    // partials are read back while the large blocks' transposes (which the rule does not change)
    // run on the stream
    a.small_work = 0;
    int nsb = 0;
    for (auto& b : a.blocks) nsb += b.small ? 1 : 0;
    const bool want_work = a.ms > 0 && M > 0;
    unsigned work_grid = 0;
    DBuf<unsigned long long> work_part;
    DBuf<uint32_t> dboff;
    if (want_work) {
        int words = (int)((a.ms + 31) / 32);
        require(words * 4 * (SW_WG / WAVE) <= 64 * 1024, "small_limit above 131072 levels is not supported");
        std::vector<uint32_t> boff(nsb);
        for (int b = 0; b < nsb; b++) boff[b] = (uint32_t)a.blocks[b].off;
        dboff.alloc(nsb);
        upload(dboff.p, boff.data(), nsb * sizeof(uint32_t), st);
        work_part.alloc(4096);
        if (nsb == 1) {
            const int L = (int)seg_lanes_for(M, a.g1);
            work_grid = grid_for(a.g1, (TPB / WAVE) * (WAVE / L), 2048);
            hipLaunchKernelGGL(k_small_work1, dim3(work_grid), dim3(TPB), 0, st, a.seg_off.p, a.cell_codes.p,
                               (size_t)a.g1, nb, L, work_part.p);
        } else {
            work_grid = grid_for(a.g1, SW_WG / WAVE, 4096);
            hipLaunchKernelGGL(k_small_work, dim3(work_grid), dim3(SW_WG), words * 4 * (SW_WG / WAVE), st,
                               a.seg_off.p, a.cell_codes.p, (size_t)a.g1, nb, nsb, dboff.p, words, work_part.p);
        }
        FH_CHECK_LAUNCH();
    }
    AsyncRead work_read(std::max<size_t>(work_grid, 1) * sizeof(unsigned long long));
    if (want_work) work_read.start(work_part.p, st);

    // bucketed level-major transpose of every block: the large blocks now, the small ones once the
    // rule has decided whether they stay small
    a.lperm.clear(); a.loff.clear(); a.lseg.clear(); a.lother.clear(); a.bperm.clear(); a.boff.clear();
    a.lperm.resize(nb); a.loff.resize(nb); a.lseg.resize(nb); a.lother.resize(nb); a.bperm.resize(nb);
    a.boff.resize(nb);
    a.nbk.assign(nb, 1);
    a.bwidth.assign(nb, 0);
    a.blanes.assign(nb, 1);
    a.llanes.assign(nb, 1);
    for (int b = 0; b < nb; b++)
        if (!a.blocks[b].small) build_block_transpose(a, b);
    if (want_work) {
        const unsigned long long* hp = static_cast<const unsigned long long*>(work_read.wait());
        for (unsigned i = 0; i < work_grid; i++) a.small_work += hp[i];
    }
    if (want_work && a.small_work > 64ull * M) {
        for (auto& b : a.blocks) b.small = false;
        a.ms = 0;
        for (int b = 0; b < nb; b++)
            if (a.bperm[b].n == 0) build_block_transpose(a, b);  // the demoted blocks
    } else {
        // the single small block of the dense small-row operator never runs a level pass; its
        // transpose is built only on request (ensure_block_transpose)
        const bool skip_small = small_dense_rule(a);
        for (int b = 0; b < nb; b++)
            if (a.blocks[b].small && !skip_small) build_block_transpose(a, b);
    }
    a.seg_lanes = seg_lanes_for(M, a.g1);
}

// Bucketed level-major transpose of block b: bperm, boff, lseg, lother (and nbk, bwidth, lanes).
void build_block_transpose(Absorber& a, int b) {
    hipStream_t st = a.st;
    StreamScope ss(st);
    const size_t M = a.M;
    const int nb = (int)a.nb;
    const uint64_t g = a.blocks[b].g;
    // FE1 buckets of bw consecutive segments. Configured: 2^bucket_log2. Automatic: one
    // bucket per XCD once the FE1 range is large (the level passes then keep each XCD's
    // gathers of segment sums within an eighth of them, see the level pass). Long levels (a
    // few levels over many cells) are cut further, to about ITEM_CELLS cells per (bucket,
    // level) item, so that every item fills a wave.
    constexpr double ITEM_CELLS = 2048.0;
    const uint64_t g1 = std::max<uint64_t>(a.g1, 1);
    uint64_t bw = g1;
    if (a.bucket_log2 > 0) bw = std::min<uint64_t>(g1, 1ull << a.bucket_log2);
    else if (a.bucket_log2 < 0 && a.g1 >= AUTO_BUCKET_G1) bw = (g1 + XCD_COUNT - 1) / XCD_COUNT;
    const double per_level = double(M) / double(std::max<uint64_t>(g, 1));
    if (per_level > ITEM_CELLS && a.g1 > 1) {
        const double want = std::ceil(per_level / ITEM_CELLS);  // buckets
        bw = std::min<uint64_t>(bw, std::max<uint64_t>(1, (uint64_t)std::ceil(double(g1) / want)));
    }
    const uint32_t nbk = (uint32_t)((g1 + bw - 1) / bw);
    a.bwidth[b] = (uint32_t)bw;
    a.nbk[b] = std::max<uint32_t>(nbk, 1);
    const uint64_t items = (uint64_t)a.nbk[b] * g;
    require(items < (1ull << 32), "too many (bucket, level) items");
    // lanes per item: ~8 positions per lane, ~4 for the short items of a bucketed block; one
    // item per level over its buckets: ~4 per lane of each bucket's range (12 columns at C2:
    // 4 lanes 368 / 558 us FP32 / FP64 against 388 / 572 us at 8)
    const double ppi = items ? double(M) / double(items) : 1.0;
    a.blanes[b] = lanes(a.nbk[b] > 1 ? ppi / 4.0 : ppi / 8.0);
    a.llanes[b] = lanes(ppi / 4.0);
    a.bperm[b].alloc(std::max<size_t>(M, 1));
    a.boff[b].alloc(items + 1);
    FH_HIP(hipMemsetAsync(a.boff[b].p, 0, a.boff[b].bytes(), st));
    a.lseg[b].alloc(std::max<size_t>(M, 1));
    a.lother[b].alloc(std::max<size_t>(M * (nb - 1), 1));
    if (M > 0) {
        DBuf<uint64_t> keys(M), kalt(M);
        DBuf<uint32_t> valt(M);
        PayloadLayout pl{};
        pl.ib = std::max(1, (int)bits_for(items));
        pl.sb = std::max(1, (int)bits_for(a.g1));
        int tot = pl.ib + pl.sb, o = 0;
        for (int bb = 0; bb < nb; bb++)
            if (bb != b) { pl.ob[o] = std::max(1, (int)bits_for(a.blocks[bb].g)); tot += pl.ob[o++]; }
        pl.pay = tot <= 64;
        hipLaunchKernelGGL(k_bucket_keys, dim3(blocks_for(M)), dim3(TPB), 0, st, a.cell_codes.p, a.cell_seg.p, M, nb,
                           b, (uint32_t)bw, g, pl, keys.p, (uint32_t*)nullptr);
        hipLaunchKernelGGL(k_iota, dim3(blocks_for(M)), dim3(TPB), 0, st, a.bperm[b].p, M);
        FH_CHECK_LAUNCH();
        // a large block's (bucket, level) keys are practically never in cell order already: sort
        // without testing that (the test costs a host round trip and a second read of the codes)
        radix_sort_pairs(keys, a.bperm[b], kalt, valt, M, pl.ib, st);
        const uint64_t mask = pl.ib >= 64 ? ~0ull : ((1ull << pl.ib) - 1);
        hipLaunchKernelGGL(k_offsets_u64, dim3(blocks_for(M)), dim3(TPB), 0, st, keys.p, M, (size_t)items,
                           a.boff[b].p, mask);
        if (pl.pay)
            hipLaunchKernelGGL(k_level_records_keys, dim3(blocks_for(M)), dim3(TPB), 0, st, keys.p, M, nb - 1, pl,
                               a.lseg[b].p, a.lother[b].p);
        else
            hipLaunchKernelGGL(k_level_records, dim3(blocks_for(M)), dim3(TPB), 0, st, a.bperm[b].p, a.cell_seg.p,
                               a.cell_codes.p, M, nb, b, a.lseg[b].p, a.lother[b].p);
        FH_CHECK_LAUNCH();
    }
}

void ensure_block_transpose(Absorber& a, int b) {
    if (a.boff[b].n == 0) build_block_transpose(a, b);
}

void ensure_level_cells(Absorber& a, size_t b) {
    if (a.lperm.size() < a.nb) { a.lperm.resize(a.nb); a.loff.resize(a.nb); }
    if (b >= a.nb) return;
    if (a.lperm[b].n) return;
    hipStream_t st = a.st;
    const size_t M = a.M;
    const int nb = (int)a.nb;
    const uint64_t g = a.blocks[b].g;
    a.lperm[b].alloc(std::max<size_t>(M, 1));
    a.loff[b].alloc(g + 1);
    FH_HIP(hipMemsetAsync(a.loff[b].p, 0, a.loff[b].bytes(), st));
    if (M == 0) return;
    DBuf<uint64_t> keys(M), kalt(M);
    DBuf<uint32_t> valt(M), flag(1);
    hipLaunchKernelGGL(k_level_keys, dim3(blocks_for(M)), dim3(TPB), 0, st, a.cell_codes.p, M, nb, (int)b, keys.p);
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(M)), dim3(TPB), 0, st, a.lperm[b].p, M);
    FH_CHECK_LAUNCH();
    check_sorted_u64(keys.p, M, flag.p, st);
    if (read_scalar(flag.p, st) != 0) radix_sort_pairs(keys, a.lperm[b], kalt, valt, M, (int)bits_for(g), st);
    hipLaunchKernelGGL(k_offsets_u64, dim3(blocks_for(M)), dim3(TPB), 0, st, keys.p, M, (size_t)g, a.loff[b].p);
    FH_CHECK_LAUNCH();
}

}  // namespace fh
