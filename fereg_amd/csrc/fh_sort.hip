// Stable LSD radix sort and exclusive scans (see fh_sort.hpp).
//
// Radix sort: 8-bit digits, tiles of 4096 keys per 256-thread workgroup. Each wave owns a
// contiguous 1024-key slice of the tile (16 coalesced loads of 64 keys), ranks its keys with the
// wave multisplit (8 ballots give the mask of lanes holding the same digit), keeps per-wave
// digit counters in LDS, and the tile's digit bases come from one exclusive scan of the
// digit-major histogram. Order inside a tile is (wave, item, lane) = input order: stable.
#include "fh_sort.hpp"

namespace fh {

namespace {

constexpr int RS_WG = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_WAVES = RS_WG / WAVE;
constexpr int RS_TILE = RS_WG * RS_ITEMS;

__global__ __launch_bounds__(RS_WG) void rs_upsweep(const uint64_t* __restrict__ keys, size_t n, int shift, uint32_t dmask,
                                                    uint32_t* __restrict__ hist, uint32_t ntiles) {
    __shared__ uint32_t cnt[RS_WAVES][256];
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_WG) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int w = threadIdx.x / WAVE, l = lane_id();
    const size_t base = (size_t)blockIdx.x * RS_TILE + (size_t)w * WAVE * RS_ITEMS;
    for (int i = 0; i < RS_ITEMS; i++) {
        size_t idx = base + (size_t)i * WAVE + l;
        bool valid = idx < n;
        uint32_t d = valid ? (uint32_t)(keys[idx] >> shift) & dmask : 0u;
        uint64_t act = __ballot(valid);
        uint64_t m = match_key(d, 8, act);
        if (valid && (m & lanemask_lt()) == 0) cnt[w][d] += (uint32_t)__popcll(m);
    }
    __syncthreads();
    const int t = threadIdx.x;
    uint32_t s = 0;
#pragma unroll
    for (int ww = 0; ww < RS_WAVES; ww++) s += cnt[ww][t];
    hist[(size_t)t * ntiles + blockIdx.x] = s;
}

// Scatter pass. The keys of a tile are first placed in LDS in their sorted-within-tile order
// (digit, then input order), then written out by consecutive threads: a digit's run of the tile
// goes to consecutive global addresses, so the writes coalesce instead of landing as single
// 8- and 4-byte stores on 256 different lines.
__global__ __launch_bounds__(RS_WG) void rs_downsweep(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      size_t n, int shift, uint32_t dmask, const uint32_t* __restrict__ hist,
                                                      uint32_t ntiles) {
    __shared__ uint32_t cnt[RS_WAVES][256];
    __shared__ uint32_t lstart[256], gstart[256];
    __shared__ uint64_t lk[RS_TILE];
    __shared__ uint32_t lv[RS_TILE];
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_WG) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int w = threadIdx.x / WAVE, l = lane_id();
    const size_t tile0 = (size_t)blockIdx.x * RS_TILE;
    const size_t base = tile0 + (size_t)w * WAVE * RS_ITEMS;
    uint64_t k[RS_ITEMS];
    uint32_t v[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
    uint32_t dig[RS_ITEMS];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
        size_t idx = base + (size_t)i * WAVE + l;
        bool valid = idx < n;
        k[i] = valid ? kin[idx] : 0;
        v[i] = valid ? vin[idx] : 0;
    }
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
        size_t idx = base + (size_t)i * WAVE + l;
        bool valid = idx < n;
        uint32_t d = (uint32_t)(k[i] >> shift) & dmask;
        uint64_t act = __ballot(valid);
        uint64_t m = match_key(d, 8, act);
        uint32_t c = cnt[w][d];
        rank[i] = c + (uint32_t)__popcll(m & lanemask_lt());
        dig[i] = d;
        if (valid && (m & lanemask_lt()) == 0) cnt[w][d] = c + (uint32_t)__popcll(m);
    }
    __syncthreads();
    {
        // per digit t: the tile's total, its global base (hist) and the per-wave local bases
        const int t = threadIdx.x;
        uint32_t tot = 0;
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ww++) {
            const uint32_t c = cnt[ww][t];
            cnt[ww][t] = tot;  // wave's offset inside the digit's run of this tile
            tot += c;
        }
        gstart[t] = hist[(size_t)t * ntiles + blockIdx.x];
        // exclusive scan of the totals over the 256 digits (one per thread)
        uint32_t x = tot;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, WAVE);
            if (l >= o) x += y;
        }
        __shared__ uint32_t wtot[RS_WAVES];
        if (l == WAVE - 1) wtot[w] = x;
        __syncthreads();
        uint32_t off = 0;
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ww++)
            if (ww < w) off += wtot[ww];
        lstart[t] = off + x - tot;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
        size_t idx = base + (size_t)i * WAVE + l;
        if (idx < n) {
            const uint32_t li = lstart[dig[i]] + cnt[w][dig[i]] + rank[i];
            lk[li] = k[i];
            lv[li] = v[i];
        }
    }
    __syncthreads();
    const size_t cnt_tile = n - tile0 < (size_t)RS_TILE ? n - tile0 : (size_t)RS_TILE;
    for (uint32_t li = threadIdx.x; li < cnt_tile; li += RS_WG) {
        const uint64_t kk = lk[li];
        const uint32_t d = (uint32_t)(kk >> shift) & dmask;
        const uint32_t pos = gstart[d] + (li - lstart[d]);
        kout[pos] = kk;
        vout[pos] = lv[li];
    }
}

// ---------------------------------------------------------------- scans

constexpr int SC_WG = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_CHUNK = SC_WG * SC_ITEMS;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[SC_WG / WAVE];
    const int l = lane_id(), w = threadIdx.x / WAVE;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t y = __shfl_up(x, o, WAVE);
        if (l >= o) x += y;
    }
    if (l == WAVE - 1) wsum[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int ww = 0; ww < SC_WG / WAVE; ww++) {
        uint32_t s = wsum[ww];
        if (ww < w) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__global__ __launch_bounds__(SC_WG) void scan_reduce(const uint32_t* __restrict__ in, size_t n,
                                                     uint32_t* __restrict__ bsum) {
    const size_t base = (size_t)blockIdx.x * SC_CHUNK;
    uint32_t s = 0;
    for (int i = 0; i < SC_ITEMS; i++) {
        size_t idx = base + (size_t)i * SC_WG + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    __shared__ uint32_t ws[SC_WG / WAVE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, WAVE);
    if (lane_id() == 0) ws[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int ww = 0; ww < SC_WG / WAVE; ww++) t += ws[ww];
        bsum[blockIdx.x] = t;
    }
}

// Exclusive scan of the block sums by one workgroup (sequential chunks per thread).
__global__ __launch_bounds__(SC_WG) void scan_top(uint32_t* __restrict__ bsum, size_t nb, uint32_t* total_dev) {
    const size_t per = (nb + SC_WG - 1) / SC_WG;
    const size_t b0 = (size_t)threadIdx.x * per;
    uint32_t s = 0;
    for (size_t i = b0; i < b0 + per && i < nb; i++) s += bsum[i];
    uint32_t tot;
    uint32_t off = block_exclusive_scan(s, &tot);
    for (size_t i = b0; i < b0 + per && i < nb; i++) {
        uint32_t x = bsum[i];
        bsum[i] = off;
        off += x;
    }
    if (threadIdx.x == 0 && total_dev) *total_dev = tot;
}

__global__ __launch_bounds__(SC_WG) void scan_down(const uint32_t* in, uint32_t* out, size_t n,
                                                   const uint32_t* __restrict__ bsum) {
    const size_t base = (size_t)blockIdx.x * SC_CHUNK + (size_t)threadIdx.x * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++) {
        size_t idx = base + i;
        v[i] = idx < n ? in[idx] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t off = block_exclusive_scan(s, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++) {
        size_t idx = base + i;
        if (idx < n) out[idx] = off;
        off += v[i];
    }
}

__global__ void sorted_check(const uint64_t* __restrict__ keys, size_t n, uint32_t* flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    bool bad = false;
    for (; i < n; i += stride) bad |= keys[i] < keys[i - 1];
    // one plain store of 1 per workgroup that saw a descent: every writer stores the same value,
    // so no atomic is needed (atomics from every workgroup on one address serialize)
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (__ballot(bad) != 0 && lane_id() == 0) any = 1;
    __syncthreads();
    if (threadIdx.x == 0 && any) *reinterpret_cast<volatile uint32_t*>(flag) = 1u;
}

}  // namespace

// Runs the passes; returns true when the sorted data ended in the alternate buffers.
static bool radix_passes(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, size_t n,
                         int end_bit, hipStream_t st) {
    if (n <= 1 || end_bit <= 0) return false;
    StreamScope ss(st);  // scratch is stream-ordered on st
    const uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
    DBuf<uint32_t> hist((size_t)256 * ntiles);
    uint64_t *ka = keys, *kb = keys_alt;
    uint32_t *va = vals, *vb = vals_alt;
    for (int shift = 0; shift < end_bit; shift += 8) {
        // digits beyond end_bit are masked: bits above it may carry payload
        const uint32_t dmask = end_bit - shift >= 8 ? 0xFFu : ((1u << (end_bit - shift)) - 1u);
        hipLaunchKernelGGL(rs_upsweep, dim3(ntiles), dim3(RS_WG), 0, st, ka, n, shift, dmask, hist.p, ntiles);
        FH_CHECK_LAUNCH();
        exclusive_scan_u32(hist.p, hist.p, (size_t)256 * ntiles, nullptr, st);
        hipLaunchKernelGGL(rs_downsweep, dim3(ntiles), dim3(RS_WG), 0, st, ka, va, kb, vb, n, shift, dmask, hist.p,
                           ntiles);
        FH_CHECK_LAUNCH();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    return ka != keys;
}

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, size_t n,
                      int end_bit, hipStream_t st) {
    if (radix_passes(keys, vals, keys_alt, vals_alt, n, end_bit, st)) {
        FH_HIP(hipMemcpyAsync(keys, keys_alt, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
        FH_HIP(hipMemcpyAsync(vals, vals_alt, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    }
}

void radix_sort_pairs(DBuf<uint64_t>& keys, DBuf<uint32_t>& vals, DBuf<uint64_t>& keys_alt, DBuf<uint32_t>& vals_alt,
                      size_t n, int end_bit, hipStream_t st) {
    require(keys.n >= n && vals.n >= n && keys_alt.n >= n && vals_alt.n >= n, "radix sort: buffers shorter than n");
    if (radix_passes(keys.p, vals.p, keys_alt.p, vals_alt.p, n, end_bit, st)) {
        std::swap(keys, keys_alt);
        std::swap(vals, vals_alt);
    }
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total_dev, hipStream_t st) {
    if (n == 0) {
        if (total_dev) FH_HIP(hipMemsetAsync(total_dev, 0, sizeof(uint32_t), st));
        return;
    }
    const size_t nb = (n + SC_CHUNK - 1) / SC_CHUNK;
    StreamScope ss(st);
    DBuf<uint32_t> bsum(nb);
    hipLaunchKernelGGL(scan_reduce, dim3((unsigned)nb), dim3(SC_WG), 0, st, in, n, bsum.p);
    FH_CHECK_LAUNCH();
    hipLaunchKernelGGL(scan_top, dim3(1), dim3(SC_WG), 0, st, bsum.p, nb, total_dev);
    FH_CHECK_LAUNCH();
    hipLaunchKernelGGL(scan_down, dim3((unsigned)nb), dim3(SC_WG), 0, st, in, out, n, bsum.p);
    FH_CHECK_LAUNCH();
}

void check_sorted_u64(const uint64_t* keys, size_t n, uint32_t* flag_dev, hipStream_t st) {
    FH_HIP(hipMemsetAsync(flag_dev, 0, sizeof(uint32_t), st));
    if (n < 2) return;
    hipLaunchKernelGGL(sorted_check, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, keys, n, flag_dev);
    FH_CHECK_LAUNCH();
}

}  // namespace fh
