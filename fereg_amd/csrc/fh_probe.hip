// Calibration of the S*p passes' limit: a random gather from an L2-resident table. The passes
// stream one 32-bit code per cell and gather that level's entry (one L2 request per lane: the
// codes of neighbouring cells are unrelated), so their ceiling is the rate at which the L2 serves
// such requests, not HBM bandwidth. This kernel does only that: streams codes (coalesced) and
// gathers table entries, summing them per thread so that no gather is optimized away.
#include <chrono>

#include "fh_fit.hpp"

namespace fh {

namespace {

__global__ void k_probe_codes(uint32_t* __restrict__ codes, size_t n, uint32_t g, uint64_t seed) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        // SplitMix64 of (seed, i): a fixed pseudo-random level per cell
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        codes[i] = (uint32_t)(z % g);
    }
}

template <class V>
__global__ __launch_bounds__(256) void k_probe_gather(const uint32_t* __restrict__ codes, size_t n,
                                                      const V* __restrict__ table, float* __restrict__ out) {
    float acc = 0.0f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const V v = table[codes[i]];
        if constexpr (sizeof(V) == 4) acc += v;
        else if constexpr (sizeof(V) == 8) acc += v.x + v.y;
        else acc += v.x + v.y + v.z + v.w;
    }
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

__global__ void k_probe_fill(float* __restrict__ t, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        t[i] = (float)(i & 1023) * 0.001f;
}

}  // namespace

}  // namespace fh

using namespace fh;

extern "C" {

int fh_time_gather(size_t levels, size_t count, int32_t entry_bytes, uint32_t reps, double* ms_per_pass) {
    return guard([&] {
        require(ms_per_pass != nullptr, "null argument");
        require(entry_bytes == 4 || entry_bytes == 8 || entry_bytes == 16, "entry_bytes must be 4, 8 or 16");
        require(levels >= 1 && levels < 0xFFFFFFFFull && count >= 1 && reps >= 1, "bad sizes");
        int device = 0;
        FH_HIP(hipGetDevice(&device));
        hipStream_t st = stream_acquire();
        {
            StreamScope ss(st);
            const size_t tw = levels * (size_t)entry_bytes / 4;
            DBuf<float> table(tw);
            DBuf<uint32_t> codes(count);
            const unsigned grid = 256u * 32u;  // 32 resident workgroups of 256 lanes per CU
            DBuf<float> out((size_t)grid * 256);
            hipLaunchKernelGGL(k_probe_fill, dim3(grid_for(tw, 256, 4096)), dim3(256), 0, st, table.p, tw);
            hipLaunchKernelGGL(k_probe_codes, dim3(grid_for(count, 256, 8192)), dim3(256), 0, st, codes.p, count,
                               (uint32_t)levels, 12345ull);
            FH_CHECK_LAUNCH();
            auto launch = [&] {
                if (entry_bytes == 4)
                    hipLaunchKernelGGL(k_probe_gather<float>, dim3(grid), dim3(256), 0, st, codes.p, count, table.p,
                                       out.p);
                else if (entry_bytes == 8)
                    hipLaunchKernelGGL(k_probe_gather<float2>, dim3(grid), dim3(256), 0, st, codes.p, count,
                                       reinterpret_cast<const float2*>(table.p), out.p);
                else
                    hipLaunchKernelGGL(k_probe_gather<float4>, dim3(grid), dim3(256), 0, st, codes.p, count,
                                       reinterpret_cast<const float4*>(table.p), out.p);
            };
            launch();  // warm the table into L2
            FH_CHECK_LAUNCH();
            hipEvent_t e0, e1;
            FH_HIP(hipEventCreate(&e0));
            FH_HIP(hipEventCreate(&e1));
            FH_HIP(hipEventRecord(e0, st));
            for (uint32_t k = 0; k < reps; k++) launch();
            FH_HIP(hipEventRecord(e1, st));
            FH_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            FH_HIP(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            *ms_per_pass = (double)ms / reps;
            FH_HIP(hipStreamSynchronize(st));
        }
        stream_release(st, device);
    });
}

}  // extern "C"
