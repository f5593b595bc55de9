// Shared host/device utilities of the fereg_hip library (gfx950: wave64, 160 KB LDS per CU).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/fereg_hip.h"

namespace fh {

// ------------------------------------------------------------------ errors

struct Error : std::runtime_error {
    int rc;
    Error(int code, const std::string& msg) : std::runtime_error(msg), rc(code) {}
};

void set_last_error(const std::string& msg);

#define FH_HIP(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            throw ::fh::Error(e_ == hipErrorOutOfMemory ? FH_OOM : FH_INTERNAL,                        \
                              std::string("HIP error: ") + hipGetErrorString(e_) + " at " + __FILE__ + \
                                  ":" + std::to_string(__LINE__));                                     \
        }                                                                                              \
    } while (0)

#define FH_CHECK_LAUNCH() FH_HIP(hipGetLastError())

inline void require(bool ok, const std::string& msg, int rc = FH_INVALID) {
    if (!ok) throw Error(rc, msg);
}

// ------------------------------------------------------------------ device buffers

// Stream-ordered caching pool (fh_alloc.hip): blocks freed on one stream are reused at once on
// that stream and on other streams once their last use has completed.
void* pool_alloc(size_t bytes);  // on the current device, for use on current_stream()
void pool_free(void* p);
void pool_release_cached();
// Marks the cached blocks last used on st as idle (call after st is synchronized, before it is
// destroyed): later requests reuse them without querying events of a dead stream.
void pool_forget_stream(hipStream_t st);
// Streams from the library's pool (non-blocking, default priority); release drains the stream
// and keeps it for the next acquire on that device.
hipStream_t stream_acquire();
void stream_release(hipStream_t st, int device);
void pool_stats(size_t* in_use, size_t* cached);
hipStream_t current_stream();

// Small host<->device transfers through pinned staging blocks (fh_alloc.hip). A copy from
// pageable memory is synchronous with the host, so every small upload of a parameter array would
// otherwise stall the host until the stream drains. upload() returns at once (the staging block
// is reused only after its copy has completed); download() waits for the stream (stream_wait).
void upload(void* dst_dev, const void* src_host, size_t bytes, hipStream_t st);
void download(void* dst_host, const void* src_dev, size_t bytes, hipStream_t st);
// Waits until all work queued on st has completed (spinning on an event, then blocking).
void stream_wait(hipStream_t st);

// A device->host read into a pinned staging block that does not wait: start() queues the copy
// on a stream, wait() spins until that copy has completed and returns the block.
class AsyncRead {
public:
    explicit AsyncRead(size_t bytes);
    ~AsyncRead();
    AsyncRead(const AsyncRead&) = delete;
    AsyncRead& operator=(const AsyncRead&) = delete;
    void start(const void* src_dev, hipStream_t st);
    const void* wait();

private:
    void* p_ = nullptr;
    size_t bytes_ = 0, cap_ = 0;
    int device_ = 0;
    hipEvent_t ev_ = nullptr;
};
// Sets the stream that pool allocations of this thread are used on, for the scope's lifetime.
struct StreamScope {
    hipStream_t prev;
    explicit StreamScope(hipStream_t s);
    ~StreamScope();
    StreamScope(const StreamScope&) = delete;
    StreamScope& operator=(const StreamScope&) = delete;
};

// Owning device allocation from the pool (current device, current_stream()).
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) p = static_cast<T*>(pool_alloc(count * sizeof(T)));
    }
    // Grows to at least `count` elements (contents not preserved).
    void reserve(size_t count) {
        if (count > n) alloc(count);
    }
    void release() {
        if (p) pool_free(p);
        p = nullptr;
        n = 0;
    }
    T* get() const { return p; }
    size_t bytes() const { return n * sizeof(T); }
};

// ------------------------------------------------------------------ device helpers

constexpr int WAVE = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
// Sum over aligned groups of G lanes (G a power of two <= 64); every lane gets its group's sum.
template <int G, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }

// Mask of the lanes whose `key` equals this lane's key among the lanes in `active`
// (8..32-bit keys matched bit by bit with ballots: the classic wave multisplit).
__device__ __forceinline__ uint64_t match_key(uint32_t key, int bits, uint64_t active) {
    uint64_t m = active;
    for (int b = 0; b < bits; b++) {
        uint64_t bal = __ballot((key >> b) & 1u);
        m &= ((key >> b) & 1u) ? bal : ~bal;
    }
    return m;
}

__device__ __forceinline__ uint64_t lanemask_lt() { return (uint64_t(1) << lane_id()) - 1; }

inline unsigned bits_for(uint64_t g) {  // structure.rs:234-236
    if (g <= 1) return 0;
    unsigned b = 0;
    uint64_t v = g - 1;
    while (v) { b++; v >>= 1; }
    return b;
}

inline unsigned grid_for(size_t work, unsigned per_block, unsigned cap = 65535u * 4) {
    size_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

}  // namespace fh
