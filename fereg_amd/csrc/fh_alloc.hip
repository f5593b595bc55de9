// Stream-ordered caching pool for the library's device buffers, and a pool of streams.
//
// hipMalloc is slow and hipFree waits for the whole device, so a fit that allocates and frees
// tens of scratch arrays would spend most of its time in the allocator. Freed blocks are kept
// per device in a size-ordered free list together with the stream they were last used on:
//   - a request on the same stream reuses a block at once (stream order makes it safe);
//   - a request on another stream reuses a block once the work queued on the block's stream up
//     to the first such request has completed (an event recorded then, queried without waiting);
//   - blocks of a stream that was drained (returned to the stream pool) are idle: reusable at once.
// Blocks are rounded to 256 B (below 1 MiB: to the next power of two) or 2 MiB multiples and a
// block is reused for requests of at least half its size. Memory goes back to HIP on
// fh_release_cached_memory() or when hipMalloc fails (completed blocks are released and the
// allocation is tried once more).
//
// Streams: creating and destroying a HIP stream costs ~100 us each, so the streams of destroyed
// absorbers are drained and kept for the next ones (never destroyed: the device pool keys its
// blocks by stream).
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "fh_common.hpp"

namespace fh {

namespace {

struct Block {
    void* p = nullptr;
    size_t bytes = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool ev_pending = false;  // ev recorded on `stream` after the free
    bool idle = true;         // no work pending on it (its stream was drained)
};

// The pool's own event queries must not leave an error for the caller's next
// hipGetLastError(): clear what they set unless an error was already pending.
struct ErrorHygiene {
    hipError_t prev;
    ErrorHygiene() : prev(hipPeekAtLastError()) {}
    ~ErrorHygiene() {
        if (prev == hipSuccess) (void)hipGetLastError();
    }
};

struct Pool {
    std::mutex mu;
    std::unordered_map<void*, Block> live;
    std::map<int, std::multimap<size_t, Block>> free_by_dev;
    size_t cached = 0, in_use = 0;
};

Pool& pool() {
    static Pool* p = new Pool();  // never destroyed: HIP may be gone at process exit
    return *p;
}

thread_local hipStream_t tls_stream = nullptr;

size_t round_bytes(size_t b) {
    if (b <= 256) return 256;
    if (b < (1u << 20)) {
        size_t r = 256;
        while (r < b) r <<= 1;
        return r;
    }
    const size_t q = 2u << 20;
    return (b + q - 1) / q * q;
}

// Whether the work that used a free block has completed (caller holds the lock; the block's
// device is current or the event API does not need it).
bool block_done(Block& b) {
    if (b.idle) return true;
    if (!b.ev_pending) {
        if (!b.ev && hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess) return false;
        if (hipEventRecord(b.ev, b.stream) != hipSuccess) return false;
        b.ev_pending = true;
    }
    return hipEventQuery(b.ev) == hipSuccess;
}

// Releases every cached block whose last use has completed (caller holds the lock).
void release_completed(Pool& P) {
    for (auto& [dev, fl] : P.free_by_dev) {
        int prev = -1;
        (void)hipGetDevice(&prev);
        if (prev != dev) (void)hipSetDevice(dev);
        for (auto it = fl.begin(); it != fl.end();) {
            Block& b = it->second;
            if (!block_done(b)) { ++it; continue; }
            if (b.ev) (void)hipEventDestroy(b.ev);
            (void)hipFree(b.p);
            P.cached -= b.bytes;
            it = fl.erase(it);
        }
        if (prev != dev && prev >= 0) (void)hipSetDevice(prev);
    }
}

struct StreamPool {
    std::mutex mu;
    std::map<int, std::vector<hipStream_t>> free_by_dev;
};

StreamPool& stream_pool() {
    static StreamPool* p = new StreamPool();
    return *p;
}

}  // namespace

hipStream_t current_stream() { return tls_stream; }

StreamScope::StreamScope(hipStream_t s) : prev(tls_stream) { tls_stream = s; }
StreamScope::~StreamScope() { tls_stream = prev; }

void* pool_alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    Pool& P = pool();
    const size_t rb = round_bytes(bytes);
    int dev = 0;
    FH_HIP(hipGetDevice(&dev));
    const hipStream_t st = tls_stream;
    std::lock_guard<std::mutex> lk(P.mu);
    ErrorHygiene eh;
    auto& fl = P.free_by_dev[dev];
    for (auto it = fl.lower_bound(rb); it != fl.end() && it->first <= 2 * rb; ++it) {
        Block& b = it->second;
        if (b.stream != st && !block_done(b)) continue;
        Block blk = b;
        fl.erase(it);
        blk.stream = st;
        blk.idle = false;
        blk.ev_pending = false;
        P.cached -= blk.bytes;
        P.in_use += blk.bytes;
        P.live.emplace(blk.p, blk);
        return blk.p;
    }
    Block blk;
    blk.bytes = rb;
    blk.device = dev;
    blk.stream = st;
    blk.idle = false;
    hipError_t e = hipMalloc(&blk.p, rb);
    if (e == hipErrorOutOfMemory) {
        release_completed(P);
        e = hipMalloc(&blk.p, rb);
    }
    if (e != hipSuccess) (void)hipGetLastError();  // reported by the throw below, not left pending
    FH_HIP(e);
    P.in_use += rb;
    P.live.emplace(blk.p, blk);
    return blk.p;
}

void pool_free(void* p) {
    if (!p) return;
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.live.find(p);
    if (it == P.live.end()) return;
    Block b = it->second;
    P.live.erase(it);
    b.ev_pending = false;  // the completion event is recorded when another stream asks for it
    P.in_use -= b.bytes;
    P.cached += b.bytes;
    P.free_by_dev[b.device].emplace(b.bytes, b);
}

void pool_release_cached() {
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    ErrorHygiene eh;
    (void)hipDeviceSynchronize();  // every cached block's last use has completed
    for (auto& [dev, fl] : P.free_by_dev)
        for (auto& [sz, b] : fl) b.idle = true;
    release_completed(P);
}

void pool_forget_stream(hipStream_t st) {
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    for (auto& [dev, fl] : P.free_by_dev)
        for (auto& [sz, b] : fl)
            if (b.stream == st) b.idle = true;
}

hipStream_t stream_acquire() {
    int dev = 0;
    FH_HIP(hipGetDevice(&dev));
    StreamPool& S = stream_pool();
    {
        std::lock_guard<std::mutex> lk(S.mu);
        auto& v = S.free_by_dev[dev];
        if (!v.empty()) {
            hipStream_t st = v.back();
            v.pop_back();
            return st;
        }
    }
    hipStream_t st = nullptr;
    FH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    return st;
}

void stream_release(hipStream_t st, int device) {
    if (!st) return;
    (void)hipStreamSynchronize(st);
    pool_forget_stream(st);  // everything queued on it has completed
    StreamPool& S = stream_pool();
    std::lock_guard<std::mutex> lk(S.mu);
    S.free_by_dev[device].push_back(st);
}

// ---------------------------------------------------------------- pinned staging
namespace {

struct Stage {
    char* p = nullptr;
    size_t bytes = 0;
    int device = 0;
    hipEvent_t ev = nullptr;  // recorded after the block's last copy (null: idle)
};

struct StagePool {
    std::mutex mu;
    std::multimap<size_t, Stage> free_;
};

StagePool& stage_pool() {
    static StagePool* p = new StagePool();  // never destroyed: HIP may be gone at process exit
    return *p;
}

constexpr size_t STAGE_MAX = 4u << 20;  // larger transfers go straight from the caller's memory

Stage stage_get(size_t bytes) {
    size_t rb = 4096;
    while (rb < bytes) rb <<= 1;
    int dev = 0;
    FH_HIP(hipGetDevice(&dev));
    StagePool& P = stage_pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        ErrorHygiene eh;
        for (auto it = P.free_.lower_bound(rb); it != P.free_.end() && it->first <= 4 * rb; ++it) {
            Stage& s = it->second;
            if (s.device != dev) continue;
            // (an event whose stream is gone reports an error, not "not ready": its copy is done,
            // since the library drains a stream before destroying it)
            if (s.ev && hipEventQuery(s.ev) == hipErrorNotReady) continue;
            Stage out = s;
            P.free_.erase(it);
            return out;
        }
    }
    Stage s;
    s.bytes = rb;
    s.device = dev;
    FH_HIP(hipHostMalloc((void**)&s.p, rb, hipHostMallocDefault));
    FH_HIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    return s;
}

void stage_put(const Stage& s) {
    StagePool& P = stage_pool();
    std::lock_guard<std::mutex> lk(P.mu);
    P.free_.emplace(s.bytes, s);
}

}  // namespace

// Spin on an event recorded behind the queued work (a query costs ~0.1 us): a blocking wait
// answers within ~15 us when the work is short but puts the thread to sleep during long waits,
// and waking it costs ~100 us. After 20 ms of spinning the wait blocks.
void stream_wait(hipStream_t st) {
    thread_local std::unordered_map<int, hipEvent_t> evs;
    int dev = 0;
    FH_HIP(hipGetDevice(&dev));
    hipEvent_t& ev = evs[dev];
    if (!ev) FH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    FH_HIP(hipEventRecord(ev, st));
    const auto t0 = std::chrono::steady_clock::now();
    const hipError_t prev = hipPeekAtLastError();
    for (int i = 0;; i++) {
        const hipError_t e = hipEventQuery(ev);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) FH_HIP(e);
        if ((i & 255) == 255 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
            FH_HIP(hipEventSynchronize(ev));
            break;
        }
    }
    if (prev == hipSuccess) (void)hipGetLastError();  // the not-ready answers are not errors
}

void upload(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return;
    if (bytes > STAGE_MAX) {
        FH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        return;
    }
    Stage s = stage_get(bytes);
    std::memcpy(s.p, src, bytes);
    FH_HIP(hipMemcpyAsync(dst, s.p, bytes, hipMemcpyHostToDevice, st));
    FH_HIP(hipEventRecord(s.ev, st));
    stage_put(s);
}

void download(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes == 0) {
        stream_wait(st);
        return;
    }
    if (bytes > STAGE_MAX) {
        FH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        stream_wait(st);
        return;
    }
    Stage s = stage_get(bytes);
    FH_HIP(hipMemcpyAsync(s.p, src, bytes, hipMemcpyDeviceToHost, st));
    FH_HIP(hipEventRecord(s.ev, st));
    stream_wait(st);
    std::memcpy(dst, s.p, bytes);
    stage_put(s);
}

AsyncRead::AsyncRead(size_t bytes) : bytes_(bytes) {
    Stage s = stage_get(bytes);
    p_ = s.p;
    cap_ = s.bytes;
    device_ = s.device;
    ev_ = s.ev;
}

AsyncRead::~AsyncRead() {
    Stage s;
    s.p = static_cast<char*>(p_);
    s.bytes = cap_;
    s.device = device_;
    s.ev = ev_;
    stage_put(s);  // reused only once the event of its last copy has completed
}

void AsyncRead::start(const void* src, hipStream_t st) {
    FH_HIP(hipMemcpyAsync(p_, src, bytes_, hipMemcpyDeviceToHost, st));
    FH_HIP(hipEventRecord(ev_, st));
}

const void* AsyncRead::wait() {
    const auto t0 = std::chrono::steady_clock::now();
    const hipError_t prev = hipPeekAtLastError();
    for (int i = 0;; i++) {
        const hipError_t e = hipEventQuery(ev_);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) FH_HIP(e);
        if ((i & 255) == 255 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
            FH_HIP(hipEventSynchronize(ev_));
            break;
        }
    }
    if (prev == hipSuccess) (void)hipGetLastError();
    return p_;
}

void pool_stats(size_t* in_use, size_t* cached) {
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    if (in_use) *in_use = P.in_use;
    if (cached) *cached = P.cached;
}

}  // namespace fh
