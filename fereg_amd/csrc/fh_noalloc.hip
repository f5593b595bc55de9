// S*p for callers that cannot allocate (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc): fh_absorber_apply_noalloc runs the reduced operator on the caller's device
// buffers with every scratch array carved from one file-scope device array, `scratch`, of
// SCRATCH_BYTES. Each GPU holds its own copy of that array (at its own address) from the
// library's load until the process ends; its device address is looked up once per device with
// hipGetSymbolAddress, and only that address is handed to kernels. Calls on one GPU share the
// array and keep its last contents: calls that overlap on one GPU race.
#include <atomic>
#include <mutex>

#include "fh_fit.hpp"

namespace fh {

namespace {

// 64 MiB: S*p at n = 1e7 with three absorbed terms (1e6, 1e5, 10 levels) needs 47 MB in FP64 with
// two columns, 25 MB in FP32.
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_addr_mu;
void* g_addr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device (looked up once per device).
void* scratch_address(int device) {
    require(device >= 0 && device < MAX_DEVICES, "device index beyond the no-alloc table", FH_INTERNAL);
    std::lock_guard<std::mutex> lk(g_addr_mu);
    if (!g_addr[device]) {
        void* ptr = nullptr;
        FH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        require(ptr != nullptr, "no device address for the no-alloc scratch array", FH_INTERNAL);
        g_addr[device] = ptr;
    }
    return g_addr[device];
}

template <class TO, class TI>
__global__ void k_cast(const TI* __restrict__ a, TO* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (TO)a[i];
}

// A Scratch whose t, ds, dl, part and ypart are views into the no-alloc array (nothing owned:
// the views are dropped, not freed, on destruction).
template <class T>
struct ScratchView {
    Scratch<T> s;
    ~ScratchView() {
        for (DBuf<T>* b : {&s.t, &s.ds, &s.dl, &s.part, &s.ypart}) {
            b->p = nullptr;
            b->n = 0;
        }
    }
};

size_t round_up(size_t bytes) { return (bytes + 255) / 256 * 256; }

// S*p in precision T on the device rows p (m x r, T) into y; every scratch array from base.
template <class T>
void apply_in(Absorber& a, const T* p, T* y, int r, unsigned char* base, size_t used, size_t avail) {
    const SpmvNeed need = spmv_scratch_need<T>(a, a.plan_for(sizeof(T) == 8, r), r);
    ScratchView<T> v;
    size_t off = used;
    auto take = [&](DBuf<T>& b, size_t count) {
        if (!count) return;
        const size_t bytes = round_up(count * sizeof(T));
        require(off + bytes <= avail,
                "apply_noalloc: the product needs more scratch than the " + std::to_string(avail >> 20) +
                    " MiB the library reserves (fewer columns or FP32 need less)");
        b.p = reinterpret_cast<T*>(base + off);
        b.n = count;
        off += bytes;
    };
    take(v.s.t, need.t);
    take(v.s.ds, need.ds);
    take(v.s.dl, need.dl);
    take(v.s.part, need.part);
    take(v.s.ypart, need.ypart);
    spmv_fixed<T>(a, p, y, r, v.s, a.st);
}

}  // namespace

}  // namespace fh

using namespace fh;

extern "C" {

int fh_absorber_apply_noalloc(fh_absorber* ab, const double* p, size_t r, double* y, int32_t precision) {
    return guard([&] {
        require(ab != nullptr && p && y, "null argument");
        require(r >= 1 && r <= 16, "1..16 columns");
        require(precision == FH_PLAIN_FP64 || precision == FH_MIXED, "unknown precision");
        Absorber& a = ab->a;
        require(!a.slope, "apply_noalloc: not available for an absorber with varying slopes");
        require(!a.irls_weights,
                "apply_noalloc: the absorber holds a Poisson fit's IRLS weights; restoring the caller's weights "
                "allocates, so call another entry point (or set_weights) first");
        DeviceGuard dg(a.device);
        unsigned char* base = static_cast<unsigned char*>(scratch_address(a.device));
        const size_t m = a.m;
        if (m == 0 || a.nb == 0) return;
        if (precision == FH_MIXED) {
            // FP32 copies of p and y in the array, then the scratch of the product after them
            const size_t vb = round_up(m * r * sizeof(float));
            require(2 * vb <= SCRATCH_BYTES, "apply_noalloc: the FP32 copies of p and y exceed the scratch array");
            float* p32 = reinterpret_cast<float*>(base);
            float* y32 = reinterpret_cast<float*>(base + vb);
            const unsigned g = grid_for(m * r, 256, 4096);
            hipLaunchKernelGGL((k_cast<float, double>), dim3(g), dim3(256), 0, a.st, p, p32, m * r);
            FH_CHECK_LAUNCH();
            apply_in<float>(a, p32, y32, (int)r, base, 2 * vb, SCRATCH_BYTES);
            hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, a.st, y32, y, m * r);
            FH_CHECK_LAUNCH();
        } else {
            apply_in<double>(a, p, y, (int)r, base, 0, SCRATCH_BYTES);
        }
        FH_HIP(hipStreamSynchronize(a.st));
    });
}

size_t fh_noalloc_scratch_bytes(void) { return SCRATCH_BYTES; }

}  // extern "C"
