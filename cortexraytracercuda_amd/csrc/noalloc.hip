// noalloc.hip - the device scratch array of crt_render_noalloc (include/cortex_rt.h).
//
// Callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc) render through crt_render_noalloc, whose streaming passes keep
// their path state, hit records, per-path radiance, traversal overflow stacks, control block,
// counters and request table in this fixed array instead of the context's pools. Every GPU the
// library runs on holds one copy (kScratchBytes) until the process ends, and every call on a GPU
// shares it: calls that overlap on one GPU race.
//
// In host code `scratch` names a host-side stand-in, not a device address: the device address is
// taken once per device with hipGetSymbolAddress (the library's init, with that device current)
// and only that address is used from then on.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "crt_scratch.h"

static __device__ __align__(256) uint8_t scratch[kCrtScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of this device's scratch array; the caller has made `device` current.
extern "C" int crt_scratch_setup(int device, void** ptr, size_t* bytes) {
    if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return (int)hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    *bytes = kCrtScratchBytes;
    return (int)hipSuccess;
}
