// Device scratch of the allocation-free trace entry point (vkhrt_trace_rays_noalloc): one
// file-scope device array per GPU, held until the process ends. Host code never passes the
// array's name to a kernel (in host code it only stands for the symbol); it uses the device
// address that hipGetSymbolAddress returns for the current device, looked up once per device
// when a context is created on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "build.hpp"

namespace {

static __device__ __align__(256) unsigned char scratch[vk::kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_address[kMaxDevices] = {};

}  // namespace

namespace vk {

hipError_t noalloc_scratch_lookup(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_address[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_address[device] = p;
    }
    *out = g_address[device];
    return hipSuccess;
}

}  // namespace vk
