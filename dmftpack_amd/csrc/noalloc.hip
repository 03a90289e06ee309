// Device scratch of the allocation-free entry point dmftgpu_greenftn_w_noalloc.
//
// One file-scope array per device, loaded with the library's code object and
// held until the process ends.  Its device address is taken once per device,
// with that device current, when a handle is created there (dmftgpu_create);
// host code uses only that address, never the array's name.  Every call on a
// device shares the array and leaves its contents behind: calls that overlap
// on one device race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "api_internal.hpp"

namespace {

static __device__ __align__(256) char scratch[dgi::kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_addr[kMaxDevices] = {};

}  // namespace

namespace dgi {

hipError_t scratch_address(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_lock);
    if (!g_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_addr[device] = p;
    }
    *out = g_addr[device];
    return hipSuccess;
}

}  // namespace dgi
