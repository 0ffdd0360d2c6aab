// The library's static scratch region: device memory for the render that allocates nothing
// (spkx_render_device_static).
//
// `scratch` is a file-scope device array, so every GPU the library runs on holds its own copy from
// the moment the code object is loaded until the process ends.  In host code the name `scratch`
// is only a host-side stand-in for the symbol, not a device address; the device address of the
// copy on each GPU comes from hipGetSymbolAddress with that GPU current, and only that address is
// ever handed to kernels or copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"

namespace spk::k {

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

}  // namespace

hipError_t static_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_mu);
    if (g_addr[device]) return hipSuccess;
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e != hipSuccess) return e;
    e = hipSetDevice(device);
    if (e != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    const hipError_t back = hipSetDevice(prev);
    if (e != hipSuccess) return e;
    if (back != hipSuccess) return back;
    if (!p) return hipErrorInvalidSymbol;
    g_addr[device] = p;
    return hipSuccess;
}

void* static_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_mu);
    return g_addr[device];
}

}  // namespace spk::k
