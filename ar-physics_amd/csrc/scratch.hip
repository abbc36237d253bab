// Device scratch of the allocation-free entry points (arp_lincomb_noalloc).
//
// One fixed array per GPU, in this file's own code object: the HIP runtime places it on
// every device the library runs on, and it stays there until the process ends.  Host code
// never passes the array's name to a kernel (on the host that name is only a stand-in, not
// a device address): arp_scratch_acquire looks up each device's address once, with that
// device current, and callers pass that address to the kernels.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's contents, so
// calls that overlap on one GPU (two streams, two threads) race on it.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

#include "scratch.h"

static __device__ __align__(256) unsigned char scratch[kArpScratchBytes];

namespace {
std::mutex g_mu;
void* g_ptr[kArpScratchMaxDevices] = {};
}  // namespace

// With `device` current: the device address of this GPU's scratch array (looked up once per
// device, then reused).  Returns 0, or the HIP error code of the lookup.
extern "C" __attribute__((visibility("hidden"))) int arp_scratch_acquire(int device, void** ptr, size_t* bytes) {
    if (device < 0 || device >= kArpScratchMaxDevices || !ptr || !bytes) return (int)hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return (int)hipErrorInvalidSymbol;
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    *bytes = kArpScratchBytes;
    return 0;
}
