// Device scratch of the allocation-free entry points (pic_particles_positions_noalloc):
// one file-scope array per GPU, so a caller that cannot allocate (code under HIP graph
// capture, a tight serving loop) still gets its staging memory.  In host code the array's
// name is only a host-side stand-in: its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device with that device current, and only that
// address is ever passed to kernels or copies.  Every call on a GPU shares the GPU's array
// and it keeps the last call's contents; calls that overlap on one GPU race.
#include <mutex>

#include "kernels.h"

namespace picmi {

constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;   // held by each GPU until the process ends
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[device] = p;
    return hipSuccess;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_scratch[device];
}

size_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

}  // namespace picmi
