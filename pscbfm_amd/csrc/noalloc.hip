// The scratch array of the allocation-free run entry point.
//
// Callers that cannot allocate device memory (code under graph capture,
// serving loops, harnesses that forbid hipMalloc) run sweeps through
// Updater::runNoAlloc, which keeps its per-call device buffers (the explicit
// draw table and the per-substep accepted counts) here instead of in
// hipMalloc'd memory.  Every call on a GPU shares that GPU's copy: calls that
// overlap on one GPU race, and each call clears what it needs zeroed.
#include "noalloc.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <sstream>
#include <stdexcept>

namespace pscbfm {

namespace {

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratchPtr[kMaxDevices] = {};

}  // namespace

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("noalloc scratch: device id out of range");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (gScratchPtr[device] == nullptr) {
        int current = 0;
        hipError_t err = hipGetDevice(&current);
        if (err == hipSuccess && current != device) err = hipSetDevice(device);
        void* ptr = nullptr;
        if (err == hipSuccess) err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (current != device) (void)hipSetDevice(current);
        if (err != hipSuccess || ptr == nullptr) {
            std::ostringstream msg;  // "HIP error" marks a device failure for the C ABI
            msg << "HIP error in hipGetSymbolAddress (noalloc scratch, device " << device << "): "
                << hipGetErrorName(err) << " (" << hipGetErrorString(err) << ")";
            throw std::runtime_error(msg.str());
        }
        gScratchPtr[device] = ptr;
    }
    return gScratchPtr[device];
}

}  // namespace pscbfm
