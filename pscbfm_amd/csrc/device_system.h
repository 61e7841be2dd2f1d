// Device-resident state of one system: one allocation carved into the
// buffers the sweep kernels read and write (layout in the .cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "layout.h"

namespace pscbfm {

[[noreturn]] void hip_fail(hipError_t e, const char* what);
inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) hip_fail(e, what);
}

struct DeviceSystem {
    void* arena = nullptr;
    uint64_t bytes = 0;
    uint64_t cells = 0;
    uint32_t accEntries = 0;  // entries of DevState::acceptedByBlock
    SysDesc desc{};
    // The slot of every monomer, for decoding coordinates on the device, and
    // (systems with their own launches, withLists) the buffer positions()
    // decodes them into.  Batch replicas decode into the batch's buffer.
    uint32_t* slotOfDev = nullptr;
    int32_t* xyzDev = nullptr;

    // Bytes allocate() would take.  `withLists`: the system runs its own
    // launches (pass-1 lists inside the allocation); batch replicas use the
    // batch's lists.  `accEntries`: per-workgroup accepted counters.
    static uint64_t footprint(const SystemLayout& L, uint32_t latticeKind, bool withLists, uint32_t accEntries);
    // Allocates, uploads layout and lattice, zeroes counters (substep 0).
    void allocate(const SystemLayout& L, uint32_t latticeKind, uint64_t seed, bool withLists, uint32_t accEntries,
                  hipStream_t stream);
    void release();
    ~DeviceSystem() { release(); }

    DevState readState(hipStream_t stream) const;
    void writeState(const DevState& s, hipStream_t stream);
    uint64_t acceptedTotal(uint32_t entries, hipStream_t stream) const;
    // Unfolded coordinates in the caller's monomer order (withLists systems).
    void positions(const SystemLayout& L, int32_t* xyzOut, hipStream_t stream) const;
    // The job that decodes monomers [first, first + count) into records
    // out.. of a shared buffer (launch_unfold_jobs; firstBlock left 0).
    UnfoldJob unfoldJob(uint32_t first, uint32_t count, uint64_t out) const;
};

// Workgroups covering the largest species with `perGroup` monomers each.
inline uint32_t tiles_for(const SystemLayout& L, uint32_t perGroup) {
    return uint32_t((uint64_t(L.maxColorCount) + perGroup - 1) / perGroup) + (L.maxColorCount == 0 ? 1u : 0u);
}

}  // namespace pscbfm
