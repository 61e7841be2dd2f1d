// Device-resident state of one system: one allocation carved into the
// buffers the sweep kernels read and write (layout in the .cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"
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
    // Systems with their own launches (withLists): the slot of every monomer
    // and a buffer of unfolded coordinates in monomer order, so positions()
    // decodes on the device.
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
    // Unfolded coordinates in the caller's monomer order.
    void positions(const SystemLayout& L, int32_t* xyzOut, hipStream_t stream) const;
    // The pieces positions() copies: pos, imgX, imgY and imgZ are consecutive
    // in the arena, so one copy of positionsBytes() bytes from positionsBase()
    // holds them; decodePositions() turns such a copy into coordinates.
    const void* positionsBase() const { return static_cast<uint32_t*>(desc.pos); }
    uint64_t positionsBytes(const SystemLayout& L) const {
        return uint64_t(reinterpret_cast<const char*>(static_cast<int32_t*>(desc.imgZ)) - reinterpret_cast<const char*>(static_cast<uint32_t*>(desc.pos))) +
               uint64_t(L.nSlots) * 4;
    }
    void decodePositions(const SystemLayout& L, const void* copy, int32_t* xyzOut) const;
};

// Workgroups covering the largest species with `perGroup` monomers each.
inline uint32_t tiles_for(const SystemLayout& L, uint32_t perGroup) {
    return uint32_t((uint64_t(L.maxColorCount) + perGroup - 1) / perGroup) + (L.maxColorCount == 0 ? 1u : 0u);
}

}  // namespace pscbfm
