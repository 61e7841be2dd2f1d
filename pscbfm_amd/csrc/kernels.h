// Launch wrappers of the gfx950 sweep kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace pscbfm {

struct LaunchShape {
    uint32_t block = 256;       // threads per workgroup
    uint32_t checkGrid = 8;     // workgroups of the pass-1 kernel (multiple of 8)
    uint32_t performGrid = 8;   // workgroups of the pass-2 kernel (multiple of 8)
};

// One substep = pass 1 (propose/check) then pass 2 (resolve/apply).  `k` is
// the offset of this substep from the device counter DevState::substep, so a
// captured sequence of substeps stays valid for every replay.
hipError_t launch_check(const SysDesc& d, uint32_t k, const LaunchShape& s, hipStream_t st);
hipError_t launch_perform(const SysDesc& d, uint32_t k, const LaunchShape& s, hipStream_t st);
// DevState::substep += count (after a block of substeps).
hipError_t launch_advance(DevState* state, uint32_t count, hipStream_t st);
// Fills n 32-bit words with v (graph-safe replacement of a memset).
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st);

// Grid shape that covers `items` with `block` threads per workgroup,
// rounded up to whole XCD rounds and capped at `cap` workgroups.
inline uint32_t grid_for(uint64_t items, uint32_t block, uint32_t cap) {
    uint64_t g = (items + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    g = (g + 7) / 8 * 8;
    return uint32_t(g);
}

}  // namespace pscbfm
