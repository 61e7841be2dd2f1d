// Device scratch of the allocation-free run entry point (noalloc.hip).
#pragma once

#include <cstdint>

namespace pscbfm {

// One call of the allocation-free run keeps, per substep, its explicit draw
// (8 bytes, R250 mode only) and its accepted count (4 bytes, when a log is
// asked for) in the scratch array: at most kNoAllocMaxSubsteps substeps of
// them, 3 MiB in all.
constexpr uint64_t kNoAllocMaxSubsteps = uint64_t(1) << 18;
constexpr uint64_t kNoAllocDrawBytes = 8 * kNoAllocMaxSubsteps;
constexpr uint64_t kNoAllocScratchBytes = 12 * kNoAllocMaxSubsteps;

// Device address of `device`'s copy of the scratch array (each GPU holds its
// own, at its own address, until the process ends).  Looked up with
// hipGetSymbolAddress on the first call for a device, with that device
// current, and kept.  Throws std::runtime_error ("HIP error ...") if the
// lookup fails.
void* noalloc_scratch(int device);

}  // namespace pscbfm
