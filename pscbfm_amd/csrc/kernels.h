// Launch wrappers of the gfx950 sweep kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace pscbfm {

// Both passes run on the same grid: workgroup w of pass 1 covers species
// indices [w*slots*block, (w+1)*slots*block) and writes list w, which
// workgroup w of pass 2 consumes.
struct LaunchShape {
    uint32_t block = 256;        // pass-1 threads per workgroup (64..256, multiple of 64)
    uint32_t slots = 2;          // monomers per pass-1 thread (1, 2 or 4)
    uint32_t grid = 8;           // pass-1 workgroups (multiple of 8)
    uint32_t performBlock = 0;   // pass-2 threads per workgroup (0: the pass-1 block)
    // Pass 2 of a single system: 0 = one pass-1 list per workgroup; else
    // each group of listGroup threads (a multiple of 64) walks one list, so
    // a workgroup walks passTwoBlock() / listGroup lists, with listDepth
    // entries (1 or 4) in flight per thread.
    uint32_t listGroup = 0;
    uint32_t listDepth = 4;
    // The pass (1 or 2) of a single system's substep that zeroes the other
    // parity's proposal buffer.  Batches clear in pass 2.
    uint32_t clearPass = 2;
    uint32_t passTwoBlock() const { return performBlock ? performBlock : block; }
    uint32_t listStride() const { return slots * block; }  // list entries per workgroup
    // Pass-1 lists each XCD produces: workgroup w runs on XCD w % 8 and
    // writes list (w % 8) * grid / 8 + w / 8.
    uint32_t listsPerXcd() const { return grid / 8; }
    uint32_t passTwoGrid() const {
        if (!listGroup) return grid;
        const uint32_t perGroup = passTwoBlock() / listGroup;
        return 8 * ((listsPerXcd() + perGroup - 1) / perGroup);
    }
};

// One substep = pass 1 (propose/check) then pass 2 (resolve/apply).  `k` is
// the offset of this substep from the device counter DevState::substep, so a
// captured sequence of substeps stays valid for every replay.
hipError_t launch_check(const SysDesc& d, uint32_t k, const LaunchShape& s, hipStream_t st);
hipError_t launch_perform(const SysDesc& d, uint32_t k, const LaunchShape& s, hipStream_t st);
// DevState::substep += count (after a block of substeps).
hipError_t launch_advance(DevState* state, uint32_t count, hipStream_t st);

// A batch of independent systems swept by one launch per pass.
struct BatchLaunch {
    const SysDesc* descs;     // device table of replica descriptors
    const Tile* tileTable;    // device, one entry per workgroup
    uint2* lists;             // tiles * slots * block entries
    uint32_t* listCounts;     // tiles entries
    uint32_t tiles;           // workgroups per launch (multiple of 8)
    uint32_t maxDegree;       // largest bond count over all replicas
    uint32_t latticeKind;     // the same for every replica
};
hipError_t launch_check_batch(const BatchLaunch& b, uint32_t k, const LaunchShape& s, hipStream_t st);
hipError_t launch_perform_batch(const BatchLaunch& b, uint32_t k, const LaunchShape& s, hipStream_t st);
// For every replica with a state: substep = min(substep + count, stopAt).
hipError_t launch_advance_batch(const SysDesc* descs, uint32_t nDescs, uint32_t count, hipStream_t st);
// stopAt of replica r = stops[r] (replicas without a state are skipped).
hipError_t launch_set_stops(const SysDesc* descs, uint32_t nDescs, const uint64_t* stops, hipStream_t st);
// out[3r..3r+2] = substep, stopAt, accepted total of replica r.
hipError_t launch_gather_state(const SysDesc* descs, uint32_t nDescs, uint64_t* out, hipStream_t st);
// xyz[3i..3i+2] = unfolded corner of monomer i (slot slotOf[i]): the
// Z-code's axis bits plus image counter times box edge, for i < n.
hipError_t launch_unfold(const SysDesc& d, const uint32_t* slotOf, uint32_t n, int32_t* xyz, hipStream_t st);
// One range of one system's monomers for launch_unfold_jobs.  Pointers are
// device addresses; ranges of one launch are in order of firstBlock, the
// first at 0, each taking ceil(count / kUnfoldBlock) workgroups.
constexpr uint32_t kUnfoldBlock = 256;
struct UnfoldJob {
    const uint32_t* pos;     // Z-curve code per slot
    const int32_t* img[3];   // image counters per slot
    const uint32_t* slotOf;  // slot of each monomer
    uint32_t mask[3];        // Z-curve axis masks
    uint32_t first;          // first monomer of the range
    uint32_t count;          // monomers in the range
    uint32_t firstBlock;     // first workgroup of the range
    uint64_t out;            // record (3 int32) of the range's first monomer in xyz
};
// xyz[3(out+i)..] = unfolded corner of monomer first+i of every job, i < count.
hipError_t launch_unfold_jobs(const UnfoldJob* jobs, uint32_t nJobs, uint32_t blocks, int32_t* xyz, hipStream_t st);
// Fills n 32-bit words with v (graph-safe replacement of a memset).
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st);
// Streaming copy of n16 16-byte elements (bandwidth and counter reference).
hipError_t launch_copy_u4(const void* src, void* dst, uint64_t n16, hipStream_t st);

// Workgroups that cover `items` with `perGroup` items each, rounded up to
// whole XCD rounds (multiples of 8).
inline uint32_t grid_for(uint64_t items, uint32_t perGroup) {
    uint64_t g = (items + perGroup - 1) / perGroup;
    if (g < 1) g = 1;
    g = (g + 7) / 8 * 8;
    return uint32_t(g);
}
// List entries needed for any shape with slots*block <= 1024.
inline uint64_t list_capacity(uint64_t maxColorCount) { return maxColorCount + 8u * 1024u; }
// Workgroups of the finest shape (slots*block = 64).
inline uint64_t max_grid(uint64_t maxColorCount) { return grid_for(maxColorCount, 64); }

}  // namespace pscbfm
