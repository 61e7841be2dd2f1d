// Many independent systems ("replicas") swept together: one launch per pass
// covers every live replica, each keeping its own substep counter, seed,
// species draw and stop point, so every replica evolves exactly as it would
// alone.  Replicas can join and leave between launches (continuous batching).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "common.h"
#include "device_system.h"
#include "kernels.h"
#include "layout.h"

namespace pscbfm {

struct BatchReplicaInfo {
    bool live = false;
    uint64_t substep = 0;   // substeps done (device counter, as of the last sync)
    uint64_t stopAt = 0;    // the replica idles at this substep
    uint64_t accepted = 0;  // accepted moves since the replica joined
    uint32_t nColors = 0;
    uint32_t nMonomers = 0;
    uint32_t tiles = 0;
    uint64_t deviceBytes = 0;
};

class Batch {
public:
    // tileCapacity: workgroups per launch (rounded up to a multiple of 8), or
    // 0 for a table that starts at 8 and grows as replicas are added (a fixed
    // table rejects a replica that does not fit); replicaCapacity: entries of
    // the descriptor table.
    Batch(int gpu, uint32_t latticeKind, uint32_t tileCapacity, uint32_t replicaCapacity, uint32_t block = 256,
          uint32_t slots = 2);
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    // Device bytes a replica with this layout takes (its own buffers only).
    uint64_t footprint(const SystemLayout& layout) const;
    // Workgroups a replica with this layout needs per launch.
    uint32_t tilesNeeded(const SystemLayout& layout) const;
    // Adds a replica (substep 0, stopped until a target is set); returns its id.
    // Throws std::runtime_error when tile or replica capacity is exhausted.
    // startMcs: age of the replica's state in sweeps; its substep counter
    // starts at startMcs * nColors, as an updater's does.
    uint32_t add(const SystemLayout& layout, uint64_t seed, uint64_t startMcs = 0);
    void remove(uint32_t id);
    // stopAt = (substeps done) + nMCS * nColors for the replica.
    void setTarget(uint32_t id, uint64_t nMCS);
    // Enqueues `substeps` substeps for every live replica (asynchronous).
    void launch(uint64_t substeps);
    // Waits for the queued work and refreshes every replica's counters.
    void synchronize();
    // setTarget(nMCS) for every live replica, launch until all reach their
    // stop, synchronize.  Returns the GPU time of the launches in ms.
    double run(uint64_t nMCS);

    const BatchReplicaInfo& info(uint32_t id) const;
    void positions(uint32_t id, int32_t* xyz, uint32_t n);
    // positions() of `count` replicas: one launch decodes the coordinates of
    // as many of them as the read-back buffer holds on the device, one copy
    // brings them into pinned host memory, and a few host threads hand each
    // replica its range; larger requests take several such rounds.
    void positionsMany(const uint32_t* ids, uint32_t count, int32_t* const* xyz, const uint32_t* n);
    uint32_t liveReplicas() const;
    uint32_t usedTiles() const { return usedTiles_; }
    uint32_t tileCapacity() const { return tileCap_; }
    uint32_t replicaCapacity() const { return replicaCap_; }
    uint64_t deviceBytes() const;
    const LaunchShape& shape() const { return shape_; }
    // Pass-2 threads per workgroup (0 = the pass-1 block).
    void setPerformBlock(uint32_t block);
    uint64_t launchedSubsteps() const { return launched_; }

private:
    struct Replica {
        bool live = false;
        SystemLayout layout;
        std::unique_ptr<DeviceSystem> dev;
        uint64_t seed = 0;
        std::vector<uint32_t> tileSlots;
        BatchReplicaInfo info;
        uint64_t pendingStop = 0;
    };
    void uploadTables();
    // Captured graph of `substeps` substeps over workgroups [0, grid) plus
    // their counter advance, instantiated and uploaded on first use.
    hipGraphExec_t graphFor(uint32_t grid, uint32_t substeps);
    // The replays that cover `substeps` substeps at the current grid:
    // `bulkCount` times the bulk graph, then each graph of `tail` once.  Every
    // graph is built here, before anything of the launch is enqueued.
    struct GraphPlan {
        hipGraphExec_t bulk = nullptr;
        uint64_t bulkCount = 0;
        std::vector<hipGraphExec_t> tail;
    };
    GraphPlan planGraphs(uint64_t substeps);
    void enqueue(const GraphPlan& plan, uint64_t substeps);
    void dropGraphs();
    void growTiles(uint32_t needed);
    // Read-back buffers for `monomers` records (at most kReadBackRecords).
    void reserveReadBack(uint64_t monomers);

    int gpu_;
    uint32_t latticeKind_;
    uint32_t tileCap_, replicaCap_;
    LaunchShape shape_;
    hipStream_t stream_ = nullptr;
    // device tables
    SysDesc* dDescs_ = nullptr;
    Tile* dTiles_ = nullptr;
    uint2* dLists_ = nullptr;
    uint32_t* dListCounts_ = nullptr;
    uint64_t* dStops_ = nullptr;
    uint64_t* dGather_ = nullptr;
    // host mirrors
    std::vector<SysDesc> descs_;
    std::vector<Tile> tiles_;
    std::vector<std::vector<uint32_t>> freeSlots_;  // per XCD
    std::vector<Replica> replicas_;
    uint32_t usedTiles_ = 0;
    bool growable_ = false;
    bool tablesDirty_ = true;
    bool stopsDirty_ = false;
    uint32_t graphDegree_ = 0;
    // Launches cover workgroups [0, activeGrid_): the highest occupied slot
    // rounded up to a power of two, so a sparsely filled table does not pay
    // for its idle workgroups.  One captured graph per grid size in use.
    uint32_t activeGrid_ = 8;
    std::map<std::pair<uint32_t, uint32_t>, hipGraphExec_t> graphs_;  // by (grid, substeps)
    // Substeps of the bulk graph; a launch replays it, then power-of-two
    // graphs cover the remainder.
    static constexpr uint32_t kGraphSubsteps = 64;
    uint64_t launched_ = 0;
    // Coordinate read-back: decoded records on the device, their pinned host
    // copy and the job table, sized to the live replicas up to 64 MiB of
    // records (a request beyond that runs in rounds).
    static constexpr uint64_t kReadBackRecords = (uint64_t(64) << 20) / 12;
    int32_t* dXyz_ = nullptr;
    int32_t* hXyz_ = nullptr;
    uint64_t xyzCap_ = 0;  // records
    UnfoldJob* dJobs_ = nullptr;
    uint32_t jobsCap_ = 0;
};

}  // namespace pscbfm
