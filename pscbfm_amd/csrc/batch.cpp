// Batched sweep of many independent replicas (see batch.h).
//
// Device tables: one SysDesc per replica slot, one Tile per workgroup of the
// fixed-size launch grid.  Workgroup w runs on XCD w%8; the tiles of a small
// replica are all taken from one XCD (its lattice then stays in that XCD's
// L2), a large replica is spread over all XCDs.  Because the grid never
// changes, a captured graph stays valid while replicas come and go; only the
// tables are rewritten between launches.
#include "batch.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace pscbfm {

namespace {
constexpr uint32_t kSpreadTiles = 64;  // replicas with at least this many workgroups use every XCD

uint32_t degree_bucket(uint32_t d) { return d <= 2 ? 2u : (d <= 4 ? 4u : 8u); }

// Pass-2 threads per workgroup of a batch with pass-1 block `block`: at most
// 192.  On the 1024-replica mix (pass 1 at 256 x 2) pass 2 with 192 threads
// took 128.6-129.3 ms for 500 MCS, with 128 130.1, with 256 131.2-134.0 and
// with 64 137.5.
uint32_t default_perform_block(uint32_t block) { return block > 192u ? 192u : block; }
}  // namespace

Batch::Batch(int gpu, uint32_t latticeKind, uint32_t tileCapacity, uint32_t replicaCapacity, uint32_t block,
             uint32_t slots)
    : gpu_(gpu), latticeKind_(latticeKind) {
    if (latticeKind != kLatticeU8 && latticeKind != kLatticeBit) throw std::invalid_argument("unknown lattice kind");
    if (block == 0 || block % 64 != 0 || block > 256) throw std::invalid_argument("block must be 64..256, multiple of 64");
    if (slots != 1 && slots != 2 && slots != 4) throw std::invalid_argument("slots must be 1, 2 or 4");
    if (replicaCapacity == 0) throw std::invalid_argument("replica capacity must be positive");
    growable_ = tileCapacity == 0;
    tileCap_ = growable_ ? 8u : (tileCapacity + 7) / 8 * 8;
    replicaCap_ = replicaCapacity;
    shape_.block = block;
    shape_.slots = slots;
    shape_.performBlock = default_perform_block(block);
    shape_.grid = tileCap_;
    hip_check(hipSetDevice(gpu_), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipMalloc(&dDescs_, sizeof(SysDesc) * replicaCap_), "hipMalloc descs");
    hip_check(hipMalloc(&dStops_, sizeof(uint64_t) * replicaCap_), "hipMalloc stops");
    hip_check(hipMalloc(&dGather_, sizeof(uint64_t) * 3 * replicaCap_), "hipMalloc gather");
    jobsCap_ = std::max<uint32_t>(replicaCap_, 64u);
    hip_check(hipMalloc(&dJobs_, sizeof(UnfoldJob) * jobsCap_), "hipMalloc unfold jobs");
    descs_.assign(replicaCap_, SysDesc{});
    replicas_.resize(replicaCap_);
    freeSlots_.assign(8, {});
    const uint32_t cap = tileCap_;
    tileCap_ = 0;
    growTiles(cap);
}

// Tile table, lists and list counts for `needed` workgroups; the new slots
// are free, the old ones keep their replicas.
void Batch::growTiles(uint32_t needed) {
    needed = (needed + 7) / 8 * 8;
    if (needed <= tileCap_) return;
    if (stream_) hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    Tile* tiles = nullptr;
    uint2* lists = nullptr;
    uint32_t* counts = nullptr;
    hip_check(hipMalloc(&tiles, sizeof(Tile) * needed), "hipMalloc tiles");
    hip_check(hipMalloc(&lists, sizeof(uint2) * uint64_t(needed) * shape_.block * shape_.slots), "hipMalloc lists");
    hip_check(hipMalloc(&counts, sizeof(uint32_t) * needed), "hipMalloc list counts");
    if (dTiles_) (void)hipFree(dTiles_);
    if (dLists_) (void)hipFree(dLists_);
    if (dListCounts_) (void)hipFree(dListCounts_);
    dTiles_ = tiles;
    dLists_ = lists;
    dListCounts_ = counts;
    tiles_.resize(needed, Tile{kNoReplica, 0, 0, 0});
    for (uint32_t w = needed; w-- > tileCap_;) freeSlots_[w % 8].push_back(w);
    for (auto& f : freeSlots_) std::sort(f.begin(), f.end(), std::greater<uint32_t>());  // pop_back = lowest
    tileCap_ = needed;
    shape_.grid = tileCap_;
    dropGraphs();
    tablesDirty_ = true;
}

Batch::~Batch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    dropGraphs();
    for (auto& r : replicas_) r.dev.reset();
    for (void* p : {static_cast<void*>(dDescs_), static_cast<void*>(dTiles_), static_cast<void*>(dLists_),
                    static_cast<void*>(dListCounts_), static_cast<void*>(dStops_), static_cast<void*>(dGather_),
                    static_cast<void*>(dJobs_), static_cast<void*>(dXyz_)})
        if (p) (void)hipFree(p);
    if (hXyz_) (void)hipHostFree(hXyz_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Batch::setPerformBlock(uint32_t block) {
    if (block != 0 && (block % 64 != 0 || block > 256)) throw std::invalid_argument("perform block must be 64..256, multiple of 64");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    shape_.performBlock = block ? block : default_perform_block(shape_.block);
    dropGraphs();
}

void Batch::reserveReadBack(uint64_t monomers) {
    const uint64_t want = std::min(monomers, kReadBackRecords);
    if (want <= xyzCap_) return;
    const uint64_t cap = std::min(kReadBackRecords, std::max(want, xyzCap_ + xyzCap_ / 2));
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    if (dXyz_) (void)hipFree(dXyz_);
    if (hXyz_) (void)hipHostFree(hXyz_);
    dXyz_ = nullptr;
    hXyz_ = nullptr;
    xyzCap_ = 0;
    hip_check(hipMalloc(&dXyz_, cap * 12), "hipMalloc read-back");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&hXyz_), cap * 12, hipHostMallocDefault), "hipHostMalloc read-back");
    xyzCap_ = cap;
}

uint32_t Batch::tilesNeeded(const SystemLayout& L) const {
    return std::max<uint32_t>(1u, tiles_for(L, shape_.block * shape_.slots));
}

uint64_t Batch::footprint(const SystemLayout& L) const {
    return DeviceSystem::footprint(L, latticeKind_, false, tilesNeeded(L));
}

uint32_t Batch::add(const SystemLayout& L, uint64_t seed, uint64_t startMcs) {
    uint32_t id = 0;
    while (id < replicaCap_ && replicas_[id].live) ++id;
    if (id == replicaCap_) throw std::runtime_error("batch: replica capacity exhausted");
    const uint32_t nT = tilesNeeded(L);
    if (growable_ && usedTiles_ + nT > tileCap_) growTiles(std::max(usedTiles_ + nT, tileCap_ + tileCap_ / 2));
    if (usedTiles_ + nT > tileCap_) {
        std::ostringstream msg;
        msg << "batch: tile capacity exhausted (" << usedTiles_ << " + " << nT << " > " << tileCap_ << ")";
        throw std::runtime_error(msg.str());
    }
    hip_check(hipSetDevice(gpu_), "hipSetDevice");
    Replica& r = replicas_[id];
    r = Replica{};
    r.layout = L;
    r.seed = seed;
    r.dev.reset(new DeviceSystem());
    r.dev->allocate(L, latticeKind_, seed, /*withLists=*/false, nT, stream_);
    const uint64_t start = startMcs * uint64_t(L.colors.size());
    DevState st = r.dev->readState(stream_);
    st.substep = start;
    st.stopAt = start;  // idle until a target is set
    r.dev->writeState(st, stream_);

    // Workgroup slots: one XCD for a small replica, all XCDs for a large one.
    r.tileSlots.reserve(nT);
    if (nT >= kSpreadTiles) {
        uint32_t x = 0;
        while (r.tileSlots.size() < nT) {
            for (uint32_t tries = 0; tries < 8 && freeSlots_[x].empty(); ++tries) x = (x + 1) % 8;
            r.tileSlots.push_back(freeSlots_[x].back());
            freeSlots_[x].pop_back();
            x = (x + 1) % 8;
        }
    } else {
        while (r.tileSlots.size() < nT) {
            uint32_t best = 0;
            for (uint32_t x = 1; x < 8; ++x)
                if (freeSlots_[x].size() > freeSlots_[best].size()) best = x;
            const uint32_t take = std::min<uint32_t>(nT - uint32_t(r.tileSlots.size()), uint32_t(freeSlots_[best].size()));
            for (uint32_t k = 0; k < take; ++k) {
                r.tileSlots.push_back(freeSlots_[best].back());
                freeSlots_[best].pop_back();
            }
        }
    }
    for (uint32_t c = 0; c < nT; ++c) tiles_[r.tileSlots[c]] = Tile{id, c, nT, 0};
    usedTiles_ += nT;
    descs_[id] = r.dev->desc;
    r.live = true;
    r.pendingStop = start;
    r.info = BatchReplicaInfo{};
    r.info.live = true;
    r.info.substep = start;
    r.info.stopAt = start;
    r.info.nColors = uint32_t(L.colors.size());
    r.info.nMonomers = L.nMonomers;
    r.info.tiles = nT;
    r.info.deviceBytes = r.dev->bytes;
    uint64_t liveMonomers = 0;
    for (const Replica& q : replicas_)
        if (q.live) liveMonomers += q.layout.nMonomers;
    reserveReadBack(liveMonomers);
    const uint32_t bucket = degree_bucket(r.dev->desc.maxDegree);
    if (bucket > graphDegree_) {
        graphDegree_ = bucket;
        // an earlier launch of a cached exec may still be running
        hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        dropGraphs();
    }
    tablesDirty_ = true;
    return id;
}

void Batch::remove(uint32_t id) {
    if (id >= replicaCap_ || !replicas_[id].live) throw std::invalid_argument("batch: no such replica");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    Replica& r = replicas_[id];
    for (uint32_t w : r.tileSlots) {
        tiles_[w] = Tile{kNoReplica, 0, 0, 0};
        freeSlots_[w % 8].push_back(w);
    }
    for (auto& f : freeSlots_) std::sort(f.begin(), f.end(), std::greater<uint32_t>());
    usedTiles_ -= uint32_t(r.tileSlots.size());
    descs_[id] = SysDesc{};
    // the tables must not point at freed memory when the next launch runs
    tablesDirty_ = true;
    uploadTables();
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    r.dev.reset();
    r = Replica{};
}

void Batch::setTarget(uint32_t id, uint64_t nMCS) {
    if (id >= replicaCap_ || !replicas_[id].live) throw std::invalid_argument("batch: no such replica");
    Replica& r = replicas_[id];
    r.pendingStop = r.info.substep + nMCS * r.info.nColors;
    r.info.stopAt = r.pendingStop;
    stopsDirty_ = true;
}

void Batch::uploadTables() {
    if (tablesDirty_) {
        hip_check(hipMemcpyAsync(dDescs_, descs_.data(), sizeof(SysDesc) * replicaCap_, hipMemcpyHostToDevice, stream_),
                  "descs H2D");
        hip_check(hipMemcpyAsync(dTiles_, tiles_.data(), sizeof(Tile) * tileCap_, hipMemcpyHostToDevice, stream_),
                  "tiles H2D");
        // the host tables may change as soon as this returns
        hip_check(hipStreamSynchronize(stream_), "sync tables");
        tablesDirty_ = false;
        uint32_t high = 0;
        for (uint32_t w = tileCap_; w-- > 0;)
            if (tiles_[w].replica != kNoReplica) {
                high = w + 1;
                break;
            }
        uint32_t grid = 8;
        while (grid < high) grid *= 2;
        activeGrid_ = std::min(grid, tileCap_);
    }
    if (stopsDirty_) {
        std::vector<uint64_t> stops(replicaCap_, 0);
        for (uint32_t i = 0; i < replicaCap_; ++i)
            if (replicas_[i].live) stops[i] = replicas_[i].pendingStop;
        hip_check(hipMemcpyAsync(dStops_, stops.data(), sizeof(uint64_t) * replicaCap_, hipMemcpyHostToDevice, stream_),
                  "stops H2D");
        hip_check(launch_set_stops(dDescs_, replicaCap_, dStops_, stream_), "set stops");
        hip_check(hipStreamSynchronize(stream_), "sync stops");  // `stops` is a local buffer
        stopsDirty_ = false;
    }
}

// Callers make sure no launch of a cached exec is still running: every path
// that drops graphs (growTiles, setPerformBlock, add raising the degree
// bucket, the destructor) synchronizes the stream first.
void Batch::dropGraphs() {
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
}

hipGraphExec_t Batch::graphFor(uint32_t grid, uint32_t substeps) {
    const auto key = std::make_pair(grid, substeps);
    auto it = graphs_.find(key);
    if (it != graphs_.end()) return it->second;
    BatchLaunch b{dDescs_, dTiles_, dLists_, dListCounts_, grid, graphDegree_, latticeKind_};
    hipGraph_t g = nullptr;
    hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipError_t err = hipSuccess;
    for (uint32_t k = 0; k < substeps && err == hipSuccess; ++k) {
        err = launch_check_batch(b, k, shape_, stream_);
        if (err == hipSuccess) err = launch_perform_batch(b, k, shape_, stream_);
    }
    if (err == hipSuccess) err = launch_advance_batch(dDescs_, replicaCap_, substeps, stream_);
    const hipError_t endErr = hipStreamEndCapture(stream_, &g);
    if (err != hipSuccess) hip_fail(err, "kernel launch during capture");
    hip_check(endErr, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hip_check(ie, "hipGraphInstantiate");
    graphs_[key] = exec;
    // the first launch of an exec would otherwise upload it
    hip_check(hipGraphUpload(exec, stream_), "hipGraphUpload");
    return exec;
}

Batch::GraphPlan Batch::planGraphs(uint64_t substeps) {
    GraphPlan p;
    const uint32_t grid = activeGrid_;
    p.bulkCount = substeps / kGraphSubsteps;
    if (p.bulkCount) p.bulk = graphFor(grid, kGraphSubsteps);
    uint64_t left = substeps - p.bulkCount * kGraphSubsteps;
    // the binary digits of the remainder, each below the bulk size
    for (uint32_t g = kGraphSubsteps / 2; g >= 1u && left > 0; g >>= 1)
        if (left >= g) {
            p.tail.push_back(graphFor(grid, g));
            left -= g;
        }
    return p;
}

void Batch::enqueue(const GraphPlan& plan, uint64_t substeps) {
    for (uint64_t r = 0; r < plan.bulkCount; ++r) hip_check(hipGraphLaunch(plan.bulk, stream_), "hipGraphLaunch");
    for (hipGraphExec_t g : plan.tail) hip_check(hipGraphLaunch(g, stream_), "hipGraphLaunch");
    launched_ += substeps;
}

void Batch::launch(uint64_t substeps) {
    hip_check(hipSetDevice(gpu_), "hipSetDevice");
    uploadTables();
    if (liveReplicas() == 0 || substeps == 0) return;
    enqueue(planGraphs(substeps), substeps);
}

void Batch::synchronize() {
    hip_check(hipSetDevice(gpu_), "hipSetDevice");
    uploadTables();
    hip_check(launch_gather_state(dDescs_, replicaCap_, dGather_, stream_), "gather");
    std::vector<uint64_t> g(3 * size_t(replicaCap_));
    hip_check(hipMemcpyAsync(g.data(), dGather_, g.size() * 8, hipMemcpyDeviceToHost, stream_), "gather D2H");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (uint32_t i = 0; i < replicaCap_; ++i) {
        Replica& r = replicas_[i];
        if (!r.live) continue;
        r.info.substep = g[3 * i];
        r.info.stopAt = g[3 * i + 1];
        r.info.accepted = g[3 * i + 2];
    }
}

double Batch::run(uint64_t nMCS) {
    uint64_t need = 0;
    for (uint32_t i = 0; i < replicaCap_; ++i)
        if (replicas_[i].live) {
            setTarget(i, nMCS);
            need = std::max(need, replicas_[i].pendingStop - replicas_[i].info.substep);
        }
    hip_check(hipSetDevice(gpu_), "hipSetDevice");
    uploadTables();
    // graphs are built before the start event: the timed window holds device work only
    const bool any = liveReplicas() != 0 && need != 0;
    GraphPlan plan;
    if (any) plan = planGraphs(need);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    hip_check(hipEventRecord(e0, stream_), "hipEventRecord");
    if (any) enqueue(plan, need);
    hip_check(hipEventRecord(e1, stream_), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    synchronize();
    for (uint32_t i = 0; i < replicaCap_; ++i)
        if (replicas_[i].live && replicas_[i].info.substep != replicas_[i].info.stopAt)
            throw std::runtime_error("batch: a replica did not reach its target");
    return ms;
}

const BatchReplicaInfo& Batch::info(uint32_t id) const {
    if (id >= replicaCap_ || !replicas_[id].live) throw std::invalid_argument("batch: no such replica");
    return replicas_[id].info;
}

void Batch::positions(uint32_t id, int32_t* xyz, uint32_t n) { positionsMany(&id, 1, &xyz, &n); }

void Batch::positionsMany(const uint32_t* ids, uint32_t count, int32_t* const* xyz, const uint32_t* n) {
    uint64_t total = 0;
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t id = ids[k];
        if (id >= replicaCap_ || !replicas_[id].live) throw std::invalid_argument("batch: no such replica");
        if (n[k] != replicas_[id].layout.nMonomers)
            throw std::invalid_argument("batch: count differs from the replica's monomers");
        if (!xyz[k] && n[k] != 0) throw std::invalid_argument("batch: null output");
        total += n[k];
    }
    if (total == 0) return;
    hip_check(hipSetDevice(gpu_), "hipSetDevice");
    reserveReadBack(total);  // already held for the live replicas, except for repeated ids
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");

    // Rounds of at most xyzCap_ records and jobsCap_ ranges: decode on the
    // device, one copy to pinned memory, then each range to its caller.
    struct Range {
        int32_t* dst;
        uint64_t out;
        uint32_t count;
    };
    std::vector<UnfoldJob> jobs;
    std::vector<Range> ranges;
    uint64_t used = 0;
    uint32_t blocks = 0;
    auto flush = [&] {
        if (jobs.empty()) return;
        hip_check(hipMemcpyAsync(dJobs_, jobs.data(), jobs.size() * sizeof(UnfoldJob), hipMemcpyHostToDevice, stream_),
                  "unfold jobs H2D");
        hip_check(launch_unfold_jobs(dJobs_, uint32_t(jobs.size()), blocks, dXyz_, stream_), "launch unfold");
        hip_check(hipMemcpyAsync(hXyz_, dXyz_, used * 12, hipMemcpyDeviceToHost, stream_), "positions D2H");
        hip_check(hipStreamSynchronize(stream_), "sync");
        auto copyOut = [&](uint32_t part, uint32_t parts) {
            for (size_t q = part; q < ranges.size(); q += parts)
                std::memcpy(ranges[q].dst, hXyz_ + 3 * ranges[q].out, uint64_t(ranges[q].count) * 12);
        };
        // a few threads for large rounds (a host copy runs at a few GB/s per thread)
        const uint32_t parts = used * 12 >= (uint64_t(8) << 20) ? std::min<uint32_t>(8u, uint32_t(ranges.size())) : 1u;
        std::vector<std::thread> pool;
        for (uint32_t p = 1; p < parts; ++p) {
            try {
                pool.emplace_back(copyOut, p, parts);
            } catch (...) {
                for (size_t q = p; q < ranges.size(); q += parts)  // this part on the calling thread
                    std::memcpy(ranges[q].dst, hXyz_ + 3 * ranges[q].out, uint64_t(ranges[q].count) * 12);
            }
        }
        copyOut(0, parts);
        for (std::thread& th : pool) th.join();
        jobs.clear();
        ranges.clear();
        used = 0;
        blocks = 0;
    };
    for (uint32_t k = 0; k < count; ++k) {
        const Replica& r = replicas_[ids[k]];
        for (uint32_t first = 0; first < n[k];) {
            if (used == xyzCap_ || jobs.size() == jobsCap_) flush();
            const uint32_t take = uint32_t(std::min<uint64_t>(n[k] - first, xyzCap_ - used));
            UnfoldJob j = r.dev->unfoldJob(first, take, used);
            j.firstBlock = blocks;
            jobs.push_back(j);
            ranges.push_back(Range{xyz[k] + 3 * size_t(first), used, take});
            used += take;
            blocks += (take + kUnfoldBlock - 1) / kUnfoldBlock;
            first += take;
        }
    }
    flush();
}

uint32_t Batch::liveReplicas() const {
    uint32_t n = 0;
    for (const auto& r : replicas_) n += r.live ? 1u : 0u;
    return n;
}

uint64_t Batch::deviceBytes() const {
    uint64_t b = sizeof(SysDesc) * uint64_t(replicaCap_) + sizeof(Tile) * uint64_t(tileCap_) +
                 sizeof(uint2) * uint64_t(tileCap_) * shape_.block * shape_.slots + 4ull * tileCap_ +
                 32ull * replicaCap_ + sizeof(UnfoldJob) * uint64_t(jobsCap_) + 12 * xyzCap_;
    for (const auto& r : replicas_)
        if (r.live) b += r.dev->bytes;
    return b;
}

}  // namespace pscbfm
