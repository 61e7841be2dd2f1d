// Host engine behind the UpdaterGPUScBFM_AB_Type surface: staging of the
// reference's setter calls, species layout, device buffers, the captured
// substep graph and the synchronous sweep driver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "common.h"
#include "device_system.h"
#include "kernels.h"
#include "layout.h"
#include "r250.h"

namespace pscbfm {

struct RunStats {
    double gpuMilliseconds = 0;     // hipEvent time of the launched substeps
    uint64_t substeps = 0;
    uint64_t attempted = 0;         // sum of species sizes over the substeps
    uint64_t accepted = 0;
    uint64_t bondTerms = 0;         // sum of bonds of the drawn species
    // Per-kernel durations, only filled by a profiled run.
    double checkMilliseconds = 0;
    double performMilliseconds = 0;
    uint64_t checkLaunches = 0;
    uint64_t performLaunches = 0;
};

// A system as staged through the setters, detached from its updater: what a
// batch or a request queue needs to build the layout later, on any thread.
struct StagedSystem {
    SystemInput input;
    std::vector<uint8_t> colors;  // empty: automatic coloring
    bool uniformColors = true;
    uint64_t seed = 0;
    uint64_t startMcs = 0;         // age of the input state in sweeps (resume point)
};

// The 108-bond check of UpdaterGPUScBFM_AB_Type.cu:514-530, then the layout.
SystemLayout layout_of(const StagedSystem& s);

class Updater {
public:
    Updater();
    ~Updater();
    Updater(const Updater&) = delete;
    Updater& operator=(const Updater&) = delete;

    // --- reference surface (UpdaterGPUScBFM_AB_Type.h:316-361) ------------
    void setGpu(int id);
    void copyBondSet(int dx, int dy, int dz, bool forbidden);
    void setNrOfAllMonomers(uint32_t n);
    void setAttribute(uint32_t i, int32_t a);
    void setMonomerCoordinates(uint32_t i, int32_t x, int32_t y, int32_t z);
    void setConnectivity(uint32_t i, uint32_t j);
    void setLatticeSize(uint32_t x, uint32_t y, uint32_t z);
    void setPeriodicity(bool x, bool y, bool z);
    void populateLattice();
    void initialize();
    void runSimulationOnGPU(int32_t nMCS);
    int32_t getMonomerPositionInX(uint32_t i) const;
    int32_t getMonomerPositionInY(uint32_t i) const;
    int32_t getMonomerPositionInZ(uint32_t i) const;
    void checkSystem() const;
    void cleanup();
    void destruct();

    // --- extensions ---------------------------------------------------------
    void setSeed(uint64_t seed);
    // Age of the input state in sweeps: the substep counter, and with it the
    // counter draw, starts at startMcs * nColors, so a run resumed from a
    // checkpoint of age m continues the uninterrupted trajectory.
    void setStartMcs(uint64_t mcs);
    uint64_t startMcs() const { return startMcs_; }
    void setLatticeKind(uint32_t kind);
    void setUniformColoring(bool on) { uniformColors_ = on; }
    void setCheckOnInitialize(bool on) { checkOnInit_ = on; }
    void setColors(const uint8_t* colors, uint32_t n);
    // Threads per workgroup (0 = 256) and monomers per pass-1 thread (0 = 2).
    void setLaunchShape(uint32_t block, uint32_t slots);
    // Substeps of the bulk graph (0 = nColors * max(1, 256 / nColors)); a
    // run replays it, then one graph of the remaining substeps.
    // invalid_argument above kMaxGraphChunk (each substep is two kernel nodes).
    static constexpr uint32_t kMaxGraphChunk = 4096;
    void setGraphChunk(uint32_t substeps);
    // Captures and instantiates every graph a run of nMCS sweeps replays, so
    // that the run's first launch starts at once (a run builds what is
    // missing itself, before its timing starts).
    void prepareGraphs(int32_t nMCS);
    // Pass-2 threads per workgroup (0 = automatic: 128 for systems at least a
    // quarter occupied, else the pass-1 block).
    void setPerformBlock(uint32_t block);
    // Pass-2 walk: `group` threads per pass-1 list (64..256, multiple of 64)
    // with `depth` entries per thread in flight (1 or 4; 0 = 4); group 0 or
    // kLegacyWalk (the default) walks one list per workgroup.
    static constexpr uint32_t kLegacyWalk = 0xFFFFFFFFu;
    void setPerformWalk(uint32_t group, uint32_t depth);
    // Slot order inside each species (before initialize): automatic (the
    // default), spatial (by starting Z-curve code) or the reference's (by
    // monomer index); see SystemInput::slotOrder.  Results do not depend on it.
    void setSlotOrder(SlotOrder order);
    // The pass (1 or 2) that zeroes the other parity's proposal buffer each
    // substep; 0 = automatic (pass 1 on the bit lattice with at most two
    // bonds per monomer, else pass 2).
    void setClearPass(uint32_t pass);
    // R250 mode: from the next run on, every substep draws its species
    // (r250 % nColors) and kernel seed (r250) from this generator in the
    // reference's order, instead of the counter draw of (seed, substep).
    void setR250State(const uint32_t* words250);
    void loadR250SeedFile(const std::string& path, uint32_t discard);
    void r250Draw(uint32_t* out, uint64_t n);
    void r250State(uint32_t* words250) const;
    bool r250Mode() const { return r250Mode_; }
    // Substeps of explicit draws kept on the device at once (0 = 2^22, 32 MiB);
    // a longer run goes through in chunks of whole sweeps.
    void setDrawChunk(uint64_t substeps) { drawChunk_ = substeps; }
    // Whether a run copies the positions back before it returns (default on,
    // as runSimulationOnGPU does).  Off: the getters keep the coordinates of
    // the last read-back until readPositionsBack().
    void setReadBack(bool on) { readBack_ = on; }
    void readPositionsBack();
    void setMonomerCoordinatesBulk(const int32_t* xyz, uint32_t n);
    void setConnectivityBulk(const uint32_t* pairs, uint64_t nPairs);
    void getPositionsBulk(int32_t* xyz, uint32_t n) const;
    // Runs nMCS sweeps; `log` (if not null) receives nMCS*nColors accepted
    // counts, one per substep.  profile=true launches eagerly with an event
    // pair around every kernel to time pass 1 and pass 2 separately.
    void run(int32_t nMCS, uint32_t* log, bool profile, bool useGraph);
    // run(nMCS, log, false, false) without any device allocation: the device
    // log and the R250 draw table of the call live in this process's scratch
    // array on the updater's GPU (noalloc.h), shared by every such call on
    // that GPU, so calls that overlap on one GPU race.  With a log or in R250
    // mode a call may cover at most kNoAllocMaxSubsteps substeps
    // (invalid_argument beyond).  Eager launches; gpuMilliseconds is host time.
    void runNoAlloc(int32_t nMCS, uint32_t* log);

    uint32_t nMonomers() const { return n_; }
    uint32_t nColors() const { return uint32_t(layout_.colors.size()); }
    const SystemLayout& layout() const { return layout_; }
    const RunStats& lastRun() const { return lastRun_; }
    uint64_t substepCounter() const { return substep_; }
    uint64_t acceptedTotal() const { return acceptedTotal_; }
    uint64_t deviceBytes() const { return dev_.bytes; }
    // Host input as staged by the setters, and the validated species layout
    // initialize() would build from it (used to hand a system to a batch).
    SystemInput stagedInput() const;
    SystemLayout buildLayout() const;
    StagedSystem staged() const;
    bool initialized() const { return initialized_; }
    const std::vector<int32_t>& attributes() const { return attr_; }
    bool forbidden(uint32_t bondIndex) const { return forbidden_[bondIndex & 511u] != 0; }
    uint32_t latticeKind() const { return latticeKind_; }
    const LaunchShape& launchShape() const { return shape_; }
    const uint32_t* box() const { return box_; }
    int gpu() const { return gpu_; }
    uint64_t seed() const { return seed_; }
    // Host view of the occupancy lattice (one byte per cell, Z-curve order)
    // as built by populateLattice().
    const std::vector<uint8_t>& hostLattice() const { return hostLattice_; }
    // Copies the device occupancy lattice to `out` (cells bytes, 0/1).
    void downloadLattice(std::vector<uint8_t>& out) const;
    // Copies both proposal lattices (2 * cells / 8 bytes).
    void downloadProposal(std::vector<uint32_t>& out) const;
    // Bonded partners in the species layout (directed) and how many of them
    // have no 16-bit partner code (kIdEscape: the uint8-lattice kernel reads
    // the full neighbour matrix for them).  From the initialized layout, else
    // from the layout initialize() would build; needs no GPU.
    void partnerCodeStats(uint64_t& partners, uint64_t& escapes) const;

private:
    void requireMonomers(const char* who) const;
    void requireIndex(uint32_t i, const char* who) const;
    void freeDevice();
    void runImpl(int32_t nMCS, uint32_t* log, bool profile, bool useGraph, void* scratch);
    void runChunk(uint64_t S, uint32_t* log, bool profile, bool useGraph, void* scratch);
    // Captured graph of `substeps` substeps plus their counter advance,
    // instantiated and uploaded on first use.
    hipGraphExec_t graphFor(uint32_t substeps);
    // The replays that cover S substeps: `bulkCount` times the bulk graph,
    // then each graph of `tail` once (the remainder's, if any).
    struct GraphPlan {
        hipGraphExec_t bulk = nullptr;
        uint64_t bulkCount = 0;
        std::vector<hipGraphExec_t> tail;
    };
    GraphPlan planGraphs(uint64_t S);
    void dropGraphs();
    uint32_t graphChunk() const;
    void syncFromDevice();
    void applyLaunchShape();

    // staging
    int gpu_ = -1;
    uint32_t n_ = 0;
    bool nSet_ = false;
    std::vector<int32_t> xyz_;
    std::vector<int32_t> attr_;
    std::vector<uint32_t> degree_;
    std::vector<uint32_t> nbrIds_;   // kMaxConnectivity per monomer
    uint32_t box_[3] = {0, 0, 0};
    uint8_t forbidden_[512];
    std::vector<uint8_t> hostLattice_;
    std::vector<uint8_t> colorOverride_;
    bool hasColorOverride_ = false;
    uint64_t seed_ = 0;
    uint64_t startMcs_ = 0;
    bool r250Mode_ = false;
    bool readBack_ = true;
    R250 r250_;
    uint32_t latticeKind_ = kLatticeBit;
    bool uniformColors_ = true;
    bool checkOnInit_ = true;
    uint32_t graphChunk_ = 0;   // 0: choose automatically
    uint32_t blockOverride_ = 0, slotsOverride_ = 0, performOverride_ = 0;
    uint32_t walkGroupOverride_ = 0, walkDepthOverride_ = 0;
    uint32_t clearPassOverride_ = 0;
    SlotOrder slotOrder_ = SlotOrder::Automatic;

    // device
    bool initialized_ = false;
    SystemLayout layout_;
    hipStream_t stream_ = nullptr;
    DeviceSystem dev_;
    LaunchShape shape_;
    std::map<uint32_t, hipGraphExec_t> graphs_;  // by substeps per replay
    uint32_t performGridMax_ = 0;  // largest pass-2 grid used (entries of acceptedByBlock)
    uint64_t substep_ = 0;
    uint64_t acceptedTotal_ = 0;
    uint64_t cells_ = 0;
    uint32_t* drawTable_ = nullptr;  // R250 mode: 2 words per substep of one run chunk
    uint64_t drawCapacity_ = 0;      // substeps drawTable_ holds
    uint64_t drawChunk_ = 0;         // 0: kMaxDrawSubsteps
    RunStats lastRun_;
};

}  // namespace pscbfm
