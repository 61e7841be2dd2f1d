// Host engine of the scBFM sweep (see updater.h).
#include "updater.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <sstream>
#include <stdexcept>

#include "autotune.h"
#include "noalloc.h"

namespace pscbfm {

// Substeps of explicit draws kept on the device at once in R250 mode (32 MiB).
constexpr uint64_t kMaxDrawSubsteps = uint64_t(1) << 22;

namespace {

bool is_pow2(uint32_t v) { return v != 0 && (v & (v - 1)) == 0; }

}  // namespace

Updater::Updater() { std::memset(forbidden_, 0, sizeof forbidden_); }

Updater::~Updater() {
    try {
        freeDevice();
    } catch (...) {
    }
}

// UpdaterGPUScBFM_AB_Type.cu:472-487
void Updater::setGpu(int id) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
    if (id < 0 || id >= count) {
        std::ostringstream msg;
        msg << "[UpdaterGPUScBFM_AB_Type::setGpu] GPU with ID " << id << " not present. Only " << count
            << " GPUs are available.";
        throw std::invalid_argument(msg.str());
    }
    hip_check(hipSetDevice(id), "hipSetDevice");
    gpu_ = id;
}

// UpdaterGPUScBFM_AB_Type.cu:693-697 (table index of CRITICAL.cu:243-261)
void Updater::copyBondSet(int dx, int dy, int dz, bool forbidden) {
    forbidden_[(uint32_t(dx) & 7u) | ((uint32_t(dy) & 7u) << 3) | ((uint32_t(dz) & 7u) << 6)] = forbidden ? 1 : 0;
}

// UpdaterGPUScBFM_AB_Type.cu:699-726
void Updater::setNrOfAllMonomers(uint32_t n) {
    if (n_ != 0 || initialized_) {
        std::ostringstream msg;
        msg << "[UpdaterGPUScBFM_AB_Type::setNrOfAllMonomers] Number of Monomers already set to " << n_ << "!";
        throw std::runtime_error(msg.str());
    }
    if (n > (1u << 28))
        throw std::invalid_argument("[UpdaterGPUScBFM_AB_Type::setNrOfAllMonomers] at most 2^28 monomers are supported");
    n_ = n;
    nSet_ = true;
    xyz_.assign(3 * size_t(n), 0);
    attr_.assign(n, 0);
    degree_.assign(n, 0);
    nbrIds_.assign(size_t(n) * kMaxConnectivity, 0);
}

void Updater::requireMonomers(const char* who) const {
    if (!nSet_) {
        std::ostringstream msg;
        msg << "[UpdaterGPUScBFM_AB_Type::" << who << "] setNrOfAllMonomers has not been called";
        throw std::runtime_error(msg.str());
    }
}

void Updater::requireIndex(uint32_t i, const char* who) const {
    if (i >= n_) {
        std::ostringstream msg;
        msg << "[UpdaterGPUScBFM_AB_Type::" << who << "] monomer index " << i << " out of range (" << n_
            << " monomers)";
        throw std::out_of_range(msg.str());
    }
}

// UpdaterGPUScBFM_AB_Type.cu:755-758 (stored, not used by athermal dynamics)
void Updater::setAttribute(uint32_t i, int32_t a) {
    requireMonomers("setAttribute");
    requireIndex(i, "setAttribute");
    attr_[i] = a;
}

// UpdaterGPUScBFM_AB_Type.cu:760-771
void Updater::setMonomerCoordinates(uint32_t i, int32_t x, int32_t y, int32_t z) {
    requireMonomers("setMonomerCoordinates");
    requireIndex(i, "setMonomerCoordinates");
    if (initialized_)
        throw std::runtime_error(
            "[UpdaterGPUScBFM_AB_Type::setMonomerCoordinates] coordinates are owned by the device after initialize");
    xyz_[3 * size_t(i) + 0] = x;
    xyz_[3 * size_t(i) + 1] = y;
    xyz_[3 * size_t(i) + 2] = z;
}

void Updater::setMonomerCoordinatesBulk(const int32_t* xyz, uint32_t n) {
    requireMonomers("setMonomerCoordinatesBulk");
    if (n != n_) throw std::invalid_argument("[setMonomerCoordinatesBulk] count differs from setNrOfAllMonomers");
    if (initialized_) throw std::runtime_error("[setMonomerCoordinatesBulk] already initialized");
    std::memcpy(xyz_.data(), xyz, sizeof(int32_t) * 3 * size_t(n));
}

// UpdaterGPUScBFM_AB_Type.cu:777-795 (directed: callers register both ways)
void Updater::setConnectivity(uint32_t i, uint32_t j) {
    requireMonomers("setConnectivity");
    requireIndex(i, "setConnectivity");
    requireIndex(j, "setConnectivity");
    if (initialized_) throw std::runtime_error("[UpdaterGPUScBFM_AB_Type::setConnectivity] already initialized");
    if (degree_[i] >= kMaxConnectivity) {
        std::ostringstream msg;
        msg << "[UpdaterGPUScBFM_AB_Type::setConnectivity] The maximum amount of bonds per monomer ("
            << kMaxConnectivity << ") has been exceeded!";
        throw std::invalid_argument(msg.str());
    }
    nbrIds_[size_t(i) * kMaxConnectivity + degree_[i]++] = j;
}

void Updater::setConnectivityBulk(const uint32_t* pairs, uint64_t nPairs) {
    for (uint64_t p = 0; p < nPairs; ++p) setConnectivity(pairs[2 * p], pairs[2 * p + 1]);
}

// UpdaterGPUScBFM_AB_Type.cu:797-848
void Updater::setLatticeSize(uint32_t x, uint32_t y, uint32_t z) {
    if (box_[0] == x && box_[1] == y && box_[2] == z) return;
    if (initialized_) throw std::runtime_error("[UpdaterGPUScBFM_AB_Type::setLatticeSize] already initialized");
    const uint32_t lim = uint32_t(std::numeric_limits<int16_t>::max());
    if (x > lim || y > lim || z > lim) {
        std::ostringstream msg;
        msg << "[UpdaterGPUScBFM_AB_Type::setLatticeSize] The box size (" << x << "," << y << "," << z
            << ") is larger than the internal integer data type for representing positions allow! (-32768 <= "
               "size <= 32767)";
        throw std::invalid_argument(msg.str());
    }
    if (!is_pow2(x) || !is_pow2(y) || !is_pow2(z)) {
        std::ostringstream msg;
        msg << "[UpdaterGPUScBFM_AB_Type::setLatticeSize] Check whether the box size is a power of 2! (" << x
            << "," << y << "," << z << ")";
        throw std::runtime_error(msg.str());
    }
    if (x < 8 || y < 8 || z < 8)
        throw std::runtime_error("[UpdaterGPUScBFM_AB_Type::setLatticeSize] every box edge must be at least 8");
    if (uint64_t(x) * y * z > (uint64_t(1) << 32))
        throw std::runtime_error("[UpdaterGPUScBFM_AB_Type::setLatticeSize] at most 2^32 lattice cells");
    box_[0] = x;
    box_[1] = y;
    box_[2] = z;
    hostLattice_.clear();
}

// UpdaterGPUScBFM_AB_Type.cu:728-753
void Updater::setPeriodicity(bool x, bool y, bool z) {
    if (!x || !y || !z) {
        std::ostringstream msg;
        msg << "[UpdaterGPUScBFM_AB_Type::setPeriodicity] Simulation is intended to use completely periodic "
               "boundary conditions, but setPeriodicity was called with ("
            << x << "," << y << "," << z << ")";
        throw std::invalid_argument(msg.str());
    }
}

// UpdaterGPUScBFM_AB_Type.cu:850-872: corner occupancy on the host.
void Updater::populateLattice() {
    requireMonomers("populateLattice");
    if (box_[0] == 0) throw std::runtime_error("[UpdaterGPUScBFM_AB_Type::populateLattice] setLatticeSize first");
    MortonCodec codec;
    codec.init(box_[0], box_[1], box_[2]);
    hostLattice_.assign(size_t(box_[0]) * box_[1] * box_[2], 0);
    for (uint32_t i = 0; i < n_; ++i) {
        const uint32_t x = uint32_t(xyz_[3 * size_t(i)]) & (box_[0] - 1);
        const uint32_t y = uint32_t(xyz_[3 * size_t(i) + 1]) & (box_[1] - 1);
        const uint32_t z = uint32_t(xyz_[3 * size_t(i) + 2]) & (box_[2] - 1);
        hostLattice_[codec.encode(x, y, z)] = 1;
    }
}

void Updater::setSeed(uint64_t seed) {
    if (initialized_) throw std::runtime_error("[setSeed] the seed is fixed at initialize");
    seed_ = seed;
}

void Updater::setStartMcs(uint64_t mcs) {
    if (initialized_) throw std::runtime_error("[setStartMcs] the start is fixed at initialize");
    startMcs_ = mcs;
}

void Updater::setLatticeKind(uint32_t kind) {
    if (initialized_) throw std::runtime_error("[setLatticeKind] the lattice kind is fixed at initialize");
    if (kind != kLatticeU8 && kind != kLatticeBit) throw std::invalid_argument("[setLatticeKind] unknown kind");
    latticeKind_ = kind;
}

void Updater::setColors(const uint8_t* colors, uint32_t n) {
    requireMonomers("setColors");
    if (n != n_) throw std::invalid_argument("[setColors] count differs from setNrOfAllMonomers");
    colorOverride_.assign(colors, colors + n);
    hasColorOverride_ = true;
}

void Updater::setLaunchShape(uint32_t block, uint32_t slots) {
    if (block != 0 && (block % 64 != 0 || block > 256))
        throw std::invalid_argument("[setLaunchShape] block must be 64, 128, 192 or 256");
    if (slots != 0 && slots != 1 && slots != 2 && slots != 4)
        throw std::invalid_argument("[setLaunchShape] slots per thread must be 1, 2 or 4");
    blockOverride_ = block;
    slotsOverride_ = slots;
    if (initialized_) applyLaunchShape();
}

void Updater::setPerformBlock(uint32_t block) {
    if (block != 0 && (block % 64 != 0 || block > 256))
        throw std::invalid_argument("[setPerformBlock] block must be 64, 128, 192 or 256");
    performOverride_ = block;
    if (initialized_) applyLaunchShape();
}

void Updater::setPerformWalk(uint32_t group, uint32_t depth) {
    if (group != kLegacyWalk && group != 0 && (group % 64 != 0 || group > 256))
        throw std::invalid_argument("[setPerformWalk] threads per list must be 64, 128, 192 or 256");
    // Depth 2 is not offered: its uint8-lattice build gave wrong results on
    // gfx950 at the 512^3 melt, while depths 1 and 4 are exact.
    if (depth != 0 && depth != 1 && depth != 4)
        throw std::invalid_argument("[setPerformWalk] entries per thread must be 1 or 4");
    walkGroupOverride_ = group;
    walkDepthOverride_ = depth;
    if (initialized_) applyLaunchShape();
}

void Updater::setSlotOrder(SlotOrder order) {
    if (initialized_) throw std::runtime_error("[setSlotOrder] call before initialize");
    if (order != SlotOrder::Reference && order != SlotOrder::Spatial && order != SlotOrder::Automatic)
        throw std::invalid_argument("[setSlotOrder] order must be 0 (reference), 1 (spatial) or 2 (automatic)");
    slotOrder_ = order;
}

void Updater::setClearPass(uint32_t pass) {
    if (pass > 2) throw std::invalid_argument("[setClearPass] pass must be 0 (automatic), 1 or 2");
    clearPassOverride_ = pass;
    if (initialized_) applyLaunchShape();
}

void Updater::applyLaunchShape() {
    shape_.block = blockOverride_ ? blockOverride_ : 256;
    shape_.slots = slotsOverride_ ? slotsOverride_ : 2;
    // Pass 2 walks one list per workgroup unless a list group is set: on the
    // 512^3 melt the multi-list walk was no faster (128 x 1: 89.5 against
    // 89.3 G attempts/s; 64 x 4: 84.1), pass 2 being bound by its writes.
    if (walkGroupOverride_ == 0 || walkGroupOverride_ == kLegacyWalk) {
        shape_.listGroup = 0;
    } else {
        shape_.listGroup = walkGroupOverride_;
        shape_.listDepth = walkDepthOverride_ ? walkDepthOverride_ : 4;
    }
    shape_.performBlock = performOverride_ ? performOverride_ : automatic_perform_block(layout_, shape_.block);
    if (shape_.listGroup) {
        // several lists per workgroup: whole lists, at most 256 threads
        if (!performOverride_) shape_.performBlock = 256;
        if (shape_.listGroup > shape_.passTwoBlock()) shape_.listGroup = shape_.passTwoBlock();
        shape_.performBlock = shape_.passTwoBlock() / shape_.listGroup * shape_.listGroup;
    }
    shape_.grid = grid_for(layout_.maxColorCount, shape_.block * shape_.slots);
    shape_.clearPass = clearPassOverride_ ? clearPassOverride_ : automatic_clear_pass(layout_, latticeKind_);
    performGridMax_ = std::max(performGridMax_, shape_.passTwoGrid());
    dropGraphs();
}

void Updater::dropGraphs() {
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
}

// UpdaterGPUScBFM_AB_Type.cu:882-986: excluded volume (8 cells per monomer,
// no overlap) and every bond in the set.
void Updater::checkSystem() const {
    requireMonomers("checkSystem");
    if (box_[0] == 0) throw std::runtime_error("[UpdaterGPUScBFM_AB_Type::checkSystem] setLatticeSize first");
    const uint64_t cells = uint64_t(box_[0]) * box_[1] * box_[2];
    std::vector<uint8_t> occ(cells, 0);
    for (uint32_t i = 0; i < n_; ++i) {
        for (uint32_t c = 0; c < 8; ++c) {
            const uint32_t x = uint32_t(xyz_[3 * size_t(i)] + int32_t(c & 1u)) & (box_[0] - 1);
            const uint32_t y = uint32_t(xyz_[3 * size_t(i) + 1] + int32_t((c >> 1) & 1u)) & (box_[1] - 1);
            const uint32_t z = uint32_t(xyz_[3 * size_t(i) + 2] + int32_t((c >> 2) & 1u)) & (box_[2] - 1);
            const uint64_t lin = x + uint64_t(box_[0]) * (y + uint64_t(box_[1]) * z);
            if (occ[lin]) {
                std::ostringstream msg;
                msg << "[UpdaterGPUScBFM_AB_Type::checkSystem] excluded volume violated: monomer " << i
                    << " at (" << xyz_[3 * size_t(i)] << "," << xyz_[3 * size_t(i) + 1] << ","
                    << xyz_[3 * size_t(i) + 2] << ") overlaps another monomer";
                throw std::runtime_error(msg.str());
            }
            occ[lin] = 1;
        }
    }
    for (uint32_t i = 0; i < n_; ++i) {
        for (uint32_t k = 0; k < degree_[i]; ++k) {
            const uint32_t j = nbrIds_[size_t(i) * kMaxConnectivity + k];
            // Minimum image: a network bond can join monomers whose unfolded
            // coordinates differ by whole box lengths; the sweep only uses the
            // vector modulo 8, which every box edge is a multiple of.
            int32_t d[3];
            for (int a = 0; a < 3; ++a) {
                const int32_t L = int32_t(box_[a]);
                const int32_t raw = xyz_[3 * size_t(j) + a] - xyz_[3 * size_t(i) + a];
                d[a] = int32_t(uint32_t(raw + L / 2) & uint32_t(L - 1)) - L / 2;
            }
            const bool tooLong = std::abs(d[0]) > 3 || std::abs(d[1]) > 3 || std::abs(d[2]) > 3;
            const uint32_t idx =
                (uint32_t(d[0]) & 7u) | ((uint32_t(d[1]) & 7u) << 3) | ((uint32_t(d[2]) & 7u) << 6);
            if (tooLong || forbidden_[idx]) {
                std::ostringstream msg;
                msg << "[UpdaterGPUScBFM_AB_Type::checkSystem] forbidden bond (" << d[0] << "," << d[1] << ","
                    << d[2] << ") between monomer " << i + 1 << " and monomer " << j + 1;
                throw std::runtime_error(msg.str());
            }
        }
    }
}

void Updater::freeDevice() {
    dropGraphs();
    dev_.release();
    if (drawTable_) (void)hipFree(drawTable_);
    drawTable_ = nullptr;
    drawCapacity_ = 0;
    if (stream_) {
        (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
    }
}

// UpdaterGPUScBFM_AB_Type.cu:490-690
SystemInput Updater::stagedInput() const {
    requireMonomers("initialize");
    if (box_[0] == 0) throw std::runtime_error("[UpdaterGPUScBFM_AB_Type::initialize] setLatticeSize first");
    SystemInput in;
    std::memcpy(in.box, box_, sizeof box_);
    in.xyz = xyz_;
    in.bonds.start.assign(size_t(n_) + 1, 0);
    for (uint32_t i = 0; i < n_; ++i) in.bonds.start[i + 1] = in.bonds.start[i] + degree_[i];
    in.bonds.ids.resize(in.bonds.start[n_]);
    for (uint32_t i = 0; i < n_; ++i)
        for (uint32_t k = 0; k < degree_[i]; ++k)
            in.bonds.ids[in.bonds.start[i] + k] = nbrIds_[size_t(i) * kMaxConnectivity + k];
    std::memcpy(in.forbidden, forbidden_, sizeof forbidden_);
    in.slotOrder = slotOrder_;
    in.latticeKind = latticeKind_;
    return in;
}

namespace {
void check_bond_set(const uint8_t forbidden[512]) {
    unsigned nAllowed = 0;
    for (int i = 0; i < 512; ++i) nAllowed += forbidden[i] ? 0u : 1u;
    if (nAllowed != 108) {
        std::ostringstream msg;
        msg << "[UpdaterGPUScBFM_AB_Type::initialize] Wrong bond-set! Expected 108 allowed bonds, but got "
            << nAllowed;
        throw std::runtime_error(msg.str());
    }
}
}  // namespace

SystemLayout layout_of(const StagedSystem& s) {
    check_bond_set(s.input.forbidden);
    return build_layout(s.input, s.colors.empty() ? nullptr : &s.colors, s.uniformColors);
}

StagedSystem Updater::staged() const {
    StagedSystem s;
    s.input = stagedInput();
    if (hasColorOverride_) s.colors = colorOverride_;
    s.uniformColors = uniformColors_;
    s.seed = seed_;
    s.startMcs = startMcs_;
    return s;
}

// The 108-bond check, the optional invariant check, then the species layout.
SystemLayout Updater::buildLayout() const {
    StagedSystem s = staged();
    check_bond_set(s.input.forbidden);
    if (checkOnInit_) checkSystem();
    return layout_of(s);
}

void Updater::initialize() {
    if (initialized_) {
        throw std::runtime_error(
            "[UpdaterGPUScBFM_AB_Type::initialize] Initialize was already called and may not be called again until "
            "cleanup was called!");
    }
    layout_ = buildLayout();

    if (gpu_ < 0) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        gpu_ = dev;
    } else {
        hip_check(hipSetDevice(gpu_), "hipSetDevice");
    }
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");

    cells_ = uint64_t(box_[0]) * box_[1] * box_[2];
    dev_.allocate(layout_, latticeKind_, seed_, /*withLists=*/true, uint32_t(max_grid(layout_.maxColorCount)),
                  stream_);

    performGridMax_ = 0;
    applyLaunchShape();
    substep_ = startMcs_ * nColors();
    if (substep_ != 0) {
        DevState st = dev_.readState(stream_);
        st.substep = substep_;
        dev_.writeState(st, stream_);
    }
    acceptedTotal_ = 0;
    initialized_ = true;
}

uint32_t Updater::graphChunk() const {
    if (graphChunk_) return graphChunk_;
    const uint32_t nc = std::max<uint32_t>(nColors(), 1u);
    return nc * std::max<uint32_t>(1u, 256u / nc);
}

hipGraphExec_t Updater::graphFor(uint32_t g) {
    const auto it = graphs_.find(g);
    if (it != graphs_.end()) return it->second;
    hipGraph_t graph = nullptr;
    hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipError_t err = hipSuccess;
    for (uint32_t k = 0; k < g && err == hipSuccess; ++k) {
        err = launch_check(dev_.desc, k, shape_, stream_);
        if (err == hipSuccess) err = launch_perform(dev_.desc, k, shape_, stream_);
    }
    if (err == hipSuccess) err = launch_advance(dev_.desc.state, g, stream_);
    const hipError_t endErr = hipStreamEndCapture(stream_, &graph);
    if (err != hipSuccess) hip_fail(err, "kernel launch during capture");
    hip_check(endErr, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t instErr = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    hip_check(instErr, "hipGraphInstantiate");
    graphs_[g] = exec;
    // the first launch of an exec would otherwise upload it
    hip_check(hipGraphUpload(exec, stream_), "hipGraphUpload");
    return exec;
}

void Updater::setGraphChunk(uint32_t substeps) {
    if (substeps > kMaxGraphChunk)
        throw std::invalid_argument("[setGraphChunk] at most 4096 substeps per captured graph");
    graphChunk_ = substeps;
}

// The bulk graph as often as it fits, then one graph of exactly the
// remaining substeps: every graph launch costs tens of microseconds of device
// time before its first kernel, so a short run is one launch.  Graphs are
// cached per size; a caller that repeats one run length builds at most two.
Updater::GraphPlan Updater::planGraphs(uint64_t S) {
    GraphPlan p;
    const uint32_t G = graphChunk();
    p.bulkCount = S / G;
    if (p.bulkCount) p.bulk = graphFor(G);
    const uint64_t left = S - p.bulkCount * G;
    if (left) p.tail.push_back(graphFor(uint32_t(left)));
    return p;
}

void Updater::prepareGraphs(int32_t nMCS) {
    if (!initialized_)
        throw std::runtime_error("[prepareGraphs] initialize has not been called");
    if (nMCS <= 0 || nColors() == 0 || n_ == 0) return;
    hip_check(hipSetDevice(gpu_), "hipSetDevice");
    const uint64_t S = uint64_t(nMCS) * nColors();
    const uint64_t drawCap = drawChunk_ ? std::min(drawChunk_, kMaxDrawSubsteps) : kMaxDrawSubsteps;
    const uint64_t chunk = r250Mode_ ? std::max<uint64_t>(1u, drawCap / nColors()) * nColors() : S;
    (void)planGraphs(std::min(chunk, S));
    if (S > chunk && S % chunk) (void)planGraphs(S % chunk);
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Updater::run(int32_t nMCS, uint32_t* log, bool profile, bool useGraph) {
    runImpl(nMCS, log, profile, useGraph, nullptr);
}

// The allocation-free form: the device log and the R250 draw table of the
// call live in the process's scratch array on this GPU (noalloc.hip) instead
// of hipMalloc'd buffers, and every substep is launched eagerly (a graph's
// instantiation allocates).
void Updater::runNoAlloc(int32_t nMCS, uint32_t* log) {
    if (!initialized_)
        throw std::runtime_error("[runNoAlloc] initialize has not been called");
    if (nMCS <= 0 || nColors() == 0 || n_ == 0) {
        lastRun_ = RunStats{};
        return;
    }
    const uint64_t S = uint64_t(nMCS) * nColors();
    if ((log != nullptr || r250Mode_) && S > kNoAllocMaxSubsteps) {
        std::ostringstream msg;
        msg << "[runNoAlloc] " << S << " substeps exceed the scratch maximum of " << kNoAllocMaxSubsteps
            << " substeps per call with a log or R250 draws";
        throw std::invalid_argument(msg.str());
    }
    runImpl(nMCS, log, false, false, noalloc_scratch(gpu_));
}

void Updater::runImpl(int32_t nMCS, uint32_t* log, bool profile, bool useGraph, void* scratch) {
    if (!initialized_)
        throw std::runtime_error("[UpdaterGPUScBFM_AB_Type::runSimulationOnGPU] initialize has not been called");
    lastRun_ = RunStats{};
    if (nMCS <= 0) return;
    hip_check(hipSetDevice(gpu_), "hipSetDevice");
    const uint32_t nc = nColors();
    if (nc == 0 || n_ == 0) return;
    const uint64_t S = uint64_t(nMCS) * nc;
    // R250 mode keeps the draw table of one chunk of whole sweeps on the
    // device at a time (at most kMaxDrawSubsteps substeps, 8 bytes each).
    const uint64_t drawCap = drawChunk_ ? std::min(drawChunk_, kMaxDrawSubsteps) : kMaxDrawSubsteps;
    const uint64_t chunk = r250Mode_ && !scratch ? std::max<uint64_t>(1u, drawCap / nc) * nc : S;
    RunStats total{};
    for (uint64_t done = 0; done < S; done += chunk) {
        const uint64_t part = std::min(chunk, S - done);
        runChunk(part, log ? log + done : nullptr, profile, useGraph, scratch);
        total.gpuMilliseconds += lastRun_.gpuMilliseconds;
        total.substeps += lastRun_.substeps;
        total.attempted += lastRun_.attempted;
        total.bondTerms += lastRun_.bondTerms;
        total.checkMilliseconds += lastRun_.checkMilliseconds;
        total.performMilliseconds += lastRun_.performMilliseconds;
        total.checkLaunches += lastRun_.checkLaunches;
        total.performLaunches += lastRun_.performLaunches;
    }
    lastRun_ = total;
    if (readBack_) syncFromDevice();
    const DevState st = dev_.readState(stream_);
    std::vector<uint64_t> perBlock(performGridMax_);
    hip_check(hipMemcpy(perBlock.data(), st.acceptedByBlock, perBlock.size() * sizeof(uint64_t), hipMemcpyDeviceToHost),
              "accepted D2H");
    uint64_t acc = 0;
    for (uint64_t v : perBlock) acc += v;
    lastRun_.accepted = acc - acceptedTotal_;
    acceptedTotal_ = acc;
    if (st.substep != substep_) throw std::runtime_error("device substep counter out of step with the host");
}

// S substeps from substep_ on; fills lastRun_ except `accepted`.  With
// `scratch` (the allocation-free form) the draw table sits at its start and
// the device log kNoAllocDrawBytes in, nothing is allocated, and the launches
// are eager and timed on the host.
void Updater::runChunk(uint64_t S, uint32_t* log, bool profile, bool useGraph, void* scratch) {
    lastRun_ = RunStats{};
    const uint32_t nc = nColors();

    // Every graph the replays need is captured, instantiated and uploaded
    // first: before the start event, so that the timed window holds device
    // work only, and before this run enqueues anything on the stream.
    GraphPlan plan;
    if (!profile && useGraph) plan = planGraphs(S);

    // Species and kernel seed of every substep, as the kernels will draw them.
    std::vector<uint32_t> species(S);
    std::vector<uint32_t> draws;
    uint32_t* drawsOnDevice = nullptr;
    if (r250Mode_) {
        // CRITICAL.cu:509-510: species first, then the kernel seed
        draws.resize(2 * S);
        for (uint64_t k = 0; k < S; ++k) {
            draws[2 * k] = r250_.next() % nc;
            draws[2 * k + 1] = r250_.next();
            species[k] = draws[2 * k];
        }
        if (!scratch && drawCapacity_ < S) {
            if (drawTable_) (void)hipFree(drawTable_);
            drawTable_ = nullptr;
            drawCapacity_ = 0;
            hip_check(hipMalloc(&drawTable_, S * 8), "hipMalloc (draw table)");
            drawCapacity_ = S;
        }
        drawsOnDevice = scratch ? static_cast<uint32_t*>(scratch) : drawTable_;
        hip_check(hipMemcpyAsync(drawsOnDevice, draws.data(), S * 8, hipMemcpyHostToDevice, stream_),
                  "draw table H2D");
    } else {
        for (uint64_t k = 0; k < S; ++k) species[k] = substep_draw(seed_, substep_ + k, nc).species;
    }

    uint32_t* dLog = nullptr;
    if (log) {
        if (scratch)  // the last call's counts are still there: cleared by a kernel
            dLog = reinterpret_cast<uint32_t*>(static_cast<char*>(scratch) + kNoAllocDrawBytes);
        else
            hip_check(hipMalloc(&dLog, S * 4), "hipMalloc (log)");
        hip_check(launch_fill_u32(dLog, S, 0u, stream_), "fill log");
    }
    const bool patched = log != nullptr || r250Mode_;
    if (patched) {
        DevState patch = dev_.readState(stream_);
        if (log) {
            patch.log = dLog;
            patch.logFirst = substep_;
            patch.logCapacity = S;
        }
        if (r250Mode_) {
            // the table covers substeps [substep_, substep_ + S); none beyond runs
            patch.drawTable = drawsOnDevice;
            patch.drawBase = substep_;
            patch.stopAt = substep_ + S;
        }
        dev_.writeState(patch, stream_);
    }

    // The allocation-free form is timed on the host (launch to completion)
    // and creates no events.
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const auto h0 = std::chrono::steady_clock::now();
    if (!scratch) {
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, stream_), "hipEventRecord");
    }
    std::vector<hipEvent_t> kev;
    if (profile) {
        kev.resize(3 * S);
        for (auto& e : kev) hip_check(hipEventCreate(&e), "hipEventCreate");
        for (uint64_t k = 0; k < S; ++k) {
            hip_check(hipEventRecord(kev[3 * k], stream_), "hipEventRecord");
            hip_check(launch_check(dev_.desc, 0, shape_, stream_), "launch pass 1");
            hip_check(hipEventRecord(kev[3 * k + 1], stream_), "hipEventRecord");
            hip_check(launch_perform(dev_.desc, 0, shape_, stream_), "launch pass 2");
            hip_check(hipEventRecord(kev[3 * k + 2], stream_), "hipEventRecord");
            hip_check(launch_advance(dev_.desc.state, 1, stream_), "launch advance");
        }
    } else {
        uint64_t left = S;
        if (useGraph) {
            // the bulk graph, then the remainder's graph, so every substep
            // runs from a graph
            for (uint64_t r = 0; r < plan.bulkCount; ++r)
                hip_check(hipGraphLaunch(plan.bulk, stream_), "hipGraphLaunch");
            for (hipGraphExec_t g : plan.tail) hip_check(hipGraphLaunch(g, stream_), "hipGraphLaunch");
            left = 0;
        }
        // Eager launches (graphs off) in blocks of at most 2^20 substeps per advance.
        while (left > 0) {
            const uint32_t blk = uint32_t(std::min<uint64_t>(left, 1u << 20));
            for (uint32_t k = 0; k < blk; ++k) {
                hip_check(launch_check(dev_.desc, k, shape_, stream_), "launch pass 1");
                hip_check(launch_perform(dev_.desc, k, shape_, stream_), "launch pass 2");
            }
            hip_check(launch_advance(dev_.desc.state, blk, stream_), "launch advance");
            left -= blk;
        }
    }
    if (scratch) {
        hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        lastRun_.gpuMilliseconds =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    } else {
        hip_check(hipEventRecord(e1, stream_), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        lastRun_.gpuMilliseconds = ms;
    }
    if (profile) {
        for (uint64_t k = 0; k < S; ++k) {
            float a = 0, b = 0;
            hip_check(hipEventElapsedTime(&a, kev[3 * k], kev[3 * k + 1]), "hipEventElapsedTime");
            hip_check(hipEventElapsedTime(&b, kev[3 * k + 1], kev[3 * k + 2]), "hipEventElapsedTime");
            lastRun_.checkMilliseconds += a;
            lastRun_.performMilliseconds += b;
        }
        lastRun_.checkLaunches = S;
        lastRun_.performLaunches = S;
        for (auto& e : kev) (void)hipEventDestroy(e);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);

    // Attempts and bond terms follow from the host copy of the draws: every
    // member of the drawn species is attempted, with all its bonds.
    std::vector<uint64_t> deg(nc, 0);
    for (uint32_t i = 0; i < n_; ++i) deg[layout_.color[i]] += degree_[i];
    for (uint64_t k = 0; k < S; ++k) {
        lastRun_.attempted += layout_.colors[species[k]].count;
        lastRun_.bondTerms += deg[species[k]];
    }
    lastRun_.substeps = S;
    substep_ += S;

    if (log) hip_check(hipMemcpy(log, dLog, S * 4, hipMemcpyDeviceToHost), "log D2H");
    if (patched) {
        DevState patch = dev_.readState(stream_);
        patch.log = nullptr;
        patch.logCapacity = 0;
        patch.drawTable = nullptr;
        patch.drawBase = 0;
        patch.stopAt = std::numeric_limits<uint64_t>::max();
        dev_.writeState(patch, stream_);
    }
    if (dLog && !scratch) (void)hipFree(dLog);
}

void Updater::setR250State(const uint32_t* words250) {
    if (!words250) throw std::invalid_argument("[setR250State] null state");
    r250_.load(words250);
    r250Mode_ = true;
}

void Updater::loadR250SeedFile(const std::string& path, uint32_t discard) {
    const std::vector<uint32_t> w = read_r250_seed_file(path);
    setR250State(w.data());
    for (uint32_t k = 0; k < discard; ++k) (void)r250_.next();
}

void Updater::r250Draw(uint32_t* out, uint64_t n) {
    if (!r250Mode_) throw std::runtime_error("[r250Draw] no R250 state was set");
    for (uint64_t k = 0; k < n; ++k) {
        const uint32_t v = r250_.next();
        if (out) out[k] = v;
    }
}

void Updater::r250State(uint32_t* words250) const {
    if (!r250Mode_) throw std::runtime_error("[r250State] no R250 state was set");
    r250_.state(words250);
}

// CRITICAL.cu:414-698 (synchronous; returns after copy-back and unsort)
void Updater::runSimulationOnGPU(int32_t nMCS) { run(nMCS, nullptr, false, true); }

void Updater::syncFromDevice() { dev_.positions(layout_, xyz_.data(), stream_); }

void Updater::readPositionsBack() {
    if (!initialized_) throw std::runtime_error("[readPositionsBack] initialize has not been called");
    hip_check(hipSetDevice(gpu_), "hipSetDevice");
    syncFromDevice();
}

void Updater::downloadLattice(std::vector<uint8_t>& out) const {
    if (!initialized_) throw std::runtime_error("[downloadLattice] not initialized");
    out.assign(cells_, 0);
    if (latticeKind_ == kLatticeU8) {
        hip_check(hipMemcpy(out.data(), dev_.desc.lattice, cells_, hipMemcpyDeviceToHost), "lattice D2H");
    } else {
        std::vector<uint32_t> w(cells_ / 32);
        hip_check(hipMemcpy(w.data(), dev_.desc.lattice, w.size() * 4, hipMemcpyDeviceToHost), "lattice D2H");
        for (uint64_t c = 0; c < cells_; ++c) out[c] = (w[c >> 5] >> (c & 31u)) & 1u;
    }
}

void Updater::downloadProposal(std::vector<uint32_t>& out) const {
    if (!initialized_) throw std::runtime_error("[downloadProposal] not initialized");
    out.resize(2 * (cells_ / 32));
    hip_check(hipMemcpy(out.data(), dev_.desc.proposal, out.size() * 4, hipMemcpyDeviceToHost), "proposal D2H");
}

void Updater::partnerCodeStats(uint64_t& partners, uint64_t& escapes) const {
    const SystemLayout built = initialized_ ? SystemLayout{} : buildLayout();
    const SystemLayout& L = initialized_ ? layout_ : built;
    partners = 0;
    escapes = 0;
    const uint32_t none = uint32_t(kIdNone) & 0xFFFFu, escape = uint32_t(kIdEscape) & 0xFFFFu;
    for (const ColorInfo& ci : L.colors)
        for (uint64_t w = 0; w < uint64_t(ci.count) * ci.idWords; ++w) {
            const uint32_t word = L.ids[ci.idsOffset + w];
            for (const uint32_t code : {word & 0xFFFFu, word >> 16}) {
                partners += code != none ? 1u : 0u;
                escapes += code == escape ? 1u : 0u;
            }
        }
}

int32_t Updater::getMonomerPositionInX(uint32_t i) const {
    requireIndex(i, "getMonomerPositionInX");
    return xyz_[3 * size_t(i)];
}
int32_t Updater::getMonomerPositionInY(uint32_t i) const {
    requireIndex(i, "getMonomerPositionInY");
    return xyz_[3 * size_t(i) + 1];
}
int32_t Updater::getMonomerPositionInZ(uint32_t i) const {
    requireIndex(i, "getMonomerPositionInZ");
    return xyz_[3 * size_t(i) + 2];
}

void Updater::getPositionsBulk(int32_t* xyz, uint32_t n) const {
    if (n != n_) throw std::invalid_argument("[getPositionsBulk] count differs from the number of monomers");
    std::memcpy(xyz, xyz_.data(), sizeof(int32_t) * 3 * size_t(n));
}

// UpdaterGPUScBFM_AB_Type.cu:455-465.  Unlike the reference this also forgets
// the monomer count, so the setter sequence can be replayed after cleanup.
void Updater::destruct() {
    freeDevice();
    initialized_ = false;
    layout_ = SystemLayout{};
    n_ = 0;
    nSet_ = false;
    xyz_.clear();
    attr_.clear();
    degree_.clear();
    nbrIds_.clear();
    hostLattice_.clear();
    colorOverride_.clear();
    hasColorOverride_ = false;
    substep_ = 0;
    acceptedTotal_ = 0;
}

// UpdaterGPUScBFM_AB_Type.cu:995-1001
void Updater::cleanup() {
    destruct();
    if (gpu_ >= 0) (void)hipDeviceSynchronize();
}

}  // namespace pscbfm
