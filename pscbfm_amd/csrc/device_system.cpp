// Device-resident state of one system (see device_system.h).
//
// One hipMalloc, 256-byte aligned pieces:
//   pos      uint32 per slot      Z-curve code of the corner
//   img[3]   int32 per slot       periodic image counters (rarely written)
//   list     uint2 per entry      pass-1 survivors (only with withLists)
//   lcount   uint32 per list
//   nbr      uint32               per-species neighbour matrices (slot ids)
//   lattice  1 bit or 1 byte per cell, Z-curve order
//   proposal 2 x 1 bit per cell   double-buffered by substep parity
//   colors   ColorInfo per species
//   bond     64 bytes             forbidden bond keys (512 bits)
//   state    DevState
//   acc      uint64 per workgroup accepted-move running totals
//   ids      16-bit partner codes, two per word (uint8 lattice only)
//   rowBase  uint32 per (species, row): base slot of the codes (uint8 lattice only)
//   slotOf   uint32 per monomer   slot of each monomer (coordinate read-back)
//   xyz      3 int32 per monomer  unfolded coordinates decoded for read-back (only with withLists)
//   key      uint32 per slot      inner move hash of the member's reference index
#include "device_system.h"

#include <algorithm>
#include <cstring>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>

#include "kernels.h"

namespace pscbfm {

void hip_fail(hipError_t e, const char* what) {
    std::ostringstream msg;
    msg << "HIP error in " << what << ": " << hipGetErrorName(e) << " (" << hipGetErrorString(e) << ")";
    throw std::runtime_error(msg.str());
}

namespace {

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

enum { P_POS, P_IMGX, P_IMGY, P_IMGZ, P_LIST, P_LCOUNT, P_NBR, P_LAT, P_PROP, P_COLORS, P_BOND, P_STATE, P_ACC, P_IDS, P_ROWB,
       P_SLOTOF, P_XYZ, P_KEY, P_N };

struct Plan {
    uint64_t bytes[P_N];
    uint64_t off[P_N];
    uint64_t total;
};

Plan plan(const SystemLayout& L, uint32_t latticeKind, bool withLists, uint32_t accEntries) {
    const uint64_t cells = (uint64_t(1) << L.codec.log2L[0]) << (L.codec.log2L[1] + L.codec.log2L[2]);
    const uint64_t latWords = cells / 32;
    const uint64_t nSlots = std::max<uint64_t>(L.nSlots, 1);
    Plan p{};
    p.bytes[P_POS] = nSlots * 4;
    p.bytes[P_IMGX] = nSlots * 4;
    p.bytes[P_IMGY] = nSlots * 4;
    p.bytes[P_IMGZ] = nSlots * 4;
    p.bytes[P_LIST] = withLists ? list_capacity(L.maxColorCount) * 8 : 8;
    p.bytes[P_LCOUNT] = withLists ? max_grid(L.maxColorCount) * 4 : 4;
    p.bytes[P_NBR] = std::max<uint64_t>(L.nbr.size(), 1) * 4;
    p.bytes[P_LAT] = latticeKind == kLatticeU8 ? cells : latWords * 4;
    p.bytes[P_PROP] = 2 * latWords * 4;
    p.bytes[P_COLORS] = std::max<uint64_t>(L.colors.size(), 1) * sizeof(ColorInfo);
    p.bytes[P_BOND] = 64;
    p.bytes[P_STATE] = sizeof(DevState);
    p.bytes[P_ACC] = std::max<uint64_t>(accEntries, 1) * sizeof(uint64_t);
    // only the uint8-lattice kernel reads the 16-bit partner codes
    const bool codes = latticeKind == kLatticeU8;
    p.bytes[P_IDS] = codes ? std::max<uint64_t>(L.ids.size(), 1) * 4 : 4;
    p.bytes[P_ROWB] = codes ? std::max<uint64_t>(L.rowBase.size(), 1) * 4 : 4;
    // coordinates are decoded on the device; a system with its own launches
    // also holds their output buffer (a batch shares one)
    const uint64_t nMono = std::max<uint64_t>(L.nMonomers, 1);
    p.bytes[P_SLOTOF] = nMono * 4;
    p.bytes[P_XYZ] = withLists ? nMono * 12 : 4;
    // draw keys: read only by systems whose slots are in spatial order
    p.bytes[P_KEY] = L.spatialOrder ? nSlots * 4 : 4;
    p.total = 0;
    for (int i = 0; i < P_N; ++i) {
        p.off[i] = p.total;
        p.total = align_up(p.total + p.bytes[i], 256);
    }
    return p;
}

}  // namespace

uint64_t DeviceSystem::footprint(const SystemLayout& L, uint32_t latticeKind, bool withLists, uint32_t accEntries) {
    return plan(L, latticeKind, withLists, accEntries).total;
}

void DeviceSystem::allocate(const SystemLayout& L, uint32_t latticeKind, uint64_t seed, bool withLists,
                            uint32_t accEntries, hipStream_t stream) {
    release();
    const Plan p = plan(L, latticeKind, withLists, accEntries);
    cells = (uint64_t(1) << L.codec.log2L[0]) << (L.codec.log2L[1] + L.codec.log2L[2]);
    const uint64_t latWords = cells / 32;
    hip_check(hipMalloc(&arena, p.total), "hipMalloc");
    bytes = p.total;
    this->accEntries = std::max<uint32_t>(accEntries, 1);
    char* base = static_cast<char*>(arena);

    desc = SysDesc{};
    desc.pos = reinterpret_cast<uint32_t*>(base + p.off[P_POS]);
    desc.imgX = reinterpret_cast<int32_t*>(base + p.off[P_IMGX]);
    desc.imgY = reinterpret_cast<int32_t*>(base + p.off[P_IMGY]);
    desc.imgZ = reinterpret_cast<int32_t*>(base + p.off[P_IMGZ]);
    desc.list = reinterpret_cast<uint2*>(base + p.off[P_LIST]);
    desc.listCount = reinterpret_cast<uint32_t*>(base + p.off[P_LCOUNT]);
    desc.nbr = reinterpret_cast<const uint32_t*>(base + p.off[P_NBR]);
    desc.lattice = reinterpret_cast<uint8_t*>(base + p.off[P_LAT]);
    desc.proposal = reinterpret_cast<uint32_t*>(base + p.off[P_PROP]);
    desc.colors = reinterpret_cast<const ColorInfo*>(base + p.off[P_COLORS]);
    desc.bondMask = reinterpret_cast<const uint32_t*>(base + p.off[P_BOND]);
    desc.ids = reinterpret_cast<const uint32_t*>(base + p.off[P_IDS]);
    desc.rowBase = reinterpret_cast<const uint32_t*>(base + p.off[P_ROWB]);
    desc.key = reinterpret_cast<const uint32_t*>(base + p.off[P_KEY]);
    desc.keyed = L.spatialOrder ? 1u : 0u;
    desc.nSlots = std::max<uint32_t>(L.nSlots, 1u);
    desc.state = reinterpret_cast<DevState*>(base + p.off[P_STATE]);
    desc.seed = seed;
    for (int a = 0; a < 3; ++a) desc.mask[a] = L.codec.mask[a];
    desc.nColors = uint32_t(L.colors.size());
    desc.latWords = uint32_t(latWords);
    desc.maxDegree = 0;
    for (const ColorInfo& ci : L.colors) desc.maxDegree = std::max(desc.maxDegree, ci.maxDegree);
    desc.latticeKind = latticeKind;
    desc.maxColorCount = L.maxColorCount;
    desc.accEntries = this->accEntries;
    slotOfDev = reinterpret_cast<uint32_t*>(base + p.off[P_SLOTOF]);
    xyzDev = withLists ? reinterpret_cast<int32_t*>(base + p.off[P_XYZ]) : nullptr;

    hip_check(hipMemsetAsync(arena, 0, p.total, stream), "hipMemsetAsync");
    auto up = [&](int piece, const void* src, uint64_t n) {
        if (n) hip_check(hipMemcpyAsync(base + p.off[piece], src, n, hipMemcpyHostToDevice, stream), "H2D");
    };
    up(P_POS, L.pos.data(), uint64_t(L.nSlots) * 4);
    if (L.spatialOrder) up(P_KEY, L.key.data(), uint64_t(L.nSlots) * 4);
    up(P_SLOTOF, L.slotOf.data(), uint64_t(L.nMonomers) * 4);
    up(P_IMGX, L.img[0].data(), uint64_t(L.nSlots) * 4);
    up(P_IMGY, L.img[1].data(), uint64_t(L.nSlots) * 4);
    up(P_IMGZ, L.img[2].data(), uint64_t(L.nSlots) * 4);
    up(P_NBR, L.nbr.data(), uint64_t(L.nbr.size()) * 4);
    up(P_COLORS, L.colors.data(), uint64_t(L.colors.size()) * sizeof(ColorInfo));
    up(P_BOND, L.bondMask, 64);
    if (latticeKind == kLatticeU8) {
        up(P_IDS, L.ids.data(), uint64_t(L.ids.size()) * 4);
        up(P_ROWB, L.rowBase.data(), uint64_t(L.rowBase.size()) * 4);
    }
    // Host staging buffers live until the stream synchronizes below: the
    // copies are asynchronous.
    std::vector<uint8_t> latBytes;
    std::vector<uint32_t> latBits;
    if (latticeKind == kLatticeU8) {
        latBytes.assign(cells, 0);
        for (uint32_t i = 0; i < L.nMonomers; ++i) latBytes[L.pos[L.slotOf[i]]] = 1;
        up(P_LAT, latBytes.data(), cells);
    } else {
        latBits.assign(latWords, 0);
        for (uint32_t i = 0; i < L.nMonomers; ++i) {
            const uint32_t c = L.pos[L.slotOf[i]];
            latBits[c >> 5] |= 1u << (c & 31u);
        }
        up(P_LAT, latBits.data(), latWords * 4);
    }
    DevState st{};
    st.stopAt = std::numeric_limits<uint64_t>::max();
    st.acceptedByBlock = reinterpret_cast<uint64_t*>(base + p.off[P_ACC]);
    up(P_STATE, &st, sizeof st);
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize (upload)");
}

void DeviceSystem::release() {
    if (arena) (void)hipFree(arena);
    arena = nullptr;
    bytes = 0;
    desc = SysDesc{};
    slotOfDev = nullptr;
    xyzDev = nullptr;
}

DevState DeviceSystem::readState(hipStream_t stream) const {
    DevState s{};
    hip_check(hipMemcpyAsync(&s, desc.state, sizeof s, hipMemcpyDeviceToHost, stream), "state D2H");
    hip_check(hipStreamSynchronize(stream), "sync");
    return s;
}

void DeviceSystem::writeState(const DevState& s, hipStream_t stream) {
    hip_check(hipMemcpyAsync(desc.state, &s, sizeof s, hipMemcpyHostToDevice, stream), "state H2D");
    hip_check(hipStreamSynchronize(stream), "sync");
}

uint64_t DeviceSystem::acceptedTotal(uint32_t entries, hipStream_t stream) const {
    const DevState s = readState(stream);
    std::vector<uint64_t> v(std::min(entries, accEntries));
    if (!v.empty()) {
        hip_check(hipMemcpyAsync(v.data(), s.acceptedByBlock, v.size() * 8, hipMemcpyDeviceToHost, stream), "D2H");
        hip_check(hipStreamSynchronize(stream), "sync");
    }
    uint64_t total = 0;
    for (uint64_t x : v) total += x;
    return total;
}

void DeviceSystem::positions(const SystemLayout& L, int32_t* xyzOut, hipStream_t stream) const {
    if (!L.nMonomers) return;
    if (!xyzDev) throw std::logic_error("positions: this system decodes into its batch's buffer");
    // decoded on the device in the caller's order: 12 bytes per monomer
    // cross the bus and the host only receives them
    hip_check(launch_unfold(desc, slotOfDev, L.nMonomers, xyzDev, stream), "launch unfold");
    hip_check(hipMemcpyAsync(xyzOut, xyzDev, uint64_t(L.nMonomers) * 12, hipMemcpyDeviceToHost, stream),
              "positions D2H");
    hip_check(hipStreamSynchronize(stream), "sync");
}

UnfoldJob DeviceSystem::unfoldJob(uint32_t first, uint32_t count, uint64_t out) const {
    UnfoldJob j{};
    j.pos = desc.pos;
    j.img[0] = desc.imgX;
    j.img[1] = desc.imgY;
    j.img[2] = desc.imgZ;
    j.slotOf = slotOfDev;
    for (int a = 0; a < 3; ++a) j.mask[a] = desc.mask[a];
    j.first = first;
    j.count = count;
    j.firstBlock = 0;
    j.out = out;
    return j;
}

}  // namespace pscbfm
