// Shared definitions of the scBFM sweep: the 3-D Z-curve (Morton) codec, the
// counter-based substep draw, the per-monomer move hash and the descriptors
// that the host hands to the gfx950 kernels.
//
// Positions on the device are the Z-curve code of the wrapped lower-left-front
// corner of each 2x2x2 monomer cube.  All arithmetic that the reference does
// on int16 box coordinates (UpdaterGPUScBFM_AB_Type.h:31-32) is done here in
// "dilated" form on that code: a step along one axis is a masked add on the
// bits of that axis, so periodic wrap-around is free and the lattice index of
// the own corner is the stored value itself.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#define PSCBFM_HD __host__ __device__ inline
#define PSCBFM_HOST_ONLY __host__ inline
#else
#define PSCBFM_HD inline
#define PSCBFM_HOST_ONLY inline
#endif

namespace pscbfm {

#if defined(__HIPCC__)
// A pointer into global memory, as the ISA's GLOBAL instructions address it.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
#endif

// A device buffer as a descriptor carries it.  Host code stores and reads the
// address (implicit conversions, host only); device code can reach the
// buffer only through global(), a global-address-space pointer.  A pointer
// read from a descriptor in memory (a batch's descriptor table) would
// otherwise be generic, and the compiler would address it with FLAT
// instructions, whose completion waits on gfx950 proved insufficient for the
// sweep's gathers under load.
template <class T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    PSCBFM_HOST_ONLY DeviceBuffer(T* p) : p_(p) {}
    PSCBFM_HOST_ONLY operator T*() const { return p_; }
#if defined(__HIPCC__)
    __device__ inline gptr<T> global() const { return (gptr<T>)p_; }
#endif

private:
    T* p_ = nullptr;
};

// Upper bound of bonds per monomer.  The reference compiles with 4 and notes
// that the hydrogel coloring example needs 8 (UpdaterGPUScBFM_AB_Type.h:38-39);
// the neighbour matrix here has per-color row counts, so 8 costs nothing for
// systems with fewer bonds.
constexpr uint32_t kMaxConnectivity = 8;
// The reference colors with uint8_t ids and refuses to use more than 254
// colors (graphColoring.tpp:88-90).
constexpr uint32_t kMaxColors = 254;
// Species blocks start on multiples of this many slots (256 B of uint32),
// as the reference aligns each species to 512 B of short4
// (UpdaterGPUScBFM_AB_Type.h:199-201).
constexpr uint32_t kSlotAlign = 64;
// Marks an unused entry of the neighbour matrix.
constexpr uint32_t kNoNeighbor = 0xFFFFFFFFu;

enum LatticeKind : uint32_t { kLatticeU8 = 0, kLatticeBit = 1 };

// ---------------------------------------------------------------------------
// Move hash of CRITICAL.cu:2-11 (Bob Jenkins' 6-shift 32-bit integer hash),
// used as rn = H(H(j) ^ seed), direction = rn % 6 (CRITICAL.cu:321-323).
PSCBFM_HD uint32_t move_hash(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

// ---------------------------------------------------------------------------
// Counter-based substep draw.  The reference draws the species and the
// 32-bit kernel seed from LeMonADE's R250 per substep (CRITICAL.cu:509-510).
// Here the t-th output of a SplitMix64 stream keyed by the system seed gives
// both: species = high word % nColors, seed = low word.  Because it is a pure
// function of (seed, t), a run can be split into any number of calls, graph
// replays or batch slots without changing a single draw.
PSCBFM_HD uint64_t splitmix64_at(uint64_t key, uint64_t t) {
    uint64_t z = key + (t + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct SubstepDraw {
    uint32_t species;
    uint32_t seed;
};

PSCBFM_HD SubstepDraw substep_draw(uint64_t key, uint64_t t, uint32_t nColors) {
    const uint64_t r = splitmix64_at(key, t);
    SubstepDraw d;
    d.species = static_cast<uint32_t>(r >> 32) % nColors;
    d.seed = static_cast<uint32_t>(r);
    return d;
}

// ---------------------------------------------------------------------------
// Dilated (Z-curve) arithmetic.  `m` is a code, `M` the bit mask of one axis.
// x + c for a dilated constant c:  ((m | ~M) + c) & M   (carries ride over
// the bits of the other axes because they are forced to 1).  Adding the
// all-ones pattern M is "minus one" modulo the axis length.
PSCBFM_HD uint32_t dil_add(uint32_t m, uint32_t M, uint32_t c) {
    return ((m | ~M) + c) & M;
}
// (a - b) on one axis, result dilated and wrapped.
PSCBFM_HD uint32_t dil_sub(uint32_t a, uint32_t b, uint32_t M) {
    return ((a & M) - (b & M)) & M;
}

// Low nine bits of a Z-curve code interleave x,y,z three times whenever every
// box edge is at least 8.  The masked differences of two codes in that window
// are the bond vector components modulo 8, i.e. exactly the
// (dx&7, dy&7, dz&7) the reference feeds into linearizeBondVectorIndex
// (CRITICAL.cu:243-261), only bit-interleaved.  The host permutes the
// 512-entry bond table into this order once.
constexpr uint32_t kBondWinX = 0x049u;  // bits 0,3,6
constexpr uint32_t kBondWinY = 0x092u;  // bits 1,4,7
constexpr uint32_t kBondWinZ = 0x124u;  // bits 2,5,8

PSCBFM_HD uint32_t bond_key(uint32_t partner, uint32_t self) {
    return (((partner & kBondWinX) - (self & kBondWinX)) & kBondWinX) |
           (((partner & kBondWinY) - (self & kBondWinY)) & kBondWinY) |
           (((partner & kBondWinZ) - (self & kBondWinZ)) & kBondWinZ);
}

// ---------------------------------------------------------------------------
// Host-side codec: bit positions are assigned round-robin x,y,z while an
// axis still has bits, so a cubic box gives the classic interleave
// D(x) | D(y)<<1 | D(z)<<2 of UpdaterGPUScBFM_AB_Type.cu:210-220 and a
// non-cubic power-of-two box is still a dense bijection onto [0, Lx*Ly*Lz).
struct MortonCodec {
    uint32_t log2L[3] = {0, 0, 0};
    uint32_t mask[3] = {0, 0, 0};

    void init(uint32_t lx, uint32_t ly, uint32_t lz) {
        const uint32_t L[3] = {lx, ly, lz};
        uint32_t rem[3];
        for (int a = 0; a < 3; ++a) {
            uint32_t b = 0;
            while ((1u << b) < L[a]) ++b;
            log2L[a] = b;
            rem[a] = b;
            mask[a] = 0;
        }
        uint32_t pos = 0;
        while (rem[0] + rem[1] + rem[2] > 0) {
            for (int a = 0; a < 3; ++a) {
                if (rem[a] > 0) {
                    mask[a] |= 1u << pos;
                    ++pos;
                    --rem[a];
                }
            }
        }
    }
    static uint32_t deposit(uint32_t v, uint32_t M) {
        uint32_t out = 0;
        for (uint32_t bit = 1; M != 0; bit <<= 1) {
            const uint32_t low = M & (~M + 1u);
            if (v & bit) out |= low;
            M &= M - 1u;
        }
        return out;
    }
    static uint32_t extract(uint32_t m, uint32_t M) {
        uint32_t out = 0;
        for (uint32_t bit = 1; M != 0; bit <<= 1) {
            const uint32_t low = M & (~M + 1u);
            if (m & low) out |= bit;
            M &= M - 1u;
        }
        return out;
    }
    // Wrapped coordinates (already reduced modulo the box) to code.
    uint32_t encode(uint32_t x, uint32_t y, uint32_t z) const {
        return deposit(x, mask[0]) | deposit(y, mask[1]) | deposit(z, mask[2]);
    }
    void decode(uint32_t m, uint32_t* x, uint32_t* y, uint32_t* z) const {
        *x = extract(m, mask[0]);
        *y = extract(m, mask[1]);
        *z = extract(m, mask[2]);
    }
};

// ---------------------------------------------------------------------------
// Device-resident descriptors.

// Per-color ("species") layout, indexed by the species drawn on the device.
// 16-bit partner codes (layout.h): no partner, and "read the full matrix".
constexpr int32_t kIdNone = 32767;
constexpr int32_t kIdEscape = -32768;

struct ColorInfo {
    uint32_t slotOffset;  // first slot of the species in pos/img/flags
    uint32_t count;       // monomers of the species
    uint32_t nbrOffset;   // first element of its neighbour matrix
    uint32_t nbrPitch;    // elements per neighbour row (>= count, padded)
    uint32_t maxDegree;   // rows of the neighbour matrix
    uint32_t idsOffset;   // first word of its members' compressed partner ids
    uint32_t idWords;     // 32-bit words per member there (two 16-bit ids each)
    uint32_t pad_;
};

// Accepted moves are not counted with atomics: workgroup w of the pass-2
// kernel adds its count to its own running total acceptedByBlock[w] (one
// writer per entry, kernels ordered on the stream), and the host sums the
// entries.  Thousands of workgroups adding to one address would serialise on
// a single memory channel.
constexpr uint32_t kMaxPerformBlocks = 1u << 16;

// Mutable per-system state that lives on the device so that a captured
// graph can replay any number of substeps without host involvement.
struct DevState {
    // The first 32 bytes are what every sweep launch reads at its start, in
    // one batch: the substep to serve, the stop point and the draw source.
    uint64_t substep;          // global substep counter t of the next substep
    uint64_t stopAt;           // substeps >= stopAt are skipped (batched runs)
    // Explicit draws (R250 mode): substep t reads (species, kernel seed) from
    // drawTable[2(t - drawBase)], [2(t - drawBase) + 1]; null = counter draw.
    const uint32_t* drawTable;
    uint64_t drawBase;
    uint32_t* log;             // optional per-substep accepted counts
    uint64_t logFirst;         // substep index stored in log[0]
    uint64_t logCapacity;      // entries in log
    uint64_t* acceptedByBlock; // kMaxPerformBlocks running totals
};
static_assert(sizeof(DevState) == 64, "DevState is one 64-byte line");
static_assert(offsetof(DevState, drawBase) == 24, "the launch words are the first 32 bytes");

// Workgroup assignment of a batched launch: chunk `chunk` of the `chunks`
// workgroups that cover replica `replica` (kNoReplica: idle workgroup).
constexpr uint32_t kNoReplica = 0xFFFFFFFFu;
struct Tile {
    uint32_t replica;
    uint32_t chunk;
    uint32_t chunks;
    uint32_t pad_;
};

// Everything a sweep kernel needs to know about one system.
struct SysDesc {
    DeviceBuffer<uint32_t> pos;           // Z-curve code per slot
    DeviceBuffer<int32_t> imgX;           // periodic image counters per slot
    DeviceBuffer<int32_t> imgY;
    DeviceBuffer<int32_t> imgZ;
    DeviceBuffer<const uint32_t> nbr;     // neighbour matrices (slot ids)
    DeviceBuffer<uint8_t> lattice;        // uint8 per cell or 1 bit per cell
    DeviceBuffer<uint32_t> proposal;      // two 1-bit proposal lattices, latWords each
    DeviceBuffer<const ColorInfo> colors; // nColors entries
    DeviceBuffer<const uint32_t> bondMask;// 16 words: bit k set <=> Z-order bond key k forbidden
    DeviceBuffer<DevState> state;
    DeviceBuffer<uint2> list;             // pass-1 survivors, one list per pass-1 workgroup
    DeviceBuffer<uint32_t> listCount;     // entries of each list
    DeviceBuffer<const uint32_t> ids;     // compressed partner ids (see layout.h)
    DeviceBuffer<const uint32_t> rowBase; // kMaxConnectivity per species: base slot of each row's deltas
    DeviceBuffer<const uint32_t> key;     // per slot: inner move hash of the member's reference index
    uint64_t seed;           // key of the substep draw
    uint32_t mask[3];        // Z-curve bit masks of x, y, z
    uint32_t nColors;
    uint32_t latWords;       // 32-bit words of one 1-bit lattice (cells / 32)
    uint32_t latticeKind;
    uint32_t maxColorCount;  // largest species
    uint32_t maxDegree;      // largest bond count of any monomer
    uint32_t accEntries;     // entries of state->acceptedByBlock
    uint32_t nSlots;         // slots of pos (decoded ids are clamped below it)
    uint32_t keyed;          // 1: slots in spatial order, read the draw keys; 0: hash the slot index
};

}  // namespace pscbfm
