// Host-side layout of one system for the sweep kernels: species-sorted slots,
// per-species neighbour matrices, Z-curve positions and image counters.
#pragma once

#include <cstdint>
#include <vector>

#include "coloring.h"
#include "common.h"

namespace pscbfm {

// Slot order inside each species (see SystemInput::slotOrder).
enum class SlotOrder : int { Reference = 0, Spatial = 1, Automatic = 2 };

struct SystemInput {
    uint32_t box[3] = {0, 0, 0};
    std::vector<int32_t> xyz;        // unfolded corner, 3 per monomer
    Adjacency bonds;                 // directed bond lists as registered
    uint8_t forbidden[512];          // reference bond table, index (dx&7)|(dy&7)<<3|(dz&7)<<6
    // Slot order inside each species: Reference = by monomer index (the
    // reference's order); Spatial = by the Z-curve code of the starting
    // corner, so that a wave's members, their partners and their lattice
    // words lie close together; Automatic = Spatial on the bit lattice, and
    // on the byte lattice only when the reference's order scatters bonded
    // partners (their species-local indices differ by more than 32766 for
    // at least 1% of the bonds).  Results do not depend on it: a move's draw
    // is keyed by the member's rank in the reference's order either way.
    SlotOrder slotOrder = SlotOrder::Automatic;
    uint32_t latticeKind = kLatticeBit;  // for SlotOrder::Automatic
};

struct SystemLayout {
    MortonCodec codec;
    uint32_t nMonomers = 0;
    uint32_t nSlots = 0;                 // padded slot count
    std::vector<uint8_t> color;          // per monomer
    std::vector<ColorInfo> colors;       // per species
    std::vector<uint32_t> slotOf;        // monomer -> slot
    std::vector<uint32_t> monomerOf;     // slot -> monomer (UINT32_MAX for padding)
    std::vector<uint32_t> pos;           // per slot, Z-curve code
    // per slot, move_hash of the member's species-local index in the
    // reference's order (CRITICAL.cu:322-323): the inner hash of its draw.
    // In the reference's order that index is the slot's own index in its
    // species, and the kernels hash it instead of reading the key.
    std::vector<uint32_t> key;
    bool spatialOrder = false;
    std::vector<int32_t> img[3];         // per slot, periodic images
    std::vector<uint32_t> nbr;           // neighbour matrices (slot ids)
    // Compressed partner ids: member j of species c has W 16-bit codes at
    // ids[colors[c].idsOffset + j*W/2 ...], two per word, W the system's
    // degree bucket.  Code v of row r means slot rowBase[c*8 + r] + j + v;
    // kIdNone: no partner; kIdEscape: read nbr (the full matrix) instead.
    std::vector<uint32_t> ids;
    std::vector<uint32_t> rowBase;
    uint32_t bondMask[16];               // Z-order bond keys, bit set = forbidden
    uint32_t maxColorCount = 0;

    // Unfolded coordinates of all monomers in original order, from per-slot
    // codes and image counters (nSlots entries each).
    void unfolded(const uint32_t* posSlots, const int32_t* const imgSlots[3], int32_t* xyzOut) const;
};

// Floor division / modulo by a power-of-two box edge.
inline int32_t floor_div_pow2(int32_t v, uint32_t log2L) {
    return v >> log2L;  // arithmetic shift is floor division for two's complement
}

// Builds the layout.  `colorOverride` (optional) replaces the automatic
// coloring; it must be a proper coloring.
SystemLayout build_layout(const SystemInput& in, const std::vector<uint8_t>* colorOverride,
                          bool uniformColors);

// Bond key (Z-order, see bond_key in common.h) -> reference table index.
uint32_t bond_key_to_reference_index(uint32_t key);

}  // namespace pscbfm
