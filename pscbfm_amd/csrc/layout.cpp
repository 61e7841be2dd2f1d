// Species sort and device layout of one system (see layout.h).
// Follows the structure of UpdaterGPUScBFM_AB_Type.cu:547-664: color the bond
// graph, count per color, align each species block, give every monomer its
// slot in stable original order, remap bonds to slots.
#include "layout.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

namespace pscbfm {

uint32_t bond_key_to_reference_index(uint32_t key) {
    // key bits: x at 0,3,6  y at 1,4,7  z at 2,5,8
    uint32_t dx = 0, dy = 0, dz = 0;
    for (uint32_t b = 0; b < 3; ++b) {
        dx |= ((key >> (3 * b + 0)) & 1u) << b;
        dy |= ((key >> (3 * b + 1)) & 1u) << b;
        dz |= ((key >> (3 * b + 2)) & 1u) << b;
    }
    return dx | (dy << 3) | (dz << 6);
}

SystemLayout build_layout(const SystemInput& in, const std::vector<uint8_t>* colorOverride,
                          bool uniformColors) {
    SystemLayout L;
    const uint32_t n = uint32_t(in.xyz.size() / 3);
    L.nMonomers = n;
    L.codec.init(in.box[0], in.box[1], in.box[2]);

    if (colorOverride) {
        if (colorOverride->size() != n) throw std::invalid_argument("color override has wrong size");
        L.color = *colorOverride;
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t k = 0; k < in.bonds.degree(i); ++k)
                if (L.color[in.bonds.neighbor(i, k)] == L.color[i])
                    throw std::invalid_argument("color override is not a proper coloring: monomer " +
                                                std::to_string(i) + " shares its color with a bond partner");
    } else {
        L.color = color_graph(in.bonds, uniformColors);
    }

    uint32_t nColors = 0;
    for (uint32_t i = 0; i < n; ++i) nColors = std::max<uint32_t>(nColors, uint32_t(L.color[i]) + 1u);
    if (nColors > kMaxColors) throw std::runtime_error("more than 254 species");
    std::vector<uint32_t> count(nColors, 0), maxDeg(nColors, 0);
    for (uint32_t i = 0; i < n; ++i) {
        ++count[L.color[i]];
        maxDeg[L.color[i]] = std::max(maxDeg[L.color[i]], in.bonds.degree(i));
    }
    for (uint32_t c = 0; c < nColors; ++c)
        if (count[c] == 0)
            throw std::invalid_argument("species " + std::to_string(c) +
                                        " has no monomers; colors must be numbered 0..k-1 without gaps");

    L.colors.resize(nColors);
    uint32_t slot = 0, nbrOff = 0;
    L.maxColorCount = 0;
    for (uint32_t c = 0; c < nColors; ++c) {
        const uint32_t padded = (count[c] + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
        ColorInfo& ci = L.colors[c];
        std::memset(&ci, 0, sizeof ci);
        ci.slotOffset = slot;
        ci.count = count[c];
        ci.nbrPitch = padded;
        ci.maxDegree = maxDeg[c];
        ci.nbrOffset = nbrOff;
        slot += padded;
        nbrOff += padded * maxDeg[c];
        L.maxColorCount = std::max(L.maxColorCount, count[c]);
    }
    L.nSlots = slot;

    // Stable placement by original index, as iToiNew in the reference.
    L.slotOf.assign(n, 0);
    L.monomerOf.assign(L.nSlots, 0xFFFFFFFFu);
    std::vector<uint32_t> next(nColors);
    for (uint32_t c = 0; c < nColors; ++c) next[c] = L.colors[c].slotOffset;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t s = next[L.color[i]]++;
        L.slotOf[i] = s;
        L.monomerOf[s] = i;
    }

    // Neighbour matrices: row r of species c holds the r-th bond partner of
    // each member, so a wavefront reads one row contiguously.
    L.nbr.assign(nbrOff, kNoNeighbor);
    for (uint32_t i = 0; i < n; ++i) {
        const ColorInfo& ci = L.colors[L.color[i]];
        const uint32_t j = L.slotOf[i] - ci.slotOffset;
        for (uint32_t k = 0; k < in.bonds.degree(i); ++k)
            L.nbr[ci.nbrOffset + k * ci.nbrPitch + j] = L.slotOf[in.bonds.neighbor(i, k)];
    }

    // Positions: wrapped corner as Z-curve code, plus image counters so the
    // unfolded coordinate can be rebuilt.
    L.pos.assign(L.nSlots, 0);
    for (int a = 0; a < 3; ++a) L.img[a].assign(L.nSlots, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t s = L.slotOf[i];
        uint32_t w[3];
        for (int a = 0; a < 3; ++a) {
            const int32_t v = in.xyz[3 * size_t(i) + a];
            const int32_t im = floor_div_pow2(v, L.codec.log2L[a]);
            L.img[a][s] = im;
            w[a] = uint32_t(v - im * int32_t(in.box[a]));
        }
        L.pos[s] = L.codec.encode(w[0], w[1], w[2]);
    }

    std::memset(L.bondMask, 0, sizeof L.bondMask);
    for (uint32_t key = 0; key < 512; ++key)
        if (in.forbidden[bond_key_to_reference_index(key)]) L.bondMask[key >> 5] |= 1u << (key & 31u);
    return L;
}

void SystemLayout::unfolded(const std::vector<uint32_t>& posSlots, const std::vector<int32_t>* imgSlots,
                            int32_t* xyzOut) const {
    const uint32_t box[3] = {1u << codec.log2L[0], 1u << codec.log2L[1], 1u << codec.log2L[2]};
    for (uint32_t i = 0; i < nMonomers; ++i) {
        const uint32_t s = slotOf[i];
        uint32_t w[3];
        codec.decode(posSlots[s], &w[0], &w[1], &w[2]);
        for (int a = 0; a < 3; ++a)
            xyzOut[3 * size_t(i) + a] = int32_t(w[a]) + imgSlots[a][s] * int32_t(box[a]);
    }
}

}  // namespace pscbfm
