// Species sort and device layout of one system (see layout.h).
// Follows the structure of UpdaterGPUScBFM_AB_Type.cu:547-664: color the bond
// graph, count per color, align each species block, give every monomer its
// slot in stable original order, remap bonds to slots.
#include "layout.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>

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

namespace {

// Byte tables of a Z-curve codec.  Each code bit belongs to one axis and
// stands for one bit of it, so a code is the OR of the deposits of the
// coordinate bytes, and an axis value the OR of the extracts of the code
// bytes.  Every table is built from its eight single-bit entries.
struct ZTables {
    uint32_t dep[3][2][256];  // coordinate byte k of axis a -> code bits (coordinates < 2^16)
    uint16_t ext[3][4][256];  // code byte k -> bits of axis a
    explicit ZTables(const MortonCodec& c) {
        for (int a = 0; a < 3; ++a) {
            for (uint32_t k = 0; k < 2; ++k) {
                uint32_t bit[8];
                for (uint32_t j = 0; j < 8; ++j) bit[j] = MortonCodec::deposit(1u << (8u * k + j), c.mask[a]);
                dep[a][k][0] = 0;
                for (uint32_t b = 1; b < 256; ++b) dep[a][k][b] = dep[a][k][b & (b - 1u)] | bit[__builtin_ctz(b)];
            }
            for (uint32_t k = 0; k < 4; ++k) {
                uint16_t bit[8];
                for (uint32_t j = 0; j < 8; ++j) bit[j] = uint16_t(MortonCodec::extract(1u << (8u * k + j), c.mask[a]));
                ext[a][k][0] = 0;
                for (uint32_t b = 1; b < 256; ++b)
                    ext[a][k][b] = uint16_t(ext[a][k][b & (b - 1u)] | bit[__builtin_ctz(b)]);
            }
        }
    }
    uint32_t encode(uint32_t x, uint32_t y, uint32_t z) const {
        return dep[0][0][x & 0xFFu] | dep[0][1][(x >> 8) & 0xFFu] | dep[1][0][y & 0xFFu] |
               dep[1][1][(y >> 8) & 0xFFu] | dep[2][0][z & 0xFFu] | dep[2][1][(z >> 8) & 0xFFu];
    }
    uint32_t axis(uint32_t m, int a) const {
        return uint32_t(ext[a][0][m & 0xFFu] | ext[a][1][(m >> 8) & 0xFFu] | ext[a][2][(m >> 16) & 0xFFu] |
                        ext[a][3][m >> 24]);
    }
};

// f(lo, hi) over [0, n): on up to 8 threads when n is large (per-monomer
// loops here are memory-bound and independent).
template <class F>
void parallel_for(uint32_t n, F&& f) {
    const uint32_t per = 1u << 20;
    const uint32_t nThreads = std::min<uint32_t>((n + per - 1) / per, 8u);
    if (nThreads <= 1) {
        f(0u, n);
        return;
    }
    std::vector<std::thread> pool;
    const uint32_t chunk = (n + nThreads - 1) / nThreads;
    for (uint32_t t = 1; t < nThreads; ++t) {
        auto part = [&f, t, chunk, n] { f(std::min(n, t * chunk), std::min(n, (t + 1) * chunk)); };
        try {
            pool.emplace_back(part);
        } catch (...) {
            part();  // no thread available: this part on the calling thread
        }
    }
    f(0u, std::min(n, chunk));
    for (std::thread& th : pool) th.join();
}

// 16-bit partner codes, two per word.  Every member of a species has W codes,
// W the system's largest degree rounded up to 2, 4 or 8.  The code of row r of
// member j in species c is the partner's slot minus (rowBase[c][r] + j), where
// rowBase[c][r] is the first slot of the species most row-r partners of c
// belong to: chains and networks built in order keep a partner near the same
// rank in its own species, so the difference is small.  A difference outside
// +-32766 stores kIdEscape (the kernel then reads the full matrix), a missing
// partner kIdNone; padding members hold kIdNone only.
void encode_partner_ids(const SystemInput& in, SystemLayout& L) {
    const uint32_t n = L.nMonomers;
    const uint32_t nColors = uint32_t(L.colors.size());
    uint32_t sysMaxDeg = 0;
    for (const ColorInfo& ci : L.colors) sysMaxDeg = std::max(sysMaxDeg, ci.maxDegree);
    const uint32_t W = sysMaxDeg <= 2u ? 2u : (sysMaxDeg <= 4u ? 4u : 8u);

    std::vector<uint32_t> votes(size_t(nColors) * kMaxConnectivity * nColors, 0u);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < in.bonds.degree(i); ++k)
            ++votes[(size_t(L.color[i]) * kMaxConnectivity + k) * nColors + L.color[in.bonds.neighbor(i, k)]];
    L.rowBase.assign(size_t(nColors) * kMaxConnectivity, 0u);
    for (uint32_t c = 0; c < nColors; ++c)
        for (uint32_t r = 0; r < kMaxConnectivity; ++r) {
            const uint32_t* v = &votes[(size_t(c) * kMaxConnectivity + r) * nColors];
            const uint32_t best = uint32_t(std::max_element(v, v + nColors) - v);
            L.rowBase[size_t(c) * kMaxConnectivity + r] = L.colors[best].slotOffset;
        }

    const uint32_t none = uint32_t(kIdNone) & 0xFFFFu;
    const uint32_t escape = uint32_t(kIdEscape) & 0xFFFFu;
    uint32_t words = 0;
    for (ColorInfo& ci : L.colors) {
        ci.idsOffset = words;
        ci.idWords = W / 2;
        words += ci.nbrPitch * (W / 2);
    }
    L.ids.assign(std::max<uint32_t>(words, 1u), none | (none << 16));
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t c = L.color[i];
        const ColorInfo& ci = L.colors[c];
        const uint32_t j = L.slotOf[i] - ci.slotOffset;
        for (uint32_t r = 0; r < in.bonds.degree(i); ++r) {
            const int64_t v = int64_t(L.slotOf[in.bonds.neighbor(i, r)]) -
                              int64_t(L.rowBase[size_t(c) * kMaxConnectivity + r]) - int64_t(j);
            const uint32_t code = (v > 32766 || v < -32767) ? escape : (uint32_t(int32_t(v)) & 0xFFFFu);
            uint32_t& w = L.ids[ci.idsOffset + size_t(j) * (W / 2) + r / 2];
            w = (r & 1u) ? ((w & 0xFFFFu) | (code << 16)) : ((w & 0xFFFF0000u) | code);
        }
    }
}

// Stable LSD radix sort of packed (code << 32 | rank) words by code, whose
// significant bits are the low `bits`: words of equal code keep their order.
void sort_by_code(std::vector<uint64_t>& a, uint32_t bits) {
    constexpr uint32_t kDigit = 11, kRadix = 1u << kDigit;
    std::vector<uint64_t> tmp(a.size());
    std::vector<size_t> start(kRadix);
    for (uint32_t shift = 0; shift < bits; shift += kDigit) {
        std::fill(start.begin(), start.end(), size_t(0));
        for (uint64_t v : a) ++start[uint32_t(v >> (32 + shift)) & (kRadix - 1)];
        size_t sum = 0;
        for (size_t& c : start) {
            const size_t k = c;
            c = sum;
            sum += k;
        }
        for (uint64_t v : a) tmp[start[uint32_t(v >> (32 + shift)) & (kRadix - 1)]++] = v;
        a.swap(tmp);
    }
}

}  // namespace

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

    // Wrapped starting corners as Z-curve codes with their image counters.
    const ZTables z(L.codec);
    std::vector<uint32_t> code(n);
    std::vector<int32_t> image(3 * size_t(n));
    parallel_for(n, [&](uint32_t lo, uint32_t hi) {
        for (uint32_t i = lo; i < hi; ++i) {
            uint32_t w[3];
            for (int a = 0; a < 3; ++a) {
                const int32_t v = in.xyz[3 * size_t(i) + a];
                const int32_t im = floor_div_pow2(v, L.codec.log2L[a]);
                image[3 * size_t(i) + a] = im;
                w[a] = uint32_t(v - im * int32_t(in.box[a]));
            }
            code[i] = z.encode(w[0], w[1], w[2]);
        }
    });
    // Members of each species in the reference's order, stable by original
    // index as iToiNew (UpdaterGPUScBFM_AB_Type.cu:601-607): a member's rank
    // there is the species-local index that keys its draws.
    std::vector<std::vector<uint32_t>> members(nColors);
    for (uint32_t c = 0; c < nColors; ++c) members[c].reserve(count[c]);
    for (uint32_t i = 0; i < n; ++i) members[L.color[i]].push_back(i);
    bool spatial = in.slotOrder == SlotOrder::Spatial;
    if (in.slotOrder == SlotOrder::Automatic) {
        spatial = in.latticeKind == kLatticeBit;
        if (!spatial) {
            // Automatic, byte lattice: spatial when the reference's order
            // scatters bonded partners (a crosslinked network); for chains
            // the key reads cost more than the locality saves (measured on
            // relaxed melts: 52.6 against 57.4 G attempts/s at 512^3).
            std::vector<uint32_t> rank(n);
            for (uint32_t c = 0; c < nColors; ++c)
                for (uint32_t r = 0; r < members[c].size(); ++r) rank[members[c][r]] = r;
            uint64_t far = 0, total = 0;
            for (uint32_t i = 0; i < n; ++i)
                for (uint32_t k = 0; k < in.bonds.degree(i); ++k) {
                    const int64_t d = int64_t(rank[in.bonds.neighbor(i, k)]) - int64_t(rank[i]);
                    far += (d > 32766 || d < -32766) ? 1u : 0u;
                    ++total;
                }
            spatial = total != 0 && far * 100 >= total;
        }
    }
    // Slots: the members in that order, or sorted by starting code when the
    // order is spatial (codes are distinct: no two corners share a cell).
    // Each slot keeps the inner hash of its member's reference rank.
    std::vector<std::vector<uint32_t>> slots(nColors);
    auto placeSpecies = [&](uint32_t c) {
        std::vector<uint32_t>& order = slots[c];
        order.resize(members[c].size());
        for (uint32_t r = 0; r < order.size(); ++r) order[r] = r;
        if (spatial) {
            std::vector<uint64_t> packed(order.size());
            for (uint32_t r = 0; r < order.size(); ++r) packed[r] = (uint64_t(code[members[c][r]]) << 32) | r;
            sort_by_code(packed, L.codec.log2L[0] + L.codec.log2L[1] + L.codec.log2L[2]);
            for (uint32_t r = 0; r < order.size(); ++r) order[r] = uint32_t(packed[r]);
        }
    };
    {
        std::vector<std::thread> pool;
        const bool threads = n >= (1u << 16);
        for (uint32_t c = 1; c < nColors; ++c) {
            if (threads) {
                try {
                    pool.emplace_back(placeSpecies, c);
                    continue;
                } catch (...) {
                }
            }
            placeSpecies(c);
        }
        placeSpecies(0);
        for (std::thread& t : pool) t.join();
    }
    L.spatialOrder = spatial;
    L.slotOf.assign(n, 0);
    L.monomerOf.assign(L.nSlots, 0xFFFFFFFFu);
    L.key.assign(L.nSlots, 0);
    for (uint32_t c = 0; c < nColors; ++c)
        for (uint32_t p = 0; p < slots[c].size(); ++p) {
            const uint32_t r = slots[c][p], i = members[c][r], s = L.colors[c].slotOffset + p;
            L.slotOf[i] = s;
            L.monomerOf[s] = i;
            L.key[s] = move_hash(r);
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

    encode_partner_ids(in, L);

    // Positions: wrapped corner as Z-curve code, plus image counters so the
    // unfolded coordinate can be rebuilt.
    L.pos.assign(L.nSlots, 0);
    for (int a = 0; a < 3; ++a) L.img[a].assign(L.nSlots, 0);
    parallel_for(n, [&](uint32_t lo, uint32_t hi) {
        for (uint32_t i = lo; i < hi; ++i) {
            const uint32_t s = L.slotOf[i];
            L.pos[s] = code[i];
            for (int a = 0; a < 3; ++a) L.img[a][s] = image[3 * size_t(i) + a];
        }
    });

    std::memset(L.bondMask, 0, sizeof L.bondMask);
    for (uint32_t key = 0; key < 512; ++key)
        if (in.forbidden[bond_key_to_reference_index(key)]) L.bondMask[key >> 5] |= 1u << (key & 31u);
    return L;
}

void SystemLayout::unfolded(const uint32_t* posSlots, const int32_t* const imgSlots[3], int32_t* xyzOut) const {
    const int32_t box[3] = {int32_t(1u << codec.log2L[0]), int32_t(1u << codec.log2L[1]),
                            int32_t(1u << codec.log2L[2])};
    const ZTables z(codec);
    parallel_for(nMonomers, [&](uint32_t lo, uint32_t hi) {
        for (uint32_t i = lo; i < hi; ++i) {
            const uint32_t s = slotOf[i];
            const uint32_t m = posSlots[s];
            for (int a = 0; a < 3; ++a) xyzOut[3 * size_t(i) + a] = int32_t(z.axis(m, a)) + imgSlots[a][s] * box[a];
        }
    });
}

}  // namespace pscbfm
