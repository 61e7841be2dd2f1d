// Seeded builders of synthetic scBFM systems (linear chains, star networks).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace pscbfm {

// The 108 bond vectors of the scBFM in a fixed order: for each base vector
// (2,0,0), (2,1,0), (2,1,1), (2,2,1), (3,0,0), (3,1,0) every distinct signed
// permutation, in lexicographic order of (dx,dy,dz).
const std::vector<std::array<int32_t, 3>>& standard_bond_vectors();

// Counter-based stream for host-side generation (SplitMix64 outputs).
struct HostRng {
    uint64_t key;
    uint64_t counter = 0;
    explicit HostRng(uint64_t k) : key(k) {}
    uint64_t next();
    uint32_t below(uint32_t n);  // uniform in [0, n)
};

// Linear chains grown one after another as self-avoiding walks on the
// periodic box: a random free start, then bond vectors drawn uniformly from
// the 108-set until the new 2x2x2 cube is free (every vector is tried in
// random order before a chain is restarted).  Returns false when more than
// maxRestarts chain restarts were needed.  xyz receives unfolded corners,
// chain c occupying monomers [c*len, (c+1)*len).
bool grow_linear_chains(const uint32_t box[3], uint32_t nChains, uint32_t len, uint64_t seed,
                        uint32_t maxRestarts, std::vector<int32_t>& xyz);

// Dense ordered packing: corners on the sublattice (2i, 2j, 3k), visited by
// a serpentine path (rows along x, rows stacked in y, layers in z), whose
// consecutive sites differ by (+-2,0,0), (0,+-2,0) or (0,0,3).  The path is
// cut into segments of `len` sites and nChains segments spread evenly over it
// become the chains.  Returns false if the box holds fewer segments.
bool place_linear_chains_ordered(const uint32_t box[3], uint32_t nChains, uint32_t len,
                                 std::vector<int32_t>& xyz);

// Star network: stars of `arms` arms of `armLength` monomers around one
// center (monomer 0 of each star), placed on an ordered lattice, then each
// arm end bonded to a free monomer of another star's arm within a valid bond
// vector when one exists.  bonds receives undirected pairs (a<b listed once).
bool build_star_network(const uint32_t box[3], uint32_t nStars, uint32_t arms, uint32_t armLength,
                        uint64_t seed, std::vector<int32_t>& xyz, std::vector<uint32_t>& bonds);

}  // namespace pscbfm
