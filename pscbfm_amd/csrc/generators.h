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

// Crosslinks of a star system (stars of 1 + arms * armLength monomers,
// center first, arm by arm): in each of `rounds` rounds every arm end is
// bonded to the first monomer, in a random order of the 108 bond vectors, of
// another star's arm interior that has no crosslink of that round yet and is
// not already its partner.  Appends the pairs to bonds.
void crosslink_arm_ends(const uint32_t box[3], const std::vector<int32_t>& xyz, uint32_t nStars, uint32_t arms,
                        uint32_t armLength, uint32_t rounds, HostRng& rng, std::vector<uint32_t>& bonds);

// Star network: stars of `arms` arms of `armLength` monomers around one
// center (monomer 0 of each star), each star grown as self-avoiding walks
// from a random free center, then each arm end bonded to a free monomer of
// another star's arm interior within a valid bond vector when one exists.
// bonds receives undirected pairs.
bool build_star_network(const uint32_t box[3], uint32_t nStars, uint32_t arms, uint32_t armLength,
                        uint64_t seed, std::vector<int32_t>& xyz, std::vector<uint32_t>& bonds);

// Dense ordered packing of four-arm stars, for occupied fractions that
// growth cannot reach: corners on the sublattice of even coordinates
// (sites).  A star's center and its +x and -x arms lie on one row of sites
// along x; its third arm starts one site up in y and runs along +x, its
// fourth starts one site up in z and runs along -x.  Rows (y, z) with y + z
// even carry centers, P sites apart; every other row carries the third arms
// of the row below it in y and the fourth arms of the row below it in z,
// the two sets interleaved by shifting the centers of neighbouring rows.
// nStars of the available star places, chosen at random, are filled, in
// row order; bonds receives the bonds inside the stars (no crosslinks: every
// bond between sites joins opposite parities, so crosslinks are added after
// a run has moved the corners off the sublattice).  Returns false if the box
// holds fewer than nStars stars.
bool place_stars_ordered(const uint32_t box[3], uint32_t nStars, uint32_t armLength, uint64_t seed,
                         std::vector<int32_t>& xyz, std::vector<uint32_t>& bonds);

}  // namespace pscbfm
