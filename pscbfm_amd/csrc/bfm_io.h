// LeMonADE BFM files (the reference's input and output format, read and
// written there by UpdaterReadBfmFile / AnalyzerWriteBfmFile,
// mainCUDASimulatorGPUScBFM_AB_Type.cpp:154-160).
//
// Keys: !number_of_monomers, !bonds (extra bonds "a b", 1-based),
// !box_x/y/z, !periodic_x/y/z, !set_of_bondvectors ("dx dy dz:id"),
// !attributes ("a-b:value"), !mcs=K followed by strands: "x y z" then one
// byte per bond, each the id of the vector to the next monomer of the strand.
// Strand lines are split positionally: after the z token, one separating
// space, then every remaining byte is a bond id.  gzip-compressed files are
// read transparently.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <ostream>
#include <string>
#include <utility>
#include <vector>

namespace pscbfm {

struct BfmData {
    uint32_t box[3] = {0, 0, 0};
    bool periodic[3] = {true, true, true};
    std::vector<int32_t> xyz;                                // last frame, unfolded, 3 per monomer
    std::vector<std::pair<uint32_t, uint32_t>> bonds;        // undirected, 0-based
    std::map<int, std::array<int32_t, 3>> bondset;           // id -> vector
    std::vector<int32_t> attributes;                         // per monomer (0 where unset)
    uint64_t mcs = 0;                                        // age of the last frame
    uint32_t frames = 0;
    uint32_t nMonomers() const { return uint32_t(xyz.size() / 3); }
};

// Reads every frame and keeps the last (READ_LAST_CONFIG_SAVE).  Throws
// std::runtime_error with the file name and line on malformed input.
BfmData read_bfm(const std::string& path);

// The 108 standard vectors with the ids LeMonADE writes for the scBFM bond
// set (17..124), so written files match reference-written ones.
std::map<int, std::array<int32_t, 3>> standard_bondset_ids();

// Writer of one file: the header (with the topology) once, then frames.
// Bonds (i, i+1) go inside strands; all others are listed under !bonds.
class BfmWriter {
public:
    BfmWriter(const uint32_t box[3], const bool periodic[3], uint32_t nMonomers,
              const std::vector<std::pair<uint32_t, uint32_t>>& bonds,
              std::map<int, std::array<int32_t, 3>> bondset, const std::vector<int32_t>& attributes);
    void header(std::ostream& out) const;
    // Throws if a strand bond's vector is not in the bond set.
    void frame(std::ostream& out, uint64_t mcs, const int32_t* xyz) const;

private:
    uint32_t box_[3];
    bool periodic_[3];
    uint32_t n_;
    std::vector<std::pair<uint32_t, uint32_t>> extra_;
    std::vector<uint8_t> linkedToPrev_;  // monomer i continues the strand of i-1
    std::map<int, std::array<int32_t, 3>> bondset_;
    std::map<std::array<int32_t, 3>, int> idOf_;
    std::vector<int32_t> attributes_;
};

}  // namespace pscbfm
