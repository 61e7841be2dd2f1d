// The host modules of the library (generators, species layout and coloring,
// BFM files, R250) and the CPU oracle, built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/sanitize/Makefile) and driven over the
// reference's fixtures and generated systems.  Any out-of-bounds access, leak
// or undefined operation aborts the program; the checks below also assert
// results: layouts decode back to their input, written BFM files read back
// equal, the R250 seed file gives the reference's known answer, and the
// oracle's multi-threaded and shuffled runs equal its single-threaded one.
//
//   host_san GOLDEN_DIR TMP_DIR
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "bfm_io.h"
#include "coloring.h"
#include "generators.h"
#include "layout.h"
#include "r250.h"

extern "C" {
#include "oracle.h"
}

using namespace pscbfm;

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        ++failures;
    }
}

struct Sys {
    uint32_t box[3];
    std::vector<int32_t> xyz;
    std::vector<std::pair<uint32_t, uint32_t>> bonds;  // undirected
    uint32_t n() const { return uint32_t(xyz.size() / 3); }
};

Adjacency adjacency(const Sys& s) {
    std::vector<std::vector<uint32_t>> rows(s.n());
    for (const auto& b : s.bonds) {
        rows[b.first].push_back(b.second);
        rows[b.second].push_back(b.first);
    }
    Adjacency a;
    a.start.assign(1, 0u);
    for (const auto& r : rows) {
        a.ids.insert(a.ids.end(), r.begin(), r.end());
        a.start.push_back(uint32_t(a.ids.size()));
    }
    return a;
}

void forbidden_table(uint8_t forbidden[512]) {
    std::memset(forbidden, 1, 512);
    for (const auto& v : standard_bond_vectors())
        forbidden[(uint32_t(v[0]) & 7u) | ((uint32_t(v[1]) & 7u) << 3) | ((uint32_t(v[2]) & 7u) << 6)] = 0;
}

// Slot order inside each species: spatial layouts have strictly increasing
// codes, reference layouts increasing monomer indices; either way every
// slot's key is move_hash of the member's rank among its species in
// monomer-index order (the draw key of CRITICAL.cu:322-323).
void check_slot_order(const SystemLayout& L, bool spatial, const std::string& name) {
    std::vector<uint32_t> rank(L.nMonomers), next(L.colors.size(), 0);
    for (uint32_t i = 0; i < L.nMonomers; ++i) rank[i] = next[L.color[i]]++;
    bool keys = true, ordered = true;
    for (uint32_t i = 0; i < L.nMonomers; ++i) keys = keys && L.key[L.slotOf[i]] == move_hash(rank[i]);
    for (const ColorInfo& ci : L.colors)
        for (uint32_t p = 1; p < ci.count; ++p) {
            const uint32_t s = ci.slotOffset + p;
            ordered = ordered && (spatial ? L.pos[s - 1] < L.pos[s] : L.monomerOf[s - 1] < L.monomerOf[s]);
        }
    expect(keys, name + ": slot keys are the hashes of the reference ranks");
    expect(ordered, name + (spatial ? ": slots in Z-curve order" : ": slots in monomer order"));
}

// Layout of the system (coloring included) with both coloring modes and
// both slot orders; the per-slot codes and image counters must decode back
// to the input.
std::vector<uint8_t> check_layout(const Sys& s, const std::string& name, bool chainsOnly) {
    SystemInput in;
    std::memcpy(in.box, s.box, sizeof in.box);
    in.xyz = s.xyz;
    in.bonds = adjacency(s);
    forbidden_table(in.forbidden);
    std::vector<uint8_t> colors;
    for (bool spatial : {false, true}) {
        in.slotOrder = spatial ? SlotOrder::Spatial : SlotOrder::Reference;
        const SystemLayout L = build_layout(in, nullptr, true);
        expect(L.spatialOrder == spatial, name + ": the slot order asked for");
        check_slot_order(L, spatial, name);
        std::vector<int32_t> back(s.xyz.size());
        const int32_t* img[3] = {L.img[0].data(), L.img[1].data(), L.img[2].data()};
        L.unfolded(L.pos.data(), img, back.data());
        expect(back == s.xyz, std::string(name) + ": layout in either slot order decodes to its input");
    }
    // the automatic order: spatial on the bit lattice; on the byte lattice
    // only for scattered partners, which chains (partners at adjacent
    // ranks) never have
    in.slotOrder = SlotOrder::Automatic;
    in.latticeKind = kLatticeBit;
    expect(build_layout(in, nullptr, true).spatialOrder, name + ": automatic order on the bit lattice is spatial");
    if (chainsOnly) {
        in.latticeKind = kLatticeU8;
        expect(!build_layout(in, nullptr, true).spatialOrder,
               name + ": automatic order of chains on the byte lattice is the reference's");
    }
    in.slotOrder = SlotOrder::Spatial;
    for (bool uniform : {true, false}) {
        const SystemLayout L = build_layout(in, nullptr, uniform);
        std::vector<int32_t> back(s.xyz.size());
        const int32_t* img[3] = {L.img[0].data(), L.img[1].data(), L.img[2].data()};
        L.unfolded(L.pos.data(), img, back.data());
        expect(back == s.xyz, std::string(name) + ": layout decodes to its input");
        expect(L.color.size() == s.n(), std::string(name) + ": one color per monomer");
        for (const auto& b : s.bonds)
            if (L.color[b.first] == L.color[b.second]) {
                expect(false, std::string(name) + ": bonded monomers share a color");
                break;
            }
        if (uniform) colors = L.color;
    }
    return colors;
}

// The oracle on the system: one thread, all threads, shuffled order; equal.
void check_oracle(const Sys& s, const std::vector<uint8_t>& color, const char* name) {
    const Adjacency a = adjacency(s);
    uint8_t forbidden[512];
    oracle_standard_bondset(forbidden);
    uint32_t nColors = 0;
    for (uint8_t c : color) nColors = std::max<uint32_t>(nColors, c + 1u);
    std::vector<std::vector<int32_t>> finals;
    std::vector<std::vector<uint32_t>> logs;
    const int threads[3] = {1, 0, 1};
    const uint64_t shuffle[3] = {0, 0, 12345};
    for (int k = 0; k < 3; ++k) {
        std::vector<int32_t> xyz = s.xyz;
        oracle_system sys{};
        std::memcpy(sys.box, s.box, sizeof sys.box);
        sys.n = s.n();
        sys.xyz = xyz.data();
        sys.nbr_start = a.start.data();
        sys.nbr = a.ids.empty() ? nullptr : a.ids.data();
        sys.color = color.data();
        sys.n_colors = nColors;
        sys.forbidden = forbidden;
        sys.seed = 42;
        expect(oracle_check_system(&sys) == 0, std::string(name) + ": valid start state");
        std::vector<uint32_t> log(20 * nColors);
        const int64_t acc = oracle_run(&sys, 0, log.size(), log.data(), threads[k], shuffle[k]);
        expect(acc >= 0, std::string(name) + ": oracle accepts the system");
        expect(oracle_check_system(&sys) == 0, std::string(name) + ": valid end state");
        finals.push_back(xyz);
        logs.push_back(log);
    }
    expect(finals[0] == finals[1] && finals[0] == finals[2] && logs[0] == logs[1] && logs[0] == logs[2],
           std::string(name) + ": oracle result independent of threads and order");
}

Sys from_bfm(const BfmData& b) {
    Sys s;
    std::memcpy(s.box, b.box, sizeof s.box);
    s.xyz = b.xyz;
    s.bonds = b.bonds;
    return s;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: host_san GOLDEN_DIR TMP_DIR\n");
        return 2;
    }
    const std::string golden = argv[1], tmp = argv[2];
    try {
        // generated systems
        {
            Sys s{{32, 32, 32}, {}, {}};
            expect(grow_linear_chains(s.box, 20, 16, 7, 200, s.xyz), "grown chains");
            for (uint32_t c = 0; c < 20; ++c)
                for (uint32_t i = 0; i + 1 < 16; ++i) s.bonds.push_back({c * 16 + i, c * 16 + i + 1});
            check_oracle(s, check_layout(s, "grown chains", true), "grown chains");
        }
        {
            Sys s{{64, 64, 64}, {}, {}};
            expect(place_linear_chains_ordered(s.box, 512, 32, s.xyz), "ordered chains");
            for (uint32_t c = 0; c < 512; ++c)
                for (uint32_t i = 0; i + 1 < 32; ++i) s.bonds.push_back({c * 32 + i, c * 32 + i + 1});
            check_oracle(s, check_layout(s, "ordered chains", true), "ordered chains");
        }
        {
            Sys s{{64, 64, 64}, {}, {}};
            std::vector<uint32_t> pairs;
            expect(build_star_network(s.box, 40, 4, 29, 9, s.xyz, pairs), "star network");
            for (size_t k = 0; k + 1 < pairs.size(); k += 2) s.bonds.push_back({pairs[k], pairs[k + 1]});
            check_oracle(s, check_layout(s, "star network", false), "star network");
        }
        // non-cubic boxes: the Z-curve then interleaves fewer bits on the
        // short axes, and the slot sort covers codes of fewer bits
        {
            Sys s{{16, 32, 64}, {}, {}};
            expect(grow_linear_chains(s.box, 12, 24, 11, 200, s.xyz), "grown chains, 16x32x64");
            for (uint32_t c = 0; c < 12; ++c)
                for (uint32_t i = 0; i + 1 < 24; ++i) s.bonds.push_back({c * 24 + i, c * 24 + i + 1});
            check_oracle(s, check_layout(s, "grown chains, 16x32x64", true), "grown chains, 16x32x64");
        }
        {
            Sys s{{128, 32, 64}, {}, {}};
            std::vector<uint32_t> pairs;
            expect(build_star_network(s.box, 30, 4, 19, 5, s.xyz, pairs), "star network, 128x32x64");
            for (size_t k = 0; k + 1 < pairs.size(); k += 2) s.bonds.push_back({pairs[k], pairs[k + 1]});
            check_oracle(s, check_layout(s, "star network, 128x32x64", false), "star network, 128x32x64");
        }
        // dense ordered stars, moved by 20 sweeps of the oracle, then two
        // rounds of crosslinks made at the moved corners
        {
            Sys s{{64, 64, 64}, {}, {}};
            std::vector<uint32_t> pairs;
            expect(place_stars_ordered(s.box, 120, 8, 13, s.xyz, pairs), "ordered stars");
            for (size_t k = 0; k + 1 < pairs.size(); k += 2) s.bonds.push_back({pairs[k], pairs[k + 1]});
            const std::vector<uint8_t> colors = check_layout(s, "ordered stars", false);
            const Adjacency a = adjacency(s);
            uint8_t forbidden[512];
            oracle_standard_bondset(forbidden);
            uint32_t nColors = 0;
            for (uint8_t c : colors) nColors = std::max<uint32_t>(nColors, c + 1u);
            expect(nColors == 2, "ordered stars: two species");
            oracle_system sys{};
            std::memcpy(sys.box, s.box, sizeof sys.box);
            sys.n = s.n();
            sys.xyz = s.xyz.data();
            sys.nbr_start = a.start.data();
            sys.nbr = a.ids.data();
            sys.color = colors.data();
            sys.n_colors = nColors;
            sys.forbidden = forbidden;
            sys.seed = 13;
            expect(oracle_check_system(&sys) == 0, "ordered stars: valid start state");
            std::vector<uint32_t> log(20 * nColors);
            expect(oracle_run(&sys, 0, log.size(), log.data(), 1, 0) >= 0, "ordered stars: oracle run");
            HostRng rng(13);
            std::vector<uint32_t> cross;
            crosslink_arm_ends(s.box, s.xyz, 120, 4, 8, 2, rng, cross);
            expect(!cross.empty(), "ordered stars: crosslinks formed after the run");
            for (size_t k = 0; k + 1 < cross.size(); k += 2) s.bonds.push_back({cross[k], cross[k + 1]});
            check_oracle(s, check_layout(s, "ordered stars, crosslinked", false), "ordered stars, crosslinked");
        }
        // the reference's fixtures: read, lay out, write and read back
        for (const char* f : {"LinearChain_l1024_n64.bfm.gz", "Ring_N6_n6.bfm", "Hydrogel_NStar_117_NoPerXYZ64.bfm"}) {
            const BfmData b = read_bfm(golden + "/" + f);
            expect(b.nMonomers() > 0, std::string(f) + ": monomers read");
            const Sys s = from_bfm(b);
            const bool chains = std::string(f).find("Hydrogel") == std::string::npos;  // chains and a ring
            const std::vector<uint8_t> colors = check_layout(s, f, chains);
            if (b.nMonomers() <= 1000) check_oracle(s, colors, f);
            const std::string out = tmp + "/round_trip.bfm";
            {
                std::ofstream os(out);
                BfmWriter w(b.box, b.periodic, b.nMonomers(), b.bonds, b.bondset, b.attributes);
                w.header(os);
                w.frame(os, 1234, b.xyz.data());
            }
            const BfmData r = read_bfm(out);
            expect(r.xyz == b.xyz && r.bonds.size() == b.bonds.size() && r.mcs == 1234,
                   std::string(f) + ": written file reads back equal");
        }
        // R250 from the reference's seed file: its known answer
        {
            const std::vector<uint32_t> w = read_r250_seed_file(golden + "/resultNormPscBFM.seeds");
            R250 r;
            r.load(w.data());
            expect(r.next() == 4223731124u, "R250 known answer of the seed file");
            uint32_t st[250];
            r.state(st);
            R250 q;
            q.load(st);
            for (int k = 0; k < 1000; ++k) (void)r.next(), (void)q.next();
            expect(r.next() == q.next(), "R250 state round trip");
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "FAIL: exception: %s\n", e.what());
        return 1;
    }
    if (failures) return 1;
    std::printf("host_san: all checks passed\n");
    return 0;
}
