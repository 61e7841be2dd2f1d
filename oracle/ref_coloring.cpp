// TEST INFRASTRUCTURE ONLY: C entry point over the reference's own
// header-only graph coloring (src/pscbfm/graphColoring.tpp:77-293), compiled
// from the reference tree where it lies (see oracle/Makefile, target ref).
// Used by the tests to compare the library's coloring with the reference's.
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include "graphColoring.tpp"

namespace {
struct Csr {
    const uint32_t* start;
    const uint32_t* ids;
};
}  // namespace

extern "C" int ref_graph_coloring(uint32_t n, const uint32_t* start, const uint32_t* ids, int uniform,
                                  uint8_t* colors) {
    try {
        const Csr adj{start, ids};
        const std::vector<uint8_t> c = graphColoring<Csr, uint32_t, uint8_t>(
            adj, n, uniform != 0,
            [](Csr const& a, uint32_t const& i) -> size_t { return a.start[i + 1] - a.start[i]; },
            [](Csr const& a, uint32_t const& i, size_t const& k) -> uint32_t { return a.ids[a.start[i] + k]; });
        if (n) std::memcpy(colors, c.data(), n);
        return 0;
    } catch (const std::exception&) {
        return 2;
    }
}
