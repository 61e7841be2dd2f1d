// Host graph coloring that splits the monomers into species whose members
// share no bond, so one species can be updated in parallel.
#pragma once

#include <cstdint>
#include <vector>

namespace pscbfm {

// Adjacency in compressed rows: the neighbours of node i are
// ids[start[i] .. start[i+1]), in the order the caller registered them.
struct Adjacency {
    std::vector<uint32_t> start;
    std::vector<uint32_t> ids;
    uint32_t size() const { return start.empty() ? 0u : uint32_t(start.size() - 1); }
    uint32_t degree(uint32_t i) const { return start[i + 1] - start[i]; }
    uint32_t neighbor(uint32_t i, uint32_t k) const { return ids[start[i] + k]; }
};

// Same coloring as graphColoring.tpp:77-293: a depth-first walk along the
// bonds that gives each node the smallest color unused by its colored
// neighbours, followed (when `uniform` is set) by moving nodes of over-full
// colors into the rarest colors that their neighbours leave free.  Throws
// std::runtime_error when more than 254 colors would be needed.
std::vector<uint8_t> color_graph(const Adjacency& adj, bool uniform);

}  // namespace pscbfm
