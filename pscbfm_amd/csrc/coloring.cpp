// Host graph coloring (see coloring.h).  The order of every decision matches
// graphColoring.tpp:77-293, so the species partition and with it the
// species-local index that keys each monomer's move draw are the ones the
// reference would compute for the same bond list.
#include "coloring.h"
#include "common.h"

#include <algorithm>
#include <stdexcept>
#include <utility>

namespace pscbfm {

namespace {

constexpr uint8_t kUncolored = 0xFF;
constexpr int64_t kNone = -1;

// Smallest color not taken by an already colored neighbour of `node`.
// `taken` is all zero on entry and on return: only the entries the
// neighbours set are cleared again, not the whole table.
uint32_t smallest_free_color(const Adjacency& adj, const std::vector<uint8_t>& color,
                             uint32_t node, std::vector<uint8_t>& taken) {
    for (uint32_t k = 0; k < adj.degree(node); ++k) {
        const uint8_t c = color[adj.neighbor(node, k)];
        if (c != kUncolored) taken[c] = 1;
    }
    uint32_t c = 0;
    while (c < kMaxColors && taken[c]) ++c;
    for (uint32_t k = 0; k < adj.degree(node); ++k) {
        const uint8_t t = color[adj.neighbor(node, k)];
        if (t != kUncolored) taken[t] = 0;
    }
    return c;
}

// Uncolored neighbours of `node`: how many, and the first in bond order.
std::pair<uint32_t, int64_t> uncolored_neighbors(const Adjacency& adj,
                                                 const std::vector<uint8_t>& color,
                                                 uint32_t node) {
    uint32_t n = 0;
    int64_t first = kNone;
    for (uint32_t k = 0; k < adj.degree(node); ++k) {
        const uint32_t nb = adj.neighbor(node, k);
        if (color[nb] == kUncolored) {
            if (first == kNone) first = nb;
            ++n;
        }
    }
    return {n, first};
}

// One pass of bubble sort on (color, count) pairs that keeps equal counts in
// place, as graphColoring.tpp:30-53 does after every recoloring.
void resort_by_count(std::vector<std::pair<uint8_t, size_t>>& v) {
    size_t end = v.size();
    while (end > 1) {
        size_t lastSwap = 0;
        for (size_t j = 0; j + 1 < end; ++j) {
            if (v[j].second > v[j + 1].second) {
                std::swap(v[j], v[j + 1]);
                lastSwap = j + 1;
            }
        }
        end = lastSwap;
    }
}

}  // namespace

std::vector<uint8_t> color_graph(const Adjacency& adj, bool uniform) {
    const uint32_t n = adj.size();
    std::vector<uint8_t> color(n, kUncolored);
    std::vector<size_t> used;  // nodes per color
    std::vector<uint8_t> taken(kMaxColors + 1, 0);

    // Depth-first walk over every connected component, started at the
    // lowest uncolored id; nodes with more than one uncolored neighbour are
    // remembered so the walk can come back to them.
    for (uint32_t root = 0; root < n; ++root) {
        if (color[root] != kUncolored) continue;
        std::vector<uint32_t> pending;
        int64_t node = root;
        for (;;) {
            const auto un = uncolored_neighbors(adj, color, uint32_t(node));
            const uint32_t c = smallest_free_color(adj, color, uint32_t(node), taken);
            if (c == kMaxColors)
                throw std::runtime_error(
                    "graph coloring needs more than 254 colors; the bond graph cannot be "
                    "split into species");
            color[node] = uint8_t(c);
            if (c >= used.size()) used.resize(c + 1, 0);
            ++used[c];

            if (un.first == 0 && pending.empty()) break;
            if (un.first > 1) pending.push_back(uint32_t(node));
            node = un.second;
            if (un.first == 0) {
                // Walk back to the most recent node that still has uncolored
                // neighbours; cycles can have colored them in the meantime.
                while (!pending.empty()) {
                    const uint32_t back = pending.back();
                    const auto un2 = uncolored_neighbors(adj, color, back);
                    if (un2.first < 2) pending.pop_back();
                    if (un2.second != kNone) {
                        node = un2.second;
                        break;
                    }
                    node = kNone;
                }
                if (pending.empty() && node == kNone) break;
            }
        }
    }

    if (uniform && !used.empty()) {
        const size_t nColors = used.size();
        std::vector<std::pair<uint8_t, size_t>> rarest(nColors);
        for (size_t c = 0; c < nColors; ++c) rarest[c] = {uint8_t(c), used[c]};
        std::sort(rarest.begin(), rarest.end(),
                  [](const std::pair<uint8_t, size_t>& a, const std::pair<uint8_t, size_t>& b) {
                      return a.second < b.second;
                  });
        const size_t target = (size_t(n) + nColors - 1) / nColors;
        for (uint32_t node = 0; node < n; ++node) {
            if (used[color[node]] <= target) continue;
            // Every rarer-than-target color that no neighbour uses takes the
            // node in turn; the scan is not cut short after the first move.
            for (size_t i = 0; i < rarest.size() && rarest[i].second < target; ++i) {
                const uint8_t candidate = rarest[i].first;
                bool free = true;
                for (uint32_t k = 0; k < adj.degree(node); ++k)
                    if (color[adj.neighbor(node, k)] == candidate) free = false;
                if (!free) continue;
                const uint8_t previous = color[node];
                --used[previous];
                ++used[candidate];
                size_t iNew = 0, iOld = 0;
                for (size_t r = 0; r < rarest.size(); ++r) {
                    if (rarest[r].first == previous) iOld = r;
                    if (rarest[r].first == candidate) iNew = r;
                }
                --rarest[iOld].second;
                ++rarest[iNew].second;
                resort_by_count(rarest);
                color[node] = candidate;
            }
        }
    }
    return color;
}

}  // namespace pscbfm
