// Seeded builders of synthetic scBFM systems (see generators.h).
#include <algorithm>
#include <array>
#include <cstdlib>
#include <stdexcept>
#include <utility>

#include "common.h"
#include "generators.h"

namespace pscbfm {

const std::vector<std::array<int32_t, 3>>& standard_bond_vectors() {
    static const std::vector<std::array<int32_t, 3>> vecs = [] {
        std::vector<std::array<int32_t, 3>> v;
        for (int x = -3; x <= 3; ++x)
            for (int y = -3; y <= 3; ++y)
                for (int z = -3; z <= 3; ++z) {
                    int a[3] = {std::abs(x), std::abs(y), std::abs(z)};
                    for (int i = 0; i < 3; ++i)
                        for (int j = i + 1; j < 3; ++j)
                            if (a[j] > a[i]) std::swap(a[i], a[j]);
                    static const int base[6][3] = {{2, 0, 0}, {2, 1, 0}, {2, 1, 1},
                                                   {2, 2, 1}, {3, 0, 0}, {3, 1, 0}};
                    for (const auto& b : base)
                        if (a[0] == b[0] && a[1] == b[1] && a[2] == b[2]) v.push_back({x, y, z});
                }
        return v;
    }();
    return vecs;
}

uint64_t HostRng::next() { return splitmix64_at(key, counter++); }

uint32_t HostRng::below(uint32_t n) {
    // Lemire's multiply-shift on the high word; bias < 2^-32 for the n used here.
    return uint32_t((uint64_t(uint32_t(next() >> 32)) * n) >> 32);
}

namespace {

struct CubeLattice {
    uint32_t box[3];
    std::vector<uint8_t> cell;
    explicit CubeLattice(const uint32_t b[3]) : box{b[0], b[1], b[2]}, cell(size_t(b[0]) * b[1] * b[2], 0) {}
    size_t index(int32_t x, int32_t y, int32_t z) const {
        return (uint32_t(x) & (box[0] - 1)) + size_t(box[0]) * ((uint32_t(y) & (box[1] - 1)) +
                                                                size_t(box[1]) * (uint32_t(z) & (box[2] - 1)));
    }
    bool free(const int32_t p[3]) const {
        for (int c = 0; c < 8; ++c)
            if (cell[index(p[0] + (c & 1), p[1] + ((c >> 1) & 1), p[2] + ((c >> 2) & 1))]) return false;
        return true;
    }
    void set(const int32_t p[3], uint8_t v) {
        for (int c = 0; c < 8; ++c) cell[index(p[0] + (c & 1), p[1] + ((c >> 1) & 1), p[2] + ((c >> 2) & 1))] = v;
    }
};

bool random_free_start(CubeLattice& lat, HostRng& rng, int32_t p[3]) {
    for (int attempt = 0; attempt < 100000; ++attempt) {
        p[0] = int32_t(rng.below(lat.box[0]));
        p[1] = int32_t(rng.below(lat.box[1]));
        p[2] = int32_t(rng.below(lat.box[2]));
        if (lat.free(p)) return true;
    }
    return false;
}

// Next monomer after `prev`: the 108 vectors in random order, first free one.
bool grow_step(CubeLattice& lat, HostRng& rng, const int32_t prev[3], int32_t out[3]) {
    const auto& vecs = standard_bond_vectors();
    uint32_t order[108];
    for (uint32_t i = 0; i < 108; ++i) order[i] = i;
    for (uint32_t i = 108; i > 1; --i) std::swap(order[i - 1], order[rng.below(i)]);
    for (uint32_t k = 0; k < 108; ++k) {
        const auto& b = vecs[order[k]];
        const int32_t c[3] = {prev[0] + b[0], prev[1] + b[1], prev[2] + b[2]};
        if (lat.free(c)) {
            out[0] = c[0];
            out[1] = c[1];
            out[2] = c[2];
            return true;
        }
    }
    return false;
}

}  // namespace

bool grow_linear_chains(const uint32_t box[3], uint32_t nChains, uint32_t len, uint64_t seed,
                        uint32_t maxRestarts, std::vector<int32_t>& xyz) {
    CubeLattice lat(box);
    HostRng rng(seed);
    xyz.assign(3 * size_t(nChains) * len, 0);
    uint32_t restarts = 0;
    for (uint32_t c = 0; c < nChains; ++c) {
        for (;;) {
            int32_t* chain = xyz.data() + 3 * size_t(c) * len;
            uint32_t placed = 0;
            if (len > 0 && random_free_start(lat, rng, chain)) {
                lat.set(chain, 1);
                placed = 1;
                while (placed < len && grow_step(lat, rng, chain + 3 * (placed - 1), chain + 3 * placed)) {
                    lat.set(chain + 3 * placed, 1);
                    ++placed;
                }
            }
            if (placed == len) break;
            for (uint32_t k = 0; k < placed; ++k) lat.set(chain + 3 * k, 0);
            if (++restarts > maxRestarts) return false;
        }
    }
    return true;
}

bool place_linear_chains_ordered(const uint32_t box[3], uint32_t nChains, uint32_t len,
                                 std::vector<int32_t>& xyz) {
    const uint64_t nx = box[0] / 2, ny = box[1] / 2, nz = box[2] / 3;
    const uint64_t sites = nx * ny * nz;
    if (len == 0) {
        xyz.clear();
        return nChains == 0;
    }
    const uint64_t segments = sites / len;
    if (nChains > segments) return false;
    xyz.assign(3 * size_t(nChains) * len, 0);
    // Site s of the serpentine path -> coordinates.
    auto site = [&](uint64_t s, int32_t p[3]) {
        const uint64_t layer = s / (nx * ny);
        uint64_t r = (s % (nx * ny)) / nx;
        uint64_t i = s % nx;
        if (layer & 1u) r = ny - 1 - r;          // layers traverse rows backwards
        const uint64_t rowInPath = (s % (nx * ny)) / nx + layer * ny;
        if (rowInPath & 1u) i = nx - 1 - i;      // rows alternate direction along the path
        p[0] = int32_t(2 * i);
        p[1] = int32_t(2 * r);
        p[2] = int32_t(3 * layer);
    };
    for (uint32_t c = 0; c < nChains; ++c) {
        const uint64_t seg = uint64_t(c) * segments / nChains;
        for (uint32_t k = 0; k < len; ++k) site(seg * len + k, xyz.data() + 3 * (size_t(c) * len + k));
    }
    // Unfold along each chain so that every bond is the short vector.
    for (uint32_t c = 0; c < nChains; ++c)
        for (uint32_t k = 1; k < len; ++k) {
            int32_t* prev = xyz.data() + 3 * (size_t(c) * len + k - 1);
            int32_t* cur = xyz.data() + 3 * (size_t(c) * len + k);
            for (int a = 0; a < 3; ++a) {
                int32_t d = cur[a] - prev[a];
                const int32_t L = int32_t(box[a]);
                d = ((d % L) + L + L / 2) % L - L / 2;
                cur[a] = prev[a] + d;
            }
        }
    return true;
}

void crosslink_arm_ends(const uint32_t box[3], const std::vector<int32_t>& xyz, uint32_t nStars, uint32_t arms,
                        uint32_t armLength, uint32_t rounds, HostRng& rng, std::vector<uint32_t>& bonds) {
    // In each round every arm end takes the first (in random vector order)
    // monomer of another star's arm interior that sits at a valid bond
    // vector, has no crosslink of this round yet and is not already its
    // partner.
    const uint32_t perStar = 1 + arms * armLength;
    const uint32_t n = nStars * perStar;
    if (xyz.size() < 3 * size_t(n)) throw std::invalid_argument("crosslink_arm_ends: fewer corners than monomers");
    if (armLength < 2 || rounds == 0) return;
    std::vector<int32_t> owner(size_t(box[0]) * box[1] * box[2], -1);
    auto cellOf = [&](const int32_t* p) {
        return (uint32_t(p[0]) & (box[0] - 1)) +
               size_t(box[0]) * ((uint32_t(p[1]) & (box[1] - 1)) + size_t(box[1]) * (uint32_t(p[2]) & (box[2] - 1)));
    };
    for (uint32_t i = 0; i < n; ++i) owner[cellOf(xyz.data() + 3 * size_t(i))] = int32_t(i);
    std::vector<uint8_t> linked(n);
    std::vector<uint32_t> partners(size_t(nStars) * arms * rounds, ~0u);  // per arm end, one per round
    const auto& vecs = standard_bond_vectors();
    for (uint32_t round = 0; round < rounds; ++round) {
        std::fill(linked.begin(), linked.end(), uint8_t(0));
        for (uint32_t s = 0; s < nStars; ++s)
            for (uint32_t a = 0; a < arms; ++a) {
                const uint32_t end = s * perStar + 1 + a * armLength + (armLength - 1);
                if (linked[end]) continue;
                uint32_t* mine = partners.data() + (size_t(s) * arms + a) * rounds;
                const int32_t* pe = xyz.data() + 3 * size_t(end);
                uint32_t order[108];
                for (uint32_t i = 0; i < 108; ++i) order[i] = i;
                for (uint32_t i = 108; i > 1; --i) std::swap(order[i - 1], order[rng.below(i)]);
                for (uint32_t v = 0; v < 108; ++v) {
                    const auto& b = vecs[order[v]];
                    const int32_t c[3] = {pe[0] + b[0], pe[1] + b[1], pe[2] + b[2]};
                    const int32_t o = owner[cellOf(c)];
                    if (o < 0) continue;
                    const uint32_t os = uint32_t(o) / perStar, ok = uint32_t(o) % perStar;
                    const bool interior = ok != 0 && ((ok - 1) % armLength) != armLength - 1;
                    if (os == s || !interior || linked[o]) continue;
                    if (std::find(mine, mine + round, uint32_t(o)) != mine + round) continue;
                    linked[end] = 1;
                    linked[o] = 1;
                    mine[round] = uint32_t(o);
                    bonds.push_back(end);
                    bonds.push_back(uint32_t(o));
                    break;
                }
            }
    }
    // A crosslink found through the periodic box can join stars whose
    // unfolded coordinates differ by whole box lengths; such a bond is valid
    // under the minimum-image convention (the kernels only use the bond
    // vector modulo 8, a divisor of every box edge).
}

bool build_star_network(const uint32_t box[3], uint32_t nStars, uint32_t arms, uint32_t armLength,
                        uint64_t seed, std::vector<int32_t>& xyz, std::vector<uint32_t>& bonds) {
    const uint32_t perStar = 1 + arms * armLength;
    CubeLattice lat(box);
    HostRng rng(seed);
    xyz.assign(3 * size_t(nStars) * perStar, 0);
    bonds.clear();
    uint32_t restarts = 0;
    for (uint32_t s = 0; s < nStars; ++s) {
        int32_t* star = xyz.data() + 3 * size_t(s) * perStar;
        for (;;) {
            uint32_t placed = 0;
            bool ok = random_free_start(lat, rng, star);
            if (ok) {
                lat.set(star, 1);
                placed = 1;
            }
            for (uint32_t a = 0; ok && a < arms; ++a)
                for (uint32_t k = 0; ok && k < armLength; ++k) {
                    const uint32_t idx = 1 + a * armLength + k;
                    const int32_t* prev = k == 0 ? star : star + 3 * (idx - 1);
                    ok = grow_step(lat, rng, prev, star + 3 * idx);
                    if (ok) {
                        lat.set(star + 3 * idx, 1);
                        ++placed;
                    }
                }
            if (ok) break;
            // Undo what was placed of this star (indices grow in order).
            for (uint32_t k = 0; k < placed; ++k) lat.set(star + 3 * k, 0);
            if (++restarts > 1000u + nStars) return false;
        }
        for (uint32_t a = 0; a < arms; ++a)
            for (uint32_t k = 0; k < armLength; ++k) {
                const uint32_t idx = 1 + a * armLength + k;
                const uint32_t prev = k == 0 ? 0 : idx - 1;
                bonds.push_back(s * perStar + prev);
                bonds.push_back(s * perStar + idx);
            }
    }
    crosslink_arm_ends(box, xyz, nStars, arms, armLength, 1, rng, bonds);
    return true;
}

bool place_stars_ordered(const uint32_t box[3], uint32_t nStars, uint32_t armLength, uint64_t seed,
                         std::vector<int32_t>& xyz, std::vector<uint32_t>& bonds) {
    const uint32_t a = armLength;
    const uint32_t nx = box[0] / 2, ny = box[1] / 2, nz = box[2] / 2;  // sites per axis (powers of two)
    xyz.clear();
    bonds.clear();
    if (nStars == 0) return true;
    if (a == 0) return false;
    // Centers P sites apart along a row, P the largest power-of-two share of
    // the row that holds a star (2a + 1 sites).
    uint32_t perRow = 1;
    while (perRow * 2 <= nx && nx / (perRow * 2) >= 2 * a + 1) perRow *= 2;
    const uint32_t P = nx / perRow;
    if (P < 2 * a + 1) return false;
    // Row (y, z) has its centers at offset(y) + j P.  An odd row holds the
    // third arms [c, c + a - 1] of row (y - 1, z) and the fourth arms
    // [c' - a + 1, c'] of row (y, z - 1), whose centers are shifted by delta:
    // the two sets are disjoint for 2a - 1 <= delta <= P - 1, and
    // offset(y) = delta * y mod P wraps consistently when delta * ny is a
    // multiple of P.
    const uint32_t step = P / std::min(P, ny);  // delta must be a multiple of this
    const uint32_t lo = 2 * a - 1, hi = P - 1, mid = (lo + hi) / 2;
    uint32_t delta = 0;
    auto dist = [&](uint32_t d) { return d > mid ? d - mid : mid - d; };
    for (uint32_t d = step; d <= hi; d += step)
        if (d >= lo && (delta == 0 || dist(d) < dist(delta))) delta = d;
    if (delta == 0) return false;

    const uint64_t places = uint64_t(perRow) * ny * nz / 2;
    if (nStars > places) return false;
    // nStars places at random (partial Fisher-Yates), filled in row order.
    HostRng rng(seed);
    std::vector<uint32_t> pick(places);
    for (uint64_t i = 0; i < places; ++i) pick[i] = uint32_t(i);
    for (uint32_t i = 0; i < nStars; ++i) {
        const uint32_t r = i + uint32_t((uint64_t(uint32_t(rng.next() >> 32)) * (places - i)) >> 32);
        std::swap(pick[i], pick[r]);
    }
    pick.resize(nStars);
    std::sort(pick.begin(), pick.end());

    const uint32_t perStar = 1 + 4 * a;
    xyz.assign(3 * size_t(nStars) * perStar, 0);
    std::vector<uint8_t> taken(size_t(nx) * ny * nz, 0);
    auto site = [&](int64_t x, int64_t y, int64_t z) {
        const uint64_t X = uint64_t(x) & (nx - 1), Y = uint64_t(y) & (ny - 1), Z = uint64_t(z) & (nz - 1);
        return size_t(X + nx * (Y + uint64_t(ny) * Z));
    };
    for (uint32_t s = 0; s < nStars; ++s) {
        // place -> (row, j): rows with y + z even in z-major order, perRow places each
        const uint32_t row = pick[s] / perRow, j = pick[s] % perRow;
        const uint32_t z = row / (ny / 2);
        const uint32_t y = 2 * (row % (ny / 2)) + (z & 1u);
        const int64_t c = (uint64_t(delta) * y % P) + uint64_t(j) * P;
        int32_t* star = xyz.data() + 3 * size_t(s) * perStar;
        auto put = [&](uint32_t idx, int64_t sx, int64_t sy, int64_t sz) {
            int32_t* p = star + 3 * size_t(idx);
            p[0] = int32_t(2 * sx);
            p[1] = int32_t(2 * sy);
            p[2] = int32_t(2 * sz);
            uint8_t& t = taken[site(sx, sy, sz)];
            if (t) throw std::logic_error("ordered stars: two monomers on one site");
            t = 1;
        };
        put(0, c, y, z);
        for (uint32_t k = 0; k < a; ++k) {
            put(1 + k, c + 1 + k, y, z);                            // arm 1: +x
            put(1 + a + k, c - 1 - int64_t(k), y, z);               // arm 2: -x
            put(1 + 2 * a + k, c + k, int64_t(y) + 1, z);           // arm 3: up in y, then +x
            put(1 + 3 * a + k, c - int64_t(k), y, int64_t(z) + 1);  // arm 4: up in z, then -x
        }
        for (uint32_t arm = 0; arm < 4; ++arm)
            for (uint32_t k = 0; k < a; ++k) {
                const uint32_t idx = 1 + arm * a + k;
                bonds.push_back(s * perStar + (k == 0 ? 0 : idx - 1));
                bonds.push_back(s * perStar + idx);
            }
    }
    return true;
}

}  // namespace pscbfm
