// LeMonADE BFM file reader and writer (see bfm_io.h).
#include "bfm_io.h"

#include <zlib.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <stdexcept>


namespace pscbfm {

namespace {

std::string slurp(const std::string& path) {
    gzFile f = gzopen(path.c_str(), "rb");  // reads plain files as they are
    if (!f) throw std::runtime_error("cannot open BFM file " + path);
    std::string data;
    char buf[1 << 16];
    for (;;) {
        const int got = gzread(f, buf, sizeof buf);
        if (got < 0) {
            gzclose(f);
            throw std::runtime_error("read error in BFM file " + path);
        }
        if (got == 0) break;
        data.append(buf, size_t(got));
    }
    gzclose(f);
    return data;
}

[[noreturn]] void malformed(const std::string& path, size_t line, const std::string& what) {
    std::ostringstream msg;
    msg << path << ":" << line + 1 << ": " << what;
    throw std::runtime_error(msg.str());
}

bool starts_with(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }

long long to_int(const std::string& s, const std::string& path, size_t line) {
    char* end = nullptr;
    const long long v = std::strtoll(s.c_str(), &end, 10);
    if (end == s.c_str()) malformed(path, line, "expected an integer in '" + s + "'");
    return v;
}

std::string value_after_eq(const std::string& s) {
    const size_t p = s.find('=');
    return p == std::string::npos ? std::string() : s.substr(p + 1);
}

bool is_data_line(const std::string& s) {
    if (s.empty() || s[0] == '!') return false;
    return s.find_first_not_of(" \t\r") != std::string::npos;
}

}  // namespace

BfmData read_bfm(const std::string& path) {
    const std::string data = slurp(path);
    std::vector<std::string> lines;
    {
        size_t start = 0;
        while (start <= data.size()) {
            size_t end = data.find('\n', start);
            if (end == std::string::npos) end = data.size();
            lines.emplace_back(data, start, end - start);
            start = end + 1;
        }
    }
    BfmData out;
    long long nHeader = -1;
    std::vector<std::pair<uint32_t, uint32_t>> extra;
    std::vector<std::array<long long, 3>> attrRanges;
    std::vector<int32_t> frameXyz;
    std::vector<std::pair<uint32_t, uint32_t>> frameBonds;
    size_t i = 0;
    while (i < lines.size()) {
        std::string line = lines[i];
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#' || line.find_first_not_of(" \t") == std::string::npos) {
            ++i;
            continue;
        }
        if (starts_with(line, "!number_of_monomers=")) {
            nHeader = to_int(value_after_eq(line), path, i);
        } else if (starts_with(line, "!box_") && line.size() > 5) {
            const int axis = line[5] - 'x';
            if (axis < 0 || axis > 2) malformed(path, i, "unknown box axis");
            out.box[axis] = uint32_t(to_int(value_after_eq(line), path, i));
        } else if (starts_with(line, "!periodic_") && line.size() > 10) {
            const int axis = line[10] - 'x';
            if (axis < 0 || axis > 2) malformed(path, i, "unknown periodic axis");
            out.periodic[axis] = to_int(value_after_eq(line), path, i) != 0;
        } else if (line == "!bonds") {
            for (++i; i < lines.size() && is_data_line(lines[i]); ++i) {
                std::istringstream in(lines[i]);
                long long a = 0, b = 0;
                if (!(in >> a >> b) || a < 1 || b < 1) malformed(path, i, "bad bond line");
                extra.emplace_back(uint32_t(a - 1), uint32_t(b - 1));
            }
            continue;
        } else if (line == "!set_of_bondvectors") {
            for (++i; i < lines.size() && is_data_line(lines[i]); ++i) {
                const std::string& s = lines[i];
                const size_t colon = s.rfind(':');
                if (colon == std::string::npos) malformed(path, i, "bad bond vector line");
                std::istringstream in(s.substr(0, colon));
                std::array<int32_t, 3> v{};
                if (!(in >> v[0] >> v[1] >> v[2])) malformed(path, i, "bad bond vector");
                out.bondset[int(to_int(s.substr(colon + 1), path, i))] = v;
            }
            continue;
        } else if (line == "!attributes") {
            for (++i; i < lines.size() && is_data_line(lines[i]); ++i) {
                const std::string& s = lines[i];
                const size_t dash = s.find('-'), colon = s.find(':');
                if (dash == std::string::npos || colon == std::string::npos || colon < dash)
                    malformed(path, i, "bad attribute range");
                attrRanges.push_back({to_int(s.substr(0, dash), path, i) - 1,
                                      to_int(s.substr(dash + 1, colon - dash - 1), path, i) - 1,
                                      to_int(s.substr(colon + 1), path, i)});
            }
            continue;
        } else if (starts_with(line, "!mcs=")) {
            out.mcs = uint64_t(to_int(value_after_eq(line), path, i));
            ++out.frames;
            frameXyz.clear();
            frameBonds.clear();
            for (++i; i < lines.size() && !lines[i].empty() && lines[i][0] != '!'; ++i) {
                const std::string& s = lines[i];
                const size_t p1 = s.find(' ');
                const size_t p2 = p1 == std::string::npos ? p1 : s.find(' ', p1 + 1);
                if (p2 == std::string::npos) malformed(path, i, "strand line needs x y z");
                const size_t p3 = s.find(' ', p2 + 1);
                int32_t x = int32_t(to_int(s.substr(0, p1), path, i));
                int32_t y = int32_t(to_int(s.substr(p1 + 1, p2 - p1 - 1), path, i));
                int32_t z = int32_t(to_int(s.substr(p2 + 1, p3 == std::string::npos ? std::string::npos : p3 - p2 - 1),
                                           path, i));
                frameXyz.insert(frameXyz.end(), {x, y, z});
                if (p3 == std::string::npos) continue;
                for (size_t c = p3 + 1; c < s.size(); ++c) {
                    const int id = int(static_cast<unsigned char>(s[c]));
                    auto it = out.bondset.find(id);
                    if (it == out.bondset.end()) {
                        if (s[c] == '\r' && c + 1 == s.size()) break;  // CRLF line end
                        malformed(path, i, "bond id " + std::to_string(id) + " not in !set_of_bondvectors");
                    }
                    x += it->second[0];
                    y += it->second[1];
                    z += it->second[2];
                    const uint32_t k = uint32_t(frameXyz.size() / 3);
                    frameXyz.insert(frameXyz.end(), {x, y, z});
                    frameBonds.emplace_back(k - 1, k);
                }
            }
            continue;
        }
        ++i;
    }
    if (out.frames == 0) throw std::runtime_error(path + ": no !mcs block");
    out.xyz = frameXyz;
    const uint32_t n = out.nMonomers();
    if (nHeader >= 0 && uint64_t(nHeader) != n) {
        std::ostringstream msg;
        msg << path << ": " << n << " monomers in the last !mcs block, header says " << nHeader;
        throw std::runtime_error(msg.str());
    }
    out.bonds = frameBonds;
    for (const auto& b : extra) {
        if (b.first >= n || b.second >= n) throw std::runtime_error(path + ": !bonds entry out of range");
        out.bonds.push_back(b);
    }
    out.attributes.assign(n, 0);
    for (const auto& r : attrRanges) {
        if (r[0] < 0 || r[1] < r[0] || r[1] >= (long long)n) throw std::runtime_error(path + ": bad attribute range");
        std::fill(out.attributes.begin() + r[0], out.attributes.begin() + r[1] + 1, int32_t(r[2]));
    }
    return out;
}

std::map<int, std::array<int32_t, 3>> standard_bondset_ids() {
    // LeMonADE's numbering of the scBFM set in BFM files: 18 seed vectors
    // (a,b,c), each followed by (b,c,a), (c,a,b), (-a,b,c), (b,c,-a),
    // (c,-a,b); ids count up from 17.
    static const int32_t seeds[18][3] = {{2, 0, 0},  {2, 1, 0},  {2, -1, 0},  {1, 2, 0}, {1, -2, 0}, {2, 1, 1},
                                         {2, -1, 1}, {2, 1, -1}, {2, -1, -1}, {3, 0, 0}, {2, 2, 1},  {2, -2, 1},
                                         {2, 2, -1}, {2, -2, -1}, {3, 1, 0},  {3, -1, 0}, {1, 3, 0}, {1, -3, 0}};
    std::map<int, std::array<int32_t, 3>> m;
    int id = 17;
    for (const auto& v : seeds) {
        const int32_t a = v[0], b = v[1], c = v[2];
        for (const std::array<int32_t, 3>& w : {std::array<int32_t, 3>{a, b, c}, std::array<int32_t, 3>{b, c, a},
                                                std::array<int32_t, 3>{c, a, b}, std::array<int32_t, 3>{-a, b, c},
                                                std::array<int32_t, 3>{b, c, -a}, std::array<int32_t, 3>{c, -a, b}})
            m[id++] = w;
    }
    return m;
}

BfmWriter::BfmWriter(const uint32_t box[3], const bool periodic[3], uint32_t nMonomers,
                     const std::vector<std::pair<uint32_t, uint32_t>>& bonds,
                     std::map<int, std::array<int32_t, 3>> bondset, const std::vector<int32_t>& attributes)
    : n_(nMonomers), bondset_(std::move(bondset)), attributes_(attributes) {
    for (int a = 0; a < 3; ++a) {
        box_[a] = box[a];
        periodic_[a] = periodic[a];
    }
    for (const auto& kv : bondset_) idOf_[kv.second] = kv.first;
    linkedToPrev_.assign(n_, 0);
    for (const auto& b : bonds) {
        const uint32_t lo = std::min(b.first, b.second), hi = std::max(b.first, b.second);
        if (hi == lo + 1 && !linkedToPrev_[hi])
            linkedToPrev_[hi] = 1;
        else
            extra_.emplace_back(lo, hi);
    }
}

void BfmWriter::header(std::ostream& out) const {
    out << "#!version=2.0\n#Bond Fluctuation Model - simulation data file\n#monomer numbering starts at 1\n\n";
    out << "!number_of_monomers=" << n_ << "\n\n!bonds\n";
    for (const auto& b : extra_) out << b.first + 1 << " " << b.second + 1 << "\n";
    out << "\n";
    for (int a = 0; a < 3; ++a) out << "!box_" << char('x' + a) << "=" << box_[a] << "\n\n";
    for (int a = 0; a < 3; ++a) out << "!periodic_" << char('x' + a) << "=" << (periodic_[a] ? 1 : 0) << "\n\n";
    out << "!set_of_bondvectors\n";
    for (const auto& kv : bondset_) out << kv.second[0] << " " << kv.second[1] << " " << kv.second[2] << ":" << kv.first << "\n";
    out << "\n";
    if (!attributes_.empty()) {
        out << "!attributes\n";
        uint32_t start = 0;
        for (uint32_t i = 1; i <= n_; ++i)
            if (i == n_ || attributes_[i] != attributes_[start]) {
                out << start + 1 << "-" << i << ":" << attributes_[start] << "\n";
                start = i;
            }
        out << "\n";
    }
}

void BfmWriter::frame(std::ostream& out, uint64_t mcs, const int32_t* xyz) const {
    out << "\n!mcs=" << mcs << "\n";
    std::string line;
    for (uint32_t i = 0; i < n_;) {
        line = std::to_string(xyz[3 * size_t(i)]) + " " + std::to_string(xyz[3 * size_t(i) + 1]) + " " +
               std::to_string(xyz[3 * size_t(i) + 2]);
        uint32_t j = i + 1;
        std::string chars;
        for (; j < n_ && linkedToPrev_[j]; ++j) {
            const std::array<int32_t, 3> v = {xyz[3 * size_t(j)] - xyz[3 * size_t(j - 1)],
                                              xyz[3 * size_t(j) + 1] - xyz[3 * size_t(j - 1) + 1],
                                              xyz[3 * size_t(j) + 2] - xyz[3 * size_t(j - 1) + 2]};
            auto it = idOf_.find(v);
            if (it == idOf_.end()) {
                std::ostringstream msg;
                msg << "bond " << j << "-" << j + 1 << " (" << v[0] << "," << v[1] << "," << v[2]
                    << ") is not in the bond set";
                throw std::runtime_error(msg.str());
            }
            chars.push_back(char(it->second));
        }
        if (!chars.empty()) line += " " + chars;
        out << line << "\n";
        i = j;
    }
}

}  // namespace pscbfm
