#include "r250.h"

#include <cstdio>
#include <stdexcept>

namespace pscbfm {

// RandomNumberGenerators::seedAll(file) as mainCUDASimulatorGPUScBFM_AB_Type.cpp:129
// calls it: state = words 7..256 of the file.
std::vector<uint32_t> read_r250_seed_file(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("[seed file] cannot open " + path);
    std::vector<uint8_t> bytes(4u * R250::kSeedFileWords);
    const size_t got = std::fread(bytes.data(), 1, bytes.size(), f);
    std::fclose(f);
    if (got < bytes.size())
        throw std::runtime_error("[seed file] " + path + " holds fewer than " +
                                 std::to_string(R250::kSeedFileWords) + " 32-bit words");
    std::vector<uint32_t> out(R250::kWords);
    for (uint32_t k = 0; k < R250::kWords; ++k) {
        const uint8_t* b = bytes.data() + 4u * (R250::kSeedFileFirst + k);
        out[k] = uint32_t(b[0]) | uint32_t(b[1]) << 8 | uint32_t(b[2]) << 16 | uint32_t(b[3]) << 24;
    }
    return out;
}

}  // namespace pscbfm
