// R250, the generator the reference draws its per-substep species and kernel
// seed with (CRITICAL.cu:509-510 through LeMonADE's RandomNumberGenerators,
// seeded from a seed file in mainCUDASimulatorGPUScBFM_AB_Type.cpp:124-139).
//
// Kirkpatrick-Stoll shift register (1981) in the 250-word ring form of
// Maier's r250.c: with the loaded words as x[0..249], draw n (counting from
// 250) is x[n] = x[n-250] ^ x[n-147].  The ring slot being replaced holds
// x[n-250]; the slot 103 ahead of it (modulo 250) was written 147 draws ago.
//
// A LeMonADE seed file is little-endian uint32: word 0 seeds std::srand,
// words 7..256 are the R250 state.  With tests/resultNormPscBFM.seeds this
// layout gives the reference's known answers 4223731124 and 0.803876
// (mainCUDASimulatorGPUScBFM_AB_Type.cpp:138-139) for the first two draws.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace pscbfm {

class R250 {
public:
    static constexpr uint32_t kWords = 250;
    static constexpr uint32_t kSeedFileFirst = 7;   // first state word in a seed file
    static constexpr uint32_t kSeedFileWords = kSeedFileFirst + kWords;

    void load(const uint32_t* words) {
        std::memcpy(buf_, words, sizeof buf_);
        index_ = 0;
    }
    uint32_t next() {
        const uint32_t j = index_ >= 147u ? index_ - 147u : index_ + 103u;
        const uint32_t v = buf_[index_] ^ buf_[j];
        buf_[index_] = v;
        index_ = index_ == kWords - 1 ? 0u : index_ + 1u;
        return v;
    }
    // The 250 ring words in draw order (the next draw replaces out[0]), so
    // that load(out) continues the same stream.
    void state(uint32_t* out) const {
        for (uint32_t k = 0; k < kWords; ++k) out[k] = buf_[(index_ + k) % kWords];
    }

private:
    uint32_t buf_[kWords] = {};
    uint32_t index_ = 0;
};

// Reads the state words of a LeMonADE seed file; throws if the file is
// missing or shorter than kSeedFileWords words.
std::vector<uint32_t> read_r250_seed_file(const std::string& path);

}  // namespace pscbfm
