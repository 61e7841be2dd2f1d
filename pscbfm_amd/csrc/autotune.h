// Launch-shape autotuner.  The sweep's results do not depend on the shape
// (threads per workgroup, monomers per pass-1 thread), so every candidate
// runs the same trajectory on a scratch copy of the system(s) and only the
// time differs; the fastest shape wins.  Results are cached per workload
// signature for the life of the process.
#pragma once

#include <cstdint>
#include <vector>

#include "kernels.h"
#include "layout.h"

namespace pscbfm {

struct TuneTrial {
    LaunchShape shape;
    double msPerMcs = 0;
};

struct TuneResult {
    LaunchShape best;
    double bestMsPerMcs = 0;
    std::vector<TuneTrial> trials;
    bool cached = false;
};

// The choices a single system's shape makes by itself once its pass-1
// block is set (the updater's defaults, also applied to every candidate
// the autotuner times):
// pass-2 threads per workgroup, 0 = the pass-1 block;
uint32_t automatic_perform_block(const SystemLayout& L, uint32_t block);
// the pass (1 or 2) that zeroes the other parity's proposal buffer.
uint32_t automatic_clear_pass(const SystemLayout& L, uint32_t latticeKind);

// Candidate shapes: block in {64, 128, 192, 256} x slots in {1, 2, 4}.
const std::vector<LaunchShape>& tune_candidates();

// One system: scratch device copy per candidate, captured substep graph,
// `mcs` timed sweeps after one warm-up sweep.
TuneResult autotune_system(int gpu, const SystemLayout& L, uint32_t latticeKind, uint64_t seed, uint32_t mcs);

// A batch: a scratch batch of the sample per candidate, `mcs` timed sweeps
// after one warm-up sweep.
TuneResult autotune_batch(int gpu, const std::vector<SystemLayout>& sample, const std::vector<uint64_t>& seeds,
                          uint32_t latticeKind, uint32_t mcs);

}  // namespace pscbfm
