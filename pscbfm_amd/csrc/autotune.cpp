// Launch-shape autotuner (see autotune.h).
#include "autotune.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

#include "batch.h"
#include "device_system.h"

namespace pscbfm {

namespace {

// Signature of a workload for the cache: lattice kind, box exponents, and
// the rounded log2 of the monomer count, species size, degree and count of
// systems (batches).
using Key = std::tuple<int, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>;

uint32_t log2_ceil(uint64_t v) {
    uint32_t r = 0;
    while ((uint64_t(1) << r) < v) ++r;
    return r;
}

Key system_key(int mode, const SystemLayout& L, uint32_t kind, uint32_t count) {
    uint32_t deg = 0;
    for (const ColorInfo& c : L.colors) deg = std::max(deg, c.maxDegree);
    return Key{mode, kind, L.codec.log2L[0], L.codec.log2L[1], L.codec.log2L[2], log2_ceil(L.nMonomers),
               log2_ceil(L.maxColorCount), deg, count};
}

std::mutex g_cacheMutex;
std::map<Key, TuneResult>& cache() {
    static std::map<Key, TuneResult> c;
    return c;
}

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events() {
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
    }
    ~Events() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    double ms() const {
        float v = 0;
        hip_check(hipEventElapsedTime(&v, a, b), "hipEventElapsedTime");
        return v;
    }
};

// Times `mcs` sweeps of a scratch copy of L at shape s (graph replay).
double time_system(const SystemLayout& L, uint32_t kind, uint64_t seed, LaunchShape s, uint32_t mcs,
                   hipStream_t st) {
    DeviceSystem dev;
    dev.allocate(L, kind, seed, /*withLists=*/true, uint32_t(max_grid(L.maxColorCount)), st);
    s.grid = grid_for(L.maxColorCount, s.block * s.slots);
    s.performBlock = automatic_perform_block(L, s.block);
    s.clearPass = automatic_clear_pass(L, kind);
    const uint32_t nc = std::max<uint32_t>(1u, uint32_t(L.colors.size()));
    const uint32_t g = nc;  // one sweep per replay
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipError_t err = hipSuccess;
    for (uint32_t k = 0; k < g && err == hipSuccess; ++k) {
        err = launch_check(dev.desc, k, s, st);
        if (err == hipSuccess) err = launch_perform(dev.desc, k, s, st);
    }
    if (err == hipSuccess) err = launch_advance(dev.desc.state, g, st);
    const hipError_t endErr = hipStreamEndCapture(st, &graph);
    if (err != hipSuccess) hip_fail(err, "kernel launch during capture");
    hip_check(endErr, "hipStreamEndCapture");
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    hip_check(ie, "hipGraphInstantiate");
    Events ev;
    hipError_t runErr = hipGraphLaunch(exec, st);  // warm-up sweep
    if (runErr == hipSuccess) runErr = hipEventRecord(ev.a, st);
    for (uint32_t r = 0; r < mcs && runErr == hipSuccess; ++r) runErr = hipGraphLaunch(exec, st);
    if (runErr == hipSuccess) runErr = hipEventRecord(ev.b, st);
    if (runErr == hipSuccess) runErr = hipStreamSynchronize(st);
    (void)hipGraphExecDestroy(exec);
    hip_check(runErr, "autotune sweep");
    return ev.ms() / std::max<uint32_t>(mcs, 1u);
}

TuneResult pick(std::vector<TuneTrial> trials) {
    TuneResult r;
    r.trials = std::move(trials);
    auto it = std::min_element(r.trials.begin(), r.trials.end(),
                               [](const TuneTrial& a, const TuneTrial& b) { return a.msPerMcs < b.msPerMcs; });
    r.best = it->shape;
    r.bestMsPerMcs = it->msPerMcs;
    return r;
}

}  // namespace

uint32_t automatic_perform_block(const SystemLayout& L, uint32_t block) {
    // Pass 2 walks one compacted pass-1 list per workgroup.  In a dense
    // system (at least a quarter of the cells occupied) a list holds about a
    // fifth of its capacity, so 128 threads have about one entry each: on
    // the 512^3 melt at 0.5 that sweeps 1.0% faster than 256 threads (1.9%
    // when pass 2 also clears the proposals) and 64 threads 3.2% slower.
    // Dilute systems have fuller lists and keep the pass-1 block (the
    // replica mix at 0.1-0.5 runs pass 2 16% faster with 256 than with 128).
    const uint32_t log2Cells = L.codec.log2L[0] + L.codec.log2L[1] + L.codec.log2L[2];
    const double occupied = 8.0 * double(L.nMonomers) / std::ldexp(1.0, int(log2Cells));
    return occupied >= 0.25 && block > 128u ? 128u : 0u;
}

uint32_t automatic_clear_pass(const SystemLayout& L, uint32_t latticeKind) {
    // The clear (16 bytes per thread and step) goes into pass 1 where that
    // pass has issue slots to spare: the bit lattice with at most two bonds
    // per monomer, whose pass 1 gathers the fewest words.  Measured, pass 1
    // against pass 2: +1.2% sweeps on the 512^3 melt at 0.5, +1.9% at 0.25,
    // +0.4% on 256^3; the 4-arm star network loses 2.3-3.4% and the uint8
    // lattice 0.6-2% (its pass 1 moves eight times the plane bytes), so
    // those clear in pass 2.
    uint32_t maxDegree = 0;
    for (const ColorInfo& ci : L.colors) maxDegree = std::max(maxDegree, ci.maxDegree);
    return latticeKind == kLatticeBit && maxDegree <= 2u ? 1u : 2u;
}

const std::vector<LaunchShape>& tune_candidates() {
    static const std::vector<LaunchShape> c = [] {
        std::vector<LaunchShape> v;
        for (uint32_t block : {64u, 128u, 192u, 256u})
            for (uint32_t slots : {1u, 2u, 4u}) {
                LaunchShape s;
                s.block = block;
                s.slots = slots;
                v.push_back(s);
            }
        return v;
    }();
    return c;
}

TuneResult autotune_system(int gpu, const SystemLayout& L, uint32_t latticeKind, uint64_t seed, uint32_t mcs) {
    if (L.nMonomers == 0) throw std::invalid_argument("autotune: empty system");
    const Key key = system_key(0, L, latticeKind, 1);
    {
        std::lock_guard<std::mutex> lock(g_cacheMutex);
        auto it = cache().find(key);
        if (it != cache().end()) {
            TuneResult r = it->second;
            r.cached = true;
            return r;
        }
    }
    hip_check(hipSetDevice(gpu), "hipSetDevice");
    hipStream_t st = nullptr;
    hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    std::vector<TuneTrial> trials;
    try {
        for (const LaunchShape& s : tune_candidates()) trials.push_back({s, time_system(L, latticeKind, seed, s, mcs, st)});
    } catch (...) {
        (void)hipStreamDestroy(st);
        throw;
    }
    (void)hipStreamDestroy(st);
    TuneResult r = pick(std::move(trials));
    std::lock_guard<std::mutex> lock(g_cacheMutex);
    cache()[key] = r;
    return r;
}

TuneResult autotune_batch(int gpu, const std::vector<SystemLayout>& sample, const std::vector<uint64_t>& seeds,
                          uint32_t latticeKind, uint32_t mcs) {
    if (sample.empty() || sample.size() != seeds.size()) throw std::invalid_argument("autotune: empty sample");
    uint64_t monomers = 0;
    for (const SystemLayout& L : sample) monomers += L.nMonomers;
    Key key = system_key(1, sample.front(), latticeKind, uint32_t(sample.size()));
    std::get<5>(key) = log2_ceil(monomers);
    {
        std::lock_guard<std::mutex> lock(g_cacheMutex);
        auto it = cache().find(key);
        if (it != cache().end()) {
            TuneResult r = it->second;
            r.cached = true;
            return r;
        }
    }
    std::vector<TuneTrial> trials;
    for (const LaunchShape& s : tune_candidates()) {
        uint64_t tiles = 0;
        for (const SystemLayout& L : sample)
            tiles += std::max<uint32_t>(1u, tiles_for(L, s.block * s.slots));
        if (tiles > (uint64_t(1) << 26)) continue;
        Batch b(gpu, latticeKind, uint32_t(tiles), uint32_t(sample.size()), s.block, s.slots);
        for (size_t i = 0; i < sample.size(); ++i) b.add(sample[i], seeds[i]);
        // The warm-up has the timed run's length, so the graphs the timed runs
        // replay are captured outside them; the faster of two runs is kept.
        b.run(mcs);
        const double ms = std::min(b.run(mcs), b.run(mcs));
        trials.push_back({s, ms / std::max<uint32_t>(mcs, 1u)});
    }
    if (trials.empty()) throw std::runtime_error("autotune: no candidate fits");
    TuneResult r = pick(std::move(trials));
    std::lock_guard<std::mutex> lock(g_cacheMutex);
    cache()[key] = r;
    return r;
}

}  // namespace pscbfm
