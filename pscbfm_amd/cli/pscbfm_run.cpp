// Command-line simulator: the counterpart of the reference's
// mainCUDASimulatorGPUScBFM_AB_Type.cpp:26-172 (same options), driving the
// MI355X sweep through the C ABI.
//
//   pscbfm_run -i in.bfm -m MAX_MCS [-s SAVE_INTERVAL] [-o out.bfm] [-g GPU]
//              [-e seeds] [--seed N] [--lattice bit|u8]
//
// Reads the last configuration of the input file, runs MAX_MCS sweeps in
// blocks of SAVE_INTERVAL and appends a frame to the output file after each
// block (the header and the initial frame first), as the reference's task
// loop with AnalyzerWriteBfmFile does.  With -e the per-substep species and
// kernel seed come from R250 seeded from the seed file (state = words 7..256),
// after the two draws the reference's main takes before its run
// (mainCUDASimulatorGPUScBFM_AB_Type.cpp:138-139), so the stream the sweep
// sees is the reference's.  Without -e they come from the counter draw keyed
// by --seed, or by a random key when neither is given (the reference seeds
// its R250 at random then).  The counter draw continues from the input
// frame's age, so a run resumed from the last frame of its own output file
// (which the reader takes) equals the uninterrupted run.
#include <getopt.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>
#include <string>

#include "pscbfm.h"

namespace {

void print_help() {
    std::cout << "usage: pscbfm_run [options]\n\n"
                 "Simple simulator for the scBFM with excluded volume and bond check on an MI355X GPU\n\n"
                 "    -e, --seeds <file path>\n"
                 "        seed file (LeMonADE layout, >= 257 32-bit words); the substep draws use its R250 stream\n"
                 "    -i, --initial-state <file path>\n"
                 "        (required) BFM file to load the configuration to simulate from (last frame)\n"
                 "    -m, --max-mcs <integer>\n"
                 "        (required) total Monte-Carlo steps to simulate\n"
                 "    -s, --save-interval <integer>\n"
                 "        save after every <integer> Monte-Carlo steps to the output file (default: max-mcs)\n"
                 "    -o, --output <file path>\n"
                 "        frames are appended to this file even if it already exists (default outfile.bfm)\n"
                 "    -g, --gpu <integer>\n"
                 "        GPU to use, 0 .. number of GPUs - 1\n"
                 "        --seed <integer>      64-bit key of the counter-based draws (used without -e)\n"
                 "        --lattice bit|u8      occupancy lattice storage (default bit)\n";
}

struct Handle {
    pscbfm_updater* h = nullptr;
    ~Handle() { pscbfm_destroy(h); }
};

int fail(const Handle& u, const std::string& what, int code) {
    std::cerr << what << " failed (" << code << "): " << pscbfm_last_error(u.h) << "\n";
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    using clock = std::chrono::steady_clock;
    std::string infile, outfile = "outfile.bfm", seedFile;
    uint64_t maxMcs = 0, saveInterval = 0, seed = 0;
    bool haveSeed = false;
    int gpu = 0, lattice = PSCBFM_LATTICE_BIT;
    if (argc <= 1) {
        print_help();
        return 0;
    }
    static struct option opts[] = {{"seeds", required_argument, 0, 'e'},         {"gpu", required_argument, 0, 'g'},
                                   {"help", no_argument, 0, 'h'},                {"initial-state", required_argument, 0, 'i'},
                                   {"max-mcs", required_argument, 0, 'm'},       {"output", required_argument, 0, 'o'},
                                   {"save-interval", required_argument, 0, 's'}, {"seed", required_argument, 0, 'S'},
                                   {"lattice", required_argument, 0, 'L'},       {0, 0, 0, 0}};
    for (;;) {
        int idx = 0;
        const int c = getopt_long(argc, argv, "e:g:hi:m:o:s:", opts, &idx);
        if (c == -1) break;
        switch (c) {
            case 'e': seedFile = optarg; break;
            case 'g': gpu = std::atoi(optarg); break;
            case 'h': print_help(); return 0;
            case 'i': infile = optarg; break;
            case 'm': maxMcs = std::strtoull(optarg, nullptr, 10); break;
            case 'o': outfile = optarg; break;
            case 's': saveInterval = std::strtoull(optarg, nullptr, 10); break;
            case 'S': seed = std::strtoull(optarg, nullptr, 0); haveSeed = true; break;
            case 'L':
                if (std::strcmp(optarg, "bit") == 0) lattice = PSCBFM_LATTICE_BIT;
                else if (std::strcmp(optarg, "u8") == 0) lattice = PSCBFM_LATTICE_U8;
                else { std::cerr << "unknown lattice " << optarg << "\n"; return 1; }
                break;
            default: std::cerr << "unknown option\n"; return 1;
        }
    }
    if (infile.empty() || maxMcs == 0) {
        std::cerr << "-i and -m are required\n";
        return 1;
    }
    if (saveInterval == 0) saveInterval = maxMcs;
    if (!seedFile.empty() && haveSeed) {
        std::cerr << "-e and --seed select different draws; give one of them\n";
        return 1;
    }
    if (seedFile.empty() && !haveSeed) seed = (uint64_t(std::random_device{}()) << 32) | std::random_device{}();
    std::cerr << "infile          = " << infile << "\noutfile         = " << outfile
              << "\nmax_mcs         = " << maxMcs << "\nsave_interval   = " << saveInterval << "\n";
    if (seedFile.empty()) std::cerr << "seed            = " << seed << " (counter-based draws)\n";

    const auto t0 = clock::now();
    Handle u;
    int rc = pscbfm_create(&u.h);
    if (rc) { std::cerr << "create failed: " << pscbfm_create_error() << "\n"; return 1; }
    // Host-side inputs first, so a bad file is reported before the GPU is touched.
    uint64_t mcs = 0;
    if ((rc = pscbfm_bfm_stage(u.h, infile.c_str(), &mcs))) return fail(u, "reading " + infile, rc);
    if (seedFile.empty()) {
        if ((rc = pscbfm_set_seed(u.h, seed))) return fail(u, "setSeed", rc);
    } else {
        std::cerr << "Use seeds from: " << seedFile << "\n";
        if ((rc = pscbfm_r250_from_seed_file(u.h, seedFile.c_str(), 0))) return fail(u, "reading " + seedFile, rc);
        uint32_t first[2];
        if ((rc = pscbfm_r250_next(u.h, first, 2))) return fail(u, "R250", rc);
        std::cerr << "RNG rand: " << first[0] << "\nRNG rand: " << first[1] / 4294967296.0 << "\n";
    }
    if ((rc = pscbfm_set_lattice_kind(u.h, lattice))) return fail(u, "setLatticeKind", rc);
    // the draws continue from the frame's age: resuming from a saved frame
    // continues the uninterrupted run
    if ((rc = pscbfm_set_start_mcs(u.h, mcs))) return fail(u, "setStartMcs", rc);
    if ((rc = pscbfm_set_gpu(u.h, gpu))) return fail(u, "setGpu", rc);
    if ((rc = pscbfm_initialize(u.h))) return fail(u, "initialize", rc);
    if ((rc = pscbfm_bfm_write(u.h, outfile.c_str(), mcs, 1))) return fail(u, "write", rc);
    const auto t1 = clock::now();
    uint64_t accepted = 0, attempted = 0;
    double gpuMs = 0;
    for (uint64_t done = 0; done + saveInterval <= maxMcs; done += saveInterval) {
        uint64_t left = saveInterval;
        while (left > 0) {  // the run call takes an int32 count, as the reference's does
            const int32_t chunk = int32_t(left > 0x7fffffffull ? 0x7fffffffull : left);
            pscbfm_run_stats st{};
            if ((rc = pscbfm_run_ex(u.h, chunk, nullptr, 0, 1, &st))) return fail(u, "runSimulationOnGPU", rc);
            accepted += st.accepted;
            attempted += st.attempted;
            gpuMs += st.gpu_ms;
            left -= uint64_t(chunk);
        }
        mcs += saveInterval;
        if ((rc = pscbfm_bfm_write(u.h, outfile.c_str(), mcs, 1))) return fail(u, "write", rc);
    }
    const auto t2 = clock::now();
    const double loop = std::chrono::duration<double>(t2 - t1).count();
    std::cerr << "tInit = " << std::chrono::duration<double>(t1 - t0).count() << "s\n"
              << "tTaskLoop = " << loop << "s\n"
              << "attempted moves = " << attempted << ", accepted = " << accepted << ", GPU time = " << gpuMs
              << " ms (" << (gpuMs > 0 ? attempted / (gpuMs * 1e-3) : 0.0) << " attempts/s)\n";
    if ((rc = pscbfm_cleanup(u.h))) return fail(u, "cleanup", rc);
    return 0;
}
