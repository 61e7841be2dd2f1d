// Continuous-batching request loop over one Batch.
//
// Jobs (a staged system + a number of sweeps) are submitted from any thread.
// A few preparation threads build their layouts (a large system does not hold
// up the small ones behind it); the GPU thread admits prepared
// jobs whenever their device footprint fits the HBM budget and their
// workgroups fit the tile table, launches the live replicas up to the next
// completion (at most `quantum` substeps at a time, so that arrivals are
// admitted while others run), retires finished replicas and keeps their
// coordinates until the caller fetches them.  Every job's trajectory is the
// one it would have alone (its own seed and substep counter).
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.h"
#include "updater.h"

namespace pscbfm {

enum class JobState : int { Queued = 0, Prepared = 1, Running = 2, Done = 3, Failed = 4 };

struct JobRecord {
    JobState state = JobState::Queued;
    uint64_t nMCS = 0;
    uint32_t nMonomers = 0;
    uint64_t footprint = 0;       // device bytes while resident
    uint64_t accepted = 0;
    double tSubmit = 0, tAdmit = 0, tDone = 0;  // seconds since the server started
    std::vector<int32_t> xyz;     // final coordinates (Done)
    std::string error;            // (Failed)
};

struct ServerStats {
    uint64_t submitted = 0, done = 0, failed = 0;
    uint32_t queued = 0, running = 0;
    uint64_t residentBytes = 0, peakResidentBytes = 0;
    uint32_t usedTiles = 0, tileCapacity = 0;
    uint64_t launches = 0, substeps = 0;
    double gpuBusySeconds = 0;  // wall time spent in launch+synchronize
};

class Server {
public:
    // hbmBudget: device bytes the resident replicas may take together;
    // tileCapacity / replicaCapacity / block / slots as for Batch;
    // quantum: most substeps launched between two admissions.
    Server(int gpu, uint32_t latticeKind, uint64_t hbmBudget, uint32_t tileCapacity, uint32_t replicaCapacity,
           uint32_t block, uint32_t slots, uint32_t quantum);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // Queues a job; returns its id.  Thread-safe.
    uint64_t submit(StagedSystem&& s, uint64_t nMCS);
    JobState state(uint64_t job) const;
    // Waits until the job is Done or Failed (timeout < 0: forever); false on timeout.
    bool wait(uint64_t job, double timeoutSeconds);
    // Copies the record (coordinates included) of a finished job.
    JobRecord record(uint64_t job) const;
    // Forgets a finished job.
    void release(uint64_t job);
    // Waits until every submitted job has finished.
    void drain();
    ServerStats stats() const;
    double now() const;

private:
    struct Pending {
        uint64_t id;
        StagedSystem staged;
        SystemLayout layout;
    };
    void prepLoop();
    void gpuLoop();
    void finish(uint64_t job, JobState st, std::string error);

    uint64_t budget_;
    uint32_t quantum_;
    std::unique_ptr<Batch> batch_;
    std::chrono::steady_clock::time_point t0_;

    mutable std::mutex mu_;
    std::condition_variable cvPrep_, cvGpu_, cvDone_;
    bool stop_ = false;
    uint64_t nextId_ = 1;
    std::deque<std::unique_ptr<Pending>> toPrepare_;
    std::deque<std::unique_ptr<Pending>> prepared_;
    std::map<uint64_t, JobRecord> jobs_;
    std::map<uint32_t, uint64_t> replicaJob_;  // batch replica id -> job id
    ServerStats stats_;
    std::string fatal_;

    static constexpr unsigned kPrepThreads = 4;
    std::vector<std::thread> prep_;
    std::thread worker_;
};

}  // namespace pscbfm
