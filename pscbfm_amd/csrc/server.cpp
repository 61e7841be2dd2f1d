// Continuous-batching request loop (see server.h).
#include "server.h"

#include <algorithm>
#include <chrono>
#include <stdexcept>

namespace pscbfm {

Server::Server(int gpu, uint32_t latticeKind, uint64_t hbmBudget, uint32_t tileCapacity, uint32_t replicaCapacity,
               uint32_t block, uint32_t slots, uint32_t quantum)
    : budget_(hbmBudget), quantum_(std::max<uint32_t>(quantum, 1u)) {
    if (tileCapacity == 0) throw std::invalid_argument("server: tile capacity must be fixed (> 0)");
    batch_.reset(new Batch(gpu, latticeKind, tileCapacity, replicaCapacity, block, slots));
    stats_.tileCapacity = batch_->tileCapacity();
    t0_ = std::chrono::steady_clock::now();
    try {
        for (unsigned k = 0; k < kPrepThreads; ++k) prep_.emplace_back([this] { prepLoop(); });
        worker_ = std::thread([this] { gpuLoop(); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        cvPrep_.notify_all();
        cvGpu_.notify_all();
        for (std::thread& t : prep_)
            if (t.joinable()) t.join();
        if (worker_.joinable()) worker_.join();
        throw;
    }
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lock(mu_);
        stop_ = true;
    }
    cvPrep_.notify_all();
    cvGpu_.notify_all();
    for (std::thread& t : prep_)
        if (t.joinable()) t.join();
    if (worker_.joinable()) worker_.join();
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

uint64_t Server::submit(StagedSystem&& s, uint64_t nMCS) {
    std::unique_ptr<Pending> p(new Pending());
    p->staged = std::move(s);
    const double t = now();
    std::lock_guard<std::mutex> lock(mu_);
    if (!fatal_.empty()) throw std::runtime_error("server stopped: " + fatal_);
    p->id = nextId_++;
    JobRecord& r = jobs_[p->id];
    r.nMCS = nMCS;
    r.nMonomers = uint32_t(p->staged.input.xyz.size() / 3);
    r.tSubmit = t;
    ++stats_.submitted;
    const uint64_t id = p->id;
    toPrepare_.push_back(std::move(p));
    cvPrep_.notify_one();
    return id;
}

void Server::finish(uint64_t job, JobState st, std::string error) {
    // caller holds mu_
    JobRecord& r = jobs_[job];
    r.state = st;
    r.error = std::move(error);
    r.tDone = now();
    if (st == JobState::Done)
        ++stats_.done;
    else
        ++stats_.failed;
    cvDone_.notify_all();
}

void Server::prepLoop() {
    for (;;) {
        std::unique_ptr<Pending> p;
        {
            std::unique_lock<std::mutex> lock(mu_);
            cvPrep_.wait(lock, [this] { return stop_ || !toPrepare_.empty(); });
            if (stop_) return;
            p = std::move(toPrepare_.front());
            toPrepare_.pop_front();
        }
        std::string err;
        try {
            p->layout = layout_of(p->staged);
            p->staged.input = SystemInput{};  // the layout holds all the sweep needs
        } catch (const std::exception& e) {
            err = e.what();
        }
        std::lock_guard<std::mutex> lock(mu_);
        if (!fatal_.empty()) continue;  // the loop has stopped; every pending job is already failed
        if (!err.empty()) {
            finish(p->id, JobState::Failed, err);
            continue;
        }
        JobRecord& r = jobs_[p->id];
        r.footprint = batch_->footprint(p->layout);
        if (r.footprint > budget_ || batch_->tilesNeeded(p->layout) > batch_->tileCapacity()) {
            finish(p->id, JobState::Failed, "job exceeds the HBM budget or the tile table on its own");
            continue;
        }
        r.state = JobState::Prepared;
        // in arrival order (ids grow with submission), whichever thread
        // finished first
        auto at = std::upper_bound(prepared_.begin(), prepared_.end(), p->id,
                                   [](uint64_t id, const std::unique_ptr<Pending>& q) { return id < q->id; });
        prepared_.insert(at, std::move(p));
        cvGpu_.notify_one();
    }
}

void Server::gpuLoop() {
    // A job that has waited this long at the head of the queue is no longer
    // overtaken by later, smaller jobs (no starvation of large systems).
    constexpr double kMaxOvertakeSeconds = 2.0;
    try {
        for (;;) {
            // 1. Pick the jobs to admit: first fit in arrival order under the
            //    byte budget and the free workgroups.
            std::vector<std::unique_ptr<Pending>> admit;
            {
                std::unique_lock<std::mutex> lock(mu_);
                cvGpu_.wait(lock, [this] { return stop_ || !prepared_.empty() || !replicaJob_.empty(); });
                if (stop_) return;
                uint64_t bytes = stats_.residentBytes;
                uint32_t tiles = batch_->usedTiles();
                uint32_t replicas = uint32_t(replicaJob_.size());
                for (auto it = prepared_.begin(); it != prepared_.end();) {
                    const JobRecord& r = jobs_[(*it)->id];
                    const uint32_t need = batch_->tilesNeeded((*it)->layout);
                    const bool fits = bytes + r.footprint <= budget_ && tiles + need <= batch_->tileCapacity() &&
                                      replicas < batch_->replicaCapacity();
                    if (!fits) {
                        if (now() - r.tSubmit > kMaxOvertakeSeconds) break;
                        ++it;
                        continue;
                    }
                    bytes += r.footprint;
                    tiles += need;
                    ++replicas;
                    admit.push_back(std::move(*it));
                    it = prepared_.erase(it);
                }
                if (admit.empty() && replicaJob_.empty()) {
                    // nothing fits and nothing runs: cannot happen for jobs that fit alone
                    cvGpu_.wait_for(lock, std::chrono::milliseconds(1));
                    continue;
                }
            }
            // 2. Admit (device allocation and upload, outside the lock).  A
            //    job whose allocation fails fails alone; the loop goes on.
            for (auto& p : admit) {
                uint32_t rid = 0;
                try {
                    rid = batch_->add(p->layout, p->staged.seed, p->staged.startMcs);
                } catch (const std::exception& e) {
                    std::lock_guard<std::mutex> lock(mu_);
                    finish(p->id, JobState::Failed, std::string("admission failed: ") + e.what());
                    continue;
                }
                std::lock_guard<std::mutex> lock(mu_);
                JobRecord& r = jobs_[p->id];
                batch_->setTarget(rid, r.nMCS);
                replicaJob_[rid] = p->id;
                r.state = JobState::Running;
                r.tAdmit = now();
                stats_.residentBytes += r.footprint;
                stats_.peakResidentBytes = std::max(stats_.peakResidentBytes, stats_.residentBytes);
                stats_.usedTiles = batch_->usedTiles();
            }
            // 3. Launch up to the nearest completion, at most `quantum` substeps.
            uint64_t step = quantum_;
            for (const auto& kv : replicaJob_) {
                const BatchReplicaInfo& in = batch_->info(kv.first);
                step = std::min<uint64_t>(step, in.stopAt - in.substep);
            }
            step = std::max<uint64_t>(step, 1u);
            const double tl = now();
            batch_->launch(step);
            batch_->synchronize();
            const double busy = now() - tl;
            // 4. Retire finished replicas (copy-back outside the lock).
            std::vector<std::pair<uint32_t, uint64_t>> finished;
            for (const auto& kv : replicaJob_) {
                const BatchReplicaInfo& in = batch_->info(kv.first);
                if (in.substep >= in.stopAt) finished.push_back(kv);
            }
            std::vector<std::vector<int32_t>> coords(finished.size());
            std::vector<uint64_t> accepted(finished.size());
            std::vector<uint32_t> ids(finished.size()), counts(finished.size());
            std::vector<int32_t*> outs(finished.size());
            for (size_t k = 0; k < finished.size(); ++k) {
                const BatchReplicaInfo& in = batch_->info(finished[k].first);
                coords[k].resize(size_t(in.nMonomers) * 3);
                ids[k] = finished[k].first;
                counts[k] = in.nMonomers;
                outs[k] = coords[k].data();
                accepted[k] = in.accepted;
            }
            // every finished replica read back with one synchronization
            batch_->positionsMany(ids.data(), uint32_t(ids.size()), outs.data(), counts.data());
            for (size_t k = 0; k < finished.size(); ++k) batch_->remove(finished[k].first);
            std::lock_guard<std::mutex> lock(mu_);
            stats_.gpuBusySeconds += busy;
            ++stats_.launches;
            stats_.substeps += step;
            for (size_t k = 0; k < finished.size(); ++k) {
                JobRecord& r = jobs_[finished[k].second];
                r.xyz = std::move(coords[k]);
                r.accepted = accepted[k];
                stats_.residentBytes -= r.footprint;
                replicaJob_.erase(finished[k].first);
                finish(finished[k].second, JobState::Done, "");
            }
            stats_.usedTiles = batch_->usedTiles();
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lock(mu_);
        fatal_ = e.what();
        toPrepare_.clear();
        prepared_.clear();
        for (auto& kv : jobs_)
            if (kv.second.state != JobState::Done && kv.second.state != JobState::Failed)
                finish(kv.first, JobState::Failed, fatal_);
    }
}

JobState Server::state(uint64_t job) const {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) throw std::invalid_argument("server: no such job");
    return it->second.state;
}

bool Server::wait(uint64_t job, double timeoutSeconds) {
    std::unique_lock<std::mutex> lock(mu_);
    auto done = [&] {
        auto it = jobs_.find(job);
        if (it == jobs_.end()) throw std::invalid_argument("server: no such job");
        return it->second.state == JobState::Done || it->second.state == JobState::Failed;
    };
    if (timeoutSeconds < 0) {
        cvDone_.wait(lock, done);
        return true;
    }
    return cvDone_.wait_for(lock, std::chrono::duration<double>(timeoutSeconds), done);
}

JobRecord Server::record(uint64_t job) const {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) throw std::invalid_argument("server: no such job");
    return it->second;
}

void Server::release(uint64_t job) {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) throw std::invalid_argument("server: no such job");
    if (it->second.state != JobState::Done && it->second.state != JobState::Failed)
        throw std::runtime_error("server: job has not finished");
    jobs_.erase(it);
}

void Server::drain() {
    std::unique_lock<std::mutex> lock(mu_);
    cvDone_.wait(lock, [this] { return stats_.done + stats_.failed >= stats_.submitted; });
}

ServerStats Server::stats() const {
    std::lock_guard<std::mutex> lock(mu_);
    ServerStats s = stats_;
    s.queued = uint32_t(toPrepare_.size() + prepared_.size());
    s.running = uint32_t(replicaJob_.size());
    return s;
}

}  // namespace pscbfm
