// extern "C" boundary of the scBFM sweep (include/pscbfm.h).  Exceptions of
// the engine become status codes plus a per-handle message.
#include "pscbfm.h"

#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "autotune.h"
#include "batch.h"
#include "bfm_io.h"
#include "coloring.h"
#include "common.h"
#include "generators.h"
#include "server.h"
#include "updater.h"

struct pscbfm_updater {
    pscbfm::Updater engine;
    std::string error;
};

struct pscbfm_batch {
    std::unique_ptr<pscbfm::Batch> engine;
    std::string error;
};

struct pscbfm_server {
    std::unique_ptr<pscbfm::Server> engine;
    std::mutex errorMutex;
    std::string error;
};

namespace {

thread_local std::string g_createError;

// Runs f, turning its exceptions into a status code and a message.
template <class F>
int translated(std::string& error, F&& f) {
    error.clear();
    try {
        f();
        return PSCBFM_OK;
    } catch (const std::invalid_argument& e) {
        error = e.what();
        return PSCBFM_INVALID_ARGUMENT;
    } catch (const std::out_of_range& e) {
        error = e.what();
        return PSCBFM_OUT_OF_RANGE;
    } catch (const std::runtime_error& e) {
        error = e.what();
        return std::strncmp(e.what(), "HIP error", 9) == 0 ? PSCBFM_DEVICE_ERROR : PSCBFM_RUNTIME_ERROR;
    } catch (const std::bad_alloc&) {
        error = "host allocation failed";
        return PSCBFM_RUNTIME_ERROR;
    } catch (const std::exception& e) {
        error = e.what();
        return PSCBFM_INTERNAL_ERROR;
    } catch (...) {
        error = "unknown exception";
        return PSCBFM_INTERNAL_ERROR;
    }
}

template <class F>
int guarded(pscbfm_updater* h, F&& f) {
    if (!h) return PSCBFM_NULL_HANDLE;
    return translated(h->error, [&] { f(h->engine); });
}

template <class F>
int batch_guarded(pscbfm_batch* b, F&& f) {
    if (!b || !b->engine) return PSCBFM_NULL_HANDLE;
    return translated(b->error, [&] { f(*b->engine); });
}

template <class F>
int server_guarded(pscbfm_server* s, F&& f) {
    if (!s || !s->engine) return PSCBFM_NULL_HANDLE;
    std::string err;
    const int rc = translated(err, [&] { f(*s->engine); });
    std::lock_guard<std::mutex> lock(s->errorMutex);
    s->error = err;
    return rc;
}

template <class F>
int guarded_const(const pscbfm_updater* h, F&& f) {
    return guarded(const_cast<pscbfm_updater*>(h), [&](pscbfm::Updater& u) { f(static_cast<const pscbfm::Updater&>(u)); });
}

void fill_stats(const pscbfm::RunStats& r, pscbfm_run_stats* s) {
    s->gpu_ms = r.gpuMilliseconds;
    s->substeps = r.substeps;
    s->attempted = r.attempted;
    s->accepted = r.accepted;
    s->bond_terms = r.bondTerms;
    s->check_ms = r.checkMilliseconds;
    s->perform_ms = r.performMilliseconds;
    s->check_launches = r.checkLaunches;
    s->perform_launches = r.performLaunches;
}

}  // namespace

extern "C" {

int pscbfm_create(pscbfm_updater** out) {
    if (!out) return PSCBFM_INVALID_ARGUMENT;
    *out = nullptr;
    try {
        *out = new pscbfm_updater();
        return PSCBFM_OK;
    } catch (const std::exception& e) {
        g_createError = e.what();
        return PSCBFM_RUNTIME_ERROR;
    }
}

void pscbfm_destroy(pscbfm_updater* h) { delete h; }

const char* pscbfm_last_error(const pscbfm_updater* h) { return h ? h->error.c_str() : "null handle"; }

const char* pscbfm_create_error(void) { return g_createError.c_str(); }

int pscbfm_destruct(pscbfm_updater* h) {
    return guarded(h, [](pscbfm::Updater& u) { u.destruct(); });
}

int pscbfm_set_gpu(pscbfm_updater* h, int gpu) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setGpu(gpu); });
}

int pscbfm_copy_bond_set(pscbfm_updater* h, int dx, int dy, int dz, int forbidden) {
    return guarded(h, [&](pscbfm::Updater& u) { u.copyBondSet(dx, dy, dz, forbidden != 0); });
}

int pscbfm_set_nr_of_all_monomers(pscbfm_updater* h, uint32_t n) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setNrOfAllMonomers(n); });
}

int pscbfm_set_attribute(pscbfm_updater* h, uint32_t i, int32_t attribute) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setAttribute(i, attribute); });
}

int pscbfm_set_monomer_coordinates(pscbfm_updater* h, uint32_t i, int32_t x, int32_t y, int32_t z) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setMonomerCoordinates(i, x, y, z); });
}

int pscbfm_set_monomer_coordinates_bulk(pscbfm_updater* h, const int32_t* xyz, uint32_t n) {
    return guarded(h, [&](pscbfm::Updater& u) {
        if (!xyz && n) throw std::invalid_argument("null coordinate array");
        u.setMonomerCoordinatesBulk(xyz, n);
    });
}

int pscbfm_set_connectivity(pscbfm_updater* h, uint32_t i, uint32_t j) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setConnectivity(i, j); });
}

int pscbfm_set_connectivity_bulk(pscbfm_updater* h, const uint32_t* pairs, uint64_t n_pairs) {
    return guarded(h, [&](pscbfm::Updater& u) {
        if (!pairs && n_pairs) throw std::invalid_argument("null pair array");
        u.setConnectivityBulk(pairs, n_pairs);
    });
}

int pscbfm_set_lattice_size(pscbfm_updater* h, uint32_t x, uint32_t y, uint32_t z) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setLatticeSize(x, y, z); });
}

int pscbfm_set_periodicity(pscbfm_updater* h, int x, int y, int z) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setPeriodicity(x != 0, y != 0, z != 0); });
}

int pscbfm_populate_lattice(pscbfm_updater* h) {
    return guarded(h, [](pscbfm::Updater& u) { u.populateLattice(); });
}

int pscbfm_initialize(pscbfm_updater* h) {
    return guarded(h, [](pscbfm::Updater& u) { u.initialize(); });
}

int pscbfm_execute(pscbfm_updater* h) { return h ? PSCBFM_OK : PSCBFM_NULL_HANDLE; }

int pscbfm_run(pscbfm_updater* h, int32_t n_mcs, uint64_t* accepted_out) {
    return guarded(h, [&](pscbfm::Updater& u) {
        u.runSimulationOnGPU(n_mcs);
        if (accepted_out) *accepted_out = u.lastRun().accepted;
    });
}

int pscbfm_get_monomer_position(const pscbfm_updater* h, uint32_t i, int32_t* x, int32_t* y, int32_t* z) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        const int32_t vx = u.getMonomerPositionInX(i);
        const int32_t vy = u.getMonomerPositionInY(i);
        const int32_t vz = u.getMonomerPositionInZ(i);
        if (x) *x = vx;
        if (y) *y = vy;
        if (z) *z = vz;
    });
}

int pscbfm_get_positions_bulk(const pscbfm_updater* h, int32_t* xyz, uint32_t n) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        if (!xyz && n) throw std::invalid_argument("null coordinate array");
        u.getPositionsBulk(xyz, n);
    });
}

int pscbfm_check_system(const pscbfm_updater* h) {
    return guarded_const(h, [](const pscbfm::Updater& u) { u.checkSystem(); });
}

int pscbfm_cleanup(pscbfm_updater* h) {
    return guarded(h, [](pscbfm::Updater& u) { u.cleanup(); });
}

int pscbfm_set_seed(pscbfm_updater* h, uint64_t seed) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setSeed(seed); });
}

int pscbfm_set_r250_state(pscbfm_updater* h, const uint32_t* words) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setR250State(words); });
}

int pscbfm_r250_from_seed_file(pscbfm_updater* h, const char* path, uint32_t discard) {
    return guarded(h, [&](pscbfm::Updater& u) {
        if (!path) throw std::invalid_argument("[seed file] null path");
        u.loadR250SeedFile(path, discard);
    });
}

int pscbfm_r250_next(pscbfm_updater* h, uint32_t* out, uint64_t n) {
    return guarded(h, [&](pscbfm::Updater& u) { u.r250Draw(out, n); });
}

int pscbfm_set_draw_chunk(pscbfm_updater* h, uint64_t substeps) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setDrawChunk(substeps); });
}

int pscbfm_r250_get_state(const pscbfm_updater* h, uint32_t* words) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        if (!words) throw std::invalid_argument("[r250State] null output");
        u.r250State(words);
    });
}

int pscbfm_set_start_mcs(pscbfm_updater* h, uint64_t mcs) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setStartMcs(mcs); });
}

int pscbfm_set_lattice_kind(pscbfm_updater* h, int kind) {
    return guarded(h, [&](pscbfm::Updater& u) {
        if (kind < 0) throw std::invalid_argument("[setLatticeKind] unknown kind");
        u.setLatticeKind(uint32_t(kind));
    });
}

int pscbfm_set_uniform_coloring(pscbfm_updater* h, int on) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setUniformColoring(on != 0); });
}

int pscbfm_set_check_on_initialize(pscbfm_updater* h, int on) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setCheckOnInitialize(on != 0); });
}

int pscbfm_set_colors(pscbfm_updater* h, const uint8_t* colors, uint32_t n) {
    return guarded(h, [&](pscbfm::Updater& u) {
        if (!colors && n) throw std::invalid_argument("null color array");
        u.setColors(colors, n);
    });
}

int pscbfm_set_perform_block(pscbfm_updater* h, uint32_t block) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setPerformBlock(block); });
}

int pscbfm_set_launch_shape(pscbfm_updater* h, uint32_t block, uint32_t slots) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setLaunchShape(block, slots); });
}

int pscbfm_run_ex_noalloc(pscbfm_updater* h, int32_t n_mcs, uint32_t* log, pscbfm_run_stats* stats) {
    return guarded(h, [&](pscbfm::Updater& u) {
        u.runNoAlloc(n_mcs, log);
        if (stats) fill_stats(u.lastRun(), stats);
    });
}

int pscbfm_set_perform_walk(pscbfm_updater* h, uint32_t threads_per_list, uint32_t entries_per_thread) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setPerformWalk(threads_per_list, entries_per_thread); });
}

int pscbfm_set_slot_order(pscbfm_updater* h, int order) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setSlotOrder(static_cast<pscbfm::SlotOrder>(order)); });
}

int pscbfm_set_clear_pass(pscbfm_updater* h, uint32_t pass) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setClearPass(pass); });
}

int pscbfm_set_read_back(pscbfm_updater* h, int on) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setReadBack(on != 0); });
}

int pscbfm_read_positions_back(pscbfm_updater* h) {
    return guarded(h, [&](pscbfm::Updater& u) { u.readPositionsBack(); });
}

int pscbfm_set_graph_chunk(pscbfm_updater* h, uint32_t substeps) {
    return guarded(h, [&](pscbfm::Updater& u) { u.setGraphChunk(substeps); });
}

int pscbfm_prepare_graphs(pscbfm_updater* h, int32_t n_mcs) {
    return guarded(h, [&](pscbfm::Updater& u) { u.prepareGraphs(n_mcs); });
}

int pscbfm_run_ex(pscbfm_updater* h, int32_t n_mcs, uint32_t* log, int profile, int use_graph,
                  pscbfm_run_stats* stats) {
    return guarded(h, [&](pscbfm::Updater& u) {
        u.run(n_mcs, log, profile != 0, use_graph != 0);
        if (stats) fill_stats(u.lastRun(), stats);
    });
}

int pscbfm_last_run_stats(const pscbfm_updater* h, pscbfm_run_stats* stats) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        if (!stats) throw std::invalid_argument("null stats");
        fill_stats(u.lastRun(), stats);
    });
}

int pscbfm_get_n_colors(const pscbfm_updater* h, uint32_t* n_colors) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        if (!u.initialized()) throw std::runtime_error("not initialized");
        if (n_colors) *n_colors = u.nColors();
    });
}

int pscbfm_get_colors(const pscbfm_updater* h, uint8_t* colors, uint32_t n) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        if (!u.initialized()) throw std::runtime_error("not initialized");
        if (n != u.nMonomers()) throw std::invalid_argument("count differs from the number of monomers");
        std::memcpy(colors, u.layout().color.data(), n);
    });
}

int pscbfm_get_color_sizes(const pscbfm_updater* h, uint32_t* sizes, uint32_t n_colors) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        if (!u.initialized()) throw std::runtime_error("not initialized");
        if (n_colors != u.nColors()) throw std::invalid_argument("wrong number of colors");
        for (uint32_t c = 0; c < n_colors; ++c) sizes[c] = u.layout().colors[c].count;
    });
}

int pscbfm_get_counters(const pscbfm_updater* h, uint64_t* substeps, uint64_t* accepted) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        if (substeps) *substeps = u.substepCounter();
        if (accepted) *accepted = u.acceptedTotal();
    });
}

int pscbfm_get_device_bytes(const pscbfm_updater* h, uint64_t* bytes) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        if (bytes) *bytes = u.deviceBytes();
    });
}

int pscbfm_download_lattice(const pscbfm_updater* h, uint8_t* cells, uint64_t n_cells) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        std::vector<uint8_t> lat;
        u.downloadLattice(lat);
        if (n_cells != lat.size()) throw std::invalid_argument("wrong number of cells");
        std::memcpy(cells, lat.data(), lat.size());
    });
}

int pscbfm_count_partner_escapes(const pscbfm_updater* h, uint64_t* partners, uint64_t* escapes) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        uint64_t p = 0, e = 0;
        u.partnerCodeStats(p, e);
        if (partners) *partners = p;
        if (escapes) *escapes = e;
    });
}

int pscbfm_count_proposal_bits(const pscbfm_updater* h, uint64_t* bits_even, uint64_t* bits_odd) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        std::vector<uint32_t> w;
        u.downloadProposal(w);
        const size_t half = w.size() / 2;
        uint64_t a = 0, b = 0;
        for (size_t i = 0; i < half; ++i) a += uint64_t(__builtin_popcount(w[i]));
        for (size_t i = half; i < w.size(); ++i) b += uint64_t(__builtin_popcount(w[i]));
        if (bits_even) *bits_even = a;
        if (bits_odd) *bits_odd = b;
    });
}

int pscbfm_bfm_stage(pscbfm_updater* h, const char* path, uint64_t* mcs) {
    return guarded(h, [&](pscbfm::Updater& u) {
        if (!path) throw std::invalid_argument("null path");
        const pscbfm::BfmData f = pscbfm::read_bfm(path);
        const uint32_t n = f.nMonomers();
        u.setPeriodicity(f.periodic[0], f.periodic[1], f.periodic[2]);
        u.setNrOfAllMonomers(n);
        u.setMonomerCoordinatesBulk(f.xyz.data(), n);
        for (uint32_t i = 0; i < n; ++i) u.setAttribute(i, f.attributes[i]);
        // both directions, grouped by source monomer as LeMonADE's molecule lists are
        std::vector<std::vector<uint32_t>> nb(n);
        for (const auto& b : f.bonds) {
            nb[b.first].push_back(b.second);
            nb[b.second].push_back(b.first);
        }
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t j : nb[i]) u.setConnectivity(i, j);
        u.setLatticeSize(f.box[0], f.box[1], f.box[2]);
        u.populateLattice();
        std::vector<uint8_t> allowed(512, 0);
        for (const auto& kv : f.bondset) {
            const auto& v = kv.second;
            allowed[uint32_t(v[0] & 7) | uint32_t(v[1] & 7) << 3 | uint32_t(v[2] & 7) << 6] = 1;
        }
        for (int dx = -4; dx < 4; ++dx)
            for (int dy = -4; dy < 4; ++dy)
                for (int dz = -4; dz < 4; ++dz)
                    u.copyBondSet(dx, dy, dz, !allowed[uint32_t(dx & 7) | uint32_t(dy & 7) << 3 | uint32_t(dz & 7) << 6]);
        if (mcs) *mcs = f.mcs;
    });
}

int pscbfm_bfm_write(const pscbfm_updater* h, const char* path, uint64_t mcs, int append) {
    return guarded_const(h, [&](const pscbfm::Updater& u) {
        if (!path) throw std::invalid_argument("null path");
        const pscbfm::SystemInput in = u.stagedInput();
        const uint32_t n = uint32_t(in.xyz.size() / 3);
        std::vector<std::pair<uint32_t, uint32_t>> bonds;
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t k = in.bonds.start[i]; k < in.bonds.start[i + 1]; ++k)
                if (i < in.bonds.ids[k]) bonds.emplace_back(i, in.bonds.ids[k]);
        std::map<int, std::array<int32_t, 3>> ids;
        for (const auto& kv : pscbfm::standard_bondset_ids()) {
            const auto& v = kv.second;
            if (!u.forbidden(uint32_t(v[0] & 7) | uint32_t(v[1] & 7) << 3 | uint32_t(v[2] & 7) << 6)) ids[kv.first] = v;
        }
        const bool periodic[3] = {true, true, true};
        pscbfm::BfmWriter w(u.box(), periodic, n, bonds, ids, u.attributes());
        std::vector<int32_t> xyz(size_t(n) * 3);
        u.getPositionsBulk(xyz.data(), n);
        bool needHeader = !append;
        if (append) {
            std::ifstream probe(path, std::ios::binary | std::ios::ate);
            needHeader = !probe || probe.tellg() <= 0;
        }
        std::ofstream out(path, needHeader ? std::ios::binary | std::ios::trunc : std::ios::binary | std::ios::app);
        if (!out) throw std::runtime_error(std::string("cannot write ") + path);
        if (needHeader) w.header(out);
        w.frame(out, mcs, xyz.data());
        if (!out) throw std::runtime_error(std::string("write failed: ") + path);
    });
}

int pscbfm_zcurve_encode(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t x, uint32_t y, uint32_t z,
                         uint32_t* code) {
    if (!code) return PSCBFM_INVALID_ARGUMENT;
    auto p2 = [](uint32_t v) { return v >= 8 && (v & (v - 1)) == 0; };
    if (!p2(lx) || !p2(ly) || !p2(lz) || uint64_t(lx) * ly * lz > (uint64_t(1) << 32)) return PSCBFM_INVALID_ARGUMENT;
    pscbfm::MortonCodec c;
    c.init(lx, ly, lz);
    *code = c.encode(x & (lx - 1), y & (ly - 1), z & (lz - 1));
    return PSCBFM_OK;
}

uint32_t pscbfm_move_hash(uint32_t a) { return pscbfm::move_hash(a); }

void pscbfm_substep_draw(uint64_t seed, uint64_t t, uint32_t n_colors, uint32_t* species, uint32_t* kernel_seed) {
    const pscbfm::SubstepDraw d = pscbfm::substep_draw(seed, t, n_colors ? n_colors : 1u);
    if (species) *species = d.species;
    if (kernel_seed) *kernel_seed = d.seed;
}

int pscbfm_color_graph(uint32_t n, const uint32_t* start, const uint32_t* ids, int uniform, uint8_t* colors) {
    try {
        pscbfm::Adjacency adj;
        adj.start.assign(start, start + n + 1);
        adj.ids.assign(ids, ids + start[n]);
        for (uint32_t id : adj.ids)
            if (id >= n) return PSCBFM_OUT_OF_RANGE;
        const std::vector<uint8_t> c = pscbfm::color_graph(adj, uniform != 0);
        std::memcpy(colors, c.data(), n);
        return PSCBFM_OK;
    } catch (const std::runtime_error&) {
        return PSCBFM_RUNTIME_ERROR;
    } catch (...) {
        return PSCBFM_INTERNAL_ERROR;
    }
}

int pscbfm_standard_bond_vectors(int32_t* xyz108) {
    if (!xyz108) return PSCBFM_INVALID_ARGUMENT;
    const auto& v = pscbfm::standard_bond_vectors();
    for (size_t i = 0; i < v.size(); ++i)
        for (int a = 0; a < 3; ++a) xyz108[3 * i + a] = v[i][size_t(a)];
    return v.size() == 108 ? PSCBFM_OK : PSCBFM_INTERNAL_ERROR;
}

namespace {
bool valid_box(uint32_t lx, uint32_t ly, uint32_t lz) {
    auto p2 = [](uint32_t v) { return v >= 8 && (v & (v - 1)) == 0; };
    return p2(lx) && p2(ly) && p2(lz) && uint64_t(lx) * ly * lz <= (uint64_t(1) << 32);
}
}  // namespace

int pscbfm_gen_linear_chains(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_chains, uint32_t len,
                             uint64_t seed, uint32_t max_restarts, int32_t* xyz) {
    if (!valid_box(lx, ly, lz) || (!xyz && n_chains && len)) return PSCBFM_INVALID_ARGUMENT;
    try {
        const uint32_t box[3] = {lx, ly, lz};
        std::vector<int32_t> out;
        if (!pscbfm::grow_linear_chains(box, n_chains, len, seed, max_restarts, out)) return PSCBFM_RUNTIME_ERROR;
        std::memcpy(xyz, out.data(), out.size() * sizeof(int32_t));
        return PSCBFM_OK;
    } catch (...) {
        return PSCBFM_INTERNAL_ERROR;
    }
}

int pscbfm_gen_linear_chains_ordered(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_chains, uint32_t len,
                                     int32_t* xyz) {
    if (!valid_box(lx, ly, lz) || (!xyz && n_chains && len)) return PSCBFM_INVALID_ARGUMENT;
    try {
        const uint32_t box[3] = {lx, ly, lz};
        std::vector<int32_t> out;
        if (!pscbfm::place_linear_chains_ordered(box, n_chains, len, out)) return PSCBFM_RUNTIME_ERROR;
        std::memcpy(xyz, out.data(), out.size() * sizeof(int32_t));
        return PSCBFM_OK;
    } catch (...) {
        return PSCBFM_INTERNAL_ERROR;
    }
}

int pscbfm_gen_star_network(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_stars, uint32_t arms,
                            uint32_t arm_length, uint64_t seed, int32_t* xyz, uint32_t* bonds,
                            uint64_t bond_capacity, uint64_t* n_bonds) {
    if (!valid_box(lx, ly, lz) || !xyz || !bonds || !n_bonds) return PSCBFM_INVALID_ARGUMENT;
    try {
        const uint32_t box[3] = {lx, ly, lz};
        std::vector<int32_t> out;
        std::vector<uint32_t> b;
        if (!pscbfm::build_star_network(box, n_stars, arms, arm_length, seed, out, b)) return PSCBFM_RUNTIME_ERROR;
        if (b.size() / 2 > bond_capacity) return PSCBFM_OUT_OF_RANGE;
        std::memcpy(xyz, out.data(), out.size() * sizeof(int32_t));
        std::memcpy(bonds, b.data(), b.size() * sizeof(uint32_t));
        *n_bonds = b.size() / 2;
        return PSCBFM_OK;
    } catch (...) {
        return PSCBFM_INTERNAL_ERROR;
    }
}

int pscbfm_gen_stars_ordered(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_stars, uint32_t arm_length,
                             uint64_t seed, int32_t* xyz, uint32_t* bonds) {
    if (!valid_box(lx, ly, lz) || (n_stars && (!xyz || !bonds))) return PSCBFM_INVALID_ARGUMENT;
    try {
        const uint32_t box[3] = {lx, ly, lz};
        std::vector<int32_t> out;
        std::vector<uint32_t> b;
        if (!pscbfm::place_stars_ordered(box, n_stars, arm_length, seed, out, b)) return PSCBFM_RUNTIME_ERROR;
        if (!out.empty()) std::memcpy(xyz, out.data(), out.size() * sizeof(int32_t));
        if (!b.empty()) std::memcpy(bonds, b.data(), b.size() * sizeof(uint32_t));
        return PSCBFM_OK;
    } catch (...) {
        return PSCBFM_INTERNAL_ERROR;
    }
}

int pscbfm_crosslink_star_ends(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_stars, uint32_t arms,
                               uint32_t arm_length, uint32_t rounds, uint64_t seed, const int32_t* xyz,
                               uint32_t* bonds, uint64_t bond_capacity, uint64_t* n_bonds) {
    if (!valid_box(lx, ly, lz) || !n_bonds || (n_stars && (!xyz || !bonds))) return PSCBFM_INVALID_ARGUMENT;
    try {
        const uint32_t box[3] = {lx, ly, lz};
        const size_t n = size_t(n_stars) * (1 + size_t(arms) * arm_length);
        const std::vector<int32_t> in(xyz, xyz + 3 * n);
        std::vector<uint32_t> b;
        pscbfm::HostRng rng(seed);
        pscbfm::crosslink_arm_ends(box, in, n_stars, arms, arm_length, rounds, rng, b);
        if (b.size() / 2 > bond_capacity) return PSCBFM_OUT_OF_RANGE;
        if (!b.empty()) std::memcpy(bonds, b.data(), b.size() * sizeof(uint32_t));
        *n_bonds = b.size() / 2;
        return PSCBFM_OK;
    } catch (...) {
        return PSCBFM_INTERNAL_ERROR;
    }
}

int pscbfm_copy_probe(int gpu, uint64_t bytes, int repeats, double* ms_per_copy) {
    if (bytes < 16 || repeats < 1 || !ms_per_copy) return PSCBFM_INVALID_ARGUMENT;
    void* src = nullptr;
    void* dst = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PSCBFM_OK;
    const uint64_t n16 = bytes / 16;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == PSCBFM_OK) rc = PSCBFM_DEVICE_ERROR;
        return rc == PSCBFM_OK;
    };
    if (ok(hipSetDevice(gpu)) && ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) &&
        ok(hipMalloc(&src, n16 * 16)) && ok(hipMalloc(&dst, n16 * 16)) && ok(hipEventCreate(&e0)) &&
        ok(hipEventCreate(&e1)) && ok(pscbfm::launch_fill_u32(static_cast<uint32_t*>(src), n16 * 4, 1u, st)) &&
        ok(pscbfm::launch_copy_u4(src, dst, n16, st)) && ok(hipEventRecord(e0, st))) {
        for (int r = 0; r < repeats && ok(pscbfm::launch_copy_u4(src, dst, n16, st)); ++r) {
        }
        float ms = 0;
        if (ok(hipEventRecord(e1, st)) && ok(hipEventSynchronize(e1)) && ok(hipEventElapsedTime(&ms, e0, e1)))
            *ms_per_copy = double(ms) / repeats;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

namespace {
void report_tune(const pscbfm::TuneResult& r, uint32_t* block, uint32_t* slots, double* trialMs) {
    if (block) *block = r.best.block;
    if (slots) *slots = r.best.slots;
    if (trialMs) {
        const auto& cand = pscbfm::tune_candidates();
        for (size_t i = 0; i < cand.size(); ++i) {
            trialMs[i] = -1.0;
            for (const auto& t : r.trials)
                if (t.shape.block == cand[i].block && t.shape.slots == cand[i].slots) trialMs[i] = t.msPerMcs;
        }
    }
}
}  // namespace

int pscbfm_autotune(pscbfm_updater* h, uint32_t mcs, uint32_t* block, uint32_t* slots, double* trial_ms) {
    return guarded(h, [&](pscbfm::Updater& u) {
        if (!u.initialized()) throw std::runtime_error("[autotune] initialize has not been called");
        if (mcs == 0) throw std::invalid_argument("[autotune] mcs must be positive");
        const pscbfm::TuneResult r = pscbfm::autotune_system(u.gpu(), u.layout(), u.latticeKind(), u.seed(), mcs);
        u.setLaunchShape(r.best.block, r.best.slots);
        report_tune(r, block, slots, trial_ms);
    });
}

// ---- batch ------------------------------------------------------------------------

int pscbfm_batch_create(int gpu, int lattice_kind, uint32_t tile_capacity, uint32_t replica_capacity,
                        uint32_t block, uint32_t slots, pscbfm_batch** out) {
    if (!out) return PSCBFM_INVALID_ARGUMENT;
    *out = nullptr;
    if (gpu < 0) {
        g_createError = "gpu must be >= 0";
        return PSCBFM_INVALID_ARGUMENT;
    }
    std::unique_ptr<pscbfm_batch> b(new (std::nothrow) pscbfm_batch());
    if (!b) return PSCBFM_RUNTIME_ERROR;
    const int rc = translated(g_createError, [&] {
        b->engine.reset(new pscbfm::Batch(gpu, uint32_t(lattice_kind), tile_capacity, replica_capacity, block, slots));
    });
    if (rc == PSCBFM_OK) *out = b.release();
    return rc;
}

void pscbfm_batch_destroy(pscbfm_batch* b) { delete b; }

const char* pscbfm_batch_last_error(const pscbfm_batch* b) { return b ? b->error.c_str() : "null handle"; }

int pscbfm_batch_footprint(pscbfm_batch* b, const pscbfm_updater* staged, uint64_t* bytes, uint32_t* tiles) {
    if (!staged) return PSCBFM_NULL_HANDLE;
    return batch_guarded(b, [&](pscbfm::Batch& e) {
        const pscbfm::SystemLayout L = staged->engine.buildLayout();
        if (bytes) *bytes = e.footprint(L);
        if (tiles) *tiles = e.tilesNeeded(L);
    });
}

// Batches and request queues draw every replica from (seed, substep); an
// updater in R250 mode draws from one sequential stream and cannot join them.
static void require_counter_draws(const pscbfm::Updater& u) {
    if (u.r250Mode())
        throw std::invalid_argument("an updater in R250 mode cannot join a batch or request queue (they use counter draws)");
}

int pscbfm_batch_add(pscbfm_batch* b, const pscbfm_updater* staged, uint32_t* id) {
    if (!staged) return PSCBFM_NULL_HANDLE;
    return batch_guarded(b, [&](pscbfm::Batch& e) {
        require_counter_draws(staged->engine);
        const uint32_t r = e.add(staged->engine.buildLayout(), staged->engine.seed(), staged->engine.startMcs());
        if (id) *id = r;
    });
}

int pscbfm_batch_set_perform_block(pscbfm_batch* b, uint32_t block) {
    return batch_guarded(b, [&](pscbfm::Batch& e) { e.setPerformBlock(block); });
}

int pscbfm_batch_remove(pscbfm_batch* b, uint32_t id) {
    return batch_guarded(b, [&](pscbfm::Batch& e) { e.remove(id); });
}

int pscbfm_batch_set_target(pscbfm_batch* b, uint32_t id, uint64_t n_mcs) {
    return batch_guarded(b, [&](pscbfm::Batch& e) { e.setTarget(id, n_mcs); });
}

int pscbfm_batch_launch(pscbfm_batch* b, uint64_t substeps) {
    return batch_guarded(b, [&](pscbfm::Batch& e) { e.launch(substeps); });
}

int pscbfm_batch_synchronize(pscbfm_batch* b) {
    return batch_guarded(b, [&](pscbfm::Batch& e) { e.synchronize(); });
}

int pscbfm_batch_run(pscbfm_batch* b, uint64_t n_mcs, double* gpu_ms) {
    return batch_guarded(b, [&](pscbfm::Batch& e) {
        const double ms = e.run(n_mcs);
        if (gpu_ms) *gpu_ms = ms;
    });
}

int pscbfm_batch_info(pscbfm_batch* b, uint32_t id, pscbfm_replica_info* info) {
    if (!info) return PSCBFM_INVALID_ARGUMENT;
    return batch_guarded(b, [&](pscbfm::Batch& e) {
        const pscbfm::BatchReplicaInfo& r = e.info(id);
        info->live = r.live ? 1 : 0;
        info->n_colors = r.nColors;
        info->n_monomers = r.nMonomers;
        info->tiles = r.tiles;
        info->substeps = r.substep;
        info->stop_at = r.stopAt;
        info->accepted = r.accepted;
        info->device_bytes = r.deviceBytes;
    });
}

int pscbfm_batch_get_positions_many(pscbfm_batch* b, const uint32_t* ids, uint32_t count, int32_t* const* xyz,
                                    const uint32_t* n) {
    if (count != 0 && (!ids || !xyz || !n)) return PSCBFM_INVALID_ARGUMENT;
    return batch_guarded(b, [&](pscbfm::Batch& e) { e.positionsMany(ids, count, xyz, n); });
}

int pscbfm_batch_get_positions(pscbfm_batch* b, uint32_t id, int32_t* xyz, uint32_t n) {
    if (!xyz && n) return PSCBFM_INVALID_ARGUMENT;
    return batch_guarded(b, [&](pscbfm::Batch& e) { e.positions(id, xyz, n); });
}

int pscbfm_batch_autotune(int gpu, int lattice_kind, const pscbfm_updater* const* staged, uint32_t n,
                          uint32_t mcs, uint32_t* block, uint32_t* slots, double* trial_ms) {
    if (!staged || n == 0 || mcs == 0 || gpu < 0) return PSCBFM_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < n; ++i)
        if (!staged[i]) return PSCBFM_NULL_HANDLE;
    return translated(g_createError, [&] {
        std::vector<pscbfm::SystemLayout> layouts;
        std::vector<uint64_t> seeds;
        layouts.reserve(n);
        for (uint32_t i = 0; i < n; ++i) {
            layouts.push_back(staged[i]->engine.buildLayout());
            seeds.push_back(staged[i]->engine.seed());
        }
        const pscbfm::TuneResult r = pscbfm::autotune_batch(gpu, layouts, seeds, uint32_t(lattice_kind), mcs);
        report_tune(r, block, slots, trial_ms);
    });
}

int pscbfm_batch_usage(pscbfm_batch* b, uint32_t* live, uint32_t* used_tiles, uint32_t* tile_capacity,
                       uint64_t* device_bytes) {
    return batch_guarded(b, [&](pscbfm::Batch& e) {
        if (live) *live = e.liveReplicas();
        if (used_tiles) *used_tiles = e.usedTiles();
        if (tile_capacity) *tile_capacity = e.tileCapacity();
        if (device_bytes) *device_bytes = e.deviceBytes();
    });
}

// ---- request loop ------------------------------------------------------------------

int pscbfm_server_create(int gpu, int lattice_kind, uint64_t hbm_budget, uint32_t tile_capacity,
                         uint32_t replica_capacity, uint32_t block, uint32_t slots, uint32_t quantum,
                         pscbfm_server** out) {
    if (!out) return PSCBFM_INVALID_ARGUMENT;
    *out = nullptr;
    if (gpu < 0) return PSCBFM_INVALID_ARGUMENT;
    std::unique_ptr<pscbfm_server> s(new (std::nothrow) pscbfm_server());
    if (!s) return PSCBFM_RUNTIME_ERROR;
    const int rc = translated(g_createError, [&] {
        s->engine.reset(new pscbfm::Server(gpu, uint32_t(lattice_kind), hbm_budget, tile_capacity, replica_capacity,
                                           block, slots, quantum));
    });
    if (rc == PSCBFM_OK) *out = s.release();
    return rc;
}

void pscbfm_server_destroy(pscbfm_server* s) { delete s; }

const char* pscbfm_server_last_error(const pscbfm_server* s) { return s ? s->error.c_str() : "null handle"; }

int pscbfm_server_submit(pscbfm_server* s, const pscbfm_updater* staged, uint64_t n_mcs, uint64_t* job) {
    if (!staged) return PSCBFM_NULL_HANDLE;
    return server_guarded(s, [&](pscbfm::Server& e) {
        require_counter_draws(staged->engine);
        const uint64_t id = e.submit(staged->engine.staged(), n_mcs);
        if (job) *job = id;
    });
}

int pscbfm_server_wait(pscbfm_server* s, uint64_t job, double timeout_s, int* finished) {
    return server_guarded(s, [&](pscbfm::Server& e) {
        const bool done = e.wait(job, timeout_s);
        if (finished) *finished = done ? 1 : 0;
    });
}

int pscbfm_server_job(pscbfm_server* s, uint64_t job, pscbfm_job_info* info) {
    if (!info) return PSCBFM_INVALID_ARGUMENT;
    return server_guarded(s, [&](pscbfm::Server& e) {
        const pscbfm::JobRecord r = e.record(job);
        info->state = int32_t(r.state);
        info->n_monomers = r.nMonomers;
        info->n_mcs = r.nMCS;
        info->accepted = r.accepted;
        info->device_bytes = r.footprint;
        info->t_submit = r.tSubmit;
        info->t_admit = r.tAdmit;
        info->t_done = r.tDone;
        if (r.state == pscbfm::JobState::Failed) throw std::runtime_error("job failed: " + r.error);
    });
}

int pscbfm_server_result(pscbfm_server* s, uint64_t job, int32_t* xyz, uint32_t n) {
    if (!xyz && n) return PSCBFM_INVALID_ARGUMENT;
    return server_guarded(s, [&](pscbfm::Server& e) {
        const pscbfm::JobRecord r = e.record(job);
        if (r.state != pscbfm::JobState::Done) throw std::runtime_error("job has no result (not done)");
        if (size_t(n) * 3 != r.xyz.size()) throw std::invalid_argument("count differs from the job's monomers");
        std::memcpy(xyz, r.xyz.data(), r.xyz.size() * sizeof(int32_t));
    });
}

int pscbfm_server_release(pscbfm_server* s, uint64_t job) {
    return server_guarded(s, [&](pscbfm::Server& e) { e.release(job); });
}

int pscbfm_server_drain(pscbfm_server* s) {
    return server_guarded(s, [&](pscbfm::Server& e) { e.drain(); });
}

int pscbfm_server_stats(pscbfm_server* s, pscbfm_server_counters* out) {
    if (!out) return PSCBFM_INVALID_ARGUMENT;
    return server_guarded(s, [&](pscbfm::Server& e) {
        const pscbfm::ServerStats st = e.stats();
        out->submitted = st.submitted;
        out->done = st.done;
        out->failed = st.failed;
        out->queued = st.queued;
        out->running = st.running;
        out->resident_bytes = st.residentBytes;
        out->peak_resident_bytes = st.peakResidentBytes;
        out->used_tiles = st.usedTiles;
        out->tile_capacity = st.tileCapacity;
        out->launches = st.launches;
        out->substeps = st.substeps;
        out->gpu_busy_seconds = st.gpuBusySeconds;
    });
}

int pscbfm_server_now(pscbfm_server* s, double* seconds) {
    if (!seconds) return PSCBFM_INVALID_ARGUMENT;
    return server_guarded(s, [&](pscbfm::Server& e) { *seconds = e.now(); });
}

const char* pscbfm_version(void) { return "pscbfm-mi355x 0.1 (gfx950)"; }

}  // extern "C"
