/*
 * UpdaterGPUScBFM_AB_Type.h - drop-in C++ class for the MI355X scBFM sweep.
 *
 * Same class name, method names, argument meaning and exception types as the
 * reference class (src/pscbfm/UpdaterGPUScBFM_AB_Type.h:124-362), header-only
 * on top of the C ABI in pscbfm.h (link with -lpscbfm).  A caller written
 * against the reference (e.g. GPUScBFM_AB_Type.h:78-196) compiles unchanged
 * except for the include path.  The per-substep species and seed draws
 * (CRITICAL.cu:509-510) come either from the caller's R250 state
 * (setR250State, the reference's stream) or from a counter draw keyed by
 * setSeed().
 *
 * Exceptions: std::invalid_argument, std::runtime_error and std::out_of_range
 * where the reference throws them; device failures, which make the reference
 * exit() (cudacommon.hpp:31-46), throw std::runtime_error instead.
 */
#ifndef PSCBFM_UPDATER_GPU_SCBFM_AB_TYPE_H
#define PSCBFM_UPDATER_GPU_SCBFM_AB_TYPE_H

#include <chrono>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pscbfm.h"

class UpdaterGPUScBFM_AB_Type {
public:
    /* UpdaterGPUScBFM_AB_Type.h:316 */
    UpdaterGPUScBFM_AB_Type() {
        if (pscbfm_create(&h_) != PSCBFM_OK) throw std::runtime_error(pscbfm_create_error());
    }
    /* UpdaterGPUScBFM_AB_Type.h:317 */
    virtual ~UpdaterGPUScBFM_AB_Type() { pscbfm_destroy(h_); }
    UpdaterGPUScBFM_AB_Type(const UpdaterGPUScBFM_AB_Type&) = delete;
    UpdaterGPUScBFM_AB_Type& operator=(const UpdaterGPUScBFM_AB_Type&) = delete;

    /* UpdaterGPUScBFM_AB_Type.h:318 */
    void destruct() { check(pscbfm_destruct(h_)); }

    /* UpdaterGPUScBFM_AB_Type.h:335-337 */
    void initialize() { check(pscbfm_initialize(h_)); }
    bool execute() { return pscbfm_execute(h_) == PSCBFM_OK; }
    void cleanup() { check(pscbfm_cleanup(h_)); }

    /* UpdaterGPUScBFM_AB_Type.h:340-346 */
    void setGpu(int iGpuToUse) { check(pscbfm_set_gpu(h_, iGpuToUse)); }
    void copyBondSet(int dx, int dy, int dz, bool bondForbidden) {
        check(pscbfm_copy_bond_set(h_, dx, dy, dz, bondForbidden ? 1 : 0));
    }
    void setNrOfAllMonomers(uint32_t nAllMonomers) { check(pscbfm_set_nr_of_all_monomers(h_, nAllMonomers)); }
    void setAttribute(uint32_t i, int32_t attribute) { check(pscbfm_set_attribute(h_, i, attribute)); }
    void setMonomerCoordinates(uint32_t i, int32_t x, int32_t y, int32_t z) {
        check(pscbfm_set_monomer_coordinates(h_, i, x, y, z));
    }
    void setConnectivity(uint32_t monoidx1, uint32_t monoidx2) {
        check(pscbfm_set_connectivity(h_, monoidx1, monoidx2));
    }
    void setLatticeSize(uint32_t boxX, uint32_t boxY, uint32_t boxZ) {
        check(pscbfm_set_lattice_size(h_, boxX, boxY, boxZ));
    }

    /* UpdaterGPUScBFM_AB_Type.h:351-358 */
    void populateLattice() { check(pscbfm_populate_lattice(h_)); }
    void runSimulationOnGPU(int32_t nrMCS_per_Call) {
        check(pscbfm_run(h_, nrMCS_per_Call, nullptr));
        mtCopyBack0 = std::chrono::high_resolution_clock::now();
    }
    int32_t getMonomerPositionInX(uint32_t i) { return position(i, 0); }
    int32_t getMonomerPositionInY(uint32_t i) { return position(i, 1); }
    int32_t getMonomerPositionInZ(uint32_t i) { return position(i, 2); }
    void setPeriodicity(bool isPeriodicX, bool isPeriodicY, bool isPeriodicZ) {
        check(pscbfm_set_periodicity(h_, isPeriodicX, isPeriodicY, isPeriodicZ));
    }

    /* UpdaterGPUScBFM_AB_Type.h:361 (for benchmarking purposes) */
    std::chrono::time_point<std::chrono::high_resolution_clock> mtCopyBack0;

    /* ---- extensions -------------------------------------------------------- */
    void checkSystem() { check(pscbfm_check_system(h_)); }
    void setSeed(uint64_t seed) { check(pscbfm_set_seed(h_, seed)); }
    void setStartMcs(uint64_t mcs) { check(pscbfm_set_start_mcs(h_, mcs)); }
    // R250 mode: the 250 state words in draw order, or a LeMonADE seed file
    // (state = words 7..256) followed by `discard` draws.
    void setR250State(const uint32_t* words250) { check(pscbfm_set_r250_state(h_, words250)); }
    void loadR250SeedFile(const std::string& path, uint32_t discard) {
        check(pscbfm_r250_from_seed_file(h_, path.c_str(), discard));
    }
    void getR250State(uint32_t* words250) { check(pscbfm_r250_get_state(h_, words250)); }
    void setLatticeKind(int kind) { check(pscbfm_set_lattice_kind(h_, kind)); }
    void setMonomerCoordinates(const std::vector<int32_t>& xyz) {
        check(pscbfm_set_monomer_coordinates_bulk(h_, xyz.data(), uint32_t(xyz.size() / 3)));
    }
    std::vector<int32_t> getMonomerPositions(uint32_t n) {
        std::vector<int32_t> out(3 * size_t(n));
        check(pscbfm_get_positions_bulk(h_, out.data(), n));
        return out;
    }
    pscbfm_run_stats run(int32_t nMCS, uint32_t* perSubstepAccepted = nullptr, bool profile = false,
                         bool useGraph = true) {
        pscbfm_run_stats s{};
        check(pscbfm_run_ex(h_, nMCS, perSubstepAccepted, profile ? 1 : 0, useGraph ? 1 : 0, &s));
        return s;
    }
    pscbfm_updater* handle() { return h_; }

private:
    int32_t position(uint32_t i, int axis) {
        int32_t v[3];
        check(pscbfm_get_monomer_position(h_, i, &v[0], &v[1], &v[2]));
        return v[axis];
    }
    void check(int code) {
        switch (code) {
            case PSCBFM_OK: return;
            case PSCBFM_INVALID_ARGUMENT: throw std::invalid_argument(pscbfm_last_error(h_));
            case PSCBFM_OUT_OF_RANGE: throw std::out_of_range(pscbfm_last_error(h_));
            default: throw std::runtime_error(pscbfm_last_error(h_));
        }
    }
    pscbfm_updater* h_ = nullptr;
};

#endif
