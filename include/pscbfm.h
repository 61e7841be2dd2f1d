/*
 * pscbfm.h - C ABI of the MI355X (gfx950) scBFM sweep.
 *
 * One handle is one UpdaterGPUScBFM_AB_Type instance of the reference
 * (src/pscbfm/UpdaterGPUScBFM_AB_Type.h:124-362).  Every entry point returns
 * PSCBFM_OK or an error code; the message of the last error of a handle is
 * pscbfm_last_error(handle).  Nothing throws across this boundary and nothing
 * calls exit(): the reference's C++ exceptions and its exit-on-CUDA-error
 * (src/pscbfm/cudacommon.hpp:31-46) both become status codes.
 *
 * Coordinates are unfolded int32 corners of the 2x2x2 monomer cubes, given
 * and returned in the caller's monomer order.  Handles are not thread-safe;
 * separate handles are independent (no process-global device tables).
 */
#ifndef PSCBFM_H
#define PSCBFM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pscbfm_updater pscbfm_updater;

enum pscbfm_status {
    PSCBFM_OK = 0,
    PSCBFM_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    PSCBFM_RUNTIME_ERROR = 2,    /* std::runtime_error in the reference */
    PSCBFM_OUT_OF_RANGE = 3,     /* std::out_of_range (vector::at) in the reference */
    PSCBFM_DEVICE_ERROR = 4,     /* HIP runtime failure (the reference exits) */
    PSCBFM_NULL_HANDLE = 5,
    PSCBFM_INTERNAL_ERROR = 6
};

enum pscbfm_lattice_kind {
    PSCBFM_LATTICE_U8 = 0,  /* one byte per cell */
    PSCBFM_LATTICE_BIT = 1  /* one bit per cell (default) */
};

/* Constructor UpdaterGPUScBFM_AB_Type() (UpdaterGPUScBFM_AB_Type.h:316). */
int pscbfm_create(pscbfm_updater** out);
/* Destructor ~UpdaterGPUScBFM_AB_Type() (UpdaterGPUScBFM_AB_Type.h:317). */
void pscbfm_destroy(pscbfm_updater* h);
/* Message of the last failed call on h (empty if none); valid until the next call. */
const char* pscbfm_last_error(const pscbfm_updater* h);
/* Message of the last failed pscbfm_create in this thread. */
const char* pscbfm_create_error(void);

/* destruct() (UpdaterGPUScBFM_AB_Type.h:318). */
int pscbfm_destruct(pscbfm_updater* h);
/* setGpu(int) (UpdaterGPUScBFM_AB_Type.h:340). */
int pscbfm_set_gpu(pscbfm_updater* h, int gpu);
/* copyBondSet(dx,dy,dz,forbidden) (UpdaterGPUScBFM_AB_Type.h:341). */
int pscbfm_copy_bond_set(pscbfm_updater* h, int dx, int dy, int dz, int forbidden);
/* setNrOfAllMonomers(n) (UpdaterGPUScBFM_AB_Type.h:342). */
int pscbfm_set_nr_of_all_monomers(pscbfm_updater* h, uint32_t n);
/* setAttribute(i, a) (UpdaterGPUScBFM_AB_Type.h:343). */
int pscbfm_set_attribute(pscbfm_updater* h, uint32_t i, int32_t attribute);
/* setMonomerCoordinates(i,x,y,z) (UpdaterGPUScBFM_AB_Type.h:344). */
int pscbfm_set_monomer_coordinates(pscbfm_updater* h, uint32_t i, int32_t x, int32_t y, int32_t z);
/* Bulk form of setMonomerCoordinates: xyz holds 3*n values. */
int pscbfm_set_monomer_coordinates_bulk(pscbfm_updater* h, const int32_t* xyz, uint32_t n);
/* setConnectivity(i, j), directed (UpdaterGPUScBFM_AB_Type.h:345). */
int pscbfm_set_connectivity(pscbfm_updater* h, uint32_t i, uint32_t j);
/* Bulk form of setConnectivity: pairs holds 2*n_pairs ids, each pair directed. */
int pscbfm_set_connectivity_bulk(pscbfm_updater* h, const uint32_t* pairs, uint64_t n_pairs);
/* setLatticeSize(x,y,z) (UpdaterGPUScBFM_AB_Type.h:346). */
int pscbfm_set_lattice_size(pscbfm_updater* h, uint32_t x, uint32_t y, uint32_t z);
/* setPeriodicity(x,y,z) (UpdaterGPUScBFM_AB_Type.h:358). */
int pscbfm_set_periodicity(pscbfm_updater* h, int x, int y, int z);
/* populateLattice() (UpdaterGPUScBFM_AB_Type.h:351). */
int pscbfm_populate_lattice(pscbfm_updater* h);
/* initialize() (UpdaterGPUScBFM_AB_Type.h:335). */
int pscbfm_initialize(pscbfm_updater* h);
/* execute() (UpdaterGPUScBFM_AB_Type.h:336): a no-op that succeeds. */
int pscbfm_execute(pscbfm_updater* h);
/* runSimulationOnGPU(nMCS) (UpdaterGPUScBFM_AB_Type.h:352; CRITICAL.cu:414-698).
 * Synchronous.  accepted_out (may be NULL) receives the accepted moves of this call. */
int pscbfm_run(pscbfm_updater* h, int32_t n_mcs, uint64_t* accepted_out);
/* getMonomerPositionInX/Y/Z(i) (UpdaterGPUScBFM_AB_Type.h:354-356). */
int pscbfm_get_monomer_position(const pscbfm_updater* h, uint32_t i, int32_t* x, int32_t* y, int32_t* z);
/* Bulk form of the getters: xyz receives 3*n values in monomer order. */
int pscbfm_get_positions_bulk(const pscbfm_updater* h, int32_t* xyz, uint32_t n);
/* checkSystem() (UpdaterGPUScBFM_AB_Type.h:313, private there): excluded
 * volume and bond-set invariants of the current coordinates. */
int pscbfm_check_system(const pscbfm_updater* h);
/* cleanup() (UpdaterGPUScBFM_AB_Type.h:337). */
int pscbfm_cleanup(pscbfm_updater* h);

/* ---- Extensions (no single reference counterpart) -------------------------- */

/* Key of the counter-based substep draw; replaces the seeding of LeMonADE's
 * R250 stream that drives CRITICAL.cu:509-510.  Before initialize. */
int pscbfm_set_seed(pscbfm_updater* h, uint64_t seed);
/* Age of the input state in sweeps (a BFM frame's !mcs), before initialize.
 * The substep counter, and with it the counter draw, starts at
 * mcs * nColors.  A run resumed from a checkpoint of age m then continues
 * the uninterrupted trajectory bit for bit; the reference cannot, since its
 * R250 state is not saved with the snapshot (mainCUDASimulatorGPUScBFM_AB_Type.cpp:155,160).
 * Batches and the request loop start a staged system's replica there too. */
int pscbfm_set_start_mcs(pscbfm_updater* h, uint64_t mcs);
/* R250 mode: the per-substep species and kernel seed come from the R250
 * stream the reference draws them from (CRITICAL.cu:509-510: species =
 * r250 % nSpecies, then seed = r250), instead of the counter draw.  Any time;
 * takes effect from the next run.  Batches and request queues reject an
 * updater in R250 mode.
 * set_r250_state: the 250 state words in draw order (RandomNumberGenerators
 *   state as LeMonADE keeps it).
 * r250_from_seed_file: RandomNumberGenerators::seedAll(file) as
 *   mainCUDASimulatorGPUScBFM_AB_Type.cpp:129 calls it (state = words 7..256
 *   of the file), then `discard` draws; the reference simulator discards 2
 *   (its known-answer prints, mainCUDASimulatorGPUScBFM_AB_Type.cpp:138-139).
 * r250_next: draws n raw outputs (out may be null to skip them).
 * r250_get_state: the current 250 words in draw order (set_r250_state of
 *   them continues the same stream). */
int pscbfm_set_r250_state(pscbfm_updater* h, const uint32_t* words);
int pscbfm_r250_from_seed_file(pscbfm_updater* h, const char* path, uint32_t discard);
int pscbfm_r250_next(pscbfm_updater* h, uint32_t* out, uint64_t n);
int pscbfm_r250_get_state(const pscbfm_updater* h, uint32_t* words);
/* R250 mode: substeps of draws kept on the device at once (0 = 2^22, 32 MiB,
 * also the maximum).  A longer run goes through in chunks of whole sweeps,
 * with the same result. */
int pscbfm_set_draw_chunk(pscbfm_updater* h, uint64_t substeps);
/* Whether pscbfm_run / pscbfm_run_ex copy the positions back before they
 * return (on = 1, the default, as runSimulationOnGPU does at
 * CRITICAL.cu:673-695).  Off: the getters return the coordinates of the last
 * read-back until pscbfm_read_positions_back, so that repeated runs whose
 * coordinates are not needed in between skip the copy. */
int pscbfm_set_read_back(pscbfm_updater* h, int on);
int pscbfm_read_positions_back(pscbfm_updater* h);
/* Occupancy lattice storage (pscbfm_lattice_kind).  Before initialize.
 * Replaces the USE_BIT_PACKING_LATTICE switch (UpdaterGPUScBFM_AB_Type.cu:13). */
int pscbfm_set_lattice_kind(pscbfm_updater* h, int kind);
/* Uniform color rebalancing on/off (UpdaterGPUScBFM_AB_Type.cu:550, default on). */
int pscbfm_set_uniform_coloring(pscbfm_updater* h, int on);
/* Run checkSystem inside initialize (default on). */
int pscbfm_set_check_on_initialize(pscbfm_updater* h, int on);
/* Replace the automatic coloring with a proper coloring given per monomer. */
int pscbfm_set_colors(pscbfm_updater* h, const uint8_t* colors, uint32_t n);
/* Threads per workgroup (64, 128, 192 or 256; 0 = 256) and monomers per
 * pass-1 thread (1, 2 or 4; 0 = 2).  Results do not depend on the shape. */
int pscbfm_set_launch_shape(pscbfm_updater* h, uint32_t block, uint32_t slots);
/* Pass-2 threads per workgroup (64..256, multiple of 64; 0 = automatic:
 * 128 when at least a quarter of the cells are occupied, else the pass-1
 * block).  Results do not depend on it. */
int pscbfm_set_perform_block(pscbfm_updater* h, uint32_t block);
/* Pass-2 walk of a single system: threads_per_list threads walk one pass-1
 * list (64..256, multiple of 64), so a pass-2 workgroup walks several lists,
 * with entries_per_thread entries in flight per thread (1 or 4; 0 = 4);
 * threads_per_list = 0 or 0xFFFFFFFF (the default) walks one list per
 * workgroup.  Results do not depend on it. */
int pscbfm_set_perform_walk(pscbfm_updater* h, uint32_t threads_per_list, uint32_t entries_per_thread);
/* Slot order inside each species, before initialize: 1 places the members
 * of a species in the Z-curve order of their starting corners, 0 in
 * monomer-index order as the reference (UpdaterGPUScBFM_AB_Type.cu:601-607,
 * iToiNew), 2 (the default) chooses: 1 on the bit lattice, and on the byte
 * lattice 1 only when the monomer order scatters bonded partners (species-
 * local indices more than 32766 apart for at least 1% of the bonds).  Each
 * move is drawn from the member's index in the reference's order either way
 * (CRITICAL.cu:322-323), so results do not depend on it.  Other values:
 * PSCBFM_INVALID_ARGUMENT. */
int pscbfm_set_slot_order(pscbfm_updater* h, int order);
/* The pass of each substep (1 or 2; 0 = automatic: pass 1 on the bit
 * lattice with at most two bonds per monomer, else pass 2) that zeroes the
 * other parity's proposal buffer, the reference's separate clear
 * (CRITICAL.cu:561-573).  Results do not depend on it. */
int pscbfm_set_clear_pass(pscbfm_updater* h, uint32_t pass);
/* Substeps per captured graph (0 = automatic, at most 4096:
 * PSCBFM_INVALID_ARGUMENT above). */
int pscbfm_set_graph_chunk(pscbfm_updater* h, uint32_t substeps);
/* After initialize: captures, instantiates and uploads every graph a run of
 * n_mcs sweeps replays, so that such a run starts its device work at once.
 * No counterpart in the reference, which launches every substep from the
 * host (CRITICAL.cu:506-583).  A run builds what is missing itself, before
 * its GPU timing starts; this moves that host work out of the run call. */
int pscbfm_prepare_graphs(pscbfm_updater* h, int32_t n_mcs);
/* Launch-shape autotuner.  Counterpart of the reference's in-run thread-count
 * search (CRITICAL.cu:445-486,600-648), which is compiled out unless
 * AUTO_CONFIGURE_BEST_SETTINGS_FOR_PSCBFM_ALGORITHM is defined (then fixed at
 * 128 threads) and which tunes on the live trajectory.  After initialize: times `mcs`
 * sweeps of a scratch copy of the system for each of the 12 shapes
 * (block 64/128/192/256 x slots 1/2/4) and applies the fastest; the system's
 * own state is not touched.  Cached per workload signature in the process.
 * trial_ms (optional, 12 entries) receives ms per sweep in candidate order. */
int pscbfm_autotune(pscbfm_updater* h, uint32_t mcs, uint32_t* block, uint32_t* slots, double* trial_ms);

/* Statistics of the last run. */
typedef struct pscbfm_run_stats {
    double gpu_ms;            /* hipEvent time of all launched substeps */
    uint64_t substeps;
    uint64_t attempted;       /* sum of species sizes over the substeps */
    uint64_t accepted;
    uint64_t bond_terms;      /* sum over substeps of the drawn species' bond count */
    double check_ms;          /* pass-1 kernel time (profiled runs only) */
    double perform_ms;        /* pass-2 kernel time (profiled runs only) */
    uint64_t check_launches;
    uint64_t perform_launches;
} pscbfm_run_stats;

/* Run with options: log (may be NULL) receives n_mcs*n_colors per-substep
 * accepted counts; profile!=0 times both kernels of every substep with HIP
 * events (eager launches); use_graph selects hipGraph replay. */
int pscbfm_run_ex(pscbfm_updater* h, int32_t n_mcs, uint32_t* log, int profile, int use_graph,
                  pscbfm_run_stats* stats);
int pscbfm_last_run_stats(const pscbfm_updater* h, pscbfm_run_stats* stats);
/* pscbfm_run_ex(h, n_mcs, log, 0, 0, stats) without any device allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The call's device log and, in R250 mode, its draw table
 * live in a static scratch array of 3 MiB that each GPU the library runs on
 * holds until the process ends, shared by every such call on that GPU: calls
 * that overlap on one GPU race.  With a log or in R250 mode a call covers at
 * most 262144 substeps (n_mcs * n_colors); beyond that it returns
 * PSCBFM_INVALID_ARGUMENT.  Launches are eager; stats->gpu_ms is host time
 * from the first launch to completion.  Results equal pscbfm_run_ex's. */
int pscbfm_run_ex_noalloc(pscbfm_updater* h, int32_t n_mcs, uint32_t* log, pscbfm_run_stats* stats);

/* Layout introspection: number of species, per-monomer species, species sizes. */
int pscbfm_get_n_colors(const pscbfm_updater* h, uint32_t* n_colors);
int pscbfm_get_colors(const pscbfm_updater* h, uint8_t* colors, uint32_t n);
int pscbfm_get_color_sizes(const pscbfm_updater* h, uint32_t* sizes, uint32_t n_colors);
/* Substep counter (substeps done since initialize) and accepted moves since initialize. */
int pscbfm_get_counters(const pscbfm_updater* h, uint64_t* substeps, uint64_t* accepted);
/* Device bytes held by the handle. */
int pscbfm_get_device_bytes(const pscbfm_updater* h, uint64_t* bytes);
/* Device occupancy lattice as one byte per cell in Z-curve order (cells bytes). */
int pscbfm_download_lattice(const pscbfm_updater* h, uint8_t* cells, uint64_t n_cells);
/* Bonded partners of the species layout (directed pairs) and how many of
 * them have no 16-bit partner code, so that the uint8-lattice pass 1 reads the
 * full neighbour matrix for them.  Staged or initialized handle; no device work. */
int pscbfm_count_partner_escapes(const pscbfm_updater* h, uint64_t* partners, uint64_t* escapes);
/* Number of set bits in both proposal lattices.  After substeps 0..T-1 the
 * buffer of parity T & 1, which substep T marks next, is empty, and the other
 * holds the marks of substep T-1. */
int pscbfm_count_proposal_bits(const pscbfm_updater* h, uint64_t* bits_even, uint64_t* bits_odd);

/* Host helpers shared with the kernels (no device work). */
/* Z-curve code of wrapped coordinates for a box (bit layout of the device lattice). */
int pscbfm_zcurve_encode(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t x, uint32_t y, uint32_t z,
                         uint32_t* code);
/* Move hash of CRITICAL.cu:2-11. */
uint32_t pscbfm_move_hash(uint32_t a);
/* Species and kernel seed of substep t (replaces CRITICAL.cu:509-510). */
void pscbfm_substep_draw(uint64_t seed, uint64_t t, uint32_t n_colors, uint32_t* species, uint32_t* kernel_seed);
/* Graph coloring of graphColoring.tpp:77-293 over a CSR adjacency. */
int pscbfm_color_graph(uint32_t n, const uint32_t* start, const uint32_t* ids, int uniform, uint8_t* colors);
/* The 108 scBFM bond vectors (3 ints each) in a fixed order. */
int pscbfm_standard_bond_vectors(int32_t* xyz108);

/* ---- BFM files (LeMonADE's format, the reference's I/O) --------------------- */

/* Reads the last configuration of a BFM file (plain or gzip) and stages it
 * with the call sequence of GPUScBFM_AB_Type.h:83-130: periodicity, monomer
 * count, coordinates, attributes, both directions of every bond, box,
 * populateLattice, and the file's !set_of_bondvectors as the allowed bonds
 * (copyBondSet over [-4,3]^3).  *mcs (optional) receives the frame's age. */
int pscbfm_bfm_stage(pscbfm_updater* h, const char* path, uint64_t* mcs);
/* Writes the updater's current coordinates as a frame with age `mcs`.
 * append = 0: a new file (header with topology, then the frame); append != 0:
 * appends the frame, writing the header first if the file is empty or
 * missing (as AnalyzerWriteBfmFile appends every save interval). */
int pscbfm_bfm_write(const pscbfm_updater* h, const char* path, uint64_t mcs, int append);

/* ---- Synthetic systems (host only, seeded, deterministic) ------------------ */

/* Linear chains grown as self-avoiding walks with bonds drawn uniformly from
 * the 108-set; chain c is monomers [c*len, (c+1)*len), bonded consecutively.
 * xyz receives 3*n_chains*len unfolded corners.  PSCBFM_RUNTIME_ERROR if the
 * growth needed more than max_restarts chain restarts (box too dense). */
int pscbfm_gen_linear_chains(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_chains, uint32_t len,
                             uint64_t seed, uint32_t max_restarts, int32_t* xyz);
/* Linear chains on an ordered serpentine packing (corners on (2i,2j,3k));
 * supports occupied-site fractions up to 2*floor(lz/3)/lz. */
int pscbfm_gen_linear_chains_ordered(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_chains, uint32_t len,
                                     int32_t* xyz);
/* Star network: n_stars stars of `arms` arms of arm_length monomers (center
 * first), arm ends crosslinked to other stars' arm interiors.  xyz receives
 * 3*n_stars*(1+arms*arm_length) corners; bonds receives undirected pairs
 * (2 ids each, capacity in pairs: n_stars*arms*(arm_length+1)). */
int pscbfm_gen_star_network(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_stars, uint32_t arms,
                            uint32_t arm_length, uint64_t seed, int32_t* xyz, uint32_t* bonds,
                            uint64_t bond_capacity, uint64_t* n_bonds);
/* Four-arm stars on a dense ordered packing, for occupied fractions that
 * growth cannot reach: n_stars stars of arm_length monomers (center first,
 * then the arms) with corners on the sites of even coordinates, n_stars of
 * the box's star places chosen at random by seed.  Up to 91% of the sites
 * hold a monomer (arm 29 on 256^3: 16,384 stars).  xyz receives
 * 3*n_stars*(1+4*arm_length) corners, bonds the 2*n_stars*4*arm_length ids
 * of the bonds inside the stars; PSCBFM_RUNTIME_ERROR if the box holds fewer
 * stars.  No crosslinks: every bond between two such sites joins sites of
 * opposite parity, so a network made there has two species; crosslinks
 * added after a run (pscbfm_crosslink_star_ends) close odd cycles. */
int pscbfm_gen_stars_ordered(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_stars, uint32_t arm_length,
                             uint64_t seed, int32_t* xyz, uint32_t* bonds);
/* Crosslinks for stars at given corners (n_stars stars of 1 + arms *
 * arm_length monomers, center first, arm by arm; unfolded coordinates, as
 * after a run).  In each of `rounds` rounds every arm end is bonded to the
 * first monomer, in a random order of the 108 bond vectors, of another
 * star's arm interior that has no crosslink of that round yet and is not
 * already its partner.  bonds receives only the crosslink pairs (capacity in
 * pairs: n_stars * arms * rounds). */
int pscbfm_crosslink_star_ends(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t n_stars, uint32_t arms,
                               uint32_t arm_length, uint32_t rounds, uint64_t seed, const int32_t* xyz,
                               uint32_t* bonds, uint64_t bond_capacity, uint64_t* n_bonds);

/* Streaming-copy reference on `gpu`: copies `bytes` (read + write) `repeats`
 * times with a plain 16-byte-per-lane kernel and returns the mean time.  Used
 * to scale profiler byte counters and as the achievable-bandwidth reference. */
int pscbfm_copy_probe(int gpu, uint64_t bytes, int repeats, double* ms_per_copy);

/* ---- Batched replicas (continuous batching) ----------------------------------
 * No single counterpart in the reference: it replaces a host loop of
 * independent UpdaterGPUScBFM_AB_Type instances, each running its own
 * runSimulationOnGPU (UpdaterGPUScBFM_AB_Type.cu:1030-1122), by one launch per
 * pass that covers every live replica.  A replica is the system staged in an
 * updater handle (setLatticeSize/.../setConnectivity and setSeed, no
 * initialize needed); each replica keeps its own seed, substep counter and
 * stop point, so its trajectory is bit-identical to the same system run
 * alone in an updater.  Replicas may be added and removed between launches. */
typedef struct pscbfm_batch pscbfm_batch;

typedef struct {
    int32_t live;
    uint32_t n_colors;
    uint32_t n_monomers;
    uint32_t tiles;         /* workgroups per launch */
    uint64_t substeps;      /* done, as of the last synchronize */
    uint64_t stop_at;       /* the replica idles from this substep on */
    uint64_t accepted;      /* accepted moves since it joined */
    uint64_t device_bytes;  /* its own device buffers */
} pscbfm_replica_info;

/* tile_capacity: workgroups per launch (rounded up to a multiple of 8), the
 * sum of the replicas' tiles may not exceed it; 0 = a table that grows as
 * replicas are added (the captured graph is then rebuilt); replica_capacity: the most
 * replicas at once; block (64..256, multiple of 64) and slots (1, 2, 4) as in
 * pscbfm_set_launch_shape; lattice_kind: PSCBFM_LATTICE_*. */
int pscbfm_batch_create(int gpu, int lattice_kind, uint32_t tile_capacity, uint32_t replica_capacity,
                        uint32_t block, uint32_t slots, pscbfm_batch** out);
void pscbfm_batch_destroy(pscbfm_batch* b);
const char* pscbfm_batch_last_error(const pscbfm_batch* b);
/* Device bytes and workgroups the system staged in `staged` would take. */
int pscbfm_batch_footprint(pscbfm_batch* b, const pscbfm_updater* staged, uint64_t* bytes, uint32_t* tiles);
/* Adds the system staged in `staged` (with its seed, coloring settings and
 * bond set); the replica idles until a target is set.  PSCBFM_RUNTIME_ERROR
 * when the tile or replica capacity is exhausted. */
int pscbfm_batch_add(pscbfm_batch* b, const pscbfm_updater* staged, uint32_t* id);
int pscbfm_batch_remove(pscbfm_batch* b, uint32_t id);
/* Pass-2 threads per workgroup of the batch (0 = the default: the pass-1
 * block, at most 192). */
int pscbfm_batch_set_perform_block(pscbfm_batch* b, uint32_t block);
/* The replica stops after n_mcs more sweeps (counted from the last synchronize). */
int pscbfm_batch_set_target(pscbfm_batch* b, uint32_t id, uint64_t n_mcs);
/* Enqueues `substeps` substeps of every live replica (returns at once). */
int pscbfm_batch_launch(pscbfm_batch* b, uint64_t substeps);
/* Waits for queued work and refreshes every replica's counters. */
int pscbfm_batch_synchronize(pscbfm_batch* b);
/* Sets every live replica's target to n_mcs more sweeps and runs to it;
 * gpu_ms (optional) receives the GPU time of the launches. */
int pscbfm_batch_run(pscbfm_batch* b, uint64_t n_mcs, double* gpu_ms);
int pscbfm_batch_info(pscbfm_batch* b, uint32_t id, pscbfm_replica_info* info);
/* Unfolded coordinates of replica `id` in its input order (3*n ints). */
int pscbfm_batch_get_positions(pscbfm_batch* b, uint32_t id, int32_t* xyz, uint32_t n);
/* pscbfm_batch_get_positions for `count` replicas at once: xyz[k] receives
 * 3*n[k] ints of replica ids[k].  The coordinates are decoded on the device
 * and copied back in rounds of up to 64 MiB.  No counterpart in the
 * reference (it reads one system back per run, CRITICAL.cu:686-695). */
int pscbfm_batch_get_positions_many(pscbfm_batch* b, const uint32_t* ids, uint32_t count, int32_t* const* xyz,
                                    const uint32_t* n);
/* Autotuner for a batch: times `mcs` sweeps of a scratch batch of the staged
 * systems for each candidate shape; returns the fastest (to pass to
 * pscbfm_batch_create) and, optionally, ms per sweep of all 12 candidates. */
int pscbfm_batch_autotune(int gpu, int lattice_kind, const pscbfm_updater* const* staged, uint32_t n,
                          uint32_t mcs, uint32_t* block, uint32_t* slots, double* trial_ms);
int pscbfm_batch_usage(pscbfm_batch* b, uint32_t* live, uint32_t* used_tiles, uint32_t* tile_capacity,
                       uint64_t* device_bytes);

/* ---- Request loop (continuous batching) -------------------------------------
 * No counterpart in the reference.  Jobs (a staged updater + a number of
 * sweeps) are submitted from any thread; four preparation threads build their
 * layouts, and a GPU thread admits them into one batch as HBM budget,
 * workgroups and replica slots allow, launches the live replicas at most
 * `quantum` substeps at a time (so arrivals join while others run), and
 * retires finished ones.  Each job's result is the one it has alone. */
typedef struct pscbfm_server pscbfm_server;

typedef struct {
    int32_t state;            /* 0 queued, 1 prepared, 2 running, 3 done, 4 failed */
    uint32_t n_monomers;
    uint64_t n_mcs;
    uint64_t accepted;
    uint64_t device_bytes;    /* footprint while resident */
    double t_submit, t_admit, t_done;  /* seconds since the server started */
} pscbfm_job_info;

typedef struct {
    uint64_t submitted, done, failed;
    uint32_t queued, running;
    uint64_t resident_bytes, peak_resident_bytes;
    uint32_t used_tiles, tile_capacity;
    uint64_t launches, substeps;
    double gpu_busy_seconds;
} pscbfm_server_counters;

/* hbm_budget: device bytes the resident replicas may take together;
 * tile_capacity (> 0), replica_capacity, block, slots as for pscbfm_batch_create. */
int pscbfm_server_create(int gpu, int lattice_kind, uint64_t hbm_budget, uint32_t tile_capacity,
                         uint32_t replica_capacity, uint32_t block, uint32_t slots, uint32_t quantum,
                         pscbfm_server** out);
void pscbfm_server_destroy(pscbfm_server* s);
const char* pscbfm_server_last_error(const pscbfm_server* s);
/* Queues the system staged in `staged` (its seed included) for n_mcs sweeps. */
int pscbfm_server_submit(pscbfm_server* s, const pscbfm_updater* staged, uint64_t n_mcs, uint64_t* job);
/* Waits for the job (timeout_s < 0: forever); *finished = 0 on timeout. */
int pscbfm_server_wait(pscbfm_server* s, uint64_t job, double timeout_s, int* finished);
int pscbfm_server_job(pscbfm_server* s, uint64_t job, pscbfm_job_info* info);
/* Final coordinates of a finished job (3*n ints, input order). */
int pscbfm_server_result(pscbfm_server* s, uint64_t job, int32_t* xyz, uint32_t n);
int pscbfm_server_release(pscbfm_server* s, uint64_t job);
int pscbfm_server_drain(pscbfm_server* s);
int pscbfm_server_stats(pscbfm_server* s, pscbfm_server_counters* stats);
int pscbfm_server_now(pscbfm_server* s, double* seconds);

/* Library version string. */
const char* pscbfm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PSCBFM_H */
