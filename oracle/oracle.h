/*
 * oracle.h - CPU restatement of the pscbfm colored two-pass scBFM sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker that the GPU path is
 * compared against, and the reported CPU baseline.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it; the library under
 * pscbfm_amd/ never calls it.
 *
 * Semantics follow the reference as cited per function; where the reference
 * depends on LeMonADE's R250 (absent), the counter-based draw of
 * oracle_substep_draw takes its place.
 *
 * Parity pins (what each function is checked against, see tests/):
 *   oracle_hash            golden vectors evaluated from the constants and
 *                          shifts of CRITICAL.cu:2-11 (tests/golden)
 *   oracle_splitmix64      published SplitMix64 known answers
 *   oracle_substep_draw    the SplitMix64 pins plus the library's own draw on
 *                          a grid of (seed, t, nColors)
 *   oracle_r250_*          the reference's known answers 4223731124 and
 *                          0.803876 for tests/resultNormPscBFM.seeds
 *                          (mainCUDASimulatorGPUScBFM_AB_Type.cpp:138-139),
 *                          and the stream recurrence on 2000 outputs; any
 *                          adjustment LeMonADE may make to other loaded
 *                          words is parity unpinned
 *   oracle_direction       the direction table of UpdaterGPUScBFM_AB_Type.cu:534-545
 *                          read from the reference text (tests/golden)
 *   oracle_standard_bondset  the 108 vectors of tests/LinearChain_l1024_n64.bfm
 *   oracle_bond_index      brute-force &7 aliasing identities
 *   oracle_zcurve_index    brute-force bit interleave
 *   oracle_front_blocked   brute-force 2x2x2 cube overlap on small boxes
 *   oracle_run             the naive O(n^2) definition (oracle_run_naive),
 *                          order independence (shuffled processing order),
 *                          excluded-volume/bond invariants, exact single
 *                          monomer and dimer acceptance; explicit draw lists
 *                          equal to the counter draw give the counter run
 *   oracle_check_system    planted overlaps and forbidden bonds on tiny systems
 *   oracle_sequential_run  statistics only (acceptance, bond lengths)
 */
#ifndef PSCBFM_ORACLE_H
#define PSCBFM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CRITICAL.cu:2-11 */
uint32_t oracle_hash(uint32_t a);
/* t-th (0-based) output of SplitMix64 seeded with `seed`. */
uint64_t oracle_splitmix64(uint64_t seed, uint64_t t);
/* Species and kernel seed of substep t: species = high word % n_colors,
 * seed = low word of oracle_splitmix64(seed, t).  Stands in for the two
 * R250 draws at CRITICAL.cu:509-510. */
void oracle_substep_draw(uint64_t seed, uint64_t t, uint32_t n_colors, uint32_t* species, uint32_t* kernel_seed);
/* CRITICAL.cu:243-261 */
uint32_t oracle_bond_index(int32_t dx, int32_t dy, int32_t dz);
/* Bond set of the scBFM: all sign/permutation images of (2,0,0), (2,1,0),
 * (2,1,1), (2,2,1), (3,0,0), (3,1,0).  Fills forbidden[512] (1 = forbidden)
 * over dx,dy,dz in [-4,3] as GPUScBFM_AB_Type.h:120-127 does; returns the
 * number of allowed entries (108). */
uint32_t oracle_standard_bondset(uint8_t forbidden[512]);
/* UpdaterGPUScBFM_AB_Type.cu:150-176 (11 bits, bit i -> bit 3i). */
uint32_t oracle_dilute_bits(uint32_t x);
/* UpdaterGPUScBFM_AB_Type.cu:210-220 with the box masks applied. */
uint32_t oracle_zcurve_index(uint32_t x, uint32_t y, uint32_t z, uint32_t lx, uint32_t ly, uint32_t lz);
/* Direction table of UpdaterGPUScBFM_AB_Type.cu:534-545. */
void oracle_direction(uint32_t d, int32_t* dx, int32_t* dy, int32_t* dz);

typedef struct oracle_system {
    uint32_t box[3];            /* powers of two, >= 8 */
    uint32_t n;                 /* monomers */
    int32_t* xyz;               /* 3n unfolded corners, updated in place */
    const uint32_t* nbr_start;  /* n+1 CSR offsets */
    const uint32_t* nbr;        /* bond partners (monomer ids) */
    const uint8_t* color;       /* species per monomer, proper coloring */
    uint32_t n_colors;
    const uint8_t* forbidden;   /* 512 entries, index oracle_bond_index */
    uint64_t seed;              /* key of the substep draw */
    const uint32_t* draws;      /* optional: 2 words per substep (species, kernel seed) from substep draw_t0 on */
    uint64_t draw_t0;
} oracle_system;

/* R250 as the reference uses it for the per-substep draws
 * (CRITICAL.cu:509-510, LeMonADE's RandomNumberGenerators, absent here),
 * restated from the published generator (Kirkpatrick and Stoll 1981) in
 * the 250-word ring form of Maier's r250.c: with the loaded words as
 * x[0..249], the stream is x[n] = x[n-250] XOR x[n-147] (the trinomial
 * x^250 + x^147 + 1, reciprocal of the 103-lag form).  The seed-file
 * layout (state = words 7..256 of the file, word 0 = srand seed) is pinned by
 * the known answers of mainCUDASimulatorGPUScBFM_AB_Type.cpp:137-139 with
 * tests/resultNormPscBFM.seeds.  Whether LeMonADE adjusts the loaded state
 * beyond that (outputs 2 and later) is parity unpinned. */
typedef struct oracle_r250 {
    uint32_t buf[250];
    uint32_t index;
} oracle_r250;
void oracle_r250_load(oracle_r250* r, const uint32_t state[250]);
uint32_t oracle_r250_rand32(oracle_r250* r);

/* Excluded volume of front plane (CRITICAL.cu:62-176): nonzero if any of
 * the nine corners p + 2e_d + a e_u + b e_v (a,b in {-1,0,1}) is occupied in
 * `occ` (one byte per cell, index x + Lx*(y + Ly*z)). */
int oracle_front_blocked(const uint8_t* occ, const uint32_t box[3], const int32_t p[3], uint32_t d);

/* Runs substeps t0 .. t0+n_substeps-1 of the colored two-pass scheme
 * (CRITICAL.cu:298-411 and :501-583).  accepted_per_substep (may be NULL)
 * receives one count per substep.  n_threads <= 0 uses all OpenMP threads.
 * shuffle_seed != 0 processes each pass in a shuffled order (the result must
 * not change).  Returns the total accepted moves, or -1 on bad input. */
int64_t oracle_run(oracle_system* sys, uint64_t t0, uint64_t n_substeps, uint32_t* accepted_per_substep,
                   int n_threads, uint64_t shuffle_seed);

/* Same substeps with pass 1 as a brute-force cube-overlap test against every
 * other monomer and pass 2 as a pairwise overlap test of all proposals:
 * O(n^2) per substep, small systems only. */
int64_t oracle_run_naive(oracle_system* sys, uint64_t t0, uint64_t n_substeps, uint32_t* accepted_per_substep);

/* 0 if no two cubes overlap and every bond is in the set and |component|<=3
 * (UpdaterGPUScBFM_AB_Type.cu:882-986); 1 for an overlap, 2 for a bad bond. */
int oracle_check_system(const oracle_system* sys);

/* LeMonADE-style sequential Metropolis sweep (random monomer, random
 * direction, immediate apply), n_mcs * n attempts, SplitMix64 stream.
 * Statistical baseline only.  Returns accepted moves. */
int64_t oracle_sequential_run(oracle_system* sys, uint64_t seed, uint64_t n_mcs);

/* Number of threads OpenMP would use. */
int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
