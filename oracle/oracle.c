/*
 * oracle.c - CPU restatement of the pscbfm colored two-pass scBFM sweep.
 * TEST INFRASTRUCTURE ONLY (see oracle.h for scope and parity pins).
 *
 * The lattice here is a plain x-fastest byte array indexed by the wrapped
 * corner, deliberately not the Z-curve layout of the GPU path, so that the
 * device's dilated-index arithmetic is checked against ordinary coordinates.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* CRITICAL.cu:2-11 */
uint32_t oracle_hash(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

/* SplitMix64 (Steele, Lea, Flood 2014; Vigna's reference code): the state
 * advances by the golden-ratio increment, the output is the mixed state.
 * The t-th output therefore mixes seed + (t+1)*gamma. */
uint64_t oracle_splitmix64(uint64_t seed, uint64_t t) {
    uint64_t z = seed + (t + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void oracle_r250_load(oracle_r250* r, const uint32_t state[250]) {
    memcpy(r->buf, state, sizeof r->buf);
    r->index = 0;
}

uint32_t oracle_r250_rand32(oracle_r250* r) {
    /* x[n] = x[n-250] ^ x[n-147]: the slot being replaced holds x[n-250] and
     * the one 103 slots ahead of it (modulo 250) was written 147 draws ago */
    const uint32_t j = r->index >= 147 ? r->index - 147 : r->index + 103;
    const uint32_t v = r->buf[r->index] ^ r->buf[j];
    r->buf[r->index] = v;
    r->index = r->index == 249 ? 0 : r->index + 1;
    return v;
}

/* Species and kernel seed of a substep: the explicit draw list when given,
 * else the counter-based draw. */
static void draw_for(const oracle_system* s, uint64_t t, uint32_t* sp, uint32_t* sigma) {
    if (s->draws) {
        *sp = s->draws[2 * (t - s->draw_t0)];
        *sigma = s->draws[2 * (t - s->draw_t0) + 1];
    } else {
        oracle_substep_draw(s->seed, t, s->n_colors, sp, sigma);
    }
}

void oracle_substep_draw(uint64_t seed, uint64_t t, uint32_t n_colors, uint32_t* species, uint32_t* kernel_seed) {
    const uint64_t r = oracle_splitmix64(seed, t);
    if (species) *species = (uint32_t)(r >> 32) % (n_colors ? n_colors : 1u);
    if (kernel_seed) *kernel_seed = (uint32_t)r;
}

/* CRITICAL.cu:243-261: components clipped to 3 bits, so +4 and -4 share an
 * entry. */
uint32_t oracle_bond_index(int32_t dx, int32_t dy, int32_t dz) {
    return ((uint32_t)dx & 7u) | (((uint32_t)dy & 7u) << 3) | (((uint32_t)dz & 7u) << 6);
}

static int is_scbfm_bond(int x, int y, int z) {
    int a[3] = {abs(x), abs(y), abs(z)};
    /* sort descending */
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (a[j] > a[i]) {
                int t = a[i];
                a[i] = a[j];
                a[j] = t;
            }
    static const int base[6][3] = {{2, 0, 0}, {2, 1, 0}, {2, 1, 1}, {2, 2, 1}, {3, 0, 0}, {3, 1, 0}};
    for (int b = 0; b < 6; ++b)
        if (a[0] == base[b][0] && a[1] == base[b][1] && a[2] == base[b][2]) return 1;
    return 0;
}

uint32_t oracle_standard_bondset(uint8_t forbidden[512]) {
    uint32_t allowed = 0;
    memset(forbidden, 1, 512);
    /* GPUScBFM_AB_Type.h:120-127 fills every (dx,dy,dz) in [-4,3]^3 */
    for (int dx = -4; dx < 4; ++dx)
        for (int dy = -4; dy < 4; ++dy)
            for (int dz = -4; dz < 4; ++dz) {
                const int ok = is_scbfm_bond(dx, dy, dz);
                forbidden[oracle_bond_index(dx, dy, dz)] = (uint8_t)!ok;
            }
    for (int i = 0; i < 512; ++i) allowed += forbidden[i] ? 0u : 1u;
    return allowed;
}

/* UpdaterGPUScBFM_AB_Type.cu:150-176: diluteBits<uint32_t,2> keeps the low
 * ceil(32/3)=11 bits and spreads bit i to bit 3i. */
uint32_t oracle_dilute_bits(uint32_t x) {
    uint32_t out = 0;
    x &= 0x7FFu;
    for (uint32_t i = 0; i < 11; ++i)
        if (x & (1u << i)) out |= 1u << (3 * i);
    return out;
}

uint32_t oracle_zcurve_index(uint32_t x, uint32_t y, uint32_t z, uint32_t lx, uint32_t ly, uint32_t lz) {
    return oracle_dilute_bits(x & (lx - 1)) + (oracle_dilute_bits(y & (ly - 1)) << 1) +
           (oracle_dilute_bits(z & (lz - 1)) << 2);
}

/* UpdaterGPUScBFM_AB_Type.cu:534-545: 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z */
void oracle_direction(uint32_t d, int32_t* dx, int32_t* dy, int32_t* dz) {
    int32_t v[3] = {0, 0, 0};
    v[d >> 1] = (d & 1u) ? 1 : -1;
    *dx = v[0];
    *dy = v[1];
    *dz = v[2];
}

static inline uint64_t lin(const uint32_t box[3], int32_t x, int32_t y, int32_t z) {
    const uint64_t wx = (uint32_t)x & (box[0] - 1);
    const uint64_t wy = (uint32_t)y & (box[1] - 1);
    const uint64_t wz = (uint32_t)z & (box[2] - 1);
    return wx + box[0] * (wy + (uint64_t)box[1] * wz);
}

int oracle_front_blocked(const uint8_t* occ, const uint32_t box[3], const int32_t p[3], uint32_t d) {
    const uint32_t axis = d >> 1;
    const int32_t s = (d & 1u) ? 1 : -1;
    const uint32_t u = axis == 0 ? 1u : 0u;
    const uint32_t v = axis == 2 ? 1u : 2u;
    int32_t f[3] = {p[0], p[1], p[2]};
    f[axis] += 2 * s;
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b) {
            int32_t c[3] = {f[0], f[1], f[2]};
            c[u] += a;
            c[v] += b;
            if (occ[lin(box, c[0], c[1], c[2])]) return 1;
        }
    return 0;
}

static int validate(const oracle_system* s) {
    for (int a = 0; a < 3; ++a)
        if (s->box[a] < 8 || (s->box[a] & (s->box[a] - 1))) return 0;
    if (s->n && (!s->xyz || !s->nbr_start || !s->color || !s->forbidden)) return 0;
    for (uint32_t i = 0; i < s->n; ++i)
        if (s->color[i] >= s->n_colors) return 0;
    return 1;
}

/* Members of each species in increasing monomer id: the slot order of the
 * reference's species sort (iToiNew, UpdaterGPUScBFM_AB_Type.cu:600-608), so
 * position in the list is the species-local index j of CRITICAL.cu:322. */
static void species_lists(const oracle_system* s, uint32_t** start_out, uint32_t** members_out) {
    uint32_t* start = (uint32_t*)calloc((size_t)s->n_colors + 1, sizeof(uint32_t));
    uint32_t* members = (uint32_t*)malloc(sizeof(uint32_t) * (s->n ? s->n : 1));
    for (uint32_t i = 0; i < s->n; ++i) start[s->color[i] + 1]++;
    for (uint32_t c = 0; c < s->n_colors; ++c) start[c + 1] += start[c];
    uint32_t* next = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_colors ? s->n_colors : 1));
    for (uint32_t c = 0; c < s->n_colors; ++c) next[c] = start[c];
    for (uint32_t i = 0; i < s->n; ++i) members[next[s->color[i]]++] = i;
    free(next);
    *start_out = start;
    *members_out = members;
}

static int bonds_ok(const oracle_system* s, uint32_t i, const int32_t q[3]) {
    for (uint32_t k = s->nbr_start[i]; k < s->nbr_start[i + 1]; ++k) {
        const int32_t* pk = s->xyz + 3 * (size_t)s->nbr[k];
        if (s->forbidden[oracle_bond_index(pk[0] - q[0], pk[1] - q[1], pk[2] - q[2])]) return 0;
    }
    return 1;
}

static void shuffle_order(uint32_t* order, uint32_t n, uint64_t key, uint64_t t) {
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    for (uint32_t i = n; i > 1; --i) {
        const uint32_t r = (uint32_t)(oracle_splitmix64(key ^ (t * 0xD1B54A32D192ED03ull), i) % i);
        const uint32_t tmp = order[i - 1];
        order[i - 1] = order[r];
        order[r] = tmp;
    }
}

/* An explicit draw list must cover the substeps and name existing species. */
static int draws_ok(const oracle_system* s, uint64_t t0, uint64_t n_substeps) {
    if (!s->draws) return 1;
    if (t0 < s->draw_t0) return 0;
    for (uint64_t k = 0; k < n_substeps; ++k)
        if (s->draws[2 * (t0 + k - s->draw_t0)] >= s->n_colors) return 0;
    return 1;
}

int64_t oracle_run(oracle_system* s, uint64_t t0, uint64_t n_substeps, uint32_t* acc_log, int n_threads,
                   uint64_t shuffle_seed) {
    if (!validate(s) || !draws_ok(s, t0, n_substeps)) return -1;
    if (s->n == 0 || s->n_colors == 0) {
        if (acc_log) memset(acc_log, 0, sizeof(uint32_t) * n_substeps);
        return 0;
    }
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
    const uint64_t cells = (uint64_t)s->box[0] * s->box[1] * s->box[2];
    uint8_t* occ = (uint8_t*)calloc(cells, 1);
    uint8_t* T = (uint8_t*)calloc(cells, 1);
    int8_t* dir = (int8_t*)malloc(s->n);
    uint32_t *start, *members;
    species_lists(s, &start, &members);
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * s->n);
    for (uint32_t i = 0; i < s->n; ++i) {
        const int32_t* p = s->xyz + 3 * (size_t)i;
        occ[lin(s->box, p[0], p[1], p[2])] = 1;
        dir[i] = -1;
    }
    int64_t total = 0;
    for (uint64_t step = 0; step < n_substeps; ++step) {
        uint32_t sp, sigma;
        draw_for(s, t0 + step, &sp, &sigma);
        const uint32_t* mem = members + start[sp];
        const uint32_t ns = start[sp + 1] - start[sp];
        if (shuffle_seed) shuffle_order(order, ns, shuffle_seed, t0 + step);
        else
            for (uint32_t j = 0; j < ns; ++j) order[j] = j;

        /* Pass 1 (CRITICAL.cu:298-374): draw, bond check, front check
         * against the old occupancy, mark the target. */
#pragma omp parallel for schedule(static) num_threads(n_threads)
        for (int64_t oi = 0; oi < (int64_t)ns; ++oi) {
            const uint32_t j = order[oi];
            const uint32_t i = mem[j];
            const uint32_t d = oracle_hash(oracle_hash(j) ^ sigma) % 6u;
            const int32_t* p = s->xyz + 3 * (size_t)i;
            int32_t e[3];
            oracle_direction(d, &e[0], &e[1], &e[2]);
            const int32_t q[3] = {p[0] + e[0], p[1] + e[1], p[2] + e[2]};
            if (bonds_ok(s, i, q) && !oracle_front_blocked(occ, s->box, p, d)) {
                dir[i] = (int8_t)d;
#pragma omp atomic write
                T[lin(s->box, q[0], q[1], q[2])] = 1;
            }
        }
        /* Pass 2 (CRITICAL.cu:376-411): a proposal survives if no other
         * proposal's target lies in its front plane; survivors move. */
        int64_t acc = 0;
#pragma omp parallel for schedule(static) num_threads(n_threads) reduction(+ : acc)
        for (int64_t oi = 0; oi < (int64_t)ns; ++oi) {
            const uint32_t i = mem[order[oi]];
            if (dir[i] < 0) continue;
            int32_t* p = s->xyz + 3 * (size_t)i;
            const uint32_t d = (uint32_t)dir[i];
            if (oracle_front_blocked(T, s->box, p, d)) continue;
            int32_t e[3];
            oracle_direction(d, &e[0], &e[1], &e[2]);
            occ[lin(s->box, p[0], p[1], p[2])] = 0;
            p[0] += e[0];
            p[1] += e[1];
            p[2] += e[2];
            occ[lin(s->box, p[0], p[1], p[2])] = 1;
            dir[i] = (int8_t)(d + 8); /* moved; remembered for the clear below */
            ++acc;
        }
        /* Clear the proposal marks (the reference memsets, CRITICAL.cu:568). */
        for (uint32_t j = 0; j < ns; ++j) {
            const uint32_t i = mem[j];
            if (dir[i] < 0) continue;
            const int32_t* p = s->xyz + 3 * (size_t)i;
            int32_t e[3];
            const int moved = dir[i] >= 8;
            oracle_direction((uint32_t)(dir[i] & 7), &e[0], &e[1], &e[2]);
            if (moved) T[lin(s->box, p[0], p[1], p[2])] = 0;
            else T[lin(s->box, p[0] + e[0], p[1] + e[1], p[2] + e[2])] = 0;
            dir[i] = -1;
        }
        if (acc_log) acc_log[step] = (uint32_t)acc;
        total += acc;
    }
    free(order);
    free(start);
    free(members);
    free(dir);
    free(T);
    free(occ);
    return total;
}

/* Periodic cube overlap: corners a and b overlap iff every component of
 * (a-b) mod L is in {-1,0,1}. */
static int cubes_overlap(const uint32_t box[3], const int32_t* a, const int32_t* b) {
    for (int k = 0; k < 3; ++k) {
        const uint32_t dlt = (uint32_t)(a[k] - b[k]) & (box[k] - 1);
        if (!(dlt == 0 || dlt == 1 || dlt == box[k] - 1)) return 0;
    }
    return 1;
}

int64_t oracle_run_naive(oracle_system* s, uint64_t t0, uint64_t n_substeps, uint32_t* acc_log) {
    if (!validate(s) || !draws_ok(s, t0, n_substeps)) return -1;
    if (s->n == 0 || s->n_colors == 0) {
        if (acc_log) memset(acc_log, 0, sizeof(uint32_t) * n_substeps);
        return 0;
    }
    uint32_t *start, *members;
    species_lists(s, &start, &members);
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)s->n);
    uint8_t* prop = (uint8_t*)malloc(s->n);
    uint8_t* keep = (uint8_t*)malloc(s->n);
    int64_t total = 0;
    for (uint64_t step = 0; step < n_substeps; ++step) {
        uint32_t sp, sigma;
        draw_for(s, t0 + step, &sp, &sigma);
        const uint32_t* mem = members + start[sp];
        const uint32_t ns = start[sp + 1] - start[sp];
        for (uint32_t j = 0; j < ns; ++j) {
            const uint32_t i = mem[j];
            const uint32_t d = oracle_hash(oracle_hash(j) ^ sigma) % 6u;
            const int32_t* p = s->xyz + 3 * (size_t)i;
            int32_t e[3];
            oracle_direction(d, &e[0], &e[1], &e[2]);
            int32_t* qi = q + 3 * (size_t)i;
            qi[0] = p[0] + e[0];
            qi[1] = p[1] + e[1];
            qi[2] = p[2] + e[2];
            int ok = bonds_ok(s, i, qi);
            for (uint32_t k = 0; ok && k < s->n; ++k)
                if (k != i && cubes_overlap(s->box, qi, s->xyz + 3 * (size_t)k)) ok = 0;
            prop[i] = (uint8_t)ok;
        }
        int64_t acc = 0;
        for (uint32_t a = 0; a < ns; ++a) {
            const uint32_t i = mem[a];
            keep[i] = 0;
            if (!prop[i]) continue;
            int ok = 1;
            for (uint32_t b = 0; ok && b < ns; ++b) {
                const uint32_t k = mem[b];
                if (k != i && prop[k] && cubes_overlap(s->box, q + 3 * (size_t)i, q + 3 * (size_t)k)) ok = 0;
            }
            keep[i] = (uint8_t)ok;
        }
        for (uint32_t a = 0; a < ns; ++a) {
            const uint32_t i = mem[a];
            if (keep[i]) {
                memcpy(s->xyz + 3 * (size_t)i, q + 3 * (size_t)i, 3 * sizeof(int32_t));
                ++acc;
            }
        }
        if (acc_log) acc_log[step] = (uint32_t)acc;
        total += acc;
    }
    free(keep);
    free(prop);
    free(q);
    free(start);
    free(members);
    return total;
}

int oracle_check_system(const oracle_system* s) {
    const uint64_t cells = (uint64_t)s->box[0] * s->box[1] * s->box[2];
    uint8_t* occ = (uint8_t*)calloc(cells, 1);
    int status = 0;
    for (uint32_t i = 0; i < s->n && !status; ++i) {
        const int32_t* p = s->xyz + 3 * (size_t)i;
        for (int c = 0; c < 8; ++c) {
            const uint64_t l = lin(s->box, p[0] + (c & 1), p[1] + ((c >> 1) & 1), p[2] + ((c >> 2) & 1));
            if (occ[l]) {
                status = 1;
                break;
            }
            occ[l] = 1;
        }
    }
    free(occ);
    if (status) return status;
    for (uint32_t i = 0; i < s->n; ++i)
        for (uint32_t k = s->nbr_start[i]; k < s->nbr_start[i + 1]; ++k) {
            const int32_t* a = s->xyz + 3 * (size_t)i;
            const int32_t* b = s->xyz + 3 * (size_t)s->nbr[k];
            /* minimum image (box edges are multiples of 8, the bond table
             * index only sees the vector modulo 8) */
            int32_t d[3];
            for (int k2 = 0; k2 < 3; ++k2) {
                const int32_t L = (int32_t)s->box[k2];
                d[k2] = (int32_t)((uint32_t)(b[k2] - a[k2] + L / 2) & (uint32_t)(L - 1)) - L / 2;
            }
            if (abs(d[0]) > 3 || abs(d[1]) > 3 || abs(d[2]) > 3) return 2;
            if (s->forbidden[oracle_bond_index(d[0], d[1], d[2])]) return 2;
        }
    return 0;
}

int64_t oracle_sequential_run(oracle_system* s, uint64_t seed, uint64_t n_mcs) {
    if (!validate(s) || s->n == 0) return s->n == 0 ? 0 : -1;
    const uint64_t cells = (uint64_t)s->box[0] * s->box[1] * s->box[2];
    uint8_t* occ = (uint8_t*)calloc(cells, 1);
    for (uint32_t i = 0; i < s->n; ++i) {
        const int32_t* p = s->xyz + 3 * (size_t)i;
        occ[lin(s->box, p[0], p[1], p[2])] = 1;
    }
    int64_t acc = 0;
    const uint64_t attempts = n_mcs * s->n;
    for (uint64_t a = 0; a < attempts; ++a) {
        const uint64_t r = oracle_splitmix64(seed, a);
        const uint32_t i = (uint32_t)((r >> 32) % s->n);
        const uint32_t d = (uint32_t)(r & 0xFFFFFFFFu) % 6u;
        int32_t* p = s->xyz + 3 * (size_t)i;
        int32_t e[3];
        oracle_direction(d, &e[0], &e[1], &e[2]);
        const int32_t q[3] = {p[0] + e[0], p[1] + e[1], p[2] + e[2]};
        if (!bonds_ok(s, i, q) || oracle_front_blocked(occ, s->box, p, d)) continue;
        occ[lin(s->box, p[0], p[1], p[2])] = 0;
        occ[lin(s->box, q[0], q[1], q[2])] = 1;
        p[0] = q[0];
        p[1] = q[1];
        p[2] = q[2];
        ++acc;
    }
    free(occ);
    return acc;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
