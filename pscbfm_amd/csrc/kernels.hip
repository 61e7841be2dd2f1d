// gfx950 kernels of the colored two-pass scBFM sweep.
//
// Pass 1 (check):   CRITICAL.cu:298-374  kernelSimulationScBFMCheckSpecies
// Pass 2 (perform): CRITICAL.cu:376-411  kernelSimulationScBFMPerformSpeciesAndApply
// Proposal clear:   CRITICAL.cu:561-573  (here folded into pass 1 or 2, other parity)
//
// Design (wave64, CDNA4):
//  * positions are Z-curve codes, so the own lattice cell is the code itself
//    and every neighbouring cell is a masked add on one axis' bits;
//  * the forbidden-bond table is a 512-bit mask in LDS, indexed by the
//    Z-order bond key of two codes;
//  * both passes are bound by the number of gather instructions and memory
//    latency more than by bandwidth: a pass-1 thread handles SLOTS monomers
//    and issues all their loads in two batches (own codes, draw keys and
//    partner ids; then partner codes and the front plane together, from one
//    basic block); the 3x3 front plane is read as at most four 64-bit words;
//  * a species' slots may be in the Z-curve order of the starting corners,
//    so that a wave's members and their partners and lattice words lie
//    close together; each slot's key is the inner move hash of the member's
//    rank in the reference's order, which keys its draw either way;
//  * pass 1 compacts the surviving proposals of its workgroup into a list
//    (wave ballot + LDS counter; the order inside a list is irrelevant
//    because the result of a substep does not depend on processing order),
//    and pass 2 walks those lists densely;
//  * the proposal lattice is 1 bit per cell and double-buffered by substep
//    parity: pass 1 of substep t sets bits in buffer t&1, pass 2 of substep t
//    reads it, and pass 1 or pass 2 of substep t zeroes buffer (t+1)&1 (last
//    read by pass 2 of substep t-1, next written by pass 1 of substep t+1);
//  * the occupancy lattice is updated with plain byte stores (no atomics):
//    device-scope atomics are executed beyond the per-XCD L2 and cost a
//    memory transaction each;
//  * workgroup w runs on XCD w%8: a single system renumbers workgroups so
//    each XCD takes one contiguous eighth of the species; a batch's tile
//    table keeps a small replica's workgroups on one XCD and spreads a large
//    one over all eight;
//  * the species and kernel seed of a substep come from the device state:
//    the counter draw of (seed, substep), or an explicit draw table (R250
//    mode), so launches carry no per-substep host data.
#include "kernels.h"

namespace pscbfm {

namespace {

// Logical workgroup id: consecutive ids on the same XCD.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x;  // multiple of 8 (grid_for)
    return (blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3);
}

// Every device buffer is reached through a global-address-space pointer.
// The batch kernels read their buffer pointers from a descriptor table in
// memory, which leaves the compiler generic (FLAT) pointers: on gfx950 the
// waits it then places before using FLAT-loaded gather results proved
// insufficient under load (a few lanes consumed words that had not arrived;
// a full s_waitcnt after the gathers removed the divergence).  GLOBAL
// instructions, which the single-system kernels always get, are tracked by
// the vector-memory counter alone.
template <class T>
__device__ __forceinline__ gptr<T> global_ptr(T* p) {
    return (gptr<T>)p;
}

// What a sweep launch reads from its system's state, taken as four 16-byte
// loads of the whole 64-byte record before any branch, so that the start of
// every workgroup waits for one round trip: the substep it serves (device
// counter + k), the stop point, the draw source and the per-substep log and
// accepted-count pointers.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct Launch {
    uint64_t t;
    SubstepDraw draw;
    gptr<uint32_t> log;
    uint64_t logFirst, logCapacity;
    gptr<uint64_t> acceptedByBlock;
};

__device__ __forceinline__ uint64_t join64(uint32_t lo, uint32_t hi) { return uint64_t(lo) | (uint64_t(hi) << 32); }

// False when substep t is at or past the stop point.  Else fills `out`; the
// draw is entry t - drawBase of the explicit table when the state holds one,
// else the counter draw of (seed, t).
__device__ __forceinline__ bool begin_substep(const SysDesc& d, uint32_t k, Launch& out) {
    const gptr<const u32x4> w = (gptr<const u32x4>)d.state.global();
    const u32x4 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];  // DevState field order
    const uint64_t t = join64(w0.x, w0.y) + k;
    // A state whose draw table, log or accepted counts are not aligned to
    // their entries is not served.  Testing them here, with the stop point,
    // also keeps all four loads ahead of the branch as one batch.
    const uint32_t misaligned = ((w1.x | w3.z) & 7u) | (w2.x & 3u);
    if (t >= join64(w0.z, w0.w) || misaligned != 0u) return false;
    out.t = t;
    const uint64_t table = join64(w1.x, w1.y);
    if (table == 0u) {
        out.draw = substep_draw(d.seed, t, d.nColors);
    } else {
        // every lane reads the same entry: keep the draw in scalar registers
        const uint64_t e = global_ptr(reinterpret_cast<const uint64_t*>(uintptr_t(table)))[t - join64(w1.z, w1.w)];
        out.draw.species = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(e));
        out.draw.seed = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(e >> 32));
    }
    out.log = (gptr<uint32_t>)(uintptr_t(join64(w2.x, w2.y)));
    out.logFirst = join64(w2.z, w2.w);
    out.logCapacity = join64(w3.x, w3.y);
    out.acceptedByBlock = (gptr<uint64_t>)(uintptr_t(join64(w3.z, w3.w)));
    return true;
}

// Species record, read word by word through a global pointer.
__device__ __forceinline__ ColorInfo load_color(gptr<const ColorInfo> colors, uint32_t species) {
    const gptr<const uint32_t> w = (gptr<const uint32_t>)(colors + species);
    ColorInfo c{};
    c.slotOffset = w[offsetof(ColorInfo, slotOffset) / 4];
    c.count = w[offsetof(ColorInfo, count) / 4];
    c.nbrOffset = w[offsetof(ColorInfo, nbrOffset) / 4];
    c.nbrPitch = w[offsetof(ColorInfo, nbrPitch) / 4];
    c.maxDegree = w[offsetof(ColorInfo, maxDegree) / 4];
    c.idsOffset = w[offsetof(ColorInfo, idsOffset) / 4];
    c.idWords = w[offsetof(ColorInfo, idWords) / 4];
    return c;
}

// A list entry as one 64-bit word: low = species-local index | direction
// << kDirShift, high = old code.
__device__ __forceinline__ uint64_t pack_entry(uint32_t x, uint32_t y) { return uint64_t(x) | (uint64_t(y) << 32); }

struct Move {
    uint32_t to;        // code of the new corner
    uint32_t cell[9];   // the 3x3 plane in front of the move
    bool wraps;         // the axis coordinate crosses the periodic boundary
};

// The Z-curve masks of x, y and z as values in scalar registers.  Passing
// the descriptor's array instead lets the compiler turn the per-lane choice
// of an axis into a per-lane load from the argument segment, one more
// dependent round trip in front of the lattice reads.
struct AxisMasks {
    uint32_t x, y, z;
};
__device__ __forceinline__ AxisMasks axis_masks(const SysDesc& d) {
    AxisMasks m;
    m.x = uint32_t(__builtin_amdgcn_readfirstlane(d.mask[0]));
    m.y = uint32_t(__builtin_amdgcn_readfirstlane(d.mask[1]));
    m.z = uint32_t(__builtin_amdgcn_readfirstlane(d.mask[2]));
    return m;
}

// Direction d: 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z (UpdaterGPUScBFM_AB_Type.cu:534-545).
// Front plane: p + 2 e_d + a e_u + b e_v, a,b in {-1,0,1} (CRITICAL.cu:62-176).
__device__ __forceinline__ Move make_move(uint32_t m, uint32_t dir, const AxisMasks& mask) {
    const uint32_t axis = dir >> 1;
    const bool plus = (dir & 1u) != 0;
    const uint32_t Ma = axis == 0 ? mask.x : (axis == 1 ? mask.y : mask.z);
    const uint32_t Mu = axis == 0 ? mask.y : mask.x;
    const uint32_t Mv = axis == 2 ? mask.y : mask.z;
    const uint32_t oneA = Ma & (~Ma + 1u);
    const uint32_t stepA = plus ? oneA : Ma;  // +1 or -1 in dilated form
    Move mv;
    const uint32_t qa = dil_add(m, Ma, stepA);
    mv.to = (m & ~Ma) | qa;
    mv.wraps = plus ? ((m & Ma) == Ma) : ((m & Ma) == 0u);
    const uint32_t fa = dil_add(qa, Ma, stepA);
    const uint32_t u0 = m & Mu, v0 = m & Mv;
    const uint32_t oneU = Mu & (~Mu + 1u), oneV = Mv & (~Mv + 1u);
    const uint32_t u[3] = {dil_add(u0, Mu, Mu), u0, dil_add(u0, Mu, oneU)};
    const uint32_t v[3] = {dil_add(v0, Mv, Mv), v0, dil_add(v0, Mv, oneV)};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) mv.cell[3 * a + b] = fa | u[a] | v[b];
    return mv;
}

// Whether any of the nine front-plane cells is set, read as 64-bit words.
// The low code bits interleave x, y and z, so a 64-bit word of the 1-bit
// lattice (64 cells, SHIFT 6) is an aligned 4x4x4 block and one of the byte
// lattice (8 cells, SHIFT 3) an aligned 2x2x2 block.  In either case the 3x3
// plane touches at most four words, those of (u-1 or u+1) x (v-1 or v+1):
// four 8-byte gathers replace nine narrow ones, and the middle row/column
// takes the word of whichever outer row/column shares its block.
template <int SHIFT>
__device__ __forceinline__ uint32_t plane_any(gptr<const uint64_t> words, const Move& mv) {
    const uint64_t w00 = words[mv.cell[0] >> SHIFT];  // (u-1, v-1)
    const uint64_t w02 = words[mv.cell[2] >> SHIFT];  // (u-1, v+1)
    const uint64_t w20 = words[mv.cell[6] >> SHIFT];  // (u+1, v-1)
    const uint64_t w22 = words[mv.cell[8] >> SHIFT];  // (u+1, v+1)
    // does the middle u (v) share the block of u-1 (v-1)?
    const bool uLow = ((mv.cell[3] ^ mv.cell[0]) >> SHIFT) == 0u;
    const bool vLow = ((mv.cell[1] ^ mv.cell[0]) >> SHIFT) == 0u;
    const uint64_t wRow0 = vLow ? w00 : w02, wRow2 = vLow ? w20 : w22;  // words of (u-1, v), (u+1, v)
    const uint64_t w10 = uLow ? w00 : w20, w12 = uLow ? w02 : w22;      // words of (u, v-1), (u, v+1)
    const uint64_t w11 = uLow ? wRow0 : wRow2;                          // word of (u, v)
    const uint64_t w[9] = {w00, wRow0, w02, w10, w11, w12, w20, wRow2, w22};
    constexpr uint32_t kLow = (1u << SHIFT) - 1u;
    constexpr uint32_t kUnit = SHIFT == 6 ? 1u : 8u;  // bits per cell inside the word
    uint64_t any = 0;
#pragma unroll
    for (int c = 0; c < 9; ++c) any |= w[c] >> ((mv.cell[c] & kLow) * kUnit);
    return uint32_t(any & (SHIFT == 6 ? 1u : 0xFFu));
}

// Rank of this lane among the lanes of its wave whose bit is set in `mask`.
__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// List entry: x = species-local index | direction << 29, y = old code.
constexpr uint32_t kDirShift = 29;

// Zeroes slice lb of nb of the proposal buffer of the other parity, 16 bytes
// per thread and step: pass 2 of substep t-1 was its last reader and pass 1
// of substep t+1 is its next writer, so either pass of substep t may do it
// (LaunchShape::clearPass).  In pass 1 the stores go out ahead of its loads
// and stream while its waves wait on them and compute the bond test; in
// pass 2 they overlap the list walk.
__device__ __forceinline__ void clear_other_parity(const SysDesc& d, uint64_t t, uint32_t lb, uint32_t nb) {
    const gptr<uint64_t> prev = (gptr<uint64_t>)(d.proposal.global() + size_t((t + 1u) & 1u) * d.latWords);
    const uint32_t quads = d.latWords >> 2;
    const uint32_t per = (quads + nb - 1) / nb;
    const uint32_t q0 = lb * per;
    const uint32_t q1 = min(q0 + per, quads);
    for (uint32_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        prev[2 * size_t(q)] = 0u;
        prev[2 * size_t(q) + 1] = 0u;
    }
}

// Pass 1 of substep t for logical workgroup lb of nb of one system.  The
// survivors go to `list` (capacity SLOTS * blockDim), their number to *count.
// With `clear` it also zeroes slice lb of nb of the other parity's proposals.
// KEYED: the system's slots are in spatial order and each slot's draw key is
// read from its `key` piece; else the key is the hash of the slot index.  A
// compile-time choice: as a runtime branch between the load and the hash
// inside the load batch, the compiler's wait for the loaded key also lands on
// the hashing path, which then waits for every load issued before it.
template <int LAT, int SLOTS, int MAXDEG, bool KEYED>
__device__ __forceinline__ void check_body(const SysDesc& d, uint64_t t, SubstepDraw draw, uint32_t lb, uint32_t nb,
                                           bool clear, gptr<uint64_t> list, gptr<uint32_t> count,
                                           uint32_t* sForbidden, uint32_t* sCount) {
    if (clear) clear_other_parity(d, t, lb, nb);
    const ColorInfo ci = load_color(d.colors.global(), draw.species);
    const gptr<uint32_t> proposal = d.proposal.global();

    // The bond table (16 words, the same for every lane: scalar loads) goes
    // to LDS after the gathers below are issued, and the barrier that
    // publishes it stands in front of the bond test.  A barrier ahead of the
    // gathers held every wave's loads back by the table's round trip.
    uint32_t forbidden[16];
#pragma unroll
    for (int w = 0; w < 16; ++w) forbidden[w] = d.bondMask.global()[w];

    const uint32_t B = blockDim.x;
    const uint32_t base = lb * (SLOTS * B);
    const gptr<const uint32_t> posAll = (gptr<const uint32_t>)d.pos.global();
    const gptr<const uint32_t> pos = posAll + ci.slotOffset;
    const gptr<const uint32_t> nbr = d.nbr.global();
    // Inner hash of each member's species-local index in the reference's
    // order, the key of its draw: read from the slot's key when the slots
    // are in spatial order, else the hash of the slot index itself.
    const gptr<const uint32_t> keys = d.key.global() + ci.slotOffset;
    // Partner ids: the bit-lattice kernel reads the full matrix rows; the
    // uint8-lattice kernel, whose plane words cost more traffic, reads the
    // 16-bit codes (half the bytes) and the full matrix only for escapes.
    const gptr<const uint32_t> ids = d.ids.global() + ci.idsOffset;
    uint32_t rb[MAXDEG];
#pragma unroll
    for (int r = 0; r < MAXDEG; ++r)
        rb[r] = LAT == kLatticeU8 ? d.rowBase.global()[draw.species * kMaxConnectivity + uint32_t(r)] : 0u;
    const gptr<const uint64_t> lattice = (gptr<const uint64_t>)d.lattice.global();
    const AxisMasks masks = axis_masks(d);

    // The load phase is branch-free: lanes without a monomer (or without a
    // partner in a row) load from an in-bounds stand-in address and discard
    // the value, so every load of a thread sits in one basic block and all of
    // them are in flight together.
    // First batch of loads: own codes and partner ids.
    uint32_t j[SLOTS], m[SLOTS], hj[SLOTS], partner[SLOTS][MAXDEG];
    bool live[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        j[s] = base + uint32_t(s) * B + threadIdx.x;
        live[s] = j[s] < ci.count;
        const uint32_t jj = live[s] ? j[s] : 0u;  // count > 0 for every species
        m[s] = pos[jj];
        if constexpr (KEYED)
            hj[s] = keys[jj];
        else
            hj[s] = move_hash(jj);
        if constexpr (LAT == kLatticeU8) {
#pragma unroll
            for (int q = 0; q < MAXDEG / 2; ++q) {
                const bool have = uint32_t(q) < ci.idWords;  // the same for every lane
                const uint32_t w = ids[have ? size_t(jj) * ci.idWords + uint32_t(q) : 0u];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int r = 2 * q + h;
                    const int32_t v = int32_t(int16_t(uint16_t(h ? (w >> 16) : (w & 0xFFFFu))));
                    uint32_t id = rb[r] + jj + uint32_t(v);
                    // an escape is read only for a real row of this species, so the index stays in its matrix
                    if (have && v == kIdEscape && uint32_t(r) < ci.maxDegree)
                        id = nbr[ci.nbrOffset + size_t(r) * ci.nbrPitch + jj];
                    id = min(id, d.nSlots - 1u);  // a valid slot whatever the code
                    const bool has = have && v != kIdNone && uint32_t(r) < ci.maxDegree;
                    partner[s][r] = (live[s] && has) ? id : kNoNeighbor;
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < MAXDEG; ++r) {
                const bool row = uint32_t(r) < ci.maxDegree;
                const uint32_t id = nbr[row ? ci.nbrOffset + size_t(r) * ci.nbrPitch + jj : 0u];
                partner[s][r] = (live[s] && row) ? id : kNoNeighbor;
            }
        }
    }
    // Second batch: partner codes and the front plane.  CRITICAL.cu:322-323: the
    // draw is keyed by the species-local index in the reference's order.
    uint32_t dir[SLOTS], to[SLOTS], partnerCode[SLOTS][MAXDEG], occ[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        dir[s] = move_hash(hj[s] ^ draw.seed) % 6u;
        const Move mv = make_move(m[s], dir[s], masks);
        to[s] = mv.to;
#pragma unroll
        for (int r = 0; r < MAXDEG; ++r)
            partnerCode[s][r] = posAll[partner[s][r] != kNoNeighbor ? partner[s][r] : 0u];
        occ[s] = plane_any<LAT == kLatticeU8 ? 3 : 6>(lattice, mv);
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < 16; ++w) sForbidden[w] = forbidden[w];
        *sCount = 0;
    }
    __syncthreads();
    bool ok[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        uint32_t bad = 0;
#pragma unroll
        for (int r = 0; r < MAXDEG; ++r) {
            const uint32_t key = bond_key(partnerCode[s][r], to[s]);
            if (partner[s][r] != kNoNeighbor) bad |= sForbidden[key >> 5] >> (key & 31u);
        }
        ok[s] = live[s] && !(bad & 1u) && occ[s] == 0u;
    }

    // Mark targets and append survivors to this workgroup's list: one LDS
    // counter update per wave for all its slots.
    const gptr<uint32_t> T = proposal + size_t(t & 1u) * d.latWords;
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t mask[SLOTS];
    uint32_t before[SLOTS], total = 0;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        mask[s] = __ballot(ok[s]);
        before[s] = total;
        total += uint32_t(__popcll(mask[s]));
    }
    uint32_t waveBase = 0;
    if (lane == 0 && total != 0) waveBase = atomicAdd(sCount, total);
    waveBase = __shfl(waveBase, 0, 64);
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        if (ok[s]) {
            __hip_atomic_fetch_or(&T[to[s] >> 5], 1u << (to[s] & 31u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            list[waveBase + before[s] + lane_rank(mask[s])] = pack_entry(j[s] | (dir[s] << kDirShift), m[s]);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *count = *sCount;
}

// Sum over the wave of small per-lane counts, bit by bit with ballots and
// scalar popcounts (a uniform loop over the bits any lane has set).
__device__ __forceinline__ uint32_t wave_sum_small(uint32_t v) {
    uint32_t total = 0;
    for (uint32_t b = 0; b < 32u && __ballot((v >> b) != 0u) != 0u; ++b)
        total += uint32_t(__popcll(__ballot(((v >> b) & 1u) != 0u))) << b;
    return total;
}

// Adds a wave's accepted moves to its workgroup's running total and to the
// substep's log entry, without waiting: atomics whose results are unused.
__device__ __forceinline__ void add_accepted(const Launch& L, gptr<uint64_t> accSlot, uint32_t accepted) {
    const uint32_t total = wave_sum_small(accepted);
    if ((threadIdx.x & 63u) == 0u && total != 0u) {
        __hip_atomic_fetch_add(&accSlot[0], uint64_t(total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (L.log != nullptr) {
            const uint64_t idx = L.t - L.logFirst;
            if (idx < L.logCapacity)
                __hip_atomic_fetch_add(&L.log[idx], total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Applies an accepted move of the monomer in `slot` from code m in
// direction dir (pass 2, CRITICAL.cu:400-408): its position, the occupancy
// of the old and new corner, and the periodic image counter on a wrap.
template <int LAT>
__device__ __forceinline__ void apply_move(const SysDesc& d, gptr<uint32_t> pos, gptr<uint8_t> lat, uint32_t slot,
                                           uint32_t m, uint32_t dir, const Move& mv) {
    pos[slot] = mv.to;
    if (LAT == kLatticeU8) {
        lat[m] = 0;
        lat[mv.to] = 1;
    } else {
        // Byte b of the bit lattice holds the aligned 2x2x2 block of cells
        // 8b..8b+7 (the low three code bits are the lowest bit of x, y and
        // z).  Two corners in one block would overlap, and an accepted move's
        // old and new corner cannot share a block with any other monomer's
        // old or new corner, so every byte touched here has exactly one
        // writer and its new value is known: 0 for the vacated block, the
        // single new bit for the entered one.
        const uint32_t bFrom = m >> 3, bTo = mv.to >> 3;
        if (bFrom != bTo) lat[bFrom] = 0;
        lat[bTo] = uint8_t(1u << (mv.to & 7u));
    }
    if (mv.wraps) {
        const uint32_t axis = dir >> 1;
        const gptr<int32_t> img = axis == 0 ? d.imgX.global() : (axis == 1 ? d.imgY.global() : d.imgZ.global());
        img[slot] += (dir & 1u) ? 1 : -1;
    }
}

// Pass 2 of substep t over one pass-1 list.  The accepted count is added to
// *accSlot, an entry only this workgroup's waves add to.  With `clear` it
// also zeroes slice lb of nb of the other parity's proposals.
template <int LAT>
__device__ __forceinline__ void perform_body(const SysDesc& d, const Launch& L, uint32_t lb, uint32_t nb, bool clear,
                                             gptr<const uint64_t> list, uint32_t count, gptr<uint64_t> accSlot) {
    const uint64_t t = L.t;
    const SubstepDraw draw = L.draw;
    const ColorInfo ci = load_color(d.colors.global(), draw.species);
    const uint32_t B = blockDim.x;

    if (clear) clear_other_parity(d, t, lb, nb);

    const AxisMasks masks = axis_masks(d);
    const gptr<const uint64_t> T64 =
        (gptr<const uint64_t>)(d.proposal.global() + size_t(t & 1u) * d.latWords);
    const gptr<uint32_t> pos = d.pos.global();
    const gptr<uint8_t> lat = d.lattice.global();

    // The list is compacted, so about a fifth of its capacity is filled: walk
    // only the entries that exist, one per thread and round.
    uint32_t accepted = 0;
    for (uint32_t idx = threadIdx.x; idx < count; idx += B) {
        const uint64_t e = list[idx];
        const uint32_t jj = uint32_t(e) & ((1u << kDirShift) - 1u);
        const uint32_t dir = uint32_t(e) >> kDirShift;
        const uint32_t m = uint32_t(e >> 32);
        const Move mv = make_move(m, dir, masks);
        if (plane_any<6>(T64, mv)) continue;
        apply_move<LAT>(d, pos, lat, ci.slotOffset + jj, m, dir, mv);
        ++accepted;
    }
    add_accepted(L, accSlot, accepted);
}

// ---- single system ----------------------------------------------------------------
template <int LAT, int SLOTS, int MAXDEG, bool KEYED>
__global__ __launch_bounds__(256) void k_check(SysDesc d, uint32_t k, uint32_t clear) {
    __shared__ uint32_t sForbidden[16];
    __shared__ uint32_t sCount;
    Launch L;
    if (!begin_substep(d, k, L)) return;
    const uint32_t lb = xcd_block();
    check_body<LAT, SLOTS, MAXDEG, KEYED>(d, L.t, L.draw, lb, gridDim.x, clear != 0u,
                                   (gptr<uint64_t>)d.list.global() + size_t(lb) * (SLOTS * blockDim.x),
                                   d.listCount.global() + lb, sForbidden, &sCount);
}

template <int LAT>
__global__ __launch_bounds__(256) void k_perform(SysDesc d, uint32_t k, uint32_t listStride, uint32_t clear) {
    Launch L;
    if (!begin_substep(d, k, L)) return;
    const uint32_t lb = xcd_block();  // same numbering as pass 1: list lb
    perform_body<LAT>(d, L, lb, gridDim.x, clear != 0u,
                      (gptr<const uint64_t>)d.list.global() + size_t(lb) * listStride, d.listCount.global()[lb],
                      L.acceptedByBlock + blockIdx.x);
}

// Pass 2 of a single system with several lists per workgroup: each group
// of `group` threads (whole waves) walks one pass-1 list, DEPTH entries per
// thread at a time.  Workgroup q runs on XCD q % 8 and takes lists of that
// XCD's share (the lists its pass-1 workgroups wrote, whose lines that XCD's
// L2 holds).  The list count and a thread's first entries do not depend on
// the state, so they load with it: the walk then waits for two round trips,
// the state with the list, then the proposal words of DEPTH moves at once.
template <int LAT, int DEPTH>
__global__ __launch_bounds__(256) void k_perform_lists(SysDesc d, uint32_t k, uint32_t listStride,
                                                       uint32_t listsPerXcd, uint32_t group, uint32_t clear) {
    const uint32_t xcd = blockIdx.x & 7u, local = blockIdx.x >> 3;
    const uint32_t gi = threadIdx.x / group, lane = threadIdx.x - gi * group;
    const uint32_t li = local * (blockDim.x / group) + gi;
    const bool haveList = li < listsPerXcd;
    const uint32_t lst = xcd * listsPerXcd + (haveList ? li : 0u);
    const gptr<const uint64_t> list = (gptr<const uint64_t>)d.list.global() + size_t(lst) * listStride;
    uint32_t count = d.listCount.global()[lst];
    uint64_t e[DEPTH];
#pragma unroll
    for (int r = 0; r < DEPTH; ++r) e[r] = list[min(lane + uint32_t(r) * group, listStride - 1u)];
    Launch L;
    if (!begin_substep(d, k, L)) return;
    if (!haveList) count = 0;
    const uint64_t t = L.t;
    const uint32_t slotOffset =
        ((gptr<const uint32_t>)(d.colors.global() + L.draw.species))[offsetof(ColorInfo, slotOffset) / 4];
    if (clear) clear_other_parity(d, t, blockIdx.x, gridDim.x);
    const AxisMasks masks = axis_masks(d);
    const gptr<const uint64_t> T64 =
        (gptr<const uint64_t>)(d.proposal.global() + size_t(t & 1u) * d.latWords);
    const gptr<uint32_t> pos = d.pos.global();
    const gptr<uint8_t> lat = d.lattice.global();

    uint32_t accepted = 0;
    for (uint32_t base = 0; base < count; base += DEPTH * group) {  // count is the same for a whole wave
        if (base != 0) {
#pragma unroll
            for (int r = 0; r < DEPTH; ++r) e[r] = list[min(base + lane + uint32_t(r) * group, listStride - 1u)];
        }
        // Entries past the count are stale: their lanes test the plane of
        // code 0 (in bounds) and drop the result.
        Move mv[DEPTH];
        uint32_t hit[DEPTH];
        bool valid[DEPTH];
#pragma unroll
        for (int r = 0; r < DEPTH; ++r) {
            valid[r] = base + lane + uint32_t(r) * group < count;
            const uint32_t m = valid[r] ? uint32_t(e[r] >> 32) : 0u;
            const uint32_t dir = valid[r] ? uint32_t(e[r]) >> kDirShift : 0u;
            mv[r] = make_move(m, dir, masks);
            hit[r] = plane_any<6>(T64, mv[r]);
        }
#pragma unroll
        for (int r = 0; r < DEPTH; ++r) {
            if (!valid[r] || hit[r]) continue;
            const uint32_t jj = uint32_t(e[r]) & ((1u << kDirShift) - 1u);
            apply_move<LAT>(d, pos, lat, slotOffset + jj, uint32_t(e[r] >> 32), uint32_t(e[r]) >> kDirShift, mv[r]);
            ++accepted;
        }
    }
    add_accepted(L, L.acceptedByBlock + blockIdx.x, accepted);
}

// ---- batch of replicas ---------------------------------------------------------------
// Workgroup w serves tiles[w]: chunk `chunk` of `chunks` of replica `replica`.
// Lists are indexed by w, accepted counts by (replica, chunk).
template <int LAT, int SLOTS, int MAXDEG>
__global__ __launch_bounds__(256) void k_check_batch(const SysDesc* __restrict__ descs, const Tile* __restrict__ tiles,
                                                     uint2* lists, uint32_t* listCounts, uint32_t k) {
    __shared__ uint32_t sForbidden[16];
    __shared__ uint32_t sCount;
    const Tile tile = tiles[blockIdx.x];
    if (tile.replica == kNoReplica) return;
    const SysDesc& d = descs[tile.replica];
    Launch L;
    if (!begin_substep(d, k, L)) return;
    const gptr<uint64_t> list = global_ptr(reinterpret_cast<uint64_t*>(lists)) + size_t(blockIdx.x) * (SLOTS * blockDim.x);
    // replicas of either slot order share a launch: one uniform branch per workgroup
    if (d.keyed != 0u)
        check_body<LAT, SLOTS, MAXDEG, true>(d, L.t, L.draw, tile.chunk, tile.chunks, false, list,
                                             global_ptr(listCounts) + blockIdx.x, sForbidden, &sCount);
    else
        check_body<LAT, SLOTS, MAXDEG, false>(d, L.t, L.draw, tile.chunk, tile.chunks, false, list,
                                              global_ptr(listCounts) + blockIdx.x, sForbidden, &sCount);
}

template <int LAT>
__global__ __launch_bounds__(256) void k_perform_batch(const SysDesc* __restrict__ descs,
                                                       const Tile* __restrict__ tiles, const uint2* lists,
                                                       const uint32_t* listCounts, uint32_t k, uint32_t listStride) {
    const Tile tile = tiles[blockIdx.x];
    if (tile.replica == kNoReplica) return;
    const SysDesc& d = descs[tile.replica];
    Launch L;
    if (!begin_substep(d, k, L)) return;
    perform_body<LAT>(d, L, tile.chunk, tile.chunks, true,
                      global_ptr(reinterpret_cast<const uint64_t*>(lists)) + size_t(blockIdx.x) * listStride,
                      global_ptr(listCounts)[blockIdx.x], L.acceptedByBlock + tile.chunk);
}

__global__ void k_advance(DevState* st, uint32_t count) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(&st->substep), (unsigned long long)count);
}

// Every live replica: substep = min(substep + count, stopAt).
__global__ void k_advance_batch(const SysDesc* descs, uint32_t nDescs, uint32_t count) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nDescs) return;
    const gptr<DevState> st = descs[r].state.global();
    if (st == nullptr) return;
    const uint64_t next = st->substep + count;
    st->substep = next < st->stopAt ? next : (st->substep > st->stopAt ? st->substep : st->stopAt);
}

__global__ void k_set_stops(const SysDesc* descs, uint32_t nDescs, const uint64_t* stops) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nDescs) return;
    const gptr<DevState> st = descs[r].state.global();
    if (st == nullptr) return;
    st->stopAt = stops[r];
}

// One workgroup per replica sums its per-workgroup accepted counters.
__global__ void k_gather_state(const SysDesc* descs, uint32_t nDescs, uint64_t* out) {
    __shared__ uint64_t part[256];
    const uint32_t r = blockIdx.x;
    if (r >= nDescs) return;
    const gptr<const DevState> st = descs[r].state.global();
    uint64_t sum = 0;
    if (st != nullptr)
        for (uint32_t i = threadIdx.x; i < descs[r].accEntries; i += blockDim.x) sum += global_ptr(st->acceptedByBlock)[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * r + 0] = st ? st->substep : 0;
        out[3 * r + 1] = st ? st->stopAt : 0;
        out[3 * r + 2] = part[0];
    }
}

// The bits of m selected by `mask`, packed from bit 0 up (mask is the same
// for every lane, so the loop is uniform).
__device__ __forceinline__ uint32_t compact_bits(uint32_t m, uint32_t mask) {
    uint32_t v = 0, k = 0;
    while (mask != 0u) {
        const uint32_t b = uint32_t(__builtin_ctz(mask));
        v |= ((m >> b) & 1u) << k;
        ++k;
        mask &= mask - 1u;
    }
    return v;
}

// Unfolded corner of the monomer in slot s: the code's axis bits plus the
// image counter times the box edge, written to out[0..2].
__device__ __forceinline__ void unfold_slot(gptr<const uint32_t> pos, const gptr<const int32_t> img[3],
                                            const uint32_t mask[3], uint32_t s, gptr<int32_t> out) {
    const uint32_t m = pos[s];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int32_t edge = int32_t(1u << __builtin_popcount(mask[a]));
        out[a] = int32_t(compact_bits(m, mask[a])) + img[a][s] * edge;
    }
}

__global__ void k_unfold(SysDesc d, const uint32_t* __restrict__ slotOf, uint32_t n, int32_t* __restrict__ xyz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const gptr<const int32_t> img[3] = {d.imgX.global(), d.imgY.global(), d.imgZ.global()};
    const uint32_t mask[3] = {d.mask[0], d.mask[1], d.mask[2]};
    unfold_slot(d.pos.global(), img, mask, global_ptr(slotOf)[i], global_ptr(xyz) + 3 * size_t(i));
}

// Several monomer ranges of several systems in one launch: workgroup b
// serves the last job whose firstBlock is at most b (jobs in order of
// firstBlock, the first at 0), monomers first + (b - firstBlock) * blockDim
// + threadIdx of that job, written as records out + ... of xyz.
__global__ __launch_bounds__(256) void k_unfold_jobs(const UnfoldJob* __restrict__ jobs, uint32_t nJobs,
                                                     int32_t* __restrict__ xyz) {
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = nJobs;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].firstBlock <= b) lo = mid;
        else hi = mid;
    }
    const UnfoldJob& J = jobs[lo];
    const uint32_t i = (b - J.firstBlock) * blockDim.x + threadIdx.x;
    if (i >= J.count) return;
    const gptr<const int32_t> img[3] = {global_ptr(J.img[0]), global_ptr(J.img[1]), global_ptr(J.img[2])};
    const uint32_t mask[3] = {J.mask[0], J.mask[1], J.mask[2]};
    unfold_slot(global_ptr(J.pos), img, mask, global_ptr(J.slotOf)[J.first + i],
                global_ptr(xyz) + 3 * (J.out + i));
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Plain streaming copy: the reference point that scales the HBM byte
// counters of a profiler run (it reads and writes exactly n*16 bytes).
__global__ void k_copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

// ---- template dispatch ------------------------------------------------------------------
template <int LAT, int SLOTS, int MAXDEG>
void check_launch(const SysDesc& d, const BatchLaunch* b, uint32_t k, const LaunchShape& s, hipStream_t st) {
    const dim3 blk(s.block);
    if (b)
        hipLaunchKernelGGL((k_check_batch<LAT, SLOTS, MAXDEG>), dim3(b->tiles), blk, 0, st, b->descs, b->tileTable,
                           b->lists, b->listCounts, k);
    else if (d.keyed != 0u)
        hipLaunchKernelGGL((k_check<LAT, SLOTS, MAXDEG, true>), dim3(s.grid), blk, 0, st, d, k,
                           uint32_t(s.clearPass == 1u));
    else
        hipLaunchKernelGGL((k_check<LAT, SLOTS, MAXDEG, false>), dim3(s.grid), blk, 0, st, d, k,
                           uint32_t(s.clearPass == 1u));
}

template <int LAT, int SLOTS>
void check_by_degree(const SysDesc& d, const BatchLaunch* b, uint32_t k, const LaunchShape& s, hipStream_t st) {
    const uint32_t deg = b ? b->maxDegree : d.maxDegree;
    if (deg <= 2)
        check_launch<LAT, SLOTS, 2>(d, b, k, s, st);
    else if (deg <= 4)
        check_launch<LAT, SLOTS, 4>(d, b, k, s, st);
    else
        check_launch<LAT, SLOTS, 8>(d, b, k, s, st);
}

template <int LAT>
void check_by_slots(const SysDesc& d, const BatchLaunch* b, uint32_t k, const LaunchShape& s, hipStream_t st) {
    switch (s.slots) {
        case 1: check_by_degree<LAT, 1>(d, b, k, s, st); break;
        case 2: check_by_degree<LAT, 2>(d, b, k, s, st); break;
        default: check_by_degree<LAT, 4>(d, b, k, s, st); break;
    }
}

template <int LAT>
void perform_launch(const SysDesc& d, const BatchLaunch* b, uint32_t k, const LaunchShape& s, hipStream_t st) {
    const dim3 blk(s.passTwoBlock());
    if (b)
        hipLaunchKernelGGL((k_perform_batch<LAT>), dim3(b->tiles), blk, 0, st, b->descs, b->tileTable, b->lists,
                           b->listCounts, k, s.listStride());
    else if (s.listGroup == 0)
        hipLaunchKernelGGL((k_perform<LAT>), dim3(s.grid), blk, 0, st, d, k, s.listStride(),
                           uint32_t(s.clearPass != 1u));
    else if (s.listDepth >= 4)
        hipLaunchKernelGGL((k_perform_lists<LAT, 4>), dim3(s.passTwoGrid()), blk, 0, st, d, k, s.listStride(),
                           s.listsPerXcd(), s.listGroup, uint32_t(s.clearPass != 1u));
    else
        hipLaunchKernelGGL((k_perform_lists<LAT, 1>), dim3(s.passTwoGrid()), blk, 0, st, d, k, s.listStride(),
                           s.listsPerXcd(), s.listGroup, uint32_t(s.clearPass != 1u));
}

}  // namespace

hipError_t launch_check(const SysDesc& d, uint32_t k, const LaunchShape& s, hipStream_t st) {
    if (d.latticeKind == kLatticeU8)
        check_by_slots<kLatticeU8>(d, nullptr, k, s, st);
    else
        check_by_slots<kLatticeBit>(d, nullptr, k, s, st);
    return hipGetLastError();
}

hipError_t launch_perform(const SysDesc& d, uint32_t k, const LaunchShape& s, hipStream_t st) {
    if (d.latticeKind == kLatticeU8)
        perform_launch<kLatticeU8>(d, nullptr, k, s, st);
    else
        perform_launch<kLatticeBit>(d, nullptr, k, s, st);
    return hipGetLastError();
}

hipError_t launch_check_batch(const BatchLaunch& b, uint32_t k, const LaunchShape& s, hipStream_t st) {
    const SysDesc none{};
    if (b.latticeKind == kLatticeU8)
        check_by_slots<kLatticeU8>(none, &b, k, s, st);
    else
        check_by_slots<kLatticeBit>(none, &b, k, s, st);
    return hipGetLastError();
}

hipError_t launch_perform_batch(const BatchLaunch& b, uint32_t k, const LaunchShape& s, hipStream_t st) {
    const SysDesc none{};
    if (b.latticeKind == kLatticeU8)
        perform_launch<kLatticeU8>(none, &b, k, s, st);
    else
        perform_launch<kLatticeBit>(none, &b, k, s, st);
    return hipGetLastError();
}

hipError_t launch_advance(DevState* state, uint32_t count, hipStream_t st) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, state, count);
    return hipGetLastError();
}

hipError_t launch_advance_batch(const SysDesc* descs, uint32_t nDescs, uint32_t count, hipStream_t st) {
    if (nDescs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_advance_batch, dim3((nDescs + 255) / 256), dim3(256), 0, st, descs, nDescs, count);
    return hipGetLastError();
}

hipError_t launch_set_stops(const SysDesc* descs, uint32_t nDescs, const uint64_t* stops, hipStream_t st) {
    if (nDescs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_stops, dim3((nDescs + 255) / 256), dim3(256), 0, st, descs, nDescs, stops);
    return hipGetLastError();
}

hipError_t launch_gather_state(const SysDesc* descs, uint32_t nDescs, uint64_t* out, hipStream_t st) {
    if (nDescs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_state, dim3(nDescs), dim3(256), 0, st, descs, nDescs, out);
    return hipGetLastError();
}

hipError_t launch_unfold(const SysDesc& d, const uint32_t* slotOf, uint32_t n, int32_t* xyz, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_unfold, dim3((n + 255) / 256), dim3(256), 0, st, d, slotOf, n, xyz);
    return hipGetLastError();
}

hipError_t launch_unfold_jobs(const UnfoldJob* jobs, uint32_t nJobs, uint32_t blocks, int32_t* xyz, hipStream_t st) {
    if (nJobs == 0 || blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_unfold_jobs, dim3(blocks), dim3(kUnfoldBlock), 0, st, jobs, nJobs, xyz);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3(uint32_t(g)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_copy_u4(const void* src, void* dst, uint64_t n16, hipStream_t st) {
    if (n16 == 0) return hipSuccess;
    uint64_t g = (n16 + 255) / 256;
    if (g > 1u << 20) g = 1u << 20;
    hipLaunchKernelGGL(k_copy_u4, dim3(uint32_t(g)), dim3(256), 0, st, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n16);
    return hipGetLastError();
}

}  // namespace pscbfm
