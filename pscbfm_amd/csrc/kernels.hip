// gfx950 kernels of the colored two-pass scBFM sweep.
//
// Pass 1 (check):   CRITICAL.cu:298-374  kernelSimulationScBFMCheckSpecies
// Pass 2 (perform): CRITICAL.cu:376-411  kernelSimulationScBFMPerformSpeciesAndApply
// Proposal clear:   CRITICAL.cu:561-573  (here folded into the next pass 1)
//
// Design (wave64, CDNA4):
//  * positions are Z-curve codes, so the own lattice cell is the code itself
//    and every neighbouring cell is a masked add on one axis' bits;
//  * the forbidden-bond table is a 512-bit mask in LDS, indexed by the
//    Z-order bond key of two codes;
//  * the proposal lattice is 1 bit per cell and double-buffered by substep
//    parity: pass 1 of substep t sets bits in buffer t&1 and zeroes buffer
//    (t+1)&1, which pass 2 of substep t-1 was the last to read;
//  * workgroup w runs on XCD w%8, so logical blocks are renumbered to give
//    each XCD one contiguous eighth of the species (its own L2 then holds the
//    lattice region of that eighth);
//  * the species and kernel seed of a substep are drawn on the device from
//    the substep counter, so launches carry no per-substep host data.
#include "kernels.h"

namespace pscbfm {

namespace {

// Logical workgroup id: consecutive ids on the same XCD.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x;  // multiple of 8 (grid_for)
    return (blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3);
}

struct Move {
    uint32_t to;        // code of the new corner
    uint32_t cell[9];   // the 3x3 plane in front of the move
    uint32_t axisMask;
    bool wraps;         // the axis coordinate crosses the periodic boundary
};

// Direction d: 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z (UpdaterGPUScBFM_AB_Type.cu:534-545).
// Front plane: p + 2 e_d + a e_u + b e_v, a,b in {-1,0,1} (CRITICAL.cu:62-176).
__device__ __forceinline__ Move make_move(uint32_t m, uint32_t dir, const uint32_t mask[3]) {
    const uint32_t axis = dir >> 1;
    const bool plus = (dir & 1u) != 0;
    const uint32_t Ma = axis == 0 ? mask[0] : (axis == 1 ? mask[1] : mask[2]);
    const uint32_t Mu = axis == 0 ? mask[1] : mask[0];
    const uint32_t Mv = axis == 2 ? mask[1] : mask[2];
    const uint32_t oneA = Ma & (~Ma + 1u);
    const uint32_t stepA = plus ? oneA : Ma;  // +1 or -1 in dilated form
    Move mv;
    const uint32_t qa = dil_add(m, Ma, stepA);
    mv.to = (m & ~Ma) | qa;
    mv.axisMask = Ma;
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

template <int LAT>
__device__ __forceinline__ bool front_occupied(const void* lattice, const Move& mv) {
    uint32_t any = 0;
    if (LAT == kLatticeU8) {
        const uint8_t* lat = static_cast<const uint8_t*>(lattice);
#pragma unroll
        for (int c = 0; c < 9; ++c) any |= lat[mv.cell[c]];
    } else {
        const uint32_t* lat = static_cast<const uint32_t*>(lattice);
#pragma unroll
        for (int c = 0; c < 9; ++c) any |= lat[mv.cell[c] >> 5] >> (mv.cell[c] & 31u);
        any &= 1u;
    }
    return any != 0u;
}

__device__ __forceinline__ bool front_proposed(const uint32_t* T, const Move& mv) {
    uint32_t any = 0;
#pragma unroll
    for (int c = 0; c < 9; ++c) any |= T[mv.cell[c] >> 5] >> (mv.cell[c] & 31u);
    return (any & 1u) != 0u;
}

template <int LAT>
__global__ __launch_bounds__(1024) void k_check(SysDesc d, uint32_t k) {
    __shared__ uint32_t sForbidden[16];
    const DevState* st = d.state;
    const uint64_t t = st->substep + k;
    if (t >= st->stopAt) return;
    const SubstepDraw draw = substep_draw(d.seed, t, d.nColors);
    const ColorInfo ci = d.colors[draw.species];
    const uint32_t lb = xcd_block();
    const uint32_t nb = gridDim.x;

    // Zero the proposal buffer of the other parity, one slice per workgroup.
    {
        uint4* prev = reinterpret_cast<uint4*>(d.proposal + size_t((t + 1u) & 1u) * d.latWords);
        const uint32_t quads = d.latWords >> 2;
        const uint32_t per = (quads + nb - 1) / nb;
        const uint32_t q0 = lb * per;
        const uint32_t q1 = min(q0 + per, quads);
        for (uint32_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) prev[q] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (threadIdx.x < 16) sForbidden[threadIdx.x] = d.bondMask[threadIdx.x];
    __syncthreads();

    uint32_t* T = d.proposal + size_t(t & 1u) * d.latWords;
    const uint32_t* nbr = d.nbr + ci.nbrOffset;
    for (uint32_t j = lb * blockDim.x + threadIdx.x; j < ci.count; j += nb * blockDim.x) {
        const uint32_t slot = ci.slotOffset + j;
        const uint32_t m = d.pos[slot];
        // CRITICAL.cu:322-323: the draw is keyed by the species-local index.
        const uint32_t dir = move_hash(move_hash(j) ^ draw.seed) % 6u;
        const Move mv = make_move(m, dir, d.mask);
        bool ok = true;
        for (uint32_t r = 0; r < ci.maxDegree; ++r) {
            const uint32_t partner = nbr[size_t(r) * ci.nbrPitch + j];
            if (partner != kNoNeighbor) {
                const uint32_t key = bond_key(d.pos[partner], mv.to);
                ok = ok && !((sForbidden[key >> 5] >> (key & 31u)) & 1u);
            }
        }
        if (ok) ok = !front_occupied<LAT>(d.lattice, mv);
        uint8_t verdict = 0;
        if (ok) {
            atomicOr(&T[mv.to >> 5], 1u << (mv.to & 31u));
            verdict = uint8_t(1u + dir);
        }
        d.flags[slot] = verdict;
    }
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* scratch) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    uint32_t total = 0;
    if (threadIdx.x == 0)
        for (uint32_t w = 0; w < (blockDim.x + 63u) >> 6; ++w) total += scratch[w];
    return total;
}

template <int LAT>
__global__ __launch_bounds__(1024) void k_perform(SysDesc d, uint32_t k) {
    __shared__ uint32_t sSum[16];
    DevState* st = d.state;
    const uint64_t t = st->substep + k;
    if (t >= st->stopAt) return;
    const SubstepDraw draw = substep_draw(d.seed, t, d.nColors);
    const ColorInfo ci = d.colors[draw.species];
    const uint32_t lb = xcd_block();
    const uint32_t nb = gridDim.x;
    const uint32_t* T = d.proposal + size_t(t & 1u) * d.latWords;

    uint32_t accepted = 0;
    for (uint32_t j = lb * blockDim.x + threadIdx.x; j < ci.count; j += nb * blockDim.x) {
        const uint32_t slot = ci.slotOffset + j;
        const uint32_t verdict = d.flags[slot];
        if (verdict == 0u) continue;
        const uint32_t dir = verdict - 1u;
        const uint32_t m = d.pos[slot];
        const Move mv = make_move(m, dir, d.mask);
        if (front_proposed(T, mv)) continue;
        d.pos[slot] = mv.to;
        if (LAT == kLatticeU8) {
            uint8_t* lat = static_cast<uint8_t*>(d.lattice);
            lat[m] = 0;
            lat[mv.to] = 1;
        } else {
            uint32_t* lat = static_cast<uint32_t*>(d.lattice);
            const uint32_t wFrom = m >> 5, wTo = mv.to >> 5;
            const uint32_t bFrom = 1u << (m & 31u), bTo = 1u << (mv.to & 31u);
            if (wFrom == wTo) {
                atomicXor(&lat[wFrom], bFrom | bTo);
            } else {
                atomicXor(&lat[wFrom], bFrom);
                atomicXor(&lat[wTo], bTo);
            }
        }
        if (mv.wraps) {
            const uint32_t axis = dir >> 1;
            int32_t* img = axis == 0 ? d.imgX : (axis == 1 ? d.imgY : d.imgZ);
            img[slot] += (dir & 1u) ? 1 : -1;
        }
        ++accepted;
    }
    const uint32_t total = block_sum(accepted, sSum);
    if (threadIdx.x == 0 && total != 0u) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&st->accepted), (unsigned long long)total);
        if (st->log != nullptr) {
            const uint64_t idx = t - st->logFirst;
            if (idx < st->logCapacity) atomicAdd(&st->log[idx], total);
        }
    }
}

__global__ void k_advance(DevState* st, uint32_t count) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(&st->substep), (unsigned long long)count);
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

}  // namespace

hipError_t launch_check(const SysDesc& d, uint32_t k, const LaunchShape& s, hipStream_t st) {
    if (d.latticeKind == kLatticeU8)
        hipLaunchKernelGGL(k_check<kLatticeU8>, dim3(s.checkGrid), dim3(s.block), 0, st, d, k);
    else
        hipLaunchKernelGGL(k_check<kLatticeBit>, dim3(s.checkGrid), dim3(s.block), 0, st, d, k);
    return hipGetLastError();
}

hipError_t launch_perform(const SysDesc& d, uint32_t k, const LaunchShape& s, hipStream_t st) {
    if (d.latticeKind == kLatticeU8)
        hipLaunchKernelGGL(k_perform<kLatticeU8>, dim3(s.performGrid), dim3(s.block), 0, st, d, k);
    else
        hipLaunchKernelGGL(k_perform<kLatticeBit>, dim3(s.performGrid), dim3(s.block), 0, st, d, k);
    return hipGetLastError();
}

hipError_t launch_advance(DevState* state, uint32_t count, hipStream_t st) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, state, count);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3(uint32_t(g)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

}  // namespace pscbfm
