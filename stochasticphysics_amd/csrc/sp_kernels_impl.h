// sp_kernels_impl.h -- kernel templates; included by the per-problem translation units.
#pragma once
#include "sp_device.h"
#include "sp_kernels.h"

namespace sp {

// Waves per SIMD a kernel class is compiled for (the register budget per wave is the SIMD's 512 registers
// divided by it): WAVES_PER_SIMD of the class, 1 by default.
template <class Cfg, class = void> struct waves_per_simd { static constexpr int value = 1; };
template <class Cfg> struct waves_per_simd<Cfg, std::void_t<decltype(Cfg::WAVES_PER_SIMD)>> {
    static constexpr int value = Cfg::WAVES_PER_SIMD;
};
// A kernel class compiled for one wave per SIMD (all registers of the SIMD to one wave): the plain launch takes it
// when the launch has no more waves than the device has SIMDs, where a second wave per SIMD could not be resident
// anyway and the smaller register budget of the class's own build would only cost spills.
template <class Cfg> struct OneWave : Cfg { static constexpr int WAVES_PER_SIMD = 1; };

// One wave per 64-particle tile of one request.
template <class Cfg, class R, bool KAHAN, bool MID>
__global__ __launch_bounds__(MAX_BLOCK) __attribute__((amdgpu_waves_per_eu(waves_per_simd<Cfg>::value))) void k_plain(const DevRequest* __restrict__ rqp, uint64_t step_begin, uint64_t step_end,
                                               double* partial, uint32_t pstride) {
    const DevRequest& rq = *rqp;
    bm_lds_fill();       // every thread, before any wave leaves
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (wave >= rq.n_tiles) return;
    const TileWork tw{0u, wave, step_begin, step_end, partial, pstride, wave};
    run_tile<Cfg, R, KAHAN, MID>(rq, tw);
}

// One 64-particle tile per work-group of 1 + NW waves: waves 1..NW compute the tile's normals into an LDS ring,
// wave 0 integrates (run_tile's split mode).  For launches that leave most SIMDs idle: the noise work moves to SIMDs
// of its own.  Tiles that do not take the fast path (replayed noise, windows not ending on group boundaries) run in
// wave 0 alone.  All waves of a block take every decision below from the same uniform values, so they reach the
// same barriers.
template <class Cfg, class R, bool KAHAN, bool MID, int NW>
__global__ __launch_bounds__(64 * (1 + NW)) void k_plain_split(const DevRequest* __restrict__ rqp, uint64_t step_begin,
                                                               uint64_t step_end, double* partial, uint32_t pstride) {
    __shared__ NoisePack<R> ring[split_ring_packs<R, NW>()];
    const DevRequest& rq = *rqp;
    bm_lds_fill();       // every thread, before any wave leaves
    const uint32_t tile = blockIdx.x;
    if (tile >= __builtin_amdgcn_readfirstlane(rq.n_tiles)) return;    // uniform over the block
    const TileWork tw{0u, tile, step_begin, step_end, partial, pstride, tile};
    const int w = (int)(threadIdx.x >> 6);
    if (w > 0) {
        if (fast_path<Cfg>(rq)) produce_noise<Cfg, R, NW>(rq, tw, ring, w - 1);
        return;
    }
    run_tile<Cfg, R, KAHAN, MID, NW>(rq, tw, ring);
}
// a plain launch of `tiles` tiles, split when it leaves at least half of the SIMDs idle (two noise waves per tile
// when it leaves two thirds)
template <class Cfg, class R, bool KAHAN, bool MID>
bool launch_split(uint32_t tiles, hipStream_t s, const DevRequest* rq, uint64_t sb, uint64_t se, double* partial,
                  uint32_t pstride) {
    const int nw = split_noise_waves(tiles);
    if (nw == 2) k_plain_split<Cfg, R, KAHAN, MID, 2><<<dim3(tiles), dim3(192), 0, s>>>(rq, sb, se, partial, pstride);
    else if (nw == 1) k_plain_split<Cfg, R, KAHAN, MID, 1><<<dim3(tiles), dim3(128), 0, s>>>(rq, sb, se, partial, pstride);
    else return false;
    return true;
}

__device__ __forceinline__ uint32_t next_work(uint32_t* counter) {
    uint32_t idx = 0;
    if ((threadIdx.x & 63) == 0) idx = atomicAdd(counter, 1u);
    return __builtin_amdgcn_readfirstlane(__shfl(idx, 0, 64));
}

// slot owning global tile idx: last s with tile_base <= idx (slots sorted by tile_base)
__device__ __forceinline__ uint32_t find_slot(const SlotWork* __restrict__ slots, uint32_t n_slots, uint32_t idx) {
    uint32_t lo = 0, hi = n_slots;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (slots[mid].tile_base <= idx) lo = mid; else hi = mid;
    }
    return lo;
}

// Persistent: every wave pulls global tile indices from the counter until all
// n_tiles tiles of the launch are taken; every wave reaches that exit.
template <class Cfg, class R, bool KAHAN, bool MID>
__global__ __launch_bounds__(MAX_BLOCK) __attribute__((amdgpu_waves_per_eu(waves_per_simd<Cfg>::value))) void k_batch(const DevRequest* __restrict__ reqs, const SlotWork* __restrict__ slots,
                                               uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    bm_lds_fill();
    for (;;) {
        const uint32_t idx = next_work(counter);
        if (idx >= n_tiles) break;
        const SlotWork sw = slots[find_slot(slots, n_slots, idx)];
        const uint32_t tile = idx - sw.tile_base;
        const TileWork tw{sw.req, tile, sw.step_begin, sw.step_end, sw.partial, sw.n_tiles, tile};
        run_tile<Cfg, R, KAHAN, MID>(reqs[sw.req], tw);
    }
}

// Fixed-order reduction of one ReduceWork entry by one 256-thread block: thread t sums slots t, t+256, ... in
// order, then a fixed LDS tree; the result depends only on the partial values, not on how the tiles were
// scheduled.  The block reduces rows blockIdx.x, blockIdx.x + gridDim.x, ... (the row index is uniform over the
// block, so every thread reaches every barrier).
__device__ __forceinline__ void reduce_rows(const ReduceWork& rw) {
    __shared__ double sh[256];
    for (uint32_t r = blockIdx.x; r < rw.n_rows; r += gridDim.x) {
        const double* row = rw.partial + (uint64_t)r * rw.pstride;
        double s = 0.0;
        for (uint32_t i = threadIdx.x; i < rw.pstride; i += 256) s += row[i];
        // the check tiles are kept one by one (n_check <= 64 <= pstride)
        if (rw.check_out && threadIdx.x < rw.n_check)
            rw.check_out[(uint64_t)(rw.row_base + r) * rw.n_check + threadIdx.x] = row[threadIdx.x];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (uint32_t w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) rw.out[rw.row_base + r] = sh[0];
        __syncthreads();     // sh is written again by the next row
    }
}

template <class Cfg, class R, bool KAHAN, bool MID>
bool fill_info(KernelInfo* info, const char* name) {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, (const void*)k_plain<Cfg, R, KAHAN, MID>) != hipSuccess) return false;
    info->dy = Cfg::DY; info->dw = Cfg::DW; info->dout = Cfg::DOUT; info->group = Cfg::G;
    info->vgprs = a.numRegs; info->sgprs = 0; info->name = name;
    return true;
}

}  // namespace sp

// expands the per-class switch cases for one (problem, solver, aniso, mixed) class
// the plain launch of a class that also has its one-wave build (OneWave)
#define SP_CASE_PLAIN_FIT(KIND, CFG, R, KAHAN, MID)                                                    \
    case KIND:                                                                                           \
        if (launch_split<CFG, R, KAHAN, MID>(grid.x * (block.x / 64), s, rq, sb, se, partial, pstride))  \
            ;                                                                                            \
        else if ((uint64_t)grid.x * block.x <= 64ull * device_simds())                                   \
            k_plain<OneWave<CFG>, R, KAHAN, MID><<<grid, block, 0, s>>>(rq, sb, se, partial, pstride);  \
        else                                                                                             \
            k_plain<CFG, R, KAHAN, MID><<<grid, block, 0, s>>>(rq, sb, se, partial, pstride);            \
        return true;
// the plain launch of a class with the split-noise kernel for small launches
#define SP_CASE_PLAIN_SPLIT(KIND, CFG, R, KAHAN, MID)                                                  \
    case KIND:                                                                                           \
        if (!launch_split<CFG, R, KAHAN, MID>(grid.x * (block.x / 64), s, rq, sb, se, partial, pstride)) \
            k_plain<CFG, R, KAHAN, MID><<<grid, block, 0, s>>>(rq, sb, se, partial, pstride);            \
        return true;
#define SP_CASE_PLAIN(KIND, CFG, R, KAHAN, MID)                                                        \
    case KIND:                                                                                           \
        k_plain<CFG, R, KAHAN, MID><<<grid, block, 0, s>>>(rq, sb, se, partial, pstride);              \
        return true;
#define SP_CASE_BATCH(KIND, CFG, R, KAHAN, MID)                                                        \
    case KIND:                                                                                           \
        k_batch<CFG, R, KAHAN, MID><<<grid, block, 0, s>>>(reqs, slots, n_slots, n_tiles, counter);    \
        return true;
#define SP_CASE_INFO(KIND, CFG, R, KAHAN, MID) \
    case KIND: return fill_info<CFG, R, KAHAN, MID>(info, #CFG);
