// sp_kernels_common.hip -- reduction, dispatch over the per-problem registries, test hooks.
#include "sp_kernels_impl.h"

#include <atomic>

namespace sp {

uint32_t device_simds() {
    static std::atomic<uint32_t> simds[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return 1024;                                   // an MI355X: 256 compute units
    }
    uint32_t v = simds[dev].load(std::memory_order_relaxed);
    if (v == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
            (void)hipGetLastError();
            cus = 256;
        }
        v = 4u * (uint32_t)cus;
        simds[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

int split_noise_waves(uint32_t tiles) {
    static const bool enabled = [] {
        const char* e = getenv("SP_PLAIN_SPLIT");
        return !(e && e[0] == '0');
    }();
    if (!enabled || tiles == 0) return 0;
    const uint64_t simds = device_simds();
    if (3ull * tiles <= simds) return 2;
    if (2ull * tiles <= simds) return 1;
    return 0;
}

bool launch_plain(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq, uint64_t sb,
                  uint64_t se, double* partial, uint32_t pstride) {
    if (kind_is_uniaxialcubic(kind) && kind / 128 <= 2) return launch_plain_uc(kind, grid, block, s, rq, sb, se, partial, pstride);
    switch (kind / 128) {
        case 1: return launch_plain_neel(kind, grid, block, s, rq, sb, se, partial, pstride);
        case 2: return launch_plain_sph(kind, grid, block, s, rq, sb, se, partial, pstride);
        case 3: return launch_plain_bn(kind, grid, block, s, rq, sb, se, partial, pstride);
        case 4:
            switch ((kind >> 1) & 3u) {   // anisotropy slot
                case 0: return launch_plain_bnes_a1(kind, grid, block, s, rq, sb, se, partial, pstride);
                case 1: return launch_plain_bnes_a2(kind, grid, block, s, rq, sb, se, partial, pstride);
                case 2: return launch_plain_bnes_a4(kind, grid, block, s, rq, sb, se, partial, pstride);
            }
            return false;
    }
    return false;
}
bool launch_batch(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs, const SlotWork* slots,
                  uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    if (kind_is_uniaxialcubic(kind) && kind / 128 <= 2) return launch_batch_uc(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter);
    switch (kind / 128) {
        case 1: return launch_batch_neel(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter);
        case 2: return launch_batch_sph(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter);
        case 3: return launch_batch_bn(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter);
        case 4:
            switch ((kind >> 1) & 3u) {
                case 0: return launch_batch_bnes_a1(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter);
                case 1: return launch_batch_bnes_a2(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter);
                case 2: return launch_batch_bnes_a4(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter);
            }
            return false;
    }
    return false;
}
bool kernel_info(uint32_t kind, KernelInfo* info) {
    if (kind_is_uniaxialcubic(kind) && kind / 128 <= 2) return info_uc(kind, info);
    switch (kind / 128) {
        case 1: return info_neel(kind, info);
        case 2: return info_sph(kind, info);
        case 3: return info_bn(kind, info);
        case 4:
            switch ((kind >> 1) & 3u) {
                case 0: return info_bnes_a1(kind, info);
                case 1: return info_bnes_a2(kind, info);
                case 2: return info_bnes_a4(kind, info);
            }
            return false;
    }
    return false;
}

__global__ __launch_bounds__(256) void k_reduce(const ReduceWork* __restrict__ work) {
    reduce_rows(work[blockIdx.y]);
}

void launch_reduce(hipStream_t s, const ReduceWork* work, uint32_t n_work, uint32_t max_rows) {
    if (!n_work || !max_rows) return;
    k_reduce<<<dim3(reduce_grid_x(max_rows), n_work), dim3(256), 0, s>>>(work);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
void launch_zero_u32(hipStream_t s, uint32_t* p, uint32_t n) {
    if (!n) return;
    k_zero_u32<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(p, n);
}

__global__ void k_philox_test(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                              uint64_t count, uint32_t* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const U4 r = philox4x32_10(U4{c0 + (uint32_t)i, c1, c2, c3}, k0, k1);
    out[4 * i + 0] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
}
void launch_philox_test(hipStream_t s, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                        uint64_t count, uint32_t* out) {
    if (!count) return;
    k_philox_test<<<dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s>>>(c0, c1, c2, c3, k0, k1, count, out);
}

// normal k of particle p's stream, one thread per Philox block
__global__ void k_normals_test(uint64_t seed, uint64_t particle, uint64_t first, uint64_t count, int mixed, double* out) {
    const uint64_t b = (first >> 2) + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t last_block = (first + count - 1) >> 2;
    if (b > last_block) return;
    const U4 w = noise_block(seed, particle, b, 0u);
    double z[4];
    if (mixed) {
        float zf[4];
        box_muller4_f32(w, zf);
        for (int j = 0; j < 4; ++j) z[j] = (double)zf[j];
    } else {
        box_muller4_f64(w, z);
    }
    for (int j = 0; j < 4; ++j) {
        const uint64_t k = 4 * b + j;
        if (k >= first && k < first + count) out[k - first] = z[j];
    }
}
void launch_normals_test(hipStream_t s, uint64_t seed, uint64_t particle, uint64_t first, uint64_t count, int mixed,
                         double* out) {
    if (!count) return;
    const uint64_t blocks = ((first + count - 1) >> 2) - (first >> 2) + 1;
    k_normals_test<<<dim3((unsigned)((blocks + 255) / 256)), dim3(256), 0, s>>>(seed, particle, first, count, mixed, out);
}

// Box-Muller of caller-given Philox words: out[4i..4i+3] from words[4i..4i+3]
__global__ void k_box_muller_test(const uint32_t* words, uint64_t n_blocks, int mixed, double* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_blocks) return;
    const U4 w{words[4 * i], words[4 * i + 1], words[4 * i + 2], words[4 * i + 3]};
    if (mixed) {
        float zf[4];
        box_muller4_f32(w, zf);
        for (int j = 0; j < 4; ++j) out[4 * i + j] = (double)zf[j];
    } else {
        double z[4];
        box_muller4_f64(w, z);
        for (int j = 0; j < 4; ++j) out[4 * i + j] = z[j];
    }
}
void launch_box_muller_test(hipStream_t s, const uint32_t* words, uint64_t n_blocks, int mixed, double* out) {
    if (!n_blocks) return;
    k_box_muller_test<<<dim3((unsigned)((n_blocks + 255) / 256)), dim3(256), 0, s>>>(words, n_blocks, mixed, out);
}

}  // namespace sp
