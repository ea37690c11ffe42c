// sp_kernels_provider.hip -- the device-side particle provider (provide_particle and particle_params of
// sp_provider.h, one thread per particle).
//
// Work per particle: up to ten Philox blocks and Box-Muller normals, a few transcendentals, the parameter
// calculators, and (NPARAM + dy) doubles written once, coalesced (param-major rows, particle fastest).
#include <hip/hip_runtime.h>

#include "sp_kernels.h"
#include "sp_provider.h"

namespace sp {

__global__ __launch_bounds__(PROV_BLOCK) void k_provide(ProvDesc d) {
    const uint64_t i = (uint64_t)blockIdx.x * PROV_BLOCK + threadIdx.x;
    double v[PROV_NSUM];
#pragma unroll
    for (int k = 0; k < PROV_NSUM; ++k) v[k] = 0.0;
    if (i < d.N) {
        ParticleProps q;
        provide_particle(d, d.types, i, q);
        double par[NPARAM];
        particle_params(d.problem, d.aniso, q, par);
#pragma unroll
        for (int k = 0; k < NPARAM; ++k) d.params[(uint64_t)k * d.N + i] = par[k];
#pragma unroll
        for (uint32_t k = 0; k < 6; ++k)       // constant trip count: q stays in registers
            if (k < d.dy) d.y0[(uint64_t)k * d.N + i] = q.y0[k];
        v[0] = q.T; v[1] = q.radius; v[2] = q.Ms; v[3] = q.alpha; v[4] = q.gamma; v[5] = q.K; v[6] = q.r_h;
        v[7] = q.visc; v[8] = par[P_MU]; v[9] = i < d.n_check ? par[P_MU] : 0.0;
    }
    // fixed-order work-group sums: a butterfly within each wave, then the waves in order
    __shared__ double part[PROV_BLOCK / 64][PROV_NSUM];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < PROV_NSUM; ++k) {
        double x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) part[wave][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < PROV_NSUM) {
        double a = 0.0;
#pragma unroll
        for (uint32_t w = 0; w < PROV_BLOCK / 64; ++w) a += part[w][threadIdx.x];
        d.sums[(uint64_t)blockIdx.x * PROV_NSUM + threadIdx.x] = a;
    }
}

void launch_provide(hipStream_t s, uint32_t blocks, const ProvDesc& d) {
    if (blocks) hipLaunchKernelGGL(k_provide, dim3(blocks), dim3(PROV_BLOCK), 0, s, d);
}

}  // namespace sp
