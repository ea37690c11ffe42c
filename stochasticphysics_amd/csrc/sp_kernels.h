// sp_kernels.h -- host-side entry points of the compiled kernel classes.
#pragma once
#include <hip/hip_runtime.h>
#include "sp_provider.h"
#include "sp_types.h"

namespace sp {

struct KernelInfo {
    int dy, dw, dout, group;   // state, noise and output sizes; steps per noise group
    int vgprs, sgprs;          // register use of the per-request kernel
    const char* name;
};

// One request over its tiles (grid = one wave per 64-particle tile).
bool launch_plain(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq,
                  uint64_t step_begin, uint64_t step_end, double* partial, uint32_t pstride);
// Persistent: waves pull global tile indices from an atomic counter until all tiles of the
// launch's slots are taken.
bool launch_batch(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                  const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter);
bool kernel_info(uint32_t kind, KernelInfo* info);
// SIMDs of the current device (four per compute unit; queried once per device)
uint32_t device_simds();
// noise waves per tile of a plain launch of `tiles` tiles that runs split (k_plain_split: noise waves beside each
// integrator wave): 2 when the launch leaves two thirds of the device's SIMDs idle, 1 when it leaves half, else 0
// (not split); SP_PLAIN_SPLIT=0 in the environment turns the split kernel off
int split_noise_waves(uint32_t tiles);

// per-problem registries (one translation unit each)
bool launch_plain_neel(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, uint64_t, uint64_t, double*, uint32_t);
bool launch_plain_sph(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, uint64_t, uint64_t, double*, uint32_t);
bool launch_plain_bn(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, uint64_t, uint64_t, double*, uint32_t);
bool launch_batch_neel(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, const SlotWork*, uint32_t, uint32_t, uint32_t*);
bool launch_batch_sph(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, const SlotWork*, uint32_t, uint32_t, uint32_t*);
bool launch_batch_bn(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, const SlotWork*, uint32_t, uint32_t, uint32_t*);
bool info_neel(uint32_t, KernelInfo*);
bool info_sph(uint32_t, KernelInfo*);
// the FP64 explicit NeelSpherical classes, a translation unit of their own (launch_plain_sph hands them on)
bool launch_plain_sph64(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, uint64_t, uint64_t, double*, uint32_t);
bool launch_batch_sph64(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, const SlotWork*, uint32_t, uint32_t, uint32_t*);
bool info_sph64(uint32_t, KernelInfo*);
// one-wave builds of the explicit NeelSpherical classes (false for a kind it does not hold)
bool launch_plain_sph1w(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, uint64_t, uint64_t, double*, uint32_t);
bool launch_batch_sph1w(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, const SlotWork*, uint32_t, uint32_t, uint32_t*);
// the FP64 and ImplicitMidpoint BrownAndNeel classes, a translation unit of their own (launch_plain_bn hands them on)
bool launch_plain_bn64(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, uint64_t, uint64_t, double*, uint32_t);
bool launch_batch_bn64(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, const SlotWork*, uint32_t, uint32_t, uint32_t*);
bool info_bn64(uint32_t, KernelInfo*);
bool info_bn(uint32_t, KernelInfo*);
bool launch_plain_uc(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, uint64_t, uint64_t, double*, uint32_t);
bool launch_batch_uc(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, const SlotWork*, uint32_t, uint32_t, uint32_t*);
bool info_uc(uint32_t, KernelInfo*);
#define SP_BNES_DECL(A)                                                                                               \
    bool launch_plain_bnes_a##A(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, uint64_t, uint64_t, double*, uint32_t); \
    bool launch_batch_bnes_a##A(uint32_t, dim3, dim3, hipStream_t, const DevRequest*, const SlotWork*, uint32_t, uint32_t,   \
                                uint32_t*);                                                                               \
    bool info_bnes_a##A(uint32_t, KernelInfo*);
SP_BNES_DECL(1)
SP_BNES_DECL(2)
SP_BNES_DECL(4)
#undef SP_BNES_DECL

// the provider kernel: particles [0, d.N) drawn into d.params / d.y0, per-group sums into d.sums
// (blocks = ceil(N / PROV_BLOCK))
void launch_provide(hipStream_t s, uint32_t blocks, const ProvDesc& d);
// deterministic second-stage reduction; one block per (entry, row)
void launch_reduce(hipStream_t s, const ReduceWork* work, uint32_t n_work, uint32_t max_rows);
// zero n counters (a kernel, so that it can sit in a captured graph)
void launch_zero_u32(hipStream_t s, uint32_t* p, uint32_t n);
// test hooks
void launch_philox_test(hipStream_t s, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                        uint64_t count, uint32_t* out);
void launch_normals_test(hipStream_t s, uint64_t seed, uint64_t particle, uint64_t first, uint64_t count, int mixed,
                         double* out);
void launch_box_muller_test(hipStream_t s, const uint32_t* words, uint64_t n_blocks, int mixed, double* out);

}  // namespace sp
