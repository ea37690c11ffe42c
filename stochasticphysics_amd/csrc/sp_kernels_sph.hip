// sp_kernels_sph.hip -- kernel classes of problem 2 (SphCfg): EulerMaruyama, Heun_Strong and
// Millstein in FP64 and mixed precision, Heun_NotConsistent and ImplicitMidpoint in FP64, uniaxial and
// cubic anisotropy.  Compiled twice, with different code-generation flags:
//   SP_SPH_PART == 1: the explicit solvers in FP64, at three waves per SIMD (SphCfg::WAVES_PER_SIMD), built
//     without machine loop-invariant code motion: hoisted out of the step loop, the polynomial constants and
//     request fields do not fit the 168 registers of that occupancy and are spilled; left in place they are
//     rematerialised where they are used.  Registry functions *_sph64.
//   SP_SPH_PART == 2: mixed precision (four waves per SIMD) and ImplicitMidpoint; its registry hands the kinds it
//     does not hold to part 1.
//   SP_SPH_PART == 3: one-wave builds (OneWave) of the explicit classes, FP64 and mixed, with machine LICM: a launch
//     of no more tiles than the device has SIMDs runs one wave per SIMD anyway, and there the hoisted constants fit
//     (65,536 particles, Heun_Strong FP64: 3.28e10 particle-steps/s against 3.06e10).  The registry of part 2 (the
//     problem's entry) hands such plain and batch launches (served iterations of a few requests) to it first.
#include "sp_kernels_impl.h"

#ifndef SP_SPH_PART
#error "sp_kernels_sph.hip is compiled with -DSP_SPH_PART=1 (FP64 explicit solvers), 2 or 3 (one-wave builds)"
#endif

namespace sp {

#define SP_COMMA ,
#if SP_SPH_PART == 1
#define SP_CLASSES(X)                                                                                       \
    X(kernel_kind(2, S_EM, 1, 0), SphCfg<1 SP_COMMA S_EM SP_COMMA double>, double, true, false)        \
    X(kernel_kind(2, S_EM, 2, 0), SphCfg<2 SP_COMMA S_EM SP_COMMA double>, double, true, false)        \
    X(kernel_kind(2, S_HEUN, 1, 0), SphCfg<1 SP_COMMA S_HEUN SP_COMMA double>, double, true, false)    \
    X(kernel_kind(2, S_HEUN, 2, 0), SphCfg<2 SP_COMMA S_HEUN SP_COMMA double>, double, true, false)    \
    X(kernel_kind(2, S_MIL, 1, 0), SphCfg<1 SP_COMMA S_MIL SP_COMMA double>, double, true, false)      \
    X(kernel_kind(2, S_MIL, 2, 0), SphCfg<2 SP_COMMA S_MIL SP_COMMA double>, double, true, false)      \
    X(kernel_kind(2, S_HNC, 1, 0), SphCfg<1 SP_COMMA S_HNC SP_COMMA double>, double, true, false)      \
    X(kernel_kind(2, S_HNC, 2, 0), SphCfg<2 SP_COMMA S_HNC SP_COMMA double>, double, true, false)

bool launch_plain_sph64(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq, uint64_t sb,
                        uint64_t se, double* partial, uint32_t pstride) {
    switch (kind) { SP_CLASSES(SP_CASE_PLAIN) default: return false; }
}
bool launch_batch_sph64(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                        const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    switch (kind) { SP_CLASSES(SP_CASE_BATCH) default: return false; }
}
bool info_sph64(uint32_t kind, KernelInfo* info) {
    switch (kind) { SP_CLASSES(SP_CASE_INFO) default: return false; }
}
#elif SP_SPH_PART == 3
#define SP_CLASSES(X)                                                                                       \
    X(kernel_kind(2, S_EM, 1, 0), OneWave<SphCfg<1 SP_COMMA S_EM SP_COMMA double>>, double, true, false)      \
    X(kernel_kind(2, S_EM, 2, 0), OneWave<SphCfg<2 SP_COMMA S_EM SP_COMMA double>>, double, true, false)      \
    X(kernel_kind(2, S_HEUN, 1, 0), OneWave<SphCfg<1 SP_COMMA S_HEUN SP_COMMA double>>, double, true, false)  \
    X(kernel_kind(2, S_HEUN, 2, 0), OneWave<SphCfg<2 SP_COMMA S_HEUN SP_COMMA double>>, double, true, false)  \
    X(kernel_kind(2, S_MIL, 1, 0), OneWave<SphCfg<1 SP_COMMA S_MIL SP_COMMA double>>, double, true, false)    \
    X(kernel_kind(2, S_MIL, 2, 0), OneWave<SphCfg<2 SP_COMMA S_MIL SP_COMMA double>>, double, true, false)    \
    X(kernel_kind(2, S_HNC, 1, 0), OneWave<SphCfg<1 SP_COMMA S_HNC SP_COMMA double>>, double, true, false)    \
    X(kernel_kind(2, S_HNC, 2, 0), OneWave<SphCfg<2 SP_COMMA S_HNC SP_COMMA double>>, double, true, false)    \
    X(kernel_kind(2, S_EM, 1, 1), OneWave<SphCfg<1 SP_COMMA S_EM SP_COMMA float>>, float, false, false)       \
    X(kernel_kind(2, S_EM, 2, 1), OneWave<SphCfg<2 SP_COMMA S_EM SP_COMMA float>>, float, false, false)       \
    X(kernel_kind(2, S_HEUN, 1, 1), OneWave<SphCfg<1 SP_COMMA S_HEUN SP_COMMA float>>, float, false, false)   \
    X(kernel_kind(2, S_HEUN, 2, 1), OneWave<SphCfg<2 SP_COMMA S_HEUN SP_COMMA float>>, float, false, false)   \
    X(kernel_kind(2, S_MIL, 1, 1), OneWave<SphCfg<1 SP_COMMA S_MIL SP_COMMA float>>, float, false, false)     \
    X(kernel_kind(2, S_MIL, 2, 1), OneWave<SphCfg<2 SP_COMMA S_MIL SP_COMMA float>>, float, false, false)

bool launch_plain_sph1w(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq, uint64_t sb,
                        uint64_t se, double* partial, uint32_t pstride) {
    switch (kind) { SP_CLASSES(SP_CASE_PLAIN) default: return false; }
}
bool launch_batch_sph1w(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                        const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    switch (kind) { SP_CLASSES(SP_CASE_BATCH) default: return false; }
}
#else
#define SP_CLASSES(X)                                                                                       \
    X(kernel_kind(2, S_EM, 1, 1), SphCfg<1 SP_COMMA S_EM SP_COMMA float>, float, false, false)        \
    X(kernel_kind(2, S_EM, 2, 1), SphCfg<2 SP_COMMA S_EM SP_COMMA float>, float, false, false)        \
    X(kernel_kind(2, S_HEUN, 1, 1), SphCfg<1 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)    \
    X(kernel_kind(2, S_HEUN, 2, 1), SphCfg<2 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)    \
    X(kernel_kind(2, S_MIL, 1, 1), SphCfg<1 SP_COMMA S_MIL SP_COMMA float>, float, false, false)      \
    X(kernel_kind(2, S_MIL, 2, 1), SphCfg<2 SP_COMMA S_MIL SP_COMMA float>, float, false, false)      \
    X(kernel_kind(2, S_IM, 1, 0), IMCfg<IMSph<1>>, double, true, true)                                \
    X(kernel_kind(2, S_IM, 2, 0), IMCfg<IMSph<2>>, double, true, true)

bool launch_plain_sph(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq, uint64_t sb,
                     uint64_t se, double* partial, uint32_t pstride) {
    if ((uint64_t)grid.x * block.x <= 64ull * device_simds() &&
        launch_plain_sph1w(kind, grid, block, s, rq, sb, se, partial, pstride))
        return true;
    switch (kind) {
        SP_CLASSES(SP_CASE_PLAIN)
        default: return launch_plain_sph64(kind, grid, block, s, rq, sb, se, partial, pstride);
    }
}
bool launch_batch_sph(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                     const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    if (n_tiles <= device_simds() &&
        launch_batch_sph1w(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter))
        return true;
    switch (kind) {
        SP_CLASSES(SP_CASE_BATCH)
        default: return launch_batch_sph64(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter);
    }
}
bool info_sph(uint32_t kind, KernelInfo* info) {
    switch (kind) { SP_CLASSES(SP_CASE_INFO) default: return info_sph64(kind, info); }
}
#endif

}  // namespace sp
