// sp_kernels_bn.hip -- kernel classes of problem 3 (BNCfg): EulerMaruyama, Heun_Strong and
// Millstein in FP64 and mixed precision, Heun_NotConsistent and ImplicitMidpoint in FP64, uniaxial and
// cubic anisotropy.  Compiled in two parts (two objects that build in parallel):
//   SP_BN_PART == 1: FP64 and ImplicitMidpoint.  The explicit FP64 classes run at two waves per SIMD
//     (BNCfg::WAVES_PER_SIMD); a plain launch with no more waves than the device has SIMDs takes their one-wave
//     build instead (OneWave).  Registry functions *_bn64.
//   SP_BN_PART == 2: mixed precision, at three waves per SIMD (168 registers; measured on the 256 BrownAndNeel
//     sets of the 512-set batching distribution, autotuned: 9.49e10 particle-steps/s against 9.04e10 at two waves
//     and 8.80e10 at four with 56-119 registers spilled); its registry hands the kinds it does not hold to part 1.
#include "sp_kernels_impl.h"

#ifndef SP_BN_PART
#error "sp_kernels_bn.hip is compiled with -DSP_BN_PART=1 (FP64, ImplicitMidpoint) or -DSP_BN_PART=2 (mixed)"
#endif

namespace sp {

#define SP_COMMA ,
#if SP_BN_PART == 1
#define SP_CLASSES_FIT(X)                                                                                   \
    X(kernel_kind(3, S_EM, 1, 0), BNCfg<1 SP_COMMA S_EM SP_COMMA double>, double, true, false)        \
    X(kernel_kind(3, S_EM, 2, 0), BNCfg<2 SP_COMMA S_EM SP_COMMA double>, double, true, false)        \
    X(kernel_kind(3, S_HEUN, 1, 0), BNCfg<1 SP_COMMA S_HEUN SP_COMMA double>, double, true, false)    \
    X(kernel_kind(3, S_HEUN, 2, 0), BNCfg<2 SP_COMMA S_HEUN SP_COMMA double>, double, true, false)    \
    X(kernel_kind(3, S_MIL, 1, 0), BNCfg<1 SP_COMMA S_MIL SP_COMMA double>, double, true, false)      \
    X(kernel_kind(3, S_MIL, 2, 0), BNCfg<2 SP_COMMA S_MIL SP_COMMA double>, double, true, false)      \
    X(kernel_kind(3, S_HNC, 1, 0), BNCfg<1 SP_COMMA S_HNC SP_COMMA double>, double, true, false)      \
    X(kernel_kind(3, S_HNC, 2, 0), BNCfg<2 SP_COMMA S_HNC SP_COMMA double>, double, true, false)
#define SP_CLASSES(X)                                                                                       \
    SP_CLASSES_FIT(X)                                                                                       \
    X(kernel_kind(3, S_IM, 1, 0), IMCfg<IMBN<1>>, double, true, true)                                   \
    X(kernel_kind(3, S_IM, 2, 0), IMCfg<IMBN<2>>, double, true, true)

bool launch_plain_bn64(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq, uint64_t sb,
                       uint64_t se, double* partial, uint32_t pstride) {
    switch (kind) {
        SP_CLASSES_FIT(SP_CASE_PLAIN_FIT)
        SP_CASE_PLAIN(kernel_kind(3, S_IM, 1, 0), IMCfg<IMBN<1>>, double, true, true)
        SP_CASE_PLAIN(kernel_kind(3, S_IM, 2, 0), IMCfg<IMBN<2>>, double, true, true)
        default: return false;
    }
}
bool launch_batch_bn64(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                       const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    switch (kind) { SP_CLASSES(SP_CASE_BATCH) default: return false; }
}
bool info_bn64(uint32_t kind, KernelInfo* info) {
    switch (kind) { SP_CLASSES(SP_CASE_INFO) default: return false; }
}
#else
#define SP_CLASSES(X)                                                                                       \
    X(kernel_kind(3, S_EM, 1, 1), BNCfg<1 SP_COMMA S_EM SP_COMMA float>, float, false, false)        \
    X(kernel_kind(3, S_EM, 2, 1), BNCfg<2 SP_COMMA S_EM SP_COMMA float>, float, false, false)        \
    X(kernel_kind(3, S_HEUN, 1, 1), BNCfg<1 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)    \
    X(kernel_kind(3, S_HEUN, 2, 1), BNCfg<2 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)    \
    X(kernel_kind(3, S_MIL, 1, 1), BNCfg<1 SP_COMMA S_MIL SP_COMMA float>, float, false, false)      \
    X(kernel_kind(3, S_MIL, 2, 1), BNCfg<2 SP_COMMA S_MIL SP_COMMA float>, float, false, false)

bool launch_plain_bn(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq, uint64_t sb,
                     uint64_t se, double* partial, uint32_t pstride) {
    switch (kind) {
        SP_CLASSES(SP_CASE_PLAIN_SPLIT)
        default: return launch_plain_bn64(kind, grid, block, s, rq, sb, se, partial, pstride);
    }
}
bool launch_batch_bn(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                     const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    switch (kind) {
        SP_CLASSES(SP_CASE_BATCH)
        default: return launch_batch_bn64(kind, grid, block, s, reqs, slots, n_slots, n_tiles, counter);
    }
}
bool info_bn(uint32_t kind, KernelInfo* info) {
    switch (kind) { SP_CLASSES(SP_CASE_INFO) default: return info_bn64(kind, info); }
}
#endif

}  // namespace sp
