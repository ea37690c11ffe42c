// sp_kernels_uc.hip -- kernel classes with uniaxialcubic anisotropy (UniaxialCubicAnisotropy.hpp):
// Neel (NeelCfg) and NeelSpherical (SphCfg), EulerMaruyama, Heun_Strong and Millstein in FP64 and
// mixed precision, Heun_NotConsistent and ImplicitMidpoint in FP64.  BrownAndNeel with uniaxialcubic
// is rejected on the host.
#include "sp_kernels_impl.h"

namespace sp {

#define SP_CLASSES(X)                                                                                  \
    X(kernel_kind(1, S_EM, 4, 0), NeelCfg<4 SP_COMMA S_EM SP_COMMA double>, double, true, false)      \
    X(kernel_kind(1, S_HEUN, 4, 0), NeelCfg<4 SP_COMMA S_HEUN SP_COMMA double>, double, true, false)  \
    X(kernel_kind(1, S_MIL, 4, 0), NeelCfg<4 SP_COMMA S_MIL SP_COMMA double>, double, true, false)    \
    X(kernel_kind(1, S_EM, 4, 1), NeelCfg<4 SP_COMMA S_EM SP_COMMA float>, float, false, false)       \
    X(kernel_kind(1, S_HEUN, 4, 1), NeelCfg<4 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)   \
    X(kernel_kind(1, S_MIL, 4, 1), NeelCfg<4 SP_COMMA S_MIL SP_COMMA float>, float, false, false)     \
    X(kernel_kind(1, S_HNC, 4, 0), NeelCfg<4 SP_COMMA S_HNC SP_COMMA double>, double, true, false)    \
    X(kernel_kind(1, S_IM, 4, 0), IMCfg<IMNeel<4>>, double, true, true)                              \
    X(kernel_kind(2, S_EM, 4, 0), SphCfg<4 SP_COMMA S_EM SP_COMMA double>, double, true, false)       \
    X(kernel_kind(2, S_HEUN, 4, 0), SphCfg<4 SP_COMMA S_HEUN SP_COMMA double>, double, true, false)   \
    X(kernel_kind(2, S_MIL, 4, 0), SphCfg<4 SP_COMMA S_MIL SP_COMMA double>, double, true, false)     \
    X(kernel_kind(2, S_EM, 4, 1), SphCfg<4 SP_COMMA S_EM SP_COMMA float>, float, false, false)        \
    X(kernel_kind(2, S_HEUN, 4, 1), SphCfg<4 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)    \
    X(kernel_kind(2, S_MIL, 4, 1), SphCfg<4 SP_COMMA S_MIL SP_COMMA float>, float, false, false)      \
    X(kernel_kind(2, S_HNC, 4, 0), SphCfg<4 SP_COMMA S_HNC SP_COMMA double>, double, true, false)     \
    X(kernel_kind(2, S_IM, 4, 0), IMCfg<IMSph<4>>, double, true, true)
#define SP_COMMA ,

bool launch_plain_uc(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq, uint64_t sb,
                     uint64_t se, double* partial, uint32_t pstride) {
    switch (kind) { SP_CLASSES(SP_CASE_PLAIN) default: return false; }
}
bool launch_batch_uc(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                     const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    switch (kind) { SP_CLASSES(SP_CASE_BATCH) default: return false; }
}
bool info_uc(uint32_t kind, KernelInfo* info) {
    switch (kind) { SP_CLASSES(SP_CASE_INFO) default: return false; }
}

}  // namespace sp
