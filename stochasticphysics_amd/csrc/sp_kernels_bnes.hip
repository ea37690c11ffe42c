// sp_kernels_bnes.hip -- kernel classes of problem 4 (BNESCfg, BrownAndNeelEulerSpherical): EulerMaruyama,
// Heun_Strong and Millstein in FP64 and mixed precision, uniaxial, cubic and uniaxialcubic anisotropy.
// Implicit_Midpoint is refused on the host (the reference has no Jacobians for this problem).
#include "sp_device_bnes.h"
#include "sp_kernels_impl.h"

namespace sp {

#define SP_COMMA ,
#define SP_CLASSES(X)                                                                               \
    X(kernel_kind(4, S_EM, 1, 0), BNESCfg<1 SP_COMMA S_EM SP_COMMA double>, double, true, false)     \
    X(kernel_kind(4, S_EM, 1, 1), BNESCfg<1 SP_COMMA S_EM SP_COMMA float>, float, false, false)      \
    X(kernel_kind(4, S_HEUN, 1, 0), BNESCfg<1 SP_COMMA S_HEUN SP_COMMA double>, double, true, false) \
    X(kernel_kind(4, S_HEUN, 1, 1), BNESCfg<1 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)  \
    X(kernel_kind(4, S_MIL, 1, 0), BNESCfg<1 SP_COMMA S_MIL SP_COMMA double>, double, true, false)   \
    X(kernel_kind(4, S_MIL, 1, 1), BNESCfg<1 SP_COMMA S_MIL SP_COMMA float>, float, false, false)    \
    X(kernel_kind(4, S_EM, 2, 0), BNESCfg<2 SP_COMMA S_EM SP_COMMA double>, double, true, false)     \
    X(kernel_kind(4, S_EM, 2, 1), BNESCfg<2 SP_COMMA S_EM SP_COMMA float>, float, false, false)      \
    X(kernel_kind(4, S_HEUN, 2, 0), BNESCfg<2 SP_COMMA S_HEUN SP_COMMA double>, double, true, false) \
    X(kernel_kind(4, S_HEUN, 2, 1), BNESCfg<2 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)  \
    X(kernel_kind(4, S_MIL, 2, 0), BNESCfg<2 SP_COMMA S_MIL SP_COMMA double>, double, true, false)   \
    X(kernel_kind(4, S_MIL, 2, 1), BNESCfg<2 SP_COMMA S_MIL SP_COMMA float>, float, false, false)    \
    X(kernel_kind(4, S_EM, 4, 0), BNESCfg<4 SP_COMMA S_EM SP_COMMA double>, double, true, false)     \
    X(kernel_kind(4, S_EM, 4, 1), BNESCfg<4 SP_COMMA S_EM SP_COMMA float>, float, false, false)      \
    X(kernel_kind(4, S_HEUN, 4, 0), BNESCfg<4 SP_COMMA S_HEUN SP_COMMA double>, double, true, false) \
    X(kernel_kind(4, S_HEUN, 4, 1), BNESCfg<4 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)  \
    X(kernel_kind(4, S_MIL, 4, 0), BNESCfg<4 SP_COMMA S_MIL SP_COMMA double>, double, true, false)   \
    X(kernel_kind(4, S_MIL, 4, 1), BNESCfg<4 SP_COMMA S_MIL SP_COMMA float>, float, false, false)

bool launch_plain_bnes(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq, uint64_t sb,
                       uint64_t se, double* partial, uint32_t pstride) {
    switch (kind) { SP_CLASSES(SP_CASE_PLAIN) default: return false; }
}
bool launch_batch_bnes(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                       const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    switch (kind) { SP_CLASSES(SP_CASE_BATCH) default: return false; }
}
bool info_bnes(uint32_t kind, KernelInfo* info) {
    switch (kind) { SP_CLASSES(SP_CASE_INFO) default: return false; }
}

}  // namespace sp
