// sp_kernels_bnes.hip -- kernel classes of problem 4 (BNESCfg, BrownAndNeelEulerSpherical): EulerMaruyama,
// Heun_Strong and Millstein in FP64 and mixed precision, Heun_NotConsistent in FP64, uniaxial, cubic and
// uniaxialcubic anisotropy.
// Implicit_Midpoint is refused on the host (the reference has no Jacobians for this problem).  Compiled once per
// anisotropy (SP_BNES_ANISO), which keeps each translation unit's compile time near the others'.
#include "sp_device_bnes.h"
#include "sp_kernels_impl.h"

namespace sp {

#define SP_COMMA ,
// one anisotropy per compilation (the build compiles this file once per SP_BNES_ANISO = 1, 2, 4)
#ifndef SP_BNES_ANISO
#error "compile with -DSP_BNES_ANISO=1, 2 or 4"
#endif
#define SP_A SP_BNES_ANISO
#define SP_CLASSES(X)                                                                                    \
    X(kernel_kind(4, S_EM, SP_A, 0), BNESCfg<SP_A SP_COMMA S_EM SP_COMMA double>, double, true, false)     \
    X(kernel_kind(4, S_EM, SP_A, 1), BNESCfg<SP_A SP_COMMA S_EM SP_COMMA float>, float, false, false)      \
    X(kernel_kind(4, S_HEUN, SP_A, 0), BNESCfg<SP_A SP_COMMA S_HEUN SP_COMMA double>, double, true, false) \
    X(kernel_kind(4, S_HEUN, SP_A, 1), BNESCfg<SP_A SP_COMMA S_HEUN SP_COMMA float>, float, false, false)  \
    X(kernel_kind(4, S_MIL, SP_A, 0), BNESCfg<SP_A SP_COMMA S_MIL SP_COMMA double>, double, true, false)   \
    X(kernel_kind(4, S_MIL, SP_A, 1), BNESCfg<SP_A SP_COMMA S_MIL SP_COMMA float>, float, false, false)    \
    X(kernel_kind(4, S_HNC, SP_A, 0), BNESCfg<SP_A SP_COMMA S_HNC SP_COMMA double>, double, true, false)
#define SP_CAT2(a, b) a##b
#define SP_CAT(a, b) SP_CAT2(a, b)

bool SP_CAT(launch_plain_bnes_a, SP_A)(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq,
                                       uint64_t sb, uint64_t se, double* partial, uint32_t pstride) {
    switch (kind) { SP_CLASSES(SP_CASE_PLAIN) default: return false; }
}
bool SP_CAT(launch_batch_bnes_a, SP_A)(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                                       const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    switch (kind) { SP_CLASSES(SP_CASE_BATCH) default: return false; }
}
bool SP_CAT(info_bnes_a, SP_A)(uint32_t kind, KernelInfo* info) {
    switch (kind) { SP_CLASSES(SP_CASE_INFO) default: return false; }
}

}  // namespace sp
