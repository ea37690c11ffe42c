// sp_kernels_neel.hip -- kernel classes of problem 1 (NeelCfg): EulerMaruyama, Heun_Strong and
// Millstein in FP64 and mixed precision, Heun_NotConsistent and ImplicitMidpoint in FP64, uniaxial and
// cubic anisotropy.
#include "sp_kernels_impl.h"

namespace sp {

#define SP_CLASSES_EXPLICIT(X)                                                          \
    X(kernel_kind(1, S_EM, 1, 0), NeelCfg<1 SP_COMMA S_EM SP_COMMA double>, double, true, false)        \
    X(kernel_kind(1, S_EM, 2, 0), NeelCfg<2 SP_COMMA S_EM SP_COMMA double>, double, true, false)        \
    X(kernel_kind(1, S_HEUN, 1, 0), NeelCfg<1 SP_COMMA S_HEUN SP_COMMA double>, double, true, false)    \
    X(kernel_kind(1, S_HEUN, 2, 0), NeelCfg<2 SP_COMMA S_HEUN SP_COMMA double>, double, true, false)    \
    X(kernel_kind(1, S_MIL, 1, 0), NeelCfg<1 SP_COMMA S_MIL SP_COMMA double>, double, true, false)      \
    X(kernel_kind(1, S_MIL, 2, 0), NeelCfg<2 SP_COMMA S_MIL SP_COMMA double>, double, true, false)      \
    X(kernel_kind(1, S_EM, 1, 1), NeelCfg<1 SP_COMMA S_EM SP_COMMA float>, float, false, false)         \
    X(kernel_kind(1, S_EM, 2, 1), NeelCfg<2 SP_COMMA S_EM SP_COMMA float>, float, false, false)         \
    X(kernel_kind(1, S_HEUN, 1, 1), NeelCfg<1 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)     \
    X(kernel_kind(1, S_HEUN, 2, 1), NeelCfg<2 SP_COMMA S_HEUN SP_COMMA float>, float, false, false)     \
    X(kernel_kind(1, S_MIL, 1, 1), NeelCfg<1 SP_COMMA S_MIL SP_COMMA float>, float, false, false)       \
    X(kernel_kind(1, S_MIL, 2, 1), NeelCfg<2 SP_COMMA S_MIL SP_COMMA float>, float, false, false)       \
    X(kernel_kind(1, S_HNC, 1, 0), NeelCfg<1 SP_COMMA S_HNC SP_COMMA double>, double, true, false)      \
    X(kernel_kind(1, S_HNC, 2, 0), NeelCfg<2 SP_COMMA S_HNC SP_COMMA double>, double, true, false)
#define SP_CLASSES_IM(X)                                                                \
    X(kernel_kind(1, S_IM, 1, 0), IMCfg<IMNeel<1>>, double, true, true)                                \
    X(kernel_kind(1, S_IM, 2, 0), IMCfg<IMNeel<2>>, double, true, true)
#define SP_CLASSES(X) SP_CLASSES_EXPLICIT(X) SP_CLASSES_IM(X)
#define SP_COMMA ,

// the explicit solvers of a small launch run split (k_plain_split)
bool launch_plain_neel(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* rq, uint64_t sb,
                     uint64_t se, double* partial, uint32_t pstride) {
    switch (kind) {
        SP_CLASSES_EXPLICIT(SP_CASE_PLAIN_SPLIT)
        SP_CLASSES_IM(SP_CASE_PLAIN)
        default: return false;
    }
}
bool launch_batch_neel(uint32_t kind, dim3 grid, dim3 block, hipStream_t s, const DevRequest* reqs,
                     const SlotWork* slots, uint32_t n_slots, uint32_t n_tiles, uint32_t* counter) {
    switch (kind) { SP_CLASSES(SP_CASE_BATCH) default: return false; }
}
bool info_neel(uint32_t kind, KernelInfo* info) {
    switch (kind) { SP_CLASSES(SP_CASE_INFO) default: return false; }
}

}  // namespace sp
