// fssh_n8.hip -- the five- to eight-state instantiations of the n-state FSSH / Ehrenfest kernel
// (fssh_n.hip), compiled as a translation unit of their own.
#include "fssh_n_kern.h"

namespace nqcd {

NQCD_NS_KERNELS(5)
NQCD_NS_KERNELS(6)
NQCD_NS_KERNELS(7)
NQCD_NS_KERNELS(8)

hipError_t ns_launch_hi(int NM, int precision, int grid, int block, hipStream_t s, const FsshNArgs* Adev) {
#define NQCD_CASE(X)                                                                                         \
    case X:                                                                                                  \
        if (precision == 0) hipLaunchKernelGGL(fssh_n_f64_##X, dim3(grid), dim3(block), 0, s, Adev);         \
        else hipLaunchKernelGGL(fssh_n_f32_##X, dim3(grid), dim3(block), 0, s, Adev);                        \
        break;
    switch (NM) {
        NQCD_CASE(5) NQCD_CASE(6) NQCD_CASE(7) NQCD_CASE(8)
    default: return hipErrorInvalidValue;
    }
#undef NQCD_CASE
    return hipGetLastError();
}

hipError_t ns_occupancy_hi(int NM, int precision, int block, int* bpc) {
#define NQCD_CASE(X)                                                                                         \
    case X:                                                                                                  \
        return precision == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, fssh_n_f64_##X, block, 0)  \
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, fssh_n_f32_##X, block, 0);
    switch (NM) {
        NQCD_CASE(5) NQCD_CASE(6) NQCD_CASE(7) NQCD_CASE(8)
    default: return hipErrorInvalidValue;
    }
#undef NQCD_CASE
}

}  // namespace nqcd
