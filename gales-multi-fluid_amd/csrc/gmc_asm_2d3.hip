// gmc_asm_2d3.hip -- element kernels instantiated for dim 2, 3 components
// (the FP64 patch kernels are in gmc_patch_2d_f64.hip).
#include "gmc_asm_kernels.hpp"

namespace gmc {
template hipError_t launch_shape<2, 3, double>(int, int, const AsmArgs&, hipStream_t);
template hipError_t launch_shape<2, 3, float>(int, int, const AsmArgs&, hipStream_t);
template hipError_t launch_patch_shape<2, 3, float>(int, const AsmArgs&, const PatchArgs&,
                                                       hipStream_t);
}  // namespace gmc
