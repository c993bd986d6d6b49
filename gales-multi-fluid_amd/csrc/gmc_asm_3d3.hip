// gmc_asm_3d3.hip -- element kernels instantiated for dim 3, 3 components.
#include "gmc_asm_kernels.hpp"

namespace gmc {
template hipError_t launch_shape<3, 3, double>(int, int, const AsmArgs&, hipStream_t);
template hipError_t launch_shape<3, 3, float>(int, int, const AsmArgs&, hipStream_t);
template hipError_t launch_patch_shape<3, 3, double>(int, const AsmArgs&, const PatchArgs&,
                                                        hipStream_t);
template hipError_t launch_patch_shape<3, 3, float>(int, const AsmArgs&, const PatchArgs&,
                                                       hipStream_t);
}  // namespace gmc
