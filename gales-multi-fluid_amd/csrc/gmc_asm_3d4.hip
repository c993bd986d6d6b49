// gmc_asm_3d4.hip -- element kernels instantiated for dim 3, 4 components.
#include "gmc_asm_kernels.hpp"

namespace gmc {
template hipError_t launch_shape<3, 4, double>(int, int, const AsmArgs&, hipStream_t);
template hipError_t launch_shape<3, 4, float>(int, int, const AsmArgs&, hipStream_t);
template hipError_t launch_patch_shape<3, 4, double>(int, const AsmArgs&, const PatchArgs&,
                                                        hipStream_t);
template hipError_t launch_patch_shape<3, 4, float>(int, const AsmArgs&, const PatchArgs&,
                                                       hipStream_t);
}  // namespace gmc
