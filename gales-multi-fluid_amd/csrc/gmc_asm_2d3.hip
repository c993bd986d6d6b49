// gmc_asm_2d3.hip -- element kernels instantiated for dim 2, 3 components
// (the FP64 patch kernels with a matrix are in gmc_patch_2d_f64.hip).
#include "gmc_asm_kernels.hpp"

namespace gmc {
template hipError_t launch_shape<2, 3, double>(int, int, const AsmArgs&, hipStream_t);
template hipError_t launch_shape<2, 3, float>(int, int, const AsmArgs&, hipStream_t);
template hipError_t launch_patch_shape<2, 3, float>(int, const AsmArgs&, const PatchArgs&,
                                                       hipStream_t);
// FP64 patch launches: the matrix variants are built in gmc_patch_2d_f64.hip,
// the residual-only one (two waves per SIMD) here with the default scheduler
extern template hipError_t launch_patch_k<2, 3, true, true, double>(const AsmArgs&, const PatchArgs&,
                                                                     hipStream_t);
extern template hipError_t launch_patch_k<2, 3, true, false, double>(const AsmArgs&, const PatchArgs&,
                                                                      hipStream_t);
template hipError_t launch_patch_k<2, 3, false, true, double>(const AsmArgs&, const PatchArgs&,
                                                              hipStream_t);
template hipError_t launch_patch_shape<2, 3, double>(int, const AsmArgs&, const PatchArgs&,
                                                        hipStream_t);
}  // namespace gmc
