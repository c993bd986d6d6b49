// gmc_patch_2d_f64.hip -- the FP64 patch kernels of the 2D shapes, in a
// translation unit of their own: they are built with the AMDGPU "max-ilp"
// machine scheduler.  At one wave per SIMD nothing else hides their latency,
// and that schedule ran the 2D channel's FP64 patch kernel 2.5 % faster
// (6.61 -> 6.44 ms on MI355X), while the two-wave FP32 kernels were 1.6 %
// slower with it, so those stay with the default scheduler.
#include "gmc_asm_kernels.hpp"

namespace gmc {
template hipError_t launch_patch_shape<2, 2, double>(int, const AsmArgs&, const PatchArgs&,
                                                        hipStream_t);
template hipError_t launch_patch_shape<2, 3, double>(int, const AsmArgs&, const PatchArgs&,
                                                        hipStream_t);
}  // namespace gmc
