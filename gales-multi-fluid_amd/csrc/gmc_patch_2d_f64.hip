// gmc_patch_2d_f64.hip -- the FP64 patch kernels with an element matrix, for
// the 2D shapes, in a translation unit of their own: they are built with the
// AMDGPU "max-ilp" machine scheduler.  At one wave per SIMD nothing else
// hides their latency, and that schedule ran the 2D channel's FP64 patch
// kernel 2.5 % faster (6.61 -> 6.44 ms on MI355X).  The two-wave kernels (the
// FP32 tier's, and the residual-only one) ran slower with it (3.73 -> 3.79 ms
// and 1.36 -> 1.52 ms), so they stay in gmc_asm_2d*.hip with the default
// scheduler.
#include "gmc_asm_kernels.hpp"

namespace gmc {
template hipError_t launch_patch_k<2, 2, true, true, double>(const AsmArgs&, const PatchArgs&, hipStream_t);
template hipError_t launch_patch_k<2, 2, true, false, double>(const AsmArgs&, const PatchArgs&, hipStream_t);
template hipError_t launch_patch_k<2, 3, true, true, double>(const AsmArgs&, const PatchArgs&, hipStream_t);
template hipError_t launch_patch_k<2, 3, true, false, double>(const AsmArgs&, const PatchArgs&, hipStream_t);
}  // namespace gmc
