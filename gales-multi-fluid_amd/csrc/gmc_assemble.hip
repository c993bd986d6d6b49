// gmc_assemble.hip -- gather kernels and the launch dispatch of the element
// assembly (the element kernels live in gmc_asm_kernels.hpp, one translation
// unit per element shape).
#include <hip/hip_runtime.h>

#include "gmc_kernels.hpp"
#include "gmc_physics.hpp"

namespace gmc {

// per-shape instantiations (gmc_asm_<dim>d<nc>.hip)
template <int DIM, int NC, typename TM>
hipError_t launch_shape(int which, int flags, const AsmArgs& a, hipStream_t s);
template <int DIM, int NC, typename TM>
hipError_t launch_patch_shape(int flags, const AsmArgs& a, const PatchArgs& pa, hipStream_t s);

// One thread per value of a shared block (or rhs node).  NB is the block
// size (ndof^2 or ndof), a compile-time constant so that the index split is a
// multiply, not a 64-bit division.
template <int NB>
__global__ __launch_bounds__(256) void k_halo_gather(long long n_groups,
                                                     const int* __restrict__ ptr,
                                                     const int* __restrict__ dest,
                                                     const double* __restrict__ halo,
                                                     double* __restrict__ out) {
  const unsigned long long g = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (unsigned long long)n_groups * NB) return;
  const unsigned long long h = g / NB;
  const int j = (int)(g - h * NB);
  const long long c0 = ptr[h], c1 = ptr[h + 1];
  // the (2-4 in 2D) partials of a shared block are contiguous; up to four
  // loads in flight, added in patch order
  double s = 0.0;
  for (long long c = c0; c < c1; c += 4) {
    const long long n = c1 - c;
    const double v0 = __builtin_nontemporal_load(halo + c * NB + j);
    const double v1 = n > 1 ? __builtin_nontemporal_load(halo + (c + 1) * NB + j) : 0.0;
    const double v2 = n > 2 ? __builtin_nontemporal_load(halo + (c + 2) * NB + j) : 0.0;
    const double v3 = n > 3 ? __builtin_nontemporal_load(halo + (c + 3) * NB + j) : 0.0;
    s += v0;
    if (n > 1) s += v1;
    if (n > 2) s += v2;
    if (n > 3) s += v3;
  }
  out[(long long)dest[h] * NB + j] = s;
}

// ---------------------------------------------------------------------------
// gather of the element buffer.  The matrix part is stored block-sorted: the
// contributions to BSR block t occupy slots cptr[t] .. cptr[t+1]-1 in
// ascending element order, so vals[t][j] is a sum over a contiguous run
// (streamed reads, fixed order).  The rhs part is element-major and summed per
// node through the (element, local node) codes.  FP64 accumulation in every
// tier.
// ---------------------------------------------------------------------------
template <int NB, typename TM>
__global__ __launch_bounds__(256) void k_gather_blocks(long long nnzb, const int* __restrict__ cptr,
                                                       const TM* __restrict__ ebm,
                                                       double* __restrict__ vals, int add) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nnzb * NB) return;
  const long long t = g / NB;
  const int j = (int)(g - t * NB);
  const long long c0 = cptr[t], c1 = cptr[t + 1];
  // four loads in flight per step; the adds keep the slot order
  double s = 0.0;
  long long c = c0;
  for (; c + 4 <= c1; c += 4) {
    const TM v0 = __builtin_nontemporal_load(ebm + c * NB + j);
    const TM v1 = __builtin_nontemporal_load(ebm + (c + 1) * NB + j);
    const TM v2 = __builtin_nontemporal_load(ebm + (c + 2) * NB + j);
    const TM v3 = __builtin_nontemporal_load(ebm + (c + 3) * NB + j);
    s += (double)v0;
    s += (double)v1;
    s += (double)v2;
    s += (double)v3;
  }
  const int left = (int)(c1 - c);
  const TM w0 = left > 0 ? __builtin_nontemporal_load(ebm + c * NB + j) : TM(0);
  const TM w1 = left > 1 ? __builtin_nontemporal_load(ebm + (c + 1) * NB + j) : TM(0);
  const TM w2 = left > 2 ? __builtin_nontemporal_load(ebm + (c + 2) * NB + j) : TM(0);
  if (left > 0) s += (double)w0;
  if (left > 1) s += (double)w1;
  if (left > 2) s += (double)w2;
  // add: a later element chunk's sum onto the values the earlier ones wrote
  vals[g] = add ? vals[g] + s : s;
}

__global__ __launch_bounds__(256) void k_gather_rhs(long long n_nodes, int ndof, int rslot,
                                                    const int* __restrict__ nptr,
                                                    const int* __restrict__ nidx,
                                                    const double* __restrict__ ebr,
                                                    double* __restrict__ rhs) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_nodes * ndof) return;
  const long long i = g / ndof;
  const int k = (int)(g - i * ndof);
  double s = 0.0;
  for (int c = nptr[i]; c < nptr[i + 1]; c++) {
    const long long code = (unsigned)nidx[c];   // element * 4 + local node
    const long long e = code / NODE_PER_ELEM_MAX;
    const int b = (int)(code - e * NODE_PER_ELEM_MAX);
    s += ebr[e * rslot + b * ndof + k];
  }
  rhs[g] = s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename TM>
static hipError_t launch_any(int dim, int nc, int which, int flags, const AsmArgs& a,
                             hipStream_t s) {
  if (dim == 2 && nc == 2) return launch_shape<2, 2, TM>(which, flags, a, s);
  if (dim == 2 && nc == 3) return launch_shape<2, 3, TM>(which, flags, a, s);
  if (dim == 3 && nc == 3) return launch_shape<3, 3, TM>(which, flags, a, s);
  if (dim == 3 && nc == 4) return launch_shape<3, 4, TM>(which, flags, a, s);
  return hipErrorInvalidValue;
}

hipError_t launch_assemble(int dim, int nc, int which, int flags, const AsmArgs& a,
                           hipStream_t s) {
  if ((flags & GMC_ASM_FP32) && !(which & (ASM_JVP | ASM_VJPT))) return launch_any<float>(dim, nc, which, flags, a, s);
  return launch_any<double>(dim, nc, which, flags, a, s);
}

template <typename TM>
static hipError_t launch_patch_any(int dim, int nc, int flags, const AsmArgs& a,
                                   const PatchArgs& pa, hipStream_t s) {
  if (dim == 2 && nc == 2) return launch_patch_shape<2, 2, TM>(flags, a, pa, s);
  if (dim == 2 && nc == 3) return launch_patch_shape<2, 3, TM>(flags, a, pa, s);
  if (dim == 3 && nc == 3) return launch_patch_shape<3, 3, TM>(flags, a, pa, s);
  if (dim == 3 && nc == 4) return launch_patch_shape<3, 4, TM>(flags, a, pa, s);
  return hipErrorInvalidValue;
}

hipError_t launch_patch(int dim, int nc, int flags, const AsmArgs& a, const PatchArgs& pa,
                        hipStream_t s) {
  if (flags & GMC_ASM_FP32) return launch_patch_any<float>(dim, nc, flags, a, pa, s);
  return launch_patch_any<double>(dim, nc, flags, a, pa, s);
}

hipError_t launch_halo_gather(long long n_groups, int nb, const int* ptr, const int* dest,
                              const double* halo, double* out, hipStream_t s) {
  if (n_groups <= 0) return hipSuccess;
  const long long n = n_groups * nb;
  const dim3 g((unsigned)((n + 255) / 256)), b(256);
#define GMC_HG(NB_) hipLaunchKernelGGL(k_halo_gather<NB_>, g, b, 0, s, n_groups, ptr, dest, halo, out)
  switch (nb) {
    case 5: GMC_HG(5); break;
    case 6: GMC_HG(6); break;
    case 7: GMC_HG(7); break;
    case 8: GMC_HG(8); break;
    case 25: GMC_HG(25); break;
    case 36: GMC_HG(36); break;
    case 49: GMC_HG(49); break;
    case 64: GMC_HG(64); break;
    default: return hipErrorInvalidValue;
  }
#undef GMC_HG
  return hipGetLastError();
}

hipError_t launch_gather(int ndof, int nen, long long nnzb, long long n_nodes, const int* cptr,
                         const int* nptr, const int* nidx, const void* ebm,
                         const double* ebr, double* vals, double* rhs, int flags, hipStream_t s) {
  const int rslot = nen * ndof;
  if (flags & GMC_ASM_MATRIX) {
    const int add = (flags & GATHER_ADD) ? 1 : 0;
    const long long n = nnzb * ndof * ndof;
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
#define GMC_GB(NB_)                                                                              \
  if (flags & GMC_ASM_FP32)                                                                      \
    hipLaunchKernelGGL((k_gather_blocks<NB_, float>), g, b, 0, s, nnzb, cptr,                    \
                       static_cast<const float*>(ebm), vals, add);                               \
  else                                                                                           \
    hipLaunchKernelGGL((k_gather_blocks<NB_, double>), g, b, 0, s, nnzb, cptr,                   \
                       static_cast<const double*>(ebm), vals, add);
    switch (ndof) {
      case 5: GMC_GB(25) break;
      case 6: GMC_GB(36) break;
      case 7: GMC_GB(49) break;
      case 8: GMC_GB(64) break;
      default: return hipErrorInvalidValue;
    }
#undef GMC_GB
  }
  if (flags & GMC_ASM_RHS) {
    const long long n = n_nodes * ndof;
    hipLaunchKernelGGL(k_gather_rhs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_nodes, ndof,
                       rslot, nptr, nidx, ebr, rhs);
  }
  return hipGetLastError();
}

}  // namespace gmc
