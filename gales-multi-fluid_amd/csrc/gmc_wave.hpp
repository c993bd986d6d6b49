// gmc_wave.hpp -- device helpers shared by the element kernels: sums over the
// four lanes of a quad (one lane per quadrature point) and the element gather.
#pragma once

#include <hip/hip_runtime.h>

#include "gmc_kernels.hpp"
#include "gmc_physics.hpp"

namespace gmc {

// quad_perm DPP: lane i reads lane i^1 (0xB1) or lane i^2 (0x4E) of its quad
template <int CTRL>
__device__ __forceinline__ double dpp_quad(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_quad(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}
template <typename T>
__device__ __forceinline__ T quad_sum(T x) {
  x += dpp_quad<0xB1>(x);
  x += dpp_quad<0x4E>(x);
  return x;
}

// model::extract_element_dofs (model.hpp:116-132): node ids, coordinates and
// the element vectors of the I (and P) state.
template <int DIM, int NC>
__device__ __forceinline__ void gather(const AsmArgs& a, long long e, int* en, double* xe,
                                       double* Ie, double* Pe) {
  using S = Shape<DIM, NC>;
  GMC_UNROLL for (int n = 0; n < S::NEN; n++) {
    en[n] = a.conn[e * S::NEN + n];
    GMC_UNROLL for (int k = 0; k < DIM; k++) xe[n * DIM + k] = a.coords[(long long)en[n] * DIM + k];
    GMC_UNROLL for (int k = 0; k < S::NDOF; k++) {
      Ie[n * S::NDOF + k] = a.I[(long long)en[n] * S::NDOF + k];
      if (Pe) Pe[n * S::NDOF + k] = a.P[(long long)en[n] * S::NDOF + k];
    }
  }
}

}  // namespace gmc
