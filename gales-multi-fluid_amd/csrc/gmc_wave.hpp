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

// Reduce-scatter over a quad: v0..v3 are this lane's values of four entries;
// the result in lane q is the quad sum of entry q, (x0 + x1) + (x2 + x3) over
// the lanes' values, the same sum (bitwise) that quad_sum gives every lane.
// Each lane keeps the half it owns and sends the other half at each of the
// two steps, so a group of four entries costs three permutes instead of
// eight, and no select of the owned entry afterwards.
template <typename T>
__device__ __forceinline__ T quad_rs4(T v0, T v1, T v2, T v3, int q) {
  const bool o = q & 1;
  T a0 = o ? v1 : v0;
  const T b0 = o ? v0 : v1;
  T a1 = o ? v3 : v2;
  const T b1 = o ? v2 : v3;
  a0 += dpp_quad<0xB1>(b0);  // lane q^1 sends entry (q&1) of each pair
  a1 += dpp_quad<0xB1>(b1);
  const bool h = q & 2;
  T c = h ? a1 : a0;
  const T d = h ? a0 : a1;
  c += dpp_quad<0x4E>(d);  // lane q^2 sends its partial of entry q
  return c;
}

// out[k] = quad sum of entry 4k + q of v[0..NV-1] (entries past NV are 0)
template <typename T, int NV>
__device__ __forceinline__ void quad_rs(const T* v, int q, T* out) {
  GMC_UNROLL for (int k = 0; k < (NV + 3) / 4; k++) {
    const T x0 = v[4 * k];
    const T x1 = 4 * k + 1 < NV ? v[4 * k + 1] : T(0);
    const T x2 = 4 * k + 2 < NV ? v[4 * k + 2] : T(0);
    const T x3 = 4 * k + 3 < NV ? v[4 * k + 3] : T(0);
    out[k] = quad_rs4(x0, x1, x2, x3, q);
  }
}

// ---------------------------------------------------------------------------
// Triads, for kernels whose lanes meet only in a few final sums: a
// triangle's three quadrature points on lanes 3g, 3g + 1, 3g + 2 of a 16-lane
// DPP row (g < 5, the row's 16th lane idle), 80 groups per 256-lane
// workgroup instead of the 64 quads whose fourth lane repeats point 2 with
// weight 0.  triad_sum adds the lanes' values as (x0 + x1) + x2, bitwise the
// quad sum's (x0 + x1) + (x2 + 0).  Row shifts: lane i reads lane i + n
// (row_shl:n) or i - n (row_shr:n); a source outside the row reads 0, and
// only the lane whose sources lie in its own triad keeps each partial sum.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ double dpp_row(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
enum { DPP_ROW_SHL1 = 0x101, DPP_ROW_SHL2 = 0x102, DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112 };

// the triad sum of x in every lane of the triad (p: the lane's point, 0..2)
__device__ __forceinline__ double triad_sum(double x, int p) {
  const double l1 = dpp_row<DPP_ROW_SHL1>(x), l2 = dpp_row<DPP_ROW_SHL2>(x);
  const double r1 = dpp_row<DPP_ROW_SHR1>(x), r2 = dpp_row<DPP_ROW_SHR2>(x);
  const double s0 = (x + l1) + l2, s1 = (r1 + x) + l1, s2 = (r2 + r1) + x;
  return p == 0 ? s0 : (p == 1 ? s1 : s2);
}

// groups (one element, or one element and direction) per 256-lane workgroup:
// triads for triangles, quads for tetrahedra
template <int DIM>
constexpr int groups_per_block() {
  return DIM == 2 ? 80 : 64;
}
// this lane's quadrature point and group within its workgroup; false for a
// triad layout's idle lane (which then mirrors group 4 of its row, point 0,
// and must neither write nor flag)
template <int DIM>
__device__ __forceinline__ bool group_lane(int lane, int& point, int& group) {
  if constexpr (DIM == 3) {
    point = lane & 3;
    group = lane >> 2;
    return true;
  } else {
    const int r = lane & 15, g = r / 3;
    point = g < 5 ? r - 3 * g : 0;
    group = (lane >> 4) * 5 + (g < 5 ? g : 4);
    return r < 15;
  }
}
template <int DIM>
__device__ __forceinline__ double group_sum(double x, int p) {
  if constexpr (DIM == 3) return quad_sum(x);
  else return triad_sum(x, p);
}

// model::extract_element_dofs (model.hpp:116-132): node ids, coordinates and
// the element vectors of the I (and, WITH_P, the P) state.  Whether P is
// gathered is a template argument, not a null test of Pe: a run-time test of
// a pointer to a private array keeps the array addressable, in scratch.
template <int DIM, int NC, bool WITH_P = true>
__device__ __forceinline__ void gather(const AsmArgs& a, long long e, int* en, double* xe,
                                       double* Ie, double* Pe) {
  using S = Shape<DIM, NC>;
  GMC_UNROLL for (int n = 0; n < S::NEN; n++) {
    en[n] = a.conn[e * S::NEN + n];
    GMC_UNROLL for (int k = 0; k < DIM; k++) xe[n * DIM + k] = a.coords[(long long)en[n] * DIM + k];
    GMC_UNROLL for (int k = 0; k < S::NDOF; k++) {
      Ie[n * S::NDOF + k] = a.I[(long long)en[n] * S::NDOF + k];
      if constexpr (WITH_P) Pe[n * S::NDOF + k] = a.P[(long long)en[n] * S::NDOF + k];
    }
  }
}

}  // namespace gmc
