// gmc_magma.hip -- magma point tables.  Before the element kernels of an
// assembly or a Jacobian product run, these kernels evaluate every magma
// component's properties (gmc_magma.hpp) at
//   - every interior quadrature point: the P state at strain rate 0 and the
//     I state at its shear rate (fluid_2d.hpp:183-194), and
//   - every boundary side point: the I state at its shear rate
//     (fluid_2d.hpp:227-235),
// and store them where the element kernels read them (AsmArgs::mt, mtb).
// The element kernels thus carry no magma code: their register allocation is
// the one of the custom / ideal-gas mixture.
#include <hip/hip_runtime.h>

#include "gmc_element.hpp"
#include "gmc_kernels.hpp"

namespace gmc {

// the magma components' values at one point and state into out[NC][kMagmaVals]
template <int NC>
__device__ __forceinline__ int magma_store(const DevParams& P, double p, double T, double sr,
                                           double* out) {
  int st = 0;
#pragma unroll 1
  for (int c = 0; c < NC; c++) {
    if (P.kind[c] != 2) continue;
    const MagmaProps m = magma_props(P.mg[c], p, T, sr);
    st |= m.st;
    double* o = out + c * kMagmaVals;
    o[0] = m.rho;
    o[1] = m.mu;
    o[2] = m.cp;
    o[3] = m.cv;
    o[4] = m.kappa;
    o[5] = m.c;
    o[6] = m.alpha;
    o[7] = m.beta;
    o[8] = m.D;
  }
  return st;
}

template <int DIM, int NC, bool WITH_P>
__device__ __forceinline__ bool magma_gather(const MagmaPointArgs& a, long long e, double* xe,
                                             double* Ie, double* Pe, Geo<DIM>& g) {
  // WITH_P: a compile-time choice (a run-time null test of a pointer to a
  // private array would keep that array in scratch)
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN;
  GMC_UNROLL for (int n = 0; n < NEN; n++) {
    const long long v = a.conn[e * NEN + n];
    GMC_UNROLL for (int k = 0; k < DIM; k++) xe[n * DIM + k] = a.coords[v * DIM + k];
    GMC_UNROLL for (int k = 0; k < N; k++) {
      Ie[n * N + k] = a.I[v * N + k];
      if constexpr (WITH_P) Pe[n * N + k] = a.P[v * N + k];
    }
  }
  return geometry<DIM>(xe, g);   // a non-positive Jacobian is the element kernels' to report
}

// one thread per (element, interior point)
template <int DIM, int NC>
__global__ __launch_bounds__(256) void k_magma_points(const MagmaPointArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NQ = S::NQ;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_elems * NQ) return;
  const long long e = t / NQ;
  const int q = (int)(t - e * NQ);
  const DevParams& P = *a.params;
  double xe[NEN * DIM], Ie[NEN * N], Pe[NEN * N];
  Geo<DIM> g;
  if (!magma_gather<DIM, NC, true>(a, e, xe, Ie, Pe, g)) return;
  double sh[NEN], Pq[N], Iq[N], dI[DIM][N];
  qp_shape<DIM>(q, sh);
  interpolate<double, DIM, NC>(sh, Pe, Pq);
  interpolate<double, DIM, NC>(sh, Ie, Iq);
  GMC_UNROLL for (int k = 0; k < DIM; k++) interpolate<double, DIM, NC>(g.dsh[k], Ie, dI[k]);
  double p, v[DIM], Tt, Y[NC];
  double* out = a.mt + t * 2 * NC * kMagmaVals;
  assign<double, DIM, NC>(Pq, p, v, Tt, Y);
  int st = magma_store<NC>(P, p, Tt, 0.0, out);
  assign<double, DIM, NC>(Iq, p, v, Tt, Y);
  const double sr = P.need_sr ? shear_rate<double, DIM, N>(dI) : 0.0;
  st |= magma_store<NC>(P, p, Tt, sr, out + NC * kMagmaVals);
  if (st) atomicOr(a.status, st);
}

// one thread per (boundary side, side point)
template <int DIM, int NC>
__global__ __launch_bounds__(256) void k_magma_side_points(const MagmaPointArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NSQ = S::NSQ;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_sides * NSQ) return;
  const long long s = t / NSQ;
  const int sq = (int)(t - s * NSQ);
  const DevParams& P = *a.params;
  double xe[NEN * DIM], Ie[NEN * N];
  Geo<DIM> g;
  if (!magma_gather<DIM, NC, false>(a, a.side_elem[s], xe, Ie, nullptr, g)) return;
  double xi[DIM], sh[NEN], Iq[N], dI[DIM][N];
  side_point<DIM>(a.side_local[s], sq, xi);
  side_shape<DIM>(xi, sh);
  interpolate<double, DIM, NC>(sh, Ie, Iq);
  GMC_UNROLL for (int k = 0; k < DIM; k++) interpolate<double, DIM, NC>(g.dsh[k], Ie, dI[k]);
  double p, v[DIM], Tt, Y[NC];
  assign<double, DIM, NC>(Iq, p, v, Tt, Y);
  const double sr = P.need_sr ? shear_rate<double, DIM, N>(dI) : 0.0;
  const int st = magma_store<NC>(P, p, Tt, sr, a.mtb + t * NC * kMagmaVals);
  if (st) atomicOr(a.status, st);
}

template <int DIM, int NC>
static hipError_t launch_points(const MagmaPointArgs& a, hipStream_t s) {
  const long long ni = a.n_elems * (DIM + 1);
  if (ni > 0)
    hipLaunchKernelGGL((k_magma_points<DIM, NC>), dim3((unsigned)((ni + 255) / 256)), dim3(256), 0, s, a);
  const long long nb = a.n_sides * DIM;
  if (nb > 0)
    hipLaunchKernelGGL((k_magma_side_points<DIM, NC>), dim3((unsigned)((nb + 255) / 256)), dim3(256), 0,
                       s, a);
  return hipGetLastError();
}

hipError_t launch_magma_points(int dim, int nc, const MagmaPointArgs& a, hipStream_t s) {
  if (dim == 2 && nc == 2) return launch_points<2, 2>(a, s);
  if (dim == 2 && nc == 3) return launch_points<2, 3>(a, s);
  if (dim == 3 && nc == 3) return launch_points<3, 3>(a, s);
  if (dim == 3 && nc == 4) return launch_points<3, 4>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace gmc
