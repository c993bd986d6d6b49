// gmc_steps.hip -- the steps either side of the assembly that evaluate the
// point physics over the mesh: the CFL-adaptive time step of
// adaptive_time_criteria::f_mc (src/fem/utils/adaptive_time_step.hpp:183-253).
#include <hip/hip_runtime.h>

#include "gmc_element.hpp"
#include "gmc_kernels.hpp"

namespace gmc {

// One thread per element: the minimum over its interior points of
//   CFL / (2 max(2 mu/rho, kappa/(rho cv)) / h^2 + v_tau / h)
// (1.2 dt_prev where |v| < eps; v_tau modified above Mach 0.3), folded into
// *gmin_bits by atomicMin on the bit pattern: for positive doubles the
// integer order is the numeric order, so the minimum is exact whatever the
// order of the updates.
template <int DIM, int NC>
__global__ __launch_bounds__(256) void k_element_dt(const double* __restrict__ coords,
                                                   const int* __restrict__ conn,
                                                   const double* __restrict__ I,
                                                   const DevParams* __restrict__ params,
                                                   long long n_elems, double cfl, double dt_prev,
                                                   unsigned long long* gmin_bits, int* status) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elems) return;
  const DevParams& P = *params;
  Material<double> M;
  load_material<double, NC>(P, M);
  double xe[NEN * DIM], Ie[NEN * N];
  GMC_UNROLL for (int a = 0; a < NEN; a++) {
    const long long n = conn[e * NEN + a];
    GMC_UNROLL for (int k = 0; k < DIM; k++) xe[a * DIM + k] = coords[n * DIM + k];
    GMC_UNROLL for (int k = 0; k < N; k++) Ie[a * N + k] = I[n * N + k];
  }
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) {
    atomicOr(status, 1);
    return;
  }
  double dI[DIM][N];
  GMC_UNROLL for (int k = 0; k < DIM; k++) interpolate<double, DIM, NC>(g.dsh[k], Ie, dI[k]);
  double dv[DIM][DIM];
  GMC_UNROLL for (int i = 0; i < DIM; i++) GMC_UNROLL for (int j = 0; j < DIM; j++) dv[i][j] = dI[j][1 + i];
  double emin = 0.0;
  GMC_UNROLL for (int q = 0; q < S::NQ; q++) {
    double sh[NEN], Iq[N], v[DIM], p, Tt, Y[NC];
    qp_shape<DIM>(q, sh);
    interpolate<double, DIM, NC>(sh, Ie, Iq);
    assign<double, DIM, NC>(Iq, p, v, Tt, Y);
    Props<double, NC> pr;
    properties<double, NC>(P, M, p, Tt, Y, pr);  // strain rate 0 (adaptive_time_step.hpp:197)
    if (pr.st) atomicOr(status, pr.st);
    double vv = 0.0;
    GMC_UNROLL for (int k = 0; k < DIM; k++) vv += v[k] * v[k];
    const double v_nrm = sqrt(vv);
    double dt_gp;
    if (v_nrm < kEps) {
      dt_gp = 1.2 * dt_prev;
    } else {
      const double h = compute_h<double, DIM>(g.G, v, dv);
      const double c = pr.c;
      double v_tau = v_nrm;
      if (v_nrm / c > 0.3) {
        if (DIM == 2) v_tau = sqrt(v_nrm * v_nrm + 1.5 * c * c + c * sqrt(16.0 * v_nrm * v_nrm + c * c));
        else v_tau = sqrt(v_nrm * v_nrm + 2.0 * c * c + c * sqrt(4.0 * v_nrm * v_nrm + c * c));
      }
      const double a = 2.0 * pr.mu / pr.rho, b = pr.kappa / (pr.rho * pr.cv);
      const double use = 2.0 * ((a < b) ? b : a) / (h * h) + v_tau / h;
      dt_gp = cfl / use;
    }
    emin = (q == 0 || dt_gp < emin) ? dt_gp : emin;
  }
  atomicMin(gmin_bits, (unsigned long long)__double_as_longlong(emin));
}

template <int DIM, int NC>
static hipError_t launch_dt(const ElementDtArgs& a, hipStream_t s) {
  const unsigned grid = (unsigned)((a.n_elems + 255) / 256);
  hipLaunchKernelGGL((k_element_dt<DIM, NC>), dim3(grid > 0 ? grid : 1), dim3(256), 0, s, a.coords,
                     a.conn, a.I, a.params, a.n_elems, a.cfl, a.dt_prev, a.gmin_bits, a.status);
  return hipGetLastError();
}

hipError_t launch_element_dt(int dim, int nc, const ElementDtArgs& a, hipStream_t s) {
  if (dim == 2 && nc == 2) return launch_dt<2, 2>(a, s);
  if (dim == 2 && nc == 3) return launch_dt<2, 3>(a, s);
  if (dim == 3 && nc == 3) return launch_dt<3, 3>(a, s);
  if (dim == 3 && nc == 4) return launch_dt<3, 4>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace gmc
