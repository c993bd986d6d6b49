// gmc_adjoint.hip -- exact vector-Jacobian products of the assembled residual.
//
// The residual of the assembly is R = sum_e r_e(I_e, P_e, theta) with the
// rows of Dirichlet-blocked dofs removed (scatter.hpp:65-78 zeroes them).
// For an adjoint vector v the kernels form, per element,
//   g_I,e = (dr_e/dI_e)^T v_e,   g_P,e = (dr_e/dP_e)^T v_e,
//   g_theta,e = (dr_e/dtheta)^T v_e
// (v_e = v on the element's dofs with blocked rows set to 0) exactly, by
// forward-mode dual numbers: one quad (4 lanes = the quadrature points) per
// (element, input direction) evaluates the element residual of
// fluid_2d.hpp:171-202 / 209-245 with the direction seeded, sums its
// directional derivative against v_e over the points, and stores the scalar in
// the element's slot.  The interior kernel covers the nen*ndof state
// directions and, when asked, the 6 material parameters (rho0, mu, cp of
// components 0 and 1); the boundary kernel adds the state and parameter
// directions of the boundary integral (it does not depend on P).  The
// previous step enters the point residual linearly, through U(P) in the time
// term and (I - P)/dt in the least-squares residual, so its derivative needs
// no dual numbers: k_vjp_prev forms it in FP64 in closed form, one quad per
// element (the nen*ndof previous-step directions cost one pass instead of
// nen*ndof).  Node sums use the same fixed-order gather as the assembly;
// parameter sums a fixed-order two-pass reduction.  No atomics, deterministic.
#include <hip/hip_runtime.h>

#include "gmc_dual.hpp"
#include "gmc_element.hpp"
#include "gmc_kernels.hpp"
#include "gmc_wave.hpp"

namespace gmc {

template <int DIM, int NC>
__global__ __launch_bounds__(256) void k_vjp_interior(const VjpArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  const int ND = NE + (a.want_params ? 6 : 0);
  const int q = threadIdx.x & 3;
  const long long gid = (long long)blockIdx.x * (blockDim.x >> 2) + (threadIdx.x >> 2);
  if (gid >= a.n_items * ND) return;  // whole quads leave together
  const long long it = gid / ND;
  const int d = (int)(gid - it * ND);
  const long long e = a.items[it];
  const DevParams& P = *a.params;
  Material<Dual> M;
  load_material<Dual, NC>(P, M);
  if (d >= NE) seed_theta<DIM, NC>(P, d - NE, M);

  int en[NEN];
  double xe[NEN * DIM];
  Dual Ie[NE], Pe[NE];
  double ve[NE];
  GMC_UNROLL for (int n = 0; n < NEN; n++) {
    en[n] = a.conn[e * NEN + n];
    GMC_UNROLL for (int k = 0; k < DIM; k++) xe[n * DIM + k] = a.coords[(long long)en[n] * DIM + k];
    const unsigned dm = a.dmask[en[n]];
    GMC_UNROLL for (int k = 0; k < N; k++) {
      const int t = n * N + k;
      const long long gi = (long long)en[n] * N + k;
      Ie[t] = Dual(a.I[gi], t == d ? 1.0 : 0.0);
      Pe[t] = Dual(a.P[gi], 0.0);
      ve[t] = ((dm >> k) & 1u) ? 0.0 : a.v[gi];
    }
  }
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) {
    if (q == 0) atomicOr(a.status, 1);
    return;
  }
  Point<Dual, DIM, NC> pt;
  const int st = point_setup<Dual, DIM, NC>(P, M, g, q < S::NQ ? q : S::NQ - 1,
                                            q < S::NQ ? 1.0 : 0.0, Ie, Pe, pt);
  if (st && q < S::NQ) atomicOr(a.status, st);
  Dual r[NE];
  point_residual<Dual, DIM, NC>(P, pt, g, r);
  double s = 0.0;
  GMC_UNROLL for (int t = 0; t < NE; t++) s += ve[t] * r[t].d;
  s = quad_sum(s);
  if (q == 0) {
    if (d < NE) a.gI[e * NE + d] = s;
    else a.gT[(long long)(d - NE) * a.n_elems + e] = s;
  }
}

// v_e^T dr_e/dP_e in FP64.  Per point (weight JxW, P_q = sum_a N_a P_a):
//   dr_b/dP_q = -(JxW/dt) (N_b DUDY(P) + first_part_b tau DUDY(I)),
// from the time term N_b (U(I) - U(P))/dt and the least-squares residual
// Res = DUDY(I) (I - P)/dt + ... (fluid_2d.hpp:256, 282); tau, nu and the
// properties depend on I only.  So with w = sum_b N_b v_b and
// f = sum_b first_part_b^T v_b,
//   g_P,q = -(JxW/dt) (DUDY(P)^T w + DUDY(I)^T tau^T f),
// first_part_b^T x = A(c_b)^T x - N_b S0^T x with
//   A(c)^T x = (c.v) DUDY^T x + c~ (U.x) + e_p (c~.x) + ((c.v) e_p + p c~) x_E,
//   S0^T x = (rho g)_v x_E,
// and node a takes sum_q N_a(q) g_P,q.
template <int DIM, int NC>
__global__ __launch_bounds__(256) void k_vjp_prev(const VjpArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE, IE = DIM + 1;
  const int q = threadIdx.x & 3;
  const long long gid = (long long)blockIdx.x * (blockDim.x >> 2) + (threadIdx.x >> 2);
  if (gid >= a.n_items) return;  // whole quads leave together
  const long long e = a.items[gid];
  const DevParams& P = *a.params;
  Material<double> M;
  load_material<double, NC>(P, M);
  int en[NEN];
  double xe[NEN * DIM], Ie[NE], Pe[NE], ve[NE];
  GMC_UNROLL for (int n = 0; n < NEN; n++) {
    en[n] = a.conn[e * NEN + n];
    GMC_UNROLL for (int k = 0; k < DIM; k++) xe[n * DIM + k] = a.coords[(long long)en[n] * DIM + k];
    const unsigned dm = a.dmask[en[n]];
    GMC_UNROLL for (int k = 0; k < N; k++) {
      const int t = n * N + k;
      const long long gi = (long long)en[n] * N + k;
      Ie[t] = a.I[gi];
      Pe[t] = a.P[gi];
      ve[t] = ((dm >> k) & 1u) ? 0.0 : a.v[gi];
    }
  }
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) {
    if (q == 0) atomicOr(a.status, 1);
    return;
  }
  Point<double, DIM, NC> pt;
  const int st = point_setup<double, DIM, NC>(P, M, g, q < S::NQ ? q : S::NQ - 1,
                                              q < S::NQ ? 1.0 : 0.0, Ie, Pe, pt);
  if (st && q < S::NQ) atomicOr(a.status, st);
  const auto& Q = pt.Q;
  // DUDY at the P state (its fluxes are formed and not used)
  QP<double, DIM, NC> QPp;
  {
    double Pq[N];
    interpolate<double, DIM, NC>(pt.sh, Pe, Pq);
    assign<double, DIM, NC>(Pq, QPp.p, QPp.v, QPp.Tt, QPp.Y);
    GMC_UNROLL for (int k = 0; k < DIM; k++) GMC_UNROLL for (int j = 0; j < N; j++) QPp.dI[k][j] = 0.0;
    properties<double, NC, MG_NONE>(P, M, QPp.p, QPp.Tt, QPp.Y, QPp.pr, 0.0, nullptr);
    flux_interior<double, DIM, NC, false>(P, QPp);
  }
  double w[N], f[N];
  GMC_UNROLL for (int j = 0; j < N; j++) {
    w[j] = 0.0;
    f[j] = 0.0;
  }
  GMC_UNROLL for (int b = 0; b < NEN; b++) {
    const double shb = pt.sh[b];
    const double* x = ve + b * N;
    double c[DIM], cv = 0.0, cx = 0.0, Ux = 0.0;
    GMC_UNROLL for (int k = 0; k < DIM; k++) {
      c[k] = g.dsh[k][b];
      cv += c[k] * Q.v[k];
      cx += c[k] * x[1 + k];
    }
    GMC_UNROLL for (int i = 0; i < N; i++) Ux += Q.U[i] * x[i];
    GMC_UNROLL for (int j = 0; j < N; j++) {
      double dtx = 0.0;
      GMC_UNROLL for (int i = 0; i < N; i++) dtx += Q.DUDY[i][j] * x[i];
      f[j] += cv * dtx;
      w[j] += shb * x[j];
    }
    f[0] += cx + cv * x[IE];
    GMC_UNROLL for (int k = 0; k < DIM; k++)
      f[1 + k] += c[k] * Ux + Q.p * c[k] * x[IE] - shb * (Q.pr.rho * P.g[k]) * x[IE];
  }
  double tf[N];
  GMC_UNROLL for (int k = 0; k < N; k++) {
    double s = 0.0;
    GMC_UNROLL for (int i = 0; i < N; i++) s += pt.tau[i][k] * f[i];
    tf[k] = s;
  }
  const double sc = -pt.JxW / P.dt;
  double gq[N];
  GMC_UNROLL for (int j = 0; j < N; j++) {
    double s = 0.0;
    GMC_UNROLL for (int i = 0; i < N; i++) s += QPp.DUDY[i][j] * w[i] + Q.DUDY[i][j] * tf[i];
    gq[j] = sc * s;
  }
  GMC_UNROLL for (int n = 0; n < NEN; n++)
    GMC_UNROLL for (int k = 0; k < N; k++) {
      const double s = quad_sum(pt.sh[n] * gq[k]);
      if (q == 0) a.gP[e * NE + n * N + k] = s;
    }
}

template <int DIM, int NC>
__global__ __launch_bounds__(256) void k_vjp_boundary(const VjpArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  const int q = threadIdx.x & 3;
  const long long gid = (long long)blockIdx.x * (blockDim.x >> 2) + (threadIdx.x >> 2);
  const int ND = NE + (a.want_params ? 6 : 0);
  if (gid >= a.n_items * ND) return;
  const long long it = gid / ND;
  const int d = (int)(gid - it * ND);
  const long long e = a.items[it];
  const DevParams& P = *a.params;
  Material<Dual> M;
  load_material<Dual, NC>(P, M);
  if (d >= NE) seed_theta<DIM, NC>(P, d - NE, M);
  int en[NEN];
  double xe[NEN * DIM];
  Dual Ie[NE];
  double ve[NE];
  GMC_UNROLL for (int n = 0; n < NEN; n++) {
    en[n] = a.conn[e * NEN + n];
    GMC_UNROLL for (int k = 0; k < DIM; k++) xe[n * DIM + k] = a.coords[(long long)en[n] * DIM + k];
    const unsigned dm = a.dmask[en[n]];
    GMC_UNROLL for (int k = 0; k < N; k++) {
      const int t = n * N + k;
      const long long gi = (long long)en[n] * N + k;
      Ie[t] = Dual(a.I[gi], t == d ? 1.0 : 0.0);
      ve[t] = ((dm >> k) & 1u) ? 0.0 : a.v[gi];
    }
  }
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) {
    if (q == 0) atomicOr(a.status, 1);
    return;
  }
  const int s0 = a.side_start[e], s1 = a.side_start[e + 1];
  const int npairs = (s1 - s0) * S::NSQ;
  double s = 0.0;
  for (int base = 0; base < npairs; base += 4) {
    const int t = base + q;
    const bool act = t < npairs;
    const int tt = act ? t : (npairs - 1);
    const int sidx = s0 + tt / S::NSQ;
    BPoint<Dual, DIM, NC> bp;
    bpoint_setup<Dual, DIM, NC>(P, M, g, xe, a.side_local[sidx], tt % S::NSQ, a.side_flag[sidx],
                                act ? 1.0 : 0.0, Ie, bp);
    Dual r[NE];
    bpoint_residual<Dual, DIM, NC>(bp, r);
    GMC_UNROLL for (int ti = 0; ti < NE; ti++) s += ve[ti] * r[ti].d;
  }
  s = quad_sum(s);
  if (q == 0) {
    if (d < NE) a.gI[e * NE + d] += s;
    else a.gT[(long long)(d - NE) * a.n_elems + e] += s;
  }
}

template <int DIM, int NC>
static hipError_t launch_vjp_shape(const VjpArgs& a, bool boundary, hipStream_t s) {
  if (a.n_items == 0) return hipSuccess;
  constexpr int NE = Shape<DIM, NC>::NE;
  const long long quads = a.n_items * (NE + (a.want_params ? 6 : 0));
  const long long blocks = (quads + 63) / 64;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks), block(256);
  if (boundary) hipLaunchKernelGGL((k_vjp_boundary<DIM, NC>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((k_vjp_interior<DIM, NC>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <int DIM, int NC>
static hipError_t launch_vjp_prev_shape(const VjpArgs& a, hipStream_t s) {
  if (a.n_items == 0) return hipSuccess;
  const long long blocks = (a.n_items + 63) / 64;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_vjp_prev<DIM, NC>), dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_vjp_prev(int dim, int nc, const VjpArgs& a, hipStream_t s) {
  if (dim == 2 && nc == 2) return launch_vjp_prev_shape<2, 2>(a, s);
  if (dim == 2 && nc == 3) return launch_vjp_prev_shape<2, 3>(a, s);
  if (dim == 3 && nc == 3) return launch_vjp_prev_shape<3, 3>(a, s);
  if (dim == 3 && nc == 4) return launch_vjp_prev_shape<3, 4>(a, s);
  return hipErrorInvalidValue;
}

hipError_t launch_vjp(int dim, int nc, bool boundary, const VjpArgs& a, hipStream_t s) {
  if (dim == 2 && nc == 2) return launch_vjp_shape<2, 2>(a, boundary, s);
  if (dim == 2 && nc == 3) return launch_vjp_shape<2, 3>(a, boundary, s);
  if (dim == 3 && nc == 3) return launch_vjp_shape<3, 3>(a, boundary, s);
  if (dim == 3 && nc == 4) return launch_vjp_shape<3, 4>(a, boundary, s);
  return hipErrorInvalidValue;
}

}  // namespace gmc
