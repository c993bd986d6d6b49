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
// the element's slot.  An interior point's residual depends on the element
// state through its value I_q and its gradients dI: the interior kernel
// covers the ndof point-value directions and, when asked, the 6 material
// parameters (rho0, mu, cp of components 0 and 1); the residual is linear in
// dI except through tau's element length and the shock-capturing
// coefficient, so k_vjp_grad forms the DIM*ndof gradient directions in FP64
// (dual numbers only for tau along its length and for that length).  The boundary kernel seeds
// the nen*ndof state and the parameter directions of the boundary integral
// (it does not depend on P).  The previous step enters the point residual
// linearly, through U(P) in the time term and (I - P)/dt in the
// least-squares residual: k_vjp_prev forms its derivative in FP64 in closed
// form, one quad per element.  Node sums use the same fixed-order gather as
// the assembly; parameter sums a fixed-order two-pass reduction.  No atomics,
// deterministic.
#include <hip/hip_runtime.h>

#include "gmc_dual.hpp"
#include "gmc_element.hpp"
#include "gmc_kernels.hpp"
#include "gmc_wave.hpp"

namespace gmc {

// One element's data for the FP64 adjoint kernels: node ids, coordinates,
// both states and v (0 on Dirichlet dofs).
template <int DIM, int NC>
__device__ __forceinline__ void vjp_gather(const VjpArgs& a, long long e, double* xe, double* Ie,
                                           double* Pe, double* ve) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN;
  GMC_UNROLL for (int n = 0; n < NEN; n++) {
    const int en = a.conn[e * NEN + n];
    GMC_UNROLL for (int k = 0; k < DIM; k++) xe[n * DIM + k] = a.coords[(long long)en * DIM + k];
    const unsigned dm = a.dmask[en];
    GMC_UNROLL for (int k = 0; k < N; k++) {
      const long long gi = (long long)en * N + k;
      Ie[n * N + k] = a.I[gi];
      Pe[n * N + k] = a.P[gi];
      ve[n * N + k] = ((dm >> k) & 1u) ? 0.0 : a.v[gi];
    }
  }
}

// The adjoint of the least-squares operator at one point:
//   f = sum_b first_part_b^T v_b,  w = sum_b N_b v_b,  y_k = sum_b dN_b,k v_b,
// with first_part_b^T x = A(c_b)^T x - N_b S0^T x, c_b = dN_b,
//   A(c)^T x = (c.v) DUDY^T x + c~ (U.x) + e_p (c~.x) + ((c.v) e_p + p c~) x_E,
//   S0^T x = (rho g)_v x_E
// (A(c) = sum_k c_k A_k and S0 = e_E (rho g)_v^T as adv_products and
// flux_interior form them).
template <int DIM, int NC>
__device__ __forceinline__ void lsq_adjoint(const DevParams& P, const QP<double, DIM, NC>& Q,
                                            const double* sh, const Geo<DIM>& g, const double* ve,
                                            double* f, double* w, double (*y)[Shape<DIM, NC>::NDOF]) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, IE = DIM + 1;
  GMC_UNROLL for (int j = 0; j < N; j++) {
    w[j] = 0.0;
    f[j] = 0.0;
    GMC_UNROLL for (int k = 0; k < DIM; k++) y[k][j] = 0.0;
  }
  GMC_UNROLL for (int b = 0; b < NEN; b++) {
    const double shb = sh[b];
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
      GMC_UNROLL for (int k = 0; k < DIM; k++) y[k][j] += c[k] * x[j];
    }
    f[0] += cx + cv * x[IE];
    GMC_UNROLL for (int k = 0; k < DIM; k++)
      f[1 + k] += c[k] * Ux + Q.p * c[k] * x[IE] - shb * (Q.pr.rho * P.g[k]) * x[IE];
  }
}

// v_e^T dr_e along the point values and the material parameters, by dual
// numbers, one instantiation per direction K: every derivative the
// direction does not seed is then a compile-time 0 and the compiler drops
// those dual parts.  K < ndof: the point value I_q moves along e_K (seeded
// after the interpolation, point_setup's SEED, so nothing of dI carries a
// derivative); lane q finds s_q = d(v_e . r_q)/dI_q,K and node a takes
// sum_q N_a(q) s_q; the gradient directions are k_vjp_grad's.  K = ndof + t:
// material parameter t (0..5; a viscosity direction, say, touches only the
// viscous terms).
template <int DIM, int NC, int K>
__global__ __launch_bounds__(256) void k_vjp_interior(const VjpArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  constexpr int THETA = K - N;  // material direction, or < 0
  int q, grp;  // quadrature point, group (triad / quad) in the workgroup
  const bool live = group_lane<DIM>(threadIdx.x, q, grp);
  const long long gid = (long long)blockIdx.x * groups_per_block<DIM>() + grp;
  if (gid >= a.n_items) return;  // whole groups leave together
  const long long e = a.items[gid];
  const DevParams& P = *a.params;
  Material<Dual> M;
  load_material<Dual, NC>(P, M);
  if constexpr (THETA >= 0) seed_theta<DIM, NC>(P, THETA, M);

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
      Ie[t] = Dual(a.I[gi], 0.0);
      Pe[t] = Dual(a.P[gi], 0.0);
      ve[t] = ((dm >> k) & 1u) ? 0.0 : a.v[gi];
    }
  }
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) {
    if (q == 0 && live) atomicOr(a.status, 1);
    return;
  }
  Point<Dual, DIM, NC> pt;
  const int st = point_setup<Dual, DIM, NC, false, (THETA < 0 ? K : -1)>(
      P, M, g, q < S::NQ ? q : S::NQ - 1, q < S::NQ ? 1.0 : 0.0, Ie, Pe, pt);
  if (live && st && q < S::NQ) atomicOr(a.status, st);
  Dual r[NE];
  point_residual<Dual, DIM, NC>(P, pt, g, r);
  double s = 0.0;
  GMC_UNROLL for (int t = 0; t < NE; t++) s += ve[t] * r[t].d;
  if constexpr (THETA < 0) {
    GMC_UNROLL for (int n = 0; n < NEN; n++) {
      const double sn = group_sum<DIM>(pt.sh[n] * s, q);
      if (q == 0 && live) a.gI[e * NE + n * N + K] = sn;
    }
  } else {
    s = group_sum<DIM>(s, q);
    if (q == 0 && live) a.gT[(long long)THETA * a.n_elems + e] = s;
  }
}

// v_e^T dr_e along the point gradients dI (dI_l = sum_a dN_a,l I_a), in FP64,
// one quad per element; added to k_vjp_interior's point-value part.  At a
// point the residual depends on dI linearly through
//   F_dif,k = sum_l K_kl dI_l,  Res = DUDY dIP + sum_l A_l dI_l - S
//   and the shock term shk sum_k dN_b,k DUDY dI_k,
// and nonlinearly only through tau (compute_h reads the velocity gradient)
// and shk = dc_nu(..., A dI, DUDY dI).  With f, y_k of lsq_adjoint,
//   d(v.r)/d dI_l,j = JxW [ (sum_k K_kl^T y_k)_j + (A_l^T tau^T f)_j
//                           + shk (DUDY^T y_l)_j + f^T (d tau/d dI_l,j) Res
//                           + (sum_k y_k . DUDY dI_k) (d shk/d dI_l,j) ],
// tau depends on dI only through its element length h = compute_h(v, dv):
// d tau/dh by one dual evaluation of tau_matrix, dh/d dv by dual evaluations
// of compute_h; d shk/d dI in closed form (dc_nu_grad).  Node a takes
// sum_l dN_a,l of it.
template <int DIM, int NC>
__global__ __launch_bounds__(256) void k_vjp_grad(const VjpArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  int q, grp;  // quadrature point, group (triad / quad) in the workgroup
  const bool live = group_lane<DIM>(threadIdx.x, q, grp);
  const long long gid = (long long)blockIdx.x * groups_per_block<DIM>() + grp;
  if (gid >= a.n_items) return;  // whole quads leave together
  const long long e = a.items[gid];
  const DevParams& P = *a.params;
  Material<double> M;
  load_material<double, NC>(P, M);
  double xe[NEN * DIM], Ie[NE], Pe[NE], ve[NE];
  vjp_gather<DIM, NC>(a, e, xe, Ie, Pe, ve);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) return;  // k_vjp_interior flags it
  Point<double, DIM, NC> pt;
  point_setup<double, DIM, NC>(P, M, g, q < S::NQ ? q : S::NQ - 1, q < S::NQ ? 1.0 : 0.0, Ie, Pe, pt);
  const auto& Q = pt.Q;
  double f[N], w[N], y[DIM][N];
  lsq_adjoint<DIM, NC>(P, Q, pt.sh, g, ve, f, w, y);
  double tf[N], Res[N];
  GMC_UNROLL for (int k = 0; k < N; k++) {
    double s = 0.0;
    GMC_UNROLL for (int i = 0; i < N; i++) s += pt.tau[i][k] * f[i];
    tf[k] = s;
  }
  // Res as point_residual forms it
  {
    const double idt = 1.0 / P.dt;
    double dIP[N];
    GMC_UNROLL for (int j = 0; j < N; j++) dIP[j] = (pt.Iq[j] - pt.Pq[j]) * idt;
    GMC_UNROLL for (int i = 0; i < N; i++) {
      double s = 0.0;
      GMC_UNROLL for (int j = 0; j < N; j++) s += Q.DUDY[i][j] * dIP[j];
      s += pt.AdI[i];
      Res[i] = s - Q.S[i];
    }
  }
  double L = 0.0;  // sum_k y_k . DUDY dI_k
  GMC_UNROLL for (int k = 0; k < DIM; k++) GMC_UNROLL for (int i = 0; i < N; i++) L += y[k][i] * pt.DdI[k][i];
  double gq[DIM][N];
  GMC_UNROLL for (int l = 0; l < DIM; l++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      double s = 0.0;
      GMC_UNROLL for (int k = 0; k < DIM; k++)
        GMC_UNROLL for (int i = 0; i < N; i++) s += Kentry<double, DIM, NC>(Q, k, l, i, j) * y[k][i];
      GMC_UNROLL for (int i = 0; i < N; i++) s += a_entry<double, DIM, NC>(Q, l, i, j) * tf[i];
      double dy = 0.0;
      GMC_UNROLL for (int i = 0; i < N; i++) dy += Q.DUDY[i][j] * y[l][i];
      gq[l][j] = s + pt.shk * dy;
    }
  // tau along its element length h (one dual evaluation), h along the
  // velocity gradient entries dI_l,(1 + i) = dv_i,l (compute_h's)
  {
    double dv[DIM][DIM];
    GMC_UNROLL for (int i = 0; i < DIM; i++) GMC_UNROLL for (int j = 0; j < DIM; j++) dv[i][j] = Q.dI[j][1 + i];
    const double h = compute_h<double, DIM>(g.G, Q.v, dv);
    QP<Dual, DIM, NC> qd;
    to_dual<DIM, NC>(Q, qd);
    Dual td[N][N];
    tau_matrix<Dual, DIM, NC, false, true>(P, qd, g, td, Dual(h, 1.0));
    double cT = 0.0;  // f^T (d tau/dh) Res
    GMC_UNROLL for (int r = 0; r < N; r++) {
      double tr = 0.0;
      GMC_UNROLL for (int k = 0; k < N; k++) tr += td[r][k].d * Res[k];
      cT += f[r] * tr;
    }
    Dual vd[DIM], dvd[DIM][DIM];
    GMC_UNROLL for (int i = 0; i < DIM; i++) {
      vd[i] = Dual(Q.v[i], 0.0);
      GMC_UNROLL for (int j = 0; j < DIM; j++) dvd[i][j] = Dual(dv[i][j], 0.0);
    }
    GMC_UNROLL for (int l = 0; l < DIM; l++)
      GMC_UNROLL for (int i = 0; i < DIM; i++) {
        dvd[i][l].d = 1.0;
        const Dual hd = compute_h<Dual, DIM>(g.G, vd, dvd);
        dvd[i][l].d = 0.0;
        gq[l][1 + i] += cT * hd.d;
      }
  }
  // the shock-capturing coefficient, in closed form
  {
    double dnu[DIM][N];
    dc_nu_grad<DIM, NC>(P, Q, g, M.uinv, pt.AdI, pt.DdI, dnu);
    GMC_UNROLL for (int l = 0; l < DIM; l++) GMC_UNROLL for (int j = 0; j < N; j++) gq[l][j] += L * dnu[l][j];
  }
  GMC_UNROLL for (int n = 0; n < NEN; n++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      double s = 0.0;
      GMC_UNROLL for (int l = 0; l < DIM; l++) s += g.dsh[l][n] * gq[l][j];
      s = group_sum<DIM>(pt.JxW * s, q);
      if (q == 0 && live) a.gI[e * NE + n * N + j] += s;
    }
}

// v_e^T dr_e/dP_e in FP64.  Per point (weight JxW, P_q = sum_a N_a P_a):
//   dr_b/dP_q = -(JxW/dt) (N_b DUDY(P) + first_part_b tau DUDY(I)),
// from the time term N_b (U(I) - U(P))/dt and the least-squares residual
// Res = DUDY(I) (I - P)/dt + ... (fluid_2d.hpp:256, 282); tau, nu and the
// properties depend on I only.  So with w, f of lsq_adjoint
//   g_P,q = -(JxW/dt) (DUDY(P)^T w + DUDY(I)^T tau^T f),
// and node a takes sum_q N_a(q) g_P,q.
template <int DIM, int NC>
__global__ __launch_bounds__(256) void k_vjp_prev(const VjpArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  int q, grp;  // quadrature point, group (triad / quad) in the workgroup
  const bool live = group_lane<DIM>(threadIdx.x, q, grp);
  const long long gid = (long long)blockIdx.x * groups_per_block<DIM>() + grp;
  if (gid >= a.n_items) return;  // whole quads leave together
  const long long e = a.items[gid];
  const DevParams& P = *a.params;
  Material<double> M;
  load_material<double, NC>(P, M);
  double xe[NEN * DIM], Ie[NE], Pe[NE], ve[NE];
  vjp_gather<DIM, NC>(a, e, xe, Ie, Pe, ve);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) return;  // k_vjp_interior flags it
  Point<double, DIM, NC> pt;
  point_setup<double, DIM, NC>(P, M, g, q < S::NQ ? q : S::NQ - 1, q < S::NQ ? 1.0 : 0.0, Ie, Pe, pt);
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
  double f[N], w[N], y[DIM][N];
  lsq_adjoint<DIM, NC>(P, Q, pt.sh, g, ve, f, w, y);
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
      const double s = group_sum<DIM>(pt.sh[n] * gq[k], q);
      if (q == 0 && live) a.gP[e * NE + n * N + k] = s;
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

// k_vjp_interior's directions K0 .. K1 - 1, one launch each
template <int DIM, int NC, int K0, int K1>
static hipError_t launch_vjp_dirs(const VjpArgs& a, dim3 grid, hipStream_t s) {
  if constexpr (K0 < K1) {
    hipLaunchKernelGGL((k_vjp_interior<DIM, NC, K0>), grid, dim3(256), 0, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_vjp_dirs<DIM, NC, K0 + 1, K1>(a, grid, s);
  } else {
    return hipSuccess;
  }
}

template <int DIM, int NC>
static hipError_t launch_vjp_shape(const VjpArgs& a, bool boundary, hipStream_t s) {
  if (a.n_items == 0) return hipSuccess;
  using S = Shape<DIM, NC>;
  // the boundary kernel seeds every nodal dof and the parameters (quads, 64
  // per workgroup); in the interior one launch per point-value direction and
  // per parameter (when asked), then k_vjp_grad for the gradients (one
  // group per element, groups_per_block triads or quads per workgroup)
  const long long per = boundary ? S::NE + (a.want_params ? 6 : 0) : S::NDOF;
  if ((a.n_items * per + 63) / 64 > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 block(256);
  if (boundary) {
    hipLaunchKernelGGL((k_vjp_boundary<DIM, NC>), dim3((unsigned)((a.n_items * per + 63) / 64)), block, 0,
                       s, a);
    return hipGetLastError();
  }
  constexpr int GB = groups_per_block<DIM>();
  const auto grid_of = [](long long groups) { return dim3((unsigned)((groups + GB - 1) / GB)); };
  const dim3 g1 = grid_of(a.n_items);
  hipError_t e = launch_vjp_dirs<DIM, NC, 0, S::NDOF>(a, g1, s);
  if (e != hipSuccess) return e;
  if (a.want_params && (e = launch_vjp_dirs<DIM, NC, S::NDOF, S::NDOF + 6>(a, g1, s)) != hipSuccess)
    return e;
  hipLaunchKernelGGL((k_vjp_grad<DIM, NC>), grid_of(a.n_items), block, 0, s, a);
  return hipGetLastError();
}

template <int DIM, int NC>
static hipError_t launch_vjp_prev_shape(const VjpArgs& a, hipStream_t s) {
  if (a.n_items == 0) return hipSuccess;
  constexpr int GB = groups_per_block<DIM>();
  const long long blocks = (a.n_items + GB - 1) / GB;
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
