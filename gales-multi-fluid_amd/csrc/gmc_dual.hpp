// gmc_dual.hpp -- forward-mode dual numbers for the exact adjoint.
//
// Dual carries a value and one directional derivative.  Its arithmetic is
// never contracted into FMAs, so the value part of a dual evaluation rounds
// exactly like the plain FP64 code where that code is contraction-free (the
// geometry, interpolation and compute_h, whose epsilon branches depend on the
// last bits): the dual evaluation then takes the same branches as the plain
// assembly it differentiates.  Instantiating the
// physics templates (gmc_physics.hpp, gmc_element.hpp) with T = Dual gives the
// exact derivative of the element residual along one seeded direction, with
// every coefficient differentiated: properties and the equation of state,
// flux matrices, tau (including the LU inverse and the min() switches),
// compute_h and the dc_2006 shock-capturing viscosity.  Branches (pivoting,
// min, the epsilon guards) are decided on the value, so the derivative is the
// one of the branch taken, as a one-sided derivative at a switch.
#pragma once

#include <cmath>

#include "gmc_physics.hpp"

namespace gmc {

struct Dual {
  double v, d;
  Dual() = default;
  GMC_HD constexpr Dual(double x) : v(x), d(0.0) {}
  GMC_HD constexpr Dual(double x, double dx) : v(x), d(dx) {}
  GMC_HD Dual& operator+=(const Dual& b) {
    GMC_NO_CONTRACT
    v += b.v;
    d += b.d;
    return *this;
  }
  GMC_HD Dual& operator-=(const Dual& b) {
    GMC_NO_CONTRACT
    v -= b.v;
    d -= b.d;
    return *this;
  }
  GMC_HD Dual& operator*=(const Dual& b) {
    GMC_NO_CONTRACT
    d = d * b.v + v * b.d;
    v *= b.v;
    return *this;
  }
  GMC_HD Dual& operator/=(const Dual& b) {
    GMC_NO_CONTRACT
    const double q = v / b.v;
    d = (d - q * b.d) / b.v;
    v = q;
    return *this;
  }
};

GMC_HD double scalar_of(const Dual& x) { return x.v; }

GMC_HD Dual operator-(const Dual& a) {
  GMC_NO_CONTRACT
  return Dual(-a.v, -a.d);
}
GMC_HD Dual operator+(const Dual& a, const Dual& b) {
  GMC_NO_CONTRACT
  return Dual(a.v + b.v, a.d + b.d);
}
GMC_HD Dual operator-(const Dual& a, const Dual& b) {
  GMC_NO_CONTRACT
  return Dual(a.v - b.v, a.d - b.d);
}
GMC_HD Dual operator*(const Dual& a, const Dual& b) {
  GMC_NO_CONTRACT
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
GMC_HD Dual operator/(const Dual& a, const Dual& b) {
  GMC_NO_CONTRACT
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
GMC_HD Dual operator+(const Dual& a, double b) {
  GMC_NO_CONTRACT
  return Dual(a.v + b, a.d);
}
GMC_HD Dual operator+(double a, const Dual& b) {
  GMC_NO_CONTRACT
  return Dual(a + b.v, b.d);
}
GMC_HD Dual operator-(const Dual& a, double b) {
  GMC_NO_CONTRACT
  return Dual(a.v - b, a.d);
}
GMC_HD Dual operator-(double a, const Dual& b) {
  GMC_NO_CONTRACT
  return Dual(a - b.v, -b.d);
}
GMC_HD Dual operator*(const Dual& a, double b) {
  GMC_NO_CONTRACT
  return Dual(a.v * b, a.d * b);
}
GMC_HD Dual operator*(double a, const Dual& b) {
  GMC_NO_CONTRACT
  return Dual(a * b.v, a * b.d);
}
GMC_HD Dual operator/(const Dual& a, double b) {
  GMC_NO_CONTRACT
  return Dual(a.v / b, a.d / b);
}
GMC_HD Dual operator/(double a, const Dual& b) {
  GMC_NO_CONTRACT
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}

#define GMC_DUAL_CMP(OP)                                                            \
  GMC_HD bool operator OP(const Dual& a, const Dual& b) { return a.v OP b.v; }     \
  GMC_HD bool operator OP(const Dual& a, double b) { return a.v OP b; }            \
  GMC_HD bool operator OP(double a, const Dual& b) { return a OP b.v; }
GMC_DUAL_CMP(<)
GMC_DUAL_CMP(>)
GMC_DUAL_CMP(<=)
GMC_DUAL_CMP(>=)
GMC_DUAL_CMP(==)
GMC_DUAL_CMP(!=)
#undef GMC_DUAL_CMP

// sqrt: derivative 0 at 0 (the physics only takes sqrt(0) of vanishing
// norms, whose guards then discard the result)
GMC_HD Dual sqrt(const Dual& a) {
  GMC_NO_CONTRACT
  const double s = std::sqrt(a.v);
  return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
GMC_HD Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
GMC_HD Dual pow(const Dual& a, double e) {
  GMC_NO_CONTRACT
  const double p = std::pow(a.v, e);
  return Dual(p, a.v != 0.0 ? e * p / a.v * a.d : 0.0);
}

// Seeds d(theta_k) = 1 in the material, theta = (rho0, mu, cp) x components
// 0, 1.  An ideal gas is parametrised by (R, mu, cp) with gamma = cp/(cp - R)
// (so rho = p/(RT), cv = R/(gamma-1) = cp - R); a custom component by its
// constant density, viscosity and heat capacity (chemicals.hpp:23-47,
// 201-264).  u_ref_inv depends on the material and is re-derived with the
// seed (its value is kept bit-identical to the plain form's).
template <int DIM, int NC>
GMC_HD void seed_theta(const DevParams& P, int k, Material<Dual>& M) {
  const int c = k / 3, w = k - 3 * (k / 3);
  // component cc = c by predicated writes at compile-time indices: a write
  // at the run-time index c would keep the Material record in scratch
  GMC_UNROLL for (int cc = 0; cc < 2; cc++) {
    if (cc != c) continue;
    if (P.kind[cc] == 0) {
      if (w == 0) M.rho[cc].d = 1.0;
      else if (w == 1) M.mu[cc].d = 1.0;
      else M.cp[cc].d = 1.0;
    } else {
      const double R = P.R[cc], gm = P.gamma[cc];
      const double cp = gm * R / (gm - 1.0), den = (cp - R) * (cp - R);
      if (w == 0) {
        M.R[cc].d = 1.0;
        M.gamma[cc].d = cp / den;
      } else if (w == 1) {
        M.mu[cc].d = 1.0;
      } else {
        M.gamma[cc].d = -R / den;
      }
    }
  }
  Dual u[8];
  uref_inv_of<Dual, DIM, NC, MG_NONE>(P, M, u);  // no magma: the exact mode rejects it
  GMC_UNROLL for (int i = 0; i < 8; i++) M.uinv[i] = Dual(P.uref_inv[i], u[i].d);
}

// The point record q as dual numbers with zero derivatives: the fields that
// tau_matrix and dc_nu read (state, properties, gradients, U, DUDY and its
// factors, S, the diffusion coefficients).  Fully unrolled, a dual
// evaluation from it along one seeded entry folds every product with a zero
// derivative, so it costs little more than the plain one.
template <int DIM, int NC>
GMC_HD void to_dual(const QP<double, DIM, NC>& a, QP<Dual, DIM, NC>& b) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  b.p = a.p;
  b.Tt = a.Tt;
  b.E = a.E;
  GMC_UNROLL for (int k = 0; k < DIM; k++) b.v[k] = a.v[k];
  GMC_UNROLL for (int c = 0; c < NC; c++) {
    b.Y[c] = a.Y[c];
    b.pr.rho_c[c] = a.pr.rho_c[c];
    b.pr.e_c[c] = a.pr.e_c[c];
    b.pr.D_c[c] = a.pr.D_c[c];
    b.rhoD[c] = a.rhoD[c];
    b.rhoDH[c] = a.rhoDH[c];
  }
  b.pr.rho = a.pr.rho;
  b.pr.mu = a.pr.mu;
  b.pr.cp = a.pr.cp;
  b.pr.cv = a.pr.cv;
  b.pr.kappa = a.pr.kappa;
  b.pr.c = a.pr.c;
  b.pr.alpha = a.pr.alpha;
  b.pr.beta = a.pr.beta;
  b.pr.st = a.pr.st;
  b.lam = a.lam;
  b.mu = a.mu;
  b.kappa = a.kappa;
  GMC_UNROLL for (int i = 0; i < N; i++) {
    b.U[i] = a.U[i];
    b.S[i] = a.S[i];
    b.d0[i] = a.d0[i];
    b.mE[i] = a.mE[i];
    GMC_UNROLL for (int j = 0; j < N; j++) b.DUDY[i][j] = a.DUDY[i][j];
    GMC_UNROLL for (int k = 0; k < DIM; k++) b.dI[k][i] = a.dI[k][i];
  }
}

}  // namespace gmc
