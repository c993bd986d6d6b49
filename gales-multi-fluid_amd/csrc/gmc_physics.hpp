// gmc_physics.hpp -- per-quadrature-point physics of the fluid_mc element.
//
// Host+device, templated on the scalar type T (double for the FP64 form and
// the adjoint's dual numbers, float for the FP32 tier's matrices), on DIM
// (2: triangles, 3: tetrahedra) and on NC (components).  All arrays are
// indexed with compile-time constants after unrolling so the compiler keeps
// them in registers and folds the structural zeros.
//
// Reference behaviour restated (GALES-multi-fluid):
//   properties          src/fem/physical_props/fluid/fluid_properties.hpp:171-226
//                       chemicals.hpp:19-49 (custom), chemicals.hpp:198-271 (ideal gas),
//                       magma submixtures in gmc_magma.hpp
//   geometry            src/fem/mesh/quad.hpp:148-383 (2D), 470-739 (3D)
//   flux matrices       sim/fluid_mc/P0/fluid_2d.hpp:391-603 (interior), 610-868 (boundary)
//   tau                 fluid_2d.hpp:877-1022 (2001, 2019, 2007, 2014 forms);
//                       compute_h quad.hpp:356-383, 701-739
//   inverse             src/fem/base/algorithm.hpp:399-414 (uBLAS LU), here DUDY^-1 in closed form
//   dc_2006             fluid_2d.hpp:1027-1065
// The 3D element is not in the reference; it is built by the same
// construction (A3, K13..K33, 3D compute_h, trace(G) in the Mach correction).
#pragma once

#include <cmath>
#include <cstdint>
#include <type_traits>

#if defined(__HIPCC__)
#define GMC_HD __host__ __device__ __forceinline__
#else
#define GMC_HD inline
#endif

#define GMC_UNROLL _Pragma("unroll")

// Quantities that feed the branch decisions of quad::compute_h
// (quad.hpp:366-375 compares the unnormalised direction s1 with an absolute
// epsilon) are computed without FMA contraction, in the reference's operation
// order, so they are bitwise reproducible across the GPU and a CPU build.
#if defined(__clang__)
#define GMC_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define GMC_NO_CONTRACT
#endif

#include "gmc_magma.hpp"

namespace gmc {

template <int DIM_, int NC_>
struct Shape {
  static constexpr int DIM = DIM_;
  static constexpr int NC = NC_;
  static constexpr int NDOF = DIM_ + 2 + NC_ - 1;
  static constexpr int NEN = DIM_ + 1;
  static constexpr int NQ = DIM_ + 1;       // interior points (3-pt triangle, 4-pt tet)
  static constexpr int NSQ = DIM_;          // points per side (2-pt edge, 3-pt face)
  static constexpr int NE = NEN * NDOF;     // element block size
  static constexpr int IT = DIM_ + 1;       // T (primitive) == rho*E (conservative) index
  static constexpr int IY = DIM_ + 2;       // first species index
};

// Device copy of the physics parameters (flattened gmc_params).
struct DevParams {
  int kind[4];
  double rho[4], mu[4], cp[4], kappa[4], alpha[4], beta[4];
  double R[4], gamma[4];      // ideal gas: R = 8.314462 / molar_mass (chemicals.hpp:243)
  double D[4];                // chemical diffusivity
  double g[3];
  double dt;
  int steady;
  int tau_variant;            // GMC_TAU_*: 2001, 2007, 2014 or 2019
  int dc;
  double dc_sharp, dc_scale;
  double uref_inv[8];
  double p_ref, v_ref[3], T_ref;   // state of u_ref (fluid_2d.hpp:123-135)
  unsigned dmask_by_flag[16];
  // Neumann overrides by side flag
  unsigned n_tau_mask[16];
  double n_tau[16][6];
  unsigned n_q_mask[16];
  double n_q[16][3];
  unsigned n_J_mask[16][3];
  double n_J[16][3][4];
  // prescribed mass flux by side flag: bit k of mflux_mask, value mflux[k]
  unsigned mflux_mask[16];
  double mflux[16][3];
  // magma submixtures (kind 2), tables in device memory on the device copy
  MagmaDev mg[4];
  int need_sr;   // a magma viscosity model depends on the strain rate
};

static constexpr double kEps = 2.220446049250313e-16;  // numeric_limits<double>::epsilon

template <typename T>
GMC_HD T tmin(T a, T b) { return (b < a) ? b : a; }

// ------------------------------------------------------------------------
// material parameters as scalars of type T (dual numbers carry seeds here)
// ------------------------------------------------------------------------
template <typename T>
struct Material {
  T rho[4], mu[4], cp[4], kappa[4], alpha[4], beta[4], R[4], gamma[4], D[4];
  T uinv[8];   // diag(U_ref)^-1 of dc_2006 (depends on the material)
};

template <typename T, int NC>
GMC_HD void load_material(const DevParams& P, Material<T>& m) {
  GMC_UNROLL for (int c = 0; c < NC; c++) {
    m.rho[c] = T(P.rho[c]);
    m.mu[c] = T(P.mu[c]);
    m.cp[c] = T(P.cp[c]);
    m.kappa[c] = T(P.kappa[c]);
    m.alpha[c] = T(P.alpha[c]);
    m.beta[c] = T(P.beta[c]);
    m.R[c] = T(P.R[c]);
    m.gamma[c] = T(P.gamma[c]);
    m.D[c] = T(P.D[c]);
  }
  GMC_UNROLL for (int i = 0; i < 8; i++) m.uinv[i] = T(P.uref_inv[i]);
}

template <typename T, int NC>
struct Props {
  T rho, mu, cp, cv, kappa, c, alpha, beta;
  T rho_c[NC], e_c[NC], D_c[NC];
  int st;   // kStTable / kStModel from a magma component, else 0
};

// the FP64 value of a scalar (the dual-number type adds its overload)
GMC_HD double scalar_of(double x) { return x; }

// Magma component values at one point and state, as the point tables hold
// them: rho, mu, cp, cv, kappa, c, alpha, beta, D (gmc_magma.hpp)
static constexpr int kMagmaVals = 9;

// How properties() obtains a magma component's values:
//   MG_NONE   no magma support (the dual-number evaluation; the library
//             rejects the exact adjoint for magma mixtures),
//   MG_CALL   magma_props at (p, T, sr),
//   MG_TABLE  read from the point table mt (kMagmaVals per component), filled
//             by the magma point kernels before the element kernels run: the
//             element kernels then carry no magma code in their register
//             allocation.
enum { MG_NONE = 0, MG_CALL = 1, MG_TABLE = 2 };

// fluid_properties::properties (fluid_properties.hpp:171-226): the
// mix_of_ch and mix_of_mix branches share their formulas; a component is a
// custom chemical, an ideal gas or a magma submixture (at strain rate sr).
template <typename T, int NC, int MG = MG_CALL>
GMC_HD void properties(const DevParams& P, const Material<T>& M, T p, T Tt, const T* Y,
                       Props<T, NC>& o, T sr = T(0.0), const double* mt = nullptr) {
  using std::sqrt;
  T rho = T(0.0), mu = T(0.0), cp = T(0.0), cv = T(0.0), kappa = T(0.0), snd = T(0.0);
  T ac[NC], bc[NC];
  o.st = 0;
  GMC_UNROLL for (int c = 0; c < NC; c++) {
    T rc, muc, cpc, cvc, kc, alc, bec, sc, Dc = M.D[c];
    if (MG != MG_NONE && P.kind[c] == 2) {  // magma_mixture (mixtures.hpp:172-191)
      double v[kMagmaVals];
      if constexpr (MG == MG_TABLE) {
        GMC_UNROLL for (int k = 0; k < kMagmaVals; k++) v[k] = mt[c * kMagmaVals + k];
      } else {
        const MagmaProps mp = magma_props(P.mg[c], scalar_of(p), scalar_of(Tt), scalar_of(sr));
        o.st |= mp.st;
        v[0] = mp.rho;
        v[1] = mp.mu;
        v[2] = mp.cp;
        v[3] = mp.cv;
        v[4] = mp.kappa;
        v[5] = mp.c;
        v[6] = mp.alpha;
        v[7] = mp.beta;
        v[8] = mp.D;
      }
      rc = T(v[0]);
      muc = T(v[1]);
      cpc = T(v[2]);
      cvc = T(v[3]);
      kc = T(v[4]);
      sc = T(v[5]);
      alc = T(v[6]);
      bec = T(v[7]);
      Dc = T(v[8]);
    } else if (P.kind[c] == 0) {  // custom_chemical
      rc = M.rho[c];
      muc = M.mu[c];
      cpc = M.cp[c];
      alc = M.alpha[c];
      bec = M.beta[c];
      kc = M.kappa[c];
      if (P.beta[c] == 0.0) {
        sc = T(1.e20);
        cvc = cpc;
      } else {
        sc = T(1.0) / sqrt(rc * bec);
        cvc = cpc - Tt * alc * alc / (rc * bec);
      }
    } else {  // ideal_gas
      const T R = M.R[c];
      cvc = R / (M.gamma[c] - T(1.0));
      cpc = M.gamma[c] * cvc;
      muc = M.mu[c];
      kc = muc * cpc / T(0.72);
      rc = p / (R * Tt);
      bec = T(1.0) / p;
      alc = T(1.0) / Tt;
      sc = sqrt(M.gamma[c] * R * Tt);
    }
    rho += Y[c] / rc;
    mu += Y[c] * muc;
    cp += Y[c] * cpc;
    cv += Y[c] * cvc;
    kappa += Y[c] * kc;
    snd += Y[c] / ((rc * rc) * (sc * sc));
    o.rho_c[c] = rc;
    o.e_c[c] = cvc * Tt;
    o.D_c[c] = Dc;
    ac[c] = alc;
    bc[c] = bec;
  }
  rho = T(1.0) / rho;
  snd = sqrt(T(1.0) / snd);
  snd /= rho;
  T al = T(0.0), be = T(0.0);
  GMC_UNROLL for (int c = 0; c < NC; c++) {
    al += Y[c] * ac[c] * rho / o.rho_c[c];
    be += Y[c] * bc[c] * rho / o.rho_c[c];
  }
  o.rho = rho;
  o.mu = mu;
  o.cp = cp;
  o.cv = cv;
  o.kappa = kappa;
  o.c = snd;
  o.alpha = al;
  o.beta = be;
}

// ------------------------------------------------------------------------
// geometry of a P1 element (constant gradients)
// ------------------------------------------------------------------------
template <int DIM>
struct Geo {
  double J;
  double dsh[DIM][DIM + 1];   // dN_a / dx_k
  double G[DIM][DIM];
};

// quad::calc + g_up_g_G; xe[a*DIM + k].  Returns false when J <= 0.
template <int DIM>
GMC_HD bool geometry(const double* xe, Geo<DIM>& g) {
  GMC_NO_CONTRACT
  if constexpr (DIM == 2) {
    // dref = [[-1,1,0],[-1,0,1]]
    const double dx_dxi = -xe[0] + xe[2];
    const double dx_deta = -xe[0] + xe[4];
    const double dy_dxi = -xe[1] + xe[3];
    const double dy_deta = -xe[1] + xe[5];
    const double J = dx_dxi * dy_deta - dy_dxi * dx_deta;
    g.J = J;
    const double dxi_dx = (1 / J) * (dy_deta);
    const double deta_dx = (1 / J) * (-dy_dxi);
    const double dxi_dy = (1 / J) * (-dx_deta);
    const double deta_dy = (1 / J) * (dx_dxi);
    g.dsh[0][0] = -dxi_dx - deta_dx;
    g.dsh[0][1] = dxi_dx;
    g.dsh[0][2] = deta_dx;
    g.dsh[1][0] = -dxi_dy - deta_dy;
    g.dsh[1][1] = dxi_dy;
    g.dsh[1][2] = deta_dy;
    g.G[0][0] = dxi_dx * dxi_dx + deta_dx * deta_dx;
    g.G[1][1] = dxi_dy * dxi_dy + deta_dy * deta_dy;
    g.G[0][1] = dxi_dx * dxi_dy + deta_dx * deta_dy;
    g.G[1][0] = g.G[0][1];
    return J > 0.0;
  } else {
    double d[3][3];  // d[c][k] = dx_c/dxi_k
    GMC_UNROLL for (int c = 0; c < 3; c++) {
      d[c][0] = -xe[c] + xe[3 + c];
      d[c][1] = -xe[c] + xe[6 + c];
      d[c][2] = -xe[c] + xe[9 + c];
    }
    const double dx_dxi = d[0][0], dx_deta = d[0][1], dx_dzeta = d[0][2];
    const double dy_dxi = d[1][0], dy_deta = d[1][1], dy_dzeta = d[1][2];
    const double dz_dxi = d[2][0], dz_deta = d[2][1], dz_dzeta = d[2][2];
    const double J = dx_dxi * (dy_deta * dz_dzeta - dz_deta * dy_dzeta) -
                     dy_dxi * (dx_deta * dz_dzeta - dz_deta * dx_dzeta) +
                     dz_dxi * (dx_deta * dy_dzeta - dy_deta * dx_dzeta);
    g.J = J;
    double inv[3][3];  // inv[k][c] = dxi_k/dx_c
    inv[0][0] = (1 / J) * (dy_deta * dz_dzeta - dz_deta * dy_dzeta);
    inv[1][0] = (1 / J) * (-dy_dxi * dz_dzeta + dz_dxi * dy_dzeta);
    inv[2][0] = (1 / J) * (dy_dxi * dz_deta - dz_dxi * dy_deta);
    inv[0][1] = (1 / J) * (-dx_deta * dz_dzeta + dz_deta * dx_dzeta);
    inv[1][1] = (1 / J) * (dx_dxi * dz_dzeta - dz_dxi * dx_dzeta);
    inv[2][1] = (1 / J) * (-dx_dxi * dz_deta + dz_dxi * dx_deta);
    inv[0][2] = (1 / J) * (dx_deta * dy_dzeta - dy_deta * dx_dzeta);
    inv[1][2] = (1 / J) * (-dx_dxi * dy_dzeta + dy_dxi * dx_dzeta);
    inv[2][2] = (1 / J) * (dx_dxi * dy_deta - dy_dxi * dx_deta);
    GMC_UNROLL for (int c = 0; c < 3; c++) {
      g.dsh[c][0] = -inv[0][c] - inv[1][c] - inv[2][c];
      g.dsh[c][1] = inv[0][c];
      g.dsh[c][2] = inv[1][c];
      g.dsh[c][3] = inv[2][c];
    }
    GMC_UNROLL for (int a = 0; a < 3; a++)
      GMC_UNROLL for (int b = 0; b < 3; b++)
        g.G[a][b] = inv[0][a] * inv[0][b] + inv[1][a] * inv[1][b] + inv[2][a] * inv[2][b];
    return J > 0.0;
  }
}

// interior quadrature (element.hpp:133-136, 446-450): shape values at qp q
template <int DIM>
GMC_HD void qp_shape(int q, double* sh) {
  if constexpr (DIM == 2) {
    const double a = 1.0 / 6.0, b = 4.0 / 6.0;
    const double xi = (q == 1) ? b : a, eta = (q == 2) ? b : a;
    sh[0] = 1.0 - xi - eta;
    sh[1] = xi;
    sh[2] = eta;
  } else {
    const double a = 0.1381966011250105, b = 0.5854101966249685;
    const double xi = (q == 1) ? b : a, eta = (q == 2) ? b : a, zeta = (q == 3) ? b : a;
    sh[0] = 1.0 - xi - eta - zeta;
    sh[1] = xi;
    sh[2] = eta;
    sh[3] = zeta;
  }
}
template <int DIM>
GMC_HD double w_in() { return DIM == 2 ? 1.0 / 6.0 : 1.0 / 24.0; }
template <int DIM>
GMC_HD double w_bd() { return DIM == 2 ? 1.0 / 2.0 : 1.0 / 6.0; }

// side quadrature (element.hpp:145-164, 455-466): reference coords of point q
// of local side s, and the area vector JN of that side (quad.hpp:271-291,
// 633-664).  P1: JN is constant over a side.
template <int DIM>
GMC_HD void side_point(int s, int q, double* xi) {
  if constexpr (DIM == 2) {
    const double gm = 0.5 * (1.0 - 1.0 / sqrt(3.0));
    const double gp = 0.5 * (1.0 + 1.0 / sqrt(3.0));
    const double t = (q == 0) ? gp : gm;
    if (s == 0) { xi[0] = t; xi[1] = 1.0 - t; }
    else if (s == 1) { xi[0] = 0.0; xi[1] = t; }
    else { xi[0] = (q == 0) ? gm : gp; xi[1] = 0.0; }
  } else {
    const double a = 1.0 / 6.0, b = 4.0 / 6.0;
    // pattern codes: 0 -> 0, 1 -> 1/6, 2 -> 4/6 ; packed 3 bits per coordinate
    const int pat[4][3] = {{0x2 | 0x1 << 2 | 0x1 << 4, 0x1 | 0x2 << 2 | 0x1 << 4, 0x1 | 0x1 << 2 | 0x2 << 4},
                           {0x0 | 0x1 << 2 | 0x1 << 4, 0x0 | 0x1 << 2 | 0x2 << 4, 0x0 | 0x2 << 2 | 0x1 << 4},
                           {0x1 | 0x0 << 2 | 0x1 << 4, 0x2 | 0x0 << 2 | 0x1 << 4, 0x1 | 0x0 << 2 | 0x2 << 4},
                           {0x1 | 0x1 << 2 | 0x0 << 4, 0x1 | 0x2 << 2 | 0x0 << 4, 0x2 | 0x1 << 2 | 0x0 << 4}};
    const int code = pat[s][q];
    for (int k = 0; k < 3; k++) {
      const int c = (code >> (2 * k)) & 3;
      xi[k] = (c == 0) ? 0.0 : ((c == 1) ? a : b);
    }
  }
}

template <int DIM>
GMC_HD void side_JN(const double* xe, const double* xi, double* JN) {
  GMC_NO_CONTRACT
  if constexpr (DIM == 2) {
    const double dx_dxi = -xe[0] + xe[2], dx_deta = -xe[0] + xe[4];
    const double dy_dxi = -xe[1] + xe[3], dy_deta = -xe[1] + xe[5];
    if (xi[0] == 0.0) { JN[0] = -dy_deta; JN[1] = dx_deta; }
    else if (xi[1] == 0.0) { JN[0] = dy_dxi; JN[1] = -dx_dxi; }
    else { JN[0] = dy_deta - dy_dxi; JN[1] = dx_dxi - dx_deta; }
  } else {
    double d[3][3];
    for (int c = 0; c < 3; c++) {
      d[c][0] = -xe[c] + xe[3 + c];
      d[c][1] = -xe[c] + xe[6 + c];
      d[c][2] = -xe[c] + xe[9 + c];
    }
    const double dx_dxi = d[0][0], dx_deta = d[0][1], dx_dzeta = d[0][2];
    const double dy_dxi = d[1][0], dy_deta = d[1][1], dy_dzeta = d[1][2];
    const double dz_dxi = d[2][0], dz_deta = d[2][1], dz_dzeta = d[2][2];
    if (xi[0] == 0.0) {
      JN[0] = dz_deta * dy_dzeta - dy_deta * dz_dzeta;
      JN[1] = dx_deta * dz_dzeta - dz_deta * dx_dzeta;
      JN[2] = dy_deta * dx_dzeta - dx_deta * dy_dzeta;
    } else if (xi[1] == 0.0) {
      JN[0] = dz_dzeta * dy_dxi - dy_dzeta * dz_dxi;
      JN[1] = dz_dxi * dx_dzeta - dx_dxi * dz_dzeta;
      JN[2] = dy_dzeta * dx_dxi - dx_dzeta * dy_dxi;
    } else if (xi[2] == 0.0) {
      JN[0] = dz_dxi * dy_deta - dy_dxi * dz_deta;
      JN[1] = dx_dxi * dz_deta - dx_deta * dz_dxi;
      JN[2] = dy_dxi * dx_deta - dx_dxi * dy_deta;
    } else {
      JN[0] = (dz_dzeta - dz_dxi) * (dy_deta - dy_dxi) - (dy_dzeta - dy_dxi) * (dz_deta - dz_dxi);
      JN[1] = (dx_dzeta - dx_dxi) * (dz_deta - dz_dxi) - (dx_deta - dx_dxi) * (dz_dzeta - dz_dxi);
      JN[2] = (dx_deta - dx_dxi) * (dy_dzeta - dy_dxi) - (dx_dzeta - dx_dxi) * (dy_deta - dy_dxi);
    }
  }
}

template <int DIM>
GMC_HD void side_shape(const double* xi, double* sh) {
  if constexpr (DIM == 2) {
    sh[0] = 1.0 - xi[0] - xi[1];
    sh[1] = xi[0];
    sh[2] = xi[1];
  } else {
    sh[0] = 1.0 - xi[0] - xi[1] - xi[2];
    sh[1] = xi[0];
    sh[2] = xi[1];
    sh[3] = xi[2];
  }
}

// ------------------------------------------------------------------------
// per-quadrature-point state and coefficient matrices
// ------------------------------------------------------------------------
template <typename T, int DIM, int NC>
struct QP {
  static constexpr int N = Shape<DIM, NC>::NDOF;
  T p, v[DIM], Tt, Y[NC];
  Props<T, NC> pr;
  T dI[DIM][N];               // gradients of the I state
  T U[N];
  T DUDY[N][N];
  T A[DIM][N][N];
  T S[N];                     // source; S0 only has (E, v_k) = rho g_k
  T Fadv[DIM][N], Fdif[DIM][N];
  // K_kl in closed form: lambda, mu, kappa, rho D_i, rho D_i H_i, v
  T lam, mu, kappa, rhoD[NC], rhoDH[NC];
  T E;
  // DUDY = c d0^T + rho M with c = [1, v, E, Y_i] (MatPoint): d0 = d rho/dY
  // and mE = dE/dY over the primitives, the velocity entries of d0 being 0
  T d0[N], mE[N];
};

// fluid::assign (fluid_2d.hpp:109-119)
template <typename T, int DIM, int NC>
GMC_HD void assign(const T* v, T& p, T* vel, T& Tt, T* Y) {
  p = v[0];
  GMC_UNROLL for (int k = 0; k < DIM; k++) vel[k] = v[1 + k];
  Tt = v[DIM + 1];
  T s = T(0.0);
  GMC_UNROLL for (int i = 0; i < NC - 1; i++) {
    Y[i] = v[DIM + 2 + i];
    s += Y[i];
  }
  Y[NC - 1] = T(1.0) - s;
}

// fluid::U (fluid_2d.hpp:372-385)
template <typename T, int DIM, int NC>
GMC_HD void U_vec(T rho, T cv, const T* vel, T Tt, const T* Y, T* u) {
  T vv = T(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) vv += vel[k] * vel[k];
  const T E = cv * Tt + T(0.5) * vv;
  u[0] = rho;
  GMC_UNROLL for (int k = 0; k < DIM; k++) u[1 + k] = rho * vel[k];
  u[DIM + 1] = rho * E;
  GMC_UNROLL for (int i = 0; i < NC - 1; i++) u[DIM + 2 + i] = rho * Y[i];
}

// u_ref_inv (fluid_2d.hpp:123-135): 1/U(p_ref, v_ref, T_ref, Y_ref) entry by
// entry (0 where U < eps), Y_ref = 0.5 for every component (fluid_2d.hpp:126).
template <typename T, int DIM, int NC, int MG = MG_CALL>
GMC_HD int uref_inv_of(const DevParams& P, const Material<T>& M, T* out) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  T Y[NC], vr[DIM], u[N];
  GMC_UNROLL for (int c = 0; c < NC; c++) Y[c] = T(0.5);
  GMC_UNROLL for (int k = 0; k < DIM; k++) vr[k] = T(P.v_ref[k]);
  Props<T, NC> pr;
  properties<T, NC, MG>(P, M, T(P.p_ref), T(P.T_ref), Y, pr);
  U_vec<T, DIM, NC>(pr.rho, pr.cv, vr, T(P.T_ref), Y, u);
  GMC_UNROLL for (int i = 0; i < 8; i++) out[i] = T(0.0);
  GMC_UNROLL for (int i = 0; i < N; i++) out[i] = (u[i] < T(kEps)) ? T(0.0) : T(1.0) / u[i];
  return pr.st;
}

// compute_shear_rate (non_newtonian.hpp:23-38): the strain rate of the
// property evaluation; only the Pal (2003) and Caricchi (2007) magma
// viscosity models depend on it.
template <typename T, int DIM, int N>
GMC_HD T shear_rate(const T (*dI)[N]) {
  using std::sqrt;
  if constexpr (DIM == 2) {
    const T a = dI[0][1], b = dI[1][2], c = dI[1][1] + dI[0][2];
    return sqrt(a * a + b * b + T(2) * (c * c));
  } else {
    const T a = dI[0][1], b = dI[1][2], c = dI[2][3];
    const T d = dI[1][1] + dI[0][2], e = dI[2][1] + dI[0][3], f = dI[2][2] + dI[1][3];
    return sqrt(a * a + b * b + c * c + T(2) * (d * d) + T(2) * (e * e) + T(2) * (f * f));
  }
}

// K_kl(i, j) from closed form (fluid_2d.hpp:531-571), compile-time indices.
// Q is any point record with lam, mu, kappa, rhoD, rhoDH and v.
template <typename T, int DIM, int NC, typename Q>
GMC_HD T Kentry(const Q& q, int k, int l, int i, int j) {
  constexpr int IE = DIM + 1, IY = DIM + 2;
  if (k == l) {
    if (i >= 1 && i <= DIM && j == i) return (i - 1 == k) ? q.lam + T(2) * q.mu : q.mu;
    if (i == IE && j >= 1 && j <= DIM)
      return (j - 1 == k) ? (q.lam + T(2) * q.mu) * q.v[j - 1] : q.mu * q.v[j - 1];
    if (i == IE && j == IE) return q.kappa;
    if (i >= IY && j == i) return q.rhoD[i - IY];
    if (i == IE && j >= IY) return q.rhoDH[j - IY];
    return T(0);
  } else {
    if (i == 1 + k && j == 1 + l) return q.lam;
    if (i == 1 + l && j == 1 + k) return q.mu;
    if (i == IE && j == 1 + k) return q.mu * q.v[l];
    if (i == IE && j == 1 + l) return q.lam * q.v[k];
    return T(0);
  }
}

// flux_matrices_and_vectors, interior (fluid_2d.hpp:391-603)
// WITH_A = false skips the dense advective Jacobians A_k: interior points use
// them only through products, which adv_products forms in closed form.
template <typename T, int DIM, int NC, bool WITH_A = true>
GMC_HD void flux_interior(const DevParams& P, QP<T, DIM, NC>& q) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  constexpr int IT = DIM + 1, IE = DIM + 1, IY = DIM + 2;
  const auto& pr = q.pr;
  const T rho = pr.rho, p = q.p, Tt = q.Tt;
  T vv = T(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) vv += q.v[k] * q.v[k];
  const T E = pr.cv * Tt + T(0.5) * vv;
  q.E = E;
  const T drho_dp = rho * pr.beta;
  const T drho_dT = -rho * pr.alpha;
  const T de_dp = (pr.beta * p - pr.alpha * Tt) / rho;
  const T de_dT = pr.cp - p * pr.alpha / rho;
  T drho_dY[NC], dE_dY[NC];
  GMC_UNROLL for (int i = 0; i < NC - 1; i++) {
    drho_dY[i] = -rho * rho * (T(1.) / pr.rho_c[i] - T(1) / pr.rho_c[NC - 1]);
    dE_dY[i] = pr.e_c[i] - pr.e_c[NC - 1];
  }
  GMC_UNROLL for (int i = 0; i < N; i++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      q.DUDY[i][j] = T(0);
      if (WITH_A) GMC_UNROLL for (int k = 0; k < DIM; k++) q.A[k][i][j] = T(0);
    }
  auto& D = q.DUDY;
  D[0][0] = drho_dp;
  D[0][IT] = drho_dT;
  GMC_UNROLL for (int c = 0; c < NC - 1; c++) D[0][IY + c] = drho_dY[c];
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    D[1 + k][0] = drho_dp * q.v[k];
    D[1 + k][1 + k] = rho;
    D[1 + k][IT] = drho_dT * q.v[k];
    GMC_UNROLL for (int c = 0; c < NC - 1; c++) D[1 + k][IY + c] = q.v[k] * drho_dY[c];
  }
  D[IE][0] = drho_dp * E + rho * de_dp;
  GMC_UNROLL for (int k = 0; k < DIM; k++) D[IE][1 + k] = rho * q.v[k];
  D[IE][IT] = drho_dT * E + rho * de_dT;
  GMC_UNROLL for (int c = 0; c < NC - 1; c++) D[IE][IY + c] = E * drho_dY[c] + rho * dE_dY[c];
  GMC_UNROLL for (int i = 0; i < NC - 1; i++) {
    D[IY + i][0] = drho_dp * q.Y[i];
    D[IY + i][IT] = drho_dT * q.Y[i];
    GMC_UNROLL for (int c = 0; c < NC - 1; c++) D[IY + i][IY + c] = q.Y[i] * drho_dY[c];
  }
  GMC_UNROLL for (int i = 0; i < NC - 1; i++) D[IY + i][IY + i] += rho;
  q.d0[0] = drho_dp;
  q.d0[IT] = drho_dT;
  q.mE[0] = de_dp;
  q.mE[IT] = de_dT;
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    q.d0[1 + k] = T(0);
    q.mE[1 + k] = q.v[k];
  }
  GMC_UNROLL for (int c = 0; c < NC - 1; c++) {
    q.d0[IY + c] = drho_dY[c];
    q.mE[IY + c] = dE_dY[c];
  }

  if (WITH_A) GMC_UNROLL for (int k = 0; k < DIM; k++) {
    auto& A = q.A[k];
    const T vk = q.v[k];
    A[0][0] = drho_dp * vk;
    A[0][1 + k] = rho;
    A[0][IT] = drho_dT * vk;
    GMC_UNROLL for (int c = 0; c < NC - 1; c++) A[0][IY + c] = drho_dY[c] * vk;
    GMC_UNROLL for (int j = 0; j < DIM; j++) {
      const int r = 1 + j;
      if (j == k) {
        A[r][0] = drho_dp * vk * vk + T(1.0);
        A[r][1 + k] = T(2.0) * rho * vk;
        A[r][IT] = drho_dT * vk * vk;
        GMC_UNROLL for (int c = 0; c < NC - 1; c++) A[r][IY + c] = drho_dY[c] * vk * vk;
      } else {
        A[r][0] = drho_dp * q.v[j] * vk;
        A[r][1 + k] = rho * q.v[j];
        A[r][1 + j] = rho * vk;
        A[r][IT] = drho_dT * q.v[j] * vk;
        GMC_UNROLL for (int c = 0; c < NC - 1; c++) A[r][IY + c] = drho_dY[c] * q.v[j] * vk;
      }
    }
    A[IE][0] = vk + vk * (drho_dp * E + rho * de_dp);
    GMC_UNROLL for (int j = 0; j < DIM; j++) {
      if (j == k) A[IE][1 + j] = p + rho * E + rho * vk * vk;
      else A[IE][1 + j] = rho * q.v[j < k ? j : k] * q.v[j < k ? k : j];
    }
    A[IE][IT] = vk * (drho_dT * E + rho * de_dT);
    GMC_UNROLL for (int c = 0; c < NC - 1; c++) A[IE][IY + c] = (E * drho_dY[c] + rho * dE_dY[c]) * vk;
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) {
      A[IY + i][0] = drho_dp * q.Y[i] * vk;
      A[IY + i][1 + k] = rho * q.Y[i];
      A[IY + i][IT] = drho_dT * q.Y[i] * vk;
      GMC_UNROLL for (int c = 0; c < NC - 1; c++) A[IY + i][IY + c] = q.Y[i] * drho_dY[c] * vk;
    }
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) A[IY + i][IY + i] += rho * vk;
  }

  // K closed-form scalars
  q.mu = pr.mu;
  q.lam = T(-2.0) * pr.mu / T(3.0);
  q.kappa = pr.kappa;
  GMC_UNROLL for (int i = 0; i < NC - 1; i++) {
    const T H = pr.e_c[i] + p / pr.rho_c[i];
    q.rhoD[i] = rho * pr.D_c[i];
    q.rhoDH[i] = rho * pr.D_c[i] * H;
  }

  U_vec<T, DIM, NC>(rho, pr.cv, q.v, Tt, q.Y, q.U);
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    q.Fadv[k][0] = rho * q.v[k];
    GMC_UNROLL for (int j = 0; j < DIM; j++) {
      const T f = rho * q.v[k] * q.v[j];
      q.Fadv[k][1 + j] = (j == k) ? f + p : f;
    }
    q.Fadv[k][IE] = q.v[k] * (rho * E + p);
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) q.Fadv[k][IY + i] = rho * q.v[k] * q.Y[i];
  }
  GMC_UNROLL for (int i = 0; i < N; i++) q.S[i] = T(0);
  T gv = T(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    q.S[1 + k] = rho * T(P.g[k]);
    gv += T(P.g[k]) * q.v[k];
  }
  q.S[IE] = rho * gv;
  // F_dif_k = sum_l K_kl dI_l
  GMC_UNROLL for (int k = 0; k < DIM; k++)
    GMC_UNROLL for (int i = 0; i < N; i++) {
      T acc = T(0);
      GMC_UNROLL for (int l = 0; l < DIM; l++) {
        T s = T(0.0);
        GMC_UNROLL for (int j = 0; j < N; j++) s += Kentry<T, DIM, NC>(q, k, l, i, j) * q.dI[l][j];
        acc = (l == 0) ? s : acc + s;
      }
      q.Fdif[k][i] = acc;
    }
}

// Products with the advective Jacobians without forming them:
//   DdI[k] = DUDY dI_k,   AdI = sum_k A_k dI_k,
// using (A_k x)_i = v_k (DUDY x)_i + U_i x_{v_k} + [i = v_k] x_p
//                   + [i = E] (v_k x_p + p x_{v_k}),
// which is entry by entry the A_k of fluid_2d.hpp:445-524 (F_k = v_k U +
// p (e_{v_k} + v_k e_E), differentiated with respect to the primitives).
template <typename T, int DIM, int NC>
GMC_HD void adv_products(const QP<T, DIM, NC>& q, T (*DdI)[Shape<DIM, NC>::NDOF], T* AdI) {
  constexpr int N = Shape<DIM, NC>::NDOF, IE = DIM + 1;
  GMC_UNROLL for (int k = 0; k < DIM; k++)
    GMC_UNROLL for (int i = 0; i < N; i++) {
      T s = T(0.0);
      GMC_UNROLL for (int j = 0; j < N; j++) s += q.DUDY[i][j] * q.dI[k][j];
      DdI[k][i] = s;
    }
  T divv = T(0.0), vdp = T(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    divv += q.dI[k][1 + k];
    vdp += q.v[k] * q.dI[k][0];
  }
  GMC_UNROLL for (int i = 0; i < N; i++) {
    T s = T(0.0);
    GMC_UNROLL for (int k = 0; k < DIM; k++) s += q.v[k] * DdI[k][i];
    s += q.U[i] * divv;
    if (i >= 1 && i <= DIM) s += q.dI[i - 1][0];
    if (i == IE) s += vdp + q.p * divv;
    AdI[i] = s;
  }
}

// Inverse by LU with partial pivoting in uBLAS order (algorithm.hpp:399-414),
// for the block-Jacobi preconditioner's diagonal blocks.  Row exchanges are
// predicated selects so every index stays compile-time.  Returns false on a
// zero pivot.
template <typename T, int N>
GMC_HD bool lu_inverse(const T (*Ain)[N], T (*Inv)[N]) {
  using std::fabs;
  T A[N][N];
  GMC_UNROLL for (int i = 0; i < N; i++) GMC_UNROLL for (int j = 0; j < N; j++) A[i][j] = Ain[i][j];
  int pm[N];
  bool ok = true;
  GMC_UNROLL for (int i = 0; i < N; i++) {
    int imax = i;
    T vmax = fabs(A[i][i]);
    GMC_UNROLL for (int r = i + 1; r < N; r++) {
      const T a = fabs(A[r][i]);
      if (a > vmax) {
        vmax = a;
        imax = r;
      }
    }
    pm[i] = imax;
    GMC_UNROLL for (int r = i + 1; r < N; r++) {
      const bool sw = (imax == r);
      GMC_UNROLL for (int c = 0; c < N; c++) {
        const T t = A[i][c];
        A[i][c] = sw ? A[r][c] : t;
        A[r][c] = sw ? t : A[r][c];
      }
    }
    if (A[i][i] == T(0)) ok = false;
    const T rp = T(1.0) / A[i][i];
    GMC_UNROLL for (int r = i + 1; r < N; r++) A[r][i] *= rp;
    GMC_UNROLL for (int r = i + 1; r < N; r++)
      GMC_UNROLL for (int c = i + 1; c < N; c++) A[r][c] -= A[r][i] * A[i][c];
  }
  GMC_UNROLL for (int r = 0; r < N; r++) GMC_UNROLL for (int c = 0; c < N; c++) Inv[r][c] = T(r == c ? 1.0 : 0.0);
  GMC_UNROLL for (int i = 0; i < N; i++)
    GMC_UNROLL for (int r = i + 1; r < N; r++) {
      const bool sw = (pm[i] == r);
      GMC_UNROLL for (int c = 0; c < N; c++) {
        const T t = Inv[i][c];
        Inv[i][c] = sw ? Inv[r][c] : t;
        Inv[r][c] = sw ? t : Inv[r][c];
      }
    }
  GMC_UNROLL for (int k = 0; k < N; k++)
    GMC_UNROLL for (int l = 0; l < N; l++) {
      const T t = Inv[k][l];
      GMC_UNROLL for (int m = k + 1; m < N; m++) Inv[m][l] -= A[m][k] * t;
    }
  GMC_UNROLL for (int k = N - 1; k >= 0; k--) {
    const T rk = T(1.0) / A[k][k];
    GMC_UNROLL for (int l = N - 1; l >= 0; l--) {
      const T t = Inv[k][l] * rk;
      Inv[k][l] = t;
      GMC_UNROLL for (int m = k - 1; m >= 0; m--) Inv[m][l] -= A[m][k] * t;
    }
  }
  return ok;
}

// tau = DUDY^-1 diag(d) in closed form.  DUDY = c d0^T + rho M (QP::d0,
// QP::mE): its mass, momentum and species rows give rho-change, velocity and
// mass-fraction unknowns directly, and the energy row with the mass row leave
// a 2 x 2 system in (p, T):
//   x_v = (b_v - v b_0)/rho,  x_Y = (b_Y - Y b_0)/rho,
//   [d0_p d0_T; mE_p mE_T] [x_p; x_T] = [b_0 - d0_Y.x_Y; (b_E - E b_0)/rho - v.x_v - mE_Y.x_Y],
// for b = d_j e_j, column by column.  Returns false if DUDY is singular
// (rho = 0 or the 2 x 2 determinant is 0), where a pivoted LU meets a zero pivot.
template <typename T, int DIM, int NC>
GMC_HD bool dudy_inverse_cols(const QP<T, DIM, NC>& q, const T* d, T (*out)[Shape<DIM, NC>::NDOF]) {
  constexpr int N = Shape<DIM, NC>::NDOF, IT = DIM + 1, IE = DIM + 1, IY = DIM + 2;
  const T rho = q.pr.rho;
  const T det = q.d0[0] * q.mE[IT] - q.d0[IT] * q.mE[0];
  const bool ok = rho != T(0) && det != T(0);
  const T irho = T(1.0) / rho, idet = T(1.0) / det;
  GMC_UNROLL for (int j = 0; j < N; j++) {
    // b = d_j e_j (compile-time zeros elsewhere)
    T b[N];
    GMC_UNROLL for (int i = 0; i < N; i++) b[i] = (i == j) ? d[j] : T(0);
    const T b0 = b[0];
    T x[N];
    T r1 = b0, r2 = (b[IE] - q.E * b0) * irho;
    GMC_UNROLL for (int k = 0; k < DIM; k++) {
      x[1 + k] = (b[1 + k] - q.v[k] * b0) * irho;
      r2 -= q.v[k] * x[1 + k];
    }
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) {
      x[IY + i] = (b[IY + i] - q.Y[i] * b0) * irho;
      r1 -= q.d0[IY + i] * x[IY + i];
      r2 -= q.mE[IY + i] * x[IY + i];
    }
    x[0] = (r1 * q.mE[IT] - q.d0[IT] * r2) * idet;
    x[IT] = (q.d0[0] * r2 - q.mE[0] * r1) * idet;
    GMC_UNROLL for (int i = 0; i < N; i++) out[i][j] = x[i];
  }
  return ok;
}

// LU factorisation with partial pivoting in uBLAS order (lu_factorize of
// algorithm.hpp:399-414): A is overwritten by L\U, pm receives the pivots, the
// reciprocal of every pivot is returned in rp.  Returns false on a zero pivot.
template <typename T, int N>
GMC_HD bool lu_factor(T (*A)[N], int* pm, T* rp) {
  using std::fabs;
  bool ok = true;
  GMC_UNROLL for (int i = 0; i < N; i++) {
    int imax = i;
    T vmax = fabs(A[i][i]);
    GMC_UNROLL for (int r = i + 1; r < N; r++) {
      const T a = fabs(A[r][i]);
      if (a > vmax) {
        vmax = a;
        imax = r;
      }
    }
    pm[i] = imax;
    GMC_UNROLL for (int r = i + 1; r < N; r++) {
      const bool sw = (imax == r);
      GMC_UNROLL for (int c = 0; c < N; c++) {
        const T t = A[i][c];
        A[i][c] = sw ? A[r][c] : t;
        A[r][c] = sw ? t : A[r][c];
      }
    }
    if (A[i][i] == T(0)) ok = false;
    rp[i] = T(1.0) / A[i][i];
    GMC_UNROLL for (int r = i + 1; r < N; r++) A[r][i] *= rp[i];
    GMC_UNROLL for (int r = i + 1; r < N; r++)
      GMC_UNROLL for (int c = i + 1; c < N; c++) A[r][c] -= A[r][i] * A[i][c];
  }
  return ok;
}

// out = A^-1 diag(d) from the factors of lu_factor, one column at a time: the
// lu_substitute steps of each column of the identity (row exchanges, unit
// lower then upper substitution, algorithm.hpp:399-414) are independent of the
// other columns, so this equals the whole-matrix substitution bit for bit
// while only one column is live.
template <typename T, int N>
GMC_HD void lu_inverse_cols(const T (*A)[N], const int* pm, const T* rp, const T* d, T (*out)[N]) {
  GMC_UNROLL for (int l = 0; l < N; l++) {
    T x[N];
    GMC_UNROLL for (int r = 0; r < N; r++) x[r] = T(r == l ? 1.0 : 0.0);
    GMC_UNROLL for (int i = 0; i < N; i++)
      GMC_UNROLL for (int r = i + 1; r < N; r++) {
        const bool sw = (pm[i] == r);
        const T t = x[i];
        x[i] = sw ? x[r] : t;
        x[r] = sw ? t : x[r];
      }
    GMC_UNROLL for (int k = 0; k < N; k++) {
      const T t = x[k];
      GMC_UNROLL for (int m = k + 1; m < N; m++) x[m] -= A[m][k] * t;
    }
    GMC_UNROLL for (int k = N - 1; k >= 0; k--) {
      const T t = x[k] * rp[k];
      x[k] = t;
      GMC_UNROLL for (int m = k - 1; m >= 0; m--) x[m] -= A[m][k] * t;
    }
    GMC_UNROLL for (int r = 0; r < N; r++) out[r][l] = x[r] * d[l];
  }
}

// quad::compute_h (quad.hpp:356-383, 701-739)
template <typename T, int DIM>
GMC_HD T compute_h(const double (*G)[DIM], const T* v, const T (*dv)[DIM]) {
  GMC_NO_CONTRACT
  using std::sqrt;
  using std::fabs;
  T nrm_v = T(0.0);
  GMC_UNROLL for (int i = 0; i < DIM; i++) nrm_v += v[i] * v[i];
  nrm_v = sqrt(nrm_v);
  T s0[DIM], s1[DIM], acc[DIM];
  GMC_UNROLL for (int i = 0; i < DIM; i++) s0[i] = v[i] / nrm_v;
  T al0 = T(0.0), al1 = T(0.0), al2 = T(0.0);
  GMC_UNROLL for (int i = 0; i < DIM; i++) {
    T gs = T(0.0);
    GMC_UNROLL for (int j = 0; j < DIM; j++) gs += T(G[i][j]) * s0[j];
    al0 += s0[i] * gs;
  }
  GMC_UNROLL for (int i = 0; i < DIM; i++) {
    acc[i] = T(0.0);
    GMC_UNROLL for (int j = 0; j < DIM; j++) acc[i] += dv[i][j] * v[j];
  }
  T as0 = T(0.0);
  GMC_UNROLL for (int i = 0; i < DIM; i++) as0 += acc[i] * s0[i];
  GMC_UNROLL for (int i = 0; i < DIM; i++) s1[i] = acc[i] - as0 * s0[i];
  T n1 = T(0.0);
  GMC_UNROLL for (int i = 0; i < DIM; i++) n1 += s1[i] * s1[i];
  n1 = sqrt(n1);
  if (n1 > T(kEps)) {
    GMC_UNROLL for (int i = 0; i < DIM; i++) s1[i] /= n1;
  }
  GMC_UNROLL for (int i = 0; i < DIM; i++) {
    T gs = T(0.0);
    GMC_UNROLL for (int j = 0; j < DIM; j++) gs += T(G[i][j]) * s1[j];
    al1 += s1[i] * gs;
  }
  if constexpr (DIM == 3) {
    T s2[3];
    s2[0] = s0[1] * s1[2] - s0[2] * s1[1];
    s2[1] = s1[0] * s0[2] - s1[2] * s0[0];
    s2[2] = s0[0] * s1[1] - s0[1] * s1[0];
    const T n2 = sqrt(s2[0] * s2[0] + s2[1] * s2[1] + s2[2] * s2[2]);
    if (n2 > T(kEps)) {
      GMC_UNROLL for (int i = 0; i < 3; i++) s2[i] /= n2;
    }
    GMC_UNROLL for (int i = 0; i < 3; i++) {
      T gs = T(0.0);
      GMC_UNROLL for (int j = 0; j < 3; j++) gs += T(G[i][j]) * s2[j];
      al2 += s2[i] * gs;
    }
  }
  T h = T(0.0);
  if (al0 != T(0.0)) h += T(1.0) / al0;
  if (al1 != T(0.0)) h += T(1.0) / al1;
  if (DIM == 3 && al2 != T(0.0)) h += T(1.0) / al2;
  return T(2) * sqrt(fabs(h));
}

// tau_matrix (fluid_2d.hpp:877-1022): 0 when |v| < eps, else the form the
// parameters select.  The non-diagonal forms (2019 :932-960, 2001 :895-929)
// are DUDY^-1 diag(t) with the low-Mach correction of row 0; 2007 (:963-991)
// and 2014 (:994-1021) are diagonal.  Returns false when the LU of DUDY fails.
// The gradients enter only through the element length h = compute_h(v, dv);
// kH takes h from h_in instead (the adjoint differentiates tau along h).
template <typename T, int DIM, int NC, bool kLU = false, bool kH = false>
GMC_HD bool tau_matrix(const DevParams& P, const QP<T, DIM, NC>& q, const Geo<DIM>& g,
                       T (*tau)[Shape<DIM, NC>::NDOF], T h_in = T(0)) {
  using std::sqrt;
  constexpr int N = Shape<DIM, NC>::NDOF;
  constexpr int IE = DIM + 1, IY = DIM + 2;
  T vv = T(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) vv += q.v[k] * q.v[k];
  const T v_nrm = sqrt(vv);
  GMC_UNROLL for (int i = 0; i < N; i++) GMC_UNROLL for (int j = 0; j < N; j++) tau[i][j] = T(0);
  if (v_nrm < T(kEps)) return true;
  const T c = q.pr.c;
  T h;
  if constexpr (kH) {
    h = h_in;
  } else {
    T dv[DIM][DIM];
    GMC_UNROLL for (int i = 0; i < DIM; i++) GMC_UNROLL for (int j = 0; j < DIM; j++) dv[i][j] = q.dI[j][1 + i];
    h = compute_h<T, DIM>(g.G, q.v, dv);
  }
  const T rho = q.pr.rho;
  const double dt = P.dt;
  double trG = 0.0;
  GMC_UNROLL for (int k = 0; k < DIM; k++) trG += g.G[k][k];
  if (P.tau_variant == 2007) {
    // diagonal, incompressible form; g_k = sum_l dxi_l/dx_k (quad.hpp:328-329)
    double use = 0.0, gg = 0.0;
    GMC_UNROLL for (int i = 0; i < DIM; i++) GMC_UNROLL for (int j = 0; j < DIM; j++) use += g.G[i][j] * g.G[i][j];
    GMC_UNROLL for (int k = 0; k < DIM; k++) {
      double gk = g.dsh[k][1];
      GMC_UNROLL for (int a = 2; a <= DIM; a++) gk += g.dsh[k][a];
      gg += gk * gk;
    }
    T vGv = T(0.0);
    GMC_UNROLL for (int i = 0; i < DIM; i++) {
      T gs = T(0.0);
      GMC_UNROLL for (int j = 0; j < DIM; j++) gs += T(g.G[i][j]) * q.v[j];
      vGv += q.v[i] * gs;
    }
    const T nm = q.pr.mu / rho, ne = q.pr.kappa / (rho * q.pr.cv);
    T tm = vGv + T(9.0 * use) * (nm * nm);
    T te = vGv + T(9.0 * use) * (ne * ne);
    if (!P.steady) {
      tm += T(4.0 / (dt * dt));
      te += T(4.0 / (dt * dt));
    }
    tm = T(1.0) / sqrt(tm);
    te = T(1.0) / sqrt(te);
    tau[0][0] = T(1.0) / (rho * tm * T(gg));
    GMC_UNROLL for (int k = 0; k < DIM; k++) tau[1 + k][1 + k] = tm / rho;
    tau[IE][IE] = te / (rho * q.pr.cv);
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) tau[IY + i][IY + i] = tm;
    return true;
  }
  if (P.tau_variant == 2014) {
    // diagonals of DUDY, A_k and K_kk (K_kl has a zero diagonal for k != l);
    // A_k(i, i) = v_k DUDY(i, i) + [i = v_k] U_i in closed form
    T td[N];
    GMC_UNROLL for (int i = 0; i < N; i++) {
      const T a0 = q.DUDY[i][i];
      const T t2 = T(4.0 / (dt * dt)) * (a0 * a0);
      T ak[DIM];
      GMC_UNROLL for (int k = 0; k < DIM; k++) ak[k] = q.v[k] * q.DUDY[i][i] + ((i == 1 + k) ? q.U[i] : T(0));
      T a2 = T(0.0), d2 = T(0.0);
      GMC_UNROLL for (int k = 0; k < DIM; k++) {
        GMC_UNROLL for (int l = 0; l < DIM; l++) a2 += T(g.G[k][l]) * (ak[k] * ak[l]);
        const T kd = Kentry<T, DIM, NC>(q, k, k, i, i);
        d2 += T(g.G[k][k] * g.G[k][k]) * (kd * kd);
      }
      td[i] = T(1.0) / (sqrt(t2) + sqrt(a2) + sqrt(d2) + T(1.e-8));
    }
    td[0] = T(1.0) / (td[0] + rho * td[1] * T(trG));
    GMC_UNROLL for (int i = 0; i < N; i++) tau[i][i] = td[i];
    return true;
  }
  T td[N];
  if (P.tau_variant == 2001) {
    // alpha (fluid_2d.hpp:885) is never assigned: lambda^2 = 4/dt^2 unsteady
    // (steady is rejected at create: lambda = 1/sqrt(0))
    const T lambda = T(1.0 / std::sqrt(4.0 / (dt * dt))) + h / (T(2.0) * v_nrm);
    const T use = tmin(T(dt * 0.5), lambda);
    td[0] = use;
    td[1] = tmin(use, rho * h * h / (q.pr.mu * T(12.0)));
    GMC_UNROLL for (int k = 1; k < DIM; k++) td[1 + k] = td[1];
    td[IE] = tmin(use, rho * h * h * q.pr.cv / (q.pr.kappa * T(12.0)));
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) td[IY + i] = tmin(use, h * h / (q.pr.D_c[i] * T(12.0)));
  } else {
    const T adv = h / (T(2.0) * (v_nrm + c)) + h / (T(2.0) * v_nrm);
    const T use = P.steady ? adv : tmin(T(dt * 0.5), adv);
    td[0] = use;
    td[1] = tmin(use, h * h * rho / (T(12.0) * q.pr.mu));
    GMC_UNROLL for (int k = 1; k < DIM; k++) td[1 + k] = td[1];
    td[IE] = tmin(use, h * h * rho * q.pr.cv / (T(12.0) * q.pr.kappa));
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) td[IY + i] = tmin(use, h * h / (T(12.0) * q.pr.D_c[i]));
  }
  // tau = DUDY^-1 diag(td) (the reference's inverse() by LU, algorithm.hpp:
  // 399-414): in closed form from DUDY's rank structure, or (kLU) by the
  // pivoted LU in uBLAS order
  bool ok;
  if constexpr (kLU) {
    T lu[N][N], rp[N];
    int pm[N];
    GMC_UNROLL for (int i = 0; i < N; i++) GMC_UNROLL for (int j = 0; j < N; j++) lu[i][j] = q.DUDY[i][j];
    ok = lu_factor<T, N>(lu, pm, rp);
    lu_inverse_cols<T, N>(lu, pm, rp, td, tau);
  } else {
    ok = dudy_inverse_cols<T, DIM, NC>(q, td, tau);
  }
  if (v_nrm / c < T(0.1)) {
    tau[0][0] = T(1) / (T(1) / tau[0][0] + T(1) / (rho * tau[1][1] * T(trG)));
    GMC_UNROLL for (int j = 1; j < N; j++) tau[0][j] = T(0.0);
  }
  return ok;
}

// dc, dc_2006 (fluid_2d.hpp:1027-1065): returns the scalar
// dc_scale_fact * nu_shoc so that shock_matrix = that * DUDY.
template <typename T, int DIM, int NC>
GMC_HD T dc_nu(const DevParams& P, const QP<T, DIM, NC>& q, const Geo<DIM>& g, const T* uinv,
                const T* AdI, const T (*DdI)[Shape<DIM, NC>::NDOF]) {
  using std::sqrt;
  using std::fabs;
  using std::pow;
  constexpr int N = Shape<DIM, NC>::NDOF;
  if (!P.dc) return T(0);
  T dU0[DIM];
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    T s = T(0.0);
    GMC_UNROLL for (int j = 0; j < N; j++) s += q.DUDY[0][j] * q.dI[k][j];
    dU0[k] = s;
  }
  T nr = T(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) nr += dU0[k] * dU0[k];
  const T del_rho_nrm = sqrt(nr);
  if (del_rho_nrm < T(kEps)) return T(0);
  T Jd[DIM];
  GMC_UNROLL for (int k = 0; k < DIM; k++) Jd[k] = dU0[k] / del_rho_nrm;
  T h = T(0.0);
  GMC_UNROLL for (int a = 0; a < DIM + 1; a++) {
    T s = Jd[0] * T(g.dsh[0][a]);
    GMC_UNROLL for (int k = 1; k < DIM; k++) s += Jd[k] * T(g.dsh[k][a]);
    h += fabs(s);
  }
  if (h > T(kEps)) h = T(1.0) / h;
  else return T(0);
  T zn = T(0.0), un = T(0.0);
  GMC_UNROLL for (int i = 0; i < N; i++) {
    const T z = AdI[i] - q.S[i];
    const T ui = uinv[i];
    zn += (ui * z) * (ui * z);
    un += (ui * q.U[i]) * (ui * q.U[i]);
  }
  zn = sqrt(zn);
  un = sqrt(un);
  T nu = T(0.0);
  if (P.dc_sharp == 2.0) {
    nu = zn * h * h / un;
  } else {
    T use1 = T(0.0);
    GMC_UNROLL for (int k = 0; k < DIM; k++) {
      T s = T(0.0);
      GMC_UNROLL for (int i = 0; i < N; i++) {
        T du = T(0.0);
        GMC_UNROLL for (int j = 0; j < N; j++) du += q.DUDY[i][j] * q.dI[k][j];
        const T t = uinv[i] * du;
        s += t * t;
      }
      const T sn = sqrt(s);
      use1 += sn * sn;
    }
    if (P.dc_sharp == 1.0) nu = zn * h / sqrt(use1);
    else if (P.dc_sharp == 1.5) nu = zn * h * (T(1) / sqrt(use1) + h / un) / T(2.0);
  }
  return T(P.dc_scale) * nu;
}

// A_k(i, j) of an interior point (compile-time k, i, j), without a stored
// matrix: the advective Jacobian of direction k (fluid_2d.hpp:445-524),
//   A_k = v_k DUDY + U e_{v_k}^T + e_{v_k} e_p^T + e_E (v_k e_p + p e_{v_k})^T
template <typename T, int DIM, int NC>
GMC_HD T a_entry(const QP<T, DIM, NC>& q, int k, int i, int j) {
  constexpr int IE = DIM + 1;
  T x = q.v[k] * q.DUDY[i][j];
  if (j == 1 + k) x += q.U[i];
  if (i == 1 + k && j == 0) x += T(1.0);
  if (i == IE && j == 0) x += q.v[k];
  if (i == IE && j == 1 + k) x += q.p;
  return x;
}

// dc_nu and its derivatives with respect to the point gradients,
// dnu[l][j] = d nu / d dI_l,j, in closed form (FP64): through
// dU0_k = DUDY row 0 . dI_k (the density-gradient direction Jd and its length
// h), z = A dI - S (|U_ref^-1 z|) and DUDY dI_k (use1 of dc_sharp 1 and 1.5).
// At a switch the derivative is that of the branch taken, with sqrt' = 0 at 0
// and |x|' = +1 at 0, as the dual numbers take them.  Returns nu.
template <int DIM, int NC>
GMC_HD double dc_nu_grad(const DevParams& P, const QP<double, DIM, NC>& q, const Geo<DIM>& g,
                         const double* uinv, const double* AdI,
                         const double (*DdI)[Shape<DIM, NC>::NDOF],
                         double (*dnu)[Shape<DIM, NC>::NDOF]) {
  using std::sqrt;
  using std::fabs;
  constexpr int N = Shape<DIM, NC>::NDOF;
  GMC_UNROLL for (int l = 0; l < DIM; l++) GMC_UNROLL for (int j = 0; j < N; j++) dnu[l][j] = 0.0;
  if (!P.dc) return 0.0;
  double dU0[DIM];
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    double s = 0.0;
    GMC_UNROLL for (int j = 0; j < N; j++) s += q.DUDY[0][j] * q.dI[k][j];
    dU0[k] = s;
  }
  double nr = 0.0;
  GMC_UNROLL for (int k = 0; k < DIM; k++) nr += dU0[k] * dU0[k];
  const double del = sqrt(nr);
  if (del < kEps) return 0.0;
  double Jd[DIM];
  GMC_UNROLL for (int k = 0; k < DIM; k++) Jd[k] = dU0[k] / del;
  // H = sum_a |Jd . dN_a|; dH/dJd_k = sum_a sign(Jd . dN_a) dN_a,k
  double H = 0.0, sig[DIM];
  GMC_UNROLL for (int k = 0; k < DIM; k++) sig[k] = 0.0;
  GMC_UNROLL for (int a = 0; a < DIM + 1; a++) {
    double s = Jd[0] * g.dsh[0][a];
    GMC_UNROLL for (int k = 1; k < DIM; k++) s += Jd[k] * g.dsh[k][a];
    H += fabs(s);
    const double sg = s < 0.0 ? -1.0 : 1.0;
    GMC_UNROLL for (int k = 0; k < DIM; k++) sig[k] += sg * g.dsh[k][a];
  }
  if (!(H > kEps)) return 0.0;
  const double h = 1.0 / H;
  // dH/d dU0_l = (sig_l - (sig . Jd) Jd_l) / del (the derivative of Jd)
  double sJ = 0.0;
  GMC_UNROLL for (int k = 0; k < DIM; k++) sJ += sig[k] * Jd[k];
  double dHdU[DIM];
  GMC_UNROLL for (int l = 0; l < DIM; l++) dHdU[l] = (sig[l] - sJ * Jd[l]) / del;
  double zn = 0.0, un = 0.0, uz[N];
  GMC_UNROLL for (int i = 0; i < N; i++) {
    const double z = AdI[i] - q.S[i];
    const double ui = uinv[i];
    uz[i] = ui * ui * z;
    zn += (ui * z) * (ui * z);
    un += (ui * q.U[i]) * (ui * q.U[i]);
  }
  zn = sqrt(zn);
  un = sqrt(un);
  // use1 = sum_k |U_ref^-1 DUDY dI_k|^2 (through sqrt and square: no
  // derivative along k where that norm is 0)
  double use1 = 0.0, sn2[DIM];
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    double s = 0.0;
    GMC_UNROLL for (int i = 0; i < N; i++) {
      const double t = uinv[i] * DdI[k][i];
      s += t * t;
    }
    const double sn = sqrt(s);
    sn2[k] = sn;
    use1 += sn * sn;
  }
  const double r = sqrt(use1);
  double nu;
  if (P.dc_sharp == 2.0) nu = zn * h * h / un;
  else if (P.dc_sharp == 1.0) nu = zn * h / r;
  else if (P.dc_sharp == 1.5) nu = zn * h * (1.0 / r + h / un) / 2.0;
  else return 0.0;
  GMC_UNROLL for (int l = 0; l < DIM; l++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      // d zn, d h, d use1 along dI_l,j
      double dz = 0.0;
      if (zn > 0.0) {
        GMC_UNROLL for (int i = 0; i < N; i++) dz += uz[i] * a_entry<double, DIM, NC>(q, l, i, j);
        dz /= zn;
      }
      const double dh = -h * h * dHdU[l] * q.DUDY[0][j];
      double du = 0.0;
      if (sn2[l] > 0.0) {
        GMC_UNROLL for (int i = 0; i < N; i++) du += uinv[i] * uinv[i] * DdI[l][i] * q.DUDY[i][j];
        du *= 2.0;
      }
      const double dr = r > 0.0 ? du / (2.0 * r) : 0.0;
      double dn;
      if (P.dc_sharp == 2.0) {
        dn = (dz * h * h + 2.0 * zn * h * dh) / un;
      } else if (P.dc_sharp == 1.0) {
        dn = ((dz * h + zn * dh) * r - zn * h * dr) / (r * r);
      } else {
        dn = ((dz * h + zn * dh) * (1.0 / r + h / un) + zn * h * (-dr / (r * r) + dh / un)) / 2.0;
      }
      dnu[l][j] = P.dc_scale * dn;
    }
  return P.dc_scale * nu;
}

// Boundary overrides (fluid_2d.hpp:760-867): rebuild F_dif from prescribed
// Neumann data and remove K entries of prescribed fluxes.  Returns the K
// suppression mask used by Kentry_b (bit layout in kmask_*).
struct KMask {
  unsigned row[3][3];   // [k][l]: bit r set -> row r of K_kl removed (velocity rows)
  unsigned eT;          // bit k: K_kk(E,T) removed
  unsigned Jrow[3];     // [k] bit i: K_kk(Y_i,Y_i) and K_kk(E,Y_i) removed
};

template <typename T, int DIM, int NC>
GMC_HD void flux_boundary(const DevParams& P, int flag, QP<T, DIM, NC>& q, KMask& km) {
  constexpr int IE = DIM + 1, IT = DIM + 1, IY = DIM + 2;
  constexpr int N = Shape<DIM, NC>::NDOF;
  // prescribed mass flux m of direction k (fluid_2d.hpp:623-758): F_k^adv is
  // replaced here, A_k by a_entry_bd (the U and DUDY it also rewrites are
  // unused by RM_b)
  if (P.mflux_mask[flag]) {
    const T p = q.p;
    GMC_UNROLL for (int k = 0; k < DIM; k++) {
      if (!(P.mflux_mask[flag] & (1u << k))) continue;
      const T m = T(P.mflux[flag][k]);
      q.Fadv[k][0] = m;
      GMC_UNROLL for (int j = 0; j < DIM; j++) q.Fadv[k][1 + j] = (j == k) ? m * q.v[j] + p : m * q.v[j];
      q.Fadv[k][IE] = m * q.E + p * q.v[k];
      GMC_UNROLL for (int i = 0; i < NC - 1; i++) q.Fadv[k][IY + i] = m * q.Y[i];
    }
  }
  const T lambda = T(-2.0) * q.pr.mu / T(3.0);
  const T mu = q.pr.mu;
  T div = T(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) div += q.dI[k][1 + k];
  T tau[DIM][DIM];
  GMC_UNROLL for (int a = 0; a < DIM; a++)
    GMC_UNROLL for (int b = 0; b < DIM; b++) {
      if (a == b) tau[a][a] = lambda * div + T(2.0) * mu * q.dI[a][1 + a];
      else tau[a][b] = mu * (q.dI[b][1 + a] + q.dI[a][1 + b]);
    }
  T qf[DIM], J[DIM][NC];
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    qf[k] = -q.pr.kappa * q.dI[k][IT];
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) J[k][i] = q.pr.rho * q.pr.D_c[i] * q.dI[k][IY + i];
  }
  const int ta[6] = {0, 1, 2, 0, 0, 1}, tb[6] = {0, 1, 2, 1, 2, 2};
  const unsigned tm = P.n_tau_mask[flag];
  for (int k = 0; k < 3; k++) {
    for (int l = 0; l < 3; l++) km.row[k][l] = 0u;
    km.Jrow[k] = 0u;
  }
  km.eT = 0u;
  GMC_UNROLL for (int t = 0; t < 6; t++) {
    if (ta[t] >= DIM || tb[t] >= DIM) continue;
    if (tm & (1u << t)) {
      tau[ta[t]][tb[t]] = T(P.n_tau[flag][t]);
      tau[tb[t]][ta[t]] = T(P.n_tau[flag][t]);
      // tau_ab -> row v_b of K_al, row v_a of K_bl (every l)
      GMC_UNROLL for (int s = 0; s < 2; s++) {
        const int a = s ? tb[t] : ta[t], b = s ? ta[t] : tb[t];
        if (s == 1 && a == b) break;
        GMC_UNROLL for (int l = 0; l < DIM; l++) km.row[a][l] |= 1u << (1 + b);
      }
    }
  }
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    if (P.n_q_mask[flag] & (1u << k)) {
      qf[k] = T(P.n_q[flag][k]);
      km.eT |= 1u << k;
    }
    GMC_UNROLL for (int i = 0; i < NC - 1; i++)
      if (P.n_J_mask[flag][k] & (1u << i)) {
        J[k][i] = T(P.n_J[flag][k][i]);
        km.Jrow[k] |= 1u << i;
      }
  }
  T H[NC];
  GMC_UNROLL for (int i = 0; i < NC - 1; i++) H[i] = q.pr.e_c[i] + q.p / q.pr.rho_c[i];
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    T HJ = T(0.0);
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) HJ += H[i] * J[k][i];
    T e = T(0);
    GMC_UNROLL for (int a = 0; a < DIM; a++) {
      q.Fdif[k][1 + a] = tau[a][k];
      e = (a == 0) ? tau[k][a] * q.v[a] : e + tau[k][a] * q.v[a];
    }
    q.Fdif[k][IE] = e - qf[k] - HJ;
    GMC_UNROLL for (int i = 0; i < NC - 1; i++) q.Fdif[k][IY + i] = J[k][i];
  }
}

// A_k(i, j) of a boundary point (compile-time k, i, j), without a stored
// matrix: the advective Jacobian
//   A_k = v_k DUDY + U e_{v_k}^T + e_{v_k} e_p^T + e_E (v_k e_p + p e_{v_k})^T
// (fluid_2d.hpp:445-524), or, where the side prescribes the mass flux m of
// direction k (mflux bit k), the Jacobian of F_k = [m, m v + p e_{v_k},
// m E + p v_k, m Y] (fluid_2d.hpp:627-758): 1 at (v_k, p), m on the velocity
// and species diagonals, and the energy row (v_k + m de/dp, m v + p e_{v_k},
// m de/dT, m (e_i - e_last)).
template <typename T, int DIM, int NC>
GMC_HD T a_entry_bd(const QP<T, DIM, NC>& q, unsigned mfmask, const double* mflux, int k, int i,
                    int j) {
  constexpr int IE = DIM + 1, IT = DIM + 1, IY = DIM + 2;
  if ((mfmask >> k) & 1u) {
    const T m = T(mflux[k]);
    const bool vi = i >= 1 && i <= DIM;
    if (vi && j == 0) return i == 1 + k ? T(1.0) : T(0);
    if ((vi || i >= IY) && j == i) return m;
    if (i != IE) return T(0);
    if (j == 0) return q.v[k] + m * ((q.pr.beta * q.p - q.pr.alpha * q.Tt) / q.pr.rho);
    if (j >= 1 && j <= DIM) return j == 1 + k ? q.p + m * q.v[k] : m * q.v[j - 1];
    if (j == IT) return m * (q.pr.cp - q.p * q.pr.alpha / q.pr.rho);
    return m * (q.pr.e_c[j - IY] - q.pr.e_c[NC - 1]);
  }
  return a_entry<T, DIM, NC>(q, k, i, j);
}

// K_kl(i, j) of a boundary point: Kentry with the Neumann suppressions.
template <typename T, int DIM, int NC>
GMC_HD T Kentry_b(const QP<T, DIM, NC>& q, const KMask& km, int k, int l, int i, int j) {
  constexpr int IE = DIM + 1, IY = DIM + 2;
  const T v = Kentry<T, DIM, NC>(q, k, l, i, j);
  if (i >= 1 && i <= DIM && ((km.row[k][l] >> i) & 1u)) return T(0);
  if (i == IE && j >= 1 && j <= DIM) {
    // the energy entry (E, v_j) of K_kl is fed by exactly one velocity row r
    int r;
    if (k == l) r = j;
    else if (j == 1 + k) r = 1 + l;
    else r = 1 + k;
    if ((km.row[k][l] >> r) & 1u) return T(0);
  }
  if (k == l && i == IE && j == IE && ((km.eT >> k) & 1u)) return T(0);
  if (k == l && j >= IY && (i == j || i == IE) && ((km.Jrow[k] >> (j - IY)) & 1u)) return T(0);
  return v;
}

}  // namespace gmc
