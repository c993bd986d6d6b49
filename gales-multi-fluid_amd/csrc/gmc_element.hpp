// gmc_element.hpp -- one quadrature point's contribution to the element
// residual and element matrix (host+device).  A kernel lane evaluates one
// point; the element result is the sum over points (the kernels sum with DPP
// quad permutes).
//
// Restates fluid_2d.hpp:171-202 (execute_i), 251-333 (RM_i), 209-245
// (execute_b) and 341-364 (RM_b) of GALES-multi-fluid, regrouped per point:
//   r_b += N_b (U - U_P)/dt - dN_b.F_adv + dN_b.F_dif - N_b S
//          + first_part_b tau Res + dN_b.(shock dI)                   (x JxW)
//   m_ba += N_b DUDY N_a/dt - dN_b.A N_a + dN_b K dN_a - N_b S0 N_a
//          + first_part_b tau R_a + (dN_b . dN_a) shock               (x JxW)
// with first_part_b = sum_k A_k dN_b,k - S0 N_b and
// R_a = DUDY N_a/dt + sum_k A_k dN_a,k - S0 N_a.
#pragma once

#include "gmc_physics.hpp"

namespace gmc {

// quad::interpolate / x_derivative_interpolate (quad.hpp:811-865): out[j] =
// sum_a w[a] ve[a*ndof + j], accumulated in node order without contraction.
template <typename T, int DIM, int NC>
GMC_HD void interpolate(const double* w, const T* ve, T* out) {
  GMC_NO_CONTRACT
  constexpr int N = Shape<DIM, NC>::NDOF, NEN = Shape<DIM, NC>::NEN;
  GMC_UNROLL for (int j = 0; j < N; j++) {
    T s = T(0);
    GMC_UNROLL for (int n = 0; n < NEN; n++) s += T(w[n]) * ve[N * n + j];
    out[j] = s;
  }
}

template <typename T, int DIM, int NC>
struct Point {
  using S = Shape<DIM, NC>;
  static constexpr int N = S::NDOF, NEN = S::NEN;
  QP<T, DIM, NC> Q;
  T tau[N][N];
  T shk;          // dc_scale_fact * nu_shoc; shock matrix = shk * DUDY
  T UP[N], Pq[N], Iq[N];
  T DdI[DIM][N];  // DUDY dI_k
  T AdI[N];       // sum_k A_k dI_k
  double sh[NEN];
  double JxW;
};

// execute_i body for interior point qp (fluid_2d.hpp:179-198).  weight 0
// makes a lane contribute nothing while keeping every value finite.
// Magma components: the FP64 evaluation reads this point's values from the
// point table mt (P state at mt, I state at mt + NC * kMagmaVals, written by
// the magma point kernels); the dual-number one has none.
// Returns status bits: 2 = DUDY singular.  kLU selects tau's inverse by the
// pivoted LU instead of the closed form (tau_matrix).  SEED >= 0 (dual
// numbers): after the interpolation the point value I_q gets the derivative
// e_SEED and the gradients dI none, whatever Ie carries, so every
// derivative of dI is a compile-time 0.
template <typename T>
constexpr int kMagmaMode = std::is_same<T, double>::value ? MG_TABLE : MG_NONE;

template <typename T, int DIM, int NC, bool kLU = false, int SEED = -1>
GMC_HD int point_setup(const DevParams& P, const Material<T>& M, const Geo<DIM>& g, int qp,
                       double weight, const T* Ie, const T* Pe, Point<T, DIM, NC>& pt,
                       const double* mt = nullptr) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN;
  qp_shape<DIM>(qp, pt.sh);
  pt.JxW = g.J * w_in<DIM>() * weight;
  interpolate<T, DIM, NC>(pt.sh, Pe, pt.Pq);
  interpolate<T, DIM, NC>(pt.sh, Ie, pt.Iq);
  int st;
  {
    T p, v[DIM], Tt, Y[NC];
    assign<T, DIM, NC>(pt.Pq, p, v, Tt, Y);
    Props<T, NC> pr;
    // strain rate 0 (fluid_2d.hpp:186)
    properties<T, NC, kMagmaMode<T>>(P, M, p, Tt, Y, pr, T(0.0), mt);
    U_vec<T, DIM, NC>(pr.rho, pr.cv, v, Tt, Y, pt.UP);
    st = pr.st;
  }
  auto& Q = pt.Q;
  GMC_UNROLL for (int k = 0; k < DIM; k++) interpolate<T, DIM, NC>(g.dsh[k], Ie, Q.dI[k]);
  if constexpr (SEED >= 0) {
    GMC_UNROLL for (int j = 0; j < N; j++) {
      pt.Iq[j].d = j == SEED ? 1.0 : 0.0;
      GMC_UNROLL for (int k = 0; k < DIM; k++) Q.dI[k][j].d = 0.0;
    }
  }
  assign<T, DIM, NC>(pt.Iq, Q.p, Q.v, Q.Tt, Q.Y);
  // the I state's magma values are at its shear rate (fluid_2d.hpp:193-194)
  properties<T, NC, kMagmaMode<T>>(P, M, Q.p, Q.Tt, Q.Y, Q.pr, T(0.0),
                                   mt ? mt + NC * kMagmaVals : nullptr);
  st |= Q.pr.st;
  flux_interior<T, DIM, NC, false>(P, Q);
  adv_products<T, DIM, NC>(Q, pt.DdI, pt.AdI);
  if (!tau_matrix<T, DIM, NC, kLU>(P, Q, g, pt.tau)) st |= 2;
  pt.shk = dc_nu<T, DIM, NC>(P, Q, g, M.uinv, pt.AdI, pt.DdI);
  return st;
}

// This point's contribution to r_ (positive sign; the element returns -r_).
// The least-squares term first_part_b tau Res is formed from
// first_part_b x = sum_k dN_b,k (A_k x) - N_b S0 x with the closed form of
// adv_products, so A_k is never stored.
template <typename T, int DIM, int NC>
GMC_HD void point_residual(const DevParams& P, const Point<T, DIM, NC>& pt, const Geo<DIM>& g,
                           T* r) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, IE = DIM + 1;
  const auto& Q = pt.Q;
  const T idt = T(1.0 / P.dt);
  const T JxW = T(pt.JxW);
  T Res[N], tRes[N], dIP[N];
  GMC_UNROLL for (int j = 0; j < N; j++) dIP[j] = (pt.Iq[j] - pt.Pq[j]) * idt;
  GMC_UNROLL for (int i = 0; i < N; i++) {
    T s = T(0.0);
    GMC_UNROLL for (int j = 0; j < N; j++) s += Q.DUDY[i][j] * dIP[j];
    s += pt.AdI[i];
    Res[i] = s - Q.S[i];
  }
  GMC_UNROLL for (int i = 0; i < N; i++) {
    T s = T(0.0);
    GMC_UNROLL for (int j = 0; j < N; j++) s += pt.tau[i][j] * Res[j];
    tRes[i] = s;
  }
  // DUDY tRes and the scalars of first_part_b tRes
  T DtR[N];
  GMC_UNROLL for (int i = 0; i < N; i++) {
    T s = T(0.0);
    GMC_UNROLL for (int j = 0; j < N; j++) s += Q.DUDY[i][j] * tRes[j];
    DtR[i] = s;
  }
  T gtR = T(0.0);  // (rho g) . tRes_v, the S0 part
  GMC_UNROLL for (int k = 0; k < DIM; k++) gtR += Q.pr.rho * T(P.g[k]) * tRes[1 + k];
  GMC_UNROLL for (int b = 0; b < NEN; b++) {
    const T shb = T(pt.sh[b]);
    T cv = T(0.0), cx = T(0.0);
    GMC_UNROLL for (int k = 0; k < DIM; k++) {
      cv += T(g.dsh[k][b]) * Q.v[k];
      cx += T(g.dsh[k][b]) * tRes[1 + k];
    }
    GMC_UNROLL for (int jb = 0; jb < N; jb++) {
      T rr = shb * (Q.U[jb] * JxW - pt.UP[jb] * JxW) * idt;
      GMC_UNROLL for (int k = 0; k < DIM; k++) rr -= T(g.dsh[k][b]) * Q.Fadv[k][jb] * JxW;
      GMC_UNROLL for (int k = 0; k < DIM; k++) rr += T(g.dsh[k][b]) * Q.Fdif[k][jb] * JxW;
      rr -= shb * Q.S[jb] * JxW;
      T l = cv * DtR[jb] + Q.U[jb] * cx;
      if (jb >= 1 && jb <= DIM) l += T(g.dsh[jb - 1][b]) * tRes[0];
      if (jb == IE) l += cv * tRes[0] + Q.p * cx - shb * gtR;
      rr += l * JxW;
      GMC_UNROLL for (int k = 0; k < DIM; k++) rr += T(g.dsh[k][b]) * (pt.shk * pt.DdI[k][jb]) * JxW;
      r[b * N + jb] = rr;
    }
  }
}

// Matrix-phase view of an interior point in the precision TM of the element
// matrix (double for the plain form, float for the FP32 tier).  The point's
// properties, flux matrices and tau are always evaluated in FP64
// (point_setup); this record is their rounded copy.  The advective Jacobians
// are not stored: for this flux, A_k = dF_k/dY satisfies (entry by entry, the
// formulas of fluid_2d.hpp:445-524)
//   A_k = v_k DUDY + U e_{v_k}^T + e_{v_k} e_p^T + e_E (v_k e_p + p e_{v_k})^T,
// so every combination sum_k c_k A_k is (c.v) DUDY plus sparse terms.
//
// DUDY is kept in its factored form DUDY = c d0^T + rho M (the entries of
// fluid_2d.hpp:407-441): row r of DUDY is d(rho phi_r)/dY = phi_r d0^T +
// rho dphi_r/dY with phi = c = [1, v, E, Y_i], d0 = d rho/dY, and
// dphi/dY = 0 (mass row), e_{v_k}, mE = dE/dY (energy row), e_{Y_i}; d0 and
// mE are indexed by the primitive [p, v, T, Y_i], and d0's velocity entries
// are 0.  3 N + 1 values instead of N^2, and products with DUDY cost O(N^2)
// instead of O(N^3).  The record also holds DUDY as a matrix: the 2D matrix
// phase uses that form (MatPhase), and the kernels keep only the fields
// their form reads.
template <typename TM, int DIM, int NC>
struct MatPoint {
  static constexpr int N = Shape<DIM, NC>::NDOF, NEN = DIM + 1;
  TM DUDY[N][N];
  TM c[N];           // [1, v, E, Y_i]
  TM d0[N], mE[N];   // d rho / dY, dE / dY over the primitives
  TM rho;
  TM U[N];
  TM tau[N][N];
  TM p;
  TM lam, mu, kappa, rhoD[NC], rhoDH[NC], v[DIM];
  TM rhog[DIM];      // S0(E, v_k) = rho g_k
  TM shk;
  TM sh[NEN];
  TM dsh[DIM][NEN];
  TM JxW, idt;
  // DUDY(i, j) for compile-time i, j (the velocity columns of d0 are 0)
  GMC_HD TM D(int i, int j) const {
    constexpr int IE = DIM + 1;
    const bool vel_col = j >= 1 && j <= DIM;
    if (i == 0) return vel_col ? TM(0) : d0[j];
    if (i == IE) return vel_col ? rho * mE[j] : c[IE] * d0[j] + rho * mE[j];
    if (vel_col) return j == i ? rho : TM(0);
    return j == i ? c[i] * d0[j] + rho : c[i] * d0[j];
  }
};

template <typename TM, int DIM, int NC>
GMC_HD void make_matpoint(const DevParams& P, const Point<double, DIM, NC>& pt, const Geo<DIM>& g,
                          MatPoint<TM, DIM, NC>& mp) {
  constexpr int N = Shape<DIM, NC>::NDOF, NEN = DIM + 1;
  const auto& Q = pt.Q;
  // the rank factors as flux_interior formed them (read by the 3D form only)
  const double* d0 = Q.d0;
  const double* mE = Q.mE;
  mp.c[0] = TM(1);
  GMC_UNROLL for (int k = 0; k < DIM; k++) mp.c[1 + k] = TM(Q.v[k]);
  mp.c[DIM + 1] = TM(Q.E);
  GMC_UNROLL for (int i = 0; i < NC - 1; i++) mp.c[DIM + 2 + i] = TM(Q.Y[i]);
  mp.rho = TM(Q.pr.rho);
  GMC_UNROLL for (int i = 0; i < N; i++) {
    mp.d0[i] = TM(d0[i]);
    mp.mE[i] = TM(mE[i]);
    mp.U[i] = TM(Q.U[i]);
    GMC_UNROLL for (int j = 0; j < N; j++) {
      mp.tau[i][j] = TM(pt.tau[i][j]);
      mp.DUDY[i][j] = TM(Q.DUDY[i][j]);
    }
  }
  mp.p = TM(Q.p);
  mp.lam = TM(Q.lam);
  mp.mu = TM(Q.mu);
  mp.kappa = TM(Q.kappa);
  GMC_UNROLL for (int i = 0; i < NC; i++) {
    mp.rhoD[i] = TM(Q.rhoD[i]);
    mp.rhoDH[i] = TM(Q.rhoDH[i]);
  }
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    mp.v[k] = TM(Q.v[k]);
    mp.rhog[k] = TM(Q.pr.rho * P.g[k]);
  }
  mp.shk = TM(pt.shk);
  GMC_UNROLL for (int a = 0; a < NEN; a++) {
    mp.sh[a] = TM(pt.sh[a]);
    GMC_UNROLL for (int k = 0; k < DIM; k++) mp.dsh[k][a] = TM(g.dsh[k][a]);
  }
  mp.JxW = TM(pt.JxW);
  mp.idt = TM(1.0 / P.dt);
}

// sum_k c_k A_k(i, j) with cv = c . v (compile-time i, j)
template <typename TM, int DIM, int NC>
GMC_HD TM a_comb(const MatPoint<TM, DIM, NC>& mp, const TM* c, TM cv, int i, int j) {
  constexpr int IE = DIM + 1;
  TM f = cv * mp.D(i, j);
  if (j >= 1 && j <= DIM) f += c[j - 1] * mp.U[i];
  if (i >= 1 && i <= DIM && j == 0) f += c[i - 1];
  if (i == IE && j == 0) f += cv;
  if (i == IE && j >= 1 && j <= DIM) f += mp.p * c[j - 1];
  return f;
}

template <typename TM, int DIM, int NC>
GMC_HD void grad_coef(const MatPoint<TM, DIM, NC>& mp, int a, TM* c, TM& cv) {
  cv = TM(0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    c[k] = mp.dsh[k][a];
    cv += c[k] * mp.v[k];
  }
}

// T_a = tau * R_a  (N x N), R_a = DUDY N_a/dt + sum_k A_k dN_a,k - S0 N_a
template <typename TM, int DIM, int NC>
GMC_HD void point_Ta(const MatPoint<TM, DIM, NC>& mp, int a, TM (*Ta)[Shape<DIM, NC>::NDOF]) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  TM c[DIM], cv;
  grad_coef<TM, DIM, NC>(mp, a, c, cv);
  const TM sa_dt = mp.sh[a] * mp.idt;
  TM R[N][N];
  GMC_UNROLL for (int k = 0; k < N; k++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      TM s = mp.D(k, j) * sa_dt + a_comb<TM, DIM, NC>(mp, c, cv, k, j);
      if (k == DIM + 1 && j >= 1 && j <= DIM) s -= mp.rhog[j - 1] * mp.sh[a];
      R[k][j] = s;
    }
  GMC_UNROLL for (int i = 0; i < N; i++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      TM s = TM(0.0);
      GMC_UNROLL for (int k = 0; k < N; k++) s += mp.tau[i][k] * R[k][j];
      Ta[i][j] = s;
    }
}

// This point's contribution to block (b, a) of m_ (N x N), given Ta = tau R_a.
template <typename TM, int DIM, int NC>
GMC_HD void point_block(const MatPoint<TM, DIM, NC>& mp, int b, int a,
                        const TM (*Ta)[Shape<DIM, NC>::NDOF], TM (*m)[Shape<DIM, NC>::NDOF]) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  const TM JxW = mp.JxW;
  const TM c0 = mp.sh[b] * mp.sh[a] * JxW * mp.idt;
  const TM cs = mp.sh[b] * mp.sh[a] * JxW;
  const TM ca = mp.sh[a] * JxW;
  TM cdc = TM(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) cdc += mp.dsh[k][b] * mp.dsh[k][a];
  cdc = cdc * mp.shk * JxW;
  TM c[DIM], cv;
  grad_coef<TM, DIM, NC>(mp, b, c, cv);
  GMC_UNROLL for (int jb = 0; jb < N; jb++) {
    // row jb of first_part_b = sum_k A_k dN_b,k - S0 N_b (its A part also
    // serves the Galerkin term -dN_b.A N_a)
    TM fa[N];
    GMC_UNROLL for (int i = 0; i < N; i++) fa[i] = a_comb<TM, DIM, NC>(mp, c, cv, jb, i);
    GMC_UNROLL for (int ja = 0; ja < N; ja++) {
      TM mm = c0 * mp.D(jb, ja);
      mm -= ca * fa[ja];
      GMC_UNROLL for (int k = 0; k < DIM; k++)
        GMC_UNROLL for (int l = 0; l < DIM; l++) {
          const TM kv = Kentry<TM, DIM, NC>(mp, k, l, jb, ja);
          mm += mp.dsh[k][b] * mp.dsh[l][a] * JxW * kv;
        }
      if (jb == DIM + 1 && ja >= 1 && ja <= DIM) mm -= cs * mp.rhog[ja - 1];
      TM lsq = TM(0.0);
      GMC_UNROLL for (int i = 0; i < N; i++) {
        TM f = fa[i];
        if (jb == DIM + 1 && i >= 1 && i <= DIM) f -= mp.rhog[i - 1] * mp.sh[b];
        lsq += f * Ta[i][ja];
      }
      mm += lsq * JxW;
      mm += cdc * mp.D(jb, ja);
      m[jb][ja] = mm;
    }
  }
}

// ---------------------------------------------------------------------------
// Matrix phase in factored form (used by the kernels).  Same terms as
// point_Ta / point_block, regrouped with the structure of the flux Jacobians
// so that no A_k, R_a or first_part_b is formed densely:
//   A(c) := sum_k c_k A_k = (c.v) D + U c~^T + c~ e_p^T + (c.v) e_E e_p^T + p e_E c~^T
// (c~ the vector with c in the velocity slots, D = DUDY), S0 = e_E (rho g)_v^T,
//   R_a = beta_a D + U c~_a^T + c~_a e_p^T + cv_a e_E e_p^T + e_E (p c~_a - sa rho g)^T,
//   beta_a = sa/dt + cv_a,
//   T_a := JxW tau R_a, column by column from TD = JxW tau D, TU = JxW tau U and
//   the columns of JxW tau,
//   first_part_b T_a = cv_b (D T_a) + U (c~_b^T T_a) + c~_b T_a[p,:]
//                      + e_E (cv_b T_a[p,:] + (p c~_b - sb rho g)^T T_a[v,:]).
// The Galerkin part per block is
//   JxW [ (sb sa/dt + (c_b.c_a) shk - sa cv_b) D - sa (U c~_b^T + c~_b e_p^T
//         + cv_b e_E e_p^T + p e_E c~_b^T) + sum_kl c_bk c_al K_kl - sb sa S0 ].
// With D = c d0^T + rho M (MatPoint), the products with D are rank-structured:
//   TD = (JxW tau c) d0^T + rho (JxW tau M),  (tau M)[:, j] = [j in v, Y] tau[:, j]
//                                                            + tau[:, E] mE_j,
//   D T_a = c s^T + rho M T_a,  s = T_a^T d0,  (M T_a) = T_a's v and Y rows,
//                                   its E row eT = T_a^T mE, 0 in the mass row,
// so neither product costs N^3 and D T_a is never stored: the block combines
// gD D + cv_b D T_a entry by entry as c_i u_j + rho (gD M + cv_b M T_a)_ij
// with u = gD d0 + cv_b s.
// ---------------------------------------------------------------------------
template <typename TM, int DIM, int NC>
struct MatPrep {
  static constexpr int N = Shape<DIM, NC>::NDOF;
  TM TD[N][N];     // JxW tau D
  TM TU[N];        // JxW tau U
  TM tv[DIM][N];   // JxW tau[:, v_k]
  TM tE[N];        // JxW tau[:, E]
};

template <typename TM, int DIM, int NC>
GMC_HD void point_prep(const MatPoint<TM, DIM, NC>& mp, MatPrep<TM, DIM, NC>& pr) {
  constexpr int N = Shape<DIM, NC>::NDOF, IE = DIM + 1;
  TM t[N][N];
  GMC_UNROLL for (int i = 0; i < N; i++)
    GMC_UNROLL for (int k = 0; k < N; k++) t[i][k] = mp.JxW * mp.tau[i][k];
  GMC_UNROLL for (int i = 0; i < N; i++) {
    TM tc = t[i][0];
    GMC_UNROLL for (int k = 1; k < N; k++) tc += t[i][k] * mp.c[k];
    GMC_UNROLL for (int j = 0; j < N; j++) {
      // [j in v, Y] t[i][j] + t[i][E] mE_j, scaled by rho, plus tc d0_j
      const bool vel = j >= 1 && j <= DIM, spec = j > IE;
      TM m = t[i][IE] * mp.mE[j];
      if (vel || spec) m += t[i][j];
      pr.TD[i][j] = vel ? mp.rho * m : tc * mp.d0[j] + mp.rho * m;
    }
    TM s = TM(0);
    GMC_UNROLL for (int k = 0; k < N; k++) s += t[i][k] * mp.U[k];
    pr.TU[i] = s;
    GMC_UNROLL for (int k = 0; k < DIM; k++) pr.tv[k][i] = t[i][1 + k];
    pr.tE[i] = t[i][IE];
  }
}

// T_a = JxW tau R_a for column node a, and the two row combinations that give
// D T_a = c s^T + rho M T_a: s = T_a^T d0, eT = T_a^T mE
template <typename TM, int DIM, int NC>
GMC_HD void point_col(const MatPoint<TM, DIM, NC>& mp, const MatPrep<TM, DIM, NC>& pr, int a,
                      TM (*T)[Shape<DIM, NC>::NDOF], TM* s, TM* eT) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  TM c[DIM], cv;
  grad_coef<TM, DIM, NC>(mp, a, c, cv);
  const TM sa = mp.sh[a];
  const TM beta = sa * mp.idt + cv;
  TM h[DIM];
  GMC_UNROLL for (int j = 0; j < DIM; j++) h[j] = mp.p * c[j] - sa * mp.rhog[j];
  GMC_UNROLL for (int i = 0; i < N; i++) {
    TM s0 = beta * pr.TD[i][0] + cv * pr.tE[i];
    GMC_UNROLL for (int k = 0; k < DIM; k++) s0 += c[k] * pr.tv[k][i];
    T[i][0] = s0;
    GMC_UNROLL for (int j = 0; j < DIM; j++)
      T[i][1 + j] = beta * pr.TD[i][1 + j] + c[j] * pr.TU[i] + h[j] * pr.tE[i];
    GMC_UNROLL for (int j = DIM + 1; j < N; j++) T[i][j] = beta * pr.TD[i][j];
  }
  GMC_UNROLL for (int j = 0; j < N; j++) {
    // the velocity entries of d0 are 0
    TM sj = mp.d0[0] * T[0][j];
    GMC_UNROLL for (int k = DIM + 1; k < N; k++) sj += mp.d0[k] * T[k][j];
    s[j] = sj;
    TM ej = TM(0);
    GMC_UNROLL for (int k = 0; k < N; k++) ej += mp.mE[k] * T[k][j];
    eT[j] = ej;
  }
}

// block (b, a) of this point's contribution to m_ (N x N).  Every entry is
// one multiply-add chain (sparse Galerkin terms, diffusion, then the D and
// least-squares terms), so the compiler contracts it into FMAs.
template <typename TM, int DIM, int NC>
GMC_HD void point_block2(const MatPoint<TM, DIM, NC>& mp, const TM (*T)[Shape<DIM, NC>::NDOF],
                         const TM* s, const TM* eT, int b, int a, TM (*m)[Shape<DIM, NC>::NDOF]) {
  constexpr int N = Shape<DIM, NC>::NDOF, IE = DIM + 1;
  const TM JxW = mp.JxW;
  TM cb[DIM], cvb;
  grad_coef<TM, DIM, NC>(mp, b, cb, cvb);
  const TM sa = mp.sh[a], sb = mp.sh[b];
  TM cbca = TM(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) cbca += cb[k] * mp.dsh[k][a];
  // Galerkin coefficients (JxW folded)
  const TM gD = (sb * sa * mp.idt + cbca * mp.shk - sa * cvb) * JxW;
  const TM gA = sa * JxW;
  const TM gS = sb * sa * JxW;
  TM gAc[DIM], gEv[DIM], ck[DIM][DIM];
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    gAc[k] = gA * cb[k];
    gEv[k] = gA * mp.p * cb[k] + gS * mp.rhog[k];
    GMC_UNROLL for (int l = 0; l < DIM; l++) ck[k][l] = cb[k] * mp.dsh[l][a] * JxW;
  }
  const TM gAcv = gA * cvb;
  // LSQ row combinations of T
  TM w[N], gE[N];
  GMC_UNROLL for (int j = 0; j < N; j++) {
    TM sw = TM(0), t = TM(0);
    GMC_UNROLL for (int k = 0; k < DIM; k++) {
      sw += cb[k] * T[1 + k][j];
      t += (mp.p * cb[k] - sb * mp.rhog[k]) * T[1 + k][j];
    }
    w[j] = sw;
    gE[j] = t + cvb * T[0][j];
  }
  // gD D + cv_b D T_a = c_i u_j + rho (gD M + cv_b M T_a)_ij, u = gD d0 + cv_b s
  const TM rg = mp.rho * gD, rc = mp.rho * cvb;
  TM u[N];
  GMC_UNROLL for (int j = 0; j < N; j++) u[j] = (j >= 1 && j <= DIM) ? cvb * s[j] : gD * mp.d0[j] + cvb * s[j];
  GMC_UNROLL for (int i = 0; i < N; i++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      // Galerkin: sparse advective and source terms, diffusion
      TM acc = TM(0.0);
      if (j >= 1 && j <= DIM) acc -= gAc[j - 1] * mp.U[i];
      if (i >= 1 && i <= DIM && j == 0) acc -= gAc[i - 1];
      if (i == IE && j == 0) acc -= gAcv;
      if (i == IE && j >= 1 && j <= DIM) acc -= gEv[j - 1];
      GMC_UNROLL for (int k = 0; k < DIM; k++)
        GMC_UNROLL for (int ll = 0; ll < DIM; ll++)
          acc += ck[k][ll] * Kentry<TM, DIM, NC>(mp, k, ll, i, j);
      // Galerkin D term and the cv_b D T_a part of the least squares
      if (i == 0) {
        acc += u[j];
      } else if (i == IE) {
        acc += mp.c[IE] * u[j];
        acc += rg * mp.mE[j];
        acc += rc * eT[j];
      } else {
        acc += mp.c[i] * u[j];
        acc += rc * T[i][j];
        if (j == i) acc += rg;
      }
      // the rest of the least squares: first_part_b T_a
      if (i >= 1 && i <= DIM) acc += cb[i - 1] * T[0][j];
      if (i == IE) acc += gE[j];
      acc += mp.U[i] * w[j];
      m[i][j] = acc;
    }
}

// The same factored phase with DUDY as an N x N matrix: TD = JxW tau D and
// DT_a = D T_a are plain products and D T_a is stored per column node.  The
// compiler folds DUDY's structural zeros (its entries are formed in the same
// kernel), which in 2D leaves fewer operations than the rank form.
template <typename TM, int DIM, int NC>
GMC_HD void point_prep_full(const MatPoint<TM, DIM, NC>& mp, MatPrep<TM, DIM, NC>& pr) {
  constexpr int N = Shape<DIM, NC>::NDOF, IE = DIM + 1;
  TM t[N][N];
  GMC_UNROLL for (int i = 0; i < N; i++)
    GMC_UNROLL for (int k = 0; k < N; k++) t[i][k] = mp.JxW * mp.tau[i][k];
  GMC_UNROLL for (int i = 0; i < N; i++) {
    GMC_UNROLL for (int j = 0; j < N; j++) {
      TM s = TM(0);
      GMC_UNROLL for (int k = 0; k < N; k++) s += t[i][k] * mp.DUDY[k][j];
      pr.TD[i][j] = s;
    }
    TM s = TM(0);
    GMC_UNROLL for (int k = 0; k < N; k++) s += t[i][k] * mp.U[k];
    pr.TU[i] = s;
    GMC_UNROLL for (int k = 0; k < DIM; k++) pr.tv[k][i] = t[i][1 + k];
    pr.tE[i] = t[i][IE];
  }
}

// T_a = JxW tau R_a (point_col) and DT_a = D T_a
template <typename TM, int DIM, int NC>
GMC_HD void point_col_full(const MatPoint<TM, DIM, NC>& mp, const MatPrep<TM, DIM, NC>& pr, int a,
                           TM (*T)[Shape<DIM, NC>::NDOF], TM (*DT)[Shape<DIM, NC>::NDOF]) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  TM s[N], eT[N];
  point_col<TM, DIM, NC>(mp, pr, a, T, s, eT);
  GMC_UNROLL for (int i = 0; i < N; i++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      TM x = TM(0);
      GMC_UNROLL for (int k = 0; k < N; k++) x += mp.DUDY[i][k] * T[k][j];
      DT[i][j] = x;
    }
}

// point_block2 with the Galerkin D term gD D and the least-squares term
// cv_b D T_a read from the matrices.  kLsq = false leaves out every
// least-squares term (T and DT unread): the Galerkin, diffusion and
// discontinuity-capturing part of the block, for kernels that form the least
// squares as the product lsq_left_col x T_a on the matrix cores.
template <typename TM, int DIM, int NC, bool kLsq = true>
GMC_HD void point_block2_full(const MatPoint<TM, DIM, NC>& mp, const TM (*T)[Shape<DIM, NC>::NDOF],
                              const TM (*DT)[Shape<DIM, NC>::NDOF], int b, int a,
                              TM (*m)[Shape<DIM, NC>::NDOF]) {
  constexpr int N = Shape<DIM, NC>::NDOF, IE = DIM + 1;
  const TM JxW = mp.JxW;
  TM cb[DIM], cvb;
  grad_coef<TM, DIM, NC>(mp, b, cb, cvb);
  const TM sa = mp.sh[a], sb = mp.sh[b];
  TM cbca = TM(0.0);
  GMC_UNROLL for (int k = 0; k < DIM; k++) cbca += cb[k] * mp.dsh[k][a];
  const TM gD = (sb * sa * mp.idt + cbca * mp.shk - sa * cvb) * JxW;
  const TM gA = sa * JxW;
  const TM gS = sb * sa * JxW;
  TM gAc[DIM], gEv[DIM], ck[DIM][DIM];
  GMC_UNROLL for (int k = 0; k < DIM; k++) {
    gAc[k] = gA * cb[k];
    gEv[k] = gA * mp.p * cb[k] + gS * mp.rhog[k];
    GMC_UNROLL for (int l = 0; l < DIM; l++) ck[k][l] = cb[k] * mp.dsh[l][a] * JxW;
  }
  const TM gAcv = gA * cvb;
  TM w[N], gE[N];
  if constexpr (kLsq) {
    GMC_UNROLL for (int j = 0; j < N; j++) {
      TM sw = TM(0), t = TM(0);
      GMC_UNROLL for (int k = 0; k < DIM; k++) {
        sw += cb[k] * T[1 + k][j];
        t += (mp.p * cb[k] - sb * mp.rhog[k]) * T[1 + k][j];
      }
      w[j] = sw;
      gE[j] = t + cvb * T[0][j];
    }
  }
  GMC_UNROLL for (int i = 0; i < N; i++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      TM acc = TM(0.0);
      if (j >= 1 && j <= DIM) acc -= gAc[j - 1] * mp.U[i];
      if (i >= 1 && i <= DIM && j == 0) acc -= gAc[i - 1];
      if (i == IE && j == 0) acc -= gAcv;
      if (i == IE && j >= 1 && j <= DIM) acc -= gEv[j - 1];
      GMC_UNROLL for (int k = 0; k < DIM; k++)
        GMC_UNROLL for (int ll = 0; ll < DIM; ll++)
          acc += ck[k][ll] * Kentry<TM, DIM, NC>(mp, k, ll, i, j);
      acc += gD * mp.DUDY[i][j];
      if constexpr (kLsq) {
        if (i >= 1 && i <= DIM) acc += cb[i - 1] * T[0][j];
        if (i == IE) acc += gE[j];
        acc += mp.U[i] * w[j];
        acc += cvb * DT[i][j];
      }
      m[i][j] = acc;
    }
}

// Column l of the least-squares left factor of row node b at this point:
// L_b[i][l] with  sum_l L_b[i][l] T_a[l][j]  the least-squares block (b, a)
// of point_block2_full (first_part_b T_a, fluid_2d.hpp:279-305):
//   L_b = cv_b D + U c~_b^T + c~_b e_p^T + cv_b e_E e_p^T + e_E (p c~_b - s_b rho g)^T
// in the row/column order of point_block2_full's terms (c~_b: its gradient
// coefficients in the velocity slots).
template <typename TM, int DIM, int NC>
GMC_HD void lsq_left_col(const MatPoint<TM, DIM, NC>& mp, int b, int l,
                         TM* col) {
  constexpr int N = Shape<DIM, NC>::NDOF, IE = DIM + 1;
  TM cb[DIM], cvb;
  grad_coef<TM, DIM, NC>(mp, b, cb, cvb);
  const TM sb = mp.sh[b];
  GMC_UNROLL for (int i = 0; i < N; i++) {
    TM x = cvb * mp.DUDY[i][l];
    if (l >= 1 && l <= DIM) x += mp.U[i] * cb[l - 1];
    if (i >= 1 && i <= DIM && l == 0) x += cb[i - 1];
    if (i == IE && l == 0) x += cvb;
    if (i == IE && l >= 1 && l <= DIM) x += mp.p * cb[l - 1] - sb * mp.rhog[l - 1];
    col[i] = x;
  }
}

// The matrix phase of one point as the kernels run it: the factored form,
// with DUDY's rank structure in 3D (point_prep / point_col / point_block2;
// measured on the 128^3 box: 3D FP64 assembly 124 -> 84 ms, register spill
// 992 -> 576 B per lane) and DUDY as a matrix in 2D (point_prep_full /
// point_col_full / point_block2_full; on the 2D channel the rank form ran
// 1 % faster in FP64 and 7 % slower in the FP32 tier).  The direct form
// (point_Ta / point_block) is selectable for comparison; it needs more
// registers in 3D and more instructions in 2D.
template <typename TM, int DIM, int NC, bool kFactored = true>
struct MatPhase {
  static constexpr int N = Shape<DIM, NC>::NDOF;
  static constexpr bool kRank = DIM == 3;
  MatPrep<TM, DIM, NC> prep;
  TM T[N][N], s[N], eT[N], DT[N][N];
  GMC_HD void init(const MatPoint<TM, DIM, NC>& mp) {
    if constexpr (kFactored && kRank) point_prep<TM, DIM, NC>(mp, prep);
    else if constexpr (kFactored) point_prep_full<TM, DIM, NC>(mp, prep);
  }
  GMC_HD void col(const MatPoint<TM, DIM, NC>& mp, int a) {
    if constexpr (kFactored && kRank) point_col<TM, DIM, NC>(mp, prep, a, T, s, eT);
    else if constexpr (kFactored) point_col_full<TM, DIM, NC>(mp, prep, a, T, DT);
    else point_Ta<TM, DIM, NC>(mp, a, T);
  }
  GMC_HD void block(const MatPoint<TM, DIM, NC>& mp, int b, int a, TM (*m)[N]) const {
    if constexpr (kFactored && kRank) point_block2<TM, DIM, NC>(mp, T, s, eT, b, a, m);
    else if constexpr (kFactored) point_block2_full<TM, DIM, NC>(mp, T, DT, b, a, m);
    else point_block<TM, DIM, NC>(mp, b, a, T, m);
  }
};

// ---------------------------------------------------------------------------
// boundary points (execute_b, fluid_2d.hpp:209-245; RM_b 341-364)
// ---------------------------------------------------------------------------
template <typename T, int DIM, int NC>
struct BPoint {
  using S = Shape<DIM, NC>;
  static constexpr int N = S::NDOF, NEN = S::NEN;
  QP<T, DIM, NC> Q;
  KMask km;
  unsigned mfmask;      // directions whose mass flux the side prescribes
  double mflux[DIM];    // those fluxes (a_entry_bd)
  double sh[NEN];
  double JNxW[DIM];
};

template <typename T, int DIM, int NC>
GMC_HD void bpoint_setup(const DevParams& P, const Material<T>& M, const Geo<DIM>& g,
                         const double* xe, int side, int sq, int flag, double weight,
                         const T* Ie, BPoint<T, DIM, NC>& bp, const double* mt = nullptr) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN;
  double xi[DIM], JN[DIM];
  side_point<DIM>(side, sq, xi);
  side_shape<DIM>(xi, bp.sh);
  side_JN<DIM>(xe, xi, JN);
  GMC_UNROLL for (int k = 0; k < DIM; k++) bp.JNxW[k] = JN[k] * w_bd<DIM>() * weight;
  auto& Q = bp.Q;
  T Iq[N];
  interpolate<T, DIM, NC>(bp.sh, Ie, Iq);
  GMC_UNROLL for (int k = 0; k < DIM; k++) interpolate<T, DIM, NC>(g.dsh[k], Ie, Q.dI[k]);
  assign<T, DIM, NC>(Iq, Q.p, Q.v, Q.Tt, Q.Y);
  // magma values from the side-point table mt, at the point's shear rate
  // (fluid_2d.hpp:233-235)
  properties<T, NC, kMagmaMode<T>>(P, M, Q.p, Q.Tt, Q.Y, Q.pr, T(0.0), mt);
  flux_interior<T, DIM, NC, false>(P, Q);
  flux_boundary<T, DIM, NC>(P, flag, Q, bp.km);
  bp.mfmask = P.mflux_mask[flag];
  GMC_UNROLL for (int k = 0; k < DIM; k++) bp.mflux[k] = P.mflux[flag][k];
}

template <typename T, int DIM, int NC>
GMC_HD void bpoint_residual(const BPoint<T, DIM, NC>& bp, T* r) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN;
  GMC_UNROLL for (int b = 0; b < NEN; b++)
    GMC_UNROLL for (int j = 0; j < N; j++) {
      T rr = T(0);
      GMC_UNROLL for (int k = 0; k < DIM; k++) rr += T(bp.sh[b]) * bp.Q.Fadv[k][j] * T(bp.JNxW[k]);
      GMC_UNROLL for (int k = 0; k < DIM; k++) rr -= T(bp.sh[b]) * bp.Q.Fdif[k][j] * T(bp.JNxW[k]);
      r[b * N + j] = rr;
    }
}

// Column a of a boundary point's blocks: block (b, a) = N_b C_a with
//   C_a = N_a sum_k JNxW_k A_k - sum_kl JNxW_k K_kl dN_a,l
// (RM_b, fluid_2d.hpp:341-364), so the NEN blocks of a column share one C_a.
template <typename T, int DIM, int NC>
GMC_HD void bpoint_col(const BPoint<T, DIM, NC>& bp, const Geo<DIM>& g, int a,
                       T (*C)[Shape<DIM, NC>::NDOF]) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  GMC_UNROLL for (int jb = 0; jb < N; jb++)
    GMC_UNROLL for (int ja = 0; ja < N; ja++) {
      T mm = T(0);
      GMC_UNROLL for (int k = 0; k < DIM; k++)
        mm += a_entry_bd<T, DIM, NC>(bp.Q, bp.mfmask, bp.mflux, k, jb, ja) * T(bp.JNxW[k]) * T(bp.sh[a]);
      GMC_UNROLL for (int k = 0; k < DIM; k++)
        GMC_UNROLL for (int l = 0; l < DIM; l++)
          mm -= Kentry_b<T, DIM, NC>(bp.Q, bp.km, k, l, jb, ja) * T(bp.JNxW[k]) * T(g.dsh[l][a]);
      C[jb][ja] = mm;
    }
}

template <typename T, int DIM, int NC>
GMC_HD void bpoint_block(const BPoint<T, DIM, NC>& bp, const Geo<DIM>& g, int b, int a,
                         T (*m)[Shape<DIM, NC>::NDOF]) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  T C[N][N];
  bpoint_col<T, DIM, NC>(bp, g, a, C);
  GMC_UNROLL for (int jb = 0; jb < N; jb++)
    GMC_UNROLL for (int ja = 0; ja < N; ja++) m[jb][ja] = T(bp.sh[b]) * C[jb][ja];
}

}  // namespace gmc
