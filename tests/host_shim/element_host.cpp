// Test shim: compiles the product's host+device element headers
// (gales-multi-fluid_amd/csrc/gmc_element.hpp) for the CPU and composes the
// per-point contributions exactly as the HIP kernels do (a plain sum over the
// points instead of the DPP quad sum), so tests can compare the kernel math
// against the oracle without a GPU.  Not part of the product.
#include <cstring>
#include <string>

#include "gmc_devparams.hpp"
#include "gmc_dual.hpp"
#include "gmc_element.hpp"

using namespace gmc;

// factored = true: the kernels' matrix phase (MatPhase: the rank form of DUDY
// in 3D, DUDY as a matrix in 2D); false: the direct form (point_Ta /
// point_block), kept as an independent check of the regrouping.
template <int DIM, int NC, typename TM>
static int interior(const gmc_params* p, const double* xe, const double* Ie, const double* Pe,
                    double* m, double* r, bool factored = true) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  DevParams P;
  std::string err;
  if (build_devparams(*p, P, err)) return -1;
  Material<double> M;
  load_material<double, NC>(P, M);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) return -3;
  std::memset(m, 0, sizeof(double) * NE * NE);
  std::memset(r, 0, sizeof(double) * NE);
  for (int q = 0; q < S::NQ; q++) {
    Point<double, DIM, NC> pt;
    if (point_setup<double, DIM, NC>(P, M, g, q, 1.0, Ie, Pe, pt)) return -4;
    double rr[NE];
    point_residual<double, DIM, NC>(P, pt, g, rr);
    for (int t = 0; t < NE; t++) r[t] -= rr[t];
    MatPoint<TM, DIM, NC> mp;
    make_matpoint<TM, DIM, NC>(P, pt, g, mp);
    MatPhase<TM, DIM, NC, true> fac;
    MatPhase<TM, DIM, NC, false> dir;
    if (factored) fac.init(mp);
    for (int a = 0; a < NEN; a++) {
      if (factored) fac.col(mp, a);
      else dir.col(mp, a);
      for (int b = 0; b < NEN; b++) {
        TM blk[N][N];
        if (factored) fac.block(mp, b, a, blk);
        else dir.block(mp, b, a, blk);
        for (int i = 0; i < N; i++)
          for (int j = 0; j < N; j++) m[(b * N + i) * NE + a * N + j] += (double)blk[i][j];
      }
    }
  }
  return 0;
}

// The element matrix as the matrix-core kernel splits it: the Galerkin,
// diffusion and DC blocks (point_block2_full without its least-squares terms)
// plus, per point, the product of the least-squares left factor (columns
// lsq_left_col) with T_a, formed here as plain loops.
template <int DIM, int NC>
static int interior_split(const gmc_params* p, const double* xe, const double* Ie, const double* Pe,
                          double* m) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  DevParams P;
  std::string err;
  if (build_devparams(*p, P, err)) return -1;
  Material<double> M;
  load_material<double, NC>(P, M);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) return -3;
  std::memset(m, 0, sizeof(double) * NE * NE);
  for (int q = 0; q < S::NQ; q++) {
    Point<double, DIM, NC> pt;
    if (point_setup<double, DIM, NC>(P, M, g, q, 1.0, Ie, Pe, pt)) return -4;
    MatPoint<double, DIM, NC> mp;
    make_matpoint<double, DIM, NC>(P, pt, g, mp);
    MatPrep<double, DIM, NC> pr;
    point_prep_full<double, DIM, NC>(mp, pr);
    double L[NEN][N][N];  // L[b][i][l]
    for (int b = 0; b < NEN; b++)
      for (int l = 0; l < N; l++) {
        double col[N];
        lsq_left_col<double, DIM, NC>(mp, b, l, col);
        for (int i = 0; i < N; i++) L[b][i][l] = col[i];
      }
    for (int a = 0; a < NEN; a++) {
      double T[N][N], s[N], eT[N];
      point_col<double, DIM, NC>(mp, pr, a, T, s, eT);
      for (int b = 0; b < NEN; b++) {
        double blk[N][N];
        point_block2_full<double, DIM, NC, false>(mp, nullptr, nullptr, b, a, blk);
        for (int i = 0; i < N; i++)
          for (int j = 0; j < N; j++) {
            double x = blk[i][j];
            for (int l = 0; l < N; l++) x += L[b][i][l] * T[l][j];
            m[(b * N + i) * NE + a * N + j] += x;
          }
      }
    }
  }
  return 0;
}

extern "C" int shim_interior_split(const gmc_params* p, const double* xe, const double* Ie,
                                   const double* Pe, double* m) {
  if (p->dim == 2 && p->nc == 2) return interior_split<2, 2>(p, xe, Ie, Pe, m);
  if (p->dim == 2 && p->nc == 3) return interior_split<2, 3>(p, xe, Ie, Pe, m);
  if (p->dim == 3 && p->nc == 3) return interior_split<3, 3>(p, xe, Ie, Pe, m);
  if (p->dim == 3 && p->nc == 4) return interior_split<3, 4>(p, xe, Ie, Pe, m);
  return -5;
}

template <int DIM, int NC>
static int boundary(const gmc_params* p, const double* xe, const double* Ie, int ns,
                    const int* sl, const int* sf, double* m, double* r) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  DevParams P;
  std::string err;
  if (build_devparams(*p, P, err)) return -1;
  Material<double> M;
  load_material<double, NC>(P, M);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) return -3;
  std::memset(m, 0, sizeof(double) * NE * NE);
  std::memset(r, 0, sizeof(double) * NE);
  for (int s = 0; s < ns; s++)
    for (int q = 0; q < S::NSQ; q++) {
      BPoint<double, DIM, NC> bp;
      bpoint_setup<double, DIM, NC>(P, M, g, xe, sl[s], q, sf[s], 1.0, Ie, bp);
      double rr[NE];
      bpoint_residual<double, DIM, NC>(bp, rr);
      for (int t = 0; t < NE; t++) r[t] -= rr[t];
      for (int b = 0; b < NEN; b++)
        for (int a = 0; a < NEN; a++) {
          double blk[N][N];
          bpoint_block<double, DIM, NC>(bp, g, b, a, blk);
          for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) m[(b * N + i) * NE + a * N + j] += blk[i][j];
        }
    }
  return 0;
}

template <typename TM>
static int interior_any(const gmc_params* p, const double* xe, const double* Ie, const double* Pe,
                        double* m, double* r, bool factored = true) {
  if (p->dim == 2 && p->nc == 2) return interior<2, 2, TM>(p, xe, Ie, Pe, m, r, factored);
  if (p->dim == 2 && p->nc == 3) return interior<2, 3, TM>(p, xe, Ie, Pe, m, r, factored);
  if (p->dim == 3 && p->nc == 3) return interior<3, 3, TM>(p, xe, Ie, Pe, m, r, factored);
  if (p->dim == 3 && p->nc == 4) return interior<3, 4, TM>(p, xe, Ie, Pe, m, r, factored);
  return -5;
}

// the direct (unfactored) matrix phase, FP64
extern "C" int shim_interior_direct(const gmc_params* p, const double* xe, const double* Ie,
                                    const double* Pe, double* m, double* r) {
  return interior_any<double>(p, xe, Ie, Pe, m, r, false);
}

extern "C" int shim_interior(const gmc_params* p, const double* xe, const double* Ie,
                             const double* Pe, double* m, double* r) {
  return interior_any<double>(p, xe, Ie, Pe, m, r);
}

// element matrix of the FP32 tier (residual stays FP64)
extern "C" int shim_interior_fp32(const gmc_params* p, const double* xe, const double* Ie,
                                  const double* Pe, double* m, double* r) {
  return interior_any<float>(p, xe, Ie, Pe, m, r);
}

extern "C" int shim_boundary(const gmc_params* p, const double* xe, const double* Ie, int ns,
                             const int* sl, const int* sf, double* m, double* r) {
  if (p->dim == 2 && p->nc == 2) return boundary<2, 2>(p, xe, Ie, ns, sl, sf, m, r);
  if (p->dim == 2 && p->nc == 3) return boundary<2, 3>(p, xe, Ie, ns, sl, sf, m, r);
  if (p->dim == 3 && p->nc == 3) return boundary<3, 3>(p, xe, Ie, ns, sl, sf, m, r);
  if (p->dim == 3 && p->nc == 4) return boundary<3, 4>(p, xe, Ie, ns, sl, sf, m, r);
  return -5;
}

// Directional derivative of the element residual (returned negated, like r)
// along dI, dP and, for k >= 0, the material parameter theta_k -- the dual
// evaluation the adjoint kernels run per direction.
template <int DIM, int NC>
static int interior_dir(const gmc_params* p, const double* xe, const double* Ie, const double* Pe,
                        const double* dI, const double* dP, int k, double* r, double* dr) {
  using S = Shape<DIM, NC>;
  constexpr int NE = S::NE;
  DevParams P;
  std::string err;
  if (build_devparams(*p, P, err)) return -1;
  Material<Dual> M;
  load_material<Dual, NC>(P, M);
  if (k >= 0) seed_theta<DIM, NC>(P, k, M);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) return -3;
  Dual I[NE], Pv[NE];
  for (int t = 0; t < NE; t++) {
    I[t] = Dual(Ie[t], dI ? dI[t] : 0.0);
    Pv[t] = Dual(Pe[t], dP ? dP[t] : 0.0);
  }
  std::memset(r, 0, sizeof(double) * NE);
  std::memset(dr, 0, sizeof(double) * NE);
  for (int q = 0; q < S::NQ; q++) {
    Point<Dual, DIM, NC> pt;
    if (point_setup<Dual, DIM, NC>(P, M, g, q, 1.0, I, Pv, pt)) return -4;
    Dual rr[NE];
    point_residual<Dual, DIM, NC>(P, pt, g, rr);
    for (int t = 0; t < NE; t++) {
      r[t] -= rr[t].v;
      dr[t] -= rr[t].d;
    }
  }
  return 0;
}

template <int DIM, int NC>
static int boundary_dir(const gmc_params* p, const double* xe, const double* Ie, const double* dI, int k,
                        int ns, const int* sl, const int* sf, double* r, double* dr) {
  using S = Shape<DIM, NC>;
  constexpr int NE = S::NE;
  DevParams P;
  std::string err;
  if (build_devparams(*p, P, err)) return -1;
  Material<Dual> M;
  load_material<Dual, NC>(P, M);
  if (k >= 0) seed_theta<DIM, NC>(P, k, M);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) return -3;
  Dual I[NE];
  for (int t = 0; t < NE; t++) I[t] = Dual(Ie[t], dI ? dI[t] : 0.0);
  std::memset(r, 0, sizeof(double) * NE);
  std::memset(dr, 0, sizeof(double) * NE);
  for (int s = 0; s < ns; s++)
    for (int q = 0; q < S::NSQ; q++) {
      BPoint<Dual, DIM, NC> bp;
      bpoint_setup<Dual, DIM, NC>(P, M, g, xe, sl[s], q, sf[s], 1.0, I, bp);
      Dual rr[NE];
      bpoint_residual<Dual, DIM, NC>(bp, rr);
      for (int t = 0; t < NE; t++) {
        r[t] -= rr[t].v;
        dr[t] -= rr[t].d;
      }
    }
  return 0;
}

extern "C" int shim_interior_dir(const gmc_params* p, const double* xe, const double* Ie,
                                 const double* Pe, const double* dI, const double* dP, int k,
                                 double* r, double* dr) {
  if (p->dim == 2 && p->nc == 2) return interior_dir<2, 2>(p, xe, Ie, Pe, dI, dP, k, r, dr);
  if (p->dim == 2 && p->nc == 3) return interior_dir<2, 3>(p, xe, Ie, Pe, dI, dP, k, r, dr);
  if (p->dim == 3 && p->nc == 3) return interior_dir<3, 3>(p, xe, Ie, Pe, dI, dP, k, r, dr);
  if (p->dim == 3 && p->nc == 4) return interior_dir<3, 4>(p, xe, Ie, Pe, dI, dP, k, r, dr);
  return -5;
}

extern "C" int shim_boundary_dir(const gmc_params* p, const double* xe, const double* Ie,
                                 const double* dI, int k, int ns, const int* sl, const int* sf,
                                 double* r, double* dr) {
  if (p->dim == 2 && p->nc == 2) return boundary_dir<2, 2>(p, xe, Ie, dI, k, ns, sl, sf, r, dr);
  if (p->dim == 2 && p->nc == 3) return boundary_dir<2, 3>(p, xe, Ie, dI, k, ns, sl, sf, r, dr);
  if (p->dim == 3 && p->nc == 3) return boundary_dir<3, 3>(p, xe, Ie, dI, k, ns, sl, sf, r, dr);
  if (p->dim == 3 && p->nc == 4) return boundary_dir<3, 4>(p, xe, Ie, dI, k, ns, sl, sf, r, dr);
  return -5;
}

// The gradient-direction pieces of the exact adjoint at interior point qp:
// nu2 = {dc_nu_grad's nu, point_setup's shk}; dc / dd = d nu / d dI_l,j from
// dc_nu_grad (closed form) / from dc_nu with dual numbers seeded on dI_l,j;
// ta / tb = d tau / d dI_l,(1+i) from tau_matrix with dual numbers seeded on
// dI_l,(1+i) / from d tau/dh (tau_matrix along h) times dh/d dv_i,l
// (compute_h), [l][i][N][N].
template <int DIM, int NC>
static int grad_pieces(const gmc_params* p, const double* xe, const double* Ie, const double* Pe,
                       int qp, double* nu2, double* dc, double* dd, double* ta, double* tb) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF;
  DevParams P;
  std::string err;
  if (build_devparams(*p, P, err)) return -1;
  Material<double> M;
  load_material<double, NC>(P, M);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) return -3;
  Point<double, DIM, NC> pt;
  point_setup<double, DIM, NC>(P, M, g, qp, 1.0, Ie, Pe, pt);
  const auto& Q = pt.Q;
  double dnu[DIM][N];
  nu2[0] = dc_nu_grad<DIM, NC>(P, Q, g, M.uinv, pt.AdI, pt.DdI, dnu);
  nu2[1] = pt.shk;
  QP<Dual, DIM, NC> qd;
  to_dual<DIM, NC>(Q, qd);
  Dual ud[8];
  for (int i = 0; i < 8; i++) ud[i] = Dual(M.uinv[i], 0.0);
  for (int l = 0; l < DIM; l++)
    for (int j = 0; j < N; j++) {
      dc[l * N + j] = dnu[l][j];
      Dual AdId[N], DdId[DIM][N];
      for (int i = 0; i < N; i++) {
        AdId[i] = Dual(pt.AdI[i], a_entry<double, DIM, NC>(Q, l, i, j));
        for (int k = 0; k < DIM; k++) DdId[k][i] = Dual(pt.DdI[k][i], k == l ? Q.DUDY[i][j] : 0.0);
      }
      qd.dI[l][j].d = 1.0;
      dd[l * N + j] = dc_nu<Dual, DIM, NC>(P, qd, g, ud, AdId, DdId).d;
      qd.dI[l][j].d = 0.0;
    }
  double dv[DIM][DIM];
  for (int i = 0; i < DIM; i++)
    for (int j = 0; j < DIM; j++) dv[i][j] = Q.dI[j][1 + i];
  const double h = compute_h<double, DIM>(g.G, Q.v, dv);
  Dual th[N][N];
  tau_matrix<Dual, DIM, NC, false, true>(P, qd, g, th, Dual(h, 1.0));
  Dual vd[DIM], dvd[DIM][DIM];
  for (int i = 0; i < DIM; i++) {
    vd[i] = Dual(Q.v[i], 0.0);
    for (int j = 0; j < DIM; j++) dvd[i][j] = Dual(dv[i][j], 0.0);
  }
  for (int l = 0; l < DIM; l++)
    for (int i = 0; i < DIM; i++) {
      qd.dI[l][1 + i].d = 1.0;
      Dual td[N][N];
      tau_matrix<Dual, DIM, NC>(P, qd, g, td);
      qd.dI[l][1 + i].d = 0.0;
      dvd[i][l].d = 1.0;
      const double hd = compute_h<Dual, DIM>(g.G, vd, dvd).d;
      dvd[i][l].d = 0.0;
      for (int r = 0; r < N; r++)
        for (int c = 0; c < N; c++) {
          const int o = ((l * DIM + i) * N + r) * N + c;
          ta[o] = td[r][c].d;
          tb[o] = th[r][c].d * hd;
        }
    }
  return 0;
}

extern "C" int shim_grad_pieces(const gmc_params* p, const double* xe, const double* Ie,
                                const double* Pe, int qp, double* nu2, double* dc, double* dd,
                                double* ta, double* tb) {
  if (p->dim == 2 && p->nc == 2) return grad_pieces<2, 2>(p, xe, Ie, Pe, qp, nu2, dc, dd, ta, tb);
  if (p->dim == 2 && p->nc == 3) return grad_pieces<2, 3>(p, xe, Ie, Pe, qp, nu2, dc, dd, ta, tb);
  if (p->dim == 3 && p->nc == 3) return grad_pieces<3, 3>(p, xe, Ie, Pe, qp, nu2, dc, dd, ta, tb);
  if (p->dim == 3 && p->nc == 4) return grad_pieces<3, 4>(p, xe, Ie, Pe, qp, nu2, dc, dd, ta, tb);
  return -5;
}
