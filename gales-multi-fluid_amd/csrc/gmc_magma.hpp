// gmc_magma.hpp -- properties of one magma submixture, host + device.
//
// A magma component of the fluid_mc mixture (mix_of_mix, fluid_properties.hpp:
// 171-197) is itself a mixture of five chemicals, indexed
//   0 melt, 1 dissolved H2O, 2 dissolved CO2, 3 gaseous H2O, 4 gaseous CO2
// whose mass fractions follow from the solubility tables at the point's
// (p, T) (magma_mixture::submix_of_comp, mixtures.hpp:197-261).  Restated:
//   melt               chemicals.hpp:56-130 (Lange-Carmichael molar volume)
//   dissolved volatile chemicals.hpp:136-185 (polynomial molar volume)
//   ideal gas          chemicals.hpp:191-271 (mu 1e-5, Prandtl 0.72)
//   table lookup       mixtures.hpp:342-412 (linear in p; bilinear in T, p)
//   viscosity          GiordanoEtAlViscosityModel.hpp:36-112 (Giordano et al.
//                      2008), then the gas and crystal corrections of
//                      mixtures.hpp:182-190, 266-336
// Evaluated in FP64 only.  A lookup outside its table, or a viscosity model
// that cannot be evaluated, sets a status bit instead of aborting as the
// reference's Error() does; the returned values are then finite placeholders.
#pragma once

#include <cmath>

#if !defined(GMC_HD)
#if defined(__HIPCC__)
#define GMC_HD __host__ __device__ __forceinline__
#else
#define GMC_HD inline
#endif
#endif

#if defined(__HIPCC__)
#define GMC_HD_NOINLINE __host__ __device__ __attribute__((noinline)) inline
#else
#define GMC_HD_NOINLINE inline
#endif

namespace gmc {

// status bits of a property evaluation (shared with the element status word)
static constexpr int kStTable = 4;     // (p, T) outside a solubility table
static constexpr int kStModel = 8;     // viscosity model division by zero / no termination

struct MagmaDev {
  double wf_comp[3];              // melt, H2O (l + g), CO2 (l + g) per magma mass
  // melt: molar mass, molar volume at 1673 K / 1e5 Pa and its T and p slopes,
  // cp (molar-fraction weighted), kappa
  double melt_M, melt_V, melt_dVdT, melt_dVdp, melt_cp, melt_kappa;
  double liq_M[2], liq_cp[2], liq_kappa[2];        // dissolved H2O, CO2
  double gas_R[2], gas_gamma[2], gas_cv[2], gas_cp[2], gas_kappa[2];  // gaseous H2O, CO2
  double crystal_vf, crystal_rho;
  int gas_mu, crystal_mu;         // GMC_MU_* models
  double gio_M;                   // Giordano: molar mass (g) of the Fe-total oxide set
  double gio[7];                  // Giordano: composition parameters
  int n_p, n_T;                   // table sizes (n_T = 0: pressure only)
  const double* p_tab;            // [n_p]
  const double* T_tab;            // [n_T]
  const double* wh;               // dissolved H2O mass fraction of the liquid
  const double* wc;               // dissolved CO2 mass fraction of the liquid
};

struct MagmaProps {
  double rho, mu, cp, cv, kappa, c, alpha, beta, D;
  int st;
};

// Dissolved-volatile molar-volume polynomial (chemicals.hpp:146)
static constexpr double kVolA[10] = {9.144e-06, 3.685e-09, 1.168e-11, -1.990e-15, 1.220e-16,
                                     -1.945e-17, -1.580e-21, 4.680e-24, 1.144e-26, -3.960e-33};

// interpolation(p, pressure_table, y) (mixtures.hpp:342-356): index
// floor((p - p_0) / (p_1 - p_0)) must leave room for its right neighbour.
GMC_HD bool magma_lerp(const double* pt, const double* y, int n, double p, double& out) {
  const double dp = pt[1] - pt[0];
  const double x = floor((p - pt[0]) / dp);
  if (!(x >= 0.0) || !(x <= double(n - 2))) return false;
  const int i = int(x);
  const double slope = (y[i + 1] - y[i]) / dp;
  out = y[i] + slope * (p - pt[i]);
  return true;
}

// interpolation(T, p, T_table, p_table, m) (mixtures.hpp:373-412): bilinear
// on the cell found from |p - p_0| and |T - T_0|; m is row-major [T][p].
GMC_HD bool magma_bilerp(const double* Tt, const double* pt, const double* m, int nT, int np,
                         double T, double p, double& out) {
  const double dp = fabs(pt[1] - pt[0]);
  const double xi = floor(fabs(p - pt[0]) / dp);
  const double dT = fabs(Tt[1] - Tt[0]);
  const double xj = floor(fabs(T - Tt[0]) / dT);
  if (!(xi >= 0.0) || !(xi <= double(np - 2))) return false;
  if (!(xj >= 0.0) || !(xj <= double(nT - 2))) return false;
  const int i = int(xi), j = int(xj);
  const double p1 = pt[i], p2 = pt[i + 1], T1 = Tt[j], T2 = Tt[j + 1];
  const double* r1 = m + (long long)j * np;
  const double* r2 = r1 + np;
  double mean = (T2 - T) * (p2 - p) * r1[i] + (T2 - T) * (p - p1) * r1[i + 1] +
                (T - T1) * (p2 - p) * r2[i] + (T - T1) * (p - p1) * r2[i + 1];
  mean /= dT * dp;
  out = mean;
  return true;
}

// GiordanoEtAlViscosityModel::computeViscosity (GiordanoEtAlViscosityModel.hpp:81-109)
GMC_HD double giordano_mu(const MagmaDev& m, double T, double wh, double wc, int& st) {
  if (1.0 - wc == 0.0) {
    st |= kStModel;
    return 1.0;
  }
  const double h2o_M = 18.015;
  const double w = wh / (1 - wc);
  double xf = 1 + h2o_M * (1 - w) / (w * m.gio_M);
  xf = 1.0 / xf;
  const double xn = 1 - xf;
  const double hp = 100.0 * xf;
  const double A = m.gio[0];
  const double B = m.gio[1] * xn + m.gio[2] * xn * xn + m.gio[3] * xn * hp + (-84.1 + 141.5) * hp +
                   141.5 * log(1 + hp);
  const double C = m.gio[4] * xn + -99.5 * log(1 + hp) + m.gio[5] * xn * xn + m.gio[6] * xn * hp;
  if (T - C == 0.0) {
    st |= kStModel;
    return 1.0;
  }
  return pow(10., A + B / (T - C));
}

// Pal (2003) eq. 21 (mixtures.hpp:266-291): the reference's Newton loop runs
// until a residual F >= 1e-9 has been evaluated; here it is also bounded.
GMC_HD double pal_2003(double vf_g, double mu, double sr, int& st) {
  if (sr < 2.220446049250313e-16) return mu;
  const double sigma = 0.36, nb = 1.e11, phim = 0.9999;
  const double r = cbrt(3.0 / (4 * 3.142 * nb));
  const double Ca = r * mu * sr / sigma;
  const double a = 2.4 * pow(Ca, 2);
  const double Tm = pow(1.0 - a, -0.8) * pow(1.0 - vf_g / phim, -phim);
  double x = 1.e-5, F;
  int it = 0;
  do {
    F = pow(x, 5) / pow(1.0 - a * x * x, 4) - pow(Tm, 5);
    const double Fp = 5 * pow(x, 4) / pow(1.0 - a * x * x, 4) + 8 * a * pow(x, 6) / pow(1.0 - a * x * x, 5);
    x -= F / Fp;
    if (++it >= 200) {
      st |= kStModel;
      return mu;
    }
  } while (F < 1.e-9);
  return mu * x;
}

// Caricchi et al. (2007) eq. 3 (mixtures.hpp:313-331)
GMC_HD double caricchi(double phi, double mu, double sr_in) {
  double sr = sr_in;
  if (sr < 2.220446049250313e-16) sr = 1.e-6;
  const double ls = log10(sr);
  const double phi_max = 0.066499 * tanh(0.913424 * ls + 3.850623) + 0.591806;
  const double delta = -6.301095 * tanh(0.818496 * ls + 2.86) + 7.462405;
  const double alpha = -0.000378 * tanh(1.148101 * ls + 3.92) + 0.999572;
  const double gamma = 3.987815 * tanh(0.8908 * ls + 3.24) + 5.099645;
  const double num = 1.0 + pow(phi / phi_max, delta);
  const double x = sqrt(3.142) * phi / (2 * alpha * phi_max) * (1.0 + pow(phi / phi_max, gamma));
  const double den = pow(1.0 - alpha * erf(x), 2.5 * phi_max);
  return mu * (num / den);
}

// magma_mixture::properties (mixtures.hpp:172-191, 197-261) at (p, T) with
// strain rate sr.  Out of line: its callers (the magma point kernels, the
// time-step kernel, host code) keep their own register allocation.
GMC_HD_NOINLINE MagmaProps magma_props(const MagmaDev& m, double p, double T, double sr) {
  MagmaProps o;
  o.st = 0;
  double wl1 = 0.0, wl2 = 0.0;
  bool ok;
  if (m.n_T == 0) {
    ok = magma_lerp(m.p_tab, m.wh, m.n_p, p, wl1) && magma_lerp(m.p_tab, m.wc, m.n_p, p, wl2);
  } else {
    ok = magma_bilerp(m.T_tab, m.p_tab, m.wh, m.n_T, m.n_p, T, p, wl1) &&
         magma_bilerp(m.T_tab, m.p_tab, m.wc, m.n_T, m.n_p, T, p, wl2);
  }
  if (!ok) {
    o.st |= kStTable;
    wl1 = wl2 = 0.0;
  }
  const double wl0 = 1.0 - wl1 - wl2;
  double wf[5];
  wf[0] = m.wf_comp[0];
  const double wph_l = wf[0] / wl0;
  wf[1] = wl1 * wph_l;
  wf[2] = wl2 * wph_l;
  wf[3] = m.wf_comp[1] - wf[1];
  wf[4] = m.wf_comp[2] - wf[2];

  double rc[5], cpc[5], cvc[5], ac[5], bc[5], kc[5], sc[5];
  {  // melt
    const double mv = m.melt_V + m.melt_dVdT * (T - 1673.0) + m.melt_dVdp * (p - 1.e5);
    ac[0] = 1. / mv * m.melt_dVdT;
    bc[0] = -1. / mv * m.melt_dVdp;
    rc[0] = m.melt_M / mv;
    sc[0] = 1.0 / sqrt(rc[0] * bc[0]);
    cpc[0] = m.melt_cp;
    cvc[0] = m.melt_cp - T * ac[0] * ac[0] / (rc[0] * bc[0]);
    kc[0] = m.melt_kappa;
  }
  const double* a = kVolA;
  for (int k = 0; k < 2; k++) {  // dissolved H2O, CO2
    const double mv = a[0] + a[1] * T + a[2] * T * T + a[3] * T * T * T + p * (a[4] + a[5] * T + a[6] * T * T) +
                      p * p * (a[7] + a[8] * T) + a[9] * p * p * p;
    const int c = 1 + k;
    rc[c] = m.liq_M[k] / mv;
    double al = a[1] + 2 * a[2] * T + 3 * a[3] * T * T + p * (a[5] + 2 * a[6] * T) + p * p * a[8];
    ac[c] = al / mv;
    double be = -((a[4] + a[5] * T + a[6] * T * T) + 2 * p * (a[7] + a[8] * T) + 3 * a[9] * p * p);
    bc[c] = be / mv;
    sc[c] = 1.0 / sqrt(rc[c] * bc[c]);
    cpc[c] = m.liq_cp[k];
    cvc[c] = m.liq_cp[k] - T * ac[c] * ac[c] / (rc[c] * bc[c]);
    kc[c] = m.liq_kappa[k];
  }
  for (int k = 0; k < 2; k++) {  // gaseous H2O, CO2
    const int c = 3 + k;
    rc[c] = p / (m.gas_R[k] * T);
    bc[c] = 1.0 / p;
    ac[c] = 1.0 / T;
    sc[c] = sqrt(m.gas_gamma[k] * m.gas_R[k] * T);
    cpc[c] = m.gas_cp[k];
    cvc[c] = m.gas_cv[k];
    kc[c] = m.gas_kappa[k];
  }
  double rho = 0.0;
  for (int c = 0; c < 5; c++) rho += wf[c] / rc[c];
  rho = 1. / rho;
  rho = (1.0 - m.crystal_vf) * rho + m.crystal_vf * m.crystal_rho;
  double cp = 0.0, cv = 0.0, al = 0.0, be = 0.0, ka = 0.0, D = 0.0, snd = 0.0;
  for (int c = 0; c < 5; c++) {
    cp += wf[c] * cpc[c];
    cv += wf[c] * cvc[c];
    al += wf[c] * ac[c] * rho / rc[c];
    be += wf[c] * bc[c] * rho / rc[c];
    ka += wf[c] * kc[c];
    D += wf[c] * 1.e-10;
    snd += wf[c] / ((rc[c] * rc[c]) * (sc[c] * sc[c]));
  }
  snd = sqrt(1. / snd);
  snd /= rho;
  const double vf_g = wf[3] * rho / rc[3] + wf[4] * rho / rc[4];

  double mu = giordano_mu(m, T, wl1, wl2, o.st);
  if (m.gas_mu == 1) mu = mu / (1.0 - vf_g);
  else if (m.gas_mu == 2) mu = pal_2003(vf_g, mu, sr, o.st);
  if (m.crystal_mu == 1) mu = mu * pow(1.0 - 1.0 * m.crystal_vf, -2.5);
  else if (m.crystal_mu == 2) mu = caricchi(m.crystal_vf, mu, sr);

  o.rho = rho;
  o.mu = mu;
  o.cp = cp;
  o.cv = cv;
  o.kappa = ka;
  o.c = snd;
  o.alpha = al;
  o.beta = be;
  o.D = D;
  if (o.st) {  // finite placeholders; the caller reports the status
    o.rho = o.mu = o.cp = o.cv = o.kappa = o.c = o.beta = 1.0;
    o.alpha = 0.0;
  }
  return o;
}

}  // namespace gmc
