// gmc_devparams.hpp -- host-side translation of gmc_params into the device
// parameter block (component constants, flag tables, u_ref_inv).
#pragma once

#include <cstring>
#include <string>

#include "../../include/gales_mc.h"
#include "gmc_physics.hpp"

namespace gmc {

template <int DIM, int NC>
inline int uref_inv_nc(DevParams& dp) {
  Material<double> M;
  load_material<double, NC>(dp, M);
  return uref_inv_of<double, DIM, NC>(dp, M, dp.uref_inv);
}

// magma_mixture's constructor (mixtures.hpp:122-157) and the constant parts
// of its chemicals: melt (chemicals.hpp:63-113), dissolved volatiles
// (chemicals.hpp:136-146), ideal gases (chemicals.hpp:201-233) and the
// Giordano et al. composition parameters (GiordanoEtAlViscosityModel.hpp:36-77).
// The operations keep the reference's order (scalings applied to the tables
// first, sums left to right).  Table pointers are copied as given.
inline int magma_setup(const gmc_magma& g, MagmaDev& m, std::string& err) {
  std::memset(&m, 0, sizeof(m));
  if (g.n_p < 2 || !g.p_table || !g.wf_h2o_l || !g.wf_co2_l || g.n_T < 0 || g.n_T == 1 ||
      (g.n_T >= 2 && !g.T_table)) {
    err = "magma needs a pressure table of n_p >= 2 (and n_T = 0 or >= 2 temperatures)";
    return GMC_ERR_ARG;
  }
  if (g.gas_on_mu < 0 || g.gas_on_mu > 2 || g.crystal_on_mu < 0 || g.crystal_on_mu > 2) {
    err = "unknown magma viscosity model";
    return GMC_ERR_ARG;
  }
  double sum = 0.0;
  for (int i = 0; i < 10; i++) sum += g.oxide_wf[i];
  if (!(sum > 0.0)) {
    err = "magma oxide mass fractions must have a positive sum";
    return GMC_ERR_ARG;
  }
  double wf[10], M[10], x[10];
  const double Mg[10] = {60.084, 79.876, 101.956, 159.692, 71.846, 70.937, 40.304, 56.078, 61.979, 94.195};
  double dVdT[10] = {0.0, 7.24, 0.0, 9.09, 2.92, 0.0, 3.27, 3.74, 7.68, 12.08};
  double dVdp[10] = {-1.89, -2.31, -2.26, -2.53, -0.45, 0.0, 0.27, 0.34, -2.4, -6.75};
  double V[10] = {26.86, 23.16, 37.42, 42.13, 13.65, 0.0, 11.69, 16.53, 28.88, 45.07};
  const double cp[10] = {1331.0, 1392.0, 1545.0, 1434.0, 1100.0, 0.0, 2424.0, 1781.0, 1651.0, 1030.0};
  for (int i = 0; i < 10; i++) {
    wf[i] = g.oxide_wf[i] / sum;
    M[i] = Mg[i] * 1.e-3;
    dVdT[i] *= 1.e-9;
    dVdp[i] *= 1.e-15;
    V[i] *= 1.e-6;
  }
  double mm = 0.0;
  for (int i = 0; i < 10; i++) mm += wf[i] / M[i];
  mm = 1. / mm;
  for (int i = 0; i < 10; i++) x[i] = wf[i] * mm / M[i];
  m.melt_M = mm;
  for (int i = 0; i < 10; i++) m.melt_dVdT += x[i] * dVdT[i];
  for (int i = 0; i < 10; i++) m.melt_dVdp += x[i] * dVdp[i];
  for (int i = 0; i < 10; i++) m.melt_V += x[i] * V[i];
  for (int i = 0; i < 10; i++) m.melt_cp += x[i] * cp[i];
  m.melt_kappa = g.melt_kappa;
  m.wf_comp[1] = g.h2o_wf;
  m.wf_comp[2] = g.co2_wf;
  m.wf_comp[0] = 1.0 - m.wf_comp[1] - m.wf_comp[2];
  m.liq_M[0] = 0.01802;
  m.liq_M[1] = 0.04401;
  m.liq_cp[0] = g.h2o_l_cp;
  m.liq_cp[1] = g.co2_l_cp;
  m.liq_kappa[0] = g.h2o_l_kappa;
  m.liq_kappa[1] = g.co2_l_kappa;
  const double gM[2] = {0.01802, 0.04401};
  for (int k = 0; k < 2; k++) {
    m.gas_gamma[k] = 1.4;
    m.gas_R[k] = 8.314462 / gM[k];
    m.gas_cv[k] = m.gas_R[k] / (m.gas_gamma[k] - 1.0);
    m.gas_cp[k] = m.gas_gamma[k] * m.gas_cv[k];
    m.gas_kappa[k] = 1.e-5 * m.gas_cp[k] / 0.72;
  }
  m.crystal_vf = g.crystal_vf;
  m.crystal_rho = g.crystal_rho;
  m.gas_mu = g.gas_on_mu;
  m.crystal_mu = g.crystal_on_mu;
  // Giordano et al.: Fe total as FeO, renormalised, mole percents
  double om[10], w[10];
  for (int i = 0; i < 10; i++) {
    om[i] = M[i] * 1.e3;
    w[i] = wf[i];
  }
  w[4] = 0.9309 * (w[3] + w[4]);
  w[3] = 0.0;
  double ws = 0.0;
  for (int i = 0; i < 10; i++) ws += w[i];
  for (int i = 0; i < 10; i++) w[i] /= ws;
  double gm = 0.0;
  for (int i = 0; i < 10; i++) gm += w[i] / om[i];
  gm = 1.0 / gm;
  double mp[10];
  for (int i = 0; i < 10; i++) mp[i] = 100.0 * w[i] * gm / om[i];
  const double b[7] = {159.6, -173.3, 72.1, 75.7, -39.0, -84.1, 141.5};
  const double b1[3] = {-2.43, -0.91, 17.6};
  const double cc[6] = {2.75, 15.7, 8.3, 10.2, -12.3, -99.5};
  m.gio_M = gm;
  m.gio[0] = -4.55;
  m.gio[1] = b[0] * (mp[0] + mp[1]) + b[1] * mp[2] + b[2] * (mp[4] + mp[5]) + b[3] * mp[6] +
             b[4] * mp[7] + b[5] * mp[8];
  m.gio[2] = b1[0] * (mp[0] + mp[1]) * (mp[4] + mp[5] + mp[6]) +
             b1[1] * (mp[0] + mp[1] + mp[2]) * (mp[8] + mp[9]) + b1[2] * mp[2] * (mp[8] + mp[9]);
  m.gio[3] = b1[1] * (mp[0] + mp[1] + mp[2]);
  m.gio[4] = cc[0] * mp[0] + cc[1] * (mp[1] + mp[2]) + cc[2] * (mp[4] + mp[5] + mp[6]) + cc[3] * mp[7] +
             cc[4] * (mp[8] + mp[9]);
  m.gio[5] = 0.3 * (mp[2] + mp[4] + mp[5] + mp[6] + mp[7]) * (mp[8] + mp[9]);
  m.gio[6] = 0.3 * (mp[2] + mp[4] + mp[5] + mp[6] + mp[7]);
  m.n_p = g.n_p;
  m.n_T = g.n_T;
  m.p_tab = g.p_table;
  m.T_tab = g.n_T ? g.T_table : nullptr;
  m.wh = g.wf_h2o_l;
  m.wc = g.wf_co2_l;
  return GMC_OK;
}

inline int build_devparams(const gmc_params& p, DevParams& dp, std::string& err) {
  std::memset(&dp, 0, sizeof(dp));
  for (int c = 0; c < p.nc; c++) {
    const gmc_component& k = p.comp[c];
    dp.kind[c] = k.kind;
    dp.rho[c] = k.rho;
    dp.mu[c] = k.mu;
    dp.cp[c] = k.cp;
    dp.kappa[c] = k.kappa;
    dp.alpha[c] = k.alpha;
    dp.beta[c] = k.beta;
    dp.D[c] = k.diffusivity;
    if (k.kind == GMC_COMP_IDEAL_GAS) {
      if (!(k.molar_mass > 0.0) || !(k.gamma > 1.0)) {
        err = "ideal gas needs molar_mass > 0 and gamma > 1";
        return GMC_ERR_ARG;
      }
      dp.R[c] = 8.314462 / k.molar_mass;
      dp.gamma[c] = k.gamma;
    } else if (k.kind == GMC_COMP_MAGMA) {
      const int st = magma_setup(k.magma, dp.mg[c], err);
      if (st) return st;
      if (k.magma.gas_on_mu == GMC_MU_PAL_2003 || k.magma.crystal_on_mu == GMC_MU_CARICCHI) dp.need_sr = 1;
    } else if (k.kind != GMC_COMP_CUSTOM) {
      err = "unknown component kind";
      return GMC_ERR_ARG;
    } else if (!(k.rho > 0.0)) {
      err = "custom component needs rho > 0";
      return GMC_ERR_ARG;
    }
  }
  for (int k = 0; k < 3; k++) dp.g[k] = (k < p.dim) ? p.gravity[k] : 0.0;
  dp.dt = p.dt;
  dp.steady = p.steady_state;
  dp.tau_variant = p.tau_variant;
  dp.dc = p.dc_2006;
  dp.dc_sharp = p.dc_sharp;
  dp.dc_scale = p.dc_scale_fact;
  for (int f = 0; f < GMC_MAX_FLAG; f++) {
    dp.dmask_by_flag[f] = p.dirichlet_mask[f];
    const gmc_neumann& nb = p.neumann[f];
    dp.n_tau_mask[f] = nb.tau_mask;
    for (int t = 0; t < 6; t++) dp.n_tau[f][t] = nb.tau[t];
    dp.n_q_mask[f] = nb.q_mask;
    for (int k = 0; k < 3; k++) {
      dp.n_q[f][k] = nb.q[k];
      dp.n_J_mask[f][k] = nb.J_mask[k];
      for (int i = 0; i < GMC_MAX_COMP; i++) dp.n_J[f][k][i] = nb.J[k][i];
      dp.mflux[f][k] = nb.mflux[k];
    }
    dp.mflux_mask[f] = nb.mflux_mask & ((1u << p.dim) - 1u);
  }
  dp.p_ref = p.p_ref;
  dp.T_ref = p.T_ref;
  for (int k = 0; k < 3; k++) dp.v_ref[k] = (k < p.dim) ? p.v_ref[k] : 0.0;
  const int key = p.dim * 10 + p.nc;
  int st;
  switch (key) {
    case 22: st = uref_inv_nc<2, 2>(dp); break;
    case 23: st = uref_inv_nc<2, 3>(dp); break;
    case 24: st = uref_inv_nc<2, 4>(dp); break;
    case 32: st = uref_inv_nc<3, 2>(dp); break;
    case 33: st = uref_inv_nc<3, 3>(dp); break;
    case 34: st = uref_inv_nc<3, 4>(dp); break;
    default: err = "dim must be 2 or 3 and nc 2..4"; return GMC_ERR_ARG;
  }
  // u_ref's properties at (p_ref, T_ref) (fluid_2d.hpp:127)
  if (st & kStTable) {
    err = "p_ref / T_ref outside a magma solubility table";
    return GMC_ERR_RANGE;
  }
  if (st) {
    err = "magma viscosity model fails at (p_ref, T_ref)";
    return GMC_ERR_NUMERIC;
  }
  return GMC_OK;
}

}  // namespace gmc
