// gmc_devparams.hpp -- host-side translation of gmc_params into the device
// parameter block (component constants, flag tables, u_ref_inv).
#pragma once

#include <cstring>
#include <string>

#include "../../include/gales_mc.h"
#include "gmc_physics.hpp"

namespace gmc {

template <int NC>
inline void uref_inv_nc(const gmc_params& p, DevParams& dp) {
  Material<double> M;
  load_material<double, NC>(dp, M);
  double Y[NC];
  for (int c = 0; c < NC; c++) Y[c] = 0.5;  // Y_ref = 0.5 for every component (fluid_2d.hpp:126)
  Props<double, NC> pr;
  properties<double, NC>(dp, M, p.p_ref, p.T_ref, Y, pr);
  double u[8] = {0};
  const int dim = p.dim;
  double vv = 0.0;
  for (int k = 0; k < dim; k++) vv += p.v_ref[k] * p.v_ref[k];
  const double E = pr.cv * p.T_ref + 0.5 * vv;
  u[0] = pr.rho;
  for (int k = 0; k < dim; k++) u[1 + k] = pr.rho * p.v_ref[k];
  u[dim + 1] = pr.rho * E;
  for (int i = 0; i < NC - 1; i++) u[dim + 2 + i] = pr.rho * Y[i];
  const int n = dim + 2 + NC - 1;
  for (int i = 0; i < 8; i++) dp.uref_inv[i] = 0.0;
  for (int i = 0; i < n; i++) dp.uref_inv[i] = (u[i] < kEps) ? 0.0 : 1.0 / u[i];
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
    }
  }
  switch (p.nc) {
    case 2: uref_inv_nc<2>(p, dp); break;
    case 3: uref_inv_nc<3>(p, dp); break;
    case 4: uref_inv_nc<4>(p, dp); break;
    default: err = "nc must be 2..4"; return GMC_ERR_ARG;
  }
  return GMC_OK;
}

}  // namespace gmc
