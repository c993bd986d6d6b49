// gmc_devparams.hpp -- host-side translation of gmc_params into the device
// parameter block (component constants, flag tables, u_ref_inv).
#pragma once

#include <cstring>
#include <string>

#include "../../include/gales_mc.h"
#include "gmc_physics.hpp"

namespace gmc {

template <int DIM, int NC>
inline void uref_inv_nc(DevParams& dp) {
  Material<double> M;
  load_material<double, NC>(dp, M);
  uref_inv_of<double, DIM, NC>(dp, M, dp.uref_inv);
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
  switch (key) {
    case 22: uref_inv_nc<2, 2>(dp); break;
    case 23: uref_inv_nc<2, 3>(dp); break;
    case 24: uref_inv_nc<2, 4>(dp); break;
    case 32: uref_inv_nc<3, 2>(dp); break;
    case 33: uref_inv_nc<3, 3>(dp); break;
    case 34: uref_inv_nc<3, 4>(dp); break;
    default: err = "dim must be 2 or 3 and nc 2..4"; return GMC_ERR_ARG;
  }
  return GMC_OK;
}

}  // namespace gmc
