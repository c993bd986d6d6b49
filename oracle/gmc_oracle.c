/*
 * gmc_oracle.c -- FP64 CPU oracle of the GALES-multi-fluid fluid_mc assembly.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this file:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and only as the checker (or as the timed CPU baseline).
 *
 * It is a plain-C restatement, written from the reference's behaviour, of:
 *   - mixture properties, mix_of_ch and mix_of_mix branches
 *       src/fem/physical_props/fluid/fluid_properties.hpp:171-226,
 *       chemicals.hpp:19-49 (custom_chemical), chemicals.hpp:198-271 (ideal_gas),
 *       base_mixture.hpp:50-65, base_chemical.hpp:80-90
 *   - the magma submixture  mixtures.hpp:115-509 (melt chemicals.hpp:56-130,
 *       dissolved volatiles :136-185, Giordano et al. viscosity
 *       GiordanoEtAlViscosityModel.hpp:36-112, solubility-table interpolation
 *       mixtures.hpp:342-412, Ishii-Zuber / Pal 2003 / Einstein-Roscoe /
 *       Caricchi corrections :266-336)
 *   - P1 geometry and quadrature  src/fem/mesh/quad.hpp:148-383,470-739,817-885,
 *       src/fem/mesh/element.hpp:63-176,360-477
 *   - the element integral  sim/fluid_mc/P0/fluid_2d.hpp:171-1065
 *       (execute_i, execute_b, RM_i, RM_b, U, flux_matrices_and_vectors,
 *        tau_matrix [tau_non_diag_comp_2019], dc [dc_2006], u_ref_inv)
 *   - inverse() by LU with partial pivoting  src/fem/base/algorithm.hpp:399-414
 *       (Boost 1.70 uBLAS lu_factorize / lu_substitute, restated)
 *   - scatter with the Dirichlet avg_val rule  src/fem/ls/scatter.hpp:37-98
 *   - the assembly loop  sim/fluid_mc/P0/fluid_assembly.hpp:39-68
 *   - residual_check  src/fem/ls/residual_check.hpp:42-80
 *   - fluid_updater + dof_trimmer  fluid_updater.hpp:18-28, model.hpp:176-184
 *   - the Krylov consumer as restarted GMRES(k) + point Jacobi (right
 *     preconditioning, x0 = 0), standing in for belos_linear_solver
 *
 * The 3D element (dim = 3) is not in the reference: it is built by the same
 * construction as fluid_2d.hpp (A3, K13..K33, 3D compute_h quad.hpp:701-739,
 * 3D shear rate non_newtonian.hpp:32-38, trace(G) in the Mach correction).
 *
 * Pinning (tests/test_oracle.py, tests/test_oracle_pins.py): closed-form
 * geometry and hand property values; DUDY and A_k as central differences of
 * the closed-form U and F_adv; the element matrix as the exact Jacobian at a
 * uniform state; compute_h on right triangles and tetrahedra (closed forms);
 * every min() branch of tau 2019 forced in turn, the diagonal by hand and
 * tau = DUDY^-1 diag(t) against an exact rational solve, the low-Mach
 * correction by hand; dc_2006 nu for dc_sharp 2, 1 and 1.5 on states varying
 * along x (Z from central differences of the closed-form flux); the interior
 * and boundary diffusive fluxes (+H.J inside, -H.J on the boundary) and the
 * Neumann overrides against the closed-form stress, Fourier and species
 * fluxes; global conservation of the assembled rhs; the 3D element reducing
 * to the 2D one on z-invariant states; brute-force dense assembly; LAPACK and
 * dense solves for the inverse and GMRES.  tests/test_oracle_mutants.py
 * plants one mistake at a time into a copy of this file and checks that a
 * pin fails for each.
 *
 * What no check here can fix: the rounding-level behaviour of uBLAS and
 * Epetra (operation order), which the restatement follows line by line.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gmc_oracle.h"

#define MD GMC_MAX_DOF
#define MN 4 /* max element nodes */

static const double TOL = 2.220446049250313e-16; /* numeric_limits<double>::epsilon */

/* ------------------------------------------------------------------------ */
/* properties                                                               */
/* ------------------------------------------------------------------------ */

/* ---- magma submixture (mixtures.hpp:115-509) ---------------------------- */

/* melt: constructor chemicals.hpp:63-113 (oxide scalings applied to the
 * tables first, sums left to right), properties chemicals.hpp:121-129 */
typedef struct {
  double wf[10], M[10], molar_mass, dV_dT, dV_dp, V, cp, kappa;
} orc_melt;

static void melt_init(const double* oxide_wf, double kappa, orc_melt* m) {
  static const double Mg[10] = {60.084, 79.876, 101.956, 159.692, 71.846,
                                70.937, 40.304, 56.078,  61.979,  94.195};
  double dVdT[10] = {0.0, 7.24, 0.0, 9.09, 2.92, 0.0, 3.27, 3.74, 7.68, 12.08};
  double dVdp[10] = {-1.89, -2.31, -2.26, -2.53, -0.45, 0.0, 0.27, 0.34, -2.4, -6.75};
  double Vm[10] = {26.86, 23.16, 37.42, 42.13, 13.65, 0.0, 11.69, 16.53, 28.88, 45.07};
  static const double cpo[10] = {1331.0, 1392.0, 1545.0, 1434.0, 1100.0,
                                 0.0,    2424.0, 1781.0, 1651.0, 1030.0};
  double sum = 0.0, x[10];
  for (int i = 0; i < 10; i++) sum += oxide_wf[i];
  for (int i = 0; i < 10; i++) m->wf[i] = oxide_wf[i] / sum;
  for (int i = 0; i < 10; i++) m->M[i] = Mg[i] * 1.e-3;
  m->molar_mass = 0.0;
  for (int i = 0; i < 10; i++) m->molar_mass += m->wf[i] / m->M[i];
  m->molar_mass = 1. / m->molar_mass;
  for (int i = 0; i < 10; i++) x[i] = m->wf[i] * m->molar_mass / m->M[i];
  for (int i = 0; i < 10; i++) dVdT[i] *= 1.e-9;
  m->dV_dT = 0.0;
  for (int i = 0; i < 10; i++) m->dV_dT += x[i] * dVdT[i];
  for (int i = 0; i < 10; i++) dVdp[i] *= 1.e-15;
  m->dV_dp = 0.0;
  for (int i = 0; i < 10; i++) m->dV_dp += x[i] * dVdp[i];
  for (int i = 0; i < 10; i++) Vm[i] *= 1.e-6;
  m->V = 0.0;
  for (int i = 0; i < 10; i++) m->V += x[i] * Vm[i];
  m->cp = 0.0;
  for (int i = 0; i < 10; i++) m->cp += x[i] * cpo[i];
  m->kappa = kappa;
}

typedef struct {
  double rho, cp, cv, kappa, alpha, beta, c;
} orc_chem;

static void melt_props(const orc_melt* m, double p, double T, orc_chem* o) {
  const double mv = m->V + m->dV_dT * (T - 1673.0) + m->dV_dp * (p - 1.e5);
  o->alpha = 1. / mv * m->dV_dT;
  o->beta = -1. / mv * m->dV_dp;
  o->rho = m->molar_mass / mv;
  o->c = 1.0 / sqrt(o->rho * o->beta);
  o->cp = m->cp;
  o->cv = m->cp - T * o->alpha * o->alpha / (o->rho * o->beta);
  o->kappa = m->kappa;
}

/* dissolved_volatiles::properties (chemicals.hpp:155-170) */
static void dissolved_props(double molar_mass, double cp, double kappa, double p, double T,
                            orc_chem* o) {
  static const double a[10] = {9.144e-06, 3.685e-09,  1.168e-11, -1.990e-15, 1.220e-16,
                               -1.945e-17, -1.580e-21, 4.680e-24, 1.144e-26,  -3.960e-33};
  const double mv = a[0] + a[1] * T + a[2] * T * T + a[3] * T * T * T +
                    p * (a[4] + a[5] * T + a[6] * T * T) + p * p * (a[7] + a[8] * T) +
                    a[9] * p * p * p;
  o->rho = molar_mass / mv;
  o->alpha = a[1] + 2 * a[2] * T + 3 * a[3] * T * T + p * (a[5] + 2 * a[6] * T) + p * p * a[8];
  o->alpha /= mv;
  o->beta = -((a[4] + a[5] * T + a[6] * T * T) + 2 * p * (a[7] + a[8] * T) + 3 * a[9] * p * p);
  o->beta /= mv;
  o->c = 1.0 / sqrt(o->rho * o->beta);
  o->cp = cp;
  o->kappa = kappa;
  o->cv = cp - T * o->alpha * o->alpha / (o->rho * o->beta);
}

/* ideal_gas(name) (chemicals.hpp:201-233) and its properties (:255-262) */
static void gas_props(double molar_mass, double p, double T, orc_chem* o) {
  const double gamma = 1.4, R = 8.314462 / molar_mass;
  o->cv = R / (gamma - 1.0);
  o->cp = gamma * o->cv;
  o->kappa = 1.e-5 * o->cp / 0.72;
  o->rho = p / (R * T);
  o->beta = 1.0 / p;
  o->alpha = 1.0 / T;
  o->c = sqrt(gamma * R * T);
}

/* interpolation(p, pressure_table, y) (mixtures.hpp:342-356) */
static int interp_p(double p, const double* pt, int n, const double* y, double* out) {
  const double dp = pt[1] - pt[0];
  const double f = floor((p - pt[0]) / dp);
  if (!(f >= 0.0) || f + 1 >= (double)n) return GMC_ERR_RANGE;
  const long i = (long)f;
  const double slope = (y[i + 1] - y[i]) / dp;
  *out = y[i] + slope * (p - pt[i]);
  return 0;
}

/* interpolation(T, p, T_table, p_table, m) (mixtures.hpp:373-412), m[T][p] */
static int interp_Tp(double T, double p, const double* Tt, int nT, const double* pt, int np,
                     const double* m, double* out) {
  const double dp = fabs(pt[1] - pt[0]);
  const double fi = floor(fabs(p - pt[0]) / dp);
  const double dT = fabs(Tt[1] - Tt[0]);
  const double fj = floor(fabs(T - Tt[0]) / dT);
  if (!(fi >= 0.0) || fi + 1 >= (double)np) return GMC_ERR_RANGE;
  if (!(fj >= 0.0) || fj + 1 >= (double)nT) return GMC_ERR_RANGE;
  const long i = (long)fi, j = (long)fj;
  const double p1 = pt[i], p2 = pt[i + 1], T1 = Tt[j], T2 = Tt[j + 1];
  double mean = (T2 - T) * (p2 - p) * m[j * np + i] + (T2 - T) * (p - p1) * m[j * np + i + 1] +
                (T - T1) * (p2 - p) * m[(j + 1) * np + i] + (T - T1) * (p - p1) * m[(j + 1) * np + i + 1];
  mean /= dT * dp;
  *out = mean;
  return 0;
}

/* GiordanoEtAlViscosityModel: set() (GiordanoEtAlViscosityModel.hpp:36-77)
 * from the melt's normalised oxides, computeViscosity() (:81-109). */
static int giordano(const orc_melt* melt, double T, double wf_h2o_l, double wf_co2_l, double* mu) {
  static const double b[7] = {159.6, -173.3, 72.1, 75.7, -39.0, -84.1, 141.5};
  static const double b1[3] = {-2.43, -0.91, 17.6};
  static const double c[6] = {2.75, 15.7, 8.3, 10.2, -12.3, -99.5};
  double oxM[10], w[10], mp[10], par[7], M = 0.0, sum = 0.0;
  for (int i = 0; i < 10; i++) oxM[i] = melt->M[i] * 1.e3;
  for (int i = 0; i < 10; i++) w[i] = melt->wf[i];
  w[4] = 0.9309 * (w[3] + w[4]);
  w[3] = 0.0;
  for (int i = 0; i < 10; i++) sum += w[i];
  for (int i = 0; i < 10; i++) w[i] /= sum;
  for (int i = 0; i < 10; i++) M += w[i] / oxM[i];
  M = 1.0 / M;
  for (int i = 0; i < 10; i++) mp[i] = 100.0 * w[i] * M / oxM[i];
  par[0] = -4.55;
  par[1] = b[0] * (mp[0] + mp[1]) + b[1] * mp[2] + b[2] * (mp[4] + mp[5]) + b[3] * mp[6] +
           b[4] * mp[7] + b[5] * mp[8];
  par[2] = b1[0] * (mp[0] + mp[1]) * (mp[4] + mp[5] + mp[6]) +
           b1[1] * (mp[0] + mp[1] + mp[2]) * (mp[8] + mp[9]) + b1[2] * mp[2] * (mp[8] + mp[9]);
  par[3] = b1[1] * (mp[0] + mp[1] + mp[2]);
  par[4] = c[0] * mp[0] + c[1] * (mp[1] + mp[2]) + c[2] * (mp[4] + mp[5] + mp[6]) + c[3] * mp[7] +
           c[4] * (mp[8] + mp[9]);
  par[5] = 0.3 * (mp[2] + mp[4] + mp[5] + mp[6] + mp[7]) * (mp[8] + mp[9]);
  par[6] = 0.3 * (mp[2] + mp[4] + mp[5] + mp[6] + mp[7]);

  if (1.0 - wf_co2_l == 0.0) return GMC_ERR_NUMERIC;
  const double new_wf_h2o = wf_h2o_l / (1 - wf_co2_l);
  double x_h2o = 1 + 18.015 * (1 - new_wf_h2o) / (new_wf_h2o * M);
  x_h2o = 1.0 / x_h2o;
  const double x_rest = 1 - x_h2o;
  const double h2o_pct = 100.0 * x_h2o;
  const double A = par[0];
  const double B = par[1] * x_rest + par[2] * x_rest * x_rest + par[3] * x_rest * h2o_pct +
                   (b[5] + b[6]) * h2o_pct + b[6] * log(1 + h2o_pct);
  const double C = par[4] * x_rest + c[5] * log(1 + h2o_pct) + par[5] * x_rest * x_rest +
                   par[6] * x_rest * h2o_pct;
  if (T - C == 0.0) return GMC_ERR_NUMERIC;
  *mu = pow(10., A + B / (T - C));
  return 0;
}

/* Pal (2003) (mixtures.hpp:266-291); the loop is bounded here */
static int pal_2003(double vf_g, double mu, double shear_rate, double* out) {
  if (shear_rate < TOL) {
    *out = mu;
    return 0;
  }
  const double surface_tension = 0.36, nd = 1.e11, max_packing = 0.9999;
  const double radius = cbrt(3.0 / (4 * 3.142 * nd));
  const double Ca = radius * mu * shear_rate / surface_tension;
  const double a = 2.4 * pow(Ca, 2);
  const double Tm = pow(1.0 - a, -0.8) * pow(1.0 - vf_g / max_packing, -max_packing);
  double x = 1.e-5, F, Fp;
  int it = 0;
  do {
    F = pow(x, 5) / pow(1.0 - a * x * x, 4) - pow(Tm, 5);
    Fp = 5 * pow(x, 4) / pow(1.0 - a * x * x, 4) + 8 * a * pow(x, 6) / pow(1.0 - a * x * x, 5);
    x -= F / Fp;
    if (++it >= 200) return GMC_ERR_NUMERIC;
  } while (F < 1.e-9);
  *out = mu * x;
  return 0;
}

double orc_pal_2003(double vf_g, double mu, double shear_rate, int* st) {
  double out = mu;
  *st = pal_2003(vf_g, mu, shear_rate, &out);
  return out;
}

/* Caricchi et al. (2007) (mixtures.hpp:313-331) */
static double caricchi(double phi, double mu, double strain_rate) {
  double sr = strain_rate;
  if (sr < TOL) sr = 1.e-6;
  const double phi_max = 0.066499 * tanh(0.913424 * log10(sr) + 3.850623) + 0.591806;
  const double delta = -6.301095 * tanh(0.818496 * log10(sr) + 2.86) + 7.462405;
  const double alpha = -0.000378 * tanh(1.148101 * log10(sr) + 3.92) + 0.999572;
  const double gamma = 3.987815 * tanh(0.8908 * log10(sr) + 3.24) + 5.099645;
  const double num = 1.0 + pow(phi / phi_max, delta);
  const double x = sqrt(3.142) * phi / (2 * alpha * phi_max) * (1.0 + pow(phi / phi_max, gamma));
  const double den = pow(1.0 - alpha * erf(x), 2.5 * phi_max);
  return mu * (num / den);
}

/* magma_mixture::properties (mixtures.hpp:172-191) = submix_of_comp
 * (:197-261) + viscosity + gas and crystal corrections */
void orc_magma(const gmc_magma* g, double p, double T, double sr, orc_magma_out* o) {
  memset(o, 0, sizeof(*o));
  orc_melt melt;
  melt_init(g->oxide_wf, g->melt_kappa, &melt);
  const double wf_comp[3] = {1.0 - g->h2o_wf - g->co2_wf, g->h2o_wf, g->co2_wf};
  double wl[3];
  int st;
  if (g->n_T == 0) {
    st = interp_p(p, g->p_table, g->n_p, g->wf_h2o_l, &wl[1]);
    if (!st) st = interp_p(p, g->p_table, g->n_p, g->wf_co2_l, &wl[2]);
  } else {
    st = interp_Tp(T, p, g->T_table, g->n_T, g->p_table, g->n_p, g->wf_h2o_l, &wl[1]);
    if (!st) st = interp_Tp(T, p, g->T_table, g->n_T, g->p_table, g->n_p, g->wf_co2_l, &wl[2]);
  }
  if (st) {
    o->st = st;
    return;
  }
  wl[0] = 1.0 - wl[1] - wl[2];
  double wf[5];
  wf[0] = wf_comp[0];
  const double wf_liq = wf[0] / wl[0];
  wf[1] = wl[1] * wf_liq;
  wf[2] = wl[2] * wf_liq;
  wf[3] = wf_comp[1] - wf[1];
  wf[4] = wf_comp[2] - wf[2];
  orc_chem ch[5];
  melt_props(&melt, p, T, &ch[0]);
  dissolved_props(0.01802, g->h2o_l_cp, g->h2o_l_kappa, p, T, &ch[1]);
  dissolved_props(0.04401, g->co2_l_cp, g->co2_l_kappa, p, T, &ch[2]);
  gas_props(0.01802, p, T, &ch[3]);
  gas_props(0.04401, p, T, &ch[4]);
  double rho = 0.0;
  for (int c = 0; c < 5; c++) rho += wf[c] / ch[c].rho;
  rho = 1. / rho;
  rho = (1.0 - g->crystal_vf) * rho + g->crystal_vf * g->crystal_rho;
  double snd = 0.0;
  for (int c = 0; c < 5; c++) {
    o->cp += wf[c] * ch[c].cp;
    o->cv += wf[c] * ch[c].cv;
    o->alpha += wf[c] * ch[c].alpha * rho / ch[c].rho;
    o->beta += wf[c] * ch[c].beta * rho / ch[c].rho;
    o->kappa += wf[c] * ch[c].kappa;
    o->D += wf[c] * 1.e-10;
    snd += wf[c] / (pow(ch[c].rho, 2) * pow(ch[c].c, 2));
  }
  snd = sqrt(1. / snd);
  snd /= rho;
  for (int c = 0; c < 5; c++) {
    o->vf[c] = wf[c] * rho / ch[c].rho;
    o->wf[c] = wf[c];
    o->rho_ch[c] = ch[c].rho;
    o->alpha_ch[c] = ch[c].alpha;
    o->beta_ch[c] = ch[c].beta;
    o->c_ch[c] = ch[c].c;
    o->cp_ch[c] = ch[c].cp;
    o->cv_ch[c] = ch[c].cv;
  }
  o->vf_gas = o->vf[3] + o->vf[4];
  o->wl_h2o = wl[1];
  o->wl_co2 = wl[2];
  o->rho = rho;
  o->c = snd;
  double mu;
  st = giordano(&melt, T, wl[1], wl[2], &mu);
  if (st) {
    o->st = st;
    return;
  }
  o->mu_melt = mu;
  if (g->gas_on_mu == GMC_MU_ISHII_ZUBER) mu = mu / (1.0 - o->vf_gas);
  else if (g->gas_on_mu == GMC_MU_PAL_2003) {
    st = pal_2003(o->vf_gas, mu, sr, &mu);
    if (st) {
      o->st = st;
      return;
    }
  }
  if (g->crystal_on_mu == GMC_MU_EINSTEIN_ROSCOE) mu = mu * pow(1.0 - 1.0 * g->crystal_vf, -2.5);
  else if (g->crystal_on_mu == GMC_MU_CARICCHI) mu = caricchi(g->crystal_vf, mu, sr);
  o->mu = mu;
}

/* One component at (p, T, strain rate): chemicals.hpp:23-47 (custom),
 * 237-264 (ideal gas), a magma submixture (mixtures.hpp:172-261).  Returns
 * 0 or the magma's status. */
static int chem_props(const gmc_component* c, double p, double T, double sr, double* rho,
                      double* mu, double* cp, double* cv, double* kappa, double* alpha,
                      double* beta, double* sound, double* D) {
  if (c->kind == GMC_COMP_MAGMA) {
    orc_magma_out m;
    orc_magma(&c->magma, p, T, sr, &m);
    if (m.st) {  /* finite placeholders; the caller reports the status */
      *rho = *mu = *cp = *cv = *kappa = *beta = *sound = 1.0;
      *alpha = 0.0;
      *D = 1.e-10;
      return m.st;
    }
    *rho = m.rho;
    *mu = m.mu;
    *cp = m.cp;
    *cv = m.cv;
    *kappa = m.kappa;
    *alpha = m.alpha;
    *beta = m.beta;
    *sound = m.c;
    *D = m.D;
    return 0;
  }
  if (c->kind == GMC_COMP_CUSTOM) {
    *rho = c->rho;
    *mu = c->mu;
    *cp = c->cp;
    *alpha = c->alpha;
    *beta = c->beta;
    *kappa = c->kappa;
    if (c->beta == 0.0) *sound = 1.e20;
    else *sound = 1.0 / sqrt(c->rho * c->beta);
    if (c->beta == 0.0) *cv = c->cp;
    else *cv = c->cp - T * c->alpha * c->alpha / (c->rho * c->beta);
  } else {
    const double R = 8.314462 / c->molar_mass;
    *cv = R / (c->gamma - 1.0);
    *cp = c->gamma * (*cv);
    *mu = c->mu;
    *kappa = (*mu) * (*cp) / 0.72;
    *rho = p / (R * T);
    *beta = 1.0 / p;
    *alpha = 1.0 / T;
    *sound = sqrt(c->gamma * R * T);
  }
  *D = c->diffusivity;
  return 0;
}

/* mix_of_ch (fluid_properties.hpp:199-226) and mix_of_mix (:171-197) share
 * these formulas; a magma component's D is its submixture's
 * chemical_diffusivity_ */
void orc_properties(const gmc_params* P, double p, double T, const double* Y, double sr,
                    orc_props* o) {
  const int nc = P->nc;
  double rc[GMC_MAX_COMP];
  double ac[GMC_MAX_COMP], bc[GMC_MAX_COMP];
  memset(o, 0, sizeof(*o));
  for (int c = 0; c < nc; c++) {
    double rho, mu, cp, cv, kappa, alpha, beta, sound, D;
    const int cs = chem_props(&P->comp[c], p, T, sr, &rho, &mu, &cp, &cv, &kappa, &alpha, &beta,
                              &sound, &D);
    if (cs && !o->st) o->st = cs;
    o->rho += Y[c] / rho;
    o->mu += Y[c] * mu;
    o->cp += Y[c] * cp;
    o->cv += Y[c] * cv;
    o->kappa += Y[c] * kappa;
    o->D += Y[c] * D;
    o->c += Y[c] / (pow(rho, 2) * pow(sound, 2));
    o->rho_comp[c] = rho;
    o->e_comp[c] = cv * T;
    o->D_comp[c] = D;
    rc[c] = rho;
    ac[c] = alpha;
    bc[c] = beta;
  }
  o->rho = 1. / o->rho;
  o->c = sqrt(1.0 / o->c);
  o->c /= o->rho;
  for (int c = 0; c < nc; c++) {
    o->alpha += Y[c] * ac[c] * o->rho / rc[c];
    o->beta += Y[c] * bc[c] * o->rho / rc[c];
  }
}

/* ------------------------------------------------------------------------ */
/* geometry                                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
  int nen;
  double sh[MN];
  double dsh[3][MN]; /* d/dx, d/dy, d/dz */
  double J;
  double G[3][3];
  double JN[3];
} orc_geo;

/* local shape derivatives of P1 triangle / tet (quad.hpp:156-168, 496-512) */
static void ref_shape(int dim, const double* xi, double* sh, double (*dref)[MN]) {
  if (dim == 2) {
    sh[0] = 1.0 - xi[0] - xi[1];
    sh[1] = xi[0];
    sh[2] = xi[1];
    dref[0][0] = -1.0; dref[1][0] = -1.0;
    dref[0][1] = 1.0;  dref[1][1] = 0.0;
    dref[0][2] = 0.0;  dref[1][2] = 1.0;
  } else {
    sh[0] = 1.0 - xi[0] - xi[1] - xi[2];
    sh[1] = xi[0];
    sh[2] = xi[1];
    sh[3] = xi[2];
    for (int k = 0; k < 3; k++)
      for (int a = 0; a < 4; a++) dref[k][a] = 0.0;
    dref[0][0] = dref[1][0] = dref[2][0] = -1.0;
    dref[0][1] = 1.0;
    dref[1][2] = 1.0;
    dref[2][3] = 1.0;
  }
}

/* quad::calc + g_up_g_G (+ JN for boundary points); returns 0 or GMC_ERR_MESH */
static int geo_at(int dim, const double* xe, const double* xi, int boundary, orc_geo* g) {
  double dref[3][MN];
  const int nen = dim + 1;
  g->nen = nen;
  ref_shape(dim, xi, g->sh, dref);
  if (dim == 2) {
    double dx_dxi = 0, dx_deta = 0, dy_dxi = 0, dy_deta = 0;
    for (int i = 0; i < nen; i++) {
      dx_dxi += dref[0][i] * xe[2 * i];
      dx_deta += dref[1][i] * xe[2 * i];
      dy_dxi += dref[0][i] * xe[2 * i + 1];
      dy_deta += dref[1][i] * xe[2 * i + 1];
    }
    const double J = dx_dxi * dy_deta - dy_dxi * dx_deta;
    if (J <= 0.0) return GMC_ERR_MESH;
    g->J = J;
    const double dxi_dx = (1 / J) * (dy_deta);
    const double deta_dx = (1 / J) * (-dy_dxi);
    const double dxi_dy = (1 / J) * (-dx_deta);
    const double deta_dy = (1 / J) * (dx_dxi);
    for (int i = 0; i < nen; i++) {
      g->dsh[0][i] = dref[0][i] * dxi_dx + dref[1][i] * deta_dx;
      g->dsh[1][i] = dref[0][i] * dxi_dy + dref[1][i] * deta_dy;
    }
    g->G[0][0] = dxi_dx * dxi_dx + deta_dx * deta_dx;
    g->G[1][1] = dxi_dy * dxi_dy + deta_dy * deta_dy;
    g->G[0][1] = dxi_dx * dxi_dy + deta_dx * deta_dy;
    g->G[1][0] = g->G[0][1];
    if (boundary) { /* quad::JN 2D triangle, quad.hpp:271-291 */
      if (xi[0] == 0.0) {
        g->JN[0] = -dy_deta;
        g->JN[1] = dx_deta;
      } else if (xi[1] == 0.0) {
        g->JN[0] = dy_dxi;
        g->JN[1] = -dx_dxi;
      } else {
        g->JN[0] = dy_deta - dy_dxi;
        g->JN[1] = dx_dxi - dx_deta;
      }
    }
  } else {
    double d[3][3] = {{0}}; /* d[c][k] = d x_c / d xi_k */
    for (int i = 0; i < nen; i++)
      for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) d[c][k] += dref[k][i] * xe[3 * i + c];
    const double dx_dxi = d[0][0], dx_deta = d[0][1], dx_dzeta = d[0][2];
    const double dy_dxi = d[1][0], dy_deta = d[1][1], dy_dzeta = d[1][2];
    const double dz_dxi = d[2][0], dz_deta = d[2][1], dz_dzeta = d[2][2];
    const double J = dx_dxi * (dy_deta * dz_dzeta - dz_deta * dy_dzeta) -
                     dy_dxi * (dx_deta * dz_dzeta - dz_deta * dx_dzeta) +
                     dz_dxi * (dx_deta * dy_dzeta - dy_deta * dx_dzeta);
    if (J <= 0.0) return GMC_ERR_MESH;
    g->J = J;
    double inv[3][3]; /* inv[k][c] = d xi_k / d x_c */
    inv[0][0] = (1 / J) * (dy_deta * dz_dzeta - dz_deta * dy_dzeta);
    inv[1][0] = (1 / J) * (-dy_dxi * dz_dzeta + dz_dxi * dy_dzeta);
    inv[2][0] = (1 / J) * (dy_dxi * dz_deta - dz_dxi * dy_deta);
    inv[0][1] = (1 / J) * (-dx_deta * dz_dzeta + dz_deta * dx_dzeta);
    inv[1][1] = (1 / J) * (dx_dxi * dz_dzeta - dz_dxi * dx_dzeta);
    inv[2][1] = (1 / J) * (-dx_dxi * dz_deta + dz_dxi * dx_deta);
    inv[0][2] = (1 / J) * (dx_deta * dy_dzeta - dy_deta * dx_dzeta);
    inv[1][2] = (1 / J) * (-dx_dxi * dy_dzeta + dy_dxi * dx_dzeta);
    inv[2][2] = (1 / J) * (dx_dxi * dy_deta - dy_dxi * dx_deta);
    for (int i = 0; i < nen; i++)
      for (int c = 0; c < 3; c++)
        g->dsh[c][i] = dref[0][i] * inv[0][c] + dref[1][i] * inv[1][c] + dref[2][i] * inv[2][c];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        g->G[a][b] = inv[0][a] * inv[0][b] + inv[1][a] * inv[1][b] + inv[2][a] * inv[2][b];
    if (boundary) { /* quad::JN 3D tet, quad.hpp:633-664 */
      if (xi[0] == 0.0) {
        g->JN[0] = dz_deta * dy_dzeta - dy_deta * dz_dzeta;
        g->JN[1] = dx_deta * dz_dzeta - dz_deta * dx_dzeta;
        g->JN[2] = dy_deta * dx_dzeta - dx_deta * dy_dzeta;
      } else if (xi[1] == 0.0) {
        g->JN[0] = dz_dzeta * dy_dxi - dy_dzeta * dz_dxi;
        g->JN[1] = dz_dxi * dx_dzeta - dx_dxi * dz_dzeta;
        g->JN[2] = dy_dzeta * dx_dxi - dx_dzeta * dy_dxi;
      } else if (xi[2] == 0.0) {
        g->JN[0] = dz_dxi * dy_deta - dy_dxi * dz_deta;
        g->JN[1] = dx_dxi * dz_deta - dx_deta * dz_dxi;
        g->JN[2] = dy_dxi * dx_deta - dx_dxi * dy_deta;
      } else {
        g->JN[0] = (dz_dzeta - dz_dxi) * (dy_deta - dy_dxi) - (dy_dzeta - dy_dxi) * (dz_deta - dz_dxi);
        g->JN[1] = (dx_dzeta - dx_dxi) * (dz_deta - dz_dxi) - (dx_deta - dx_dxi) * (dz_dzeta - dz_dxi);
        g->JN[2] = (dx_deta - dx_dxi) * (dy_dzeta - dy_dxi) - (dx_dzeta - dx_dxi) * (dy_deta - dy_dxi);
      }
    }
  }
  return 0;
}

/* interior rules (element.hpp:133-136, 446-450) */
int orc_n_qp(int dim) { return dim == 2 ? 3 : 4; }
void orc_qp(int dim, int q, double* xi) {
  if (dim == 2) {
    const double a = 1.0 / 6.0, b = 4.0 / 6.0;
    xi[0] = (q == 1) ? b : a;
    xi[1] = (q == 2) ? b : a;
  } else {
    const double a = 0.1381966011250105, b = 0.5854101966249685;
    xi[0] = (q == 1) ? b : a;
    xi[1] = (q == 2) ? b : a;
    xi[2] = (q == 3) ? b : a;
  }
}
double orc_w_in(int dim) { return dim == 2 ? 1.0 / 6.0 : 1.0 / 24.0; }
double orc_w_bd(int dim) { return dim == 2 ? 1.0 / 2.0 : 1.0 / 6.0; }
int orc_n_side_qp(int dim) { return dim == 2 ? 2 : 3; }

/* side rules (element.hpp:157-164, 461-466) */
void orc_side_qp(int dim, int side, int q, double* xi) {
  if (dim == 2) {
    const double gm = 0.5 * (1.0 - 1.0 / sqrt(3.0));
    const double gp = 0.5 * (1.0 + 1.0 / sqrt(3.0));
    const double t = (q == 0) ? gp : gm;
    if (side == 0) {
      xi[0] = t;
      xi[1] = 1.0 - t;
    } else if (side == 1) {
      xi[0] = 0.0;
      xi[1] = t;
    } else {
      xi[0] = (q == 0) ? gm : gp;
      xi[1] = 0.0;
    }
  } else {
    const double a = 1.0 / 6.0, b = 4.0 / 6.0;
    static const int pat[4][3][3] = {
        {{2, 1, 1}, {1, 2, 1}, {1, 1, 2}}, /* xi+eta+zeta = 1 */
        {{0, 1, 1}, {0, 1, 2}, {0, 2, 1}}, /* xi = 0 */
        {{1, 0, 1}, {2, 0, 1}, {1, 0, 2}}, /* eta = 0 */
        {{1, 1, 0}, {1, 2, 0}, {2, 1, 0}}, /* zeta = 0 */
    };
    for (int k = 0; k < 3; k++) {
      const int c = pat[side][q][k];
      xi[k] = (c == 0) ? 0.0 : ((c == 1) ? a : b);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* dense helpers and uBLAS-style LU inverse                                 */
/* ------------------------------------------------------------------------ */

/* lu_factorize + lu_substitute(identity), algorithm.hpp:399-414.  Partial
 * pivoting on the first entry of largest magnitude; the column below the
 * pivot is multiplied by 1/pivot; substitution is column-wise axpy. */
int orc_inverse(int n, const double* A_in, double* Inv) {
  double A[MD * MD];
  int pm[MD];
  memcpy(A, A_in, sizeof(double) * n * n);
  int singular = 0;
  for (int i = 0; i < n; i++) {
    int imax = i;
    double vmax = fabs(A[i * n + i]);
    for (int r = i + 1; r < n; r++)
      if (fabs(A[r * n + i]) > vmax) {
        vmax = fabs(A[r * n + i]);
        imax = r;
      }
    pm[i] = imax;
    if (A[imax * n + i] != 0.0) {
      if (imax != i)
        for (int c = 0; c < n; c++) {
          double t = A[i * n + c];
          A[i * n + c] = A[imax * n + c];
          A[imax * n + c] = t;
        }
      const double rp = 1.0 / A[i * n + i];
      for (int r = i + 1; r < n; r++) A[r * n + i] *= rp;
    } else if (!singular) {
      singular = i + 1;
    }
    for (int r = i + 1; r < n; r++)
      for (int c = i + 1; c < n; c++) A[r * n + c] -= A[r * n + i] * A[i * n + c];
  }
  if (singular) return GMC_ERR_NUMERIC;
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) Inv[r * n + c] = (r == c) ? 1.0 : 0.0;
  for (int i = 0; i < n; i++)
    if (pm[i] != i)
      for (int c = 0; c < n; c++) {
        double t = Inv[i * n + c];
        Inv[i * n + c] = Inv[pm[i] * n + c];
        Inv[pm[i] * n + c] = t;
      }
  for (int k = 0; k < n; k++)
    for (int l = 0; l < n; l++) {
      const double t = Inv[k * n + l];
      if (t != 0.0)
        for (int m = k + 1; m < n; m++) Inv[m * n + l] -= A[m * n + k] * t;
    }
  for (int k = n - 1; k >= 0; k--)
    for (int l = n - 1; l >= 0; l--) {
      const double t = Inv[k * n + l] /= A[k * n + k];
      if (t != 0.0)
        for (int m = k - 1; m >= 0; m--) Inv[m * n + l] -= A[m * n + k] * t;
    }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* per quadrature point: flux matrices and vectors                           */
/* ------------------------------------------------------------------------ */

typedef struct {
  int dim, nc, ndof;
  double p, v[3], T, Y[GMC_MAX_COMP];
  orc_props pr;
  double U[MD];
  double DUDY[MD][MD];
  double A[3][MD][MD];
  double K[3][3][MD][MD];
  double S0[MD][MD];
  double Fadv[3][MD], Fdif[3][MD], S[MD];
  double dI[3][MD];
} orc_qp_t;

/* fluid::assign (fluid_2d.hpp:109-119) */
static void assign(int dim, int nc, const double* v, double* p, double* vel, double* T,
                   double* Y) {
  *p = v[0];
  for (int k = 0; k < dim; k++) vel[k] = v[1 + k];
  *T = v[dim + 1];
  for (int i = 0; i < nc - 1; i++) Y[i] = v[dim + 2 + i];
  double s = 0.0;
  for (int i = 0; i < nc - 1; i++) s += Y[i];
  Y[nc - 1] = 1.0 - s;
}

/* fluid::U (fluid_2d.hpp:372-385) */
static void U_vec(int dim, int nc, double rho, double cv, const double* vel, double T,
                  const double* Y, double* u) {
  double vv = 0.0;
  for (int k = 0; k < dim; k++) vv += vel[k] * vel[k];
  const double E = cv * T + 0.5 * vv;
  u[0] = rho;
  for (int k = 0; k < dim; k++) u[1 + k] = rho * vel[k];
  u[dim + 1] = rho * E;
  for (int i = 0; i < nc - 1; i++) u[dim + 2 + i] = rho * Y[i];
}

/* compute_shear_rate (non_newtonian.hpp:23-38) */
static double shear_rate(int dim, double (*dI)[MD]) {
  if (dim == 2) {
    const double d11 = dI[0][1], d21 = dI[0][2], d12 = dI[1][1], d22 = dI[1][2];
    return sqrt(pow(d11, 2) + pow(d22, 2) + 2 * pow(d12 + d21, 2));
  }
  const double d11 = dI[0][1], d21 = dI[0][2], d31 = dI[0][3];
  const double d12 = dI[1][1], d22 = dI[1][2], d32 = dI[1][3];
  const double d13 = dI[2][1], d23 = dI[2][2], d33 = dI[2][3];
  return sqrt(pow(d11, 2) + pow(d22, 2) + pow(d33, 2) + 2 * pow(d12 + d21, 2) +
              2 * pow(d13 + d31, 2) + 2 * pow(d23 + d32, 2));
}

/* flux_matrices_and_vectors, interior form (fluid_2d.hpp:391-603), dim-generic */
static void flux_interior(const gmc_params* P, orc_qp_t* q) {
  const int dim = q->dim, nc = q->nc, n = q->ndof;
  const int iT = dim + 1, iE = dim + 1, iY = dim + 2;
  const orc_props* pr = &q->pr;
  const double rho = pr->rho, p = q->p, T = q->T;
  const double* v = q->v;
  const double* Y = q->Y;
  double vv = 0.0;
  for (int k = 0; k < dim; k++) vv += v[k] * v[k];
  const double E = pr->cv * T + 0.5 * vv;
  const double drho_dp = rho * pr->beta;
  const double drho_dT = -rho * pr->alpha;
  const double de_dp = (pr->beta * p - pr->alpha * T) / rho;
  const double de_dT = pr->cp - p * pr->alpha / rho;
  double drho_dY[GMC_MAX_COMP], dE_dY[GMC_MAX_COMP];
  for (int i = 0; i < nc - 1; i++) {
    drho_dY[i] = -rho * rho * (1. / pr->rho_comp[i] - 1 / pr->rho_comp[nc - 1]);
    dE_dY[i] = pr->e_comp[i] - pr->e_comp[nc - 1];
  }
  memset(q->DUDY, 0, sizeof(q->DUDY));
  memset(q->A, 0, sizeof(q->A));
  memset(q->K, 0, sizeof(q->K));
  memset(q->S0, 0, sizeof(q->S0));

  /* DUDY (A0), fluid_2d.hpp:407-441 */
  double(*D)[MD] = q->DUDY;
  D[0][0] = drho_dp;
  D[0][iT] = drho_dT;
  for (int c = 0; c < nc - 1; c++) D[0][iY + c] = drho_dY[c];
  for (int k = 0; k < dim; k++) {
    D[1 + k][0] = drho_dp * v[k];
    D[1 + k][1 + k] = rho;
    D[1 + k][iT] = drho_dT * v[k];
    for (int c = 0; c < nc - 1; c++) D[1 + k][iY + c] = v[k] * drho_dY[c];
  }
  D[iE][0] = drho_dp * E + rho * de_dp;
  for (int k = 0; k < dim; k++) D[iE][1 + k] = rho * v[k];
  D[iE][iT] = drho_dT * E + rho * de_dT;
  for (int c = 0; c < nc - 1; c++) D[iE][iY + c] = E * drho_dY[c] + rho * dE_dY[c];
  for (int i = 0; i < nc - 1; i++) {
    D[iY + i][0] = drho_dp * Y[i];
    D[iY + i][iT] = drho_dT * Y[i];
    for (int c = 0; c < nc - 1; c++) D[iY + i][iY + c] = Y[i] * drho_dY[c];
  }
  for (int i = 0; i < nc - 1; i++) D[iY + i][iY + i] += rho;

  /* A_k, fluid_2d.hpp:445-524 */
  for (int k = 0; k < dim; k++) {
    double(*A)[MD] = q->A[k];
    const double vk = v[k];
    A[0][0] = drho_dp * vk;
    A[0][1 + k] = rho;
    A[0][iT] = drho_dT * vk;
    for (int c = 0; c < nc - 1; c++) A[0][iY + c] = drho_dY[c] * vk;
    for (int j = 0; j < dim; j++) {
      const int r = 1 + j;
      if (j == k) {
        A[r][0] = drho_dp * vk * vk + 1.0;
        A[r][1 + k] = 2.0 * rho * vk;
        A[r][iT] = drho_dT * vk * vk;
        for (int c = 0; c < nc - 1; c++) A[r][iY + c] = drho_dY[c] * vk * vk;
      } else {
        A[r][0] = drho_dp * v[j] * vk;
        A[r][1 + k] = rho * v[j];
        A[r][1 + j] = rho * vk;
        A[r][iT] = drho_dT * v[j] * vk;
        for (int c = 0; c < nc - 1; c++) A[r][iY + c] = drho_dY[c] * v[j] * vk;
      }
    }
    A[iE][0] = vk + vk * (drho_dp * E + rho * de_dp);
    for (int j = 0; j < dim; j++) {
      if (j == k) A[iE][1 + j] = p + rho * E + rho * vk * vk;
      else {
        const int lo = j < k ? j : k, hi = j < k ? k : j;
        A[iE][1 + j] = rho * v[lo] * v[hi];
      }
    }
    A[iE][iT] = vk * (drho_dT * E + rho * de_dT);
    for (int c = 0; c < nc - 1; c++) A[iE][iY + c] = (E * drho_dY[c] + rho * dE_dY[c]) * vk;
    for (int i = 0; i < nc - 1; i++) {
      A[iY + i][0] = drho_dp * Y[i] * vk;
      A[iY + i][1 + k] = rho * Y[i];
      A[iY + i][iT] = drho_dT * Y[i] * vk;
      for (int c = 0; c < nc - 1; c++) A[iY + i][iY + c] = Y[i] * drho_dY[c] * vk;
    }
    for (int i = 0; i < nc - 1; i++) A[iY + i][iY + i] += rho * vk;
  }

  /* K_kl, fluid_2d.hpp:531-571 */
  double H[GMC_MAX_COMP];
  for (int i = 0; i < nc; i++) H[i] = pr->e_comp[i] + p / pr->rho_comp[i];
  const double mu = pr->mu;
  const double lambda = -2.0 * mu / 3.0;
  for (int k = 0; k < dim; k++) {
    double(*Kk)[MD] = q->K[k][k];
    for (int j = 0; j < dim; j++) {
      if (j == k) {
        Kk[1 + j][1 + j] = lambda + 2 * mu;
        Kk[iE][1 + j] = (lambda + 2 * mu) * v[j];
      } else {
        Kk[1 + j][1 + j] = mu;
        Kk[iE][1 + j] = mu * v[j];
      }
    }
    Kk[iE][iT] = pr->kappa;
    for (int i = 0; i < nc - 1; i++) {
      Kk[iY + i][iY + i] = rho * pr->D_comp[i];
      Kk[iE][iY + i] = rho * pr->D_comp[i] * H[i];
    }
    for (int l = 0; l < dim; l++) {
      if (l == k) continue;
      double(*Kl)[MD] = q->K[k][l];
      Kl[1 + k][1 + l] = lambda;
      Kl[1 + l][1 + k] = mu;
      Kl[iE][1 + k] = mu * v[l];
      Kl[iE][1 + l] = lambda * v[k];
    }
  }

  /* S0, U, F_adv, S, F_dif  (fluid_2d.hpp:573-602) */
  for (int k = 0; k < dim; k++) q->S0[iE][1 + k] = rho * P->gravity[k];
  U_vec(dim, nc, rho, pr->cv, v, T, Y, q->U);
  for (int k = 0; k < dim; k++) {
    double* F = q->Fadv[k];
    F[0] = rho * v[k];
    for (int j = 0; j < dim; j++) F[1 + j] = rho * v[k] * v[j] + ((j == k) ? p : 0.0);
    F[iE] = v[k] * (rho * E + p);
    for (int i = 0; i < nc - 1; i++) F[iY + i] = rho * v[k] * Y[i];
  }
  for (int i = 0; i < n; i++) q->S[i] = 0.0;
  double gv = 0.0;
  for (int k = 0; k < dim; k++) {
    q->S[1 + k] = rho * P->gravity[k];
    gv += P->gravity[k] * v[k];
  }
  q->S[iE] = rho * gv;
  for (int k = 0; k < dim; k++)
    for (int i = 0; i < n; i++) {
      double acc = 0.0;
      for (int l = 0; l < dim; l++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += q->K[k][l][i][j] * q->dI[l][j];
        acc = (l == 0) ? s : acc + s;
      }
      q->Fdif[k][i] = acc;
    }
}

/* boundary overrides of flux_matrices_and_vectors (fluid_2d.hpp:610-868) */
static void flux_boundary(const gmc_params* P, int side_flag, orc_qp_t* q) {
  const int dim = q->dim, nc = q->nc;
  const int iT = dim + 1, iE = dim + 1, iY = dim + 2;
  const gmc_neumann* nb = &P->neumann[side_flag];
  const orc_props* pr = &q->pr;
  const double mu = pr->mu, rho = pr->rho, p = q->p;
  const double* v = q->v;
  double(*dI)[MD] = q->dI;
  /* prescribed mass flux m of direction k (fluid_2d.hpp:623-758): F_k^adv and
   * A_k are replaced; the U and DUDY it also rewrites are not used by RM_b */
  if (nb->mflux_mask) {
    double vv = 0.0;
    for (int k = 0; k < dim; k++) vv += v[k] * v[k];
    const double E = pr->cv * q->T + 0.5 * vv;
    const double de_dp = (pr->beta * p - pr->alpha * q->T) / rho;
    const double de_dT = pr->cp - p * pr->alpha / rho;
    double dE_dY[GMC_MAX_COMP];
    for (int i = 0; i < nc - 1; i++) dE_dY[i] = pr->e_comp[i] - pr->e_comp[nc - 1];
    for (int k = 0; k < dim; k++) {
      if (!(nb->mflux_mask & (1u << k))) continue;
      const double m = nb->mflux[k];
      double(*A)[MD] = q->A[k];
      for (int i = 0; i < q->ndof; i++)
        for (int j = 0; j < q->ndof; j++) A[i][j] = 0.0;
      q->Fadv[k][0] = m;
      for (int j = 0; j < dim; j++) q->Fadv[k][1 + j] = (j == k) ? m * v[j] + p : m * v[j];
      q->Fadv[k][iE] = m * E + p * v[k];
      for (int i = 0; i < nc - 1; i++) q->Fadv[k][iY + i] = m * q->Y[i];
      A[1 + k][0] = 1.0;
      for (int j = 0; j < dim; j++) A[1 + j][1 + j] = m;
      A[iE][0] = v[k] + m * de_dp;
      for (int j = 0; j < dim; j++) A[iE][1 + j] = (j == k) ? p + m * v[k] : m * v[j];
      A[iE][iT] = m * de_dT;
      for (int c = 0; c < nc - 1; c++) A[iE][iY + c] = m * dE_dY[c];
      for (int i = 0; i < nc - 1; i++) A[iY + i][iY + i] = m;
    }
  }
  const double lambda = -2.0 * mu / 3.0;
  double div = 0.0;
  for (int k = 0; k < dim; k++) div += dI[k][1 + k];
  double tau[3][3];
  for (int a = 0; a < dim; a++)
    for (int b = 0; b < dim; b++) {
      if (a == b) tau[a][a] = lambda * div + 2.0 * mu * dI[a][1 + a];
      else tau[a][b] = mu * (dI[b][1 + a] + dI[a][1 + b]);
    }
  double qf[3], J[3][GMC_MAX_COMP];
  for (int k = 0; k < dim; k++) {
    qf[k] = -pr->kappa * dI[k][iT];
    for (int i = 0; i < nc - 1; i++) J[k][i] = rho * pr->D_comp[i] * dI[k][iY + i];
  }
  /* tau index order 11,22,33,12,13,23 */
  static const int ta[6] = {0, 1, 2, 0, 0, 1}, tb[6] = {0, 1, 2, 1, 2, 2};
  for (int t = 0; t < 6; t++) {
    if (ta[t] >= dim || tb[t] >= dim) continue;
    if (nb->tau_mask & (1u << t)) {
      tau[ta[t]][tb[t]] = nb->tau[t];
      tau[tb[t]][ta[t]] = nb->tau[t];
    }
  }
  for (int k = 0; k < dim; k++) {
    if (nb->q_mask & (1u << k)) qf[k] = nb->q[k];
    for (int i = 0; i < nc - 1; i++)
      if (nb->J_mask[k] & (1u << i)) J[k][i] = nb->J[k][i];
  }
  double H[GMC_MAX_COMP];
  for (int i = 0; i < nc - 1; i++) H[i] = pr->e_comp[i] + p / pr->rho_comp[i];
  for (int k = 0; k < dim; k++) {
    double HJ = 0.0;
    for (int i = 0; i < nc - 1; i++) HJ += H[i] * J[k][i];
    double e = 0.0;
    for (int a = 0; a < dim; a++) {
      q->Fdif[k][1 + a] = tau[a][k];
      e = (a == 0) ? tau[k][a] * v[a] : e + tau[k][a] * v[a];
    }
    q->Fdif[k][iE] = e - qf[k] - HJ;
    for (int i = 0; i < nc - 1; i++) q->Fdif[k][iY + i] = J[k][i];
  }
  /* K entries of Neumann-fixed fluxes are zeroed (fluid_2d.hpp:820-867).
   * Overriding tau_ab removes the row v_b of every K_al and the row v_a of
   * every K_bl (the rows that produce F_a[v_b] and F_b[v_a]), together with
   * the energy-row entries those rows feed.  Structural nonzeros of row v_b of
   * K_al: (v_b, v_b) if a == l; (v_a, v_l) and (v_l, v_a) if a != l. */
  for (int t = 0; t < 6; t++) {
    if (ta[t] >= dim || tb[t] >= dim) continue;
    if (!(nb->tau_mask & (1u << t))) continue;
    for (int s = 0; s < 2; s++) {
      const int a = s ? tb[t] : ta[t], b = s ? ta[t] : tb[t];
      if (s == 1 && a == b) break;
      for (int l = 0; l < dim; l++) {
        double(*Kal)[MD] = q->K[a][l];
        if (a == l) {
          Kal[1 + b][1 + b] = 0.0;
          Kal[iE][1 + b] = 0.0;
        } else if (b == a) {
          Kal[1 + a][1 + l] = 0.0;
          Kal[iE][1 + l] = 0.0;
        } else if (b == l) {
          Kal[1 + l][1 + a] = 0.0;
          Kal[iE][1 + a] = 0.0;
        }
      }
    }
  }
  for (int k = 0; k < dim; k++) {
    if (nb->q_mask & (1u << k)) q->K[k][k][iE][iT] = 0.0;
    for (int i = 0; i < nc - 1; i++)
      if (nb->J_mask[k] & (1u << i)) {
        q->K[k][k][iY + i][iY + i] = 0.0;
        q->K[k][k][iE][iY + i] = 0.0;
      }
  }
}

/* quad::compute_h (quad.hpp:356-383 2D, 701-739 3D) */
static double compute_h(int dim, const double (*G)[3], const double* v, const double (*dv)[3]) {
  double nrm_v = 0.0;
  for (int i = 0; i < dim; i++) nrm_v += v[i] * v[i];
  nrm_v = sqrt(nrm_v);
  double s0[3], s1[3], s2[3], Gs[3], acc[3];
  for (int i = 0; i < dim; i++) s0[i] = v[i] / nrm_v;
  double al[3] = {0, 0, 0};
  for (int i = 0; i < dim; i++) {
    Gs[i] = 0.0;
    for (int j = 0; j < dim; j++) Gs[i] += G[i][j] * s0[j];
  }
  for (int i = 0; i < dim; i++) al[0] += s0[i] * Gs[i];
  for (int i = 0; i < dim; i++) {
    acc[i] = 0.0;
    for (int j = 0; j < dim; j++) acc[i] += dv[i][j] * v[j];
  }
  double as0 = 0.0;
  for (int i = 0; i < dim; i++) as0 += acc[i] * s0[i];
  for (int i = 0; i < dim; i++) s1[i] = acc[i] - as0 * s0[i];
  double n1 = 0.0;
  for (int i = 0; i < dim; i++) n1 += s1[i] * s1[i];
  n1 = sqrt(n1);
  if (n1 > TOL)
    for (int i = 0; i < dim; i++) s1[i] /= n1;
  for (int i = 0; i < dim; i++) {
    Gs[i] = 0.0;
    for (int j = 0; j < dim; j++) Gs[i] += G[i][j] * s1[j];
  }
  for (int i = 0; i < dim; i++) al[1] += s1[i] * Gs[i];
  if (dim == 3) {
    s2[0] = s0[1] * s1[2] - s0[2] * s1[1];
    s2[1] = s1[0] * s0[2] - s1[2] * s0[0];
    s2[2] = s0[0] * s1[1] - s0[1] * s1[0];
    double n2 = sqrt(s2[0] * s2[0] + s2[1] * s2[1] + s2[2] * s2[2]);
    if (n2 > TOL)
      for (int i = 0; i < 3; i++) s2[i] /= n2;
    for (int i = 0; i < 3; i++) {
      Gs[i] = 0.0;
      for (int j = 0; j < 3; j++) Gs[i] += G[i][j] * s2[j];
    }
    for (int i = 0; i < 3; i++) al[2] += s2[i] * Gs[i];
  }
  double h = 0.0;
  for (int i = 0; i < dim; i++)
    if (al[i] != 0.0) h += 1.0 / al[i];
  return 2 * sqrt(fabs(h));
}

static double dmin(double a, double b) { return (b < a) ? b : a; }

/* tau_matrix (fluid_2d.hpp:877-1022): tau = 0 when |v| < eps, else the variant
 * setup.txt selects.  h_out / td_out report h and the diagonal before DUDY^-1
 * (the diagonal variants report their diagonal). */
static int tau_matrix(const gmc_params* P, const orc_qp_t* q, const orc_geo* g,
                      double (*tau)[MD], double* h_out, double* td_out) {
  const int dim = q->dim, nc = q->nc, n = q->ndof;
  const int iE = dim + 1, iY = dim + 2;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) tau[i][j] = 0.0;
  *h_out = 0.0;
  for (int i = 0; i < n; i++) td_out[i] = 0.0;
  double vv = 0.0;
  for (int k = 0; k < dim; k++) vv += q->v[k] * q->v[k];
  const double v_nrm = sqrt(vv);
  if (v_nrm < TOL) return 0;
  const double c = q->pr.c;
  double dv[3][3];
  for (int i = 0; i < dim; i++)
    for (int j = 0; j < dim; j++) dv[i][j] = q->dI[j][1 + i];
  const double h = compute_h(dim, (const double(*)[3])g->G, q->v, (const double(*)[3])dv);
  const double rho = q->pr.rho, dt = P->dt;
  const int steady = P->steady_state;
  *h_out = h;
  double td[MD];
  if (P->tau_variant == GMC_TAU_DIAG_INCOMP_2007) {
    /* fluid_2d.hpp:963-991: diagonal, incompressible form */
    double use = 0.0;
    for (int i = 0; i < dim; i++)
      for (int j = 0; j < dim; j++) use += g->G[i][j] * g->G[i][j];
    double Gv[3], vGv = 0.0, gvec[3], gg = 0.0;
    for (int i = 0; i < dim; i++) {
      Gv[i] = 0.0;
      for (int j = 0; j < dim; j++) Gv[i] += g->G[i][j] * q->v[j];
    }
    for (int i = 0; i < dim; i++) vGv += q->v[i] * Gv[i];
    for (int k = 0; k < dim; k++) { /* g = sum_k dxi_k/dx (quad.hpp:328-329, 678-680) */
      gvec[k] = g->dsh[k][1];
      for (int a = 2; a <= dim; a++) gvec[k] += g->dsh[k][a];
    }
    for (int k = 0; k < dim; k++) gg += gvec[k] * gvec[k];
    double tau_m = vGv + 9.0 * use * pow(q->pr.mu / rho, 2);
    if (!steady) tau_m += 4.0 / (dt * dt);
    tau_m = 1.0 / sqrt(tau_m);
    const double tau_c = 1.0 / (rho * tau_m * gg);
    double tau_e = vGv + 9.0 * use * pow(q->pr.kappa / (rho * q->pr.cv), 2);
    if (!steady) tau_e += 4.0 / (dt * dt);
    tau_e = 1.0 / sqrt(tau_e);
    td[0] = tau_c;
    for (int k = 0; k < dim; k++) td[1 + k] = tau_m / rho;
    td[iE] = tau_e / (rho * q->pr.cv);
    for (int i = 0; i < nc - 1; i++) td[iY + i] = tau_m;
    for (int i = 0; i < n; i++) {
      tau[i][i] = td[i];
      td_out[i] = td[i];
    }
    return 0;
  }
  if (P->tau_variant == GMC_TAU_DIAG_2014) {
    /* fluid_2d.hpp:994-1021: diagonals of DUDY, A_k, K_kl; no steady check */
    for (int i = 0; i < n; i++) {
      const double a0 = q->DUDY[i][i];
      const double t2 = 4.0 / (dt * dt) * (a0 * a0);
      double a2 = 0.0, d2 = 0.0;
      for (int k = 0; k < dim; k++)
        for (int l = 0; l < dim; l++) {
          a2 += g->G[k][l] * (q->A[k][i][i] * q->A[l][i][i]);
          d2 += g->G[k][l] * g->G[k][l] * (q->K[k][l][i][i] * q->K[k][l][i][i]);
        }
      td[i] = pow(sqrt(t2) + sqrt(a2) + sqrt(d2) + 1.e-8, -1);
    }
    double trG = 0.0;
    for (int k = 0; k < dim; k++) trG += g->G[k][k];
    td[0] = pow(td[0] + rho * td[1] * trG, -1);
    for (int i = 0; i < n; i++) {
      tau[i][i] = td[i];
      td_out[i] = td[i];
    }
    return 0;
  }
  double use;
  if (P->tau_variant == GMC_TAU_NON_DIAG_COMP_2001) {
    /* fluid_2d.hpp:895-929.  alpha (:885) is never assigned, so lambda_sqr
     * reduces to 4/dt^2 (unsteady) */
    const double alpha0 = 0.0, alpha1 = 0.0;
    double lambda_sqr = alpha0 * (v_nrm * v_nrm + c * c) + (alpha1) * 0.5 * c * c;
    lambda_sqr += 0.5 * c * sqrt(c * c * alpha1 * alpha1 + 16 * alpha0 * alpha0 * v_nrm * v_nrm);
    if (!steady) lambda_sqr += 4 / (dt * dt);
    const double lambda = 1.0 / sqrt(lambda_sqr) + h / (2 * v_nrm);
    use = steady ? lambda : dmin(dt * 0.5, lambda);
    td[0] = use;
    td[1] = dmin(use, rho * h * h / (q->pr.mu * 12));
    for (int k = 1; k < dim; k++) td[1 + k] = td[1];
    td[iE] = dmin(use, rho * h * h * q->pr.cv / (q->pr.kappa * 12));
    for (int i = 0; i < nc - 1; i++) td[iY + i] = dmin(use, h * h / (q->pr.D_comp[i] * 12));
  } else {
    /* tau_non_diag_comp_2019, fluid_2d.hpp:932-960 */
    if (steady) use = h / (2.0 * (v_nrm + c)) + h / (2.0 * v_nrm);
    else use = dmin(dt * 0.5, h / (2.0 * (v_nrm + c)) + h / (2.0 * v_nrm));
    td[0] = use;
    td[1] = dmin(use, h * h * rho / (12.0 * q->pr.mu));
    for (int k = 1; k < dim; k++) td[1 + k] = td[1];
    td[iE] = dmin(use, h * h * rho * q->pr.cv / (12.0 * q->pr.kappa));
    for (int i = 0; i < nc - 1; i++) td[iY + i] = dmin(use, h * h / (12.0 * q->pr.D_comp[i]));
  }
  for (int i = 0; i < n; i++) td_out[i] = td[i];
  double inv[MD * MD], Dm[MD * MD];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Dm[i * n + j] = q->DUDY[i][j];
  if (orc_inverse(n, Dm, inv)) return GMC_ERR_NUMERIC;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) tau[i][j] = inv[i * n + j] * td[j];
  /* weak-compressibility correction (M < 0.1), both non-diagonal forms */
  if (v_nrm / c < 0.1) {
    double trG = 0.0;
    for (int k = 0; k < dim; k++) trG += g->G[k][k];
    tau[0][0] = pow(pow(tau[0][0], -1) + pow(rho * tau[1][1] * trG, -1), -1);
    for (int j = 1; j < n; j++) tau[0][j] = 0.0;
  }
  return 0;
}

/* dc, dc_2006 (fluid_2d.hpp:1027-1065) */
static void dc_2006(const gmc_params* P, const orc_qp_t* q, const orc_geo* g,
                    const double* uref_inv, double (*shock)[MD], double* nu_out, double* h_out) {
  const int dim = q->dim, n = q->ndof, nen = dim + 1;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) shock[i][j] = 0.0;
  *nu_out = 0.0;
  *h_out = 0.0;
  double dU[3][MD];
  for (int k = 0; k < dim; k++)
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = 0; j < n; j++) s += q->DUDY[i][j] * q->dI[k][j];
      dU[k][i] = s;
    }
  double nr = 0.0;
  for (int k = 0; k < dim; k++) nr += pow(dU[k][0], 2);
  const double del_rho_nrm = sqrt(nr);
  if (del_rho_nrm < TOL) return;
  double Jd[3];
  for (int k = 0; k < dim; k++) Jd[k] = dU[k][0] / del_rho_nrm;
  double h = 0.0;
  for (int a = 0; a < nen; a++) {
    double s = 0.0;
    for (int k = 0; k < dim; k++) s = (k == 0) ? Jd[0] * g->dsh[0][a] : s + Jd[k] * g->dsh[k][a];
    h += fabs(s);
  }
  if (h > TOL) h = 1.0 / h;
  else return;
  double Z[MD];
  for (int i = 0; i < n; i++) {
    double z = 0.0;
    for (int k = 0; k < dim; k++) {
      double s = 0.0;
      for (int j = 0; j < n; j++) s += q->A[k][i][j] * q->dI[k][j];
      z = (k == 0) ? s : z + s;
    }
    Z[i] = z - q->S[i];
  }
  double zn = 0.0, un = 0.0, use1 = 0.0;
  for (int i = 0; i < n; i++) {
    zn += (uref_inv[i] * Z[i]) * (uref_inv[i] * Z[i]);
    un += (uref_inv[i] * q->U[i]) * (uref_inv[i] * q->U[i]);
  }
  zn = sqrt(zn);
  un = sqrt(un);
  for (int k = 0; k < dim; k++) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += (uref_inv[i] * dU[k][i]) * (uref_inv[i] * dU[k][i]);
    use1 += pow(sqrt(s), 2);
  }
  const double a = P->dc_sharp;
  double nu = 0.0;
  if (a == 1.0) nu = zn * h * pow(use1, -0.5);
  else if (a == 2.0) nu = zn * h * h / un;
  else if (a == 1.5) nu = zn * h * (pow(use1, -0.5) + h / un) / 2.0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) shock[i][j] = P->dc_scale_fact * (nu * q->DUDY[i][j]);
  *nu_out = P->dc_scale_fact * nu;
  *h_out = h;
}

/* u_ref_inv (fluid_2d.hpp:123-135): Y_ref = 0.5 for every component */
void orc_uref_inv(const gmc_params* P, double* uref_inv) {
  const int dim = P->dim, nc = P->nc, n = dim + 2 + nc - 1;
  double Y[GMC_MAX_COMP], u[MD];
  orc_props pr;
  for (int c = 0; c < nc; c++) Y[c] = 0.5;
  orc_properties(P, P->p_ref, P->T_ref, Y, 0.0, &pr);
  U_vec(dim, nc, pr.rho, pr.cv, P->v_ref, P->T_ref, Y, u);
  for (int i = 0; i < n; i++) uref_inv[i] = (u[i] < TOL) ? 0.0 : 1.0 / u[i];
}

/* ------------------------------------------------------------------------ */
/* element integrals                                                        */
/* ------------------------------------------------------------------------ */

static void interp(int nen, int n, const double* sh, const double* ve, double* out) {
  for (int j = 0; j < n; j++) out[j] = 0.0;
  for (int i = 0; i < nen; i++)
    for (int j = 0; j < n; j++) out[j] += sh[i] * ve[n * i + j];
}

/* execute_i (fluid_2d.hpp:171-202) + RM_i (251-333).  m: (nen*ndof)^2 row-major,
 * r: nen*ndof, returned as -r_.  When probe is given, the quantities of
 * interior point probe_q are copied into it on the way (orc_point_probe). */
static int element_interior(const gmc_params* P, const double* xe, const double* Ie,
                            const double* Pe, const double* uref_inv, double* m, double* r,
                            int probe_q, orc_probe* probe) {
  const int dim = P->dim, nc = P->nc, n = dim + 2 + nc - 1, nen = dim + 1;
  const int N = nen * n;
  const double dt = P->dt;
  for (int i = 0; i < N * N; i++) m[i] = 0.0;
  for (int i = 0; i < N; i++) r[i] = 0.0;
  orc_qp_t Q;
  Q.dim = dim;
  Q.nc = nc;
  Q.ndof = n;
  for (int iq = 0; iq < orc_n_qp(dim); iq++) {
    double xi[3];
    orc_geo g;
    orc_qp(dim, iq, xi);
    int st = geo_at(dim, xe, xi, 0, &g);
    if (st) return st;
    const double JxW = g.J * orc_w_in(dim);
    double Pq[MD], Iq[MD], UP[MD];
    double p, vel[3], T, Y[GMC_MAX_COMP];
    orc_props prP;
    interp(nen, n, g.sh, Pe, Pq);
    assign(dim, nc, Pq, &p, vel, &T, Y);
    orc_properties(P, p, T, Y, 0.0, &prP);
    if (prP.st) return prP.st;
    U_vec(dim, nc, prP.rho, prP.cv, vel, T, Y, UP);

    interp(nen, n, g.sh, Ie, Iq);
    for (int k = 0; k < dim; k++) interp(nen, n, g.dsh[k], Ie, Q.dI[k]);
    assign(dim, nc, Iq, &Q.p, Q.v, &Q.T, Q.Y);
    const double sr = shear_rate(dim, Q.dI);
    orc_properties(P, Q.p, Q.T, Q.Y, sr, &Q.pr);
    if (Q.pr.st) return Q.pr.st;
    flux_interior(P, &Q);

    /* RM_i: Galerkin residual */
    for (int b = 0; b < nen; b++)
      for (int jb = 0; jb < n; jb++) {
        double* rr = &r[n * b + jb];
        *rr += g.sh[b] * (Q.U[jb] * JxW - UP[jb] * JxW) / dt;
        for (int k = 0; k < dim; k++) *rr -= g.dsh[k][b] * Q.Fadv[k][jb] * JxW;
        for (int k = 0; k < dim; k++) *rr += g.dsh[k][b] * Q.Fdif[k][jb] * JxW;
        *rr -= g.sh[b] * Q.S[jb] * JxW;
      }
    /* Galerkin matrix */
    for (int b = 0; b < nen; b++)
      for (int jb = 0; jb < n; jb++)
        for (int a = 0; a < nen; a++)
          for (int ja = 0; ja < n; ja++) {
            double* mm = &m[(n * b + jb) * N + n * a + ja];
            *mm += g.sh[b] * Q.DUDY[jb][ja] * g.sh[a] * JxW / dt;
            for (int k = 0; k < dim; k++) *mm -= g.dsh[k][b] * Q.A[k][jb][ja] * g.sh[a] * JxW;
            for (int k = 0; k < dim; k++)
              for (int l = 0; l < dim; l++)
                *mm += g.dsh[k][b] * Q.K[k][l][jb][ja] * g.dsh[l][a] * JxW;
            *mm -= g.sh[b] * Q.S0[jb][ja] * g.sh[a] * JxW;
          }
    /* least squares */
    double tau[MD][MD], h_tau, td[MD];
    st = tau_matrix(P, &Q, &g, tau, &h_tau, td);
    if (st) return st;
    double Res[MD];
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = 0; j < n; j++) s += Q.DUDY[i][j] * ((Iq[j] - Pq[j]) / dt);
      for (int k = 0; k < dim; k++) {
        double t = 0.0;
        for (int j = 0; j < n; j++) t += Q.A[k][i][j] * Q.dI[k][j];
        s += t;
      }
      Res[i] = s - Q.S[i];
    }
    const int probed = probe && iq == probe_q;
    if (probed) {
      probe->J = g.J;
      probe->JxW = JxW;
      for (int a = 0; a < nen; a++) {
        probe->sh[a] = g.sh[a];
        for (int k = 0; k < dim; k++) probe->dsh[k][a] = g.dsh[k][a];
      }
      for (int k = 0; k < dim; k++)
        for (int l = 0; l < dim; l++) probe->G[k][l] = g.G[k][l];
      probe->pr = Q.pr;
      probe->h = h_tau;
      for (int i = 0; i < n; i++) {
        probe->I[i] = Iq[i];
        probe->U[i] = Q.U[i];
        probe->S[i] = Q.S[i];
        probe->td[i] = td[i];
        probe->Res[i] = Res[i];
        for (int k = 0; k < dim; k++) {
          probe->dI[k][i] = Q.dI[k][i];
          probe->Fadv[k][i] = Q.Fadv[k][i];
          probe->Fdif[k][i] = Q.Fdif[k][i];
        }
        for (int j = 0; j < n; j++) {
          probe->DUDY[i][j] = Q.DUDY[i][j];
          probe->tau[i][j] = tau[i][j];
          for (int k = 0; k < dim; k++) {
            probe->A[k][i][j] = Q.A[k][i][j];
            for (int l = 0; l < dim; l++) probe->K[k][l][i][j] = Q.K[k][l][i][j];
          }
        }
      }
      probe->nu = 0.0;
      probe->h_dc = 0.0;
    }
    double fpt[MN][MD][MD];
    for (int b = 0; b < nen; b++) {
      double fp[MD][MD];
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
          double s = Q.A[0][i][j] * g.dsh[0][b];
          for (int k = 1; k < dim; k++) s += Q.A[k][i][j] * g.dsh[k][b];
          fp[i][j] = s - Q.S0[i][j] * g.sh[b];
        }
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
          double s = 0.0;
          for (int k = 0; k < n; k++) s += fp[i][k] * tau[k][j];
          fpt[b][i][j] = s;
        }
      for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int k = 0; k < n; k++) s += fpt[b][i][k] * Res[k];
        r[n * b + i] += s * JxW;
      }
    }
    for (int a = 0; a < nen; a++) {
      double Rm[MD][MD];
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
          double s = Q.DUDY[i][j] * g.sh[a] / dt;
          for (int k = 0; k < dim; k++) s += Q.A[k][i][j] * g.dsh[k][a];
          Rm[i][j] = s - Q.S0[i][j] * g.sh[a];
        }
      for (int b = 0; b < nen; b++)
        for (int i = 0; i < n; i++)
          for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += fpt[b][i][k] * Rm[k][j];
            m[(n * b + i) * N + n * a + j] += s * JxW;
          }
    }
    /* discontinuity capturing */
    if (P->dc_2006) {
      double shock[MD][MD], nu, h_dc;
      dc_2006(P, &Q, &g, uref_inv, shock, &nu, &h_dc);
      if (probed) {
        probe->nu = nu;
        probe->h_dc = h_dc;
      }
      double app[3][MD];
      for (int k = 0; k < dim; k++)
        for (int i = 0; i < n; i++) {
          double s = 0.0;
          for (int j = 0; j < n; j++) s += shock[i][j] * Q.dI[k][j];
          app[k][i] = s;
        }
      for (int b = 0; b < nen; b++)
        for (int jb = 0; jb < n; jb++)
          for (int k = 0; k < dim; k++) r[n * b + jb] += g.dsh[k][b] * app[k][jb] * JxW;
      for (int b = 0; b < nen; b++)
        for (int jb = 0; jb < n; jb++)
          for (int a = 0; a < nen; a++)
            for (int ja = 0; ja < n; ja++)
              for (int k = 0; k < dim; k++)
                m[(n * b + jb) * N + n * a + ja] += g.dsh[k][b] * shock[jb][ja] * g.dsh[k][a] * JxW;
    }
  }
  for (int i = 0; i < N; i++) r[i] = -r[i];
  return 0;
}

int orc_element_interior(const gmc_params* P, const double* xe, const double* Ie,
                         const double* Pe, const double* uref_inv, double* m, double* r) {
  return element_interior(P, xe, Ie, Pe, uref_inv, m, r, -1, NULL);
}

/* The per-point quantities of interior point q of one element, from the same
 * code path as orc_element_interior. */
int orc_point_probe(const gmc_params* P, const double* xe, const double* Ie, const double* Pe,
                    const double* uref_inv, int q, orc_probe* out) {
  const int n = P->dim + 2 + P->nc - 1, N = (P->dim + 1) * n;
  double m[(MN * MD) * (MN * MD)], r[MN * MD];
  memset(out, 0, sizeof(*out));
  if (q < 0 || q >= orc_n_qp(P->dim) || N > MN * MD) return GMC_ERR_ARG;
  return element_interior(P, xe, Ie, Pe, uref_inv, m, r, q, out);
}

/* quad::compute_h on explicit data: G and dv are row-major 3x3 (the leading
 * dim x dim part is used), dv[i][j] = d v_i / d x_j. */
double orc_compute_h(int dim, const double* G, const double* v, const double* dv) {
  double Gm[3][3], dvm[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      Gm[i][j] = G[3 * i + j];
      dvm[i][j] = dv[3 * i + j];
    }
  return compute_h(dim, (const double(*)[3])Gm, v, (const double(*)[3])dvm);
}

/* execute_b (fluid_2d.hpp:209-245) + RM_b (341-364) over the listed sides. */
int orc_element_boundary(const gmc_params* P, const double* xe, const double* Ie, int nsides,
                         const int* side_local, const int* side_flag, double* m, double* r) {
  const int dim = P->dim, nc = P->nc, n = dim + 2 + nc - 1, nen = dim + 1;
  const int N = nen * n;
  for (int i = 0; i < N * N; i++) m[i] = 0.0;
  for (int i = 0; i < N; i++) r[i] = 0.0;
  orc_qp_t Q;
  Q.dim = dim;
  Q.nc = nc;
  Q.ndof = n;
  for (int s = 0; s < nsides; s++) {
    for (int iq = 0; iq < orc_n_side_qp(dim); iq++) {
      double xi[3];
      orc_geo g;
      orc_side_qp(dim, side_local[s], iq, xi);
      int st = geo_at(dim, xe, xi, 1, &g);
      if (st) return st;
      double JNxW[3];
      for (int k = 0; k < dim; k++) JNxW[k] = g.JN[k] * orc_w_bd(dim);
      double Iq[MD];
      interp(nen, n, g.sh, Ie, Iq);
      for (int k = 0; k < dim; k++) interp(nen, n, g.dsh[k], Ie, Q.dI[k]);
      assign(dim, nc, Iq, &Q.p, Q.v, &Q.T, Q.Y);
      const double sr = shear_rate(dim, Q.dI);
      orc_properties(P, Q.p, Q.T, Q.Y, sr, &Q.pr);
      if (Q.pr.st) return Q.pr.st;
      flux_interior(P, &Q);
      flux_boundary(P, side_flag[s], &Q);
      for (int b = 0; b < nen; b++)
        for (int j = 0; j < n; j++) {
          double* rr = &r[n * b + j];
          for (int k = 0; k < dim; k++) *rr += g.sh[b] * Q.Fadv[k][j] * JNxW[k];
          for (int k = 0; k < dim; k++) *rr -= g.sh[b] * Q.Fdif[k][j] * JNxW[k];
        }
      for (int b = 0; b < nen; b++)
        for (int jb = 0; jb < n; jb++)
          for (int a = 0; a < nen; a++)
            for (int ja = 0; ja < n; ja++) {
              double* mm = &m[(n * b + jb) * N + n * a + ja];
              for (int k = 0; k < dim; k++) *mm += g.sh[b] * Q.A[k][jb][ja] * JNxW[k] * g.sh[a];
              for (int k = 0; k < dim; k++)
                for (int l = 0; l < dim; l++)
                  *mm -= g.sh[b] * Q.K[k][l][jb][ja] * JNxW[k] * g.dsh[l][a];
            }
    }
  }
  for (int i = 0; i < N; i++) r[i] = -r[i];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* steps either side of the assembly                                         */
/* ------------------------------------------------------------------------ */

/* adaptive_time_criteria::f_mc (adaptive_time_step.hpp:183-253): the element
 * minimum over the interior points of CFL / (2 max(2 mu/rho, kappa/(rho cv))
 * / h^2 + v_tau / h), 1.2 dt_prev at rest, v_tau modified above Mach 0.3;
 * then set_dt (:380-394): at most 1 % growth, clamped to [dt_min, dt_max].
 * elem_dt (optional) receives the element minima. */
double orc_adaptive_dt(const gmc_params* P, const orc_mesh_view* M, long long n_elems,
                       const double* I, double cfl, double dt_min, double dt_max, double dt_prev,
                       double* elem_dt) {
  const int dim = P->dim, nc = P->nc, n = dim + 2 + nc - 1, nen = dim + 1;
  double gmin = 0.0;
  for (long long e = 0; e < n_elems; e++) {
    double xe[MN * 3], Ie[MN * MD];
    for (int a = 0; a < nen; a++) {
      const int g = M->conn[e * nen + a];
      for (int k = 0; k < dim; k++) xe[a * dim + k] = M->coords[(long long)g * dim + k];
      for (int k = 0; k < n; k++) Ie[a * n + k] = I[(long long)g * n + k];
    }
    double emin = 0.0;
    for (int iq = 0; iq < orc_n_qp(dim); iq++) {
      double xi[3], Iq[MD], dI[3][MD], p, v[3], T, Y[GMC_MAX_COMP];
      orc_geo g;
      orc_qp(dim, iq, xi);
      geo_at(dim, xe, xi, 0, &g);
      interp(nen, n, g.sh, Ie, Iq);
      for (int k = 0; k < dim; k++) interp(nen, n, g.dsh[k], Ie, dI[k]);
      assign(dim, nc, Iq, &p, v, &T, Y);
      orc_props pr;
      orc_properties(P, p, T, Y, 0.0, &pr);
      double vv = 0.0;
      for (int k = 0; k < dim; k++) vv += v[k] * v[k];
      const double v_nrm = sqrt(vv);
      double dt_gp;
      if (v_nrm < TOL) {
        dt_gp = 1.2 * dt_prev;
      } else {
        double dv[3][3];
        for (int i = 0; i < dim; i++)
          for (int j = 0; j < dim; j++) dv[i][j] = dI[j][1 + i];
        const double h = compute_h(dim, (const double(*)[3])g.G, v, (const double(*)[3])dv);
        const double c = pr.c;
        double v_tau = v_nrm;
        if (v_nrm / c > 0.3) {
          if (dim == 2) v_tau = sqrt(v_nrm * v_nrm + 1.5 * c * c + c * sqrt(16.0 * v_nrm * v_nrm + c * c));
          else v_tau = sqrt(v_nrm * v_nrm + 2.0 * c * c + c * sqrt(4.0 * v_nrm * v_nrm + c * c));
        }
        const double a = 2.0 * pr.mu / pr.rho, b = pr.kappa / (pr.rho * pr.cv);
        const double use = 2.0 * ((a < b) ? b : a) / (h * h) + v_tau / h;
        dt_gp = cfl / use;
      }
      emin = (iq == 0 || dt_gp < emin) ? dt_gp : emin;
    }
    if (elem_dt) elem_dt[e] = emin;
    gmin = (e == 0 || emin < gmin) ? emin : gmin;
  }
  double dt = (gmin > 1.01 * dt_prev) ? 1.01 * dt_prev : gmin;
  dt = (dt_max < dt) ? dt_max : dt;
  return (dt < dt_min) ? dt_min : dt;
}

/* pressure_profile_mc::Runge_Kutta (pressure_profile.hpp:487-501): depth
 * steps of h from y_top, dp/ds = 9.81 rho(p, T, Y) by the classical RK4 with
 * T, Y at y_i and at the half step y_i - h/2 (T_half, Y_half); Y holds the
 * first nc - 1 fractions per point.  y_out[i] = y_top - h i. */
void orc_hydrostatic_profile(const gmc_params* P, double p_top, double y_top, double h, int n_steps,
                             const double* T, const double* T_half, const double* Y,
                             const double* Y_half, double* y_out, double* p_out) {
  const int nc = P->nc;
  p_out[0] = p_top;
  for (int i = 0; i <= n_steps; i++) y_out[i] = y_top - h * i;
  for (int i = 0; i < n_steps; i++) {
    double A[GMC_MAX_COMP], A1[GMC_MAX_COMP], A2[GMC_MAX_COMP];
    double s = 0.0, s1 = 0.0, s2 = 0.0;
    for (int c = 0; c < nc - 1; c++) {
      A[c] = Y[i * (nc - 1) + c];
      A1[c] = Y_half[i * (nc - 1) + c];
      A2[c] = Y[(i + 1) * (nc - 1) + c];
      s += A[c];
      s1 += A1[c];
      s2 += A2[c];
    }
    A[nc - 1] = 1.0 - s;
    A1[nc - 1] = 1.0 - s1;
    A2[nc - 1] = 1.0 - s2;
    orc_props pr;
    orc_properties(P, p_out[i], T[i], A, 0.0, &pr);
    const double m1 = 9.81 * pr.rho;
    orc_properties(P, p_out[i] + 0.5 * h * m1, T_half[i], A1, 0.0, &pr);
    const double m2 = 9.81 * pr.rho;
    orc_properties(P, p_out[i] + 0.5 * h * m2, T_half[i], A1, 0.0, &pr);
    const double m3 = 9.81 * pr.rho;
    orc_properties(P, p_out[i] + h * m3, T[i + 1], A2, 0.0, &pr);
    const double m4 = 9.81 * pr.rho;
    p_out[i + 1] = p_out[i] + h / 6. * (m1 + 2.0 * (m2 + m3) + m4);
  }
}

/* p_interpolation (pressure_profile.hpp:26-41): linear interpolation in the
 * table (y ascending or descending); GMC_ERR_ARG when y leaves the table. */
int orc_pressure_at(const double* y_tab, const double* p_tab, int n, double y, double* p) {
  long long i = 0;
  const int asc = y_tab[0] < y_tab[n - 1];
  /* lower_bound: first index whose entry is not before y in the table order */
  while (i < n && (asc ? (y_tab[i] < y) : (y_tab[i] > y))) i++;
  if (i > 0) i--;
  if (i >= n - 1) return GMC_ERR_ARG;
  const double f = (y - y_tab[i]) / (y_tab[i + 1] - y_tab[i]);
  *p = f * p_tab[i + 1] + (1.0 - f) * p_tab[i];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* BSR pattern, scatter, assembly                                            */
/* ------------------------------------------------------------------------ */

static int cmp_int(const void* a, const void* b) {
  const int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/* linear_system graph (linear_system.hpp:46-63): block (i, j) for every pair of
 * nodes that share an element.  rowptr[n_nodes+1]; colidx sorted per row.
 * Returns nnzb; colidx may be NULL to count only. */
long long orc_pattern(int n_nodes, int n_elems, int nen, const int* conn, int* rowptr,
                      int* colidx) {
  /* node -> element incidence */
  int* cnt = (int*)calloc((size_t)n_nodes + 1, sizeof(int));
  for (long long e = 0; e < n_elems; e++)
    for (int a = 0; a < nen; a++) cnt[conn[e * nen + a] + 1]++;
  for (int i = 0; i < n_nodes; i++) cnt[i + 1] += cnt[i];
  int* inc = (int*)malloc(sizeof(int) * (size_t)cnt[n_nodes]);
  int* pos = (int*)malloc(sizeof(int) * (size_t)n_nodes);
  for (int i = 0; i < n_nodes; i++) pos[i] = cnt[i];
  for (int e = 0; e < n_elems; e++)
    for (int a = 0; a < nen; a++) inc[pos[conn[(long long)e * nen + a]]++] = e;
  int* buf = (int*)malloc(sizeof(int) * 4096);
  long long nnzb = 0;
  rowptr[0] = 0;
  for (int i = 0; i < n_nodes; i++) {
    int nb = 0;
    for (int t = cnt[i]; t < cnt[i + 1]; t++)
      for (int a = 0; a < nen; a++) buf[nb++] = conn[(long long)inc[t] * nen + a];
    qsort(buf, nb, sizeof(int), cmp_int);
    int u = 0;
    for (int t = 0; t < nb; t++)
      if (t == 0 || buf[t] != buf[t - 1]) buf[u++] = buf[t];
    if (colidx)
      for (int t = 0; t < u; t++) colidx[nnzb + t] = buf[t];
    nnzb += u;
    rowptr[i + 1] = (int)nnzb;
  }
  free(buf);
  free(pos);
  free(inc);
  free(cnt);
  return nnzb;
}

static long long find_blk(const int* rowptr, const int* colidx, int i, int j) {
  int lo = rowptr[i], hi = rowptr[i + 1] - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) / 2;
    if (colidx[mid] == j) return mid;
    if (colidx[mid] < j) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

/* scatter::execute (scatter.hpp:37-98): Dirichlet rows get avg_val =
 * (1 + sum_i m_ii)/n on the diagonal, zeros elsewhere, r_i = avg_val * 0. */
void orc_scatter(int nen, int n, const int* enodes, const unsigned* dmask, double* m, double* r,
                 const int* rowptr, const int* colidx, double* vals, double* rhs) {
  const int N = nen * n;
  double avg = 1.0;
  for (int i = 0; i < N; i++) avg += m[i * N + i];
  avg /= N;
  for (int i = 0; i < N; i++) {
    const int a = i / n, k = i % n;
    if (dmask[enodes[a]] & (1u << k)) {
      for (int j = 0; j < N; j++) m[i * N + j] = 0.0;
      m[i * N + i] = avg;
      r[i] = avg * 0.0;
    }
  }
  for (int b = 0; b < nen; b++) {
    for (int jb = 0; jb < n; jb++) rhs[(long long)enodes[b] * n + jb] += r[b * n + jb];
    for (int a = 0; a < nen; a++) {
      const long long blk = find_blk(rowptr, colidx, enodes[b], enodes[a]);
      double* V = vals + blk * n * n;
      for (int jb = 0; jb < n; jb++)
        for (int ja = 0; ja < n; ja++) V[jb * n + ja] += m[(b * n + jb) * N + a * n + ja];
    }
  }
}

/* fluid_assembly::execute (fluid_assembly.hpp:39-68) over elements [e0, e1) into
 * vals/rhs (not cleared here).  Sides must be grouped by element (any order of
 * groups); side_start[e]..side_start[e+1] indexes them (CSR by element). */
int orc_assemble_range(const gmc_params* P, const orc_mesh_view* M, const double* I,
                       const double* Pst, const int* rowptr, const int* colidx, double* vals,
                       double* rhs, long long e0, long long e1) {
  const int dim = P->dim, n = dim + 2 + P->nc - 1, nen = dim + 1, N = nen * n;
  double uref[MD];
  orc_uref_inv(P, uref);
  double* m = (double*)malloc(sizeof(double) * N * N);
  double r[MN * MD], xe[MN * 3], Ie[MN * MD], Pe[MN * MD];
  int st = 0;
  for (long long e = e0; e < e1 && !st; e++) {
    const int* en = M->conn + e * nen;
    for (int a = 0; a < nen; a++) {
      for (int k = 0; k < dim; k++) xe[a * dim + k] = M->coords[(long long)en[a] * dim + k];
      for (int k = 0; k < n; k++) {
        Ie[a * n + k] = I[(long long)en[a] * n + k];
        Pe[a * n + k] = Pst[(long long)en[a] * n + k];
      }
    }
    st = orc_element_interior(P, xe, Ie, Pe, uref, m, r);
    if (st) break;
    orc_scatter(nen, n, en, M->dmask, m, r, rowptr, colidx, vals, rhs);
    const int s0 = M->side_start[e], s1 = M->side_start[e + 1];
    if (s1 > s0) {
      st = orc_element_boundary(P, xe, Ie, s1 - s0, M->side_local + s0, M->side_flag + s0, m, r);
      if (st) break;
      orc_scatter(nen, n, en, M->dmask, m, r, rowptr, colidx, vals, rhs);
    }
  }
  free(m);
  return st;
}

/* ------------------------------------------------------------------------ */
/* Krylov consumer, residual check, Newton update                            */
/* ------------------------------------------------------------------------ */

void orc_spmv(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
              const double* x, double* y) {
  for (int i = 0; i < n_nodes; i++) {
    double acc[MD];
    for (int k = 0; k < n; k++) acc[k] = 0.0;
    for (int t = rowptr[i]; t < rowptr[i + 1]; t++) {
      const double* V = vals + (long long)t * n * n;
      const double* xx = x + (long long)colidx[t] * n;
      for (int k = 0; k < n; k++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += V[k * n + j] * xx[j];
        acc[k] += s;
      }
    }
    for (int k = 0; k < n; k++) y[(long long)i * n + k] = acc[k];
  }
}

static double dot(long long n, const double* a, const double* b) {
  double s = 0.0;
  for (long long i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* ---- the device solver's arithmetic order (ordered mode of the GMRES) ----
 * The library's Krylov kernels sum in a fixed order: a dot product of length N
 * is split over nb = clamp(ceil(N / 1024), 1, 1024) blocks of 256 lanes; lane
 * t of block b accumulates a[i] b[i] with fused multiply-adds over
 * i = 256 b + t, stepping by 256 nb; each block sums its 256 lane values by a
 * pairwise tree (stride 128, 64, ..., 1); the nb block sums are folded the
 * same way (lane t adds parts t, t + 256, ... then the tree).  SpMV rows,
 * the Gram-Schmidt updates and the solution update are fused multiply-add
 * chains.  Restating that order lets a test compare the device's update with
 * this one on the same matrix to rounding level, not to the conditioning of
 * the system. */
static double tree256(double* sh) {
  for (int w = 128; w > 0; w >>= 1)
    for (int t = 0; t < w; t++) sh[t] += sh[t + w];
  return sh[0];
}

static double dot_ordered(long long N, const double* a, const double* b) {
  long long nb = (N + 1023) / 1024;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  double part[1024], sh[256];
  for (long long blk = 0; blk < nb; blk++) {
    for (int t = 0; t < 256; t++) {
      double acc = 0.0;
      for (long long i = blk * 256 + t; i < N; i += nb * 256) acc = fma(a[i], b[i], acc);
      sh[t] = acc;
    }
    part[blk] = tree256(sh);
  }
  for (int t = 0; t < 256; t++) {
    double acc = 0.0;
    for (long long i = t; i < nb; i += 256) acc += part[i];
    sh[t] = acc;
  }
  return tree256(sh);
}

static void spmv_ordered(int n_nodes, int n, const int* rowptr, const int* colidx,
                         const double* vals, const double* x, double* y) {
  for (int i = 0; i < n_nodes; i++)
    for (int k = 0; k < n; k++) {
      double acc = 0.0;
      for (int t = rowptr[i]; t < rowptr[i + 1]; t++) {
        const double* V = vals + (long long)t * n * n + k * n;
        const double* xx = x + (long long)colidx[t] * n;
        double s = 0.0;
        for (int j = 0; j < n; j++) s = fma(V[j], xx[j], s);
        acc += s;
      }
      y[(long long)i * n + k] = acc;
    }
}

/* Restarted GMRES(k), right preconditioning M^-1 = diag(J)^-1 (or identity),
 * x0 = 0, classical Gram-Schmidt applied twice (CGS2), Givens rotations.
 * Stops when the recurrence residual <= rel_tol*||b|| or after max_iters.
 * ordered != 0: the device's arithmetic order (above); point Jacobi or none. */
static int gmres_impl(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
                      const double* b, double* x, int restart, int max_iters, double rel_tol,
                      int precond, int* iters_out, double* rel_res_out, int ordered);

int orc_gmres(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
              const double* b, double* x, int restart, int max_iters, double rel_tol,
              int precond, int* iters_out, double* rel_res_out) {
  return gmres_impl(n_nodes, n, rowptr, colidx, vals, b, x, restart, max_iters, rel_tol, precond,
                    iters_out, rel_res_out, 0);
}

int orc_gmres_ordered(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
                      const double* b, double* x, int restart, int max_iters, double rel_tol,
                      int precond, int* iters_out, double* rel_res_out) {
  if (precond == 2 || precond == 3) return GMC_ERR_ARG;
  return gmres_impl(n_nodes, n, rowptr, colidx, vals, b, x, restart, max_iters, rel_tol, precond,
                    iters_out, rel_res_out, 1);
}

/* Block ILU(0) on the BSR pattern, the preconditioner of the reference's
 * shipped solver set-up (Ifpack "ILU", belos_linear_solver_factory.hpp:
 * 141-160, restated at fill level 0 on one subdomain).  Row by row (IKJ):
 * L_ik = A_ik U_kk^-1 for k < i in increasing k, then A_ij -= L_ik U_kj for
 * every j > k of row k that row i also holds.  lu: strictly lower blocks L
 * (unit block diagonal implied), diagonal and upper blocks U; dinv: U_ii^-1.
 * Columns must be sorted within each row.  Returns GMC_ERR_NUMERIC for a
 * singular diagonal block. */
int orc_bilu0(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
              double* lu, double* dinv) {
  const long long nb = (long long)n * n;
  memcpy(lu, vals, sizeof(double) * (size_t)rowptr[n_nodes] * nb);
  double* tmp = (double*)malloc(sizeof(double) * nb);
  int st = 0;
  for (int i = 0; i < n_nodes && !st; i++) {
    for (int t = rowptr[i]; t < rowptr[i + 1] && colidx[t] < i; t++) {
      const int k = colidx[t];
      double* Lik = lu + (long long)t * nb;
      /* L_ik = A_ik U_kk^-1 */
      for (int r = 0; r < n; r++)
        for (int c2 = 0; c2 < n; c2++) {
          double sacc = 0.0;
          for (int m = 0; m < n; m++) sacc += Lik[r * n + m] * dinv[(long long)k * nb + m * n + c2];
          tmp[r * n + c2] = sacc;
        }
      memcpy(Lik, tmp, sizeof(double) * nb);
      /* row i minus L_ik times the upper part of row k, on row i's pattern */
      int q = t + 1;
      for (int t2 = rowptr[k]; t2 < rowptr[k + 1]; t2++) {
        const int j = colidx[t2];
        if (j <= k) continue;
        while (q < rowptr[i + 1] && colidx[q] < j) q++;
        if (q >= rowptr[i + 1] || colidx[q] != j) continue;
        double* Aij = lu + (long long)q * nb;
        const double* Ukj = lu + (long long)t2 * nb;
        for (int r = 0; r < n; r++)
          for (int c2 = 0; c2 < n; c2++) {
            double sacc = 0.0;
            for (int m = 0; m < n; m++) sacc += Lik[r * n + m] * Ukj[m * n + c2];
            Aij[r * n + c2] -= sacc;
          }
      }
    }
    const double* Dii = NULL;
    for (int t = rowptr[i]; t < rowptr[i + 1]; t++)
      if (colidx[t] == i) Dii = lu + (long long)t * nb;
    if (!Dii || orc_inverse(n, Dii, dinv + (long long)i * nb)) st = GMC_ERR_NUMERIC;
  }
  free(tmp);
  return st;
}

/* x = U^-1 L^-1 r with the factors of orc_bilu0 */
void orc_bilu0_apply(int n_nodes, int n, const int* rowptr, const int* colidx, const double* lu,
                     const double* dinv, const double* r, double* x) {
  const long long nb = (long long)n * n;
  double* y = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n_nodes; i++) {
    for (int c2 = 0; c2 < n; c2++) x[(long long)i * n + c2] = r[(long long)i * n + c2];
    for (int t = rowptr[i]; t < rowptr[i + 1] && colidx[t] < i; t++) {
      const double* L = lu + (long long)t * nb;
      const double* xk = x + (long long)colidx[t] * n;
      for (int a = 0; a < n; a++) {
        double sacc = 0.0;
        for (int m = 0; m < n; m++) sacc += L[a * n + m] * xk[m];
        x[(long long)i * n + a] -= sacc;
      }
    }
  }
  for (int i = n_nodes - 1; i >= 0; i--) {
    for (int a = 0; a < n; a++) y[a] = x[(long long)i * n + a];
    for (int t = rowptr[i]; t < rowptr[i + 1]; t++) {
      if (colidx[t] <= i) continue;
      const double* U = lu + (long long)t * nb;
      const double* xj = x + (long long)colidx[t] * n;
      for (int a = 0; a < n; a++) {
        double sacc = 0.0;
        for (int m = 0; m < n; m++) sacc += U[a * n + m] * xj[m];
        y[a] -= sacc;
      }
    }
    for (int a = 0; a < n; a++) {
      double sacc = 0.0;
      for (int m = 0; m < n; m++) sacc += dinv[(long long)i * nb + a * n + m] * y[m];
      x[(long long)i * n + a] = sacc;
    }
  }
  free(y);
}

static int gmres_impl(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
                      const double* b, double* x, int restart, int max_iters, double rel_tol,
                      int precond, int* iters_out, double* rel_res_out, int ordered) {
#define ODOT(N_, a_, b_) (ordered ? dot_ordered(N_, a_, b_) : dot(N_, a_, b_))
#define OSPMV(x_, y_) \
  (ordered ? spmv_ordered(n_nodes, n, rowptr, colidx, vals, x_, y_) \
           : orc_spmv(n_nodes, n, rowptr, colidx, vals, x_, y_))
  const long long N = (long long)n_nodes * n;
  const int k = restart;
  double* V = (double*)malloc(sizeof(double) * N * (k + 1));
  double* w = (double*)malloc(sizeof(double) * N);
  double* z = (double*)malloc(sizeof(double) * N);
  double* dinv = (double*)malloc(sizeof(double) * N);
  double* H = (double*)calloc((size_t)(k + 1) * k, sizeof(double));
  double* cs = (double*)malloc(sizeof(double) * k);
  double* sn = (double*)malloc(sizeof(double) * k);
  double* g = (double*)malloc(sizeof(double) * (k + 1));
  double* y = (double*)malloc(sizeof(double) * k);
  double* hcol = (double*)malloc(sizeof(double) * (k + 1));
  /* precond 1: point Jacobi; 2: block Jacobi (inverse ndof x ndof diagonal
   * blocks); 3: block ILU(0) */
  double* binv = (precond == 2) ? (double*)malloc(sizeof(double) * n_nodes * n * n) : NULL;
  double* ilu = NULL;
  double* idinv = NULL;
  if (precond == 3) {
    ilu = (double*)malloc(sizeof(double) * (size_t)rowptr[n_nodes] * n * n);
    idinv = (double*)malloc(sizeof(double) * (size_t)n_nodes * n * n);
    if (orc_bilu0(n_nodes, n, rowptr, colidx, vals, ilu, idinv)) {
      free(ilu);
      free(idinv);
      free(binv);
      free(V); free(w); free(z); free(dinv); free(H); free(cs); free(sn); free(g); free(y); free(hcol);
      return GMC_ERR_NUMERIC;
    }
  }
  for (int i = 0; i < n_nodes; i++) {
    const double* D = NULL;
    for (int t = rowptr[i]; t < rowptr[i + 1]; t++)
      if (colidx[t] == i) D = vals + (long long)t * n * n;
    for (int c = 0; c < n; c++) dinv[(long long)i * n + c] = (precond == 1) ? 1.0 / D[c * n + c] : 1.0;
    if (binv) orc_inverse(n, D, binv + (long long)i * n * n);
  }
  for (long long i = 0; i < N; i++) x[i] = 0.0;
  const double bnorm = sqrt(ODOT(N, b, b));
  int it = 0;
  double res = bnorm;
  if (bnorm == 0.0) goto done;
  while (it < max_iters) {
    /* r = b - J x */
    OSPMV(x, w);
    for (long long i = 0; i < N; i++) w[i] = b[i] - w[i];
    const double beta = sqrt(ODOT(N, w, w));
    res = beta;
    if (beta <= rel_tol * bnorm) break;
    for (long long i = 0; i < N; i++) V[i] = w[i] / beta;
    for (int i = 0; i <= k; i++) g[i] = 0.0;
    g[0] = beta;
    int j;
    for (j = 0; j < k && it < max_iters; j++, it++) {
      double* vj = V + (long long)j * N;
      if (ilu) {
        orc_bilu0_apply(n_nodes, n, rowptr, colidx, ilu, idinv, vj, z);
      } else if (binv) {
        for (long long i = 0; i < n_nodes; i++)
          for (int r = 0; r < n; r++) {
            double sacc = 0.0;
            for (int c = 0; c < n; c++) sacc += binv[i * n * n + r * n + c] * vj[i * n + c];
            z[i * n + r] = sacc;
          }
      } else {
        for (long long i = 0; i < N; i++) z[i] = dinv[i] * vj[i];
      }
      OSPMV(z, w);
      for (int i = 0; i <= j + 1; i++) hcol[i] = 0.0;
      for (int pass = 0; pass < 2; pass++) {
        double* hh = y; /* scratch of k entries; y is rebuilt after the cycle */
        for (int i = 0; i <= j; i++) hh[i] = ODOT(N, V + (long long)i * N, w);
        for (int i = 0; i <= j; i++) {
          const double* vi = V + (long long)i * N;
          if (ordered)
            for (long long t = 0; t < N; t++) w[t] = fma(-hh[i], vi[t], w[t]);
          else
            for (long long t = 0; t < N; t++) w[t] -= hh[i] * vi[t];
          hcol[i] += hh[i];
        }
      }
      const double hn = sqrt(ODOT(N, w, w));
      hcol[j + 1] = hn;
      double* vn = V + (long long)(j + 1) * N;
      for (long long t = 0; t < N; t++) vn[t] = (hn != 0.0) ? w[t] / hn : 0.0;
      for (int i = 0; i < j; i++) {
        const double t0 = cs[i] * hcol[i] + sn[i] * hcol[i + 1];
        hcol[i + 1] = -sn[i] * hcol[i] + cs[i] * hcol[i + 1];
        hcol[i] = t0;
      }
      const double den = sqrt(hcol[j] * hcol[j] + hcol[j + 1] * hcol[j + 1]);
      cs[j] = (den != 0.0) ? hcol[j] / den : 1.0;
      sn[j] = (den != 0.0) ? hcol[j + 1] / den : 0.0;
      hcol[j] = den;
      hcol[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int i = 0; i <= j; i++) H[i * k + j] = hcol[i];
      res = fabs(g[j + 1]);
      if (res <= rel_tol * bnorm) {
        j++;
        it++;
        break;
      }
    }
    /* y = R^-1 g, x += M^-1 V y */
    for (int i = j - 1; i >= 0; i--) {
      double s = g[i];
      for (int c = i + 1; c < j; c++) s -= H[i * k + c] * y[c];
      y[i] = s / H[i * k + i];
    }
    for (long long t = 0; t < N; t++) {
      double s = 0.0;
      for (int i = 0; i < j; i++)
        s = ordered ? fma(V[(long long)i * N + t], y[i], s) : s + V[(long long)i * N + t] * y[i];
      z[t] = s;
    }
    if (ilu) {
      orc_bilu0_apply(n_nodes, n, rowptr, colidx, ilu, idinv, z, w);
      for (long long t = 0; t < N; t++) x[t] += w[t];
    } else if (binv) {
      for (long long i = 0; i < n_nodes; i++)
        for (int r = 0; r < n; r++) {
          double sacc = 0.0;
          for (int c = 0; c < n; c++) sacc += binv[i * n * n + r * n + c] * z[i * n + c];
          x[i * n + r] += sacc;
        }
    } else {
      for (long long t = 0; t < N; t++) x[t] = ordered ? fma(dinv[t], z[t], x[t]) : x[t] + dinv[t] * z[t];
    }
    if (res <= rel_tol * bnorm) break;
  }
  /* true residual */
  OSPMV(x, w);
  for (long long i = 0; i < N; i++) w[i] = b[i] - w[i];
  res = sqrt(ODOT(N, w, w));
done:
  if (iters_out) *iters_out = it;
  if (rel_res_out) *rel_res_out = (bnorm > 0.0) ? res / bnorm : 0.0;
  free(binv);
  free(ilu);
  free(idinv);
  free(V);
  free(w);
  free(z);
  free(dinv);
  free(H);
  free(cs);
  free(sn);
  free(g);
  free(y);
  free(hcol);
  return 0;
#undef ODOT
#undef OSPMV
}

/* residual_check::execute (residual_check.hpp:42-80).  state[0] holds ||b_0||.
 * Returns 1 when converged. */
int orc_residual_check(const double* b, long long n, int it, double* state, double* res_out) {
  const double nrm = sqrt(dot(n, b, b));
  if (it == 0) {
    state[0] = nrm;
    *res_out = nrm;
    return 0;
  }
  *res_out = nrm / state[0];
  return nrm <= ((1.e-8 * state[0] > 1.e-5) ? 1.e-8 * state[0] : 1.e-5);
}

/* fluid_updater::corrector + model::dof_trimmer(trim_dof, 0, 1) */
void orc_newton_update(long long n_nodes, int n, int trim_dof, double* I, const double* delta) {
  for (long long i = 0; i < n_nodes * n; i++) I[i] += delta[i];
  for (long long i = 0; i < n_nodes; i++) {
    const double d = I[i * n + trim_dof];
    const double mn = (d < 1.0) ? d : 1.0;
    I[i * n + trim_dof] = (0.0 < mn) ? mn : 0.0;
  }
}
