/*
 * gmc_oracle.h -- FP64 CPU oracle of the fluid_mc assembly path.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, never by the product library.
 */
#ifndef GMC_ORACLE_H_
#define GMC_ORACLE_H_

#include "../include/gales_mc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_props {
  double rho, mu, cp, cv, kappa, c, alpha, beta, D;
  double rho_comp[GMC_MAX_COMP], e_comp[GMC_MAX_COMP], D_comp[GMC_MAX_COMP];
  int st;   /* 0, or GMC_ERR_RANGE / GMC_ERR_NUMERIC from a magma component */
} orc_props;

/* One magma submixture (mixtures.hpp:115-509) at (p, T, strain rate): the
 * submixture's properties, its five chemicals' mass fractions (melt, H2O(l),
 * CO2(l), H2O(g), CO2(g)), the gas volume fraction and the dissolved mass
 * fractions of the liquid from the tables. */
typedef struct orc_magma_out {
  double rho, mu, cp, cv, kappa, c, alpha, beta, D;
  double wf[5], vf[5], rho_ch[5], vf_gas, wl_h2o, wl_co2, mu_melt;
  double alpha_ch[5], beta_ch[5], c_ch[5], cp_ch[5], cv_ch[5];
  int st;
} orc_magma_out;
void orc_magma(const gmc_magma* g, double p, double T, double sr, orc_magma_out* o);
/* Block ILU(0) (Ifpack "ILU" at fill level 0, belos_linear_solver_factory.hpp:
 * 141-160): lu = L (strictly lower blocks) and U (diagonal and upper blocks),
 * dinv = U_ii^-1; apply: x = U^-1 L^-1 r.  gmres precond 3 uses them. */
int orc_bilu0(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
              double* lu, double* dinv);
void orc_bilu0_apply(int n_nodes, int n, const int* rowptr, const int* colidx, const double* lu,
                     const double* dinv, const double* r, double* x);
/* Pal (2003) viscosity correction alone (mixtures.hpp:266-291) */
double orc_pal_2003(double vf_g, double mu, double shear_rate, int* st);

typedef struct orc_mesh_view {
  const double* coords;      /* [n_nodes*dim] */
  const int* conn;           /* [n_elems*nen] */
  const unsigned* dmask;     /* [n_nodes] Dirichlet dof bitmask */
  const int* side_start;     /* [n_elems+1] CSR of boundary sides by element */
  const int* side_local;     /* [n_sides] */
  const int* side_flag;      /* [n_sides] */
} orc_mesh_view;

/* One interior quadrature point of orc_element_interior, for the pins of
 * the per-point quantities (tau, compute_h, dc_2006, flux matrices). */
typedef struct orc_probe {
  double J, JxW;
  double sh[4], dsh[3][4], G[3][3];
  double I[GMC_MAX_DOF], dI[3][GMC_MAX_DOF];
  orc_props pr;
  double U[GMC_MAX_DOF], S[GMC_MAX_DOF], Res[GMC_MAX_DOF];
  double DUDY[GMC_MAX_DOF][GMC_MAX_DOF];
  double A[3][GMC_MAX_DOF][GMC_MAX_DOF];
  double K[3][3][GMC_MAX_DOF][GMC_MAX_DOF];
  double Fadv[3][GMC_MAX_DOF], Fdif[3][GMC_MAX_DOF];
  double h;                                /* compute_h of tau (0 when tau = 0) */
  double td[GMC_MAX_DOF];                  /* tau diagonal before DUDY^-1 */
  double tau[GMC_MAX_DOF][GMC_MAX_DOF];    /* after the Mach correction */
  double nu;                               /* dc_scale_fact * nu_shoc (0: no DC) */
  double h_dc;                             /* DC length scale (0: no DC) */
} orc_probe;

void orc_properties(const gmc_params* P, double p, double T, const double* Y, double sr,
                    orc_props* o);
int orc_n_qp(int dim);
void orc_qp(int dim, int q, double* xi);
double orc_w_in(int dim);
double orc_w_bd(int dim);
int orc_n_side_qp(int dim);
void orc_side_qp(int dim, int side, int q, double* xi);
int orc_inverse(int n, const double* A, double* Inv);
void orc_uref_inv(const gmc_params* P, double* uref_inv);
int orc_element_interior(const gmc_params* P, const double* xe, const double* Ie,
                         const double* Pe, const double* uref_inv, double* m, double* r);
int orc_point_probe(const gmc_params* P, const double* xe, const double* Ie, const double* Pe,
                    const double* uref_inv, int q, orc_probe* out);
double orc_compute_h(int dim, const double* G, const double* v, const double* dv);
int orc_element_boundary(const gmc_params* P, const double* xe, const double* Ie, int nsides,
                         const int* side_local, const int* side_flag, double* m, double* r);
double orc_adaptive_dt(const gmc_params* P, const orc_mesh_view* M, long long n_elems,
                       const double* I, double cfl, double dt_min, double dt_max, double dt_prev,
                       double* elem_dt);
void orc_hydrostatic_profile(const gmc_params* P, double p_top, double y_top, double h, int n_steps,
                             const double* T, const double* T_half, const double* Y,
                             const double* Y_half, double* y_out, double* p_out);
int orc_pressure_at(const double* y_tab, const double* p_tab, int n, double y, double* p);
long long orc_pattern(int n_nodes, int n_elems, int nen, const int* conn, int* rowptr,
                      int* colidx);
void orc_scatter(int nen, int n, const int* enodes, const unsigned* dmask, double* m, double* r,
                 const int* rowptr, const int* colidx, double* vals, double* rhs);
int orc_assemble_range(const gmc_params* P, const orc_mesh_view* M, const double* I,
                       const double* Pst, const int* rowptr, const int* colidx, double* vals,
                       double* rhs, long long e0, long long e1);
void orc_spmv(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
              const double* x, double* y);
int orc_gmres(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
              const double* b, double* x, int restart, int max_iters, double rel_tol,
              int precond, int* iters_out, double* rel_res_out);
/* orc_gmres in the device solver's summation order (point Jacobi or none) */
int orc_gmres_ordered(int n_nodes, int n, const int* rowptr, const int* colidx, const double* vals,
                      const double* b, double* x, int restart, int max_iters, double rel_tol,
                      int precond, int* iters_out, double* rel_res_out);
int orc_residual_check(const double* b, long long n, int it, double* state, double* res_out);
void orc_newton_update(long long n_nodes, int n, int trim_dof, double* I, const double* delta);

#ifdef __cplusplus
}
#endif

#endif
