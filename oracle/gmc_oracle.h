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
} orc_props;

typedef struct orc_mesh_view {
  const double* coords;      /* [n_nodes*dim] */
  const int* conn;           /* [n_elems*nen] */
  const unsigned* dmask;     /* [n_nodes] Dirichlet dof bitmask */
  const int* side_start;     /* [n_elems+1] CSR of boundary sides by element */
  const int* side_local;     /* [n_sides] */
  const int* side_flag;      /* [n_sides] */
} orc_mesh_view;

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
int orc_element_boundary(const gmc_params* P, const double* xe, const double* Ie, int nsides,
                         const int* side_local, const int* side_flag, double* m, double* r);
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
int orc_residual_check(const double* b, long long n, int it, double* state, double* res_out);
void orc_newton_update(long long n_nodes, int n, int trim_dof, double* I, const double* delta);

#ifdef __cplusplus
}
#endif

#endif
