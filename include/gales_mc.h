/*
 * gales_mc.h -- C-ABI of the MI355X-native fluid_mc element-assembly library.
 *
 * The library replaces the multicomponent compressible-flow (fluid_mc) assembly
 * path of GALES-multi-fluid with hand-written HIP kernels for gfx950.  Every
 * entry point takes plain pointers and sizes; no C++ or torch type crosses this
 * boundary.  Each entry point below names the reference interface it replaces
 * (paths relative to the GALES-multi-fluid source tree).
 *
 * Conventions
 *  - State layout is node-major interleaved, dof gid = node * ndof + k, with
 *    k = [p, v_1..v_dim, T, Y_1..Y_{nc-1}]  (src/fem/mesh/node.hpp:73,
 *    sim/fluid_mc/P0/fluid_2d.hpp:86-99).  ndof = dim + 2 + nc - 1.
 *  - Two state slots: I (current Newton iterate, slot 0) and P (previous time
 *    step, slot 1)  (sim/fluid_mc/P0/main.cpp:100, fluid_2d.hpp:143-144).
 *  - The assembled system is J * Delta = rhs with rhs = -R
 *    (fluid_2d.hpp:200-201, src/fem/ls/scatter.hpp:96-97).  J is stored as a
 *    block-CSR (BSR) matrix with ndof x ndof row-major blocks; block (i, j)
 *    holds the rows of node i and the columns of node j, so it is the same
 *    matrix as the reference's point CSR (src/fem/ls/linear_system.hpp:40-66).
 *  - Every call returns GMC_OK (0) or a negative status; gmc_last_error()
 *    gives the message.  This replaces Error() -> MPI_Abort
 *    (src/fem/base/algorithm.hpp:37-42).
 *  - One context is bound to one HIP stream and is not thread-safe, as the
 *    reference integral object is not re-entrant.  Distinct contexts are
 *    independent.
 */
#ifndef GALES_MC_H_
#define GALES_MC_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMC_OK 0
#define GMC_ERR_ARG (-1)        /* invalid argument or size */
#define GMC_ERR_HIP (-2)        /* HIP runtime failure */
#define GMC_ERR_MESH (-3)       /* non-positive element Jacobian (quad.hpp:215-218) */
#define GMC_ERR_NUMERIC (-4)    /* singular DUDY: rho = 0 or a zero (p, T) determinant of its closed-form
                                   inverse, a zero pivot where the LU is kept (algorithm.hpp:407-410) */
#define GMC_ERR_UNSUPPORTED (-5)
#define GMC_ERR_STATE (-6)      /* call order (e.g. solve before assemble) */
#define GMC_ERR_RANGE (-7)      /* (p, T) outside a magma solubility table (mixtures.hpp:349-353, 379-389) */

#define GMC_MAX_COMP 4
#define GMC_MAX_DOF 8
#define GMC_MAX_FLAG 16

/* component models (src/fem/physical_props/fluid/chemicals.hpp) */
#define GMC_COMP_CUSTOM 0     /* custom_chemical, chemicals.hpp:19-49 */
#define GMC_COMP_IDEAL_GAS 1  /* ideal_gas, chemicals.hpp:198-271 */
#define GMC_COMP_MAGMA 2      /* magma_mixture submixture (mix_of_mix), mixtures.hpp:115-509 */

/* magma_mixture models of gas and crystals on the viscosity (mixtures.hpp:161-166) */
#define GMC_MU_NONE 0
#define GMC_MU_ISHII_ZUBER 1      /* gas: mu / (1 - vf_g) */
#define GMC_MU_PAL_2003 2         /* gas, strain-rate dependent */
#define GMC_MU_EINSTEIN_ROSCOE 1  /* crystals: mu (1 - vf_c)^-2.5 */
#define GMC_MU_CARICCHI 2         /* crystals, strain-rate dependent */

/* One magma submixture of five chemicals (melt, H2O(l), CO2(l), H2O(g),
 * CO2(g)): the magma_mix block of props.txt (fluid_properties_reader.hpp:
 * 293-390; defaults melt_kappa 1.5, h2o_l/co2_l cp 2278 and kappa 0.64) and
 * its solubility tables (input/lookuptable, mixtures.hpp:416-490): the
 * dissolved H2O and CO2 mass fractions of the liquid on an equispaced pressure
 * grid (n_T = 0) or on an n_T x n_p temperature-pressure grid (row-major, T
 * rows).  The table pointers are host memory read by gmc_create. */
typedef struct gmc_magma {
  double oxide_wf[10];                  /* SiO2 TiO2 Al2O3 Fe2O3 FeO MnO MgO CaO Na2O K2O */
  double melt_kappa;
  double h2o_wf, co2_wf;                /* total H2O and CO2 (liquid + gas) per magma mass */
  double h2o_l_cp, h2o_l_kappa, co2_l_cp, co2_l_kappa;
  double crystal_vf, crystal_rho;
  int gas_on_mu;                        /* GMC_MU_NONE / ISHII_ZUBER / PAL_2003 */
  int crystal_on_mu;                    /* GMC_MU_NONE / EINSTEIN_ROSCOE / CARICCHI */
  int n_p, n_T;
  const double* p_table;                /* [n_p] */
  const double* T_table;                /* [n_T] (n_T > 0) */
  const double* wf_h2o_l;               /* [n_p] or [n_T * n_p] */
  const double* wf_co2_l;
} gmc_magma;

/* tau variants (fluid_2d.hpp:877-1022; setup.txt:25-28 selects one) */
#define GMC_TAU_NON_DIAG_COMP_2001 2001  /* fluid_2d.hpp:895-929 (steady_state unsupported: tau infinite) */
#define GMC_TAU_NON_DIAG_COMP_2019 2019  /* fluid_2d.hpp:932-960, the active one in the cases */
#define GMC_TAU_DIAG_INCOMP_2007 2007    /* fluid_2d.hpp:963-991 */
#define GMC_TAU_DIAG_2014 2014           /* fluid_2d.hpp:994-1021 */

/* precision tiers of the element kernels */
#define GMC_PREC_FP64 0       /* plain FP64 form */
#define GMC_PREC_FP32 1       /* FP32 element matrices (VALU FP32 FMAs; least squares on MFMA with GMC_FP32_MFMA=1), FP64 EOS/CSR/Krylov */

/* neumann override bits (fluid_ic_bc.hpp:131-217) */
#define GMC_N_TAU11 (1u << 0)
#define GMC_N_TAU22 (1u << 1)
#define GMC_N_TAU33 (1u << 2)
#define GMC_N_TAU12 (1u << 3)
#define GMC_N_TAU13 (1u << 4)
#define GMC_N_TAU23 (1u << 5)

typedef struct gmc_component {
  int kind;                  /* GMC_COMP_* */
  /* custom_chemical(rho, mu, cp, kappa, alpha, beta)  chemicals.hpp:23-33 */
  double rho, mu, cp, kappa, alpha, beta;
  /* ideal_gas(molar_mass, gamma)  chemicals.hpp:237-246; mu = 1e-5, Pr = 0.72 */
  double molar_mass, gamma;
  /* chemical_diffusivity_ (base_chemical.hpp:86, default 1e-10) */
  double diffusivity;
  gmc_magma magma;                      /* kind GMC_COMP_MAGMA */
} gmc_component;

/* Neumann data of one side flag: the pairs returned by the ic_bc neumann_*
 * hooks (fluid_ic_bc.hpp:131-195).  tau index order 11,22,33,12,13,23. */
typedef struct gmc_neumann {
  unsigned tau_mask;                    /* GMC_N_TAU* bits */
  double tau[6];
  unsigned q_mask;                      /* bit k: q_{k+1} prescribed */
  double q[3];
  unsigned J_mask[3];                   /* [k] bit i: J_{k+1} of component i */
  double J[3][GMC_MAX_COMP];
  /* prescribed mass flux of direction k+1 (mass_flux1/2/3, fluid_ic_bc.hpp:
   * 202-217): replaces F_k^adv and A_k at the side's points
   * (fluid_2d.hpp:627-758) */
  unsigned mflux_mask;                  /* bit k: mass_flux_{k+1} prescribed */
  double mflux[3];
} gmc_neumann;

typedef struct gmc_params {
  int dim;                              /* 2 or 3 */
  int nc;                               /* number of components, 2..4 */
  gmc_component comp[GMC_MAX_COMP];
  double gravity[3];                    /* body_force, fluid_ic_bc.hpp:44-48 */
  double p_ref, v_ref[3], T_ref;        /* dc_2006 reference state, fluid_ic_bc.hpp:30-33 */
  double dt;                            /* time::get().delta_t() */
  int steady_state;                     /* setup.txt:24 */
  int tau_variant;                      /* GMC_TAU_* */
  int dc_2006;                          /* setup.txt:29 */
  double dc_sharp;                      /* 1.0, 1.5 or 2.0 (setup.txt:30) */
  double dc_scale_fact;                 /* setup.txt:31 */
  /* node flag -> bitmask of Dirichlet dofs (dof_constraint,
   * fluid_dofs_ic_bc.hpp:86-117); flags >= GMC_MAX_FLAG are an error */
  unsigned dirichlet_mask[GMC_MAX_FLAG];
  gmc_neumann neumann[GMC_MAX_FLAG];    /* side flag -> Neumann overrides */
  int precision;                        /* GMC_PREC_* */
} gmc_params;

/* Mesh of linear triangles (dim 2) or tetrahedra (dim 3), nodes counter-
 * clockwise / positive Jacobian.  Boundary sides are (element, local side,
 * flag) with local side numbering of element.hpp:145-149 (triangle:
 * {1,2},{2,0},{0,1}) and element.hpp:455-459 (tet: {1,2,3},{0,3,2},{0,1,3},
 * {0,2,1}). */
typedef struct gmc_mesh {
  int dim;
  int n_nodes;
  int n_elems;
  int n_sides;
  const double* coords;                 /* [n_nodes * dim] */
  const int* conn;                      /* [n_elems * (dim + 1)] */
  const int* node_flag;                 /* [n_nodes] */
  const int* side_elem;                 /* [n_sides] */
  const int* side_local;                /* [n_sides] */
  const int* side_flag;                 /* [n_sides] */
} gmc_mesh;

typedef struct gmc_info {
  int dim, nc, ndof, nen;
  int n_nodes, n_elems, n_sides;
  long long nnzb;                       /* blocks in the BSR matrix */
  int n_colors;                         /* interior element colours (0 unless FP64 assemblies colour-scatter) */
  int n_bd_elems;                       /* elements with boundary sides */
  int n_bd_colors;
  size_t device_bytes;                  /* device memory held by the context */
  int scatter;                          /* of FP64 assemblies: 2 patch-local (LDS) + halo gather, 1 element buffer + gather, 0 colour scatter
                                           (FP32-tier assemblies also use an FP32-only element buffer when FP64 ones use 0) */
  int ebuf_chunks;                      /* element chunks through the element buffer (1 when it holds the whole mesh) */
  int fp32_mfma;                        /* 1: FP32-tier interior element matrices form their least-squares term with
                                           v_mfma_f32_16x16x4f32 (environment GMC_FP32_MFMA=1 at gmc_create; patch-local
                                           or element-buffer scatter, 2D with 2 components); 0: on the vector ALUs */
} gmc_info;

typedef struct gmc_solver {
  int restart;                          /* GMRES(k) subspace size */
  int max_iters;                        /* total iteration cap */
  double rel_tol;                       /* stop when ||r|| <= rel_tol * ||b|| */
  int precond;                          /* 0 none, 1 point Jacobi, 2 block Jacobi (ndof x ndof),
                                           3 block ILU(0) (GMC_PRECOND_ILU0) */
} gmc_solver;

/* Block ILU(0) of the assembled matrix on its own BSR pattern: the reference's
 * preconditioner is Ifpack "ILU" (belos_linear_solver_factory.hpp:141-160,
 * ls_precond ILU in sim/fluid_mc/P0/setup.txt), an ILU(k) per MPI subdomain;
 * on one GPU the subdomain is the whole mesh and the fill level is 0. */
#define GMC_PRECOND_ILU0 3

typedef struct gmc_solve_info {
  int iterations;                       /* belos num_it() */
  double rel_residual;                  /* true ||b - J x|| / ||b||  (RRE) */
  double abs_residual;                  /* true ||b - J x||          (ARE) */
  double seconds;
} gmc_solve_info;

typedef struct gmc_timing {
  double assemble_s;                    /* whole gmc_assemble span */
  double clear_s;                       /* matrix/rhs clear kernels */
  double interior_s;                    /* sum of interior element kernels */
  double boundary_s;                    /* sum of boundary element kernels */
  double gather_s;                      /* element-buffer gather kernels */
  int interior_launches;
  int boundary_launches;
} gmc_timing;

typedef struct gmc_ctx gmc_ctx;

/* ---- lifecycle ---------------------------------------------------------- */

/* Builds everything the reference builds once before the time loop: the
 * linear_system graph (linear_system.hpp:40-66) as a device BSR pattern, the
 * element colouring used by the deterministic scatter, the per-element block
 * map, the boundary-side lists (mesh.hpp:468-495 precomputes quad data; here
 * geometry is recomputed in-kernel), device state buffers, and u_ref_inv
 * (fluid_2d.hpp:102,123-135).  It also picks the scatter: the patch-local
 * plan in 2D (Morton-ordered patches of 64 elements whose LDS fits in 64 KB),
 * else the element buffer when it fits in half the free device memory (FP64
 * entries, or FP32 entries for the precision tier alone), else the colour
 * scatter; the environment variable GMC_SCATTER = patch | ebuf | ebuf32 |
 * colour forces one.  Returns GMC_ERR_MESH for a non-positive element
 * Jacobian. */
int gmc_create(const gmc_mesh* mesh, const gmc_params* params, gmc_ctx** out);
int gmc_destroy(gmc_ctx* ctx);
/* Message of the last failure of ctx (or of the last failed gmc_create when
 * ctx is NULL). */
const char* gmc_last_error(const gmc_ctx* ctx);
int gmc_get_info(const gmc_ctx* ctx, gmc_info* info);
/* Binds the context to a caller-owned HIP stream (hipStream_t as void*); NULL
 * restores the context's own stream. */
int gmc_set_stream(gmc_ctx* ctx, void* hip_stream);
/* Changes dt (time::get().delta_t(), fluid_2d.hpp:141) between calls. */
int gmc_set_dt(gmc_ctx* ctx, double dt);
/* Replaces the six adjoint material parameters theta = (rho0, mu, cp) of
 * components 0 and 1 (see gmc_vjp). */
int gmc_set_material(gmc_ctx* ctx, const double theta[6]);

/* ---- state (dof_state, src/fem/dofs/dof_state.hpp:17-49) ----------------- */

/* Host arrays of n_nodes * ndof doubles; either pointer may be NULL. */
int gmc_set_state(gmc_ctx* ctx, const double* I, const double* P);
int gmc_get_state(const gmc_ctx* ctx, double* I, double* P);
/* Device-resident variants (pointers to device memory of the same size). */
int gmc_set_state_device(gmc_ctx* ctx, const double* dI, const double* dP);

/* ---- assembly ------------------------------------------------------------ */

#define GMC_ASM_MATRIX (1 << 0)
#define GMC_ASM_RHS (1 << 1)
#define GMC_ASM_FP32 (1 << 2)           /* element matrices in FP32 (the precision tier; VALU, or its least-squares
                                           product on MFMA where gmc_info.fp32_mfma says so) */

/* = fluid_assembly::execute()  (sim/fluid_mc/P0/fluid_assembly.hpp:39-68):
 * clear, every interior element (execute_i), the Dirichlet avg_val rule and
 * sum-into the device BSR + rhs (scatter.hpp:37-98), then every boundary
 * element (execute_b) with its own scatter.  Writes the elapsed seconds of the
 * device span to *seconds (may be NULL). */
int gmc_assemble(gmc_ctx* ctx, int flags, double* seconds);
/* rhs only, FP64, and ||rhs||_2 (residual_check::execute's MvNorm,
 * src/fem/ls/residual_check.hpp:42-55). */
int gmc_residual(gmc_ctx* ctx, double* norm);
/* Timings of the last gmc_assemble, from HIP events on the context stream. */
int gmc_get_timing(const gmc_ctx* ctx, gmc_timing* t);

/* Single-element integrals = fluid<ic_bc,dim>::execute_i / execute_b
 * (fluid_2d.hpp:171-245) for element e, run on the device; m is
 * (nen*ndof)^2 row-major, r has nen*ndof entries and is returned negated as in
 * the reference (r = -r_).  No Dirichlet rule is applied (that is scatter's).
 * execute_b of an element without boundary sides returns zeros. */
int gmc_element(gmc_ctx* ctx, int e, int boundary, double* m, double* r);

/* Sizes and host copies of the assembled system (linear_system matrix()/rhs()). */
int gmc_get_pattern(const gmc_ctx* ctx, int* rowptr, int* colidx);
int gmc_get_system(const gmc_ctx* ctx, double* vals, double* rhs);
/* y = J x with host vectors of n_nodes*ndof doubles. */
int gmc_spmv(gmc_ctx* ctx, const double* x, double* y);

/* ---- solver and Newton update --------------------------------------------- */

/* Solves J * Delta = rhs on the device: restarted GMRES(k), k <= 200, right
 * preconditioning, x0 = 0 (belos_linear_solver_factory.hpp:103-218), with
 * point or block Jacobi or block ILU(0).  Delta stays on the device until
 * gmc_newton_update or gmc_get_solution. */
int gmc_solve(gmc_ctx* ctx, const gmc_solver* s, gmc_solve_info* info);
int gmc_get_solution(const gmc_ctx* ctx, double* delta);
/* fluid_updater::corrector (fluid_updater.hpp:25-28) followed by
 * model::dof_trimmer(first Y dof, 0, 1) (model.hpp:176-184, main.cpp:234-235). */
int gmc_newton_update(gmc_ctx* ctx);
/* fluid_updater::predictor: P <- I (fluid_updater.hpp:18-21). */
int gmc_predictor(gmc_ctx* ctx);
/* fluid_updater::corrector alone: I <- I + Delta (fluid_updater.hpp:25-28),
 * Delta being the device solution of the last gmc_solve / gmc_refine. */
int gmc_corrector(gmc_ctx* ctx);
/* model::dof_trimmer(dof_index, lo, hi) (model.hpp:176-184): every node's dof
 * dof_index of I becomes max(lo, min(value, hi)). */
int gmc_dof_trimmer(gmc_ctx* ctx, int dof_index, double lo, double hi);
/* ||rhs||_2 of the assembled rhs without re-assembling (the MvNorm of
 * residual_check::execute, src/fem/ls/residual_check.hpp:42-55). */
int gmc_rhs_norm(const gmc_ctx* ctx, double* norm);
/* gmc_rhs_norm that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory per call (code under HIP
 * graph capture, tight serving loops).  The reduction's partial sums live in a
 * fixed 10 KiB device array per GPU (every rhs length fits), shared by all
 * calls on that GPU: calls that overlap on one GPU race.  Same result as
 * gmc_rhs_norm, bit for bit. */
int gmc_rhs_norm_noalloc(const gmc_ctx* ctx, double* norm);

/* ---- steps either side of the assembly ----------------------------------- */

/* CFL-adaptive time step: adaptive_time_criteria::f_mc + set_dt
 * (src/fem/utils/adaptive_time_step.hpp:183-253, 380-394; called from
 * sim/fluid_mc/P0/main.cpp:257-260).  Over the interior points of every
 * element, on the I state: CFL / (2 max(2 mu/rho, kappa/(rho cv)) / h^2 +
 * v_tau / h), 1.2 dt where |v| < eps.  The mesh minimum, grown by at most 1 %
 * over the context's dt and clamped to [dt_min, dt_max], becomes the
 * context's dt (time::get().delta_t) and is returned in *dt. */
int gmc_adaptive_dt(gmc_ctx* ctx, double cfl, double dt_min, double dt_max, double* dt);

/* Hydrostatic pressure table: pressure_profile_mc::Runge_Kutta
 * (src/fem/utils/pressure_profile.hpp:487-501), no device work.  Depth steps of
 * h from y_top (y_out[i] = y_top - h i, i = 0..n_steps), dp/ds = 9.81 rho(p,
 * T, Y) by classical RK4, p_out[0] = p_top.  T[n_steps + 1] and Y[(n_steps +
 * 1) * (nc - 1)] at y_out, T_half[n_steps] and Y_half[n_steps * (nc - 1)] at
 * y_out[i] - h/2 (the ic_bc initial_T / initial_Y of those points).  Node
 * pressures follow by p_interpolation (pressure_profile.hpp:26-41), in
 * gales_mc::pressure_from_profile of the host layer. */
int gmc_hydrostatic_profile(const gmc_params* params, double p_top, double y_top, double h,
                            int n_steps, const double* T, const double* T_half, const double* Y,
                            const double* Y_half, double* y_out, double* p_out);

/* ---- precision tier: mixed-precision iterative refinement ----------------- */

/* Matrix-free FP64 Jacobian product y = J(I, P) x (host vectors of
 * n_nodes*ndof doubles): the element matrices of the plain form are recomputed
 * in FP64 and applied to x without storing them, Dirichlet rule included, so
 * the result equals gmc_spmv after an FP64 assembly.  No reference
 * counterpart; it is the FP64 operator of the refinement below. */
int gmc_jvp(gmc_ctx* ctx, const double* x, double* y);

/* Newton update in the FP32 tier.  rhs = -R(I) is assembled in FP64 and the
 * working matrix J32 with FP32 element matrices (FP64 properties and closures,
 * FP64 accumulation).  The Newton equation J64 x = rhs is then solved by
 * mixed-precision iterative refinement, n_refine >= 1 passes of
 *   solve J32 d = rhs - J64 x  (GMRES with s);  x += d
 * with J64 x formed matrix-free in FP64 (gmc_jvp).  x is left as the update
 * for gmc_newton_update / gmc_get_solution; the assembled rhs is -R(I) in
 * FP64 and the stored matrix is J32.  res_norms (n_refine + 1 entries, may be
 * NULL) receives ||rhs - J64 x|| before the first and after every pass. */
int gmc_refine(gmc_ctx* ctx, const gmc_solver* s, int n_refine, double* res_norms);

/* ---- adjoint ------------------------------------------------------------- */

#define GMC_VJP_EXACT 0     /* exact derivative of R: tau, nu, properties, S, u_ref all differentiated */
#define GMC_VJP_FROZEN 1    /* J^T v with J the assembled (frozen-coefficient) matrix */

/* Vector-Jacobian products of the assembled residual R (= -rhs) of the current
 * state; no reference counterpart (the reference has no adjoint).
 * GMC_VJP_EXACT: g_state = v^T dR/dI, g_prev = v^T dR/dP (may be NULL) and
 * g_params = v^T dR/dtheta (6 entries, may be NULL) for theta = (rho0, mu, cp)
 * of components 0 and 1, in the order of gmc_set_material (rho0 of an ideal gas
 * is its specific gas constant R, since rho = p/(RT), chemicals.hpp:260, and
 * its gamma = cp/(cp - R)).  R is differentiated with every coefficient
 * (properties, flux matrices, tau with its LU inverse and min() switches,
 * compute_h, dc_2006 nu, u_ref); at a switch the branch taken by the value is
 * differentiated.  Rows of Dirichlet-blocked dofs are not part of R (the
 * assembled rhs holds 0 there), so v on them is ignored.
 * GMC_VJP_FROZEN: g_state = J^T v with J exactly the matrix gmc_assemble builds
 * in FP64 (Dirichlet rows: avg_val on the diagonal), formed matrix-free;
 * g_prev and g_params must be NULL.  Host pointers (n_nodes*ndof doubles). */
int gmc_vjp(gmc_ctx* ctx, const double* v, double* g_state, double* g_prev,
            double* g_params, int mode);
/* Device-pointer variant (v, g_state, g_prev of n_nodes*ndof doubles; g_params
 * of 6 doubles; all in device memory, ordered on the context's stream). */
int gmc_vjp_device(gmc_ctx* ctx, const double* v, double* g_state, double* g_prev,
                   double* g_params, int mode);
/* R(I, P) = -rhs with I, P and R in device memory (I or P NULL: keep the
 * context's state); the forward of a differentiable residual binding
 * (fluid_assembly::execute's rhs, fluid_assembly.hpp:39-68, residual only). */
int gmc_residual_device(gmc_ctx* ctx, const double* I, const double* P, double* R);

#ifdef __cplusplus
}
#endif

#endif /* GALES_MC_H_ */
