// gales_mc.hpp -- C++ host classes with the call surface of the GALES-multi-fluid
// fluid_mc assembly path, over the C-ABI of libgales_mc.so (include/gales_mc.h).
//
// A GALES driver (sim/fluid_mc/P0/main.cpp) builds a model, an integral, a
// scatter, a linear_system, fluid_assembly, residual_check, a Belos solver and
// fluid_updater, then runs the time/Newton loop of main.cpp:206-267.  This
// header keeps those class names and member calls so the loop reads the same:
//
//   fluid_updater.predictor();                                   // main.cpp:213
//   double fill_time = fluid_assembly.execute();                 // main.cpp:222
//   bool conv = fluid_res_check.execute(fluid_lp, it, res);      // main.cpp:225
//   double solver_time = fluid_ls_solver.execute();              // main.cpp:231
//   fluid_updater.corrector(fluid_ls_solver.solution_device());  // main.cpp:234
//   fluid_model.dof_trimmer(4, 0, 1);                            // main.cpp:235
//
// Everything numerical runs on the GPU; host copies are made only when a
// caller asks for one (matrix(), rhs(), solution(), dofs()).  Errors throw
// gales_mc::error (status code + message) where the reference calls Error()
// -> MPI_Abort (src/fem/base/algorithm.hpp:37-42).  As in the reference, the
// classes hold their collaborators by reference and are neither copyable nor
// movable; one model is bound to one HIP stream and is not thread-safe.
#pragma once

#include <chrono>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gales_mc.h"

namespace gales_mc {

// boost::numeric::ublas::vector<double> / matrix<double> of the reference
// signatures (fluid_2d.hpp:26-27), reduced to what the path uses.
using vec = std::vector<double>;

class mat {
 public:
  mat() = default;
  mat(std::size_t n1, std::size_t n2, double v = 0.0) : n1_(n1), n2_(n2), a_(n1 * n2, v) {}
  double& operator()(std::size_t i, std::size_t j) { return a_[i * n2_ + j]; }
  double operator()(std::size_t i, std::size_t j) const { return a_[i * n2_ + j]; }
  std::size_t size1() const { return n1_; }
  std::size_t size2() const { return n2_; }
  void clear() { a_.assign(a_.size(), 0.0); }
  double* data() { return a_.data(); }
  const double* data() const { return a_.data(); }

 private:
  std::size_t n1_ = 0, n2_ = 0;
  std::vector<double> a_;
};

class error : public std::runtime_error {
 public:
  error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int status, const gmc_ctx* ctx) {
  if (status != GMC_OK) throw error(status, gmc_last_error(ctx));
}

#define GALES_MC_NO_COPY(T)        \
  T(const T&) = delete;            \
  T& operator=(const T&) = delete; \
  T(T&&) = delete;                 \
  T& operator=(T&&) = delete;

// model<dim> + dof_state + maps (src/fem/dofs/model.hpp:40-184,
// dof_state.hpp:17-49): the mesh, the two state slots (0 = I, 1 = P) and the
// device context that owns them.
class model {
 public:
  model(const gmc_mesh& mesh, const gmc_params& params) {
    check(gmc_create(&mesh, &params, &ctx_), nullptr);
    check(gmc_get_info(ctx_, &info_), ctx_);
  }
  ~model() { gmc_destroy(ctx_); }
  GALES_MC_NO_COPY(model)

  gmc_ctx* ctx() const { return ctx_; }
  const gmc_info& info() const { return info_; }
  int nb_dofs() const { return info_.ndof; }
  std::size_t state_size() const { return (std::size_t)info_.n_nodes * info_.ndof; }

  // dof_state::dofs(slot) as host copies; set_dofs uploads a whole slot.
  vec dofs(int slot) const {
    vec I(state_size()), P(state_size());
    check(gmc_get_state(ctx_, I.data(), P.data()), ctx_);
    return slot == 0 ? I : P;
  }
  void set_dofs(int slot, const vec& v) {
    if (v.size() != state_size()) throw error(GMC_ERR_ARG, "state size mismatch");
    check(gmc_set_state(ctx_, slot == 0 ? v.data() : nullptr, slot == 1 ? v.data() : nullptr), ctx_);
  }

  // model::dof_trimmer (model.hpp:176-184): clamp dof index k of every node
  // to [lo, hi] (main.cpp:235 trims the first mass fraction to [0, 1]).
  void dof_trimmer(int dof_index, double lo, double hi) {
    check(gmc_dof_trimmer(ctx_, dof_index, lo, hi), ctx_);
  }

  // time::get().delta_t() (src/fem/base/time.hpp:47-51) as read by the integral.
  void set_delta_t(double dt) { check(gmc_set_dt(ctx_, dt), ctx_); }

 private:
  gmc_ctx* ctx_ = nullptr;
  gmc_info info_{};
};

// fluid<ic_bc_type, dim> (sim/fluid_mc/P0/fluid_2d.hpp:22-245): the element
// integral.  The library evaluates every element on the device inside
// fluid_assembly::execute; execute_i / execute_b here run one element on the
// device from the model's current state (the dofs of model::extract_element_dofs)
// and return m (nen*ndof square) and r = -r_ (fluid_2d.hpp:200-201), without
// the Dirichlet rule, for tests and single-element checks.
class fluid {
 public:
  explicit fluid(model& m) : model_(m) {}
  GALES_MC_NO_COPY(fluid)

  void execute_i(int el, mat& m, vec& r) { run(el, 0, m, r); }
  void execute_b(int el, mat& m, vec& r) { run(el, 1, m, r); }

 private:
  void run(int el, int boundary, mat& m, vec& r) {
    const std::size_t ne = (std::size_t)model_.info().nen * model_.info().ndof;
    if (m.size1() != ne || m.size2() != ne) m = mat(ne, ne);
    r.assign(ne, 0.0);
    check(gmc_element(model_.ctx(), el, boundary, m.data(), r.data()), model_.ctx());
  }
  model& model_;
};

// scatter<dofs_ic_bc_type> (src/fem/ls/scatter.hpp:37-98): the Dirichlet
// avg_val rule and the sum into the global matrix run inside the device
// assembly, so the scatter object carries no work of its own.
struct scatter {};

// Host copy of the assembled block-CSR matrix (ndof x ndof row-major blocks);
// the same matrix as the reference's point CSR (linear_system.hpp:40-66).
struct bsr_matrix {
  int n_block_rows = 0, block_size = 0;
  std::vector<int> rowptr, colidx;
  std::vector<double> vals;
};

// linear_system (src/fem/ls/linear_system.hpp:40-100): J and rhs = -R.  The
// device assembly overwrites both, so clear() and assembled() have nothing to
// do; they are kept for the call surface.
class linear_system {
 public:
  explicit linear_system(model& m) : model_(m) {}
  GALES_MC_NO_COPY(linear_system)

  void clear() {}
  void assembled() {}
  bsr_matrix matrix() const {
    const auto& in = model_.info();
    bsr_matrix a;
    a.n_block_rows = in.n_nodes;
    a.block_size = in.ndof;
    a.rowptr.resize(in.n_nodes + 1);
    a.colidx.resize((std::size_t)in.nnzb);
    a.vals.resize((std::size_t)in.nnzb * in.ndof * in.ndof);
    check(gmc_get_pattern(model_.ctx(), a.rowptr.data(), a.colidx.data()), model_.ctx());
    check(gmc_get_system(model_.ctx(), a.vals.data(), nullptr), model_.ctx());
    return a;
  }
  vec rhs() const {
    vec b(model_.state_size());
    check(gmc_get_system(model_.ctx(), nullptr, b.data()), model_.ctx());
    return b;
  }
  model& owner() const { return model_; }

 private:
  model& model_;
};

// fluid_assembly<integral_type, scatter_type, dim> (fluid_assembly.hpp:15-68):
// execute() clears, integrates every interior and boundary element, applies
// the Dirichlet rule and sums into the linear system; returns wall seconds.
template <typename integral_type = fluid, typename scatter_type = scatter>
class fluid_assembly {
 public:
  fluid_assembly(integral_type& integral, scatter_type& sct, model& m, linear_system& lp)
      : integral_(integral), scatter_(sct), model_(m), lp_(lp) {}
  GALES_MC_NO_COPY(fluid_assembly)

  double execute(int flags = GMC_ASM_MATRIX | GMC_ASM_RHS) {
    const auto t0 = std::chrono::steady_clock::now();
    lp_.clear();
    check(gmc_assemble(model_.ctx(), flags, nullptr), model_.ctx());
    lp_.assembled();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

 private:
  integral_type& integral_;
  scatter_type& scatter_;
  model& model_;
  linear_system& lp_;
};

// residual_check (src/fem/ls/residual_check.hpp:42-80): ||rhs||_2; at
// it_count 0 it stores the norm and returns false, later it returns
// ||rhs|| <= max(1e-8 ||rhs_0||, 1e-5) with non_linear_res = ||rhs|| / ||rhs_0||.
class residual_check {
 public:
  residual_check() = default;
  GALES_MC_NO_COPY(residual_check)

  bool execute(const linear_system& lp, int it_count, double& non_linear_res) {
    double nrm = 0.0;
    check(gmc_rhs_norm(lp.owner().ctx(), &nrm), lp.owner().ctx());
    if (it_count == 0) {
      res0_ = nrm;
      non_linear_res = nrm;
      return false;
    }
    non_linear_res = nrm / res0_;
    return nrm <= std::max(1.e-8 * res0_, 1.e-5);
  }

 private:
  double res0_ = 0.0;
};

// belos_linear_solver (src/fem/ls/belos_linear_solver_factory.hpp:103-248)
// with the Krylov method replaced by restarted GMRES(k) + point or block
// Jacobi on the device: right preconditioning, x0 = 0, true residuals for
// RRE()/ARE().
class linear_solver {
 public:
  linear_solver(linear_system& lp, const gmc_solver& s) : lp_(lp), s_(s) {}
  GALES_MC_NO_COPY(linear_solver)

  double execute() {
    const auto t0 = std::chrono::steady_clock::now();
    check(gmc_solve(lp_.owner().ctx(), &s_, &info_), lp_.owner().ctx());
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  vec solution() const {
    vec x(lp_.owner().state_size());
    check(gmc_get_solution(lp_.owner().ctx(), x.data()), lp_.owner().ctx());
    return x;
  }
  // the solution stays on the device; fluid_updater::corrector consumes it there
  const linear_solver& solution_device() const { return *this; }
  double RRE() const { return info_.rel_residual; }
  double ARE() const { return info_.abs_residual; }
  int num_it() const { return info_.iterations; }
  model& owner() const { return lp_.owner(); }

 private:
  linear_system& lp_;
  gmc_solver s_;
  gmc_solve_info info_{};
};

// fluid_updater (sim/fluid_mc/P0/fluid_updater.hpp:18-28):
// predictor P <- I, corrector I <- I + correction.
class fluid_updater {
 public:
  explicit fluid_updater(model& m) : model_(m) {}
  GALES_MC_NO_COPY(fluid_updater)

  void predictor() { check(gmc_predictor(model_.ctx()), model_.ctx()); }
  void corrector(const linear_solver&) { check(gmc_corrector(model_.ctx()), model_.ctx()); }

 private:
  model& model_;
};

#undef GALES_MC_NO_COPY

// p_interpolation (src/fem/utils/pressure_profile.hpp:26-41): linear
// interpolation of a pressure table (y ascending or descending, e.g. the one
// gmc_hydrostatic_profile returns) at height y; throws where the reference
// calls Error() (y outside the table).
inline double pressure_from_profile(const std::vector<double>& y_tab, const std::vector<double>& p_tab,
                                    double y) {
  const std::size_t n = y_tab.size();
  if (n < 2 || p_tab.size() != n) throw error(GMC_ERR_ARG, "pressure table needs two or more points");
  const bool asc = y_tab.front() < y_tab.back();
  std::size_t i = 0;
  while (i < n && (asc ? (y_tab[i] < y) : (y_tab[i] > y))) i++;   // std::lower_bound
  if (i > 0) i--;
  if (i >= n - 1) throw error(GMC_ERR_ARG, "index i goes out of the vector scope");
  const double f = (y - y_tab[i]) / (y_tab[i + 1] - y_tab[i]);
  return f * p_tab[i + 1] + (1.0 - f) * p_tab[i];
}

}  // namespace gales_mc
