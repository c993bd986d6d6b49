// fluid_mc_driver.cpp -- the time / Newton loop of a GALES fluid_mc case
// (sim/fluid_mc/P0/main.cpp:206-267) written against the host classes of
// gales_mc.hpp, which reach the MI355X kernels through the C-ABI.
//
//   fluid_mc_driver CASE_FILE N_STEPS N_MAX_IT RESTART MAX_ITERS REL_TOL PRECOND OUT_STATE [KEEP_P]
//
// CASE_FILE holds the mesh, the gmc_params struct and the I and P states
// (written by gales_mc.casefile.write_case).  Every step starts with
// fluid_updater.predictor() (P <- I), as main.cpp:212 does, the first one
// included; KEEP_P = 1 skips the first step's predictor so that the case's P
// is the previous-step state (a restart that carries both states).  Per Newton iteration it prints
// the line of main.cpp:238-242 (it, assembly s, non-linear residual, solver s,
// ARE, RRE, iterations, dt) and, at the end, writes the I state as raw
// float64 in global dof order (node * ndof + k) to OUT_STATE.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "gales_mc.hpp"

namespace {

struct case_data {
  gmc_params params{};
  int dim = 0, n_nodes = 0, n_elems = 0, n_sides = 0;
  std::vector<double> coords, I, P;
  std::vector<int> conn, node_flag, side_elem, side_local, side_flag;
};

template <typename T>
void read_into(std::ifstream& f, std::vector<T>& v, std::size_t n) {
  v.resize(n);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
}

case_data read_case(const char* path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw gales_mc::error(GMC_ERR_ARG, std::string("cannot open ") + path);
  char magic[8];
  f.read(magic, 8);
  if (std::memcmp(magic, "GMCCASE1", 8) != 0) throw gales_mc::error(GMC_ERR_ARG, "not a case file");
  int32_t hdr[5];
  f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
  case_data c;
  c.dim = hdr[0];
  c.n_nodes = hdr[1];
  c.n_elems = hdr[2];
  c.n_sides = hdr[3];
  if (hdr[4] != (int32_t)sizeof(gmc_params))
    throw gales_mc::error(GMC_ERR_ARG, "gmc_params size differs from the case file's");
  f.read(reinterpret_cast<char*>(&c.params), sizeof(gmc_params));
  const int nen = c.dim + 1, ndof = c.dim + 2 + c.params.nc - 1;
  read_into(f, c.coords, (std::size_t)c.n_nodes * c.dim);
  read_into(f, c.conn, (std::size_t)c.n_elems * nen);
  read_into(f, c.node_flag, (std::size_t)c.n_nodes);
  read_into(f, c.side_elem, (std::size_t)c.n_sides);
  read_into(f, c.side_local, (std::size_t)c.n_sides);
  read_into(f, c.side_flag, (std::size_t)c.n_sides);
  read_into(f, c.I, (std::size_t)c.n_nodes * ndof);
  read_into(f, c.P, (std::size_t)c.n_nodes * ndof);
  if (!f) throw gales_mc::error(GMC_ERR_ARG, "truncated case file");
  return c;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 9 && argc != 10) {
    std::cerr << "usage: " << argv[0]
              << " CASE_FILE N_STEPS N_MAX_IT RESTART MAX_ITERS REL_TOL PRECOND OUT_STATE [KEEP_P]\n";
    return 2;
  }
  const bool keep_p = argc == 10 && std::atoi(argv[9]) != 0;
  try {
    case_data cs = read_case(argv[1]);
    const int n_steps = std::atoi(argv[2]), n_max_it = std::atoi(argv[3]);
    gmc_solver sv{std::atoi(argv[4]), std::atoi(argv[5]), std::atof(argv[6]), std::atoi(argv[7])};

    gmc_mesh mesh{cs.dim,          cs.n_nodes,        cs.n_elems,          cs.n_sides,
                  cs.coords.data(), cs.conn.data(),    cs.node_flag.data(), cs.side_elem.data(),
                  cs.side_local.data(), cs.side_flag.data()};

    // the objects of main.cpp:73-145
    gales_mc::model fluid_model(mesh, cs.params);
    fluid_model.set_dofs(0, cs.I);
    fluid_model.set_dofs(1, cs.P);
    gales_mc::fluid fluid_integral(fluid_model);
    gales_mc::scatter fluid_scatter;
    gales_mc::linear_system fluid_lp(fluid_model);
    gales_mc::fluid_assembly<> fluid_assembly(fluid_integral, fluid_scatter, fluid_model, fluid_lp);
    gales_mc::residual_check fluid_res_check;
    gales_mc::linear_solver fluid_ls_solver(fluid_lp, sv);
    gales_mc::fluid_updater fluid_updater(fluid_model);
    const int y_dof = cs.dim + 2;  // first mass fraction (main.cpp:235 trims dof 4 in 2D)

    // the time loop of main.cpp:206-267
    double non_linear_res = 0.0;
    for (int step = 0; step < n_steps; step++) {
      if (step > 0 || !keep_p) fluid_updater.predictor();
      for (int it_count = 0; it_count < n_max_it; it_count++) {
        const double fill_time = fluid_assembly.execute();
        const bool converged = fluid_res_check.execute(fluid_lp, it_count, non_linear_res);
        if (converged) {
          std::printf("FLUID  it: %5d Assembly: %15.6e NonLinearRes: %15.6e\n", it_count, fill_time,
                      non_linear_res);
          break;
        }
        const double solver_time = fluid_ls_solver.execute();
        fluid_updater.corrector(fluid_ls_solver.solution_device());
        fluid_model.dof_trimmer(y_dof, 0.0, 1.0);
        std::printf(
            "FLUID  it: %5d Assembly: %15.6e NonLinearRes: %15.6e Solver: %15.6e AbsResErr: %15.6e "
            "RelResErr: %15.6e Num_it: %8d dt: %g\n",
            it_count, fill_time, non_linear_res, solver_time, fluid_ls_solver.ARE(),
            fluid_ls_solver.RRE(), fluid_ls_solver.num_it(), cs.params.dt);
      }
    }
    const gales_mc::vec I = fluid_model.dofs(0);
    std::ofstream out(argv[8], std::ios::binary);
    out.write(reinterpret_cast<const char*>(I.data()), (std::streamsize)(I.size() * sizeof(double)));
    if (!out) throw gales_mc::error(GMC_ERR_ARG, "cannot write the output state");
    std::fflush(stdout);
  } catch (const gales_mc::error& e) {
    std::cerr << "fluid_mc_driver: " << e.what() << " (status " << e.code() << ")\n";
    return 1;
  }
  return 0;
}
