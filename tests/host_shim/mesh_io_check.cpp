// Test helper: reads a GALES text mesh with gales_mc::io::read_mesh and dumps
// the arrays for tests/test_meshio.py to compare with the Python reader.
//   mesh_io_check MESH.txt PID OUT.bin REWRITTEN.txt
// OUT.bin: int32 dim, n_nodes, n_elems, n_sides, then float64 coords and
// int32 conn, node_flag, node_gid, side_elem, side_local, side_flag.
// REWRITTEN.txt: the mesh written back by write_mesh.  Also prints the
// restart file names of a few times, one per line.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "gales_mc_io.hpp"

template <typename T>
static void put(std::ofstream& f, const std::vector<T>& v) {
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  try {
    const auto m = gales_mc::io::read_mesh(argv[1], std::atoi(argv[2]));
    std::ofstream f(argv[3], std::ios::binary);
    const int32_t hdr[4] = {m.dim, m.n_nodes(), m.n_elems(), m.n_sides()};
    f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    put(f, m.coords);
    put(f, m.conn);
    put(f, m.node_flag);
    put(f, m.node_gid);
    put(f, m.side_elem);
    put(f, m.side_local);
    put(f, m.side_flag);
    gales_mc::io::write_mesh(argv[4], m.view(), m.elem_pid);
    for (double t : {0.0, 0.01, 1.0 / 3.0, 2.5, 1e-5, 123456789.0})
      std::cout << gales_mc::io::state_file("fluid_dofs/", t) << "\n";
    std::vector<double> s = {1.5, -2.25, 3e300};
    gales_mc::io::write_state(std::string(argv[3]) + ".state", s.data(), 3);
    std::vector<double> back(3);
    gales_mc::io::read_state(std::string(argv[3]) + ".state", back.data(), 3);
    if (back != s) return 3;
    try {
      gales_mc::io::read_state(std::string(argv[3]) + ".state", back.data(), 4);
      return 4;
    } catch (const gales_mc::error&) {
    }
  } catch (const gales_mc::error& e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
  return 0;
}
