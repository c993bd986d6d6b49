// gales_mc_io.hpp -- the reference's on-disk formats for a host program that
// drives the C-ABI library:
//
//  - the text mesh of src/fem/mesh/mesh.hpp:141-456:
//        MESH! 2D                       (or 3D)
//        nodes <N>
//        elements <E>
//        sides <S>
//        X <min> <max>                  (one line per dimension)
//        Y <min> <max>
//        Node <gid> <x> <y> [<z>] <flag>                  N lines
//        Element <gid> <pid> <nnodes> <n_0> ... <flag>    E lines (flag > 0: on the boundary)
//        Side <gid> <pid> <nsn> <n_0> ... <flag>          S lines
//    tokens separated by spaces or commas (algorithm.hpp:219-225).  A side is
//    matched to the local side of a boundary element whose node set equals its
//    own (mesh.hpp:420-446, element.hpp:145-149 / 455-459 numbering); sides
//    with flag 0 carry no boundary integral.  Reading one partition (pid >= 0)
//    keeps that pid's elements and renumbers their nodes compactly, the gid of
//    each local node in node_gid;
//
//  - the restart dumps of src/fem/base/io.hpp:84-181: the global state vector
//    (dof gid = node gid * ndof + k) as raw native float64, in the file
//    "results/<dir><t>" with t printed as an ostream prints a double.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "gales_mc.hpp"

namespace gales_mc {
namespace io {

struct text_mesh {
  int dim = 0;
  std::vector<double> coords;                  // [n_nodes * dim]
  std::vector<int> conn;                       // [n_elems * (dim + 1)], local node ids
  std::vector<int> node_flag;                  // [n_nodes]
  std::vector<int> node_gid;                   // [n_nodes]: gid in the file
  std::vector<int> elem_gid, elem_pid;         // [n_elems]
  std::vector<int> side_elem, side_local, side_flag;
  int n_nodes() const { return (int)node_flag.size(); }
  int n_elems() const { return (int)elem_gid.size(); }
  int n_sides() const { return (int)side_elem.size(); }
  // the C-ABI view (arrays stay owned by this object)
  gmc_mesh view() const {
    return gmc_mesh{dim,          n_nodes(),          n_elems(),        n_sides(),
                    coords.data(), conn.data(),       node_flag.data(), side_elem.data(),
                    side_local.data(), side_flag.data()};
  }
};

// local side tables (element.hpp:145-149, 455-459)
inline const std::vector<std::vector<int>>& side_table(int dim) {
  static const std::vector<std::vector<int>> tri = {{1, 2}, {2, 0}, {0, 1}};
  static const std::vector<std::vector<int>> tet = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
  return dim == 2 ? tri : tet;
}

inline std::vector<std::string> tokens(const std::string& line) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : line) {
    if (ch == ' ' || ch == ',' || ch == '\t' || ch == '\r' || ch == '\n') {
      if (!cur.empty()) out.push_back(cur), cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

inline std::vector<std::string> next_line(std::ifstream& f, const char* what) {
  std::string s;
  while (std::getline(f, s)) {
    auto t = tokens(s);
    if (!t.empty()) return t;
  }
  throw error(GMC_ERR_ARG, std::string("mesh file ends before ") + what);
}

inline int to_int(const std::string& s, const char* what) {
  try {
    size_t pos = 0;
    const int v = std::stoi(s, &pos);
    if (pos == s.size()) return v;
  } catch (...) {
  }
  throw error(GMC_ERR_ARG, std::string("mesh file: bad integer '") + s + "' in " + what);
}

inline double to_double(const std::string& s, const char* what) {
  try {
    size_t pos = 0;
    const double v = std::stod(s, &pos);
    if (pos == s.size()) return v;
  } catch (...) {
  }
  throw error(GMC_ERR_ARG, std::string("mesh file: bad number '") + s + "' in " + what);
}

// Reads a text mesh; pid < 0 reads every element, else the elements of that
// partition.
inline text_mesh read_mesh(const std::string& path, int pid = -1) {
  std::ifstream f(path);
  if (!f) throw error(GMC_ERR_ARG, "cannot open mesh file " + path);
  auto t = next_line(f, "the header");
  if (t.size() < 2 || t[0] != "MESH!" || (t[1] != "2D" && t[1] != "3D"))
    throw error(GMC_ERR_ARG, "not a GALES mesh file (first line 'MESH! 2D' or 'MESH! 3D')");
  text_mesh m;
  m.dim = t[1] == "2D" ? 2 : 3;
  const int nen = m.dim + 1;
  int counts[3];
  const char* names[3] = {"nodes", "elements", "sides"};
  for (int k = 0; k < 3; k++) {
    t = next_line(f, names[k]);
    if (t.size() < 2 || t[0] != names[k]) throw error(GMC_ERR_ARG, std::string("expected '") + names[k] + " <n>'");
    counts[k] = to_int(t[1], names[k]);
    if (counts[k] < 0) throw error(GMC_ERR_ARG, "negative count in the header");
  }
  for (int k = 0; k < m.dim; k++) next_line(f, "the extent lines");
  const int N = counts[0], E = counts[1], S = counts[2];
  std::vector<double> X((size_t)N * m.dim);
  std::vector<int> flag(N), seen(N, 0);
  for (int i = 0; i < N; i++) {
    t = next_line(f, "a Node line");
    if (t[0] != "Node" || (int)t.size() < m.dim + 3) throw error(GMC_ERR_ARG, "expected 'Node gid x y [z] flag'");
    const int g = to_int(t[1], "Node");
    if (g < 0 || g >= N || seen[g]) throw error(GMC_ERR_ARG, "Node gid out of range or repeated");
    seen[g] = 1;
    for (int k = 0; k < m.dim; k++) X[(size_t)g * m.dim + k] = to_double(t[2 + k], "Node");
    flag[g] = to_int(t[2 + m.dim], "Node");
  }
  std::vector<int> econn, egid, epid, eflag;
  for (int e = 0; e < E; e++) {
    t = next_line(f, "an Element line");
    if (t[0] != "Element" || t.size() < 4) throw error(GMC_ERR_ARG, "expected 'Element gid pid n nodes... flag'");
    const int nn = to_int(t[3], "Element");
    if (nn != nen || (int)t.size() < 5 + nn)
      throw error(GMC_ERR_ARG, "Element with " + std::to_string(nn) + " nodes (linear triangles / tetrahedra only)");
    const int p = to_int(t[2], "Element");
    if (pid >= 0 && p != pid) continue;
    egid.push_back(to_int(t[1], "Element"));
    epid.push_back(p);
    for (int a = 0; a < nn; a++) {
      const int g = to_int(t[4 + a], "Element");
      if (g < 0 || g >= N) throw error(GMC_ERR_ARG, "Element node out of range");
      econn.push_back(g);
    }
    eflag.push_back(to_int(t[4 + nn], "Element"));
  }
  // compact node numbering (identity when every element is read)
  std::vector<int> local(N, -1);
  if (pid < 0) {
    for (int g = 0; g < N; g++) local[g] = g;
  } else {
    std::vector<int> used(econn);
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    for (size_t k = 0; k < used.size(); k++) local[used[k]] = (int)k;
  }
  for (int g = 0; g < N; g++) {
    if (local[g] < 0) continue;
    m.node_gid.push_back(g);
  }
  m.coords.resize(m.node_gid.size() * m.dim);
  m.node_flag.resize(m.node_gid.size());
  for (size_t k = 0; k < m.node_gid.size(); k++) {
    const int g = m.node_gid[k];
    for (int d = 0; d < m.dim; d++) m.coords[k * m.dim + d] = X[(size_t)g * m.dim + d];
    m.node_flag[k] = flag[g];
  }
  m.conn.resize(econn.size());
  for (size_t k = 0; k < econn.size(); k++) m.conn[k] = local[econn[k]];
  m.elem_gid = egid;
  m.elem_pid = epid;
  // sides: node sets -> (element, local side) among boundary elements whose
  // side nodes all have a flag > 0 (mesh.hpp:428-442)
  std::map<std::vector<int>, std::pair<int, int>> where;
  const auto& tab = side_table(m.dim);
  for (int e = 0; e < m.n_elems(); e++) {
    if (eflag[e] <= 0) continue;
    for (int s = 0; s < (int)tab.size(); s++) {
      std::vector<int> key;
      bool bd = true;
      for (int a : tab[s]) {
        const int g = econn[(size_t)e * nen + a];
        key.push_back(g);
        bd = bd && flag[g] > 0;
      }
      if (!bd) continue;
      std::sort(key.begin(), key.end());
      where.emplace(key, std::make_pair(e, s));
    }
  }
  for (int k = 0; k < S; k++) {
    t = next_line(f, "a Side line");
    if (t[0] != "Side" || t.size() < 4) throw error(GMC_ERR_ARG, "expected 'Side gid pid n nodes... flag'");
    const int nsn = to_int(t[3], "Side");
    if (nsn != m.dim || (int)t.size() < 5 + nsn) throw error(GMC_ERR_ARG, "Side with the wrong node count");
    if (pid >= 0 && to_int(t[2], "Side") != pid) continue;
    std::vector<int> key(nsn);
    for (int a = 0; a < nsn; a++) key[a] = to_int(t[4 + a], "Side");
    const int sflag = to_int(t[4 + nsn], "Side");
    std::sort(key.begin(), key.end());
    auto it = where.find(key);
    if (it == where.end() || sflag <= 0) continue;   // no boundary element carries it
    m.side_elem.push_back(it->second.first);
    m.side_local.push_back(it->second.second);
    m.side_flag.push_back(sflag);
  }
  return m;
}

// Writes a text mesh read back by read_mesh (pid of every element 0 unless
// given; an element is flagged when it carries a side).
inline void write_mesh(const std::string& path, const gmc_mesh& m, const std::vector<int>& elem_pid = {}) {
  std::ofstream f(path);
  if (!f) throw error(GMC_ERR_ARG, "cannot write mesh file " + path);
  const int nen = m.dim + 1;
  f << "MESH! " << m.dim << "D\n";
  f << "nodes " << m.n_nodes << "\nelements " << m.n_elems << "\nsides " << m.n_sides << "\n";
  const char* axes = "XYZ";
  for (int d = 0; d < m.dim; d++) {
    double lo = m.coords[d], hi = m.coords[d];
    for (int i = 0; i < m.n_nodes; i++) {
      lo = std::min(lo, m.coords[(size_t)i * m.dim + d]);
      hi = std::max(hi, m.coords[(size_t)i * m.dim + d]);
    }
    f << axes[d] << " " << lo << " " << hi << "\n";
  }
  f.precision(17);
  for (int i = 0; i < m.n_nodes; i++) {
    f << "Node " << i;
    for (int d = 0; d < m.dim; d++) f << " " << m.coords[(size_t)i * m.dim + d];
    f << " " << m.node_flag[i] << "\n";
  }
  std::vector<int> on_bd(m.n_elems, 0);
  for (int s = 0; s < m.n_sides; s++) on_bd[m.side_elem[s]] = 1;
  for (int e = 0; e < m.n_elems; e++) {
    f << "Element " << e << " " << (elem_pid.empty() ? 0 : elem_pid[e]) << " " << nen;
    for (int a = 0; a < nen; a++) f << " " << m.conn[(size_t)e * nen + a];
    f << " " << on_bd[e] << "\n";
  }
  const auto& tab = side_table(m.dim);
  for (int s = 0; s < m.n_sides; s++) {
    const int e = m.side_elem[s];
    f << "Side " << s << " " << (elem_pid.empty() ? 0 : elem_pid[e]) << " " << m.dim;
    for (int a : tab[m.side_local[s]]) f << " " << m.conn[(size_t)e * nen + a];
    f << " " << m.side_flag[s] << "\n";
  }
  if (!f) throw error(GMC_ERR_HIP, "write failed: " + path);
}

// "results/<dir><t>" with t as an ostream prints a double (io.hpp:112-113)
inline std::string state_file(const std::string& dir, double t) {
  std::ostringstream ss;
  ss << "results/" << dir << t;
  return ss.str();
}

inline void write_state(const std::string& path, const double* v, long long n) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw error(GMC_ERR_ARG, "cannot write state file " + path);
  f.write(reinterpret_cast<const char*>(v), (std::streamsize)(n * (long long)sizeof(double)));
  if (!f) throw error(GMC_ERR_HIP, "write failed: " + path);
}

inline void read_state(const std::string& path, double* v, long long n) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw error(GMC_ERR_ARG, "the file '" + path + "' to read does not exist");
  const long long bytes = (long long)f.tellg();
  if (bytes != n * (long long)sizeof(double))
    throw error(GMC_ERR_ARG, path + " holds " + std::to_string(bytes) + " bytes, expected " +
                                 std::to_string(n * (long long)sizeof(double)));
  f.seekg(0);
  f.read(reinterpret_cast<char*>(v), (std::streamsize)bytes);
}

}  // namespace io
}  // namespace gales_mc
