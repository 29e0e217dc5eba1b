// OBJ input/output and the reference mesh helpers shared by the CLI
// drop-ins (rigid_deform, cad_deform).
#pragma once

#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace mo_cli {

// mo_lm_stats.solver_path, spelled for the "LM:" line; a solve that fell
// back from a refused cooperative launch is also reported on stderr.
inline const char* solver_path_name(int p) {
  switch (p) {
    case 0: return "lds";
    case 1: return "one-workgroup";
    case 2: return "team";
    case 3: return "one-workgroup (cooperative launch refused)";
    case 4: return "lds-team";
    default: return "unknown";
  }
}

inline void warn_if_degraded(int p) {
  if (p == 3)
    std::cerr << "warning: the runtime refused the multi-workgroup (cooperative) launch; the solve ran on one "
                 "workgroup per problem (slower; low-bit differences from the team result)"
              << std::endl;
}

struct Mesh {
  std::vector<double> V;   // [n,3]
  std::vector<int> F;      // [m,3]
  size_t nv() const { return V.size() / 3; }
  size_t nf() const { return F.size() / 3; }
};

// Mesh::ReadOBJ (mesh.cc:14-44): 256-byte lines, first token selects the
// record, "f" keeps the digits before '/' of three tokens (a missing token
// reuses the previous one), other records are ignored.
inline bool read_obj(const char* path, Mesh* m) {
  std::ifstream is(path);
  if (!is) return false;
  char buffer[256];
  while (is.getline(buffer, 256)) {
    std::istringstream str(buffer);
    std::string tok;
    if (!(str >> tok)) continue;
    if (tok == "v") {
      double x = 0, y = 0, z = 0;
      str >> x >> y >> z;
      m->V.push_back(x);
      m->V.push_back(y);
      m->V.push_back(z);
    } else if (tok == "f") {
      std::string cur = tok;
      for (int j = 0; j < 3; ++j) {
        std::string t;
        if (str >> t) cur = t;
        int id = 0;
        for (char c : cur) {
          if (c == '/') break;
          id = id * 10 + (c - '0');
        }
        m->F.push_back(id - 1);
      }
    }
  }
  return true;
}

// Mesh::ReflectionSymmetrize (mesh.cc:234-246).
inline void reflection_symmetrize(Mesh* m) {
  const size_t n = m->nv(), f = m->nf();
  for (size_t i = 0; i < n; ++i) {
    m->V.push_back(-m->V[i * 3]);
    m->V.push_back(m->V[i * 3 + 1]);
    m->V.push_back(m->V[i * 3 + 2]);
  }
  for (size_t i = 0; i < f * 3; ++i) m->F.push_back(m->F[i] + (int)n);
}

// Mesh::WriteOBJ (mesh.cc:46-64); V is already in the output frame.
inline void write_obj(const char* path, const std::vector<double>& V, const std::vector<int>& F) {
  std::ofstream os(path);
  for (size_t i = 0; i < V.size() / 3; ++i)
    os << "v " << V[i * 3] << " " << V[i * 3 + 1] << " " << V[i * 3 + 2] << "\n";
  for (size_t i = 0; i < F.size() / 3; ++i)
    os << "f " << F[i * 3] + 1 << " " << F[i * 3 + 1] + 1 << " " << F[i * 3 + 2] + 1 << "\n";
}

}  // namespace mo_cli
