// CAD preprocessing chain of cad_deform (host code): mesh clean-up, the
// Delaunay subdivision, geometry neighbours and the representative graph.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace mo {
namespace cad {

struct Mesh {
  std::vector<double> V;  // [n,3]
  std::vector<int> F;     // [m,3]
  size_t nv() const { return V.size() / 3; }
  size_t nf() const { return F.size() / 3; }
};

// Mesh::RemoveDegenerated (mesh.cc:320-333): drop faces whose normal has
// zero norm.
void remove_degenerated(Mesh* m);

// Mesh::MergeDuplex (mesh.cc:181-232): merge vertices with equal truncated
// 1e-6-scaled coordinates (first occurrence kept), then drop faces with a
// repeated vertex and faces whose sorted vertex triple was already seen.
void merge_duplex(Mesh* m);

struct Subdivision {
  Mesh mesh;                                   // subdivided mesh
  std::vector<char> internal;                  // lattice (interior) vertices
  std::vector<std::pair<int, int>> neighbors;  // geometry neighbour pairs (sorted, unique)
  std::vector<int> ref;                        // vertex -> representative vertex
  std::vector<double> graph_v;                 // representative vertices [g,3]
  std::vector<std::pair<int, int>> graph_e;    // representative edges (sorted, unique)
};

// Subdivision::Subdivide (subdivision.cc:28-129) with DelaunaySubdivision
// (:131-250).
void subdivide(const Mesh& in, double len_thres, Subdivision* s);

// Subdivision::ComputeGeometryNeighbors (subdivision.cc:252-341).
void geometry_neighbors(Subdivision* s, double step, int threads);

// Subdivision::ComputeRepresentativeGraph (subdivision.cc:343-399).
void representative_graph(Subdivision* s, double step);

// The whole chain as cad_deform.cc:43-51 and LoadCadMesh (mesh_tensor.cc:
// 102-115) run it: clean, merge, Subdivide(2e-2),
// ComputeGeometryNeighbors(1.5e-2), ComputeRepresentativeGraph(1e-2).
void prepare(Mesh cad, Subdivision* s, int threads);

}  // namespace cad
}  // namespace mo
