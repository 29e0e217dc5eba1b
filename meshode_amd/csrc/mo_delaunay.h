// Delaunay triangulations for the CAD subdivision chain (host code).
//
// The reference calls CGAL's Delaunay_triangulation_2 (projection traits,
// exact predicates) on a face's boundary + lattice points
// (delaunay.cc:50-74) and Delaunay_triangulation_3 on the vertices of one
// neighbourhood cell, keeping its finite edges (delaunay.cc:76-101).
//
// Here: incremental Bowyer-Watson with an infinite vertex, exact
// predicates (floating-point filter, then exact expansion arithmetic after
// Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast Robust
// Geometric Predicates", 1997) and the symbolic perturbation of Devillers and
// Teillaud ("Perturbations for Delaunay and weighted Delaunay 3D
// triangulations", CGTA 2011): a point exactly on a circum-circle/sphere is
// decided by the lexicographic order of the points, so the triangulation is
// unique and does not depend on the insertion order.  Point sets of lower
// dimension are handled as CGAL does: a coplanar 3D set is triangulated in
// its plane, a collinear set yields the chain of consecutive points.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace mo {
namespace geom {

// Exact signs (+1, 0, -1).
// orient2(a,b,c) > 0  <=>  a,b,c counter-clockwise.
int orient2(const double* a, const double* b, const double* c);
// incircle(a,b,c,d) > 0  <=>  d strictly inside the circle through a,b,c
// (a,b,c counter-clockwise).
int incircle(const double* a, const double* b, const double* c, const double* d);
// orient3(a,b,c,d) > 0  <=>  d on the side of plane abc that (b-a)x(c-a)
// points to.
int orient3(const double* a, const double* b, const double* c, const double* d);
// insphere(a,b,c,d,e) > 0  <=>  e strictly inside the sphere through a,b,c,d
// (orient3(a,b,c,d) > 0).
int insphere(const double* a, const double* b, const double* c, const double* d, const double* e);

// Triangles (CCW) of the Delaunay triangulation of n 2D points xy[2n].
// Exact duplicates are one vertex (the lowest index).  Fewer than three
// non-collinear points give no triangle.
void delaunay2(const double* xy, int n, std::vector<int>* tris);

// Simplices of the Delaunay triangulation of n 3D points xyz[3n]: *dim = 3
// (tetrahedra, positively oriented), 2 (coplanar set: triangles), 1
// (collinear set: segments between consecutive points), 0 or -1 (one or no
// distinct point: nothing).  Exact duplicates are one vertex (the lowest
// index).
void delaunay3(const double* xyz, int n, int* dim, std::vector<int>* simplices);

// Edges (i<j) of the Delaunay triangulation of n 3D points xyz[3n]
// (tetrahedra, or triangles/segments for coplanar/collinear sets).  Exact
// duplicates are one vertex (the lowest index).
void delaunay3_edges(const double* xyz, int n, std::vector<std::pair<int, int>>* edges);

}  // namespace geom
}  // namespace mo
