// CAD preprocessing chain (host code); see mo_cad.h.
//
// Arithmetic follows the reference expressions operation by operation
// (Eigen's fixed-size 3-vector expressions evaluated per component); dot
// products and norms sum their three terms left to right.  Built with
// -ffp-contract=off.
//
// Where the reference's result depends on an unspecified order (CGAL's face
// iteration, unordered_set iteration), this code fixes one: the faces of a
// subdivided triangle are listed with their smallest vertex first, sorted;
// cell members are visited in increasing vertex index (so a 1e-7 key shared
// by two points resolves to the larger index, and a cluster mean sums in
// index order).
#include "mo_cad.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <map>
#include <numeric>
#include <set>
#include <thread>
#include <unordered_map>

#include "mo_delaunay.h"

namespace mo {
namespace cad {
namespace {

struct V3 {
  double x, y, z;
  double operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
};
inline V3 at(const std::vector<double>& V, int i) { return V3{V[3 * i], V[3 * i + 1], V[3 * i + 2]}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 div(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
inline void push(std::vector<double>& V, V3 p) {
  V.push_back(p.x);
  V.push_back(p.y);
  V.push_back(p.z);
}

// double -> int conversion as the x86 truncating instruction does it
// (out-of-range and NaN give INT_MIN).
inline int trunc_i32(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
  return (int)v;
}

inline uint64_t ekey(int a, int b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

struct Key3 {
  int a, b, c;
  bool operator==(const Key3& o) const { return a == o.a && b == o.b && c == o.c; }
  bool operator<(const Key3& o) const {
    return a < o.a || (a == o.a && (b < o.b || (b == o.b && c < o.c)));
  }
};
struct Key3Hash {
  size_t operator()(const Key3& k) const {
    uint64_t h = (uint64_t)(uint32_t)k.a * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)k.b * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)k.c * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// Subdivision::calculateBarycentricCoordinate (subdivision.cc:448-460).
inline double area2(double ax, double ay, double bx, double by, double cx, double cy) {
  return (cx - ax) * (by - ay) - (bx - ax) * (cy - ay);
}

// Subdivision::DelaunaySubdivision (subdivision.cc:131-250): boundary points
// bent onto arcs of radius ~1e3 bulging out of the face, lattice points of
// spacing len_thres strictly inside the face in its (tx, ty) frame, the 2D
// Delaunay triangulation of both, minus triangles with a 3D edge longer than
// 3 len_thres.
void delaunay_subdivision(const std::vector<int>* boundary, std::vector<double>& V, std::vector<int>& F,
                          const int* face, double len) {
  const V3 v0 = at(V, face[0]), v1 = at(V, face[1]), v2 = at(V, face[2]);
  V3 n = cross(v1 - v0, v2 - v0);
  n = div(n, norm(n));
  std::map<int, V3> curved;
  for (int i = 0; i < 3; ++i) {
    const V3 x = at(V, face[i]), y = at(V, face[(i + 1) % 3]);
    V3 dir = cross(n, y - x);
    dir = div(dir, norm(dir));
    const V3 c{(y.x + x.x) * 0.5 + 1e3 * dir.x, (y.y + x.y) * 0.5 + 1e3 * dir.y, (y.z + x.z) * 0.5 + 1e3 * dir.z};
    const double l = norm(c - x);
    for (int p : boundary[i]) {
      const V3 d = at(V, p) - c;
      const double dn = norm(d);
      curved[p] = V3{d.x / dn * l + c.x, d.y / dn * l + c.y, d.z / dn * l + c.z};
    }
  }
  // the reference compares each axis with n[min_axis] but stores 1, not the
  // axis index (subdivision.cc:166-171)
  int min_axis = 0;
  for (int i = 1; i < 3; ++i)
    if (std::fabs(n[i]) < std::fabs(n[min_axis])) min_axis = 1;
  V3 e{0, 0, 0};
  if (min_axis == 0) e.x = 1; else e.y = 1;
  V3 tx = cross(e, n);
  tx = div(tx, norm(tx));
  const V3 ty = cross(n, tx);

  std::vector<int> vindices;
  std::vector<V3> points;
  for (const auto& kv : curved) {
    vindices.push_back(kv.first);
    points.push_back(kv.second);
  }
  const double v0x = 0, v0y = 0;
  const double v1x = dot(v1 - v0, tx) / len, v1y = dot(v1 - v0, ty) / len;
  const double v2x = dot(v2 - v0, tx) / len, v2y = dot(v2 - v0, ty) / len;
  const int minX = trunc_i32(std::min(v0x, std::min(v1x, v2x)));
  const int minY = trunc_i32(std::min(v0y, std::min(v1y, v2y)));
  const int maxX = trunc_i32(std::max(v0x, std::max(v1x, v2x)) + (double)0.999999f);
  const int maxY = trunc_i32(std::max(v0y, std::max(v1y, v2y)) + (double)0.999999f);
  const double tri_inv = 1.0 / area2(v0x, v0y, v1x, v1y, v2x, v2y);
  for (int py = minY; py <= maxY; ++py) {
    for (int px = minX; px <= maxX; ++px) {
      const double beta = area2(v0x, v0y, px, py, v2x, v2y) * tri_inv;
      const double gamma = area2(v0x, v0y, v1x, v1y, px, py) * tri_inv;
      const double alpha = 1.0 - beta - gamma;
      if (alpha > 0.0 && alpha < 1.0 && beta > 0.0 && beta < 1.0 && gamma > 0.0 && gamma < 1.0) {
        const double sx = px * len, sy = py * len;
        const V3 r{v0.x + tx.x * sx + ty.x * sy, v0.y + tx.y * sx + ty.y * sy, v0.z + tx.z * sx + ty.z * sy};
        vindices.push_back((int)(V.size() / 3));
        push(V, r);
        points.push_back(r);
      }
    }
  }

  const int np = (int)points.size();
  std::vector<double> xy(2 * (size_t)np);
  std::unordered_map<uint64_t, int> p2k;  // 1e7-truncated key -> last point with it
  auto key2 = [](double x, double y) {
    return ((uint64_t)(uint32_t)trunc_i32(x * 1e7) << 32) | (uint32_t)trunc_i32(y * 1e7);
  };
  for (int i = 0; i < np; ++i) {
    xy[2 * i] = dot(points[i] - v0, tx);
    xy[2 * i + 1] = dot(points[i] - v0, ty);
    p2k[key2(xy[2 * i], xy[2 * i + 1])] = i;
  }
  std::vector<int> tris;
  geom::delaunay2(xy.data(), np, &tris);
  std::vector<std::array<int, 3>> out;
  for (size_t t = 0; t < tris.size(); t += 3) {
    int v[3];
    for (int j = 0; j < 3; ++j) {
      const int k = tris[t + j];
      v[j] = vindices[p2k[key2(xy[2 * k], xy[2 * k + 1])]];
    }
    bool drop = false;
    for (int j = 0; j < 3; ++j)
      for (int k = j + 1; k < 3; ++k)
        if (norm(at(V, v[j]) - at(V, v[k])) > 3 * len) drop = true;
    if (drop) continue;
    int r = 0;
    if (v[1] < v[r]) r = 1;
    if (v[2] < v[r]) r = 2;
    out.push_back({v[r], v[(r + 1) % 3], v[(r + 2) % 3]});
  }
  std::sort(out.begin(), out.end());
  for (const auto& f : out) F.insert(F.end(), f.begin(), f.end());
}

}  // namespace

void remove_degenerated(Mesh* m) {
  std::vector<int> F;
  for (size_t i = 0; i < m->nf(); ++i) {
    const int* f = &m->F[3 * i];
    const V3 a = at(m->V, f[0]), b = at(m->V, f[1]), c = at(m->V, f[2]);
    if (norm(cross(b - a, c - a)) > 0) F.insert(F.end(), f, f + 3);
  }
  m->F.swap(F);
}

void merge_duplex(Mesh* m) {
  std::map<Key3, int> grid;
  std::vector<int> shrink(m->nv());
  std::vector<double> V;
  int top = 0;
  for (size_t i = 0; i < m->nv(); ++i) {
    const Key3 k{trunc_i32(m->V[3 * i] * 1e6), trunc_i32(m->V[3 * i + 1] * 1e6), trunc_i32(m->V[3 * i + 2] * 1e6)};
    auto it = grid.find(k);
    if (it == grid.end()) {
      shrink[i] = top;
      grid[k] = top;
      V.insert(V.end(), &m->V[3 * i], &m->V[3 * i] + 3);
      ++top;
    } else {
      shrink[i] = it->second;
    }
  }
  m->V.swap(V);
  std::set<std::array<int, 3>> seen;
  std::vector<int> F;
  for (size_t i = 0; i < m->F.size() / 3; ++i) {
    int f[3] = {shrink[m->F[3 * i]], shrink[m->F[3 * i + 1]], shrink[m->F[3 * i + 2]]};
    if (f[0] == f[1] || f[1] == f[2] || f[2] == f[0]) continue;
    std::array<int, 3> s{f[0], f[1], f[2]};
    std::sort(s.begin(), s.end());
    if (!seen.insert(s).second) continue;
    F.insert(F.end(), f, f + 3);
  }
  m->F.swap(F);
}

void subdivide(const Mesh& in, double len, Subdivision* s) {
  s->mesh.V = in.V;
  s->mesh.F.clear();
  std::vector<double>& V = s->mesh.V;
  const int nf = (int)in.nf();
  const int v_num = (int)in.nv();

  // faces connected through shared (undirected) edges, numbered in order of
  // their first face; faces are then visited component by component
  std::unordered_map<uint64_t, std::vector<int>> e2f;
  for (int i = 0; i < nf; ++i)
    for (int j = 0; j < 3; ++j) e2f[ekey(in.F[3 * i + j], in.F[3 * i + (j + 1) % 3])].push_back(i);
  std::vector<int> group(nf, -1);
  int ng = 0;
  std::vector<int> queue;
  for (int i = 0; i < nf; ++i) {
    if (group[i] >= 0) continue;
    queue.assign(1, i);
    group[i] = ng;
    for (size_t h = 0; h < queue.size(); ++h) {
      const int f = queue[h];
      for (int j = 0; j < 3; ++j)
        for (int g : e2f[ekey(in.F[3 * f + j], in.F[3 * f + (j + 1) % 3])])
          if (group[g] < 0) {
            group[g] = ng;
            queue.push_back(g);
          }
    }
    ++ng;
  }
  std::vector<int> order(nf);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return group[a] < group[b]; });

  std::unordered_map<uint64_t, std::vector<int>> split;
  std::vector<char> boundary(V.size() / 3, 0);
  for (int f : order) {
    const int face[3] = {in.F[3 * f], in.F[3 * f + 1], in.F[3 * f + 2]};
    std::vector<int> lists[3];
    for (int j = 0; j < 3; ++j) {
      const int a = face[j], b = face[(j + 1) % 3];
      const uint64_t h = ekey(a, b);
      auto it = split.find(h);
      if (it == split.end()) {
        const V3 pa = at(V, a);
        V3 diff = at(V, b) - pa;
        const int ns = trunc_i32(norm(diff) / len + 1);
        diff = div(diff, (double)ns);
        std::vector<int> idx{a};
        for (int t = 1; t < ns; ++t) {
          idx.push_back((int)(V.size() / 3));
          boundary.push_back(1);
          push(V, V3{pa.x + diff.x * t, pa.y + diff.y * t, pa.z + diff.z * t});
        }
        idx.push_back(b);
        it = split.emplace(h, std::move(idx)).first;
      }
      lists[j] = it->second;
    }
    delaunay_subdivision(lists, V, s->mesh.F, face, len);
    boundary.resize(V.size() / 3, 0);
  }
  s->internal.assign(V.size() / 3, 0);
  for (size_t i = v_num; i < V.size() / 3; ++i) s->internal[i] = !boundary[i];
}

void geometry_neighbors(Subdivision* s, double step, int threads) {
  const std::vector<double>& V = s->mesh.V;
  const int nv = (int)s->mesh.nv();
  // every vertex joins the cells of its position shifted by 0 or step per
  // axis (subdivision.cc:264-288)
  std::unordered_map<Key3, std::vector<int>, Key3Hash> cells;
  for (int i = 0; i < nv; ++i) {
    for (int j = 0; j < 8; ++j) {
      const double d0 = (j & 4) ? step : 0, d1 = (j & 2) ? step : 0, d2 = (j & 1) ? step : 0;
      const Key3 k{trunc_i32((V[3 * i] + d0) / step), trunc_i32((V[3 * i + 1] + d1) / step),
                   trunc_i32((V[3 * i + 2] + d2) / step)};
      auto& l = cells[k];
      if (l.empty() || l.back() != i) l.push_back(i);
    }
  }
  std::vector<const std::vector<int>*> work;
  std::vector<std::pair<int, int>> pairs;
  for (const auto& kv : cells) {
    if (kv.second.size() == 2) pairs.emplace_back(kv.second[0], kv.second[1]);
    else if (kv.second.size() >= 3) work.push_back(&kv.second);
  }
  // larger cells first so the threads finish together
  std::sort(work.begin(), work.end(), [](const std::vector<int>* a, const std::vector<int>* b) {
    return a->size() > b->size() || (a->size() == b->size() && (*a) < (*b));
  });
  threads = std::max(1, std::min(threads, 64));
  std::vector<std::vector<std::pair<int, int>>> part(threads);
  auto body = [&](int tid) {
    std::vector<double> xyz;
    std::vector<std::pair<int, int>> e;
    std::unordered_map<Key3, int, Key3Hash> p2k;
    for (size_t w = tid; w < work.size(); w += threads) {
      const std::vector<int>& l = *work[w];
      xyz.resize(3 * l.size());
      p2k.clear();
      std::vector<Key3> keys(l.size());
      for (size_t i = 0; i < l.size(); ++i) {
        for (int k = 0; k < 3; ++k) xyz[3 * i + k] = V[3 * l[i] + k];
        keys[i] = Key3{trunc_i32(xyz[3 * i] * 1e7), trunc_i32(xyz[3 * i + 1] * 1e7), trunc_i32(xyz[3 * i + 2] * 1e7)};
        p2k[keys[i]] = l[i];
      }
      geom::delaunay3_edges(xyz.data(), (int)l.size(), &e);
      for (const auto& ab : e) {
        int a = p2k[keys[ab.first]], b = p2k[keys[ab.second]];
        if (a == b) continue;
        if (a > b) std::swap(a, b);
        part[tid].emplace_back(a, b);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(body, t);
  body(0);
  for (auto& th : pool) th.join();
  for (auto& p : part) pairs.insert(pairs.end(), p.begin(), p.end());
  for (auto& p : pairs)
    if (p.first > p.second) std::swap(p.first, p.second);
  for (size_t i = 0; i < s->mesh.nf(); ++i)
    for (int j = 0; j < 3; ++j) {
      int a = s->mesh.F[3 * i + j], b = s->mesh.F[3 * i + (j + 1) % 3];
      if (a == b) continue;
      if (a > b) std::swap(a, b);
      pairs.emplace_back(a, b);
    }
  pairs.erase(std::remove_if(pairs.begin(), pairs.end(), [](const std::pair<int, int>& p) { return p.first == p.second; }),
              pairs.end());
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  s->neighbors.swap(pairs);
}

void representative_graph(Subdivision* s, double step) {
  const std::vector<double>& V = s->mesh.V;
  const int nv = (int)s->mesh.nv();
  std::map<Key3, std::vector<int>> cells;
  for (int i = 0; i < nv; ++i)
    cells[Key3{trunc_i32(V[3 * i] / step), trunc_i32(V[3 * i + 1] / step), trunc_i32(V[3 * i + 2] / step)}].push_back(i);
  s->ref.assign(nv, 0);
  s->graph_v.clear();
  int g = 0;
  for (const auto& kv : cells) {
    double p[3] = {0, 0, 0};
    for (int id : kv.second) {
      for (int k = 0; k < 3; ++k) p[k] += V[3 * id + k];
      s->ref[id] = g;
    }
    for (int k = 0; k < 3; ++k) s->graph_v.push_back(p[k] / (double)kv.second.size());
    ++g;
  }
  std::vector<std::pair<int, int>> e;
  auto add = [&](int a, int b) {
    a = s->ref[a];
    b = s->ref[b];
    if (a == b) return;
    if (a > b) std::swap(a, b);
    e.emplace_back(a, b);
  };
  for (const auto& p : s->neighbors) add(p.first, p.second);
  for (size_t i = 0; i < s->mesh.nf(); ++i)
    for (int j = 0; j < 3; ++j) add(s->mesh.F[3 * i + j], s->mesh.F[3 * i + (j + 1) % 3]);
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  s->graph_e.swap(e);
}

void prepare(Mesh cad, Subdivision* s, int threads) {
  remove_degenerated(&cad);
  merge_duplex(&cad);
  subdivide(cad, 2e-2, s);
  geometry_neighbors(s, 1.5e-2, threads);
  representative_graph(s, 1e-2);
}

}  // namespace cad
}  // namespace mo
