// Exact predicates and Delaunay triangulations (host code); see mo_delaunay.h.
// Built with -ffp-contract=off: the error-free transformations below rely on
// every floating-point operation being rounded on its own.
#include "mo_delaunay.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace mo {
namespace geom {
namespace {

// ---------------------------------------------------------------------------
// Expansion arithmetic (nonoverlapping expansions, increasing magnitude,
// zero components eliminated; the empty expansion is zero).
using Ex = std::vector<double>;

inline void two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  const double bv = x - a;
  const double av = x - bv;
  y = (a - av) + (b - bv);
}
inline void two_diff(double a, double b, double& x, double& y) {
  x = a - b;
  const double bv = a - x;
  const double av = x + bv;
  y = (a - av) + (bv - b);
}
inline void two_prod(double a, double b, double& x, double& y) {
  x = a * b;
  y = std::fma(a, b, -x);
}

Ex ex_diff(double a, double b) {
  double x, y;
  two_diff(a, b, x, y);
  Ex e;
  if (y != 0) e.push_back(y);
  if (x != 0) e.push_back(x);
  return e;
}

// e + b
void grow(Ex& e, double b) {
  Ex h;
  h.reserve(e.size() + 1);
  double q = b;
  for (double ei : e) {
    double s, hh;
    two_sum(q, ei, s, hh);
    if (hh != 0) h.push_back(hh);
    q = s;
  }
  if (q != 0) h.push_back(q);
  e.swap(h);
}

Ex ex_add(const Ex& e, const Ex& f) {
  Ex r = e;
  for (double fi : f) grow(r, fi);
  return r;
}

Ex ex_neg(Ex e) {
  for (double& x : e) x = -x;
  return e;
}

Ex ex_sub(const Ex& e, const Ex& f) { return ex_add(e, ex_neg(f)); }

// e * b
Ex ex_scale(const Ex& e, double b) {
  Ex h;
  if (e.empty() || b == 0) return h;
  h.reserve(2 * e.size());
  double q, hh;
  two_prod(e[0], b, q, hh);
  if (hh != 0) h.push_back(hh);
  for (size_t i = 1; i < e.size(); ++i) {
    double p1, p0, s;
    two_prod(e[i], b, p1, p0);
    two_sum(q, p0, s, hh);
    if (hh != 0) h.push_back(hh);
    two_sum(p1, s, q, hh);
    if (hh != 0) h.push_back(hh);
  }
  if (q != 0) h.push_back(q);
  return h;
}

Ex ex_mul(const Ex& e, const Ex& f) {
  Ex r;
  for (double fi : f) r = ex_add(r, ex_scale(e, fi));
  return r;
}

inline int ex_sign(const Ex& e) { return e.empty() ? 0 : (e.back() > 0 ? 1 : -1); }

constexpr double kEps = 1.1102230246251565e-16;  // 2^-53

inline int sgn(double x) { return (x > 0) - (x < 0); }

// ---------------------------------------------------------------------------
// Exact determinants.

int orient2_exact(const double* a, const double* b, const double* c) {
  const Ex acx = ex_diff(a[0], c[0]), bcy = ex_diff(b[1], c[1]);
  const Ex acy = ex_diff(a[1], c[1]), bcx = ex_diff(b[0], c[0]);
  return ex_sign(ex_sub(ex_mul(acx, bcy), ex_mul(acy, bcx)));
}

int incircle_exact(const double* a, const double* b, const double* c, const double* d) {
  const Ex adx = ex_diff(a[0], d[0]), ady = ex_diff(a[1], d[1]);
  const Ex bdx = ex_diff(b[0], d[0]), bdy = ex_diff(b[1], d[1]);
  const Ex cdx = ex_diff(c[0], d[0]), cdy = ex_diff(c[1], d[1]);
  const Ex alift = ex_add(ex_mul(adx, adx), ex_mul(ady, ady));
  const Ex blift = ex_add(ex_mul(bdx, bdx), ex_mul(bdy, bdy));
  const Ex clift = ex_add(ex_mul(cdx, cdx), ex_mul(cdy, cdy));
  const Ex bc = ex_sub(ex_mul(bdx, cdy), ex_mul(cdx, bdy));
  const Ex ca = ex_sub(ex_mul(cdx, ady), ex_mul(adx, cdy));
  const Ex ab = ex_sub(ex_mul(adx, bdy), ex_mul(bdx, ady));
  return ex_sign(ex_add(ex_add(ex_mul(alift, bc), ex_mul(blift, ca)), ex_mul(clift, ab)));
}

// sign of det[a-d; b-d; c-d]
int det3_exact(const double* a, const double* b, const double* c, const double* d) {
  Ex ad[3], bd[3], cd[3];
  for (int k = 0; k < 3; ++k) {
    ad[k] = ex_diff(a[k], d[k]);
    bd[k] = ex_diff(b[k], d[k]);
    cd[k] = ex_diff(c[k], d[k]);
  }
  const Ex m0 = ex_sub(ex_mul(bd[1], cd[2]), ex_mul(bd[2], cd[1]));
  const Ex m1 = ex_sub(ex_mul(cd[1], ad[2]), ex_mul(cd[2], ad[1]));
  const Ex m2 = ex_sub(ex_mul(ad[1], bd[2]), ex_mul(ad[2], bd[1]));
  return ex_sign(ex_add(ex_add(ex_mul(ad[0], m0), ex_mul(bd[0], m1)), ex_mul(cd[0], m2)));
}

// sign of the lifted 4x4 determinant with rows (p - e, |p - e|^2), p = a,b,c,d
int lift4_exact(const double* a, const double* b, const double* c, const double* d, const double* e) {
  Ex A[3], B[3], C[3], D[3];
  for (int k = 0; k < 3; ++k) {
    A[k] = ex_diff(a[k], e[k]);
    B[k] = ex_diff(b[k], e[k]);
    C[k] = ex_diff(c[k], e[k]);
    D[k] = ex_diff(d[k], e[k]);
  }
  auto m2 = [](const Ex* p, const Ex* q) { return ex_sub(ex_mul(p[0], q[1]), ex_mul(q[0], p[1])); };
  const Ex ab = m2(A, B), bc = m2(B, C), cd = m2(C, D), da = m2(D, A), ac = m2(A, C), bd = m2(B, D);
  const Ex abc = ex_add(ex_sub(ex_mul(A[2], bc), ex_mul(B[2], ac)), ex_mul(C[2], ab));
  const Ex bcd = ex_add(ex_sub(ex_mul(B[2], cd), ex_mul(C[2], bd)), ex_mul(D[2], bc));
  const Ex cda = ex_add(ex_add(ex_mul(C[2], da), ex_mul(D[2], ac)), ex_mul(A[2], cd));
  const Ex dab = ex_add(ex_add(ex_mul(D[2], ab), ex_mul(A[2], bd)), ex_mul(B[2], da));
  auto lift = [](const Ex* p) { return ex_add(ex_add(ex_mul(p[0], p[0]), ex_mul(p[1], p[1])), ex_mul(p[2], p[2])); };
  const Ex la = lift(A), lb = lift(B), lc = lift(C), ld = lift(D);
  const Ex det = ex_add(ex_sub(ex_mul(ld, abc), ex_mul(lc, dab)), ex_sub(ex_mul(lb, cda), ex_mul(la, bcd)));
  return ex_sign(det);
}

}  // namespace

// ---------------------------------------------------------------------------
// Filtered predicates: the floating-point value decides when its magnitude
// exceeds a bound on its rounding error (a constant times eps times the
// permanent, at least twice the constants of Shewchuk's first-stage bounds);
// otherwise the exact expansion decides.

int orient2(const double* a, const double* b, const double* c) {
  const double l = (a[0] - c[0]) * (b[1] - c[1]);
  const double r = (a[1] - c[1]) * (b[0] - c[0]);
  const double det = l - r;
  const double bound = 8 * kEps * (std::fabs(l) + std::fabs(r));
  if (det > bound || -det > bound) return sgn(det);
  return orient2_exact(a, b, c);
}

int incircle(const double* a, const double* b, const double* c, const double* d) {
  const double adx = a[0] - d[0], ady = a[1] - d[1];
  const double bdx = b[0] - d[0], bdy = b[1] - d[1];
  const double cdx = c[0] - d[0], cdy = c[1] - d[1];
  const double bc1 = bdx * cdy, bc2 = cdx * bdy;
  const double ca1 = cdx * ady, ca2 = adx * cdy;
  const double ab1 = adx * bdy, ab2 = bdx * ady;
  const double al = adx * adx + ady * ady, bl = bdx * bdx + bdy * bdy, cl = cdx * cdx + cdy * cdy;
  const double det = al * (bc1 - bc2) + bl * (ca1 - ca2) + cl * (ab1 - ab2);
  const double perm = (std::fabs(bc1) + std::fabs(bc2)) * al + (std::fabs(ca1) + std::fabs(ca2)) * bl +
                      (std::fabs(ab1) + std::fabs(ab2)) * cl;
  const double bound = 24 * kEps * perm;
  if (det > bound || -det > bound) return sgn(det);
  return incircle_exact(a, b, c, d);
}

int orient3(const double* a, const double* b, const double* c, const double* d) {
  const double adx = a[0] - d[0], ady = a[1] - d[1], adz = a[2] - d[2];
  const double bdx = b[0] - d[0], bdy = b[1] - d[1], bdz = b[2] - d[2];
  const double cdx = c[0] - d[0], cdy = c[1] - d[1], cdz = c[2] - d[2];
  const double m0a = bdy * cdz, m0b = bdz * cdy;
  const double m1a = cdy * adz, m1b = cdz * ady;
  const double m2a = ady * bdz, m2b = adz * bdy;
  const double det = adx * (m0a - m0b) + bdx * (m1a - m1b) + cdx * (m2a - m2b);
  const double perm = std::fabs(adx) * (std::fabs(m0a) + std::fabs(m0b)) +
                      std::fabs(bdx) * (std::fabs(m1a) + std::fabs(m1b)) +
                      std::fabs(cdx) * (std::fabs(m2a) + std::fabs(m2b));
  const double bound = 16 * kEps * perm;
  // det = det[a-d; b-d; c-d] = -det[b-a, c-a, d-a]
  if (det > bound || -det > bound) return -sgn(det);
  return -det3_exact(a, b, c, d);
}

int insphere(const double* a, const double* b, const double* c, const double* d, const double* e) {
  double A[3], B[3], C[3], D[3];
  for (int k = 0; k < 3; ++k) {
    A[k] = a[k] - e[k];
    B[k] = b[k] - e[k];
    C[k] = c[k] - e[k];
    D[k] = d[k] - e[k];
  }
  struct M { double v, p; };
  auto m2 = [](const double* p, const double* q) {
    const double x = p[0] * q[1], y = q[0] * p[1];
    return M{x - y, std::fabs(x) + std::fabs(y)};
  };
  const M ab = m2(A, B), bc = m2(B, C), cd = m2(C, D), da = m2(D, A), ac = m2(A, C), bd = m2(B, D);
  const double abc = A[2] * bc.v - B[2] * ac.v + C[2] * ab.v;
  const double bcd = B[2] * cd.v - C[2] * bd.v + D[2] * bc.v;
  const double cda = C[2] * da.v + D[2] * ac.v + A[2] * cd.v;
  const double dab = D[2] * ab.v + A[2] * bd.v + B[2] * da.v;
  const double abcp = std::fabs(A[2]) * bc.p + std::fabs(B[2]) * ac.p + std::fabs(C[2]) * ab.p;
  const double bcdp = std::fabs(B[2]) * cd.p + std::fabs(C[2]) * bd.p + std::fabs(D[2]) * bc.p;
  const double cdap = std::fabs(C[2]) * da.p + std::fabs(D[2]) * ac.p + std::fabs(A[2]) * cd.p;
  const double dabp = std::fabs(D[2]) * ab.p + std::fabs(A[2]) * bd.p + std::fabs(B[2]) * da.p;
  auto lift = [](const double* p) { return p[0] * p[0] + p[1] * p[1] + p[2] * p[2]; };
  const double la = lift(A), lb = lift(B), lc = lift(C), ld = lift(D);
  const double det = (ld * abc - lc * dab) + (lb * cda - la * bcd);
  const double perm = ld * abcp + lc * dabp + lb * cdap + la * bcdp;
  const double bound = 48 * kEps * perm;
  // the lifted determinant is negative for e inside when orient3(a,b,c,d) > 0
  if (det > bound || -det > bound) return -sgn(det);
  return -lift4_exact(a, b, c, d, e);
}

namespace {

// ---------------------------------------------------------------------------
// 2D Bowyer-Watson over a predicate set P:
//   int  orient(a, b, c)        exact orientation (>0: counter-clockwise)
//   int  incircle(a, b, c, d)   exact, >0: d inside the circle of the CCW abc
//   bool less(a, b)             lexicographic order of the points
//   bool between(a, b, p)       p strictly inside segment ab (p collinear)
template <class P>
struct Tri2 {
  struct T {
    int v[3];
    int nb[3];
  };
  const P& P_;
  int inf;
  std::vector<T> t;
  std::vector<char> dead;
  std::vector<int> mark;
  int stamp = 0;

  Tri2(const P& p, int n) : P_(p), inf(n) {}

  // Devillers-Teillaud symbolic perturbation for a point exactly on the
  // circle: the points are lifted by decreasing infinitesimals in
  // lexicographic order, and the sign is that of the first non-vanishing
  // orientation term, taken from the largest point down.
  int incircle_pert(int p0, int p1, int p2, int p) const {
    const int s = P_.incircle(p0, p1, p2, p);
    if (s) return s;
    int pts[4] = {p0, p1, p2, p};
    std::sort(pts, pts + 4, [&](int x, int y) { return P_.less(x, y); });
    for (int i = 3; i >= 1; --i) {
      const int q = pts[i];
      int o = 0;
      if (q == p) return -1;
      if (q == p2) o = P_.orient(p0, p1, p);
      else if (q == p1) o = P_.orient(p0, p, p2);
      else if (q == p0) o = P_.orient(p, p1, p2);
      if (o) return o;
    }
    return -1;
  }

  bool conflict(const T& x, int p) const {
    int k = -1;
    for (int i = 0; i < 3; ++i)
      if (x.v[i] == inf) k = i;
    if (k < 0) return incircle_pert(x.v[0], x.v[1], x.v[2], p) > 0;
    int w[3] = {x.v[0], x.v[1], x.v[2]};
    w[k] = p;
    const int o = P_.orient(w[0], w[1], w[2]);
    if (o) return o > 0;
    return P_.between(x.v[(k + 1) % 3], x.v[(k + 2) % 3], p);
  }

  int add(int a, int b, int c) {
    t.push_back(T{{a, b, c}, {-1, -1, -1}});
    dead.push_back(0);
    mark.push_back(0);
    return (int)t.size() - 1;
  }

  void link_all() {
    // adjacency of a freshly built set of triangles
    std::unordered_map<long long, std::pair<int, int>> open;
    for (int i = 0; i < (int)t.size(); ++i) {
      for (int j = 0; j < 3; ++j) {
        const int a = t[i].v[(j + 1) % 3], b = t[i].v[(j + 2) % 3];
        const long long key = (long long)std::min(a, b) * (inf + 1) + std::max(a, b);
        auto it = open.find(key);
        if (it == open.end()) {
          open[key] = {i, j};
        } else {
          t[i].nb[j] = it->second.first;
          t[it->second.first].nb[it->second.second] = i;
          open.erase(it);
        }
      }
    }
  }

  bool init(const std::vector<int>& order, int* used) {
    const int i0 = order[0], i1 = order[1];
    for (size_t k = 2; k < order.size(); ++k) {
      const int o = P_.orient(i0, i1, order[k]);
      if (!o) continue;
      int a = i0, b = i1, c = order[k];
      if (o < 0) std::swap(a, b);
      add(a, b, c);
      add(b, a, inf);
      add(c, b, inf);
      add(a, c, inf);
      link_all();
      *used = (int)k;
      return true;
    }
    return false;
  }

  void insert(int p) {
    int seed = -1;
    for (int i = (int)t.size() - 1; i >= 0; --i)
      if (!dead[i] && conflict(t[i], p)) { seed = i; break; }
    if (seed < 0) return;  // cannot happen for a distinct point
    ++stamp;
    std::vector<int> cav{seed};
    mark[seed] = stamp;
    for (size_t h = 0; h < cav.size(); ++h) {
      const T& x = t[cav[h]];
      for (int j = 0; j < 3; ++j) {
        const int nb = x.nb[j];
        if (mark[nb] == stamp || mark[nb] == -stamp) continue;
        if (conflict(t[nb], p)) { mark[nb] = stamp; cav.push_back(nb); }
        else mark[nb] = -stamp;
      }
    }
    // one new triangle per boundary edge: the conflict triangle with the
    // vertex opposite that edge replaced by p
    std::unordered_map<int, std::pair<int, int>> open;
    for (int c : cav) {
      for (int j = 0; j < 3; ++j) {
        const int nb = t[c].nb[j];
        if (mark[nb] == stamp) continue;
        T nt = t[c];
        nt.v[j] = p;
        const int id = add(nt.v[0], nt.v[1], nt.v[2]);
        t[id].nb[j] = nb;
        for (int q = 0; q < 3; ++q)
          if (t[nb].nb[q] == c) t[nb].nb[q] = id;
        for (int s = 1; s <= 2; ++s) {
          const int slot = (j + s) % 3;             // edge opposite `slot` joins p and x
          const int x = t[id].v[(j + 3 - s) % 3];
          auto it = open.find(x);
          if (it == open.end()) {
            open[x] = {id, slot};
          } else {
            t[id].nb[slot] = it->second.first;
            t[it->second.first].nb[it->second.second] = id;
            open.erase(it);
          }
        }
      }
    }
    for (int c : cav) dead[c] = 1;
  }
};

struct Plane2 {
  const double* xy;
  int orient(int a, int b, int c) const { return orient2(xy + 2 * a, xy + 2 * b, xy + 2 * c); }
  int incircle(int a, int b, int c, int d) const {
    return geom::incircle(xy + 2 * a, xy + 2 * b, xy + 2 * c, xy + 2 * d);
  }
  bool less(int a, int b) const {
    const double* p = xy + 2 * a;
    const double* q = xy + 2 * b;
    return p[0] < q[0] || (p[0] == q[0] && p[1] < q[1]);
  }
  bool between(int a, int b, int p) const {
    const int k = xy[2 * a] != xy[2 * b] ? 0 : 1;
    const double lo = std::min(xy[2 * a + k], xy[2 * b + k]), hi = std::max(xy[2 * a + k], xy[2 * b + k]);
    return xy[2 * p + k] > lo && xy[2 * p + k] < hi;
  }
};

bool lex3(const double* p, const double* q) {
  return p[0] < q[0] || (p[0] == q[0] && (p[1] < q[1] || (p[1] == q[1] && p[2] < q[2])));
}

// Exactly coplanar 3D points: orientation as seen from an apex q off the
// plane, in-circle through the sphere on the triangle and q (the sphere
// meets the plane in the triangle's circum-circle).
struct Coplanar3 {
  const double* xyz;
  double q[3];
  int orient(int a, int b, int c) const { return orient3(xyz + 3 * a, xyz + 3 * b, xyz + 3 * c, q); }
  int incircle(int a, int b, int c, int d) const {
    const double *A = xyz + 3 * a, *B = xyz + 3 * b, *C = xyz + 3 * c;
    return insphere(A, B, C, q, xyz + 3 * d) * orient3(A, B, C, q);
  }
  bool less(int a, int b) const { return lex3(xyz + 3 * a, xyz + 3 * b); }
  bool between(int a, int b, int p) const {
    int k = 0;
    while (k < 2 && xyz[3 * a + k] == xyz[3 * b + k]) ++k;
    const double lo = std::min(xyz[3 * a + k], xyz[3 * b + k]), hi = std::max(xyz[3 * a + k], xyz[3 * b + k]);
    return xyz[3 * p + k] > lo && xyz[3 * p + k] < hi;
  }
};

// An apex off the plane through a, b, c (exactly non-collinear): a moved
// along a coordinate axis by more than the points' extent, trying the axes
// in decreasing order of the (rounded) normal's components until the exact
// orientation is non-zero (a plane cannot contain all three axes).
void apex_off_plane(const double* a, const double* b, const double* c, double extent, double* q) {
  const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double n[3] = {std::fabs(u[1] * v[2] - u[2] * v[1]), std::fabs(u[2] * v[0] - u[0] * v[2]),
                       std::fabs(u[0] * v[1] - u[1] * v[0])};
  int ax[3] = {0, 1, 2};
  std::sort(ax, ax + 3, [&](int i, int j) { return n[i] > n[j]; });
  const double off = std::ldexp(1.0, std::ilogb(extent + 1.0) + 2);
  for (int t = 0; t < 3; ++t) {
    q[0] = a[0];
    q[1] = a[1];
    q[2] = a[2];
    q[ax[t]] = a[ax[t]] + off;
    if (orient3(a, b, c, q) != 0) return;
  }
}

// Unique points (exact duplicates dropped, keeping the lowest index).
template <int D>
std::vector<int> unique_points(const double* p, int n) {
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  auto lt = [&](int a, int b) {
    for (int k = 0; k < D; ++k) {
      if (p[D * a + k] != p[D * b + k]) return p[D * a + k] < p[D * b + k];
    }
    return a < b;
  };
  std::sort(idx.begin(), idx.end(), lt);
  std::vector<int> out;
  for (int i = 0; i < n; ++i) {
    if (i > 0 && std::memcmp(p + D * idx[i], p + D * idx[i - 1], sizeof(double) * D) == 0) continue;
    out.push_back(idx[i]);
  }
  std::sort(out.begin(), out.end());
  return out;
}

template <class P>
void run_tri2(const P& pred, int n, const std::vector<int>& pts, std::vector<int>* tris) {
  tris->clear();
  if (pts.size() < 3) return;
  Tri2<P> tr(pred, n);
  int used = 0;
  if (!tr.init(pts, &used)) return;
  for (size_t k = 2; k < pts.size(); ++k)
    if ((int)k != used) tr.insert(pts[k]);
  for (size_t i = 0; i < tr.t.size(); ++i) {
    if (tr.dead[i]) continue;
    const auto& x = tr.t[i];
    if (x.v[0] == n || x.v[1] == n || x.v[2] == n) continue;
    tris->insert(tris->end(), {x.v[0], x.v[1], x.v[2]});
  }
}

// ---------------------------------------------------------------------------
// 3D Bowyer-Watson with an infinite vertex.
struct Tet3 {
  struct T {
    int v[4];
    int nb[4];
  };
  const double* X;
  int inf;
  double extent;
  std::vector<T> t;
  std::vector<char> dead;
  std::vector<int> mark;
  int stamp = 0;

  const double* P(int i) const { return X + 3 * i; }

  int insphere_pert(const int* v, int p) const {
    const int s = insphere(P(v[0]), P(v[1]), P(v[2]), P(v[3]), P(p));
    if (s) return s;
    int pts[5] = {v[0], v[1], v[2], v[3], p};
    std::sort(pts, pts + 5, [&](int a, int b) { return lex3(P(a), P(b)); });
    for (int i = 4; i >= 1; --i) {
      const int q = pts[i];
      if (q == p) return -1;
      int w[4] = {v[0], v[1], v[2], v[3]};
      for (int j = 0; j < 4; ++j)
        if (w[j] == q) w[j] = p;
      const int o = orient3(P(w[0]), P(w[1]), P(w[2]), P(w[3]));
      if (o) return o;
    }
    return -1;
  }

  // p exactly in the plane of the hull facet (a, b, c): inside its
  // circum-circle under the same perturbation (orientations relative to the
  // facet's own orientation seen from an apex off the plane).
  int coplanar_incircle_pert(int a, int b, int c, int p) const {
    double q[3];
    apex_off_plane(P(a), P(b), P(c), extent, q);
    const int of = orient3(P(a), P(b), P(c), q);
    const int s = insphere(P(a), P(b), P(c), q, P(p)) * of;
    if (s) return s;
    int pts[4] = {a, b, c, p};
    std::sort(pts, pts + 4, [&](int x, int y) { return lex3(P(x), P(y)); });
    for (int i = 3; i >= 1; --i) {
      const int r = pts[i];
      if (r == p) return -1;
      int w[3] = {a, b, c};
      for (int j = 0; j < 3; ++j)
        if (w[j] == r) w[j] = p;
      const int o = orient3(P(w[0]), P(w[1]), P(w[2]), q) * of;
      if (o) return o;
    }
    return -1;
  }

  bool conflict(const T& x, int p) const {
    int k = -1;
    for (int i = 0; i < 4; ++i)
      if (x.v[i] == inf) k = i;
    if (k < 0) return insphere_pert(x.v, p) > 0;
    int w[4] = {x.v[0], x.v[1], x.v[2], x.v[3]};
    w[k] = p;
    const int o = orient3(P(w[0]), P(w[1]), P(w[2]), P(w[3]));
    if (o) return o > 0;
    int f[3], m = 0;
    for (int i = 0; i < 4; ++i)
      if (i != k) f[m++] = x.v[i];
    return coplanar_incircle_pert(f[0], f[1], f[2], p) > 0;
  }

  int add(const int* v) {
    t.push_back(T{{v[0], v[1], v[2], v[3]}, {-1, -1, -1, -1}});
    dead.push_back(0);
    mark.push_back(0);
    return (int)t.size() - 1;
  }

  static long long fkey(int a, int b, int c, int base) {
    int s[3] = {a, b, c};
    std::sort(s, s + 3);
    return ((long long)s[0] * base + s[1]) * base + s[2];
  }

  void link_all() {
    std::unordered_map<long long, std::pair<int, int>> open;
    for (int i = 0; i < (int)t.size(); ++i) {
      for (int j = 0; j < 4; ++j) {
        int f[3], m = 0;
        for (int q = 0; q < 4; ++q)
          if (q != j) f[m++] = t[i].v[q];
        const long long key = fkey(f[0], f[1], f[2], inf + 1);
        auto it = open.find(key);
        if (it == open.end()) {
          open[key] = {i, j};
        } else {
          t[i].nb[j] = it->second.first;
          t[it->second.first].nb[it->second.second] = i;
          open.erase(it);
        }
      }
    }
  }

  void init(int a, int b, int c, int d) {
    if (orient3(P(a), P(b), P(c), P(d)) < 0) std::swap(a, b);
    const int v[4] = {a, b, c, d};
    add(v);
    for (int j = 0; j < 4; ++j) {
      // the finite tet with vertex j replaced by the infinite vertex and two
      // others swapped: outside points then orient positively
      int w[4] = {v[0], v[1], v[2], v[3]};
      w[j] = inf;
      std::swap(w[(j + 1) % 4], w[(j + 2) % 4]);
      add(w);
    }
    link_all();
  }

  void insert(int p) {
    int seed = -1;
    for (int i = (int)t.size() - 1; i >= 0; --i)
      if (!dead[i] && conflict(t[i], p)) { seed = i; break; }
    if (seed < 0) return;
    ++stamp;
    std::vector<int> cav{seed};
    mark[seed] = stamp;
    for (size_t h = 0; h < cav.size(); ++h) {
      for (int j = 0; j < 4; ++j) {
        const int nb = t[cav[h]].nb[j];
        if (mark[nb] == stamp || mark[nb] == -stamp) continue;
        if (conflict(t[nb], p)) { mark[nb] = stamp; cav.push_back(nb); }
        else mark[nb] = -stamp;
      }
    }
    std::unordered_map<long long, std::pair<int, int>> open;
    for (int c : cav) {
      for (int j = 0; j < 4; ++j) {
        const int nb = t[c].nb[j];
        if (mark[nb] == stamp) continue;
        int v[4] = {t[c].v[0], t[c].v[1], t[c].v[2], t[c].v[3]};
        v[j] = p;
        const int id = add(v);
        t[id].nb[j] = nb;
        for (int q = 0; q < 4; ++q)
          if (t[nb].nb[q] == c) t[nb].nb[q] = id;
        for (int s = 0; s < 4; ++s) {
          if (s == j) continue;
          // facet opposite s contains p and the two vertices other than s, j
          int o[2], m = 0;
          for (int q = 0; q < 4; ++q)
            if (q != s && q != j) o[m++] = v[q];
          const long long key = (long long)std::min(o[0], o[1]) * (inf + 1) + std::max(o[0], o[1]);
          auto it = open.find(key);
          if (it == open.end()) {
            open[key] = {id, s};
          } else {
            t[id].nb[s] = it->second.first;
            t[it->second.first].nb[it->second.second] = id;
            open.erase(it);
          }
        }
      }
    }
    for (int c : cav) dead[c] = 1;
  }
};

bool collinear3(const double* a, const double* b, const double* c) {
  // all three coordinate-plane projections have zero orientation
  const double a01[2] = {a[0], a[1]}, b01[2] = {b[0], b[1]}, c01[2] = {c[0], c[1]};
  const double a12[2] = {a[1], a[2]}, b12[2] = {b[1], b[2]}, c12[2] = {c[1], c[2]};
  const double a20[2] = {a[2], a[0]}, b20[2] = {b[2], b[0]}, c20[2] = {c[2], c[0]};
  return orient2(a01, b01, c01) == 0 && orient2(a12, b12, c12) == 0 && orient2(a20, b20, c20) == 0;
}

}  // namespace

void delaunay2(const double* xy, int n, std::vector<int>* tris) {
  const std::vector<int> pts = unique_points<2>(xy, n);
  run_tri2(Plane2{xy}, n, pts, tris);
}

void delaunay3(const double* xyz, int n, int* dim, std::vector<int>* simplices) {
  simplices->clear();
  const std::vector<int> pts = unique_points<3>(xyz, n);
  *dim = pts.empty() ? -1 : 0;
  if (pts.size() < 2) return;
  double extent = 0;
  for (int i : pts)
    for (int k = 0; k < 3; ++k) extent = std::max(extent, std::fabs(xyz[3 * i + k]));
  const int i0 = pts[0], i1 = pts[1];
  size_t k2 = 2;
  while (k2 < pts.size() && collinear3(xyz + 3 * i0, xyz + 3 * i1, xyz + 3 * pts[k2])) ++k2;
  if (k2 == pts.size()) {
    // collinear: consecutive points in lexicographic order
    std::vector<int> s = pts;
    std::sort(s.begin(), s.end(), [&](int a, int b) { return lex3(xyz + 3 * a, xyz + 3 * b); });
    *dim = 1;
    for (size_t i = 0; i + 1 < s.size(); ++i) simplices->insert(simplices->end(), {s[i], s[i + 1]});
    return;
  }
  const int i2 = pts[k2];
  size_t k3 = 2;
  while (k3 < pts.size() &&
         (k3 == k2 || orient3(xyz + 3 * i0, xyz + 3 * i1, xyz + 3 * i2, xyz + 3 * pts[k3]) == 0))
    ++k3;
  if (k3 == pts.size()) {
    Coplanar3 cp{xyz, {0, 0, 0}};
    apex_off_plane(xyz + 3 * i0, xyz + 3 * i1, xyz + 3 * i2, extent, cp.q);
    *dim = 2;
    run_tri2(cp, n, pts, simplices);
    return;
  }
  Tet3 tr{xyz, n, extent, {}, {}, {}, 0};
  tr.init(i0, i1, i2, pts[k3]);
  for (size_t k = 2; k < pts.size(); ++k)
    if (k != k2 && k != k3) tr.insert(pts[k]);
  *dim = 3;
  for (size_t i = 0; i < tr.t.size(); ++i) {
    if (tr.dead[i]) continue;
    const auto& x = tr.t[i];
    if (x.v[0] == n || x.v[1] == n || x.v[2] == n || x.v[3] == n) continue;
    simplices->insert(simplices->end(), {x.v[0], x.v[1], x.v[2], x.v[3]});
  }
}

void delaunay3_edges(const double* xyz, int n, std::vector<std::pair<int, int>>* edges) {
  int dim = 0;
  std::vector<int> s;
  delaunay3(xyz, n, &dim, &s);
  edges->clear();
  if (dim < 1) return;
  const int k = dim + 1;
  for (size_t i = 0; i + k <= s.size(); i += k)
    for (int a = 0; a < k; ++a)
      for (int b = a + 1; b < k; ++b)
        edges->emplace_back(std::min(s[i + a], s[i + b]), std::max(s[i + a], s[i + b]));
  std::sort(edges->begin(), edges->end());
  edges->erase(std::unique(edges->begin(), edges->end()), edges->end());
}

}  // namespace geom
}  // namespace mo
