// 3-D alpha complex boundary matrices with exact predicates — host part of libssq.
//
// Replaces phat-highclear/addons/alpha_3.cpp (CGAL Delaunay_triangulation_3 over an exact
// kernel) without CGAL:
//  * Delaunay triangulation by incremental insertion (Bowyer-Watson: locate by a visibility
//    walk, conflict region by breadth-first search, star re-triangulation), with an infinite
//    vertex for the hull;
//  * every predicate is exact: a floating-point evaluation with a forward error bound, and when
//    that cannot decide, the same polynomial evaluated exactly in fixed-width integers (the
//    input doubles scaled by a common power of two are integers of at most 120 bits);
//  * cospherical and cocircular ties are broken by symbolic perturbation of the lifted
//    coordinate, each point perturbed by a power of epsilon ranked by its lexicographic
//    (x, y, z) order (Devillers and Teillaud, "Perturbations and vertex removal in a 3D
//    Delaunay triangulation", SODA 2003).  The perturbed triangulation is unique: it does not
//    depend on the insertion order;
//  * alpha values follow alpha_3.cpp:260-352: a tetrahedron takes its squared circumradius; a
//    triangle takes the value of a tetrahedron it is attached to (the tetrahedron's fourth
//    vertex strictly inside the triangle's smallest circumsphere, :185-201,278-300), else its
//    own squared circumradius; an edge takes the smallest value of the triangles it is
//    attached to (the triangle's third vertex strictly inside the edge's diametral sphere,
//    :153-181,313-345), else a quarter of its squared length; vertices take 0.
//  * cells are sorted by (value, dimension) (alpha_3.cpp:141-151,354-356).  Values are
//    computed in binary128 and values that agree to 1e-26 relative are one value (the exact
//    kernel's ties: copied values, cospherical tetrahedra); ties of (value, dimension) are
//    broken by the vertex ids in lexicographic order, where the reference's std::sort leaves
//    them unspecified.
// Duplicate points are inserted once, as the triangulation does; vertex ids follow the first
// occurrence of each distinct point in the input.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ssq_internal.h"

namespace {

// ---- fixed-width two's complement integers for the exact fallback ----------------------
constexpr int kLimbs = 14;  // 896 bits: the largest polynomial below needs about 740
struct Big {
  uint64_t w[kLimbs];
};
using u128 = unsigned __int128;
using i128 = __int128;

Big big_from(i128 v) {
  Big r;
  const uint64_t fill = v < 0 ? ~0ull : 0ull;
  r.w[0] = (uint64_t)v;
  r.w[1] = (uint64_t)((u128)v >> 64);
  for (int i = 2; i < kLimbs; i++) r.w[i] = fill;
  return r;
}
Big big_add(const Big& a, const Big& b) {
  Big r;
  u128 c = 0;
  for (int i = 0; i < kLimbs; i++) {
    c += (u128)a.w[i] + b.w[i];
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}
Big big_neg(const Big& a) {
  Big r;
  u128 c = 1;
  for (int i = 0; i < kLimbs; i++) {
    c += (u128)(~a.w[i]);
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}
Big big_sub(const Big& a, const Big& b) { return big_add(a, big_neg(b)); }
bool big_negative(const Big& a) { return (a.w[kLimbs - 1] >> 63) != 0; }
int big_sign(const Big& a) {
  if (big_negative(a)) return -1;
  for (int i = 0; i < kLimbs; i++)
    if (a.w[i]) return 1;
  return 0;
}
Big big_mul(const Big& a, const Big& b) {
  const bool na = big_negative(a), nb = big_negative(b);
  const Big x = na ? big_neg(a) : a, y = nb ? big_neg(b) : b;
  int lx = kLimbs, ly = kLimbs;
  while (lx > 0 && x.w[lx - 1] == 0) lx--;
  while (ly > 0 && y.w[ly - 1] == 0) ly--;
  Big r;
  std::memset(r.w, 0, sizeof(r.w));
  for (int i = 0; i < lx; i++) {
    u128 c = 0;
    for (int j = 0; j < ly && i + j < kLimbs; j++) {
      c += (u128)x.w[i] * y.w[j] + r.w[i + j];
      r.w[i + j] = (uint64_t)c;
      c >>= 64;
    }
    if (i + ly < kLimbs) r.w[i + ly] = (uint64_t)c;
  }
  return na != nb ? big_neg(r) : r;
}

// ---- points: doubles, and the same points as exact integers ----------------------------
struct Geom {
  std::vector<double> x;  // 3 per point
  std::vector<i128> X;    // x * 2^-lo, exact
};

struct V3d {
  double x, y, z;
};
struct V3b {
  Big x, y, z;
};

inline V3d dsub(const Geom& G, int a, int b) {
  return {G.x[3 * a] - G.x[3 * b], G.x[3 * a + 1] - G.x[3 * b + 1], G.x[3 * a + 2] - G.x[3 * b + 2]};
}
inline V3b bsub(const Geom& G, int a, int b) {
  return {big_from(G.X[3 * a] - G.X[3 * b]), big_from(G.X[3 * a + 1] - G.X[3 * b + 1]),
          big_from(G.X[3 * a + 2] - G.X[3 * b + 2])};
}
inline Big bdot(const V3b& a, const V3b& b) {
  return big_add(big_add(big_mul(a.x, b.x), big_mul(a.y, b.y)), big_mul(a.z, b.z));
}
inline V3b bcross(const V3b& a, const V3b& b) {
  return {big_sub(big_mul(a.y, b.z), big_mul(a.z, b.y)), big_sub(big_mul(a.z, b.x), big_mul(a.x, b.z)),
          big_sub(big_mul(a.x, b.y), big_mul(a.y, b.x))};
}
inline Big bdet3(const V3b& a, const V3b& b, const V3b& c) { return bdot(a, bcross(b, c)); }

constexpr double kEps = 1.1102230246251565e-16;  // 2^-53

inline int dsign(double v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }

// sign det[q-p; r-p; s-p]: positive when (q-p, r-p, s-p) is right-handed
int orient3d(const Geom& G, int p, int q, int r, int s) {
  const V3d a = dsub(G, q, p), b = dsub(G, r, p), c = dsub(G, s, p);
  const double m1 = b.y * c.z - b.z * c.y, m2 = b.x * c.z - b.z * c.x, m3 = b.x * c.y - b.y * c.x;
  const double det = a.x * m1 - a.y * m2 + a.z * m3;
  const double mag = std::fabs(a.x) * (std::fabs(b.y * c.z) + std::fabs(b.z * c.y)) +
                     std::fabs(a.y) * (std::fabs(b.x * c.z) + std::fabs(b.z * c.x)) +
                     std::fabs(a.z) * (std::fabs(b.x * c.y) + std::fabs(b.y * c.x));
  if (std::fabs(det) > 64 * kEps * mag) return dsign(det);
  return big_sign(bdet3(bsub(G, q, p), bsub(G, r, p), bsub(G, s, p)));
}

// 2-D orientation of the projections on (x, y), else (y, z), else (x, z): a consistent
// orientation of the plane of three non-collinear points.
int orient2d_axes(const Geom& G, int p, int q, int r, int i, int j) {
  const double ax = G.x[3 * q + i] - G.x[3 * p + i], ay = G.x[3 * q + j] - G.x[3 * p + j];
  const double bx = G.x[3 * r + i] - G.x[3 * p + i], by = G.x[3 * r + j] - G.x[3 * p + j];
  const double det = ax * by - ay * bx, mag = std::fabs(ax * by) + std::fabs(ay * bx);
  if (std::fabs(det) > 16 * kEps * mag) return dsign(det);
  const Big Ax = big_from(G.X[3 * q + i] - G.X[3 * p + i]), Ay = big_from(G.X[3 * q + j] - G.X[3 * p + j]);
  const Big Bx = big_from(G.X[3 * r + i] - G.X[3 * p + i]), By = big_from(G.X[3 * r + j] - G.X[3 * p + j]);
  return big_sign(big_sub(big_mul(Ax, By), big_mul(Ay, Bx)));
}
int coplanar_orient(const Geom& G, int p, int q, int r) {
  int o = orient2d_axes(G, p, q, r, 0, 1);
  if (o) return o;
  o = orient2d_axes(G, p, q, r, 1, 2);
  if (o) return o;
  return orient2d_axes(G, p, q, r, 0, 2);
}

// sign of det[ p_i - p, |p_i - p|^2 ] (i = 0..3): negative when p is inside the sphere of a
// positively oriented p0..p3
int insphere_det(const Geom& G, const int* v, int p) {
  V3d d[4];
  double l[4], la[4];
  for (int i = 0; i < 4; i++) {
    d[i] = dsub(G, v[i], p);
    l[i] = d[i].x * d[i].x + d[i].y * d[i].y + d[i].z * d[i].z;
    la[i] = l[i];
  }
  // expansion along the lift column: sum_i (-1)^(i+3) l_i * det3(rows != i)
  auto det3 = [&](int a, int b, int c, double* mag) {
    const V3d &A = d[a], &B = d[b], &C = d[c];
    *mag = std::fabs(A.x) * (std::fabs(B.y * C.z) + std::fabs(B.z * C.y)) +
           std::fabs(A.y) * (std::fabs(B.x * C.z) + std::fabs(B.z * C.x)) +
           std::fabs(A.z) * (std::fabs(B.x * C.y) + std::fabs(B.y * C.x));
    return A.x * (B.y * C.z - B.z * C.y) - A.y * (B.x * C.z - B.z * C.x) + A.z * (B.x * C.y - B.y * C.x);
  };
  double m0, m1, m2, m3;
  const double D0 = det3(1, 2, 3, &m0), D1 = det3(0, 2, 3, &m1), D2 = det3(0, 1, 3, &m2), D3 = det3(0, 1, 2, &m3);
  const double det = -l[0] * D0 + l[1] * D1 - l[2] * D2 + l[3] * D3;
  const double mag = la[0] * m0 + la[1] * m1 + la[2] * m2 + la[3] * m3;
  if (std::fabs(det) > 256 * kEps * mag) return dsign(det);
  V3b B[4];
  Big L[4];
  for (int i = 0; i < 4; i++) {
    B[i] = bsub(G, v[i], p);
    L[i] = bdot(B[i], B[i]);
  }
  Big r = big_neg(big_mul(L[0], bdet3(B[1], B[2], B[3])));
  r = big_add(r, big_mul(L[1], bdet3(B[0], B[2], B[3])));
  r = big_sub(r, big_mul(L[2], bdet3(B[0], B[1], B[3])));
  r = big_add(r, big_mul(L[3], bdet3(B[0], B[1], B[2])));
  return big_sign(r);
}

// sign of |w-a|^2 |n|^2 - (w-a).U with n = ab x ac, U = |ac|^2 (n x ab) + |ab|^2 (ac x n):
// negative when w is strictly inside the smallest sphere through a, b, c
int tri_ball_sign(const Geom& G, int a, int b, int c, int w) {
  const V3d ab = dsub(G, b, a), ac = dsub(G, c, a), wa = dsub(G, w, a);
  auto cross = [](const V3d& u, const V3d& v) {
    return V3d{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
  };
  auto across = [](const V3d& u, const V3d& v) {
    return V3d{std::fabs(u.y * v.z) + std::fabs(u.z * v.y), std::fabs(u.z * v.x) + std::fabs(u.x * v.z),
               std::fabs(u.x * v.y) + std::fabs(u.y * v.x)};
  };
  auto absv = [](const V3d& u) { return V3d{std::fabs(u.x), std::fabs(u.y), std::fabs(u.z)}; };
  auto dot = [](const V3d& u, const V3d& v) { return u.x * v.x + u.y * v.y + u.z * v.z; };
  const V3d n = cross(ab, ac), nA = across(ab, ac);
  const double lab = dot(ab, ab), lac = dot(ac, ac), lwa = dot(wa, wa), nn = dot(n, n);
  const V3d c1 = cross(n, ab), c2 = cross(ac, n);
  const V3d U{lac * c1.x + lab * c2.x, lac * c1.y + lab * c2.y, lac * c1.z + lab * c2.z};
  const double val = lwa * nn - dot(wa, U);
  const V3d c1A = across(nA, absv(ab)), c2A = across(absv(ac), nA);
  const V3d UA{lac * c1A.x + lab * c2A.x, lac * c1A.y + lab * c2A.y, lac * c1A.z + lab * c2A.z};
  const double mag = lwa * dot(nA, nA) + dot(absv(wa), UA);
  if (std::fabs(val) > 512 * kEps * mag) return dsign(val);
  const V3b Ab = bsub(G, b, a), Ac = bsub(G, c, a), Wa = bsub(G, w, a);
  const V3b N = bcross(Ab, Ac);
  const Big Lab = bdot(Ab, Ab), Lac = bdot(Ac, Ac);
  const V3b C1 = bcross(N, Ab), C2 = bcross(Ac, N);
  const V3b Ub{big_add(big_mul(Lac, C1.x), big_mul(Lab, C2.x)), big_add(big_mul(Lac, C1.y), big_mul(Lab, C2.y)),
               big_add(big_mul(Lac, C1.z), big_mul(Lab, C2.z))};
  return big_sign(big_sub(big_mul(bdot(Wa, Wa), bdot(N, N)), bdot(Wa, Ub)));
}

// sign of (w-a).(w-b): negative when w is strictly inside the sphere with diameter ab
int edge_ball_sign(const Geom& G, int a, int b, int w) {
  const V3d p = dsub(G, w, a), q = dsub(G, w, b);
  const double val = p.x * q.x + p.y * q.y + p.z * q.z;
  const double mag = std::fabs(p.x * q.x) + std::fabs(p.y * q.y) + std::fabs(p.z * q.z);
  if (std::fabs(val) > 16 * kEps * mag) return dsign(val);
  return big_sign(bdot(bsub(G, w, a), bsub(G, w, b)));
}

// ---- Delaunay triangulation ------------------------------------------------------------
constexpr int kInf = -1;

struct Cell {
  int v[4];
  int nb[4];  // neighbour across the facet opposite v[i]
};

struct Delaunay {
  const Geom& G;
  std::vector<Cell> cells;
  std::vector<char> alive;
  std::vector<int> free_list;
  std::vector<uint32_t> conf_mark, done_mark;
  std::vector<int> rank;  // lexicographic rank of each point (the perturbation order)
  uint32_t stamp = 0;
  uint64_t rng = 0x2545F4914F6CDD1Dull;

  explicit Delaunay(const Geom& g) : G(g) {}

  int new_cell() {
    int c;
    if (!free_list.empty()) {
      c = free_list.back();
      free_list.pop_back();
      alive[c] = 1;
    } else {
      c = (int)cells.size();
      cells.push_back(Cell{});
      alive.push_back(1);
      conf_mark.push_back(0);
      done_mark.push_back(0);
    }
    return c;
  }
  static int index_of(const Cell& c, int v) {
    for (int i = 0; i < 4; i++)
      if (c.v[i] == v) return i;
    return -1;
  }

  // side of p with respect to the circumsphere of a positively oriented finite cell, with the
  // symbolic perturbation when p is on it: +1 inside (conflict), -1 outside
  int sphere_side(const int* v, int p) {
    const int d = insphere_det(G, v, p);
    if (d) return -d;
    // the dominant perturbation belongs to the lexicographically largest of the five points;
    // two terms always decide when v is not coplanar
    int pts[5] = {v[0], v[1], v[2], v[3], p};
    std::sort(pts, pts + 5, [&](int a, int b) { return rank[a] < rank[b]; });
    for (int k = 4; k > 2; k--) {
      const int q = pts[k];
      if (q == p) return -1;
      const int i = q == v[0] ? 0 : q == v[1] ? 1 : q == v[2] ? 2 : 3;
      int w[4] = {v[0], v[1], v[2], v[3]};
      w[i] = p;
      const int o = orient3d(G, w[0], w[1], w[2], w[3]);
      if (o) return o;
    }
    return -1;  // not reached for a non-degenerate cell
  }

  // p coplanar with the triangle a, b, c: +1 inside its circumcircle (with the perturbation
  // on the circle), -1 outside
  int circle_side(int a, int b, int c, int p) {
    const int s = tri_ball_sign(G, a, b, c, p);
    if (s) return -s;
    int pts[4] = {a, b, c, p};
    std::sort(pts, pts + 4, [&](int x, int y) { return rank[x] < rank[y]; });
    const int local = coplanar_orient(G, a, b, c);
    for (int k = 3; k > 0; k--) {
      const int q = pts[k];
      if (q == p) return -1;
      int o = 0;
      if (q == c) o = coplanar_orient(G, a, b, p);
      else if (q == b) o = coplanar_orient(G, a, p, c);
      else o = coplanar_orient(G, p, b, c);
      if (o) return o * local;
    }
    return -local;
  }

  bool in_conflict(int ci, int p) {
    const Cell& c = cells[ci];
    const int i = index_of(c, kInf);
    if (i < 0) return sphere_side(c.v, p) > 0;
    int w[4] = {c.v[0], c.v[1], c.v[2], c.v[3]};
    w[i] = p;
    const int o = orient3d(G, w[0], w[1], w[2], w[3]);
    if (o) return o > 0;
    int f[3], k = 0;
    for (int j = 0; j < 4; j++)
      if (j != i) f[k++] = c.v[j];
    return circle_side(f[0], f[1], f[2], p) > 0;
  }

  uint32_t next_rand() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (uint32_t)rng;
  }

  // a cell in conflict with p, by a visibility walk from `start` (facets tried from a random
  // one, which ends with probability 1 in any triangulation; past a step budget, a scan)
  int locate(int start, int p) {
    int c = start;
    for (int64_t steps = 0;; steps++) {
      if (steps > 64 + 4 * (int64_t)cells.size()) {
        for (int k = 0; k < (int)cells.size(); k++)
          if (alive[k] && in_conflict(k, p)) return k;
        return start;  // not reached: every point conflicts with some cell
      }
      const Cell& C = cells[c];
      const int i = index_of(C, kInf);
      if (i >= 0) {
        if (in_conflict(c, p)) return c;
        c = C.nb[i];
        continue;
      }
      const int r0 = (int)(next_rand() & 3);
      int moved = -1;
      for (int t = 0; t < 4 && moved < 0; t++) {
        const int k = (r0 + t) & 3;
        int w[4] = {C.v[0], C.v[1], C.v[2], C.v[3]};
        w[k] = p;
        if (orient3d(G, w[0], w[1], w[2], w[3]) < 0) moved = C.nb[k];
      }
      if (moved < 0) return c;  // p in the closed cell: strictly inside its circumsphere
      c = moved;
    }
  }

  struct Bnd {
    int v[4];   // the new cell's vertices
    int k;      // position of p
    int out;    // neighbour outside the cavity
    int out_i;  // its index of the shared facet
  };
  std::vector<int> stack, conflict;
  std::vector<Bnd> bnd;

  int insert(int p, int hint) {
    const int c0 = locate(hint, p);
    ++stamp;
    stack.assign(1, c0);
    conflict.assign(1, c0);
    conf_mark[c0] = stamp;
    bnd.clear();
    while (!stack.empty()) {
      const int c = stack.back();
      stack.pop_back();
      for (int k = 0; k < 4; k++) {
        const int n = cells[c].nb[k];
        if (conf_mark[n] == stamp) continue;
        if (done_mark[n] != stamp) {
          done_mark[n] = stamp;
          if (in_conflict(n, p)) {
            conf_mark[n] = stamp;
            stack.push_back(n);
            conflict.push_back(n);
            continue;
          }
        }
        Bnd b;
        std::memcpy(b.v, cells[c].v, sizeof(b.v));
        b.v[k] = p;
        b.k = k;
        b.out = n;
        b.out_i = -1;
        for (int j = 0; j < 4; j++)
          if (cells[n].nb[j] == c) b.out_i = j;
        bnd.push_back(b);
      }
    }
    for (int c : conflict) {
      alive[c] = 0;
      free_list.push_back(c);
    }
    // new cells; the facets through p are matched by their other two vertices
    struct Half {
      int64_t key;
      int cell, idx;
    };
    std::vector<Half> halves;
    halves.reserve(bnd.size() * 3);
    int any = -1;
    for (const Bnd& b : bnd) {
      const int nc = new_cell();
      Cell& C = cells[nc];
      std::memcpy(C.v, b.v, sizeof(C.v));
      C.nb[b.k] = b.out;
      cells[b.out].nb[b.out_i] = nc;
      for (int j = 0; j < 4; j++) {
        if (j == b.k) continue;
        int e[2], t = 0;
        for (int q = 0; q < 4; q++)
          if (q != j && q != b.k) e[t++] = C.v[q];
        const int lo = std::min(e[0], e[1]), hi = std::max(e[0], e[1]);
        halves.push_back({((int64_t)(lo + 1) << 32) | (int64_t)(uint32_t)(hi + 1), nc, j});
      }
      if (any < 0 || index_of(C, kInf) < 0) any = nc;
    }
    std::sort(halves.begin(), halves.end(), [](const Half& x, const Half& y) { return x.key < y.key; });
    if (halves.size() & 1) return -1;
    for (size_t i = 0; i + 1 < halves.size(); i += 2) {
      if (halves[i].key != halves[i + 1].key) return -1;
      cells[halves[i].cell].nb[halves[i].idx] = halves[i + 1].cell;
      cells[halves[i + 1].cell].nb[halves[i + 1].idx] = halves[i].cell;
    }
    return any;
  }

  // the first tetrahedron and its four hull cells
  void init(const int* q) {
    int v[4] = {q[0], q[1], q[2], q[3]};
    if (orient3d(G, v[0], v[1], v[2], v[3]) < 0) std::swap(v[2], v[3]);
    const int f = new_cell();
    std::memcpy(cells[f].v, v, sizeof(v));
    int inf[4];
    for (int i = 0; i < 4; i++) {
      inf[i] = new_cell();
      Cell& C = cells[inf[i]];
      std::memcpy(C.v, v, sizeof(v));
      C.v[i] = kInf;
      // replacing the infinite vertex by a point beyond the facet must give a positive cell
      const int a = (i + 1) & 3, b = (i + 2) & 3;
      std::swap(C.v[a], C.v[b]);
      cells[f].nb[i] = inf[i];
    }
    for (int i = 0; i < 4; i++) {
      Cell& C = cells[inf[i]];
      C.nb[index_of(C, kInf)] = f;
    }
    // hull cells share the facets through the infinite vertex
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        if (i == j) continue;
        Cell& A = cells[inf[i]];
        for (int k = 0; k < 4; k++) {
          if (A.v[k] == kInf) continue;
          // facet of A opposite A.v[k] contains kInf and two finite vertices; B shares it if B
          // has those two vertices (B lacks exactly the finite vertex v[j])
          int e[2], t = 0;
          for (int m = 0; m < 4; m++)
            if (m != k && A.v[m] != kInf) e[t++] = A.v[m];
          const Cell& B = cells[inf[j]];
          if (index_of(B, e[0]) >= 0 && index_of(B, e[1]) >= 0) A.nb[k] = inf[j];
        }
      }
  }
};

// Morton order of points quantized to 21 bits per axis: nearby insertions, short walks
std::vector<int> spatial_order(const std::vector<double>& x, const std::vector<int>& ids) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int p : ids)
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], x[3 * p + k]);
      hi[k] = std::max(hi[k], x[3 * p + k]);
    }
  auto spread = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
  };
  std::vector<std::pair<uint64_t, int>> key;
  key.reserve(ids.size());
  for (int p : ids) {
    uint64_t m = 0;
    for (int k = 0; k < 3; k++) {
      const double span = hi[k] > lo[k] ? hi[k] - lo[k] : 1.0;
      const uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, (x[3 * p + k] - lo[k]) / span * 2097151.0));
      m |= spread(q) << k;
    }
    key.push_back({m, p});
  }
  std::sort(key.begin(), key.end());
  std::vector<int> out;
  out.reserve(ids.size());
  for (auto& kv : key) out.push_back(kv.second);
  return out;
}

// distinct points (first occurrence order) and their exact integer copies; false when the
// coordinates span more than 120 bits of one binary grid or are not finite
bool prepare(const double* pts, int64_t n, Geom& G) {
  std::vector<int64_t> idx((size_t)n);
  std::iota(idx.begin(), idx.end(), 0);
  for (int64_t i = 0; i < 3 * n; i++)
    if (!std::isfinite(pts[i])) return false;
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    for (int k = 0; k < 3; k++)
      if (pts[3 * a + k] != pts[3 * b + k]) return pts[3 * a + k] < pts[3 * b + k];
    return false;
  });
  std::vector<char> first((size_t)n, 0);
  for (int64_t t = 0; t < n; t++) {
    const int64_t i = idx[t];
    const int64_t j = t ? idx[t - 1] : 0;
    first[i] = t == 0 || pts[3 * i] != pts[3 * j] || pts[3 * i + 1] != pts[3 * j + 1] || pts[3 * i + 2] != pts[3 * j + 2];
  }
  G.x.clear();
  for (int64_t i = 0; i < n; i++)
    if (first[i]) G.x.insert(G.x.end(), pts + 3 * i, pts + 3 * i + 3);
  int lo = 1 << 30, hi = -(1 << 30);
  for (double v : G.x) {
    if (v == 0) continue;
    int e;
    (void)std::frexp(v, &e);  // v = m 2^e, 0.5 <= |m| < 1: an integer times 2^(e-53)
    lo = std::min(lo, e - 53);
    hi = std::max(hi, e);
  }
  if (lo > hi) lo = hi = 0;
  if (hi - lo > 120) return false;
  G.X.resize(G.x.size());
  for (size_t i = 0; i < G.x.size(); i++) {
    const double v = G.x[i];
    if (v == 0) {
      G.X[i] = 0;
      continue;
    }
    int e;
    const double m = std::frexp(v, &e);
    const int64_t M = (int64_t)std::ldexp(m, 53);  // exact
    const u128 mag = (u128)(M < 0 ? -M : M) << ((e - 53) - lo);  // below 2^120
    G.X[i] = M < 0 ? -(i128)mag : (i128)mag;
  }
  return true;
}

// Delaunay tetrahedra (finite cells, positively oriented) of the distinct points in G
int triangulate(const Geom& G, std::vector<Cell>& out_cells, std::vector<char>& out_alive, std::string* err) {
  const int n = (int)(G.x.size() / 3);
  Delaunay D(G);
  D.rank.resize(n);
  {
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
      for (int k = 0; k < 3; k++)
        if (G.x[3 * a + k] != G.x[3 * b + k]) return G.x[3 * a + k] < G.x[3 * b + k];
      return false;
    });
    for (int i = 0; i < n; i++) D.rank[ord[i]] = i;
  }
  // four affinely independent points
  int q[4] = {0, -1, -1, -1};
  for (int i = 1; i < n && q[1] < 0; i++) q[1] = i;  // distinct already
  for (int i = 1; i < n && q[2] < 0 && q[1] >= 0; i++) {
    if (i == q[1]) continue;
    if (coplanar_orient(G, q[0], q[1], i) != 0) q[2] = i;
  }
  for (int i = 1; i < n && q[3] < 0 && q[2] >= 0; i++) {
    if (i == q[1] || i == q[2]) continue;
    if (orient3d(G, q[0], q[1], q[2], i) != 0) q[3] = i;
  }
  if (q[3] < 0) {
    *err = "points do not span three dimensions";
    return SSQ_EINVAL;
  }
  D.init(q);
  std::vector<int> rest;
  rest.reserve(n);
  for (int i = 0; i < n; i++)
    if (i != q[0] && i != q[1] && i != q[2] && i != q[3]) rest.push_back(i);
  int hint = 0;
  for (int p : spatial_order(G.x, rest)) {
    hint = D.insert(p, hint);
    if (hint < 0) {
      *err = "inconsistent cavity (internal error)";
      return SSQ_EDEVICE;
    }
  }
  // every finite cell positively oriented, every neighbour relation symmetric
  for (size_t c = 0; c < D.cells.size(); c++) {
    if (!D.alive[c]) continue;
    const Cell& C = D.cells[c];
    for (int k = 0; k < 4; k++) {
      const int n = C.nb[k];
      if (n < 0 || !D.alive[n]) {
        *err = "dangling neighbour (internal error)";
        return SSQ_EDEVICE;
      }
      bool back = false;
      for (int j = 0; j < 4; j++) back |= D.cells[n].nb[j] == (int)c;
      if (!back) {
        *err = "asymmetric neighbours (internal error)";
        return SSQ_EDEVICE;
      }
    }
    if (Delaunay::index_of(C, kInf) < 0 && orient3d(G, C.v[0], C.v[1], C.v[2], C.v[3]) <= 0) {
      *err = "a flat or inverted tetrahedron (internal error)";
      return SSQ_EDEVICE;
    }
  }
  out_cells.swap(D.cells);
  out_alive.swap(D.alive);
  return SSQ_OK;
}

using f128 = __float128;

f128 sq_circumradius_tet(const Geom& G, const int* v) {
  f128 a[3], b[3], c[3];
  for (int k = 0; k < 3; k++) {
    const f128 o = G.x[3 * v[0] + k];
    a[k] = (f128)G.x[3 * v[1] + k] - o;
    b[k] = (f128)G.x[3 * v[2] + k] - o;
    c[k] = (f128)G.x[3 * v[3] + k] - o;
  }
  auto cross = [](const f128* u, const f128* w, f128* r) {
    r[0] = u[1] * w[2] - u[2] * w[1];
    r[1] = u[2] * w[0] - u[0] * w[2];
    r[2] = u[0] * w[1] - u[1] * w[0];
  };
  f128 bc[3], ca[3], ab[3];
  cross(b, c, bc);
  cross(c, a, ca);
  cross(a, b, ab);
  const f128 la = a[0] * a[0] + a[1] * a[1] + a[2] * a[2], lb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2],
             lc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  const f128 den = 2 * (a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2]);
  f128 r2 = 0;
  for (int k = 0; k < 3; k++) {
    const f128 u = (la * bc[k] + lb * ca[k] + lc * ab[k]) / den;
    r2 += u * u;
  }
  return r2;
}
f128 sq_circumradius_tri(const Geom& G, int a, int b, int c) {
  f128 u[3], w[3], z[3];
  for (int k = 0; k < 3; k++) {
    u[k] = (f128)G.x[3 * b + k] - G.x[3 * a + k];
    w[k] = (f128)G.x[3 * c + k] - G.x[3 * a + k];
    z[k] = (f128)G.x[3 * c + k] - G.x[3 * b + k];
  }
  const f128 n0 = u[1] * w[2] - u[2] * w[1], n1 = u[2] * w[0] - u[0] * w[2], n2 = u[0] * w[1] - u[1] * w[0];
  const f128 lu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], lw = w[0] * w[0] + w[1] * w[1] + w[2] * w[2],
             lz = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
  return lu * lw * lz / (4 * (n0 * n0 + n1 * n1 + n2 * n2));
}
f128 sq_half_length(const Geom& G, int a, int b) {
  f128 s = 0;
  for (int k = 0; k < 3; k++) {
    const f128 d = (f128)G.x[3 * b + k] - G.x[3 * a + k];
    s += d * d;
  }
  return s / 4;
}

}  // namespace

namespace ssq {

int delaunay_tets(const double* pts, int64_t npts, std::vector<int32_t>& tets, std::string* err) {
  Geom G;
  if (npts < 4 || npts >= (int64_t(1) << 30) || !prepare(pts, npts, G)) {
    *err = "need at least 4 finite points whose coordinates span at most 120 bits of one binary grid";
    return SSQ_EINVAL;
  }
  std::vector<Cell> cells;
  std::vector<char> alive;
  int rc = triangulate(G, cells, alive, err);
  if (rc) return rc;
  tets.clear();
  for (size_t c = 0; c < cells.size(); c++) {
    if (!alive[c] || Delaunay::index_of(cells[c], kInf) >= 0) continue;
    int v[4] = {cells[c].v[0], cells[c].v[1], cells[c].v[2], cells[c].v[3]};
    std::sort(v, v + 4);
    tets.insert(tets.end(), v, v + 4);
  }
  return SSQ_OK;
}

int alpha_build(const double* pts, int64_t npts, double max_value, int dim_major, ssq_matrix* out, double** values,
                std::string* err) {
  Geom G;
  if (npts < 4 || npts >= (int64_t(1) << 30) || !prepare(pts, npts, G)) {
    *err = "need at least 4 finite points whose coordinates span at most 120 bits of one binary grid";
    return SSQ_EINVAL;
  }
  std::vector<Cell> cells;
  std::vector<char> alive;
  int rc = triangulate(G, cells, alive, err);
  if (rc) return rc;
  const int nv = (int)(G.x.size() / 3);
  // finite cells
  std::vector<int> cell_id(cells.size(), -1), tet_cell;
  for (size_t c = 0; c < cells.size(); c++)
    if (alive[c] && Delaunay::index_of(cells[c], kInf) < 0) {
      cell_id[c] = (int)tet_cell.size();
      tet_cell.push_back((int)c);
    }
  const int nt = (int)tet_cell.size();
  std::vector<f128> vT((size_t)nt);
  for (int t = 0; t < nt; t++) vT[t] = sq_circumradius_tet(G, cells[tet_cell[t]].v);
  // triangles: each finite facet once; its value by attachment to either cell
  std::vector<std::array<int, 3>> F;
  std::vector<f128> vF;
  std::vector<std::array<int, 4>> tet_faces((size_t)nt);
  for (int t = 0; t < nt; t++) {
    const int c = tet_cell[t];
    const Cell& C = cells[c];
    for (int k = 0; k < 4; k++) {
      const int nbc = C.nb[k];
      const int nt2 = cell_id[nbc];
      if (nt2 >= 0 && nt2 < t) continue;  // the other cell owns it
      std::array<int, 3> f;
      int m = 0;
      for (int j = 0; j < 4; j++)
        if (j != k) f[m++] = C.v[j];
      std::sort(f.begin(), f.end());
      const int fid = (int)F.size();
      F.push_back(f);
      f128 val;
      if (tri_ball_sign(G, f[0], f[1], f[2], C.v[k]) < 0) {
        val = vT[t];
      } else if (nt2 >= 0) {
        const Cell& N = cells[nbc];
        int apex = -1;
        for (int j = 0; j < 4; j++)
          if (N.v[j] != f[0] && N.v[j] != f[1] && N.v[j] != f[2]) apex = N.v[j];
        val = tri_ball_sign(G, f[0], f[1], f[2], apex) < 0 ? vT[nt2] : sq_circumradius_tri(G, f[0], f[1], f[2]);
      } else {
        val = sq_circumradius_tri(G, f[0], f[1], f[2]);
      }
      vF.push_back(val);
      tet_faces[t][k] = fid;
      if (nt2 >= 0) {
        const Cell& N = cells[nbc];
        for (int j = 0; j < 4; j++)
          if (N.nb[j] == c) tet_faces[nt2][j] = fid;
      }
    }
  }
  const int nf = (int)F.size();
  // edges: each with its triangles and their third vertices
  struct EInc {
    int64_t key;
    int tri, third;
  };
  std::vector<EInc> inc;
  inc.reserve((size_t)nf * 3);
  for (int f = 0; f < nf; f++) {
    const auto& t = F[f];
    inc.push_back({((int64_t)t[0] << 32) | t[1], f, t[2]});
    inc.push_back({((int64_t)t[0] << 32) | t[2], f, t[1]});
    inc.push_back({((int64_t)t[1] << 32) | t[2], f, t[0]});
  }
  std::sort(inc.begin(), inc.end(), [](const EInc& a, const EInc& b) {
    return a.key != b.key ? a.key < b.key : a.tri < b.tri;
  });
  std::vector<std::array<int, 2>> E;
  std::vector<f128> vE;
  std::vector<std::array<int, 3>> tri_edges((size_t)nf);
  std::vector<int> tri_edge_fill((size_t)nf, 0);
  for (size_t i = 0; i < inc.size();) {
    size_t j = i;
    while (j < inc.size() && inc[j].key == inc[i].key) j++;
    const int a = (int)(inc[i].key >> 32), b = (int)(inc[i].key & 0xffffffff);
    const int eid = (int)E.size();
    E.push_back({a, b});
    bool att = false;
    f128 val = 0;
    for (size_t k = i; k < j; k++) {
      if (edge_ball_sign(G, a, b, inc[k].third) < 0) {
        val = att ? std::min(val, vF[inc[k].tri]) : vF[inc[k].tri];
        att = true;
      }
      tri_edges[inc[k].tri][tri_edge_fill[inc[k].tri]++] = eid;
    }
    vE.push_back(att ? val : sq_half_length(G, a, b));
    i = j;
  }
  const int ne = (int)E.size();
  // values are made non-decreasing from faces to cofaces (a no-op on the reference's data;
  // it keeps every kept cell's faces before it)
  for (int f = 0; f < nf; f++)
    for (int e : tri_edges[f]) vF[f] = std::max(vF[f], vE[e]);
  for (int t = 0; t < nt; t++)
    for (int f : tet_faces[t]) vT[t] = std::max(vT[t], vF[f]);
  // order: (value class, dimension, vertex ids)
  struct Key {
    f128 val;
    int dim, id;
  };
  std::vector<Key> keys;
  const f128 cut = (f128)max_value;
  const bool all = !(max_value >= 0) || std::isinf(max_value);
  for (int v = 0; v < nv; v++) keys.push_back({0, 0, v});
  for (int e = 0; e < ne; e++)
    if (all || vE[e] <= cut) keys.push_back({vE[e], 1, e});
  for (int f = 0; f < nf; f++)
    if (all || vF[f] <= cut) keys.push_back({vF[f], 2, f});
  for (int t = 0; t < nt; t++)
    if (all || vT[t] <= cut) keys.push_back({vT[t], 3, t});
  // value classes: values in increasing order (as double-double pairs, 106 bits), a new class
  // where the relative gap exceeds 1e-26
  struct VK {
    double hi, lo;
    uint32_t i;
  };
  std::vector<VK> vk(keys.size());
  for (size_t i = 0; i < keys.size(); i++) {
    const double hi = (double)keys[i].val;
    vk[i] = {hi, (double)(keys[i].val - (f128)hi), (uint32_t)i};
  }
  std::sort(vk.begin(), vk.end(), [](const VK& x, const VK& y) { return x.hi != y.hi ? x.hi < y.hi : x.lo < y.lo; });
  std::vector<int64_t> cls(keys.size());
  int64_t cur = 0;
  for (size_t t = 0; t < vk.size(); t++) {
    if (t > 0) {
      const double gap = (vk[t].hi - vk[t - 1].hi) + (vk[t].lo - vk[t - 1].lo);
      if (gap > 1e-26 * std::fabs(vk[t].hi)) cur++;
    }
    cls[vk[t].i] = cur;
  }
  auto class_of_key = [&](size_t i) { return cls[i]; };
  std::vector<int> tvs((size_t)nt * 4);
  for (int t = 0; t < nt; t++) {
    const Cell& C = cells[tet_cell[t]];
    int v[4] = {C.v[0], C.v[1], C.v[2], C.v[3]};
    std::sort(v, v + 4);
    std::copy(v, v + 4, &tvs[(size_t)t * 4]);
  }
  auto verts = [&](const Key& k, int* v) {
    switch (k.dim) {
      case 0: v[0] = k.id; return 1;
      case 1: v[0] = E[k.id][0]; v[1] = E[k.id][1]; return 2;
      case 2: std::copy(F[k.id].begin(), F[k.id].end(), v); return 3;
      default: std::copy(&tvs[(size_t)k.id * 4], &tvs[(size_t)k.id * 4] + 4, v); return 4;
    }
  };
  std::vector<std::pair<int64_t, size_t>> order;
  order.reserve(keys.size());
  for (size_t i = 0; i < keys.size(); i++) order.push_back({class_of_key(i), i});
  std::sort(order.begin(), order.end(), [&](const std::pair<int64_t, size_t>& a, const std::pair<int64_t, size_t>& b) {
    const Key &x = keys[a.second], &y = keys[b.second];
    if (dim_major && x.dim != y.dim) return x.dim < y.dim;
    if (a.first != b.first) return a.first < b.first;
    if (x.dim != y.dim) return x.dim < y.dim;
    int vx[4], vy[4];
    const int n = verts(x, vx);
    verts(y, vy);
    return std::lexicographical_compare(vx, vx + n, vy, vy + n);
  });
  // positions and boundary columns
  std::vector<int64_t> posV((size_t)nv, -1), posE((size_t)ne, -1), posF((size_t)nf, -1), posT((size_t)nt, -1);
  for (size_t i = 0; i < order.size(); i++) {
    const Key& k = keys[order[i].second];
    (k.dim == 0 ? posV : k.dim == 1 ? posE : k.dim == 2 ? posF : posT)[k.id] = (int64_t)i;
  }
  const int64_t N = (int64_t)order.size();
  if (N >= 0x7fffffffLL) {
    *err = "more than 2^31 - 1 cells";
    return SSQ_EINVAL;
  }
  std::vector<int64_t> colptr((size_t)N + 1, 0);
  std::vector<int32_t> rows, dims((size_t)N);
  rows.reserve((size_t)N * 3);
  for (int64_t i = 0; i < N; i++) {
    const Key& k = keys[order[i].second];
    dims[i] = k.dim;
    int64_t b[4];
    int m = 0;
    if (k.dim == 1) {
      b[m++] = posV[E[k.id][0]];
      b[m++] = posV[E[k.id][1]];
    } else if (k.dim == 2) {
      for (int e : tri_edges[k.id]) b[m++] = posE[e];
    } else if (k.dim == 3) {
      for (int f : tet_faces[k.id]) b[m++] = posF[f];
    }
    std::sort(b, b + m);
    for (int j = 0; j < m; j++) {
      if (b[j] < 0 || b[j] >= i) {
        *err = "a face of a kept cell is missing or ordered after it";
        return SSQ_EFORMAT;
      }
      rows.push_back((int32_t)b[j]);
    }
    colptr[i + 1] = (int64_t)rows.size();
  }
  if (values) {
    *values = (double*)std::malloc(sizeof(double) * (size_t)std::max<int64_t>(N, 1));
    if (!*values) return SSQ_ENOMEM;
    for (int64_t i = 0; i < N; i++) (*values)[i] = (double)keys[order[i].second].val;
  }
  const int rc2 = matrix_from_vectors(colptr, rows, dims, out);
  if (rc2 && values) {
    std::free(*values);
    *values = nullptr;
  }
  return rc2;
}

}  // namespace ssq

extern "C" int ssq_alpha_build(const double* pts, int64_t npts, double max_value, int dim_major, ssq_matrix* out,
                               double** values) {
  if (!out || (npts > 0 && !pts) || npts < 0 || (dim_major & ~1)) return SSQ_EINVAL;
  std::memset(out, 0, sizeof(*out));
  if (values) *values = nullptr;
  std::string err;
  return ssq::alpha_build(pts, npts, max_value, dim_major, out, values, &err);
}

extern "C" void ssq_free(void* p) { std::free(p); }

extern "C" int ssq_delaunay_3(const double* pts, int64_t npts, int32_t* tets, int64_t cap, int64_t* ntets) {
  if (!ntets || (npts > 0 && !pts) || npts < 0 || cap < 0 || (cap > 0 && !tets)) return SSQ_EINVAL;
  std::vector<int32_t> t;
  std::string err;
  const int rc = ssq::delaunay_tets(pts, npts, t, &err);
  if (rc) return rc;
  *ntets = (int64_t)(t.size() / 4);
  if (*ntets > cap) return SSQ_ENOMEM;
  std::copy(t.begin(), t.end(), tets);
  return SSQ_OK;
}
