// FP64 CPU oracle for the visibility-edge and plane-score definitions.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this
// library; tests, the smoke check and the bench's CPU-baseline leg load it
// as the checker.  It restates, on float32 inputs promoted exactly to
// float64 and decided with exact orientation predicates:
//
//   * lib/ptx/ray_segments.cu:49-111 and lib/src/intersections.cpp:67-73,
//     327-385: pair order k -> (i<j), the new-vertex mask (plus the cap
//     rule of mask bits 1-7, no counterpart in the reference),
//     zero-length segments, and the combine rule
//     edge <=> hits(cage) && !hits(self) in cage mode;
//   * the midpoint-parity form of the same rule (cage absent):
//     edge <=> !hits(self) && midpoint outside the closed self mesh;
//   * lib/cuda/plane_intersections.cu:43-66: dead zone 1e-6, strict
//     crossing, score = sum of crossing edge lengths (FP64 form and the
//     float32 / Q32.32 fixed-point form);
//   * lib/src/process.cpp:132-134: argmax, first maximum wins;
//   * lib/src/process.cpp:28-72: bisector candidate planes drawn with
//     std::mt19937 + std::uniform_int_distribution<int>.
//
// "hits" is defined without tolerances (the reference's t in [0.001, len]
// window belongs to the closed OptiX traversal and is not reproducible):
// a segment hits a triangle iff the closed segment and the closed triangle
// share a point other than an endpoint that is itself a vertex of that
// triangle; a segment equal to a triangle edge hits it; triangles with
// collinear vertices are never hit.  The midpoint parity ray runs along +x
// with the query point perturbed symbolically by (0, eps, eps^2).
//
// Exact signs: every orientation determinant of float32 coordinates is a
// signed sum of products of at most three float32 values; each product is
// represented exactly by at most two doubles (two-product via fma), and the
// sum is formed exactly with Shewchuk's Grow-Expansion (J. R. Shewchuk,
// "Adaptive Precision Floating-Point Arithmetic and Fast Robust Geometric
// Predicates", DCG 18, 1997, section 2).  A float64 filter answers first
// when its error bound allows.
//
// Parity status: the predicates are pinned against exact rational
// arithmetic (tests/test_oracle_predicates.py); the visibility definition
// against closed-form meshes, brute force, symmetry and rotation
// equivariance; the scores against closed forms, the FP64 form and
// permutation invariance.  OptiX's own hit decisions and the float
// atomicAdd order of the reference are not reproducible and are not pinned.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- exact sums
struct ExactSum {
  double c[192];
  int n = 0;
  // Grow-Expansion with zero elimination; c[] stays nonoverlapping and
  // ordered by increasing magnitude, so the last component has the sign of
  // the exact sum.
  void add(double b) {
    if (b == 0.0) return;
    double q = b;
    int m = 0;
    for (int i = 0; i < n; ++i) {
      double x = q + c[i];
      double bv = x - q;
      double av = x - bv;
      double err = (q - av) + (c[i] - bv);
      if (err != 0.0) c[m++] = err;
      q = x;
    }
    if (q != 0.0) c[m++] = q;
    n = m;
  }
  void add_prod2(double x, double y) { add(x * y); }  // exact for float32 x, y
  void add_prod3(double x, double y, double z) {       // x, y, z float32
    double p = x * y;                                  // exact (48 bits)
    double hi = p * z;
    double lo = std::fma(p, z, -hi);                   // exact remainder
    add(hi);
    add(lo);
  }
  int sign() const { return n == 0 ? 0 : (c[n - 1] > 0.0 ? 1 : -1); }
};

struct P3 {
  double x, y, z;
};
inline P3 ld(const float* v) { return P3{(double)v[0], (double)v[1], (double)v[2]}; }
inline double co(const P3& p, int k) { return k == 0 ? p.x : (k == 1 ? p.y : p.z); }
inline bool same(const P3& a, const P3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// 4x4 permutations: perm[t][col] = row used for column col (col 3 is the
// all-ones column), with their signs.
struct Perm4 {
  int r[24][4];
  int s[24];
  Perm4() {
    int t = 0;
    int a[4] = {0, 1, 2, 3};
    do {
      int inv = 0;
      for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j)
          if (a[i] > a[j]) ++inv;
      for (int i = 0; i < 4; ++i) r[t][i] = a[i];
      s[t] = (inv & 1) ? -1 : 1;
      ++t;
    } while (std::next_permutation(a, a + 4));
  }
};
const Perm4 PERM4;

// Adds sgn * det[[a 1],[b 1],[c 1],[d 1]] (rows x, y, z, 1) to acc.
void acc_det4(ExactSum& acc, const P3& a, const P3& b, const P3& c, const P3& d, int sgn) {
  const P3* row[4] = {&a, &b, &c, &d};
  for (int t = 0; t < 24; ++t) {
    int s = PERM4.s[t] * sgn;
    double x = row[PERM4.r[t][0]]->x, y = row[PERM4.r[t][1]]->y, z = row[PERM4.r[t][2]]->z;
    if (s > 0)
      acc.add_prod3(x, y, z);
    else
      acc.add_prod3(-x, y, z);
  }
}

// Float64 evaluation of det[a-d, b-d, c-d] with a bound on its error.
double o3_approx(const P3& a, const P3& b, const P3& c, const P3& d, double& bound) {
  double adx = a.x - d.x, ady = a.y - d.y, adz = a.z - d.z;
  double bdx = b.x - d.x, bdy = b.y - d.y, bdz = b.z - d.z;
  double cdx = c.x - d.x, cdy = c.y - d.y, cdz = c.z - d.z;
  double m1 = bdy * cdz - bdz * cdy, m2 = cdy * adz - cdz * ady, m3 = ady * bdz - adz * bdy;
  double det = adx * m1 + bdx * m2 + cdx * m3;
  double perm = std::fabs(adx) * (std::fabs(bdy * cdz) + std::fabs(bdz * cdy)) +
                std::fabs(bdx) * (std::fabs(cdy * adz) + std::fabs(cdz * ady)) +
                std::fabs(cdx) * (std::fabs(ady * bdz) + std::fabs(adz * bdy));
  bound = 1e-14 * perm;
  return det;
}

// Exact sign of orient3d(a,b,c,d) = det[a-d, b-d, c-d].
int orient3d(const P3& a, const P3& b, const P3& c, const P3& d) {
  double bound;
  double det = o3_approx(a, b, c, d, bound);
  if (det > bound) return 1;
  if (-det > bound) return -1;
  ExactSum acc;
  acc_det4(acc, a, b, c, d, 1);
  return acc.sign();
}

// Exact sign of orient3d(a,b,c,(p+q)/2) (equal to that of o(p) + o(q)).
int orient3d_mid(const P3& a, const P3& b, const P3& c, const P3& p, const P3& q) {
  double b1, b2;
  double d1 = o3_approx(a, b, c, p, b1);
  double d2 = o3_approx(a, b, c, q, b2);
  double s = d1 + d2;
  double bound = (b1 + b2) * 1.0000001 + 1e-15 * std::fabs(s);
  if (s > bound) return 1;
  if (-s > bound) return -1;
  ExactSum acc;
  acc_det4(acc, a, b, c, p, 1);
  acc_det4(acc, a, b, c, q, 1);
  return acc.sign();
}

// Projection dropping axis k: (u, v) = (k==0: y,z  k==1: z,x  k==2: x,y).
inline int ax_u(int k) { return k == 0 ? 1 : (k == 1 ? 2 : 0); }
inline int ax_v(int k) { return k == 0 ? 2 : (k == 1 ? 0 : 1); }

// Exact sign of det[[a.u a.v 1],[b.u b.v 1],[c.u c.v 1]] (standard orient2d).
void acc_det3(ExactSum& acc, const P3& a, const P3& b, const P3& c, int k, int sgn) {
  int u = ax_u(k), v = ax_v(k);
  const P3* row[3] = {&a, &b, &c};
  static const int pr[6][2] = {{0, 1}, {0, 2}, {1, 0}, {1, 2}, {2, 0}, {2, 1}};
  static const int ps[6] = {1, -1, -1, 1, 1, -1};
  for (int t = 0; t < 6; ++t) {
    double x = co(*row[pr[t][0]], u), y = co(*row[pr[t][1]], v);
    acc.add_prod2(ps[t] * sgn > 0 ? x : -x, y);
  }
}

double o2_approx(const P3& a, const P3& b, const P3& c, int k, double& bound) {
  int u = ax_u(k), v = ax_v(k);
  double acu = co(a, u) - co(c, u), acv = co(a, v) - co(c, v);
  double bcu = co(b, u) - co(c, u), bcv = co(b, v) - co(c, v);
  double det = acu * bcv - acv * bcu;
  bound = 1e-14 * (std::fabs(acu * bcv) + std::fabs(acv * bcu));
  return det;
}

int orient2d(const P3& a, const P3& b, const P3& c, int k) {
  double bound;
  double det = o2_approx(a, b, c, k, bound);
  if (det > bound) return 1;
  if (-det > bound) return -1;
  ExactSum acc;
  acc_det3(acc, a, b, c, k, 1);
  return acc.sign();
}

// orient2d(a, b, (p+q)/2) in projection k, exact sign.
int orient2d_mid(const P3& a, const P3& b, const P3& p, const P3& q, int k) {
  double b1, b2;
  double d1 = o2_approx(a, b, p, k, b1);
  double d2 = o2_approx(a, b, q, k, b2);
  double s = d1 + d2;
  double bound = (b1 + b2) * 1.0000001 + 1e-15 * std::fabs(s);
  if (s > bound) return 1;
  if (-s > bound) return -1;
  ExactSum acc;
  acc_det3(acc, a, b, p, k, 1);
  acc_det3(acc, a, b, q, k, 1);
  return acc.sign();
}

// ------------------------------------------------- segment/triangle contact
// Projection axis with exact nonzero orient2d(a,b,c); -1 if degenerate.
int proj_axis(const P3& a, const P3& b, const P3& c) {
  double nx = (b.y - a.y) * (c.z - a.z) - (b.z - a.z) * (c.y - a.y);
  double ny = (b.z - a.z) * (c.x - a.x) - (b.x - a.x) * (c.z - a.z);
  double nz = (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
  int order[3] = {0, 1, 2};
  double mag[3] = {std::fabs(nx), std::fabs(ny), std::fabs(nz)};
  std::sort(order, order + 3, [&](int i, int j) { return mag[i] > mag[j]; });
  for (int t = 0; t < 3; ++t)
    if (orient2d(a, b, c, order[t]) != 0) return order[t];
  return -1;
}

bool tri_contains_closed(const P3& a, const P3& b, const P3& c, const P3& x, int k) {
  int s = orient2d(a, b, c, k);
  return s * orient2d(a, b, x, k) >= 0 && s * orient2d(b, c, x, k) >= 0 &&
         s * orient2d(c, a, x, k) >= 0;
}

bool on_seg_collinear(const P3& p, const P3& q, const P3& x, int k) {
  int u = ax_u(k), v = ax_v(k);
  return std::min(co(p, u), co(q, u)) <= co(x, u) && co(x, u) <= std::max(co(p, u), co(q, u)) &&
         std::min(co(p, v), co(q, v)) <= co(x, v) && co(x, v) <= std::max(co(p, v), co(q, v));
}

bool seg_seg_closed(const P3& p, const P3& q, const P3& a, const P3& b, int k) {
  int o1 = orient2d(p, q, a, k), o2 = orient2d(p, q, b, k);
  int o3 = orient2d(a, b, p, k), o4 = orient2d(a, b, q, k);
  if (o1 * o2 < 0 && o3 * o4 < 0) return true;
  if (o1 == 0 && on_seg_collinear(p, q, a, k)) return true;
  if (o2 == 0 && on_seg_collinear(p, q, b, k)) return true;
  if (o3 == 0 && on_seg_collinear(a, b, p, k)) return true;
  if (o4 == 0 && on_seg_collinear(a, b, q, k)) return true;
  return false;
}

// Does the closed segment pq touch the closed triangle abc anywhere other
// than at an endpoint that is a vertex of the triangle?
bool contact(const P3& p, const P3& q, const P3& a, const P3& b, const P3& c) {
  const P3* t[3] = {&a, &b, &c};
  int ip = -1, iq = -1;
  for (int i = 0; i < 3; ++i) {
    if (ip < 0 && same(p, *t[i])) ip = i;
    if (iq < 0 && same(q, *t[i])) iq = i;
  }
  if (ip >= 0 && iq >= 0) return true;  // the segment is a triangle edge
  if (ip >= 0 || iq >= 0) {
    const int iv = ip >= 0 ? ip : iq;
    const P3& o = ip >= 0 ? q : p;  // the endpoint that is not a vertex
    if (orient3d(a, b, c, o) != 0) return false;
    int k = proj_axis(a, b, c);
    if (k < 0) return false;
    const P3& v0 = *t[iv];
    const P3& v1 = *t[(iv + 1) % 3];
    const P3& v2 = *t[(iv + 2) % 3];
    int s = orient2d(v0, v1, v2, k);
    return s * orient2d(v0, v1, o, k) >= 0 && s * orient2d(v0, v2, o, k) <= 0;
  }
  int op = orient3d(a, b, c, p), oq = orient3d(a, b, c, q);
  if (op != 0 && op == oq) return false;
  if (op != 0 && oq != 0) {
    int s1 = orient3d(p, q, a, b), s2 = orient3d(p, q, b, c), s3 = orient3d(p, q, c, a);
    bool pos = s1 > 0 || s2 > 0 || s3 > 0, neg = s1 < 0 || s2 < 0 || s3 < 0;
    return !(pos && neg);
  }
  if (op == 0 && oq != 0) {
    int k = proj_axis(a, b, c);
    return k >= 0 && tri_contains_closed(a, b, c, p, k);
  }
  if (oq == 0 && op != 0) {
    int k = proj_axis(a, b, c);
    return k >= 0 && tri_contains_closed(a, b, c, q, k);
  }
  int k = proj_axis(a, b, c);
  if (k < 0) return false;
  return tri_contains_closed(a, b, c, p, k) || tri_contains_closed(a, b, c, q, k) ||
         seg_seg_closed(p, q, a, b, k) || seg_seg_closed(p, q, b, c, k) ||
         seg_seg_closed(p, q, c, a, k);
}

// Symbolically perturbed orient2d in the (y, z) plane of (a, b, m + (0,e,e^2)),
// m = (p+q)/2.
int sos_yz(const P3& a, const P3& b, const P3& p, const P3& q) {
  int s = orient2d_mid(a, b, p, q, 0);
  if (s != 0) return s;
  if (a.z != b.z) return a.z > b.z ? 1 : -1;
  if (a.y != b.y) return b.y > a.y ? 1 : -1;
  return 0;
}

// Does the ray m + t*(1,0,0), t > 0, m = (p+q)/2, cross triangle abc?
bool ray_cross(const P3& p, const P3& q, const P3& a, const P3& b, const P3& c) {
  int s1 = sos_yz(a, b, p, q);
  if (s1 == 0) return false;
  int s2 = sos_yz(b, c, p, q);
  if (s2 != s1) return false;
  int s3 = sos_yz(c, a, p, q);
  if (s3 != s1) return false;
  return orient3d_mid(a, b, c, p, q) == s1;
}

// ------------------------------------------------------------------ CPU BVH
struct Bvh {
  struct Node {
    double lo[3], hi[3];
    int left, right;  // children when count == 0
    int start, count;  // leaf range into order[]
  };
  std::vector<Node> nodes;
  std::vector<int> order;
  std::vector<P3> a, b, c;  // per triangle (original order)
  double eps = 0.0;

  void build(const float* V, const uint32_t* T, int64_t nT) {
    a.resize(nT), b.resize(nT), c.resize(nT);
    order.resize(nT);
    double glo[3] = {1e300, 1e300, 1e300}, ghi[3] = {-1e300, -1e300, -1e300};
    for (int64_t t = 0; t < nT; ++t) {
      a[t] = ld(V + 3 * (size_t)T[3 * t]);
      b[t] = ld(V + 3 * (size_t)T[3 * t + 1]);
      c[t] = ld(V + 3 * (size_t)T[3 * t + 2]);
      order[t] = (int)t;
      for (int k = 0; k < 3; ++k) {
        glo[k] = std::min({glo[k], co(a[t], k), co(b[t], k), co(c[t], k)});
        ghi[k] = std::max({ghi[k], co(a[t], k), co(b[t], k), co(c[t], k)});
      }
    }
    double ext = 0;
    for (int k = 0; k < 3; ++k) ext = std::max(ext, std::fabs(glo[k]) + std::fabs(ghi[k]));
    eps = 1e-9 * (ext + 1.0);
    nodes.clear();
    if (nT > 0) rec(0, (int)nT);
  }
  int rec(int s, int e) {
    Node nd;
    for (int k = 0; k < 3; ++k) nd.lo[k] = 1e300, nd.hi[k] = -1e300;
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int i = s; i < e; ++i) {
      int t = order[i];
      for (int k = 0; k < 3; ++k) {
        double x0 = co(a[t], k), x1 = co(b[t], k), x2 = co(c[t], k);
        nd.lo[k] = std::min({nd.lo[k], x0, x1, x2});
        nd.hi[k] = std::max({nd.hi[k], x0, x1, x2});
        double cc = (x0 + x1 + x2) / 3.0;
        clo[k] = std::min(clo[k], cc);
        chi[k] = std::max(chi[k], cc);
      }
    }
    for (int k = 0; k < 3; ++k) nd.lo[k] -= eps, nd.hi[k] += eps;
    int id = (int)nodes.size();
    nodes.push_back(nd);
    if (e - s <= 4) {
      nodes[id].left = nodes[id].right = -1;
      nodes[id].start = s;
      nodes[id].count = e - s;
      return id;
    }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
    int mid = (s + e) / 2;
    std::nth_element(order.begin() + s, order.begin() + mid, order.begin() + e, [&](int x, int y) {
      double cx = co(a[x], ax) + co(b[x], ax) + co(c[x], ax);
      double cy = co(a[y], ax) + co(b[y], ax) + co(c[y], ax);
      return cx < cy || (cx == cy && x < y);
    });
    int l = rec(s, mid);
    int r = rec(mid, e);
    nodes[id].left = l;
    nodes[id].right = r;
    nodes[id].start = 0;
    nodes[id].count = 0;
    return id;
  }
  static bool seg_box(const P3& p, const P3& q, const Node& n) {
    double t0 = 0.0, t1 = 1.0;
    double pp[3] = {p.x, p.y, p.z}, dd[3] = {q.x - p.x, q.y - p.y, q.z - p.z};
    for (int k = 0; k < 3; ++k) {
      if (dd[k] == 0.0) {
        if (pp[k] < n.lo[k] || pp[k] > n.hi[k]) return false;
        continue;
      }
      double ta = (n.lo[k] - pp[k]) / dd[k], tb = (n.hi[k] - pp[k]) / dd[k];
      if (ta > tb) std::swap(ta, tb);
      t0 = std::max(t0, ta - 1e-12);
      t1 = std::min(t1, tb + 1e-12);
      if (t0 > t1) return false;
    }
    return true;
  }
  static bool ray_box(const P3& p, const P3& q, const Node& n) {
    double my = 0.5 * (p.y + q.y), mz = 0.5 * (p.z + q.z), mx = 0.5 * (p.x + q.x);
    double tol = 1e-12 * (1.0 + std::fabs(mx) + std::fabs(my) + std::fabs(mz));
    return my >= n.lo[1] - tol && my <= n.hi[1] + tol && mz >= n.lo[2] - tol &&
           mz <= n.hi[2] + tol && n.hi[0] + tol >= mx;
  }
  template <class F>
  bool any_seg(const P3& p, const P3& q, F&& f) const {
    if (nodes.empty()) return false;
    int st[256];
    int sp = 0;
    st[sp++] = 0;
    while (sp) {
      const Node& n = nodes[st[--sp]];
      if (!seg_box(p, q, n)) continue;
      if (n.count) {
        for (int i = n.start; i < n.start + n.count; ++i)
          if (f(order[i])) return true;
      } else {
        st[sp++] = n.left;
        st[sp++] = n.right;
      }
    }
    return false;
  }
  template <class F>
  void each_ray(const P3& p, const P3& q, F&& f) const {
    if (nodes.empty()) return;
    int st[256];
    int sp = 0;
    st[sp++] = 0;
    while (sp) {
      const Node& n = nodes[st[--sp]];
      if (!ray_box(p, q, n)) continue;
      if (n.count) {
        for (int i = n.start; i < n.start + n.count; ++i) f(order[i]);
      } else {
        st[sp++] = n.left;
        st[sp++] = n.right;
      }
    }
  }
};

struct Mesh {
  const float* V;
  const uint32_t* T;
  int64_t nT;
};

// k -> (i, j), i < j, row-major upper triangle without the diagonal
// (lib/ptx/ray_segments.cu:39-46), computed with exact integer fix-up.
inline int64_t row_start(int64_t i, int64_t n) { return i * (2 * n - i - 1) / 2; }
void k_to_ij(int64_t k, int64_t n, int64_t& i, int64_t& j) {
  double nd = (double)n, kd = (double)k;
  double t = std::sqrt(std::max(0.0, -8.0 * kd + 4.0 * nd * (nd - 1.0) - 7.0));
  int64_t ii = (int64_t)(nd - 2.0 - std::floor(t / 2.0 - 0.5));
  if (ii < 0) ii = 0;
  if (ii > n - 2) ii = n - 2;
  while (ii > 0 && row_start(ii, n) > k) --ii;
  while (ii < n - 2 && row_start(ii + 1, n) <= k) ++ii;
  i = ii;
  j = k - row_start(ii, n) + ii + 1;
}

enum { ST_NOEDGE = 0, ST_EDGE = 1, ST_SKIP = 2 };
// Per-pair outcome detail: 0 none, 1 edge, 2 skipped by mask, 3 zero length,
// 4 blocked by self.
constexpr int kFanMax = 16;  // incident triangles above which an end is not tested

struct Ctx {
  const float* pts;
  int64_t n;
  const uint8_t* mask;
  Mesh self, cage;
  bool has_cage;
  bool brute;
  Bvh bself, bcage;
  // cap rule (mask bits 1-7): triangle groups, the axis each group's 2-D
  // projection drops, whether the self mesh is closed, and each point's fan
  const int8_t* tri_group = nullptr;
  int axis[7] = {3, 3, 3, 3, 3, 3, 3};
  bool closed = false;
  std::vector<std::vector<int>> fan;
};

// Mask of a point: bit 0 new (ray_segments.cu:61-75: a pair is skipped when
// neither point is new); bit 1 + g: the point lies on the plane of triangle
// group g (no counterpart in the reference; see enters_cap).
bool mask_skips(uint8_t a, uint8_t b) { return !((a | b) & 1); }

// The cap rule: for a pair whose points both lie on the plane of group g
// (shared mask bit 1 + g), in parity mode on a closed self mesh, the pair is
// not an edge when its direction at either end enters the closed 2-D angle,
// in the projection that drops the axis of the plane normal's largest
// component, of a group-g triangle incident to that end (vertex identity:
// equal coordinates; an end with more than kFanMax incident triangles is not
// tested).  With the group's triangles exactly in their plane this is the
// contact of the segment with them; with rounded (cut) vertices it decides
// the in-plane pairs that rounding alone would decide.
bool enters_cap(const Ctx& C, int64_t i, const P3& p, const P3& q, unsigned groups) {
  const std::vector<int>& fan = C.fan[i];
  if ((int)fan.size() > kFanMax) return false;
  for (int t : fan) {
    const int g = C.tri_group[t];
    if (g < 0 || g >= 7 || !((groups >> g) & 1u) || C.axis[g] == 3) continue;
    const int k = C.axis[g];
    const P3 &a = C.bself.a[t], &b = C.bself.b[t], &c = C.bself.c[t];
    const bool pa = same(p, a), pb = same(p, b);
    const P3& v1 = pa ? b : (pb ? c : a);
    const P3& v2 = pa ? c : (pb ? a : b);
    const int o = orient2d(p, v1, v2, k);
    if (o == 0) continue;
    const int o1 = orient2d(p, v1, q, k), o2 = orient2d(p, q, v2, k);
    if ((o1 != 0 || o2 != 0) && o * o1 >= 0 && o * o2 >= 0) return true;
  }
  return false;
}

int pair_detail(const Ctx& C, int64_t i, int64_t j) {
  if (C.mask && mask_skips(C.mask[i], C.mask[j])) return 2;
  P3 p = ld(C.pts + 3 * i), q = ld(C.pts + 3 * j);
  if (same(p, q)) return 3;
  if (C.mask && C.tri_group && C.closed && !C.has_cage) {
    const unsigned shared = (unsigned)(C.mask[i] & C.mask[j]) >> 1;
    if (shared && (enters_cap(C, i, p, q, shared) || enters_cap(C, j, q, p, shared))) return 4;
  }
  auto self_hit = [&](int t) {
    return contact(p, q, C.bself.a[t], C.bself.b[t], C.bself.c[t]);
  };
  bool blocked;
  if (C.brute) {
    blocked = false;
    for (int64_t t = 0; t < C.self.nT && !blocked; ++t) blocked = self_hit((int)t);
  } else {
    blocked = C.bself.any_seg(p, q, self_hit);
  }
  if (blocked) return 4;
  if (C.has_cage) {
    auto cage_hit = [&](int t) {
      return contact(p, q, C.bcage.a[t], C.bcage.b[t], C.bcage.c[t]);
    };
    bool hit;
    if (C.brute) {
      hit = false;
      for (int64_t t = 0; t < C.cage.nT && !hit; ++t) hit = cage_hit((int)t);
    } else {
      hit = C.bcage.any_seg(p, q, cage_hit);
    }
    return hit ? 1 : 0;
  }
  int cnt = 0;
  auto cross = [&](int t) {
    if (ray_cross(p, q, C.bself.a[t], C.bself.b[t], C.bself.c[t])) ++cnt;
  };
  if (C.brute) {
    for (int64_t t = 0; t < C.self.nT; ++t) cross((int)t);
  } else {
    C.bself.each_ray(p, q, cross);
  }
  return (cnt & 1) ? 0 : 1;
}

// Axis dropped by the 2-D projection of a plane with normal (a, b, c): the
// largest |component|, ties to the lower axis.
int plane_axis(const double* p) {
  const double a = std::fabs(p[0]), b = std::fabs(p[1]), c = std::fabs(p[2]);
  int k = b > a ? 1 : 0;
  if (c > std::max(a, b)) k = 2;
  return k;
}

void setup(Ctx& C, const float* pts, int64_t n, const uint8_t* mask, const float* V,
           const uint32_t* T, int64_t nT, const float* cV, const uint32_t* cT, int64_t cnT,
           int brute, const int8_t* tri_group = nullptr, const double* group_planes = nullptr,
           int n_groups = 0) {
  C.pts = pts;
  C.n = n;
  C.mask = mask;
  C.self = Mesh{V, T, nT};
  C.has_cage = cV != nullptr && cT != nullptr;
  C.cage = Mesh{cV, cT, cnT};
  C.brute = brute != 0;
  C.bself.build(V, T, nT);
  if (C.has_cage) C.bcage.build(cV, cT, cnT);
  if (!tri_group || n_groups <= 0 || !mask) return;
  C.tri_group = tri_group;
  for (int g = 0; g < n_groups && g < 7; ++g) C.axis[g] = plane_axis(group_planes + 4 * g);
  // closed: every undirected edge (by vertex index) has an even number of triangles
  std::vector<std::pair<uint32_t, uint32_t>> es;
  es.reserve(3 * (size_t)nT);
  for (int64_t t = 0; t < nT; ++t)
    for (int e = 0; e < 3; ++e) {
      uint32_t u = T[3 * t + e], v = T[3 * t + (e + 1) % 3];
      es.push_back({std::min(u, v), std::max(u, v)});
    }
  std::sort(es.begin(), es.end());
  C.closed = true;
  for (size_t a = 0; a < es.size();) {
    size_t b = a;
    while (b < es.size() && es[b] == es[a]) ++b;
    if ((b - a) & 1) C.closed = false;
    a = b;
  }
  // each point's fan: the triangles with a vertex of equal coordinates
  std::vector<std::pair<std::array<uint32_t, 3>, int>> keyed;
  keyed.reserve(3 * (size_t)nT);
  auto key = [](const float* v) {
    std::array<uint32_t, 3> k;
    std::memcpy(k.data(), v, 12);
    for (auto& x : k) x = (x == 0x80000000u) ? 0u : x;  // -0 equals +0
    return k;
  };
  for (int64_t t = 0; t < nT; ++t)
    for (int c = 0; c < 3; ++c) keyed.push_back({key(V + 3 * (size_t)T[3 * t + c]), (int)t});
  std::sort(keyed.begin(), keyed.end());
  C.fan.assign((size_t)n, {});
  for (int64_t i = 0; i < n; ++i) {
    const auto k = key(pts + 3 * i);
    auto it = std::lower_bound(keyed.begin(), keyed.end(), std::make_pair(k, -1));
    for (; it != keyed.end() && it->first == k; ++it)
      if (C.fan[i].empty() || C.fan[i].back() != it->second) C.fan[i].push_back(it->second);
  }
}


// ------------------------------------------- the reference's ray window
// lib/ptx/ray_segments.cu:77-111 restated in float32 for the mismatch
// count between that rule and the exact contact definition above: the
// ray starts at p with direction (q - p) * (1 / len), len = sqrtf of the
// squared length (contracted into fmaf as nvcc does by default), and a
// triangle is hit when the ray meets it at t in the closed window
// [0.001, len].  OptiX's own triangle test and rounding are closed; the
// test here is the published watertight one (S. Woop, C. Benthin, I. Wald,
// "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013), float32, no
// culling, its double-precision fallback for zero edge functions.
struct RayF {
  float o[3], d[3];
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

RayF make_ray(const float* p, const float* dir) {
  RayF r;
  for (int k = 0; k < 3; ++k) r.o[k] = p[k], r.d[k] = dir[k];
  int kz = 0;
  for (int k = 1; k < 3; ++k)
    if (std::fabs(dir[k]) > std::fabs(dir[kz])) kz = k;
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (dir[kz] < 0.0f) std::swap(kx, ky);
  r.kx = kx, r.ky = ky, r.kz = kz;
  r.Sx = dir[kx] / dir[kz];
  r.Sy = dir[ky] / dir[kz];
  r.Sz = 1.0f / dir[kz];
  return r;
}

// true (and t) when the ray meets triangle abc.
bool watertight_hit(const RayF& r, const float* a, const float* b, const float* c, float& t) {
  const float A[3] = {a[0] - r.o[0], a[1] - r.o[1], a[2] - r.o[2]};
  const float B[3] = {b[0] - r.o[0], b[1] - r.o[1], b[2] - r.o[2]};
  const float C[3] = {c[0] - r.o[0], c[1] - r.o[1], c[2] - r.o[2]};
  const float Ax = A[r.kx] - r.Sx * A[r.kz], Ay = A[r.ky] - r.Sy * A[r.kz];
  const float Bx = B[r.kx] - r.Sx * B[r.kz], By = B[r.ky] - r.Sy * B[r.kz];
  const float Cx = C[r.kx] - r.Sx * C[r.kz], Cy = C[r.ky] - r.Sy * C[r.kz];
  float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float Az = r.Sz * A[r.kz], Bz = r.Sz * B[r.kz], Cz = r.Sz * C[r.kz];
  const float T = U * Az + V * Bz + W * Cz;
  t = T * (1.0f / det);
  return true;
}

// float64 distances for the grazing class
double dot3(const P3& a, const P3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
P3 sub3(const P3& a, const P3& b) { return P3{a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 axpy3(double s, const P3& d, const P3& o) { return P3{o.x + s * d.x, o.y + s * d.y, o.z + s * d.z}; }
double dist3(const P3& a, const P3& b) { return std::sqrt(dot3(sub3(a, b), sub3(a, b))); }

// Distance between segments p1q1 and p2q2 (closest points by clamped
// parameters, C. Ericson, Real-Time Collision Detection, 5.1.9).
double seg_seg_dist(const P3& p1, const P3& q1, const P3& p2, const P3& q2) {
  const P3 d1 = sub3(q1, p1), d2 = sub3(q2, p2), r = sub3(p1, p2);
  const double a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
  double s = 0.0, t = 0.0;
  if (a <= 1e-300 && e <= 1e-300) return dist3(p1, p2);
  if (a <= 1e-300) {
    t = std::clamp(f / e, 0.0, 1.0);
  } else {
    const double c = dot3(d1, r);
    if (e <= 1e-300) {
      s = std::clamp(-c / a, 0.0, 1.0);
    } else {
      const double b = dot3(d1, d2), den = a * e - b * b;
      s = den > 0.0 ? std::clamp((b * f - c * e) / den, 0.0, 1.0) : 0.0;
      t = (b * s + f) / e;
      if (t < 0.0) t = 0.0, s = std::clamp(-c / a, 0.0, 1.0);
      else if (t > 1.0) t = 1.0, s = std::clamp((b - c) / a, 0.0, 1.0);
    }
  }
  return dist3(axpy3(s, d1, p1), axpy3(t, d2, p2));
}

// Distance from x to the closed triangle abc (Ericson 5.1.5).
double point_tri_dist(const P3& x, const P3& a, const P3& b, const P3& c) {
  const P3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(x, a);
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return dist3(x, a);
  const P3 bp = sub3(x, b);
  const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return dist3(x, b);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return dist3(x, axpy3(d1 / (d1 - d3), ab, a));
  const P3 cp = sub3(x, c);
  const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return dist3(x, c);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return dist3(x, axpy3(d2 / (d2 - d6), ac, a));
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
    return dist3(x, axpy3((d4 - d3) / ((d4 - d3) + (d5 - d6)), sub3(c, b), b));
  const double den = 1.0 / (va + vb + vc);
  const double v = vb * den, w = vc * den;
  return dist3(x, P3{a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w,
                     a.z + ab.z * v + ac.z * w});
}

// Within tau of grazing: the segment passes within tau of the triangle's
// boundary, or an endpoint lies within tau of the triangle.
bool grazing(const P3& p, const P3& q, const P3& a, const P3& b, const P3& c, double tau) {
  return seg_seg_dist(p, q, a, b) <= tau || seg_seg_dist(p, q, b, c) <= tau ||
         seg_seg_dist(p, q, c, a) <= tau || point_tri_dist(p, a, b, c) <= tau ||
         point_tri_dist(q, a, b, c) <= tau;
}

// Per pair (i, j), i < j, of points on or off the mesh: bit 0 exact
// contact with some triangle; bit 1 a window hit, t in [0.001, len]; bit 2
// a hit with t in [0, len]; bit 3 a hit with t in [0.001, len - tau];
// bit 4 every triangle on which contact and the window rule disagree is
// within tau of grazing.
uint8_t window_bits(const Bvh& B, const float* V, const uint32_t* T, const float* pf,
                    const float* qf, double tau, std::vector<int>& cand) {
  const P3 p = ld(pf), q = ld(qf);
  const float dx = qf[0] - pf[0], dy = qf[1] - pf[1], dz = qf[2] - pf[2];
  const float len = std::sqrt(std::fma(dz, dz, std::fma(dy, dy, dx * dx)));
  uint8_t bits = 0;
  if (len == 0.0f) return 0;
  const float inv = 1.0f / len;
  const float dir[3] = {dx * inv, dy * inv, dz * inv};
  const RayF r = make_ray(pf, dir);
  // candidates: triangles whose box, widened by tau, meets the segment
  cand.clear();
  if (!B.nodes.empty()) {
    int st[256];
    int sp = 0;
    st[sp++] = 0;
    while (sp) {
      Bvh::Node n = B.nodes[st[--sp]];
      for (int k = 0; k < 3; ++k) n.lo[k] -= tau, n.hi[k] += tau;
      if (!Bvh::seg_box(p, q, n)) continue;
      if (n.count) {
        for (int i = n.start; i < n.start + n.count; ++i) cand.push_back(B.order[i]);
      } else {
        st[sp++] = n.left;
        st[sp++] = n.right;
      }
    }
  }
  bool all_graze = true;
  const float tmin = 0.001f, tcut = (float)((double)len - tau);
  for (int t : cand) {
    const float* a = V + 3 * (size_t)T[3 * t];
    const float* b = V + 3 * (size_t)T[3 * t + 1];
    const float* c = V + 3 * (size_t)T[3 * t + 2];
    const bool ex = contact(p, q, ld(a), ld(b), ld(c));
    float th = 0.0f;
    const bool h = watertight_hit(r, a, b, c, th);
    const bool win = h && th >= tmin && th <= len;
    if (ex) bits |= 1;
    if (win) bits |= 2;
    if (h && th >= 0.0f && th <= len) bits |= 4;
    if (h && th >= tmin && th <= tcut) bits |= 8;
    if (ex != win && all_graze && !grazing(p, q, ld(a), ld(b), ld(c), tau)) all_graze = false;
  }
  if (all_graze) bits |= 16;
  return bits;
}
}  // namespace

extern "C" {

int orc_orient3d(const float* a, const float* b, const float* c, const float* d) {
  return orient3d(ld(a), ld(b), ld(c), ld(d));
}
int orc_orient3d_mid(const float* a, const float* b, const float* c, const float* p,
                     const float* q) {
  return orient3d_mid(ld(a), ld(b), ld(c), ld(p), ld(q));
}
int orc_orient2d(const float* a, const float* b, const float* c, int k) {
  return orient2d(ld(a), ld(b), ld(c), k);
}
int orc_orient2d_mid(const float* a, const float* b, const float* p, const float* q, int k) {
  return orient2d_mid(ld(a), ld(b), ld(p), ld(q), k);
}
int orc_contact(const float* p, const float* q, const float* a, const float* b, const float* c) {
  return contact(ld(p), ld(q), ld(a), ld(b), ld(c)) ? 1 : 0;
}
int orc_ray_cross(const float* p, const float* q, const float* a, const float* b,
                  const float* c) {
  return ray_cross(ld(p), ld(q), ld(a), ld(b), ld(c)) ? 1 : 0;
}
void orc_k_to_ij(int64_t k, int64_t n, int64_t* i, int64_t* j) { k_to_ij(k, n, *i, *j); }

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// All pairs in k order.  Writes visibility edges (i, j) to pairs_out (up to
// cap pairs) and returns the edge count; stats[0..4] receive the counts of
// details 0..4 (none, edge, masked, zero length, blocked).  Returns -1 if
// the edge count exceeds cap.
// tri_group (nullable): per triangle its group (-1 none, 0..6) for the cap
// rule of mask bits 1-7, with group_planes [n_groups][4] (a, b, c, d).
int64_t orc_visibility(const float* pts, int64_t n, const uint8_t* mask, const float* V,
                       int64_t nV, const uint32_t* T, int64_t nT, const float* cV,
                       int64_t cnV, const uint32_t* cT, int64_t cnT, int brute,
                       uint32_t* pairs_out, int64_t cap, int64_t* stats,
                       const int8_t* tri_group, const double* group_planes, int n_groups) {
  (void)nV;
  (void)cnV;
  Ctx C;
  setup(C, pts, n, mask, V, T, nT, cV, cT, cnT, brute, tri_group, group_planes, n_groups);
  std::vector<std::vector<uint32_t>> rows(n > 0 ? n : 0);
  int64_t st[5] = {0, 0, 0, 0, 0};
#pragma omp parallel
  {
    int64_t lst[5] = {0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
      for (int64_t j = i + 1; j < n; ++j) {
        int d = pair_detail(C, i, j);
        lst[d]++;
        if (d == 1) rows[i].push_back((uint32_t)j);
      }
    }
#pragma omp critical
    for (int k = 0; k < 5; ++k) st[k] += lst[k];
  }
  if (stats)
    for (int k = 0; k < 5; ++k) stats[k] = st[k];
  int64_t e = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (uint32_t j : rows[i]) {
      if (e < cap) {
        pairs_out[2 * e] = (uint32_t)i;
        pairs_out[2 * e + 1] = j;
      }
      ++e;
    }
  }
  return e <= cap ? e : -1;
}

// Per-pair details for an explicit list of pair indices k (any order).
void orc_visibility_sample(const float* pts, int64_t n, const uint8_t* mask, const float* V,
                           int64_t nV, const uint32_t* T, int64_t nT, const float* cV,
                           int64_t cnV, const uint32_t* cT, int64_t cnT, int brute,
                           const int64_t* ks, int64_t m, uint8_t* detail_out,
                           const int8_t* tri_group, const double* group_planes, int n_groups) {
  (void)nV;
  (void)cnV;
  Ctx C;
  setup(C, pts, n, mask, V, T, nT, cV, cT, cnT, brute, tri_group, group_planes, n_groups);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t s = 0; s < m; ++s) {
    int64_t i, j;
    k_to_ij(ks[s], n, i, j);
    detail_out[s] = (uint8_t)pair_detail(C, i, j);
  }
}

// FP64 plane scores: side by a*x+b*y+c*z+d in float64 with dead zone 1e-6,
// score = sum of float64 edge lengths over strictly crossing edges.
void orc_score_fp64(const float* planes, int64_t P, const float* pts, int64_t n,
                    const uint32_t* edges, int64_t E, double* scores) {
  // edges with an endpoint index >= n contribute nothing
  // (lib/cuda/plane_intersections.cu:34-36)
  std::vector<double> len(E);
  for (int64_t e = 0; e < E; ++e) {
    if ((int64_t)edges[2 * e] >= n || (int64_t)edges[2 * e + 1] >= n) {
      len[e] = -1.0;
      continue;
    }
    P3 u = ld(pts + 3 * (size_t)edges[2 * e]), v = ld(pts + 3 * (size_t)edges[2 * e + 1]);
    double dx = v.x - u.x, dy = v.y - u.y, dz = v.z - u.z;
    len[e] = std::sqrt(dx * dx + dy * dy + dz * dz);
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t p = 0; p < P; ++p) {
    double a = planes[4 * p], b = planes[4 * p + 1], c = planes[4 * p + 2], d = planes[4 * p + 3];
    long double acc = 0.0L;
    for (int64_t e = 0; e < E; ++e) {
      if (len[e] < 0.0) continue;
      P3 u = ld(pts + 3 * (size_t)edges[2 * e]), v = ld(pts + 3 * (size_t)edges[2 * e + 1]);
      double vu = a * u.x + b * u.y + c * u.z + d;
      double vv = a * v.x + b * v.y + c * v.z + d;
      int su = (vu > 1e-6) - (vu < -1e-6), sv = (vv > 1e-6) - (vv < -1e-6);
      if (su * sv < 0) acc += len[e];
    }
    scores[p] = (double)acc;
  }
}

// Float32 side test and Q32.32 fixed-point score (the kernels' definition):
//   val = fmaf(c, z, fmaf(b, y, a*x)) + d          (float32)
//   side = (val > 1e-6f) - (val < -1e-6f); crossing iff side_u * side_v < 0
//   len  = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx*dx))) (d = second - first)
//   q    = round-to-nearest-even(len * 2^32) as int64; score = sum of q.
void orc_score_q32(const float* planes, int64_t P, const float* pts, int64_t n,
                   const uint32_t* edges, int64_t E, int64_t* scores) {
  std::vector<int64_t> lq(E);
  std::vector<char> ok(E);
  for (int64_t e = 0; e < E; ++e) {
    ok[e] = (int64_t)edges[2 * e] < n && (int64_t)edges[2 * e + 1] < n;
    if (!ok[e]) {
      lq[e] = 0;
      continue;
    }
    const float* u = pts + 3 * (size_t)edges[2 * e];
    const float* v = pts + 3 * (size_t)edges[2 * e + 1];
    float dx = v[0] - u[0], dy = v[1] - u[1], dz = v[2] - u[2];
    float l = std::sqrt(std::fma(dz, dz, std::fma(dy, dy, dx * dx)));
    lq[e] = (int64_t)std::nearbyint((double)l * 4294967296.0);
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t p = 0; p < P; ++p) {
    float a = planes[4 * p], b = planes[4 * p + 1], c = planes[4 * p + 2], d = planes[4 * p + 3];
    int64_t acc = 0;
    for (int64_t e = 0; e < E; ++e) {
      if (!ok[e]) continue;
      const float* u = pts + 3 * (size_t)edges[2 * e];
      const float* v = pts + 3 * (size_t)edges[2 * e + 1];
      float vu = std::fmaf(c, u[2], std::fmaf(b, u[1], a * u[0])) + d;
      float vv = std::fmaf(c, v[2], std::fmaf(b, v[1], a * v[0])) + d;
      int su = (vu > 1e-6f) - (vu < -1e-6f), sv = (vv > 1e-6f) - (vv < -1e-6f);
      if (su * sv < 0) acc += lq[e];
    }
    scores[p] = acc;
  }
}

int64_t orc_argmax_i64(const int64_t* s, int64_t P) {
  int64_t best = 0;
  for (int64_t p = 1; p < P; ++p)
    if (s[p] > s[best]) best = p;
  return P > 0 ? best : -1;
}

// Batch form of the exact predicates (same kinds and input layout as the
// engine's host predicate check): kind 0 orient3d, 1 orient3d with the
// midpoint of two points as the last, 2 orient2d in projection k, 3
// orient2d with a midpoint.
void orc_predicates_batch(int kind, int64_t n, const float* in, int k, int* out) {
  const int stride = kind == 0 ? 12 : (kind == 1 ? 15 : (kind == 2 ? 9 : 12));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const float* x = in + stride * i;
    if (kind == 0) out[i] = orient3d(ld(x), ld(x + 3), ld(x + 6), ld(x + 9));
    else if (kind == 1) out[i] = orient3d_mid(ld(x), ld(x + 3), ld(x + 6), ld(x + 9), ld(x + 12));
    else if (kind == 2) out[i] = orient2d(ld(x), ld(x + 3), ld(x + 6), k);
    else out[i] = orient2d_mid(ld(x), ld(x + 3), ld(x + 6), ld(x + 9), k);
  }
}

// Raw outputs of std::mt19937(seed) (the reference's random_engine,
// lib/include/core.hpp:13) and draws of std::uniform_int_distribution<int>
// (lo, hi) on it (lib/src/process.cpp:38-43), for the known-answer tests.
void orc_mt19937(uint32_t seed, int64_t n, uint32_t* out) {
  std::mt19937 rng(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = (uint32_t)rng();
}
void orc_uniform_int(uint32_t seed, int lo, int hi, int64_t n, int32_t* out) {
  std::mt19937 rng(seed);
  std::uniform_int_distribution<> dis(lo, hi);
  for (int64_t i = 0; i < n; ++i) out[i] = dis(rng);
}

// Bisector candidate planes (lib/src/process.cpp:28-72) with a fresh
// std::mt19937(seed).  Vertices are float64.  Returns the plane count.
int64_t orc_bisector_planes(const double* V, const uint32_t* edges, int64_t E, int num_planes,
                            uint32_t seed, double* planes_out) {
  if (E == 0) return 0;
  std::mt19937 rng(seed);
  std::uniform_int_distribution<> dis(0, (int)E - 1);
  const double normal_eps = std::cos(5.0 * 3.14159265358979323846 / 180.0);
  const double dist_eps = 1e-3;
  int64_t np = 0;
  for (int i = 0; i < num_planes * 5 && np < num_planes; ++i) {
    int idx = dis(rng);
    const double* p1 = V + 3 * (size_t)edges[2 * idx];
    const double* p2 = V + 3 * (size_t)edges[2 * idx + 1];
    double n[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l < 1e-6) continue;
    for (double& x : n) x /= l;
    double m[3] = {(p1[0] + p2[0]) * 0.5, (p1[1] + p2[1]) * 0.5, (p1[2] + p2[2]) * 0.5};
    double d = -(n[0] * m[0] + n[1] * m[1] + n[2] * m[2]);
    bool similar = false;
    for (int64_t q = 0; q < np; ++q) {
      const double* pl = planes_out + 4 * q;
      double dot = std::fabs(n[0] * pl[0] + n[1] * pl[1] + n[2] * pl[2]);
      if (dot > normal_eps && std::fabs(d - pl[3]) < dist_eps) {
        similar = true;
        break;
      }
    }
    if (!similar) {
      double* pl = planes_out + 4 * np;
      pl[0] = n[0], pl[1] = n[1], pl[2] = n[2], pl[3] = d;
      ++np;
    }
  }
  return np;
}

// The reference's ray-window rule against the exact contact definition,
// for the pairs k of `ks` (k order indices, or all pairs when ks is NULL):
// out[t] = window_bits of pair ks[t] (bits in the comment above it).
void orc_window_hits(const float* pts, int64_t n, const float* V, const uint32_t* T,
                     int64_t nT, const int64_t* ks, int64_t m, double tau, uint8_t* out) {
  Bvh B;
  B.build(V, T, nT);
  const int64_t total = ks ? m : n * (n - 1) / 2;
#pragma omp parallel
  {
    std::vector<int> cand;
#pragma omp for schedule(dynamic, 256)
    for (int64_t t = 0; t < total; ++t) {
      int64_t i, j;
      k_to_ij(ks ? ks[t] : t, n, i, j);
      out[t] = window_bits(B, V, T, pts + 3 * i, pts + 3 * j, tau, cand);
    }
  }
}

}  // extern "C"
