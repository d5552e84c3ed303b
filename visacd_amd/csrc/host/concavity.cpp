// Concavity of a part against its convex hull, and flat-surface planes.
//
// h = max(0.3 * rv, hb) (lib/src/cost.cpp:107-123):
//   rv = cbrt(3 |V(part) - V(hull)| / (4 pi))                (cost.cpp:59-67)
//   hb = symmetric sampled surface distance: each mesh is sampled with
//        max(1000, resolution * area) area-weighted points plus one point
//        in every triangle left unsampled (core.cpp:144-232); every sample
//        of one mesh is matched to its 10 nearest samples of the other and
//        measured exactly against their triangles; hb is the largest such
//        distance in either direction (hausdorff.hpp:214-298).
// Sampling uses a generator seeded per call, so a part's concavity does not
// depend on the order in which parts are evaluated.
#include <algorithm>
#include <cmath>
#include <queue>
#include <unordered_map>

#include "mesh.hpp"

namespace vhost {
namespace {

struct Samples {
  std::vector<Vec3> p;
  std::vector<int> tri;
};

Samples sample_surface(const Mesh &m, int resolution, std::mt19937 &rng) {
  Samples s;
  const size_t T = m.triangles.size();
  if (T == 0) return s;
  std::vector<double> area(T);
  double total = 0;
  for (size_t t = 0; t < T; ++t) {
    const auto &f = m.triangles[t];
    area[t] = tri_area(m.vertices[f[0]], m.vertices[f[1]], m.vertices[f[2]]);
    total += area[t];
  }
  const size_t want = (size_t)std::max(1000, (int)(resolution * total));
  std::discrete_distribution<size_t> pick(area.begin(), area.end());
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  std::vector<char> hit(T, 0);
  s.p.reserve(want + T);
  s.tri.reserve(want + T);
  auto emit = [&](size_t t) {
    const auto &f = m.triangles[t];
    const double a = std::sqrt(u01(rng)), b = u01(rng);
    const Vec3 &A = m.vertices[f[0]], &B = m.vertices[f[1]], &C = m.vertices[f[2]];
    s.p.push_back(A * (1 - a) + B * (a * (1 - b)) + C * (a * b));
    s.tri.push_back((int)t);
    hit[t] = 1;
  };
  if (total > 0)
    for (size_t k = 0; k < want; ++k) emit(pick(rng));
  for (size_t t = 0; t < T; ++t)
    if (!hit[t]) emit(t);
  return s;
}

// Static kd-tree over points with k-nearest queries.
struct KdTree {
  const std::vector<Vec3> *pts = nullptr;
  std::vector<int> idx;
  struct KNode {
    int lo, hi, axis, left, right;
    double split;
  };
  std::vector<KNode> nodes;
  void build(const std::vector<Vec3> &p) {
    pts = &p;
    idx.resize(p.size());
    for (size_t i = 0; i < p.size(); ++i) idx[i] = (int)i;
    nodes.clear();
    if (!p.empty()) rec(0, (int)p.size());
  }
  int rec(int lo, int hi) {
    KNode nd{lo, hi, -1, -1, -1, 0.0};
    const int id = (int)nodes.size();
    nodes.push_back(nd);
    if (hi - lo <= 8) return id;
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = lo; i < hi; ++i)
      for (int k = 0; k < 3; ++k) mn[k] = std::min(mn[k], (*pts)[idx[i]][k]), mx[k] = std::max(mx[k], (*pts)[idx[i]][k]);
    int ax = 0;
    for (int k = 1; k < 3; ++k)
      if (mx[k] - mn[k] > mx[ax] - mn[ax]) ax = k;
    const int mid = (lo + hi) / 2;
    std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi,
                     [&](int a, int b) { return (*pts)[a][ax] < (*pts)[b][ax]; });
    const double split = (*pts)[idx[mid]][ax];
    const int l = rec(lo, mid), r = rec(mid, hi);
    nodes[id].axis = ax;
    nodes[id].split = split;
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
  }
  // k nearest (squared distance, index), nearest first
  void knn(const Vec3 &q, int k, std::vector<std::pair<double, int>> &out) const {
    std::priority_queue<std::pair<double, int>> heap;  // max-heap
    search(0, q, k, heap);
    out.resize(heap.size());
    for (int i = (int)heap.size() - 1; i >= 0; --i) out[i] = heap.top(), heap.pop();
  }
  void search(int n, const Vec3 &q, int k, std::priority_queue<std::pair<double, int>> &heap) const {
    const KNode &nd = nodes[n];
    if (nd.axis < 0) {
      for (int i = nd.lo; i < nd.hi; ++i) {
        const Vec3 d = (*pts)[idx[i]] - q;
        const double d2 = dot(d, d);
        if ((int)heap.size() < k) heap.push({d2, idx[i]});
        else if (d2 < heap.top().first) heap.pop(), heap.push({d2, idx[i]});
      }
      return;
    }
    const double diff = q[nd.axis] - nd.split;
    const int first = diff < 0 ? nd.left : nd.right, second = diff < 0 ? nd.right : nd.left;
    search(first, q, k, heap);
    if ((int)heap.size() < k || diff * diff < heap.top().first) search(second, q, k, heap);
  }
};

// Euclidean distance from p to triangle abc (closest-point regions; C. Ericson,
// Real-Time Collision Detection, 2005, section 5.1.5).
double point_tri_dist(const Vec3 &p, const Vec3 &a, const Vec3 &b, const Vec3 &c) {
  const Vec3 ab = b - a, ac = c - a, ap = p - a;
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0 && d2 <= 0) return norm(p - a);
  const Vec3 bp = p - b;
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0 && d4 <= d3) return norm(p - b);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) return norm(p - (a + ab * (d1 / (d1 - d3))));
  const Vec3 cp = p - c;
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0 && d5 <= d6) return norm(p - c);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) return norm(p - (a + ac * (d2 / (d2 - d6))));
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0)
    return norm(p - (b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)))));
  const double den = 1.0 / (va + vb + vc);
  const double v = vb * den, w = vc * den;
  return norm(p - (a + ab * v + ac * w));
}

double one_side(const Mesh &ma, const Samples &sa, const KdTree &ta, const Samples &sb) {
  double cmax = 0.0;
#pragma omp parallel for reduction(max : cmax) schedule(dynamic, 256)
  for (long i = 0; i < (long)sb.p.size(); ++i) {
    std::vector<std::pair<double, int>> nn;
    ta.knn(sb.p[i], 10, nn);
    double cmin = 1e300;
    for (auto &r : nn) {
      const auto &f = ma.triangles[sa.tri[r.second]];
      const double d = point_tri_dist(sb.p[i], ma.vertices[f[0]], ma.vertices[f[1]], ma.vertices[f[2]]);
      if (d < cmin) {
        cmin = d;
        if (cmin < 1e-14) break;
      }
    }
    if (cmin > 10 && !nn.empty()) cmin = std::sqrt(nn[0].first);
    if (cmin < 1e300 && cmin > cmax) cmax = cmin;
  }
  return cmax;
}

}  // namespace

std::vector<Vec3> surface_samples(const Mesh &m, int resolution, std::mt19937 &rng) {
  return sample_surface(m, resolution, rng).p;
}

double hausdorff_boundary(const Mesh &a, const Mesh &b, int resolution, uint32_t seed) {
  std::mt19937 rng(seed), rng_b(second_sample_seed(seed));
  const Samples sa = sample_surface(a, resolution, rng);
  const Samples sb = sample_surface(b, resolution, rng_b);
  if (sa.p.empty() || sb.p.empty()) return 1e300;
  KdTree ta, tb;
  ta.build(sa.p);
  tb.build(sb.p);
  return std::max(one_side(a, sa, ta, sb), one_side(b, sb, tb, sa));
}

double concavity(const Mesh &part, const Mesh &hull, int resolution, uint32_t seed) {
  const double rv = std::cbrt(3.0 * std::fabs(mesh_volume(part) - mesh_volume(hull)) / (4.0 * M_PI));
  const double hb = hausdorff_boundary(part, hull, resolution, seed);
  return std::max(0.3 * rv, hb);
}

// Flat regions (lib/src/support_surface.cpp:14-182): grow over edge-adjacent
// triangles whose normal has cosine > 0.999 with the running mean normal;
// keep regions larger than min_area (the 40 largest); each gives the plane
// through its centroid pushed 0.02 along its mean normal; near-duplicate
// planes (cosine of the (a,b,c,d) vectors > 1 - 1e-4) are dropped.
std::vector<Plane> flat_surface_planes(const Mesh &m, double min_area) {
  const int T = (int)m.triangles.size();
  // unit normals, their lengths (as the growth test recomputes them) and
  // the triangle areas
  std::vector<Vec3> nrm(T);
  std::vector<double> nlen(T), tarea(T);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < T; ++i) {
    const auto &t = m.triangles[i];
    Vec3 n = cross(m.vertices[t[1]] - m.vertices[t[0]], m.vertices[t[2]] - m.vertices[t[0]]);
    const double l = norm(n);
    nrm[i] = l > 0 ? n * (1.0 / l) : Vec3{0, 0, 0};
    nlen[i] = norm(nrm[i]);
    tarea[i] = 0.5 * l;
  }
  // triangle adjacency across shared edges, from a vertex -> triangle index
  // (for each edge (a, b) of t: the other triangles of a that contain b)
  const int nv = (int)m.vertices.size();
  std::vector<int> inc_off(nv + 1, 0), inc;
  for (const auto &t : m.triangles)
    for (int k = 0; k < 3; ++k) ++inc_off[t[k] + 1];
  for (int v = 0; v < nv; ++v) inc_off[v + 1] += inc_off[v];
  inc.resize(inc_off[nv]);
  {
    std::vector<int> fill(inc_off.begin(), inc_off.end() - 1);
    for (int i = 0; i < T; ++i)
      for (int k = 0; k < 3; ++k) inc[fill[m.triangles[i][k]]++] = i;
  }
  // neighbours of every triangle, counted then written (both passes in
  // parallel, the same order as one sequential pass)
  std::vector<int> adj_off(T + 1, 0), adj;
  auto each_neighbour = [&](int i, auto &&emit) {
    const auto &f = m.triangles[i];
    for (int e = 0; e < 3; ++e) {
      const int a = f[e], b = f[(e + 1) % 3];
      if (a == b) continue;
      for (int x = inc_off[a]; x < inc_off[a + 1]; ++x) {
        const int u = inc[x];
        if (u == i) continue;
        const auto &g = m.triangles[u];
        if (g[0] == b || g[1] == b || g[2] == b) emit(u);
      }
    }
  };
#pragma omp parallel for schedule(static)
  for (int i = 0; i < T; ++i) {
    int c = 0;
    each_neighbour(i, [&](int) { ++c; });
    adj_off[i + 1] = c;
  }
  for (int i = 0; i < T; ++i) adj_off[i + 1] += adj_off[i];
  adj.resize(adj_off[T]);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < T; ++i) {
    int w = adj_off[i];
    each_neighbour(i, [&](int u) { adj[w++] = u; });
  }
  struct Region {
    std::vector<int> tris;
    double area = 0;
  };
  std::vector<Region> regions;
  std::vector<char> used(T, 0);
  std::vector<int> grown, queue;  // region being grown (reused)
  for (int s = 0; s < T; ++s) {
    if (used[s]) continue;
    grown.assign(1, s);
    used[s] = 1;
    Vec3 mean = nrm[s];
    double lm = norm(mean);  // |mean|, recomputed whenever mean changes
    queue.assign(1, s);
    for (size_t qh = 0; qh < queue.size(); ++qh) {
      const int t = queue[qh];
      for (int x = adj_off[t]; x < adj_off[t + 1]; ++x) {
        const int nb = adj[x];
        if (used[nb]) continue;
        const double ln = nlen[nb];
        const double c = (lm > 0 && ln > 0) ? dot(mean, nrm[nb]) / (lm * ln) : 0.0;
        if (c > 0.999) {
          used[nb] = 1;
          grown.push_back(nb);
          const double k = (double)grown.size();
          mean = (mean * (k - 1) + nrm[nb]) * (1.0 / k);
          lm = norm(mean);
          queue.push_back(nb);
        }
      }
    }
    double area = 0.0;
    for (int t : grown) area += tarea[t];
    if (area > min_area) regions.push_back(Region{grown, area});
  }
  if (regions.size() > 40) {
    std::stable_sort(regions.begin(), regions.end(),
                     [](const Region &a, const Region &b) { return a.area > b.area; });
    regions.resize(40);
  }
  std::vector<Plane> planes;
  for (auto &r : regions) {
    Vec3 n{0, 0, 0}, c{0, 0, 0};
    for (int t : r.tris) {
      const auto &f = m.triangles[t];
      n = n + nrm[t];
      c = c + (m.vertices[f[0]] + m.vertices[f[1]] + m.vertices[f[2]]) * (1.0 / 3.0);
    }
    n = n * (1.0 / norm(n));
    c = c * (1.0 / (double)r.tris.size()) + n * 2e-2;
    Plane p(n[0], n[1], n[2], -dot(n, c));
    bool dup = false;
    for (auto &q : planes) {
      const double d = p.a * q.a + p.b * q.b + p.c * q.c + p.d * q.d;
      const double m1 = std::sqrt(p.a * p.a + p.b * p.b + p.c * p.c + p.d * p.d);
      const double m2 = std::sqrt(q.a * q.a + q.b * q.b + q.c * q.c + q.d * q.d);
      if (m1 > 1e-9 && m2 > 1e-9 && std::fabs(d) / (m1 * m2) > 1.0 - 1e-4) {
        dup = true;
        break;
      }
    }
    if (!dup) planes.push_back(p);
  }
  return planes;
}

}  // namespace vhost
