// Normalisation, volumes, convex hull and connected components.
#include <algorithm>
#include <functional>
#include <limits>
#include <numeric>
#include <unordered_map>
#include <unordered_set>

#include "mesh.hpp"

namespace vhost {

std::mt19937 &random_engine() {
  static std::mt19937 eng(std::random_device{}());
  return eng;
}

void set_seed(unsigned int seed) { random_engine().seed(seed); }

// Bbox centre to the origin, longest side to 2 (lib/src/core.cpp:262-291);
// returns the original bbox {xmin, xmax, ymin, ymax, zmin, zmax}.
std::vector<double> normalize(Mesh &m) {
  const double inf = std::numeric_limits<double>::max();
  double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
  for (const Vec3 &v : m.vertices)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], v[k]), hi[k] = std::max(hi[k], v[k]);
  const double len = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
  const double mid[3] = {(hi[0] + lo[0]) / 2, (hi[1] + lo[1]) / 2, (hi[2] + lo[2]) / 2};
  for (Vec3 &v : m.vertices)
    for (int k = 0; k < 3; ++k) v[k] = 2.0 * (v[k] - mid[k]) / len;
  return {lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]};
}

void unnormalize(Mesh &m, const std::vector<double> &b) {
  const double len = std::max({b[1] - b[0], b[3] - b[2], b[5] - b[4]});
  const double mid[3] = {(b[1] + b[0]) / 2, (b[3] + b[2]) / 2, (b[5] + b[4]) / 2};
  for (Vec3 &v : m.vertices)
    for (int k = 0; k < 3; ++k) v[k] = v[k] / 2 * len + mid[k];
}

double mesh_volume(const Mesh &m) {
  double vol = 0.0;
  for (const auto &t : m.triangles)
    vol += dot(m.vertices[t[0]], cross(m.vertices[t[1]], m.vertices[t[2]]));
  return vol / 6.0;
}

double tri_area(const Vec3 &a, const Vec3 &b, const Vec3 &c) { return 0.5 * norm(cross(b - a, c - a)); }

double surface_area(const Mesh &m) {
  double s = 0.0;
  for (const auto &t : m.triangles) s += tri_area(m.vertices[t[0]], m.vertices[t[1]], m.vertices[t[2]]);
  return s;
}

// ------------------------------------------------------------ convex hull
// Quickhull (Barber, Dobkin, Huhdanpaa 1996): start from a tetrahedron of
// extreme points, repeatedly take the farthest outside point of a face,
// delete the faces it sees, and cone the horizon to it.
namespace {
struct HFace {
  int v[3];
  Vec3 n;
  double off;
  std::vector<int> outside;
  bool alive = true;
};

HFace make_face(const std::vector<Vec3> &P, int a, int b, int c) {
  HFace f;
  f.v[0] = a, f.v[1] = b, f.v[2] = c;
  Vec3 n = cross(P[b] - P[a], P[c] - P[a]);
  const double l = norm(n);
  f.n = l > 0 ? n * (1.0 / l) : Vec3{0, 0, 0};
  f.off = dot(f.n, P[a]);
  return f;
}
}  // namespace

Mesh convex_hull(const std::vector<Vec3> &P) {
  Mesh out;
  const int n = (int)P.size();
  if (n < 4) return out;
  double scale = 0.0;
  for (const Vec3 &p : P) scale = std::max({scale, std::fabs(p[0]), std::fabs(p[1]), std::fabs(p[2])});
  const double eps = 1e-12 * std::max(1.0, scale);
  // initial tetrahedron
  int i0 = 0, i1 = 0;
  for (int i = 1; i < n; ++i) {
    if (P[i][0] < P[i0][0]) i0 = i;
    if (P[i][0] > P[i1][0]) i1 = i;
  }
  if (i0 == i1) return out;
  const Vec3 dir = P[i1] - P[i0];
  int i2 = -1;
  double best = 0.0;
  for (int i = 0; i < n; ++i) {
    const double d = norm(cross(P[i] - P[i0], dir));
    if (d > best) best = d, i2 = i;
  }
  if (i2 < 0 || best <= eps) return out;
  const Vec3 nrm = cross(P[i1] - P[i0], P[i2] - P[i0]);
  int i3 = -1;
  best = 0.0;
  for (int i = 0; i < n; ++i) {
    const double d = std::fabs(dot(P[i] - P[i0], nrm));
    if (d > best) best = d, i3 = i;
  }
  if (i3 < 0 || best <= eps * norm(nrm)) return out;
  std::vector<HFace> F;
  if (dot(P[i3] - P[i0], nrm) > 0) std::swap(i1, i2);  // i3 below face (i0, i1, i2)
  F.push_back(make_face(P, i0, i1, i2));
  F.push_back(make_face(P, i0, i3, i1));
  F.push_back(make_face(P, i1, i3, i2));
  F.push_back(make_face(P, i2, i3, i0));
  for (int i = 0; i < n; ++i) {
    if (i == i0 || i == i1 || i == i2 || i == i3) continue;
    for (auto &f : F)
      if (dot(f.n, P[i]) - f.off > eps) {
        f.outside.push_back(i);
        break;
      }
  }
  std::vector<int> stack = {0, 1, 2, 3};
  while (!stack.empty()) {
    const int fi = stack.back();
    stack.pop_back();
    if (!F[fi].alive || F[fi].outside.empty()) continue;
    // farthest point of the face
    int apex = -1;
    double far = -1.0;
    for (int i : F[fi].outside) {
      const double d = dot(F[fi].n, P[i]) - F[fi].off;
      if (d > far) far = d, apex = i;
    }
    // faces visible from the apex (flood over edge adjacency via a scan)
    std::vector<int> visible;
    for (int g = 0; g < (int)F.size(); ++g)
      if (F[g].alive && dot(F[g].n, P[apex]) - F[g].off > eps) visible.push_back(g);
    std::unordered_set<uint64_t> dedges;
    auto key = [](int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };
    for (int g : visible)
      for (int e = 0; e < 3; ++e) dedges.insert(key(F[g].v[e], F[g].v[(e + 1) % 3]));
    std::vector<int> orphans;
    std::vector<std::pair<int, int>> horizon;
    for (int g : visible) {
      for (int e = 0; e < 3; ++e) {
        const int a = F[g].v[e], b = F[g].v[(e + 1) % 3];
        if (!dedges.count(key(b, a))) horizon.push_back({a, b});
      }
      F[g].alive = false;
      for (int i : F[g].outside)
        if (i != apex) orphans.push_back(i);
      F[g].outside.clear();
    }
    const int first_new = (int)F.size();
    for (auto &h : horizon) F.push_back(make_face(P, h.first, h.second, apex));
    for (int i : orphans)
      for (int g = first_new; g < (int)F.size(); ++g)
        if (dot(F[g].n, P[i]) - F[g].off > eps) {
          F[g].outside.push_back(i);
          break;
        }
    for (int g = first_new; g < (int)F.size(); ++g) stack.push_back(g);
    // compact dead faces now and then
    if (F.size() > 4 * 1024 && F.size() > 8 * (size_t)n) break;
  }
  std::unordered_map<int, int> remap;
  for (const auto &f : F) {
    if (!f.alive) continue;
    std::array<int, 3> t;
    for (int e = 0; e < 3; ++e) {
      auto it = remap.find(f.v[e]);
      if (it == remap.end()) {
        it = remap.emplace(f.v[e], (int)out.vertices.size()).first;
        out.vertices.push_back(P[f.v[e]]);
      }
      t[e] = it->second;
    }
    out.triangles.push_back(t);
  }
  return out;
}

// ------------------------------------------------------ connected components
// Triangles sharing a vertex are connected; each component becomes a part
// with its vertices renumbered in increasing old-id order, edges and flags
// carried over (lib/src/postprocess.cpp:202-326).  Components with
// |volume| < 1e-6 are dropped when drop_flat (lib/src/postprocess.cpp:319).
MeshList separate_disjoint(const Mesh &m, bool drop_flat) {
  const int nv = (int)m.vertices.size();
  std::vector<int> parent(nv);
  std::iota(parent.begin(), parent.end(), 0);
  std::function<int(int)> find = [&](int x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  for (const auto &t : m.triangles) {
    const int a = find(t[0]), b = find(t[1]), c = find(t[2]);
    parent[b] = a;
    parent[find(c)] = find(a);
  }
  std::unordered_map<int, int> comp_of_root;
  std::vector<int> comp(nv, -1);
  std::vector<char> used(nv, 0);
  for (const auto &t : m.triangles)
    for (int e = 0; e < 3; ++e) used[t[e]] = 1;
  int nc = 0;
  for (int v = 0; v < nv; ++v) {
    if (!used[v]) continue;
    const int r = find(v);
    auto it = comp_of_root.find(r);
    if (it == comp_of_root.end()) it = comp_of_root.emplace(r, nc++).first;
    comp[v] = it->second;
  }
  if (nc <= 1 && (int)std::count(used.begin(), used.end(), 1) == nv) {
    MeshList one = {m};
    if (drop_flat && std::fabs(mesh_volume(m)) < 1e-6) one.clear();
    return one;
  }
  MeshList parts(nc);
  std::vector<int> local(nv, -1);
  for (int v = 0; v < nv; ++v) {
    if (comp[v] < 0) continue;
    Mesh &p = parts[comp[v]];
    local[v] = (int)p.vertices.size();
    p.vertices.push_back(m.vertices[v]);
    p.is_new.push_back(v < (int)m.is_new.size() ? (bool)m.is_new[v] : false);
  }
  for (const auto &t : m.triangles)
    parts[comp[t[0]]].triangles.push_back({local[t[0]], local[t[1]], local[t[2]]});
  for (const auto &e : m.intersecting_edges) {
    if (comp[e.first] >= 0 && comp[e.first] == comp[e.second])
      parts[comp[e.first]].intersecting_edges.push_back(
          {(unsigned)local[e.first], (unsigned)local[e.second]});
  }
  for (int v : m.points)
    if (comp[v] >= 0) parts[comp[v]].points.push_back(local[v]);
  MeshList out;
  for (auto &p : parts) {
    if (drop_flat && std::fabs(mesh_volume(p)) < 1e-6) continue;
    out.push_back(std::move(p));
  }
  return out;
}

}  // namespace vhost
