// Normalisation, volumes, convex hull and connected components.
#include <algorithm>
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
// Quickhull (Barber, Dobkin, Huhdanpaa 1996) with face adjacency: start from
// a tetrahedron of extreme points; repeatedly take the farthest outside
// point of a face, collect the faces it sees by a flood over neighbours,
// cone the horizon to it and hand the orphaned points to the new faces.
// Outside sets are singly linked lists through one per-point array (points
// keep their insertion order), faces are plain records, and the horizon's
// adjacency uses per-vertex slots: no allocation per face or per step.
namespace {
struct HFace {
  int v[3];
  int nb[3];  // face across edge (v[e], v[e+1])
  Vec3 n;
  double off;
  int head, tail;  // outside points, in insertion order (-1: none)
  int mark;
  bool alive;
};

void set_plane(HFace &f, const std::vector<Vec3> &P) {
  Vec3 n = cross(P[f.v[1]] - P[f.v[0]], P[f.v[2]] - P[f.v[0]]);
  const double l = norm(n);
  f.n = l > 0 ? n * (1.0 / l) : Vec3{0, 0, 0};
  f.off = dot(f.n, P[f.v[0]]);
}

HFace new_face(int a, int b, int c, const std::vector<Vec3> &P) {
  HFace f;
  f.v[0] = a, f.v[1] = b, f.v[2] = c;
  f.nb[0] = f.nb[1] = f.nb[2] = -1;
  f.head = f.tail = -1;
  f.mark = -1;
  f.alive = true;
  set_plane(f, P);
  return f;
}

// Hull triangles (ids into Pall, outward) of the points Pall[ids[k]]; empty
// when they are fewer than four or flat.
std::vector<std::array<int, 3>> quickhull(const std::vector<Vec3> &Pall, const std::vector<int> &ids) {
  std::vector<std::array<int, 3>> out;
  const int n = (int)ids.size();
  if (n < 4) return out;
  std::vector<Vec3> P(n);
  for (int i = 0; i < n; ++i) P[i] = Pall[ids[i]];
  double scale = 0.0;
  for (const Vec3 &p : P) scale = std::max({scale, std::fabs(p[0]), std::fabs(p[1]), std::fabs(p[2])});
  const double eps = 1e-12 * std::max(1.0, scale);
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
  if (dot(P[i3] - P[i0], nrm) > 0) std::swap(i1, i2);  // i3 below face (i0, i1, i2)
  std::vector<HFace> F;
  F.reserve(2 * (size_t)n + 64);  // about two faces are made per point: no regrowth
  const int tv[4][3] = {{i0, i1, i2}, {i0, i3, i1}, {i1, i3, i2}, {i2, i3, i0}};
  for (int f = 0; f < 4; ++f) F.push_back(new_face(tv[f][0], tv[f][1], tv[f][2], P));
  // tetrahedron adjacency by matching reversed edges
  for (int f = 0; f < 4; ++f)
    for (int e = 0; e < 3; ++e) {
      const int a = F[f].v[e], b = F[f].v[(e + 1) % 3];
      for (int g = 0; g < 4; ++g)
        for (int h = 0; h < 3; ++h)
          if (F[g].v[h] == b && F[g].v[(h + 1) % 3] == a) F[f].nb[e] = g;
    }
  std::vector<int> next(n, -1);  // outside-list links
  auto append = [&](HFace &f, int i) {
    next[i] = -1;
    if (f.tail < 0) f.head = i;
    else next[f.tail] = i;
    f.tail = i;
  };
  for (int i = 0; i < n; ++i) {
    if (i == i0 || i == i1 || i == i2 || i == i3) continue;
    for (auto &f : F)
      if (dot(f.n, P[i]) - f.off > eps) {
        append(f, i);
        break;
      }
  }
  std::vector<int> work = {0, 1, 2, 3};
  std::vector<int> visible, queue, orphans;
  struct Hz {
    int a, b, other, other_e;
  };
  std::vector<Hz> horizon;
  // per vertex: the new face whose horizon edge starts / ends there (set for
  // every horizon vertex before it is read in the same step)
  std::vector<int> start_at(n, -1), end_at(n, -1);
  int iter = 0;
  while (!work.empty()) {
    const int fi = work.back();
    work.pop_back();
    if (!F[fi].alive || F[fi].head < 0) continue;
    ++iter;
    int apex = -1;
    double far = -1.0;
    {
      const Vec3 fn = F[fi].n;
      const double off = F[fi].off;
      for (int i = F[fi].head; i >= 0; i = next[i]) {
        const double d = dot(fn, P[i]) - off;
        if (d > far) far = d, apex = i;
      }
    }
    const Vec3 A = P[apex];
    visible.clear();
    horizon.clear();
    queue.assign(1, fi);
    F[fi].mark = iter;
    visible.push_back(fi);
    for (size_t q = 0; q < queue.size(); ++q) {
      const int f = queue[q];
      for (int e = 0; e < 3; ++e) {
        const int g = F[f].nb[e];
        if (F[g].mark == iter) continue;
        if (dot(F[g].n, A) - F[g].off > eps) {
          F[g].mark = iter;
          visible.push_back(g);
          queue.push_back(g);
        }
      }
    }
    for (int f : visible)
      for (int e = 0; e < 3; ++e) {
        const int g = F[f].nb[e];
        if (F[g].mark == iter) continue;
        int ge = 0;
        while (F[g].nb[ge] != f) ++ge;
        horizon.push_back({F[f].v[e], F[f].v[(e + 1) % 3], g, ge});
      }
    orphans.clear();
    for (int f : visible) {
      F[f].alive = false;
      for (int i = F[f].head; i >= 0; i = next[i])
        if (i != apex) orphans.push_back(i);
      F[f].head = F[f].tail = -1;
    }
    const int first = (int)F.size();
    for (const Hz &h : horizon) {
      const int id = (int)F.size();
      F.push_back(new_face(h.a, h.b, apex, P));
      F[id].nb[0] = h.other;
      F[h.other].nb[h.other_e] = id;
      start_at[h.a] = id;
      end_at[h.b] = id;
    }
    for (int id = first; id < (int)F.size(); ++id) {
      F[id].nb[1] = start_at[F[id].v[1]];  // edge (b, apex): face whose horizon edge starts at b
      F[id].nb[2] = end_at[F[id].v[0]];    // edge (apex, a): face whose horizon edge ends at a
    }
    const int last = (int)F.size();
    for (int i : orphans) {
      const Vec3 &p = P[i];
      for (int g = first; g < last; ++g)
        if (dot(F[g].n, p) - F[g].off > eps) {
          append(F[g], i);
          break;
        }
    }
    for (int g = first; g < last; ++g)
      if (F[g].head >= 0) work.push_back(g);
  }
  for (const auto &f : F)
    if (f.alive) out.push_back({ids[f.v[0]], ids[f.v[1]], ids[f.v[2]]});
  return out;
}

}  // namespace

// Hull of a point set (vertices in order of first use by the faces that
// survive, faces in creation order).
Mesh convex_hull(const std::vector<Vec3> &P) {
  const int n = (int)P.size();
  std::vector<int> all(n);
  std::iota(all.begin(), all.end(), 0);
  const std::vector<std::array<int, 3>> tris = quickhull(P, all);
  Mesh out;
  std::vector<int> remap(n, -1);
  out.triangles.reserve(tris.size());
  for (const auto &t : tris) {
    std::array<int, 3> r;
    for (int e = 0; e < 3; ++e) {
      int &m = remap[t[e]];
      if (m < 0) {
        m = (int)out.vertices.size();
        out.vertices.push_back(P[t[e]]);
      }
      r[e] = m;
    }
    out.triangles.push_back(r);
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
  auto find = [&](int x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  for (const auto &t : m.triangles) {
    const int a = find(t[0]), b = find(t[1]);
    parent[b] = a;
    parent[find(t[2])] = find(a);
  }
  // components numbered in order of their lowest used vertex
  std::vector<int> comp(nv, -1), comp_of_root(nv, -1);
  std::vector<char> used(nv, 0);
  for (const auto &t : m.triangles)
    for (int e = 0; e < 3; ++e) used[t[e]] = 1;
  int nc = 0, nused = 0;
  for (int v = 0; v < nv; ++v) {
    if (!used[v]) continue;
    ++nused;
    int &c = comp_of_root[find(v)];
    if (c < 0) c = nc++;
    comp[v] = c;
  }
  if (nc <= 1 && nused == nv) {
    MeshList one;
    if (!(drop_flat && std::fabs(mesh_volume(m)) < 1e-6)) one.push_back(m);
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

MeshList separate_disjoint(Mesh &&m, bool drop_flat) {
  const int nv = (int)m.vertices.size();
  std::vector<char> used(nv, 0);
  for (const auto &t : m.triangles)
    for (int e = 0; e < 3; ++e) used[t[e]] = 1;
  bool all_used = std::count(used.begin(), used.end(), 1) == nv;
  if (all_used && nv > 0) {  // one component? then the mesh moves through unchanged
    std::vector<int> parent(nv);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int x) {
      while (parent[x] != x) x = parent[x] = parent[parent[x]];
      return x;
    };
    for (const auto &t : m.triangles) {
      const int a = find(t[0]), b = find(t[1]);
      parent[b] = a;
      parent[find(t[2])] = find(a);
    }
    const int r = find(0);
    bool one = true;
    for (int v = 1; v < nv && one; ++v) one = find(v) == r;
    if (one) {
      MeshList out;
      if (!(drop_flat && std::fabs(mesh_volume(m)) < 1e-6)) out.push_back(std::move(m));
      return out;
    }
  }
  return separate_disjoint(static_cast<const Mesh &>(m), drop_flat);
}

}  // namespace vhost
