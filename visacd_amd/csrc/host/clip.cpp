// Plane clip of a closed triangle mesh with capped cross-sections.
//
// Semantics follow the reference clip (lib/src/clip.cpp:190-741): vertices
// with side +1 / -1 (dead zone 1e-6) go to the positive / negative part in
// increasing id order with 1-based maps (0 = absent); vertices on the plane
// and edge/plane intersection points become cap vertices appended after
// them and flagged is_new (lib/src/clip.cpp:664-673).  The cap polygons are
// triangulated here by ear clipping with hole bridging (D. Eberly,
// "Triangulation by Ear Clipping", Geometric Tools, 2002) instead of the
// reference's constrained Delaunay triangulation.  New vertices closer than
// float32 resolution are then welded (weld_short_new_edges).
#include <algorithm>
#include <cmath>
#include <exception>
#include <iterator>
#include <map>
#include <thread>
#include <unordered_map>

#include "mesh.hpp"

namespace vhost {
namespace {

struct P2 {
  double x, y;
  int id;  // cap vertex index
};

inline double orient(const P2 &a, const P2 &b, const P2 &c) {
  return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
}

double signed_area(const std::vector<P2> &poly) {
  double s = 0;
  for (size_t i = 0, n = poly.size(); i < n; ++i) {
    const P2 &a = poly[i], &b = poly[(i + 1) % n];
    s += a.x * b.y - b.x * a.y;
  }
  return 0.5 * s;
}

bool point_in_poly(const std::vector<P2> &poly, double x, double y) {
  bool in = false;
  for (size_t i = 0, n = poly.size(), j = n - 1; i < n; j = i++) {
    const P2 &a = poly[i], &b = poly[j];
    if ((a.y > y) != (b.y > y)) {
      const double xi = a.x + (y - a.y) * (b.x - a.x) / (b.y - a.y);
      if (x < xi) in = !in;
    }
  }
  return in;
}

bool in_tri(const P2 &a, const P2 &b, const P2 &c, const P2 &p) {
  return orient(a, b, p) >= 0 && orient(b, c, p) >= 0 && orient(c, a, p) >= 0;
}

// Splices a CW hole into a CCW outer polygon through a mutually visible pair.
void bridge_hole(std::vector<P2> &outer, const std::vector<P2> &hole) {
  int mi = 0;
  for (int i = 1; i < (int)hole.size(); ++i)
    if (hole[i].x > hole[mi].x) mi = i;
  const P2 M = hole[mi];
  // nearest outer edge hit by the ray M + t(1, 0)
  int best_e = -1;
  double best_x = 1e300;
  const int n = (int)outer.size();
  for (int i = 0; i < n; ++i) {
    const P2 &a = outer[i], &b = outer[(i + 1) % n];
    if ((a.y > M.y) == (b.y > M.y) && !(a.y == M.y || b.y == M.y)) continue;
    if (a.y == b.y) continue;
    const double xi = a.x + (M.y - a.y) * (b.x - a.x) / (b.y - a.y);
    if (xi >= M.x && xi < best_x) best_x = xi, best_e = i;
  }
  int pi;
  if (best_e < 0) {
    // fall back to the nearest outer vertex
    pi = 0;
    double bd = 1e300;
    for (int i = 0; i < n; ++i) {
      const double d = std::hypot(outer[i].x - M.x, outer[i].y - M.y);
      if (d < bd) bd = d, pi = i;
    }
  } else {
    const P2 &a = outer[best_e], &b = outer[(best_e + 1) % n];
    pi = a.x > b.x ? best_e : (best_e + 1) % n;
    const P2 I{best_x, M.y, -1};
    // a reflex vertex inside (M, I, P) is a better candidate: smallest angle
    const P2 P = outer[pi];
    double best_ang = 1e300;
    for (int i = 0; i < n; ++i) {
      const P2 &prev = outer[(i + n - 1) % n], &cur = outer[i], &next = outer[(i + 1) % n];
      if (orient(prev, cur, next) >= 0) continue;  // convex
      const bool inside = orient(M, I, P) > 0 ? in_tri(M, I, P, cur) : in_tri(M, P, I, cur);
      if (!inside || (cur.x == P.x && cur.y == P.y)) continue;
      const double ang = std::atan2(std::fabs(cur.y - M.y), cur.x - M.x);
      if (ang < best_ang) best_ang = ang, pi = i;
    }
  }
  std::vector<P2> merged;
  merged.reserve(outer.size() + hole.size() + 2);
  for (int i = 0; i <= pi; ++i) merged.push_back(outer[i]);
  for (int k = 0; k <= (int)hole.size(); ++k) merged.push_back(hole[(mi + k) % hole.size()]);
  for (int i = pi; i < n; ++i) merged.push_back(outer[i]);
  outer.swap(merged);
}

// Ear clipping of a simple CCW polygon (bridges allowed); appends triangles
// of cap ids.  Returns false if it gets stuck.  The ear test ("does any other
// remaining vertex lie in the candidate ear") looks only at the vertices in
// the cells of a uniform grid that the ear's box overlaps, widened by a cell
// (a point on a cell border may round into the next cell); the answer, and so
// the triangulation, is the one of a scan over every remaining vertex, at a
// fraction of the cost on the long cap polygons of large parts.
bool ear_clip(std::vector<P2> poly, std::vector<std::array<int, 3>> &tris) {
  int n = (int)poly.size();
  if (n < 3) return n == 0;
  std::vector<int> prv(n), nxt(n);
  for (int i = 0; i < n; ++i) prv[i] = (i + n - 1) % n, nxt[i] = (i + 1) % n;
  int remaining = n, cur = 0, guard = 0;
  double scale = 0;
  for (auto &p : poly) scale = std::max({scale, std::fabs(p.x), std::fabs(p.y)});
  const double eps = 1e-14 * std::max(1.0, scale * scale);
  // the grid: about one vertex per cell; short polygons scan the ring
  const bool use_grid = n >= 64;
  double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
  for (auto &p : poly) x0 = std::min(x0, p.x), y0 = std::min(y0, p.y), x1 = std::max(x1, p.x), y1 = std::max(y1, p.y);
  const int G = use_grid ? std::min(1024, (int)std::sqrt((double)n)) : 1;
  const double cw = (x1 - x0) / G, ch = (y1 - y0) / G;
  auto cx = [&](double x) { return cw > 0 ? std::min(G - 1, std::max(0, (int)((x - x0) / cw))) : 0; };
  auto cy = [&](double y) { return ch > 0 ? std::min(G - 1, std::max(0, (int)((y - y0) / ch))) : 0; };
  std::vector<int> cell_start, cell_items;
  std::vector<char> alive;
  if (use_grid) {
    cell_start.assign(G * G + 1, 0);
    cell_items.resize(n);
    alive.assign(n, 1);
    for (int i = 0; i < n; ++i) ++cell_start[cy(poly[i].y) * G + cx(poly[i].x) + 1];
    for (int c = 0; c < G * G; ++c) cell_start[c + 1] += cell_start[c];
    std::vector<int> fill(cell_start.begin(), cell_start.end() - 1);
    for (int i = 0; i < n; ++i) cell_items[fill[cy(poly[i].y) * G + cx(poly[i].x)]++] = i;
  }
  while (remaining > 3) {
    if (guard++ > 4 * remaining + 16) {
      // no strict ear: clip the most convex vertex (degenerate input)
      int best = cur;
      double bo = -1e300;
      int v = cur;
      for (int k = 0; k < remaining; ++k, v = nxt[v]) {
        const double o = orient(poly[prv[v]], poly[v], poly[nxt[v]]);
        if (o > bo) bo = o, best = v;
      }
      if (bo < -eps) return false;
      cur = best;
    } else {
      const int ip = prv[cur], in = nxt[cur];
      const P2 &a = poly[ip], &b = poly[cur], &c = poly[in];
      bool ear = orient(a, b, c) > eps;
      if (ear && !use_grid) {
        for (int v = nxt[in]; v != ip; v = nxt[v]) {
          const P2 &p = poly[v];
          if (p.id == a.id || p.id == b.id || p.id == c.id) continue;
          if (in_tri(a, b, c, p)) {
            ear = false;
            break;
          }
        }
      } else if (ear) {
        const int gx0 = std::max(0, cx(std::min({a.x, b.x, c.x})) - 1);
        const int gx1 = std::min(G - 1, cx(std::max({a.x, b.x, c.x})) + 1);
        const int gy0 = std::max(0, cy(std::min({a.y, b.y, c.y})) - 1);
        const int gy1 = std::min(G - 1, cy(std::max({a.y, b.y, c.y})) + 1);
        for (int gy = gy0; gy <= gy1 && ear; ++gy)
          for (int gx = gx0; gx <= gx1 && ear; ++gx)
            for (int k = cell_start[gy * G + gx]; k < cell_start[gy * G + gx + 1]; ++k) {
              const int v = cell_items[k];
              if (!alive[v] || v == ip || v == cur || v == in) continue;
              const P2 &p = poly[v];
              if (p.id == a.id || p.id == b.id || p.id == c.id) continue;
              if (in_tri(a, b, c, p)) {
                ear = false;
                break;
              }
            }
      }
      if (!ear) {
        cur = nxt[cur];
        continue;
      }
    }
    tris.push_back({poly[prv[cur]].id, poly[cur].id, poly[nxt[cur]].id});
    const int p = prv[cur], q = nxt[cur];
    nxt[p] = q;
    prv[q] = p;
    if (use_grid) alive[cur] = 0;
    --remaining;
    cur = p;
    guard = 0;
  }
  const int a = prv[cur], c = nxt[cur];
  tris.push_back({poly[a].id, poly[cur].id, poly[c].id});
  return true;
}

// Cap triangles (cap ids) closing loops of directed plane edges.  Edges are
// given as the part's own boundary direction; the cap runs them backwards.
bool build_cap(const std::vector<std::pair<int, int>> &edges, const std::vector<Vec3> &cap_pts,
               const Vec3 &cap_normal, std::vector<std::array<int, 3>> &tris) {
  if (edges.empty()) return true;
  std::unordered_multimap<int, int> nxt;  // reversed: y -> x
  for (auto &e : edges) nxt.emplace(e.second, e.first);
  // 2D basis with e1 x e2 = cap_normal
  Vec3 e1 = std::fabs(cap_normal[0]) < 0.9 ? cross(cap_normal, Vec3{1, 0, 0}) : cross(cap_normal, Vec3{0, 1, 0});
  e1 = e1 * (1.0 / norm(e1));
  const Vec3 e2 = cross(cap_normal, e1);
  std::vector<std::vector<P2>> loops;
  while (!nxt.empty()) {
    auto it = nxt.begin();
    const int start = it->first;
    std::vector<P2> loop;
    int v = start;
    for (size_t guard = 0; guard <= edges.size(); ++guard) {
      auto jt = nxt.find(v);
      if (jt == nxt.end()) return false;  // open chain
      const Vec3 &p = cap_pts[v];
      loop.push_back(P2{dot(p, e1), dot(p, e2), v});
      const int w = jt->second;
      nxt.erase(jt);
      v = w;
      if (v == start) break;
    }
    if (v != start) return false;
    if (loop.size() >= 3) loops.push_back(std::move(loop));
  }
  std::vector<int> outer, holes;
  for (int i = 0; i < (int)loops.size(); ++i)
    (signed_area(loops[i]) > 0 ? outer : holes).push_back(i);
  if (outer.empty()) return false;
  std::vector<std::vector<int>> holes_of(loops.size());
  for (int h : holes) {
    int owner = -1;
    double owner_area = 1e300;
    for (int o : outer) {
      const double a = signed_area(loops[o]);
      if (a < owner_area && point_in_poly(loops[o], loops[h][0].x, loops[h][0].y))
        owner = o, owner_area = a;
    }
    if (owner < 0) return false;
    holes_of[owner].push_back(h);
  }
  for (int o : outer) {
    std::vector<P2> poly = loops[o];
    auto &hs = holes_of[o];
    std::sort(hs.begin(), hs.end(), [&](int a, int b) {
      double ma = -1e300, mb = -1e300;
      for (auto &p : loops[a]) ma = std::max(ma, p.x);
      for (auto &p : loops[b]) mb = std::max(mb, p.x);
      return ma > mb;
    });
    for (int h : hs) bridge_hole(poly, loops[h]);
    if (!ear_clip(poly, tris)) return false;
  }
  return true;
}

// Collapses the edges between two new (cap) vertices shorter than 4e-7 of the
// part's largest coordinate: crossing points that close round to one float32
// position, and the triangles between them would have zero area in the
// visibility engine's float32 geometry.  An edge (u, v) is collapsed into u
// when its two triangles are the only ones holding both (the link
// condition), which keeps the mesh closed and manifold; the two triangles go.
// Only new vertices are removed, and they follow the old ones, so the parent
// maps stay valid.  Returns the number of collapsed edges.
int weld_short_new_edges(Mesh &m) {
  const int nv = (int)m.vertices.size();
  double scale = 0.0;
  for (const Vec3 &v : m.vertices)
    for (int k = 0; k < 3; ++k) scale = std::max(scale, std::fabs(v[k]));
  const double delta = 4e-7 * scale;
  auto fresh = [&](int v) { return v < (int)m.is_new.size() && m.is_new[v]; };
  std::vector<std::pair<double, std::pair<int, int>>> cand;
  for (const auto &t : m.triangles)
    for (int k = 0; k < 3; ++k) {
      const int u = t[k], v = t[(k + 1) % 3];
      if (u < v && fresh(u) && fresh(v)) {
        const double d = norm(m.vertices[u] - m.vertices[v]);
        if (d < delta) cand.push_back({d, {u, v}});
      }
    }
  if (cand.empty()) return 0;
  std::sort(cand.begin(), cand.end());
  std::vector<std::vector<int>> vt(nv);  // vertex -> triangles (may hold dead ones)
  for (int t = 0; t < (int)m.triangles.size(); ++t)
    for (int k = 0; k < 3; ++k) vt[m.triangles[t][k]].push_back(t);
  std::vector<char> dead_t(m.triangles.size(), 0), dead_v(nv, 0);
  auto holds = [&](int t, int x) {
    const auto &r = m.triangles[t];
    return r[0] == x || r[1] == x || r[2] == x;
  };
  auto ring = [&](int u) {
    std::vector<int> r;
    for (int t : vt[u])
      if (!dead_t[t] && holds(t, u))
        for (int k = 0; k < 3; ++k)
          if (m.triangles[t][k] != u) r.push_back(m.triangles[t][k]);
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    return r;
  };
  int collapsed = 0;
  for (const auto &c : cand) {
    const int u = c.second.first, v = c.second.second;
    if (dead_v[u] || dead_v[v]) continue;
    std::vector<int> both;  // live triangles holding u and v
    for (int t : vt[u])
      if (!dead_t[t] && holds(t, u) && holds(t, v)) both.push_back(t);
    std::sort(both.begin(), both.end());
    both.erase(std::unique(both.begin(), both.end()), both.end());
    if (both.size() != 2) continue;
    std::vector<int> opp;
    for (int t : both)
      for (int k = 0; k < 3; ++k)
        if (m.triangles[t][k] != u && m.triangles[t][k] != v) opp.push_back(m.triangles[t][k]);
    std::sort(opp.begin(), opp.end());
    if (opp.size() != 2 || opp[0] == opp[1]) continue;
    const std::vector<int> ru = ring(u), rv = ring(v);
    std::vector<int> common;
    std::set_intersection(ru.begin(), ru.end(), rv.begin(), rv.end(), std::back_inserter(common));
    if (common != opp) continue;  // link condition
    for (int t : both) dead_t[t] = 1;
    for (int t : vt[v]) {
      if (dead_t[t] || !holds(t, v)) continue;
      for (int k = 0; k < 3; ++k)
        if (m.triangles[t][k] == v) m.triangles[t][k] = u;
      vt[u].push_back(t);
    }
    dead_v[v] = 1;
    ++collapsed;
  }
  if (!collapsed) return 0;
  std::vector<int> id(nv, -1);
  Mesh out;
  for (int i = 0; i < nv; ++i)
    if (!dead_v[i]) {
      id[i] = (int)out.vertices.size();
      out.vertices.push_back(m.vertices[i]);
      out.is_new.push_back(fresh(i));
    }
  for (int t = 0; t < (int)m.triangles.size(); ++t)
    if (!dead_t[t])
      out.triangles.push_back({id[m.triangles[t][0]], id[m.triangles[t][1]], id[m.triangles[t][2]]});
  m.vertices.swap(out.vertices);
  m.triangles.swap(out.triangles);
  m.is_new.swap(out.is_new);
  return collapsed;
}

}  // namespace

MeshList clip(const Mesh &m, const Plane &pl, std::vector<int> &map_pos, std::vector<int> &map_neg) {
  const int N = (int)m.vertices.size();
  const int64_t T = (int64_t)m.triangles.size();
  // large meshes: the per-vertex and per-triangle passes on several threads,
  // the two halves built at once (the output does not depend on it).
  // Triangle kinds: 0 on the plane (covered by the caps), 1 wholly on the +
  // side, 2 wholly on the - side, 3 touching or crossing the plane.
  const bool big = T >= 65536;
  std::vector<int> side(N);
  std::vector<uint8_t> kind(T);
  auto classify = [&](int64_t t) {
    const auto &tr = m.triangles[t];
    const int s0 = side[tr[0]], s1 = side[tr[1]], s2 = side[tr[2]];
    kind[t] = (s0 == 0 && s1 == 0 && s2 == 0) ? 0
              : (s0 == 1 && s1 == 1 && s2 == 1) ? 1
              : (s0 == -1 && s1 == -1 && s2 == -1) ? 2 : 3;
  };
  if (big) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) side[i] = pl.side(m.vertices[i]);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < T; ++t) classify(t);
  } else {
    for (int i = 0; i < N; ++i) side[i] = pl.side(m.vertices[i]);
    for (int64_t t = 0; t < T; ++t) classify(t);
  }
  map_pos.assign(N, 0);
  map_neg.assign(N, 0);
  // cap points: on-plane vertices and edge crossings, keyed, numbered in
  // order of first use over the touching triangles in triangle order
  std::map<std::pair<int, int>, int> cap_key;  // (a, b) a<=b; (v, v) for a vertex
  std::vector<Vec3> cap_pts;
  auto cap_vertex = [&](int v) {
    auto it = cap_key.find({v, v});
    if (it != cap_key.end()) return it->second;
    cap_key[{v, v}] = (int)cap_pts.size();
    cap_pts.push_back(m.vertices[v]);
    return (int)cap_pts.size() - 1;
  };
  auto cap_cross = [&](int a, int b) {
    const int lo = std::min(a, b), hi = std::max(a, b);
    auto it = cap_key.find({lo, hi});
    if (it != cap_key.end()) return it->second;
    const Vec3 &p = m.vertices[lo], &q = m.vertices[hi];
    const double fp = pl.eval(p), fq = pl.eval(q);
    const double t = fp / (fp - fq);
    cap_key[{lo, hi}] = (int)cap_pts.size();
    cap_pts.push_back(p + (q - p) * t);
    return (int)cap_pts.size() - 1;
  };
  // pieces of the touching triangles: polygon corners (at most four: a
  // triangle cut by a plane) encoded as vertex id >= 0 or ~cap id; tri: the
  // triangle it comes from
  struct Piece {
    int c[4];
    int k;
    int64_t tri;
  };
  std::vector<Piece> pieces[2];
  std::vector<std::pair<int, int>> plane_edges[2];
  for (int64_t t = 0; t < T; ++t) {
    if (kind[t] != 3) continue;
    const auto &tr = m.triangles[t];
    const int s[3] = {side[tr[0]], side[tr[1]], side[tr[2]]};
    for (int w = 0; w < 2; ++w) {
      const int want = w == 0 ? 1 : -1;
      if (s[0] != want && s[1] != want && s[2] != want) continue;
      Piece pc;
      pc.k = 0;
      pc.tri = t;
      for (int e = 0; e < 3; ++e) {
        const int a = tr[e], b = tr[(e + 1) % 3];
        const int sa = s[e], sb = s[(e + 1) % 3];
        if (sa == want) pc.c[pc.k++] = a;
        else if (sa == 0) pc.c[pc.k++] = ~cap_vertex(a);
        if (sa != 0 && sb != 0 && sa != sb) pc.c[pc.k++] = ~cap_cross(a, b);
      }
      for (int e = 0; e < pc.k; ++e) {
        const int a = pc.c[e], b = pc.c[(e + 1) % pc.k];
        if (a < 0 && b < 0 && a != b) plane_edges[w].push_back({~a, ~b});
      }
      pieces[w].push_back(pc);
    }
  }
  MeshList out;
  int npos = 0, nneg = 0;
  for (int i = 0; i < N; ++i) npos += side[i] > 0, nneg += side[i] < 0;
  if (npos == 0 || nneg == 0) {
    out.push_back(m);
    return out;
  }
  const Vec3 n = {pl.a, pl.b, pl.c};
  Mesh parts[2];
  bool ok[2] = {true, true};
  // half w: its vertices in increasing id order with 1-based maps, then its
  // triangles in triangle order (whole ones, then the fan of each piece),
  // cap vertices appended in order of first use, the cap, the weld
  auto build = [&](int w) {
    Mesh &part = parts[w];
    std::vector<int> &map = w == 0 ? map_pos : map_neg;
    const int want = w == 0 ? 1 : -1;
    for (int i = 0; i < N; ++i)
      if (side[i] == want) {
        map[i] = (int)part.vertices.size() + 1;
        part.vertices.push_back(m.vertices[i]);
        part.is_new.push_back(false);
      }
    std::vector<int> cap_local(cap_pts.size(), -1);
    auto local = [&](int code) {
      if (code >= 0) return map[code] - 1;
      const int c = ~code;
      if (cap_local[c] < 0) {
        cap_local[c] = (int)part.vertices.size();
        part.vertices.push_back(cap_pts[c]);
        part.is_new.push_back(true);
      }
      return cap_local[c];
    };
    part.triangles.reserve((size_t)(T / 2) + 2 * pieces[w].size() + 2 * plane_edges[w].size());
    const uint8_t whole = w == 0 ? 1 : 2;
    size_t pi = 0;
    for (int64_t t = 0; t < T; ++t) {
      if (kind[t] == whole) {
        const auto &tr = m.triangles[t];
        part.triangles.push_back({map[tr[0]] - 1, map[tr[1]] - 1, map[tr[2]] - 1});
      } else if (kind[t] == 3) {
        for (; pi < pieces[w].size() && pieces[w][pi].tri == t; ++pi) {
          const Piece &pc = pieces[w][pi];
          for (int k = 1; k + 1 < pc.k; ++k)
            part.triangles.push_back({local(pc.c[0]), local(pc.c[k]), local(pc.c[k + 1])});
        }
      }
    }
    std::vector<std::array<int, 3>> cap;
    const Vec3 cap_normal = w == 0 ? n * -1.0 : n;
    if (!build_cap(plane_edges[w], cap_pts, cap_normal * (1.0 / norm(cap_normal)), cap)) {
      ok[w] = false;
      return;
    }
    for (auto &t : cap) part.triangles.push_back({local(~t[0]), local(~t[1]), local(~t[2])});
    // a collapse can leave a new short edge next to it: repeat until none
    for (int pass = 0; pass < 8 && weld_short_new_edges(part) > 0; ++pass) {
    }
  };
  if (big) {
    std::exception_ptr err;
    std::thread other([&] {
      try {
        build(1);
      } catch (...) {
        err = std::current_exception();
      }
    });
    try {
      build(0);
    } catch (...) {
      other.join();
      throw;
    }
    other.join();
    if (err) std::rethrow_exception(err);
  } else {
    build(0);
    build(1);
  }
  if (!ok[0] || !ok[1]) {
    // a failed + side leaves the - side's map unset, as a serial build would
    if (!ok[0]) map_neg.assign(N, 0);
    out.push_back(m);
    return out;
  }
  out.push_back(std::move(parts[0]));
  out.push_back(std::move(parts[1]));
  return out;
}

}  // namespace vhost
