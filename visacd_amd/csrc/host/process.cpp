// Decomposition driver: the caller of the visibility and scoring engine.
//
// Restates the control flow of lib/src/process.cpp:28-347 over the C ABI of
// include/vacd.h:
//   * candidate planes: perpendicular bisectors of visibility edges drawn
//     with the global std::mt19937 and std::uniform_int_distribution<int>,
//     5-degree / 1e-3 duplicate rejection (process.cpp:28-72), then the
//     flat-surface planes (process.cpp:90-94);
//   * scores from vacd_score_planes (exact Q32.32), flat planes times
//     config.flat_surface_k, first maximum wins (process.cpp:101-134);
//   * clip, edge inheritance through the 1-based maps (process.cpp:136-159),
//     connected components, parts under 10 vertices dropped, visibility of
//     the new parts for pairs touching a cap vertex (process.cpp:164-178);
//   * part selection by concavity or by total edge length, stop below the
//     threshold, hulls, final concavity, return hulls or parts
//     (process.cpp:183-347).
// Differences, by design: the remesh and the cages are the engine's own
// signed-distance surfaces (OpenVDB is absent), an input is remeshed only
// when it is not closed (engine.remesh), and visibility runs in
// midpoint-parity mode unless engine.visibility = "cage"; candidate planes
// and concavity samples come from generators seeded per part, so a part
// splits the same way whenever it is split and its concavity is computed
// once and cached; engine.max_points bounds the points whose pairs are
// tested by a seeded subsample.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <exception>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "mesh.hpp"
#include "proposal.hpp"
#include "vacd.h"

namespace vhost {
namespace {

// start of the running process() call, for the verbose log's timestamps
std::chrono::steady_clock::time_point process_start = std::chrono::steady_clock::now();

void log_line(const std::string &msg) {
  if (!engine.verbose) return;
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - process_start).count();
  char stamp[32];
  std::snprintf(stamp, sizeof(stamp), "[visacd %8.3f s] ", t);
  std::cout << stamp << msg << "\n";
  std::cout.flush();
}

// Adds the lifetime of the object to a phase of last_profile.
struct PhaseTimer {
  const char *phase;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit PhaseTimer(const char *p) : phase(p) {}
  bool done = false;
  void stop() {
    if (done) return;
    done = true;
    last_profile.add(phase, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  ~PhaseTimer() { stop(); }
};

struct Engine {
  int device = 0;
  vacd_ctx *ctx = nullptr;
  vacd_mlp *mlp = nullptr;  // proposal model, loaded on first use
  // extra contexts (own streams) for BVH builds run from several host threads
  std::vector<vacd_ctx *> aux;
  explicit Engine(int dev) : device(dev) {
    if (vacd_ctx_create(device, &ctx) != VACD_OK)
      throw std::runtime_error(std::string("visacd: ") + vacd_last_error(nullptr));
    if (engine.front_max >= 0 &&
        vacd_ctx_set_option(ctx, VACD_OPT_FRONT_MAX, engine.front_max) != VACD_OK) {
      const std::string msg = vacd_last_error(ctx);
      vacd_ctx_destroy(ctx);
      throw std::runtime_error("visacd: " + msg);
    }
  }
  ~Engine() {
    vacd_mlp_destroy(mlp);
    for (vacd_ctx *c : aux) vacd_ctx_destroy(c);
    vacd_ctx_destroy(ctx);
  }
  // context i: 0 is ctx, the others are created on first use (call from one thread)
  vacd_ctx *context(size_t i) {
    if (i == 0) return ctx;
    while (aux.size() < i) {
      vacd_ctx *c = nullptr;
      if (vacd_ctx_create(device, &c) != VACD_OK)
        throw std::runtime_error(std::string("visacd: ") + vacd_last_error(nullptr));
      aux.push_back(c);
    }
    return aux[i - 1];
  }
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;
  void check(int rc) const {
    if (rc != VACD_OK) throw std::runtime_error(std::string("visacd: ") + vacd_last_error(ctx));
  }
  // the proposal model with the parameters of engine.proposal_weights
  vacd_mlp *proposal_model() {
    if (mlp) return mlp;
    const std::string &path = engine.proposal_weights;
    if (path.empty()) throw std::runtime_error("visacd: engine.proposal_weights is not set");
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("visacd: cannot open proposal weights " + path);
    const int64_t want = vacd_mlp_param_count();
    const std::streamoff bytes = f.tellg();
    if (bytes != (std::streamoff)(4 * want))
      throw std::runtime_error("visacd: " + path + " holds " + std::to_string(bytes) +
                               " bytes, the proposal model has " + std::to_string(4 * want));
    std::vector<float> w((size_t)want);
    f.seekg(0);
    f.read(reinterpret_cast<char *>(w.data()), bytes);
    if (!f) throw std::runtime_error("visacd: cannot read proposal weights " + path);
    check(vacd_mlp_create(ctx, 0, &mlp));
    check(vacd_mlp_set_params(ctx, mlp, w.data()));
    return mlp;
  }
};

std::vector<float> to_f32(const std::vector<Vec3> &v) {
  std::vector<float> f(3 * v.size());
  for (size_t i = 0; i < v.size(); ++i)
    for (int k = 0; k < 3; ++k) f[3 * i + k] = (float)v[i][k];
  return f;
}

std::vector<int> seeded_subsample(const std::vector<int> &from, size_t k, uint32_t seed) {
  if (from.size() <= k) return from;
  std::vector<int> v = from;
  std::mt19937 rng(seed);
  for (size_t i = 0; i < k; ++i) {
    std::uniform_int_distribution<size_t> d(i, v.size() - 1);
    std::swap(v[i], v[d(rng)]);
  }
  v.resize(k);
  std::sort(v.begin(), v.end());
  return v;
}

// Points of a part whose pairs are tested.  Without a limit: every vertex.
void choose_points(Mesh &part, bool inherited, uint32_t seed) {
  const int V = (int)part.vertices.size();
  const int limit = engine.max_points;
  std::vector<int> all(V);
  for (int i = 0; i < V; ++i) all[i] = i;
  if (limit <= 0 || V <= limit) {
    part.points = all;
    return;
  }
  if (!inherited) {
    part.points = seeded_subsample(all, (size_t)limit, seed);
    return;
  }
  std::vector<int> fresh;
  for (int i = 0; i < V; ++i)
    if (i < (int)part.is_new.size() && part.is_new[i]) fresh.push_back(i);
  const int keep = (int)part.points.size();
  const int budget = std::max(limit - keep, limit / 4);
  std::vector<int> add = seeded_subsample(fresh, (size_t)std::max(budget, 0), seed);
  std::vector<int> pts = part.points;
  pts.insert(pts.end(), add.begin(), add.end());
  std::sort(pts.begin(), pts.end());
  pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
  part.points = pts;
}

// Host staging buffer for visibility pairs, reused across calls (grown, never
// zero-filled again).
struct PairBuffer {
  std::unique_ptr<uint32_t[]> data;
  int64_t cap = 0;  // pairs
  uint32_t *get(int64_t pairs) {
    if (pairs > cap) {
      data.reset(new uint32_t[2 * (size_t)pairs]);
      cap = pairs;
    }
    return data.get();
  }
};

vacd_bvh *build_bvh(const Engine &E, const Mesh &m, bool planar = false);

// Planar groups of a part's BVH: the caps of its last 7 cuts, each the
// triangles with all three vertices on the cut plane (|n.v + d| <= 1e-9 |n|),
// kept when at least 16 triangles lie on it.  grp[t] = group of triangle t
// (-1 none); planes = (a, b, c, d) per group.
void cap_groups(const Mesh &m, std::vector<int8_t> &grp, std::vector<double> &planes) {
  const size_t nT = m.triangles.size();
  grp.assign(nT, -1);
  planes.clear();
  std::vector<size_t> mine;
  for (int c = (int)m.cut_planes.size() - 1; c >= 0 && planes.size() < 4 * 7; --c) {
    const Plane &q = m.cut_planes[c];
    const double l = std::sqrt(q.a * q.a + q.b * q.b + q.c * q.c);
    if (!(l > 0.0)) continue;
    const double tol = 1e-9 * l;
    mine.clear();
    for (size_t t = 0; t < nT; ++t) {
      if (grp[t] >= 0) continue;
      const auto &tr = m.triangles[t];
      if (std::fabs(q.eval(m.vertices[tr[0]])) <= tol && std::fabs(q.eval(m.vertices[tr[1]])) <= tol &&
          std::fabs(q.eval(m.vertices[tr[2]])) <= tol)
        mine.push_back(t);
    }
    if (mine.size() < 16) continue;
    const int8_t g = (int8_t)(planes.size() / 4);
    for (size_t t : mine) grp[t] = g;
    planes.insert(planes.end(), {q.a, q.b, q.c, q.d});
  }
}

// Visibility mask of a part's points (vacd_visibility): bit 0 = new vertex
// (the reference's mask); with engine.cap_rule (and planar caps), bit 1 + g
// = the vertex lies on the plane of the part's planar group g, so that pairs
// of two points on one cap plane whose segment runs into the cap at an end
// are decided by the engine's cap rule.
std::vector<uint8_t> point_mask(const Mesh &part) {
  const size_t n = part.points.size();
  std::vector<uint8_t> mask(n, 0);
  std::vector<int8_t> grp;
  std::vector<double> pl;
  if (engine.cap_rule && engine.planar_caps) cap_groups(part, grp, pl);
  const size_t G = pl.size() / 4;
  for (size_t i = 0; i < n; ++i) {
    const int v = part.points[i];
    const bool fresh = v < (int)part.is_new.size() && part.is_new[v];
    uint8_t m = fresh ? 1 : 0;
    for (size_t g = 0; g < G; ++g) {
      const Plane q(pl[4 * g], pl[4 * g + 1], pl[4 * g + 2], pl[4 * g + 3]);
      const double tol = 1e-9 * std::sqrt(q.a * q.a + q.b * q.b + q.c * q.c);
      if (std::fabs(q.eval(part.vertices[v])) <= tol) m |= (uint8_t)(2u << g);
    }
    mask[i] = m;
  }
  return mask;
}

// The surface at distance `level` outside a mesh on a 1 / scale grid
// (vacd_sdf_remesh; lib/src/preprocess.cpp:18-61).
Mesh sdf_remesh_mesh(vacd_ctx *ctx, const Mesh &m, double scale, double level) {
  auto check = [ctx](int rc) {
    if (rc != VACD_OK) throw std::runtime_error(std::string("visacd: ") + vacd_last_error(ctx));
  };
  const std::vector<float> V = to_f32(m.vertices);
  std::vector<uint32_t> T(3 * m.triangles.size());
  for (size_t t = 0; t < m.triangles.size(); ++t)
    for (int k = 0; k < 3; ++k) T[3 * t + k] = (uint32_t)m.triangles[t][k];
  int64_t vc = 4 * (int64_t)m.vertices.size() + 4096, tc = 8 * (int64_t)m.triangles.size() + 8192;
  std::vector<float> vo;
  std::vector<uint32_t> to;
  int64_t nv = 0, nt = 0;
  for (;;) {
    vo.resize(3 * (size_t)vc);
    to.resize(3 * (size_t)tc);
    const int rc = vacd_sdf_remesh(ctx, V.data(), (int64_t)m.vertices.size(), T.data(),
                                   (int64_t)m.triangles.size(), scale, level, vo.data(), vc,
                                   to.data(), tc, &nv, &nt);
    if (rc == VACD_ERR_CAPACITY) {
      vc = std::max(vc, nv), tc = std::max(tc, nt);
      continue;
    }
    check(rc);
    break;
  }
  Mesh out;
  out.vertices.resize((size_t)nv);
  for (int64_t v = 0; v < nv; ++v) out.vertices[v] = {vo[3 * v], vo[3 * v + 1], vo[3 * v + 2]};
  out.triangles.resize((size_t)nt);
  for (int64_t t = 0; t < nt; ++t) out.triangles[t] = {(int)to[3 * t], (int)to[3 * t + 1], (int)to[3 * t + 2]};
  return out;
}

// The cage a part's pairs are tested against in cage mode (engine.visibility
// = "cage"): its surface at distance `level` on a 1/40 grid
// (lib/src/process.cpp:170-171 per part with 0.02, 260-261 for the input
// with 0.03), as a BVH; nullptr in parity mode.
vacd_bvh *build_cage(vacd_ctx *ctx, const Mesh &src, double level) {
  if (engine.visibility != "cage") return nullptr;
  PhaseTimer t("visibility.cage");
  const Mesh cage = sdf_remesh_mesh(ctx, src, 40.0, level);
  if (cage.triangles.empty()) throw std::runtime_error("visacd: empty cage");
  const std::vector<float> V = to_f32(cage.vertices);
  std::vector<uint32_t> T(3 * cage.triangles.size());
  for (size_t k = 0; k < cage.triangles.size(); ++k)
    for (int c = 0; c < 3; ++c) T[3 * k + c] = (uint32_t)cage.triangles[k][c];
  vacd_bvh *b = nullptr;
  if (vacd_bvh_build(ctx, V.data(), (int64_t)cage.vertices.size(), T.data(), (int64_t)cage.triangles.size(),
                     &b) != VACD_OK)
    throw std::runtime_error(std::string("visacd: ") + vacd_last_error(ctx));
  return b;
}

// Visibility of a part in parity mode; with_mask restricts to pairs touching
// a new vertex (lib/ptx/ray_segments.cu:61-75).  Edges are appended in k order
// as pairs of part vertex ids.
void add_visibility(const Engine &E, Mesh &part, bool with_mask, PairBuffer &buf,
                    const vacd_bvh *cage = nullptr) {
  if (part.points.size() < 2) return;
  PhaseTimer timer("visibility");
  std::vector<float> V = to_f32(part.vertices);
  vacd_bvh *bvh = nullptr;
  {
    PhaseTimer t("visibility.bvh");
    bvh = build_bvh(E, part, true);
  }
  const int64_t n = (int64_t)part.points.size();
  last_profile.count("visibility.points", (long)(n));
  std::vector<float> P(3 * n);
  const std::vector<uint8_t> mask = point_mask(part);
  bool any_new = false;
  for (int64_t i = 0; i < n; ++i) {
    const int v = part.points[i];
    for (int k = 0; k < 3; ++k) P[3 * i + k] = V[3 * (size_t)v + k];
    any_new |= (mask[i] & 1) != 0;
  }
  const bool use_mask = with_mask;
  if (use_mask && !any_new) {
    vacd_bvh_destroy(bvh);
    return;
  }
  int64_t cap = std::max<int64_t>(buf.cap, std::min<int64_t>(n * (n - 1) / 2, int64_t(1) << 22));
  uint32_t *pairs = nullptr;
  int64_t count = 0;
  PhaseTimer call_timer("visibility.call");
  for (;;) {
    pairs = buf.get(std::max<int64_t>(cap, 1));
    const int rc = vacd_visibility(E.ctx, P.data(), n, use_mask ? mask.data() : nullptr, bvh,
                                   cage, pairs, cap, &count);
    if (rc == VACD_ERR_CAPACITY) {
      cap = count;
      continue;
    }
    if (rc != VACD_OK) {
      vacd_bvh_destroy(bvh);
      E.check(rc);
    }
    break;
  }
  vacd_bvh_destroy(bvh);
  vacd_vis_stats st;
  if (vacd_visibility_stats(E.ctx, &st) == VACD_OK) {
    last_profile.count("visibility.segments", (long)(st.segments));
    last_profile.count("visibility.masked", (long)(st.masked));
    last_profile.count("visibility.walked", (long)(st.walked));
    last_profile.count("visibility.exact_pairs", (long)(st.exact_calls));
  }
  PhaseTimer append_timer("visibility.append");
  last_profile.count("visibility.edges", (long)(count));
  part.intersecting_edges.reserve(part.intersecting_edges.size() + (size_t)count);
  for (int64_t e = 0; e < count; ++e)
    part.intersecting_edges.push_back(
        {(unsigned)part.points[pairs[2 * e]], (unsigned)part.points[pairs[2 * e + 1]]});
}

// BVH of a part.  planar: the caps of the part's cuts (triangles with all
// three vertices on one of its last 7 cut planes, at least 16 of them) are
// passed as planar groups, which the contact walks prune by plane distance.
vacd_bvh *build_bvh_on(vacd_ctx *ctx, const Mesh &m, bool planar);

vacd_bvh *build_bvh(const Engine &E, const Mesh &m, bool planar) { return build_bvh_on(E.ctx, m, planar); }

vacd_bvh *build_bvh_on(vacd_ctx *ctx, const Mesh &m, bool planar) {
  auto check = [ctx](int rc) {
    if (rc != VACD_OK) throw std::runtime_error(std::string("visacd: ") + vacd_last_error(ctx));
  };
  const std::vector<float> V = to_f32(m.vertices);
  const size_t nT = m.triangles.size();
  std::vector<uint32_t> T(3 * nT);
  for (size_t t = 0; t < nT; ++t)
    for (int k = 0; k < 3; ++k) T[3 * t + k] = (uint32_t)m.triangles[t][k];
  std::vector<int8_t> grp;
  std::vector<double> pl;
  if (planar && engine.planar_caps && !m.cut_planes.empty()) cap_groups(m, grp, pl);
  vacd_bvh *b = nullptr;
  if (!pl.empty()) {
    check(vacd_bvh_build_planar(ctx, V.data(), (int64_t)m.vertices.size(), T.data(), (int64_t)nT,
                                grp.data(), pl.data(), (int)(pl.size() / 4), &b));
    last_profile.count("visibility.planar_groups", (long)((long)(pl.size() / 4)));
  } else {
    check(vacd_bvh_build(ctx, V.data(), (int64_t)m.vertices.size(), T.data(), (int64_t)nT, &b));
  }
  return b;
}

// One point-to-mesh distance query in the layout of vacd_max_distance_batch.
struct DistQuery {
  std::vector<float> pts, V;
  std::vector<uint32_t> T;
  void set(const std::vector<Vec3> &samples, const Mesh &mesh) {
    pts = to_f32(samples);
    V = to_f32(mesh.vertices);
    T.resize(3 * mesh.triangles.size());
    for (size_t t = 0; t < mesh.triangles.size(); ++t)
      for (int k = 0; k < 3; ++k) T[3 * t + k] = (uint32_t)mesh.triangles[t][k];
  }
};

// Largest distance of each query's samples to its mesh, all queries in one
// batched GPU call (every sample against every triangle of its mesh).
std::vector<double> max_distances(const Engine &E, const std::vector<DistQuery> &qs) {
  std::vector<vacd_distance_query> q(qs.size());
  for (size_t k = 0; k < qs.size(); ++k)
    q[k] = vacd_distance_query{qs[k].pts.data(), (int64_t)(qs[k].pts.size() / 3), qs[k].V.data(),
                               (int64_t)(qs[k].V.size() / 3), qs[k].T.data(),
                               (int64_t)(qs[k].T.size() / 3)};
  std::vector<float> out(qs.size(), 0.0f);
  if (!q.empty()) E.check(vacd_max_distance_batch(E.ctx, q.data(), (int32_t)q.size(), out.data()));
  return std::vector<double>(out.begin(), out.end());
}

// Symmetric sampled surface distance with exact point-to-mesh distances
// measured on the GPU (same samples as the host form).
double gpu_hausdorff(const Engine &E, const Mesh &a, const Mesh &b, int resolution, uint32_t seed) {
  std::mt19937 rng(seed), rng_b(second_sample_seed(seed));
  const std::vector<Vec3> sa = surface_samples(a, resolution, rng);
  const std::vector<Vec3> sb = surface_samples(b, resolution, rng_b);
  if (sa.empty() || sb.empty()) return 1e300;
  std::vector<DistQuery> qs(2);
  qs[0].set(sa, b);
  qs[1].set(sb, a);
  const std::vector<double> d = max_distances(E, qs);
  return std::max(d[0], d[1]);
}

// Final concavity of a decomposition (lib/src/process.cpp:318-327): the
// largest symmetric sampled surface distance (10,000 samples) between a
// part and its hull.  Samples on all host threads, distances of every part
// in one batched call.
double final_concavity(const Engine &E, const MeshList &parts, const MeshList &hulls) {
  const size_t N = parts.size();
  std::vector<DistQuery> qs(2 * N);
  std::vector<double> host(N, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (long i = 0; i < (long)N; ++i) {
    if (!engine.gpu_distance || hulls[i].triangles.empty()) {
      host[i] = hausdorff_boundary(parts[i], hulls[i], 10000, engine.sample_seed);
      continue;
    }
    std::mt19937 rng(engine.sample_seed), rng_b(second_sample_seed(engine.sample_seed));
    const std::vector<Vec3> sa = surface_samples(parts[i], 10000, rng);
    const std::vector<Vec3> sb = surface_samples(hulls[i], 10000, rng_b);
    if (sa.empty() || sb.empty()) {
      host[i] = 1e300;
      continue;
    }
    qs[2 * i].set(sa, hulls[i]);
    qs[2 * i + 1].set(sb, parts[i]);
  }
  const std::vector<double> d = max_distances(E, qs);
  double h = 0.0;
  for (size_t i = 0; i < N; ++i) h = std::max(h, std::max(host[i], std::max(d[2 * i], d[2 * i + 1])));
  return h;
}

double part_concavity(const Engine &E, const Mesh &part) {
  PhaseTimer timer("concavity");
  const Mesh hull = convex_hull(part.vertices);
  if (!engine.gpu_distance || hull.triangles.empty())
    return concavity(part, hull, engine.hausdorff_resolution, engine.sample_seed);
  const double rv = std::cbrt(3.0 * std::fabs(mesh_volume(part) - mesh_volume(hull)) / (4.0 * M_PI));
  return std::max(0.3 * rv, gpu_hausdorff(E, part, hull, engine.hausdorff_resolution, engine.sample_seed));
}

std::vector<Plane> bisector_planes(const Mesh &m, int num_planes, std::mt19937 &rng) {
  std::vector<Plane> planes;
  const auto &E = m.intersecting_edges;
  if (E.empty()) return planes;
  std::uniform_int_distribution<> dis(0, (int)E.size() - 1);
  const double cos_tol = std::cos(5.0 * M_PI / 180.0), dist_tol = 1e-3;
  for (int i = 0; i < num_planes * 5 && (int)planes.size() < num_planes; ++i) {
    const auto &e = E[dis(rng)];
    const Vec3 &p1 = m.vertices[e.first], &p2 = m.vertices[e.second];
    Vec3 n = p2 - p1;
    const double l = norm(n);
    if (l < 1e-6) continue;
    n = Vec3{n[0] / l, n[1] / l, n[2] / l};  // normalize_vector (lib/include/core.hpp:102-108)
    const Vec3 mid = (p1 + p2) * 0.5;
    const double d = -dot(n, mid);
    bool similar = false;
    for (const Plane &q : planes)
      if (std::fabs(n[0] * q.a + n[1] * q.b + n[2] * q.c) > cos_tol && std::fabs(d - q.d) < dist_tol) {
        similar = true;
        break;
      }
    if (!similar) planes.emplace_back(n[0], n[1], n[2], d);
  }
  return planes;
}

// Normals uniform on S^2 and offsets uniform over the part's extent along
// each normal (the benchmark workloads' candidate sets).  The draws are taken
// in order (normal x, y, z, then the offset fraction, per plane); the
// extents of every part of a round come from one GPU call
// (vacd_plane_extents_batch, float64 rounded as a host loop would).
struct SphereDraws {
  std::vector<double> normals;  // [P][3]
  std::vector<double> frac;
};

SphereDraws sphere_draws(int num_planes, std::mt19937 &rng) {
  std::normal_distribution<double> g(0.0, 1.0);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  SphereDraws d;
  for (int i = 0; i < num_planes; ++i) {
    Vec3 n{g(rng), g(rng), g(rng)};
    const double l = norm(n);
    if (l == 0) continue;
    n = n * (1.0 / l);
    d.normals.insert(d.normals.end(), {n[0], n[1], n[2]});
    d.frac.push_back(u(rng));
  }
  return d;
}

std::vector<std::vector<Plane>> sphere_planes(const Engine &E, const std::vector<const Mesh *> &ms,
                                              const std::vector<SphereDraws> &dr) {
  const size_t F = ms.size();
  std::vector<std::vector<double>> pts(F);
  std::vector<vacd_extent_query> q(F);
  int64_t total = 0;
  for (size_t f = 0; f < F; ++f) {
    const Mesh &m = *ms[f];
    pts[f].resize(3 * m.points.size());
    for (size_t i = 0; i < m.points.size(); ++i)
      for (int k = 0; k < 3; ++k) pts[f][3 * i + k] = m.vertices[m.points[i]][k];
    q[f] = vacd_extent_query{dr[f].normals.data(), (int64_t)dr[f].frac.size(), pts[f].data(),
                             (int64_t)m.points.size()};
    total += q[f].P;
  }
  std::vector<double> lh(2 * (size_t)total);
  E.check(vacd_plane_extents_batch(E.ctx, q.data(), (int32_t)F, lh.data()));
  std::vector<std::vector<Plane>> planes(F);
  size_t o = 0;
  for (size_t f = 0; f < F; ++f) {
    for (size_t i = 0; i < dr[f].frac.size(); ++i, ++o) {
      const double lo = lh[2 * o], hi = lh[2 * o + 1];
      const double *n = &dr[f].normals[3 * i];
      planes[f].emplace_back(n[0], n[1], n[2], -(lo + dr[f].frac[i] * (hi - lo)));
    }
  }
  return planes;
}

double edge_score(const Mesh &m) {
  double s = 0;
  for (auto &e : m.intersecting_edges) s += norm(m.vertices[e.second] - m.vertices[e.first]);
  return s;
}


// ---------------------------------------------------------------- batched
// Visibility of many parts in one batched call (masked: pairs touching a
// vertex of the last cut).  Edges are appended to each part in k order.
// Verbose diagnostic: which part of a batch most of the exactly redone pairs
// came from, with its size and cut planes.
void log_exact_hotspot(const Engine &E, const std::vector<Mesh *> &parts,
                       const std::vector<size_t> &idx, const std::vector<vacd_part> &vp,
                       const vacd_vis_stats &st) {
  std::vector<int64_t> ids((size_t)st.exact_calls);
  int64_t cnt = 0;
  if (vacd_visibility_deferred(E.ctx, ids.data(), (int64_t)ids.size(), &cnt) != VACD_OK) return;
  std::vector<int64_t> first(vp.size() + 1, 0);  // first pair id of each part
  for (size_t j = 0; j < vp.size(); ++j)
    first[j + 1] = first[j] + 64 * ((vp[j].n * (vp[j].n - 1) / 2 + 63) / 64);
  std::vector<int64_t> per(vp.size(), 0);
  for (int64_t t = 0; t < std::min<int64_t>(cnt, (int64_t)ids.size()); ++t) {
    const size_t j = (size_t)(std::upper_bound(first.begin(), first.end(), ids[t]) - first.begin()) - 1;
    if (j < per.size()) ++per[j];
  }
  const size_t j = (size_t)(std::max_element(per.begin(), per.end()) - per.begin());
  const Mesh &m = *parts[idx[j]];
  std::ostringstream o;
  o << "visibility: " << st.exact_calls << " pairs redone exactly; " << per[j] << " in a part of "
    << m.vertices.size() << " vertices, " << m.triangles.size() << " triangles, " << vp[j].n
    << " points; cut planes:";
  o.precision(9);
  for (const Plane &q : m.cut_planes) o << " (" << q.a << ", " << q.b << ", " << q.c << ", " << q.d << ")";
  log_line(o.str());
}

void add_visibility_batch(Engine &E, const std::vector<Mesh *> &parts, bool with_mask,
                          PairBuffer &buf, const Mesh *cage_src = nullptr, double cage_level = 0.02) {
  PhaseTimer timer("visibility");
  const size_t np = parts.size();
  std::vector<vacd_bvh *> bvhs(np, nullptr), cages(np, nullptr);
  std::vector<std::vector<float>> P(np);
  std::vector<std::vector<uint8_t>> M(np);
  std::vector<vacd_part> vp;
  std::vector<size_t> idx;
  auto release = [&] {
    for (auto *b : bvhs) vacd_bvh_destroy(b);
    for (auto *b : cages) vacd_bvh_destroy(b);
  };
  try {
    {  // points and masks of every part, on all host threads
      PhaseTimer t("visibility.masks");
      std::vector<char> keep(np, 0);
#pragma omp parallel for schedule(dynamic)
      for (long k = 0; k < (long)np; ++k) {
        Mesh &part = *parts[k];
        const int64_t n = (int64_t)part.points.size();
        if (n < 2) continue;
        P[k].resize(3 * n);
        M[k] = point_mask(part);
        bool any_new = false;
        for (int64_t i = 0; i < n; ++i) {
          const int v = part.points[i];
          for (int c = 0; c < 3; ++c) P[k][3 * i + c] = (float)part.vertices[v][c];
          any_new |= (M[k][i] & 1) != 0;
        }
        keep[k] = !(with_mask && !any_new);
      }
      for (size_t k = 0; k < np; ++k)
        if (keep[k]) {
          idx.push_back(k);
          last_profile.count("visibility.points", (long)parts[k]->points.size());
        }
    }
    {  // the BVHs, from up to four host threads with a context each
      PhaseTimer t("visibility.bvh");
      const size_t W = std::min<size_t>(4, idx.size());
      std::vector<vacd_ctx *> cs(W);
      for (size_t w = 0; w < W; ++w) cs[w] = E.context(w);
      std::atomic<size_t> next{0};
      std::vector<std::exception_ptr> errs(W);
      auto work = [&](size_t w) {
        try {
          for (size_t j = next++; j < idx.size(); j = next++) {
            bvhs[idx[j]] = build_bvh_on(cs[w], *parts[idx[j]], true);
            cages[idx[j]] = build_cage(cs[w], cage_src ? *cage_src : *parts[idx[j]], cage_level);
          }
        } catch (...) {
          errs[w] = std::current_exception();
        }
      };
      std::vector<std::thread> pool;
      for (size_t w = 1; w < W; ++w) pool.emplace_back(work, w);
      if (W > 0) work(0);
      for (auto &th : pool) th.join();
      for (auto &e : errs)
        if (e) std::rethrow_exception(e);
    }
    for (size_t k : idx)
      vp.push_back(vacd_part{P[k].data(), (int64_t)parts[k]->points.size(),
                             with_mask ? M[k].data() : nullptr, bvhs[k], cages[k]});
    if (!vp.empty()) {
      std::vector<int64_t> off(vp.size() + 1, 0);
      int64_t cap = std::max<int64_t>(buf.cap, int64_t(1) << 22), count = 0;
      uint32_t *pairs = nullptr;
      PhaseTimer call_timer("visibility.call");
      for (;;) {
        pairs = buf.get(cap);
        const int rc = vacd_visibility_batch(E.ctx, vp.data(), (int32_t)vp.size(), pairs, cap,
                                             &count, off.data());
        if (rc == VACD_ERR_CAPACITY) {
          cap = count;
          continue;
        }
        E.check(rc);
        break;
      }
      call_timer.stop();
      vacd_vis_stats st;
      if (vacd_visibility_stats(E.ctx, &st) == VACD_OK) {
        last_profile.count("visibility.segments", (long)(st.segments));
        last_profile.count("visibility.walked", (long)(st.walked));
        last_profile.count("visibility.exact_pairs", (long)(st.exact_calls));
        if (engine.verbose && st.exact_calls > 100000) log_exact_hotspot(E, parts, idx, vp, st);
      }
      last_profile.count("visibility.edges", (long)(count));
      PhaseTimer append_timer("visibility.append");
#pragma omp parallel for schedule(dynamic)
      for (long j = 0; j < (long)idx.size(); ++j) {
        Mesh &part = *parts[idx[j]];
        part.intersecting_edges.reserve(part.intersecting_edges.size() + (size_t)(off[j + 1] - off[j]));
        for (int64_t e = off[j]; e < off[j + 1]; ++e)
          part.intersecting_edges.push_back({(unsigned)part.points[pairs[2 * e]],
                                             (unsigned)part.points[pairs[2 * e + 1]]});
      }
    }
  } catch (...) {
    release();
    throw;
  }
  release();
}

// Host half of the concavity of a list of parts: hulls, volumes, surface
// samples and the two distance queries of each part, on all host threads.
struct ConcavityPrep {
  std::vector<double> rv, host_value;
  std::vector<char> gpu;  // 1: distances from the queries, 2: no samples
  std::vector<DistQuery> qs;  // 2 per part (empty for host-measured parts)
};

ConcavityPrep concavity_prep(const std::vector<const Mesh *> &ms) {
  PhaseTimer host_timer("concavity.host");
  ConcavityPrep cp;
  const size_t J = ms.size();
  cp.rv.assign(J, 0.0);
  cp.host_value.assign(J, 0.0);
  cp.gpu.assign(J, 0);
  cp.qs.resize(2 * J);
#pragma omp parallel for schedule(dynamic)
  for (long j = 0; j < (long)J; ++j) {
    const Mesh &part = *ms[j];
    // a large part's samples are drawn on a thread of their own while its
    // hull is built (both take tens of ms on a 100k-vertex part)
    std::vector<Vec3> sa;
    const bool early = engine.gpu_distance && part.vertices.size() > 20000;
    std::thread sampler;
    std::exception_ptr sample_err;
    if (early)
      sampler = std::thread([&] {
        try {
          std::mt19937 rng(engine.sample_seed);
          sa = surface_samples(part, engine.hausdorff_resolution, rng);
        } catch (...) {
          sample_err = std::current_exception();
        }
      });
    Mesh hull;
    try {
      hull = convex_hull(part.vertices);
    } catch (...) {
      if (sampler.joinable()) sampler.join();
      throw;
    }
    if (sampler.joinable()) sampler.join();
    if (sample_err) std::rethrow_exception(sample_err);
    if (!engine.gpu_distance || hull.triangles.empty()) {
      cp.host_value[j] = concavity(part, hull, engine.hausdorff_resolution, engine.sample_seed);
      continue;
    }
    cp.rv[j] = std::cbrt(3.0 * std::fabs(mesh_volume(part) - mesh_volume(hull)) / (4.0 * M_PI));
    if (!early) {
      std::mt19937 rng(engine.sample_seed);
      sa = surface_samples(part, engine.hausdorff_resolution, rng);
    }
    std::mt19937 rng_b(second_sample_seed(engine.sample_seed));
    const std::vector<Vec3> sb = surface_samples(hull, engine.hausdorff_resolution, rng_b);
    if (sa.empty() || sb.empty()) {
      cp.gpu[j] = 2;
      continue;
    }
    cp.gpu[j] = 1;
    cp.qs[2 * j].set(sa, hull);
    cp.qs[2 * j + 1].set(sb, part);
  }
  return cp;
}

// GPU half: every part's two one-sided distances in one batched call.
std::vector<double> concavity_finish(const Engine &E, const ConcavityPrep &cp) {
  PhaseTimer gpu_timer("concavity.distance");
  const std::vector<double> d = max_distances(E, cp.qs);
  std::vector<double> out(cp.gpu.size());
  for (size_t j = 0; j < cp.gpu.size(); ++j) {
    if (cp.gpu[j] == 0)
      out[j] = cp.host_value[j];
    else if (cp.gpu[j] == 2)
      out[j] = std::max(0.3 * cp.rv[j], 1e300);
    else
      out[j] = std::max(0.3 * cp.rv[j], std::max(d[2 * j], d[2 * j + 1]));
  }
  return out;
}

// ------------------------------------------------------- plane selection
// Scoring inputs of a part, converted once per selection.
struct ScoreInput {
  std::vector<float> pts;
  std::vector<uint32_t> ed;
  void fill(const Mesh &m) {
    if (!pts.empty() || m.intersecting_edges.empty()) return;
    pts = to_f32(m.vertices);
    ed.resize(2 * m.intersecting_edges.size());
    for (size_t e = 0; e < m.intersecting_edges.size(); ++e)
      ed[2 * e] = m.intersecting_edges[e].first, ed[2 * e + 1] = m.intersecting_edges[e].second;
  }
};

// Exact Q32.32 scores of the listed planes of the listed parts in one batched
// call: q[f][j] is the score of planes[f][sub[f][j]].
void score_subsets(const Engine &E, const std::vector<const Mesh *> &ms,
                   const std::vector<std::vector<Plane>> &planes,
                   const std::vector<std::vector<int>> &sub, const std::vector<size_t> &parts,
                   std::vector<ScoreInput> &in, std::vector<std::vector<int64_t>> &q) {
  std::vector<std::vector<float>> hp(parts.size());
  std::vector<vacd_score_part> sp;
  int64_t total = 0;
  for (size_t j = 0; j < parts.size(); ++j) {
    const size_t f = parts[j];
    in[f].fill(*ms[f]);
    const std::vector<int> &idx = sub[f];
    hp[j].resize(4 * idx.size());
    for (size_t i = 0; i < idx.size(); ++i) {
      const Plane &p = planes[f][idx[i]];
      hp[j][4 * i] = (float)p.a, hp[j][4 * i + 1] = (float)p.b;
      hp[j][4 * i + 2] = (float)p.c, hp[j][4 * i + 3] = (float)p.d;
    }
    sp.push_back(vacd_score_part{hp[j].data(), (int64_t)idx.size(), in[f].pts.data(),
                                 (int64_t)ms[f]->vertices.size(), in[f].ed.data(),
                                 (int64_t)ms[f]->intersecting_edges.size()});
    total += (int64_t)idx.size();
  }
  if (sp.empty()) return;
  std::vector<int64_t> q32(total);
  std::vector<int32_t> am(sp.size());
  E.check(vacd_score_planes_batch(E.ctx, sp.data(), (int)sp.size(), q32.data(), am.data()));
  int64_t off = 0;
  for (size_t j = 0; j < parts.size(); ++j) {
    q[parts[j]].assign(q32.begin() + off, q32.begin() + off + sp[j].P);
    off += sp[j].P;
  }
  last_profile.count("scoring.planes", (long)(total));
}

// First maximum of the scores in plane order; planes from ss_offset on (the
// flat-surface planes) count config.flat_surface_k times
// (lib/src/process.cpp:128-134).
int first_max(const std::vector<int> &idx, const std::vector<int64_t> &q, size_t ss_offset) {
  int best = -1;
  double best_s = -1.0;
  for (size_t j = 0; j < idx.size(); ++j) {
    double sc = (double)q[j] * 0x1p-32;
    if ((size_t)idx[j] >= ss_offset) sc *= config.flat_surface_k;
    if (sc > best_s) best_s = sc, best = idx[j];
  }
  return best;
}

// The plane each part is cut with (index into planes[f], or -1 when the part
// has no planes or no edges).  engine.proposal "exact" scores every plane;
// "learned" scores the model's top-ranked planes, an audit sample and the
// flat-surface planes, and falls back to every plane when the best of the top
// is under engine.proposal_audit_ratio times the best of the audit.
// seeds[f] seeds the part's edge sample and audit draw.
std::vector<int> choose_planes(Engine &E, const std::vector<const Mesh *> &ms,
                               const std::vector<std::vector<Plane>> &planes,
                               const std::vector<size_t> &ss_offset,
                               const std::vector<uint64_t> &seeds) {
  const size_t F = ms.size();
  std::vector<int> best(F, -1);
  std::vector<ScoreInput> in(F);
  std::vector<std::vector<int>> sub(F);
  std::vector<std::vector<int64_t>> q(F);
  std::vector<size_t> live, ranked;
  const bool learned = engine.proposal == "learned";
  if (!learned && engine.proposal != "exact")
    throw std::invalid_argument("visacd: engine.proposal must be \"exact\" or \"learned\"");
  const size_t T = (size_t)std::max(1, engine.proposal_top);
  const size_t A = (size_t)std::max(1, engine.proposal_audit);
  for (size_t f = 0; f < F; ++f) {
    if (planes[f].empty() || ms[f]->intersecting_edges.empty()) continue;
    live.push_back(f);
    if (learned && ss_offset[f] > T + A) {
      ranked.push_back(f);
    } else {
      sub[f].resize(planes[f].size());
      std::iota(sub[f].begin(), sub[f].end(), 0);
    }
  }
  std::vector<std::vector<uint8_t>> role(F);  // 1 = top-ranked, 2 = audit
  if (!ranked.empty()) {
    PhaseTimer t("proposal");
    size_t Pmax = 0;
    for (size_t f : ranked) Pmax = std::max(Pmax, ss_offset[f]);
    const size_t chunk = std::max<size_t>(1, (size_t(1) << 21) / Pmax);
    std::vector<float> feat, pin, pred;
    for (size_t c0 = 0; c0 < ranked.size(); c0 += chunk) {
      const size_t B = std::min(chunk, ranked.size() - c0);
      feat.assign(B * kProposalEdges * kEdgeFeatures, 0.0f);
      pin.assign(B * Pmax * 4, 0.0f);
      pred.resize(B * Pmax);
#pragma omp parallel for schedule(dynamic)
      for (long b = 0; b < (long)B; ++b) {
        const size_t f = ranked[c0 + b];
        const PartFrame fr = part_frame(ms[f]->vertices);
        edge_features(ms[f]->vertices, ms[f]->intersecting_edges, mix64(seeds[f] ^ 0xfea7ull), fr,
                      feat.data() + b * kProposalEdges * kEdgeFeatures);
        float *pq = pin.data() + b * Pmax * 4;
        plane_inputs(planes[f], ss_offset[f], fr, pq);
        for (size_t i = ss_offset[f]; i < Pmax; ++i)  // padding rows: copies of plane 0
          for (int k = 0; k < 4; ++k) pq[4 * i + k] = pq[k];
      }
      E.check(vacd_mlp_forward(E.ctx, E.proposal_model(), (int)B, kProposalEdges, feat.data(),
                               (int)Pmax, pin.data(), pred.data()));
#pragma omp parallel for schedule(dynamic)
      for (long b = 0; b < (long)B; ++b) {
        const size_t f = ranked[c0 + b], n = ss_offset[f];
        std::vector<float> y(pred.begin() + b * Pmax, pred.begin() + b * Pmax + n);
        for (float &v : y)
          if (!(v == v)) v = -std::numeric_limits<float>::infinity();
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::partial_sort(order.begin(), order.begin() + T, order.end(),
                          [&](int i, int j) { return y[i] > y[j] || (y[i] == y[j] && i < j); });
        std::vector<uint8_t> &r = role[f];
        r.assign(planes[f].size(), 0);
        for (size_t i = 0; i < T; ++i) r[order[i]] |= 1;
        std::mt19937_64 rng(mix64(seeds[f] ^ 0xa0d17ull));
        std::iota(order.begin(), order.end(), 0);
        for (size_t i = 0; i < A; ++i) {  // partial Fisher-Yates: A distinct planes
          std::uniform_int_distribution<size_t> d(i, n - 1);
          std::swap(order[i], order[d(rng)]);
          r[order[i]] |= 2;
        }
        for (size_t i = 0; i < planes[f].size(); ++i)
          if (r[i] || i >= n) sub[f].push_back((int)i);
      }
    }
    last_profile.count("proposal.parts", (long)((long)ranked.size()));
    for (size_t f : ranked) last_profile.count("proposal.ranked_planes", (long)((long)ss_offset[f]));
  }
  {
    PhaseTimer t("scoring");
    score_subsets(E, ms, planes, sub, live, in, q);
    std::vector<size_t> redo;
    for (size_t f : ranked) {
      double top = -1.0, audit = -1.0;
      for (size_t j = 0; j < sub[f].size(); ++j) {
        const int i = sub[f][j];
        if ((size_t)i >= ss_offset[f]) continue;
        const double sc = (double)q[f][j];
        if (role[f][i] & 1) top = std::max(top, sc);
        if (role[f][i] & 2) audit = std::max(audit, sc);
      }
      if (top < engine.proposal_audit_ratio * audit) redo.push_back(f);
    }
    if (!redo.empty()) {  // full sweep of the parts whose ranking missed
      for (size_t f : redo) {
        sub[f].resize(planes[f].size());
        std::iota(sub[f].begin(), sub[f].end(), 0);
      }
      score_subsets(E, ms, planes, sub, redo, in, q);
      last_profile.count("proposal.fallbacks", (long)((long)redo.size()));
    }
  }
  for (size_t f : live) best[f] = first_max(sub[f], q[f], ss_offset[f]);
  return best;
}

// ------------------------------------------------------------ part lineage
// Every part carries a 64-bit seed.  The parts of the input take
// mix64(root + i) (root: two draws of the global generator, so set_seed
// fixes a run); the k-th piece of a split (1-based over the pieces of both
// halves, dropped ones included) takes kid_seed(parent, k).  A part's
// candidate planes, point subsample and proposal draws come from its own
// seed, so a part splits the same way whenever it is split: the plain loop
// and the batched frontier, which splits parts in another order, make the
// same parts.
uint64_t kid_seed(uint64_t parent, uint32_t k) { return mix64(parent * 31 + k); }
uint32_t point_seed(uint64_t s) { return (uint32_t)mix64(s ^ 0x9017ull); }
// generator of the candidate planes of a part's attempt-th split (a failed
// clip is retried with fresh planes, as the reference's next step draws
// fresh ones from its global generator)
uint32_t plane_seed(uint64_t s, int attempt) { return (uint32_t)mix64(s ^ (0x5eedull + (uint64_t)attempt)); }

struct Driver {
  Engine E;
  std::vector<Plane> flat;
  PairBuffer pair_buf;
  explicit Driver(int device) : E(device) {}

  // Candidate planes of each part (its own generator), then the flat-surface
  // planes (lib/src/process.cpp:87-94); ss_offset[f] = first flat plane.
  std::vector<std::vector<Plane>> candidates(const std::vector<const Mesh *> &ms,
                                             const std::vector<uint64_t> &seeds,
                                             const std::vector<int> &attempts,
                                             std::vector<size_t> &ss_offset) {
    PhaseTimer t("planes");
    const size_t F = ms.size();
    std::vector<std::vector<Plane>> planes(F);
    std::vector<SphereDraws> dr(F);
#pragma omp parallel for schedule(dynamic)
    for (long f = 0; f < (long)F; ++f) {
      std::mt19937 rng(plane_seed(seeds[f], attempts[f]));
      if (engine.planes == "sphere") dr[f] = sphere_draws(engine.num_planes, rng);
      else planes[f] = bisector_planes(*ms[f], engine.num_planes, rng);
    }
    if (engine.planes == "sphere" && F) planes = sphere_planes(E, ms, dr);
    ss_offset.assign(F, 0);
    for (size_t f = 0; f < F; ++f) {
      ss_offset[f] = planes[f].size();
      planes[f].insert(planes[f].end(), flat.begin(), flat.end());
    }
    return planes;
  }
};

// The parent's edges whose both ends a half keeps, renumbered by the half's
// 1-based map, in the parent's order.  Long lists are split over threads:
// each chunk counts its kept edges, then writes them at its offset.
void inherit_edges(const std::vector<std::pair<unsigned, unsigned>> &src, const std::vector<int> &map,
                   std::vector<std::pair<unsigned, unsigned>> &dst) {
  const size_t E = src.size();
  const int nth = E >= 262144 ? 4 : 1;
  std::vector<size_t> cnt(nth + 1, 0);
  auto chunk = [&](int c, bool fill) {
    const size_t lo = E * c / nth, hi = E * (c + 1) / nth;
    size_t o = fill ? cnt[c] : 0;
    for (size_t i = lo; i < hi; ++i) {
      const int a = map[src[i].first], b = map[src[i].second];
      if (a && b) {
        if (fill) dst[o] = {(unsigned)a - 1, (unsigned)b - 1};
        ++o;
      }
    }
    if (!fill) cnt[c + 1] = o;
  };
  auto run = [&](bool fill) {
    std::vector<std::thread> th;
    for (int c = 1; c < nth; ++c) th.emplace_back(chunk, c, fill);
    chunk(0, fill);
    for (auto &t : th) t.join();
  };
  run(false);  // cnt[c + 1] = kept edges of chunk c
  for (int c = 0; c < nth; ++c) cnt[c + 1] += cnt[c];
  dst.resize(cnt[nth]);
  run(true);
}

// Clips a part by a plane (lib/src/process.cpp:136-168): the pieces of both
// halves with at least 10 vertices, each with the parent's edges whose both
// ends it keeps (1-based maps), its points, cut planes and seed.  False when
// the clip fails.  Pieces are numbered over the + half, then the - half; a
// large part's two halves are prepared on two threads.
bool split_part(const Mesh &mesh, uint64_t seed, const Plane &plane, MeshList &kids,
                std::vector<uint64_t> &kid_seeds) {
  std::vector<int> mp, mn;
  MeshList halves = clip(mesh, plane, mp, mn);
  if (halves.size() < 2) return false;
  MeshList pieces[2];
  auto half = [&](int w) {
    const std::vector<int> &map = w == 0 ? mp : mn;
    inherit_edges(mesh.intersecting_edges, map, halves[w].intersecting_edges);
    for (int v : mesh.points)
      if (map[v]) halves[w].points.push_back(map[v] - 1);
    pieces[w] = separate_disjoint(std::move(halves[w]));
  };
  if (mesh.triangles.size() >= 20000) {
    std::exception_ptr err;
    std::thread other([&] {
      try {
        half(1);
      } catch (...) {
        err = std::current_exception();
      }
    });
    try {
      half(0);
    } catch (...) {
      other.join();
      throw;
    }
    other.join();
    if (err) std::rethrow_exception(err);
  } else {
    half(0);
    half(1);
  }
  uint32_t k = 0;
  for (auto &list : pieces)
    for (Mesh &piece : list) {
      ++k;
      if (piece.vertices.size() < 10) continue;
      piece.cut_planes = mesh.cut_planes;
      piece.cut_planes.push_back(plane);
      const uint64_t ks = kid_seed(seed, k);
      choose_points(piece, true, point_seed(ks));
      kids.push_back(std::move(piece));
      kid_seeds.push_back(ks);
    }
  return true;
}

// Index of the part the loop splits next (lib/src/process.cpp:286-299): the
// first maximum of the concavity (or of the total edge length); -1 for none.
int pick_part(const std::vector<double> &key) {
  int idx = -1;
  double best = -1.0;
  for (int i = 0; i < (int)key.size(); ++i)
    if (key[i] > best) best = key[i], idx = i;
  return idx;
}

ProcessResult finish(Driver &D, MeshList parts, const std::vector<double> &bbox) {
  log_line("Computing convex hulls for " + std::to_string(parts.size()) + " parts...");
  PhaseTimer final_timer("final_hulls");
  MeshList hulls(parts.size());
  PhaseTimer fh_timer("final_hulls.hulls");
#pragma omp parallel for schedule(dynamic)
  for (long i = 0; i < (long)parts.size(); ++i) hulls[i] = convex_hull(parts[i].vertices);
  fh_timer.stop();
  PhaseTimer fd_timer("final_hulls.distance");
  const double final_h = final_concavity(D.E, parts, hulls);
  fd_timer.stop();
  if (config.use_merging) log_line("Hull merging is not part of this engine; parts are kept.");
  for (auto &h : hulls) unnormalize(h, bbox);
  for (auto &p : parts) unnormalize(p, bbox);
  ProcessResult r;
  r.parts = config.return_parts ? parts : hulls;
  r.concavity = final_h;
  r.num_parts = (int)r.parts.size();
  std::ostringstream o;
  o.setf(std::ios::fixed);
  o.precision(4);
  o << "Done. parts=" << r.num_parts << "  concavity=" << final_h;
  log_line(o.str());
  return r;
}

// Whether every undirected edge (by vertex index) has an even number of
// triangles: a closed mesh, whose inside midpoint parity decides.
// Every undirected edge in an even number of triangles (mod 2 closedness):
// each edge key is hashed by two independent 64-bit mixers and all hashes
// are XOR-ed, so edges met an even number of times cancel; an open mesh
// reads as closed only if its odd edges' hashes XOR to zero in both halves
// (probability 2^-128 for random mixers).  One pass on all host threads; a
// sort of the 983,040 edge keys of a 327,680-triangle mesh took ~50 ms.
uint64_t edge_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

bool closed_mesh(const Mesh &m) {
  uint64_t h1 = 0, h2 = 0;
  const long nT = (long)m.triangles.size();
#pragma omp parallel for reduction(^ : h1, h2) schedule(static)
  for (long t = 0; t < nT; ++t)
    for (int k = 0; k < 3; ++k) {
      const uint64_t u = (uint32_t)m.triangles[t][k], v = (uint32_t)m.triangles[t][(k + 1) % 3];
      if (u == v) continue;
      const uint64_t key = (std::min(u, v) << 32) | std::max(u, v);
      h1 ^= edge_mix(key + 0x9E3779B97F4A7C15ull);
      h2 ^= edge_mix((key ^ 0xD6E8FEB86659FD93ull) * 0xA0761D6478BD642Full + 0x2545F4914F6CDD1Dull);
    }
  return h1 == 0 && h2 == 0;
}

// The input remesh (lib/src/process.cpp:250-258): with engine.remesh
// "always" as the reference does (scale 30, level 0.55 / 30; scale 20 when
// that gives more than 15,000 vertices), with "auto" only for inputs that
// are not closed, with "never" not at all.  The mesh is in the normalised
// frame.
void maybe_remesh(Driver &D, Mesh &mesh) {
  const std::string &r = engine.remesh;
  if (r != "always" && r != "auto" && r != "never")
    throw std::invalid_argument("visacd: engine.remesh must be \"auto\", \"always\" or \"never\"");
  if (r == "never" || (r == "auto" && closed_mesh(mesh))) return;
  PhaseTimer t("remesh");
  Mesh out = sdf_remesh_mesh(D.E.ctx, mesh, 30.0, 0.55 / 30.0);
  if (out.vertices.size() > 15000) out = sdf_remesh_mesh(D.E.ctx, mesh, 20.0, 0.55 / 20.0);
  log_line("Remeshed to " + std::to_string(out.vertices.size()) + " verts.");
  mesh = std::move(out);
}

// The reference's loop (lib/src/process.cpp:238-347): one part per step,
// its planes scored, clipped, and the new parts' visibility computed at once.
ProcessResult process_plain(Driver &D, Mesh mesh, double concavity_threshold, int num_parts,
                            uint64_t root) {
  const std::vector<double> bbox = normalize(mesh);
  log_line("Preprocessing mesh (" + std::to_string(mesh.vertices.size()) + " verts)...");
  mesh.is_new.clear();
  mesh.intersecting_edges.clear();
  mesh.points.clear();
  maybe_remesh(D, mesh);
  if (config.use_flat_surfaces) {
    D.flat = flat_surface_planes(mesh, config.flat_surface_min_area);
    log_line("Detected " + std::to_string(D.flat.size()) + " flat surface(s).");
  }
  MeshList parts = separate_disjoint(mesh);
  std::vector<uint64_t> seed;
  std::vector<int> attempt;
  {
    vacd_bvh *root_cage = build_cage(D.E.ctx, mesh, 0.03);  // cage mode only
    try {
      for (size_t i = 0; i < parts.size(); ++i) {
        seed.push_back(mix64(root + i));
        attempt.push_back(0);
        choose_points(parts[i], false, point_seed(seed[i]));
        add_visibility(D.E, parts[i], false, D.pair_buf, root_cage);
      }
    } catch (...) {
      vacd_bvh_destroy(root_cage);
      throw;
    }
    vacd_bvh_destroy(root_cage);
  }
  std::ostringstream hdr;
  hdr << "Starting decomposition (max parts=" << num_parts << ", concavity threshold="
      << std::to_string(concavity_threshold) << ", mode=" << config.score_mode << ").";
  log_line(hdr.str());
  const double nan = std::numeric_limits<double>::quiet_NaN();
  std::vector<double> conc(parts.size(), nan);  // cached per part
  for (int step = 0; step < num_parts - 1; ++step) {
    std::vector<double> key(parts.size());
    if (config.score_mode == "edge") {
      for (size_t i = 0; i < parts.size(); ++i) key[i] = edge_score(parts[i]);
    } else if (config.score_mode == "concavity") {
      for (size_t i = 0; i < parts.size(); ++i) {
        if (std::isnan(conc[i])) conc[i] = part_concavity(D.E, parts[i]);
        key[i] = conc[i];
      }
    }
    const int idx = pick_part(key);
    if (config.score_mode == "concavity" && idx >= 0 && key[idx] < concavity_threshold) {
      log_line("Step " + std::to_string(step + 1) + ": concavity " +
               std::to_string(key[idx]).substr(0, 6) + " < threshold, stopping.");
      break;
    }
    if (idx < 0) break;
    log_line("Step " + std::to_string(step + 1) + "/" + std::to_string(num_parts - 1) +
             ": splitting part " + std::to_string(idx) + " (" + std::to_string(parts.size()) +
             " parts total).");
    Mesh part = std::move(parts[idx]);
    const uint64_t s = seed[idx];
    const int at = attempt[idx];
    const double c = conc[idx];
    parts.erase(parts.begin() + idx);
    seed.erase(seed.begin() + idx);
    attempt.erase(attempt.begin() + idx);
    conc.erase(conc.begin() + idx);
    std::vector<size_t> ss;
    std::vector<std::vector<Plane>> planes = D.candidates({&part}, {s}, {at}, ss);
    if (planes[0].empty() || part.intersecting_edges.empty()) {
      parts.push_back(std::move(part));
      seed.push_back(s), attempt.push_back(at), conc.push_back(c);
      log_line("No more intersecting edges, stopping early.");
      break;
    }
    const int best = choose_planes(D.E, {&part}, planes, ss, {s})[0];
    MeshList kids;
    std::vector<uint64_t> ks;
    bool ok;
    {
      PhaseTimer timer("clip");
      ok = split_part(part, s, planes[0][best], kids, ks);
    }
    if (!ok) {  // clipping failed: the part goes back, to be retried with fresh planes
      parts.push_back(std::move(part));
      seed.push_back(s), attempt.push_back(at + 1), conc.push_back(c);
      continue;
    }
    for (size_t k = 0; k < kids.size(); ++k) {
      vacd_bvh *cage = build_cage(D.E.ctx, kids[k], 0.02);  // cage mode only
      try {
        add_visibility(D.E, kids[k], true, D.pair_buf, cage);
      } catch (...) {
        vacd_bvh_destroy(cage);
        throw;
      }
      vacd_bvh_destroy(cage);
      parts.push_back(std::move(kids[k]));
      seed.push_back(ks[k]), attempt.push_back(0), conc.push_back(nan);
    }
  }
  return finish(D, std::move(parts), bbox);
}

// ------------------------------------------------------------- batched
// The batched frontier gives exactly the plain loop's decomposition.  It
// keeps the tree of parts split so far and replays the plain loop on it
// (which part is split at each step, in which order parts are listed); when
// the replay reaches a part not yet split, it splits in one round that part
// and, speculatively, the other unsplit parts the loop is likely to split
// next (the highest concavities of its list, up to engine.batch_width and
// the steps left): their visibility, planes, scores, clips and the new
// parts' concavities each in one batched call or on all host threads.  A
// speculative split the loop never reaches is discarded work; parts split
// the same way whenever they are split, so the result does not depend on
// which parts a round took.
struct Node {
  Mesh mesh;
  uint64_t seed = 0;
  int attempt = 0;
  double conc = std::numeric_limits<double>::quiet_NaN();
  int state = 0;  // 0 not split yet, 1 split (kids), 2 nothing to split (no planes or edges)
  std::vector<int> kids;
  bool edges = false;  // visibility of its new pairs computed
};

// Visibility (masked when with_mask) of the listed nodes on a helper thread,
// overlapped with the host half of the concavity of the nodes in `conc_of`;
// then their distances on the GPU.
void visibility_and_concavity(Driver &D, std::vector<Node> &nodes, const std::vector<int> &vis_of,
                              bool with_mask, const std::vector<int> &conc_of,
                              const Mesh *cage_src = nullptr, double cage_level = 0.02) {
  std::vector<Mesh *> vm;
  for (int x : vis_of) vm.push_back(&nodes[x].mesh);
  std::vector<const Mesh *> cm;
  for (int x : conc_of) cm.push_back(&nodes[x].mesh);
  std::exception_ptr err;
  std::thread helper;
  if (!vm.empty())
    helper = std::thread([&] {
      try {
        add_visibility_batch(D.E, vm, with_mask, D.pair_buf, cage_src, cage_level);
      } catch (...) {
        err = std::current_exception();
      }
    });
  ConcavityPrep cp;
  try {
    PhaseTimer timer("concavity");
    if (!cm.empty()) cp = concavity_prep(cm);
  } catch (...) {
    if (helper.joinable()) helper.join();
    throw;
  }
  if (helper.joinable()) helper.join();
  if (err) std::rethrow_exception(err);
  for (int x : vis_of) nodes[x].edges = true;
  if (!cm.empty()) {
    PhaseTimer timer("concavity");
    const std::vector<double> c = concavity_finish(D.E, cp);
    for (size_t j = 0; j < conc_of.size(); ++j) nodes[conc_of[j]].conc = c[j];
  }
}

ProcessResult process_batched(Driver &D, Mesh mesh, double concavity_threshold, int num_parts,
                              uint64_t root) {
  const std::vector<double> bbox = normalize(mesh);
  log_line("Preprocessing mesh (" + std::to_string(mesh.vertices.size()) + " verts), batched...");
  mesh.is_new.clear();
  mesh.intersecting_edges.clear();
  mesh.points.clear();
  maybe_remesh(D, mesh);
  // the flat-surface planes (needed from the first split on) on a thread of
  // their own, beside the input parts' visibility and concavity
  std::exception_ptr flat_err;
  std::thread flat_th;
  if (config.use_flat_surfaces)
    flat_th = std::thread([&] {
      try {
        D.flat = flat_surface_planes(mesh, config.flat_surface_min_area);
      } catch (...) {
        flat_err = std::current_exception();
      }
    });
  struct Joiner {
    std::thread &t;
    ~Joiner() {
      if (t.joinable()) t.join();
    }
  } flat_join{flat_th};
  const bool by_edges = config.score_mode == "edge";
  const bool by_conc = config.score_mode == "concavity";
  std::vector<Node> nodes;
  std::vector<int> L;  // the plain loop's list of parts
  {
    MeshList roots = separate_disjoint(mesh);
    for (size_t i = 0; i < roots.size(); ++i) {
      Node nd;
      nd.mesh = std::move(roots[i]);
      nd.seed = mix64(root + i);
      choose_points(nd.mesh, false, point_seed(nd.seed));
      nodes.push_back(std::move(nd));
      L.push_back((int)i);
    }
  }
  log_line("  " + std::to_string(L.size()) + " connected part(s)");
  visibility_and_concavity(D, nodes, L, false, by_conc ? L : std::vector<int>{}, &mesh, 0.03);
  if (flat_th.joinable()) flat_th.join();
  if (flat_err) std::rethrow_exception(flat_err);
  log_line("  " + std::to_string(D.flat.size()) + " flat surface(s)");
  auto key_of = [&](int x) {
    if (by_edges) return edge_score(nodes[x].mesh);
    if (by_conc) return nodes[x].conc;
    return 0.0;
  };
  int step = 0, round = 0;
  for (;;) {
    // replay the plain loop as far as the tree goes
    int need = -1;
    while (step < num_parts - 1) {
      std::vector<double> key(L.size());
      for (size_t i = 0; i < L.size(); ++i) key[i] = key_of(L[i]);
      const int idx = pick_part(key);
      if (idx < 0 || (by_conc && key[idx] < concavity_threshold)) break;
      const int x = L[idx];
      if (nodes[x].state == 0) {
        need = x;
        break;
      }
      L.erase(L.begin() + idx);
      if (nodes[x].state == 2) {  // no planes or no edges: the loop stops
        L.push_back(x);
        step = num_parts;
        break;
      }
      for (int k : nodes[x].kids) L.push_back(k);
      nodes[x].mesh = Mesh();  // split for good: its mesh is not needed again
      ++step;
    }
    if (need < 0) break;
    // this round: the part the loop needs and the unsplit parts most likely
    // next -- the leaves of the split tree under the loop's list (parts the
    // loop has not reached yet but already split count through their
    // pieces), highest key first
    const int left = num_parts - 1 - step;
    const int width = std::max(1, std::min(engine.batch_width, left));
    std::vector<std::pair<double, int>> cand;
    std::vector<int> walk(L.rbegin(), L.rend());
    while (!walk.empty()) {
      const int x = walk.back();
      walk.pop_back();
      if (nodes[x].state == 1) {
        for (auto k = nodes[x].kids.rbegin(); k != nodes[x].kids.rend(); ++k) walk.push_back(*k);
        continue;
      }
      if (x != need && nodes[x].state == 0 && (!by_conc || nodes[x].conc >= concavity_threshold))
        cand.push_back({key_of(x), x});
    }
    std::stable_sort(cand.begin(), cand.end(),
                     [](const auto &p, const auto &q) { return p.first > q.first; });
    std::vector<int> batch{need};
    for (size_t c = 0; c < cand.size() && (int)batch.size() < width; ++c) batch.push_back(cand[c].second);
    ++round;
    log_line("Round " + std::to_string(round) + ": step " + std::to_string(step + 1) + ", splitting " +
             std::to_string(batch.size()) + " part(s) (" + std::to_string(L.size()) + " listed).");
    // visibility of the parts whose new pairs are still untested
    std::vector<int> novis;
    for (int x : batch)
      if (!nodes[x].edges) novis.push_back(x);
    if (!novis.empty()) visibility_and_concavity(D, nodes, novis, true, {});
    // planes and the cut of every part, in batched calls
    const size_t F = batch.size();
    std::vector<const Mesh *> ms(F);
    std::vector<uint64_t> seeds(F);
    std::vector<int> attempts(F);
    for (size_t f = 0; f < F; ++f) {
      ms[f] = &nodes[batch[f]].mesh;
      seeds[f] = nodes[batch[f]].seed;
      attempts[f] = nodes[batch[f]].attempt;
    }
    std::vector<size_t> ss;
    const std::vector<std::vector<Plane>> planes = D.candidates(ms, seeds, attempts, ss);
    std::vector<int> best(F, -1);
    {
      std::vector<const Mesh *> live_m;
      std::vector<std::vector<Plane>> live_p;
      std::vector<size_t> live_ss, live_f;
      std::vector<uint64_t> live_s;
      for (size_t f = 0; f < F; ++f) {
        if (planes[f].empty() || ms[f]->intersecting_edges.empty()) continue;
        live_m.push_back(ms[f]), live_p.push_back(planes[f]), live_ss.push_back(ss[f]);
        live_s.push_back(seeds[f]), live_f.push_back(f);
      }
      const std::vector<int> b = choose_planes(D.E, live_m, live_p, live_ss, live_s);
      for (size_t j = 0; j < live_f.size(); ++j) best[live_f[j]] = b[j];
    }
    // clips on all host threads
    std::vector<MeshList> kids(F);
    std::vector<std::vector<uint64_t>> kseeds(F);
    std::vector<char> ok(F, 0);
    {
      PhaseTimer t("clip");
#pragma omp parallel for schedule(dynamic)
      for (long f = 0; f < (long)F; ++f)
        if (best[f] >= 0) ok[f] = split_part(*ms[f], seeds[f], planes[f][best[f]], kids[f], kseeds[f]);
    }
    std::vector<int> fresh;
    for (size_t f = 0; f < F; ++f) {
      const int x = batch[f];
      if (best[f] < 0) {
        nodes[x].state = 2;
        continue;
      }
      nodes[x].state = 1;
      if (!ok[f]) {  // clipping failed: retried later with fresh planes
        Node retry;
        retry.mesh = std::move(nodes[x].mesh);
        retry.seed = nodes[x].seed;
        retry.attempt = nodes[x].attempt + 1;
        retry.conc = nodes[x].conc;
        retry.edges = true;
        nodes.push_back(std::move(retry));
        nodes[x].kids = {(int)nodes.size() - 1};
        continue;
      }
      for (size_t k = 0; k < kids[f].size(); ++k) {
        Node kid;
        kid.mesh = std::move(kids[f][k]);
        kid.seed = kseeds[f][k];
        nodes.push_back(std::move(kid));
        nodes[x].kids.push_back((int)nodes.size() - 1);
        fresh.push_back((int)nodes.size() - 1);
      }
    }
    log_line("  parts clipped: " + std::to_string(fresh.size()) + " new");
    // the new parts' concavity; their visibility now (overlapped) when
    // engine.eager_visibility or when the edge score needs it
    std::vector<int> vis_now;
    if (engine.eager_visibility || by_edges) vis_now = fresh;
    visibility_and_concavity(D, nodes, vis_now, true, by_conc ? fresh : std::vector<int>{});
    log_line("  visibility and concavity done");
  }
  MeshList parts;
  for (int x : L) parts.push_back(std::move(nodes[x].mesh));
  last_profile.count("batched.rounds", (long)round);
  last_profile.count("batched.parts_split", (long)std::count_if(nodes.begin(), nodes.end(),
                                                                [](const Node &n) { return n.state == 1; }));
  return finish(D, std::move(parts), bbox);
}

}  // namespace

std::vector<Plane> bisector_planes_seeded(const Mesh &m, int num_planes, uint32_t seed) {
  std::mt19937 rng(seed);
  return bisector_planes(m, num_planes, rng);
}

bool split_part_seeded(const Mesh &mesh, uint64_t seed, const Plane &plane, MeshList &kids,
                       std::vector<uint64_t> &kid_seeds) {
  return split_part(mesh, seed, plane, kids, kid_seeds);
}

void sphere_draws_seeded(int num_planes, uint32_t seed, std::vector<double> &normals,
                         std::vector<double> &frac) {
  std::mt19937 rng(seed);
  SphereDraws d = sphere_draws(num_planes, rng);
  normals.swap(d.normals);
  frac.swap(d.frac);
}

int select_plane(const std::vector<int64_t> &q32, size_t ss_offset) {
  std::vector<int> idx(q32.size());
  std::iota(idx.begin(), idx.end(), 0);
  return first_max(idx, q32, ss_offset);
}

ProcessResult process(Mesh mesh, double concavity_threshold, int num_parts) {
  last_profile.clear();
  process_start = std::chrono::steady_clock::now();
  PhaseTimer total("total");
  // the run's root seed: two draws of the global generator (set_seed)
  const uint64_t root = ((uint64_t)random_engine()() << 32) | (uint64_t)random_engine()();
  Driver D(engine.device);
  if (engine.mode == "batched")
    return process_batched(D, std::move(mesh), concavity_threshold, num_parts, root);
  if (engine.mode != "plain") throw std::invalid_argument("visacd: engine.mode must be \"plain\" or \"batched\"");
  return process_plain(D, std::move(mesh), concavity_threshold, num_parts, root);
}

}  // namespace vhost
