// Host-side mesh types and geometry services of the decomposition driver.
//
// Mirrors the reference's data model (lib/include/core.hpp:18-65,
// lib/include/config.hpp:6-30, lib/include/process.hpp:9-15) so the pybind
// surface is signature-identical; the algorithms are written independently.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <random>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace vhost {

using Vec3 = std::array<double, 3>;

inline Vec3 operator+(const Vec3 &a, const Vec3 &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline Vec3 operator-(const Vec3 &a, const Vec3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline Vec3 operator*(const Vec3 &a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline double dot(const Vec3 &a, const Vec3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline Vec3 cross(const Vec3 &a, const Vec3 &b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double norm(const Vec3 &a) { return std::sqrt(dot(a, a)); }

// Plane a*x + b*y + c*z + d = 0 (lib/include/core.hpp:18-35).
struct Plane {
  double a = 0, b = 0, c = 0, d = 0;
  Plane() = default;
  Plane(double a_, double b_, double c_, double d_) : a(a_), b(b_), c(c_), d(d_) {}
  double eval(const Vec3 &p) const { return p[0] * a + p[1] * b + p[2] * c + d; }
  // +1 / -1 / 0 with dead zone eps (lib/include/core.hpp:178-185)
  int side(const Vec3 &p, double eps = 1e-6) const {
    const double r = eval(p);
    return r > eps ? 1 : (r < -eps ? -1 : 0);
  }
};

struct Mesh {
  std::vector<Vec3> vertices;
  std::vector<std::array<int, 3>> triangles;
  // per vertex: created by the last cut (cap vertex); pairs of two old
  // vertices are not re-tested (lib/src/clip.cpp:664-673)
  std::vector<bool> is_new;
  // visibility edges as pairs of vertex ids of this mesh, k order first
  std::vector<std::pair<unsigned, unsigned>> intersecting_edges;
  // vertex ids whose pairs are tested; empty = all vertices
  std::vector<int> points;
  // planes of the cuts that made this part, oldest first (their caps are
  // planar groups of the part's BVH)
  std::vector<Plane> cut_planes;

  Mesh copy() const {  // geometry only (lib/src/core.cpp:239-246)
    Mesh m;
    m.vertices = vertices;
    m.triangles = triangles;
    return m;
  }
};
using MeshList = std::vector<Mesh>;

// Process-global knobs, same fields and defaults as the reference.
struct Config {
  bool return_parts = false;
  std::string score_mode = "concavity";  // "edge" or "concavity"
  double flat_surface_min_area = 0.1;
  bool use_flat_surfaces = true;
  double flat_surface_k = 2.0;
  bool use_merging = false;
};
inline Config config;

// Engine knobs that the reference does not have (kept off its Config so the
// Config surface stays identical).
struct EngineOptions {
  int device = 0;
  int max_points = 0;            // visibility points per part (seeded subsample); 0 = all
  int num_planes = 2500;         // candidate planes per cut
  std::string planes = "bisector";  // "bisector" (reference) or "sphere"
  int hausdorff_resolution = 3000;
  uint32_t sample_seed = 42;     // seeds subsampling and surface sampling
  bool gpu_distance = true;      // Hausdorff term on the GPU (exact point-to-mesh)
  // "parity": a pair is an edge when its segment touches no triangle and its
  // midpoint is outside the part; "cage": when it touches no triangle and
  // touches the part's cage, the surface at distance 0.02 (0.03 for the
  // input) on a 1/40 grid, as the reference does (lib/ptx/ray_segments.cu:
  // 96-110, lib/src/process.cpp:170-171, 260-261)
  std::string visibility = "parity";
  // input remesh (lib/src/process.cpp:250-258): "auto" remeshes inputs that
  // are not closed, "always" every input (as the reference), "never" none
  std::string remesh = "auto";
  // cap rule: a pair of two points on one cut plane whose segment runs into
  // that cut's cap at an end touches the cap and is not an edge, decided by
  // exact 2-D orientations in the plane instead of by the rounded cap
  // vertices (needs planar_caps; the reference tests such pairs like any)
  bool cap_rule = true;
  // "plain": the reference's loop, one part split per step; "batched": the
  // same decomposition with the parts the loop will split next split
  // together, up to batch_width per round
  std::string mode = "plain";
  int batch_width = 64;
  // walk queues of up to this many pairs are walked breadth first, longer
  // ones depth first (VACD_OPT_FRONT_MAX); -1 keeps the library's default
  int64_t front_max = -1;
  // batched: compute the new parts' visibility as soon as they are made,
  // overlapped with their concavity on the host (true), or only when a part
  // is about to be split (false: no visibility for parts never split)
  bool eager_visibility = false;
  // plane selection: "exact" scores every candidate plane (the reference);
  // "learned" ranks the candidates with the proposal model, scores the
  // proposal_top best-ranked and proposal_audit random ones exactly, and
  // scores them all when the best of the top is under proposal_audit_ratio
  // times the best of the audit
  std::string proposal = "exact";
  std::string proposal_weights;  // float32 parameter file of the model
  int proposal_top = 256;
  int proposal_audit = 256;
  double proposal_audit_ratio = 0.98;
  // cap triangles of earlier cuts go to the visibility BVH as planar groups
  bool planar_caps = true;
  bool verbose = true;
};
inline EngineOptions engine;

// Wall time per phase of the last process() call, in seconds.
// Phase times and counters of the last process() call; add and count may be
// called from several host threads.
struct Profile {
  std::map<std::string, double> seconds;
  std::map<std::string, long> calls;
  std::mutex mu;
  void add(const std::string &phase, double s) {
    std::lock_guard<std::mutex> g(mu);
    seconds[phase] += s;
    calls[phase] += 1;
  }
  void count(const std::string &name, long v) {
    std::lock_guard<std::mutex> g(mu);
    calls[name] += v;
  }
  void clear() {
    std::lock_guard<std::mutex> g(mu);
    seconds.clear();
    calls.clear();
  }
};
inline Profile last_profile;

struct ProcessResult {
  MeshList parts;
  double concavity = 0.0;
  int num_parts = 0;
};

// global generator of the candidate-plane draws (lib/include/core.hpp:13)
std::mt19937 &random_engine();
void set_seed(unsigned int seed);

// geometry services
std::vector<double> normalize(Mesh &m);                        // returns the original bbox
void unnormalize(Mesh &m, const std::vector<double> &bbox);
double mesh_volume(const Mesh &m);
double surface_area(const Mesh &m);
double tri_area(const Vec3 &a, const Vec3 &b, const Vec3 &c);
Mesh convex_hull(const std::vector<Vec3> &pts);
MeshList separate_disjoint(const Mesh &m, bool drop_flat = true);
MeshList separate_disjoint(Mesh &&m, bool drop_flat = true);  // moves a single component
// Splits m by the plane; returns 2 parts (positive side first) or fewer on
// failure.  map_pos/map_neg receive 1-based vertex maps (0 = not in part).
MeshList clip(const Mesh &m, const Plane &pl, std::vector<int> &map_pos,
              std::vector<int> &map_neg);
double concavity(const Mesh &part, const Mesh &hull, int resolution, uint32_t seed);
// max(1000, resolution * area) area-weighted surface points plus one point in
// every triangle left unsampled (lib/src/core.cpp:144-232)
std::vector<Vec3> surface_samples(const Mesh &m, int resolution, std::mt19937 &rng);
// A sampled distance between two meshes draws the first mesh's samples from
// std::mt19937(seed) and the second's (the hull's) from
// std::mt19937(second_sample_seed(seed)), so that the two can be drawn
// concurrently (the part's while its hull is being built).
inline uint32_t second_sample_seed(uint32_t seed) { return seed ^ 0x6a09e667u; }
double hausdorff_boundary(const Mesh &a, const Mesh &b, int resolution, uint32_t seed);
std::vector<Plane> flat_surface_planes(const Mesh &m, double min_area);

ProcessResult process(Mesh mesh, double concavity, int num_parts);
// The driver's candidate sampler with a fresh std::mt19937(seed) over
// m.intersecting_edges (lib/src/process.cpp:28-72), and its plane choice
// over Q32.32 scores: first maximum, planes from ss_offset on (flat
// surfaces) times config.flat_surface_k (lib/src/process.cpp:128-134).
std::vector<Plane> bisector_planes_seeded(const Mesh &m, int num_planes, uint32_t seed);
int select_plane(const std::vector<int64_t> &q32, size_t ss_offset);
// The driver's split of a part (clip, inherited edges through the 1-based
// maps, connected pieces of at least 10 vertices with points, cut planes and
// seeds; lib/src/process.cpp:136-168); false when the clip fails.
bool split_part_seeded(const Mesh &mesh, uint64_t seed, const Plane &plane, MeshList &kids,
                       std::vector<uint64_t> &kid_seeds);
// The sphere sampler's draws (normals, offset fractions) with std::mt19937(seed).
void sphere_draws_seeded(int num_planes, uint32_t seed, std::vector<double> &normals,
                         std::vector<double> &frac);

}  // namespace vhost
