// Inputs of the learned plane proposal; declarations and conventions in
// proposal.hpp.
#include "proposal.hpp"

#include <algorithm>
#include <cmath>

namespace vhost {

// splitmix64: per-part seeds derived from the parent's, independent of the
// order in which parts are processed.
uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

PartFrame part_frame(const std::vector<Vec3> &v) {
  PartFrame fr;
  if (v.empty()) return fr;
  Vec3 lo = v[0], hi = v[0];
  for (const Vec3 &p : v)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], p[k]), hi[k] = std::max(hi[k], p[k]);
  double ext = 0.0;
  for (int k = 0; k < 3; ++k) fr.c[k] = 0.5 * (lo[k] + hi[k]), ext = std::max(ext, hi[k] - lo[k]);
  fr.s = ext > 0.0 ? 0.5 * ext : 1.0;
  return fr;
}

void edge_features(const std::vector<Vec3> &v, const std::vector<std::pair<unsigned, unsigned>> &edges,
                   uint64_t seed, const PartFrame &fr, float *out) {
  const uint64_t E = edges.size();
  for (int k = 0; k < kProposalEdges; ++k) {
    uint64_t e = 0;
    if (E == 0) {
      for (int j = 0; j < kEdgeFeatures; ++j) out[kEdgeFeatures * k + j] = 0.0f;
      continue;
    }
    if (E <= (uint64_t)kProposalEdges) {
      e = (uint64_t)k % E;
    } else {
      const uint64_t lo = (uint64_t)k * E / kProposalEdges, hi = (uint64_t)(k + 1) * E / kProposalEdges;
      e = lo + mix64(seed + (uint64_t)k) % (hi - lo);
    }
    Vec3 a = v[edges[e].first], b = v[edges[e].second];
    for (int j = 0; j < 3; ++j) a[j] = (a[j] - fr.c[j]) / fr.s, b[j] = (b[j] - fr.c[j]) / fr.s;
    if (b < a) std::swap(a, b);
    float *o = out + kEdgeFeatures * k;
    for (int j = 0; j < 3; ++j) o[j] = (float)a[j], o[3 + j] = (float)b[j];
    o[6] = (float)norm(b - a);
  }
}

void plane_inputs(const std::vector<Plane> &planes, size_t count, const PartFrame &fr, float *out) {
  for (size_t i = 0; i < count; ++i) {
    const Plane &p = planes[i];
    double l = std::sqrt(p.a * p.a + p.b * p.b + p.c * p.c);
    if (!(l > 0.0)) l = 1.0;
    double a = p.a / l, b = p.b / l, c = p.c / l, d = p.d / l;
    d = (a * fr.c[0] + b * fr.c[1] + c * fr.c[2] + d) / fr.s;
    // (n, d) and (-n, -d) are one plane: the sign is fixed by d > 0, or by the
    // first non-zero normal component when d = 0
    const bool flip = d < 0.0 || (d == 0.0 && (a < 0.0 || (a == 0.0 && (b < 0.0 || (b == 0.0 && c < 0.0)))));
    if (flip) a = -a, b = -b, c = -c, d = -d;
    out[4 * i] = (float)a, out[4 * i + 1] = (float)b, out[4 * i + 2] = (float)c;
    out[4 * i + 3] = (float)d;
  }
}

}  // namespace vhost
