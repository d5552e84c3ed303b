// Segment-visibility kernels for gfx950.
//
// Replaces the OptiX raygen/any-hit programs (lib/ptx/ray_segments.cu:49-125),
// the dense uint32 result matrix and the host combine loop of
// compute_intersection_matrix (lib/src/intersections.cpp:327-385).
//
// Pairs are numbered k = i(2n-i-1)/2 + j-i-1; a wave owns 64 consecutive k,
// so its ballot word holds 64 results in k order.  A segment that touches
// nothing lies wholly inside or outside, so its midpoint parity equals the
// parity at p + e(q-p), e -> 0+; on a closed mesh with p one of its vertices
// that is (crossings of p's own fan) xor (crossings of the +x ray from p
// with all other triangles).
// k_far_parity: the second term, once per point (a wave per point).
// k_fan: the triangles having each point as a vertex.
// k_visibility: one pair per lane.  With fans (closed mesh, parity mode) it
//   only tests p's fan: a pair leaving p into the solid is never an edge and
//   is settled there; the rest are queued.  Otherwise (cage mode, open mesh)
//   each lane walks its pair itself.
// k_walk_split: the queued pairs, 2^k lanes per pair, any contact with the
//   self BVH (short stack of node ids in LDS, lane-interleaved so the 64
//   lanes hit 64 banks), then the local parity, or the midpoint ray when a
//   fan crossing is degenerate or the point's far parity is.
// Triangle tests use float64 filters; a pair whose walk meets a test the
// filter cannot decide is appended to a deferred list, and k_walk_exact
// redoes those pairs one per lane with exact predicates (expansion
// arithmetic) in place of the undecided tests.
// k_word_scan / k_scan_blocks / k_emit: popcount prefix over the ballot
//   words and emission of (i, j) in k order.
#include "internal.hpp"
#include "predicates.hpp"

#include <algorithm>
#include <cstring>

namespace vacd {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kScanTile = 1024;  // ballot words per scan block (256 threads x 4)

__device__ __forceinline__ int64_t row_start(int64_t i, int64_t n) {
  return i * (2 * n - i - 1) / 2;
}

// k -> (i, j), i < j: closed form in float64 (lib/ptx/ray_segments.cu:39-46),
// then exact integer correction.
__device__ __forceinline__ void pair_of(int64_t k, int64_t n, int64_t &i, int64_t &j) {
  const double nd = (double)n, kd = (double)k;
  const double t = sqrt(fmax(0.0, -8.0 * kd + 4.0 * nd * (nd - 1.0) - 7.0));
  int64_t ii = (int64_t)(nd - 2.0 - floor(t * 0.5 - 0.5));
  ii = ii < 0 ? 0 : (ii > n - 2 ? n - 2 : ii);
  while (ii > 0 && row_start(ii, n) > k) --ii;
  while (ii < n - 2 && row_start(ii + 1, n) <= k) ++ii;
  i = ii;
  j = k - row_start(ii, n) + ii + 1;
}

struct Seg {
  float p[3], q[3];
  float inv[3], pinv[3];  // slab test: t = lo*inv - p*inv
  float m[3];             // midpoint, float32 (box tests only)
};

__device__ __forceinline__ void make_seg(Seg &s) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float d = s.q[k] - s.p[k];
    s.inv[k] = fabsf(d) > 1e-30f ? 1.0f / d : 1e30f;
    s.pinv[k] = s.p[k] * s.inv[k];
    s.m[k] = 0.5f * (s.p[k] + s.q[k]);
  }
}

__device__ __forceinline__ bool seg_box(const Seg &s, float lx, float ly, float lz, float hx,
                                        float hy, float hz) {
  const float ax = fmaf(lx, s.inv[0], -s.pinv[0]), bx = fmaf(hx, s.inv[0], -s.pinv[0]);
  const float ay = fmaf(ly, s.inv[1], -s.pinv[1]), by = fmaf(hy, s.inv[1], -s.pinv[1]);
  const float az = fmaf(lz, s.inv[2], -s.pinv[2]), bz = fmaf(hz, s.inv[2], -s.pinv[2]);
  const float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), 0.0f));
  const float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), 1.0f));
  return t0 <= t1;
}

// seg_box over the parameter range [t0, t1] of the segment only
__device__ __forceinline__ bool seg_box_span(const Seg &s, float lx, float ly, float lz, float hx,
                                             float hy, float hz, float t0, float t1) {
  const float ax = fmaf(lx, s.inv[0], -s.pinv[0]), bx = fmaf(hx, s.inv[0], -s.pinv[0]);
  const float ay = fmaf(ly, s.inv[1], -s.pinv[1]), by = fmaf(hy, s.inv[1], -s.pinv[1]);
  const float az = fmaf(lz, s.inv[2], -s.pinv[2]), bz = fmaf(hz, s.inv[2], -s.pinv[2]);
  const float u0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), t0));
  const float u1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), t1));
  return u0 <= u1;
}

// Parameters [t0, t1] (within [0, 1]) of the part of the segment within tau
// of planar group `code`'s plane (grp: float4 pairs (n, d), (tau, ...)), with
// room for the float32 rounding of this evaluation: every point of the
// segment on a triangle of the group lies in that part.  False when empty.
__device__ __forceinline__ bool group_span(const Seg &s, const float4 *__restrict__ grp, int code,
                                           float &t0, float &t1) {
  const float4 pl = grp[2 * (code - 1)];
  const float tau = grp[2 * (code - 1) + 1].x;
  const float dx = s.q[0] - s.p[0], dy = s.q[1] - s.p[1], dz = s.q[2] - s.p[2];
  const float a = fmaf(pl.x, s.p[0], fmaf(pl.y, s.p[1], fmaf(pl.z, s.p[2], pl.w)));
  const float b = fmaf(pl.x, dx, fmaf(pl.y, dy, pl.z * dz));
  // |n| = 1: each rounding error is below 2^-23 of these magnitudes
  const float err = 1e-6f * (fabsf(s.p[0]) + fabsf(s.p[1]) + fabsf(s.p[2]) + fabsf(pl.w) +
                             fabsf(dx) + fabsf(dy) + fabsf(dz));
  const float T = tau + err;
  if (!(fabsf(b) > 1e-30f)) {
    t0 = 0.0f, t1 = 1.0f;
    return fabsf(a) <= T;
  }
  float u0 = (-T - a) / b, u1 = (T - a) / b;
  if (u0 > u1) {
    const float x = u0;
    u0 = u1, u1 = x;
  }
  u0 -= 1e-6f * (1.0f + fabsf(u0));
  u1 += 1e-6f * (1.0f + fabsf(u1));
  t0 = fmaxf(u0, 0.0f), t1 = fminf(u1, 1.0f);
  return t0 <= t1;
}

__device__ __forceinline__ bool ray_box(const Seg &s, float lx, float ly, float lz, float hx,
                                        float hy, float hz) {
  (void)lx;
  return s.m[1] >= ly && s.m[1] <= hy && s.m[2] >= lz && s.m[2] <= hz && hx >= s.m[0];
}

// orient2d in (y, z), float64, with its error bound.
__device__ __forceinline__ double o2yz(const float *a, const float *b, const float *c,
                                       double &bound) {
  return o2_fast(a, b, c, 0, bound);
}

// Contact of the segment with a triangle: 0 none, 1 contact, 2 undecided.
// For a triangle of p's fan (having p, not q, as a vertex): 4 contact
// undecided, 5 no contact but the crossing below undecided.
// For a triangle incident to p (and not to q) it also toggles xpar when the
// +x ray from p + e(q-p), e -> 0+, crosses the triangle, and counts it in
// ninc.  Over the full fan of p on a closed mesh, xpar is then the inside
// parity of the segment's interior (the same as that of its midpoint when
// the segment touches nothing).
__device__ __forceinline__ int contact_fast(const Seg &s, const Tri &t, int &xpar, int &ninc) {
  const float a[3] = {t.a.x, t.a.y, t.a.z}, b[3] = {t.b.x, t.b.y, t.b.z},
              c[3] = {t.c.x, t.c.y, t.c.z};
  const bool pa = same3(s.p, a), pb = same3(s.p, b), pc = same3(s.p, c);
  const bool ip = pa || pb || pc;
  const bool iq = same3(s.q, a) || same3(s.q, b) || same3(s.q, c);
  if (ip && iq) return 1;  // the segment is an edge of the triangle
  if (iq) {
    double bound;
    const double det = o3_fast(a, b, c, s.p, bound);
    return (det > bound || -det > bound) ? 0 : 2;
  }
  if (ip) {
    double bound;
    const double det = o3_fast(a, b, c, s.q, bound);  // orient3d(p, b', c', q)
    // a sliced triangle (contact tree) counts its crossing in the slice
    // holding p only: the apex's slice is the last, the other vertices' the
    // first; unsliced records have meta 0
    const unsigned meta = __float_as_uint(t.b.w);
    const unsigned sl = meta & 0xffu, nsl = (meta >> 8) & 0xffu, apex = (meta >> 16) & 3u;
    const unsigned vp = pa ? 0u : (pb ? 1u : 2u);
    const bool counts = nsl == 0u || (vp == apex ? sl == nsl - 1u : sl == 0u);
    if (!(det > bound || -det > bound)) return counts ? 4 : 2;
    if (!counts) return 0;  // q is off the plane: the segment meets it at p only
    ++ninc;
    // the triangle rotated to start at p: (p, v1, v2)
    const float *v1 = pa ? b : (pb ? c : a);
    const float *v2 = pa ? c : (pb ? a : b);
    double b0, b1, b2;
    const double o0 = o2yz(s.p, v1, v2, b0);
    const double o1 = o2yz(s.p, v1, s.q, b1);
    const double o2 = o2yz(v2, s.p, s.q, b2);
    // q is off the triangle's plane: the segment meets it at p only
    if (!(o0 > b0 || -o0 > b0) || !(o1 > b1 || -o1 > b1) || !(o2 > b2 || -o2 > b2)) return 5;
    const bool s0 = o0 > 0.0;
    if ((o1 > 0.0) == s0 && (o2 > 0.0) == s0 && (det > 0.0) == s0) xpar ^= 1;
    return 0;
  }
  // plane sides of p and q: det[a-x, b-a, c-a] = orient3d(a, b, c, x)
  const double e1x = (double)b[0] - a[0], e1y = (double)b[1] - a[1], e1z = (double)b[2] - a[2];
  const double e2x = (double)c[0] - a[0], e2y = (double)c[1] - a[1], e2z = (double)c[2] - a[2];
  const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const double mx = fabs(e1y * e2z) + fabs(e1z * e2y), my = fabs(e1z * e2x) + fabs(e1x * e2z),
               mz = fabs(e1x * e2y) + fabs(e1y * e2x);
  const double apx = (double)a[0] - s.p[0], apy = (double)a[1] - s.p[1], apz = (double)a[2] - s.p[2];
  const double aqx = (double)a[0] - s.q[0], aqy = (double)a[1] - s.q[1], aqz = (double)a[2] - s.q[2];
  const double dp = apx * nx + apy * ny + apz * nz;
  const double dq = aqx * nx + aqy * ny + aqz * nz;
  const double bp = kFilter * (fabs(apx) * mx + fabs(apy) * my + fabs(apz) * mz);
  const double bq = kFilter * (fabs(aqx) * mx + fabs(aqy) * my + fabs(aqz) * mz);
  const bool cp = dp > bp || -dp > bp, cq = dq > bq || -dq > bq;
  if (cp && cq && ((dp > 0.0) == (dq > 0.0))) return 0;  // strictly one side
  // does the line pq pass through the triangle?  Asked also when a side is
  // undecided: a line that certainly misses the closed triangle rules out
  // contact whatever the sides (an endpoint lying in the plane of a
  // triangle it is not a vertex of, as on a flat face, is the common case).
  // u = D.(BxC), v = D.(CxA), w = D.(AxB) with A = a-p, B = b-p, C = c-p, D = q-p;
  // u + v + w = D.N, so (u, v, w) / D.N are the barycentric coordinates of
  // where the line meets the plane; with D.N = 0 the line is parallel to it.
  const double Ax = apx, Ay = apy, Az = apz;
  const double Bx = (double)b[0] - s.p[0], By = (double)b[1] - s.p[1], Bz = (double)b[2] - s.p[2];
  const double Cx = (double)c[0] - s.p[0], Cy = (double)c[1] - s.p[1], Cz = (double)c[2] - s.p[2];
  const double Dx = (double)s.q[0] - s.p[0], Dy = (double)s.q[1] - s.p[1],
               Dz = (double)s.q[2] - s.p[2];
  const double u = Dx * (By * Cz - Bz * Cy) + Dy * (Bz * Cx - Bx * Cz) + Dz * (Bx * Cy - By * Cx);
  const double v = Dx * (Cy * Az - Cz * Ay) + Dy * (Cz * Ax - Cx * Az) + Dz * (Cx * Ay - Cy * Ax);
  const double w = Dx * (Ay * Bz - Az * By) + Dy * (Az * Bx - Ax * Bz) + Dz * (Ax * By - Ay * Bx);
  const double adx = fabs(Dx), ady = fabs(Dy), adz = fabs(Dz);
  const double bu = kFilter * (adx * (fabs(By * Cz) + fabs(Bz * Cy)) +
                               ady * (fabs(Bz * Cx) + fabs(Bx * Cz)) +
                               adz * (fabs(Bx * Cy) + fabs(By * Cx)));
  const double bv = kFilter * (adx * (fabs(Cy * Az) + fabs(Cz * Ay)) +
                               ady * (fabs(Cz * Ax) + fabs(Cx * Az)) +
                               adz * (fabs(Cx * Ay) + fabs(Cy * Ax)));
  const double bw = kFilter * (adx * (fabs(Ay * Bz) + fabs(Az * By)) +
                               ady * (fabs(Az * Bx) + fabs(Ax * Bz)) +
                               adz * (fabs(Ax * By) + fabs(Ay * Bx)));
  const bool pos = u > bu || v > bv || w > bw;
  const bool neg = -u > bu || -v > bv || -w > bw;
  if (pos && neg) return 0;  // the line misses the triangle
  if (!(cp && cq)) return 2;
  // opposite sides: contact iff the line passes through the triangle
  const bool all = (u > bu || -u > bu) && (v > bv || -v > bv) && (w > bw || -w > bw);
  return all ? 1 : 2;
}

// +x ray from m = (p+q)/2 against a triangle: 0 no crossing, 1 crossing,
// 2 undecided.
__device__ __forceinline__ int cross_fast(const Seg &s, const Tri &t) {
  const double my = 0.5 * ((double)s.p[1] + (double)s.q[1]);
  const double mz = 0.5 * ((double)s.p[2] + (double)s.q[2]);
  const double mag = fmax(fabs(my), fabs(mz));
  const double ay = (double)t.a.y - my, az = (double)t.a.z - mz;
  const double by = (double)t.b.y - my, bz = (double)t.b.z - mz;
  const double cy = (double)t.c.y - my, cz = (double)t.c.z - mz;
  // orient2d(a, b, m), orient2d(b, c, m), orient2d(c, a, m) in (y, z)
  const double o1 = ay * bz - az * by, o2 = by * cz - bz * cy, o3 = cy * az - cz * ay;
  const double b1 = kFilter * (fabs(ay * bz) + fabs(az * by) +
                               mag * (fabs((double)t.a.z - t.b.z) + fabs((double)t.b.y - t.a.y)));
  const double b2 = kFilter * (fabs(by * cz) + fabs(bz * cy) +
                               mag * (fabs((double)t.b.z - t.c.z) + fabs((double)t.c.y - t.b.y)));
  const double b3 = kFilter * (fabs(cy * az) + fabs(cz * ay) +
                               mag * (fabs((double)t.c.z - t.a.z) + fabs((double)t.a.y - t.c.y)));
  const bool pos = o1 > b1 || o2 > b2 || o3 > b3;
  const bool neg = -o1 > b1 || -o2 > b2 || -o3 > b3;
  if (pos && neg) return 0;
  const bool all = (o1 > b1 || -o1 > b1) && (o2 > b2 || -o2 > b2) && (o3 > b3 || -o3 > b3);
  if (!all) return 2;
  // the ray line passes inside; crossing at t > 0 iff orient3d(a,b,c,m) has
  // the sign of the projected orientation.
  const double mx = 0.5 * ((double)s.p[0] + (double)s.q[0]);
  const double ax = (double)t.a.x - mx, bx = (double)t.b.x - mx, cx = (double)t.c.x - mx;
  const double det = ax * (by * cz - bz * cy) + bx * (cy * az - cz * ay) + cx * (ay * bz - az * by);
  const double mag3 = fmax(mag, fabs(mx));
  const double bd = kFilter * (fabs(ax) * (fabs(by * cz) + fabs(bz * cy)) +
                               fabs(bx) * (fabs(cy * az) + fabs(cz * ay)) +
                               fabs(cx) * (fabs(ay * bz) + fabs(az * by)) +
                               mag3 * ((fabs(by) + fabs(cy)) * (fabs(bz) + fabs(cz)) + (fabs(ay) + fabs(cy)) * (fabs(az) + fabs(cz)) + (fabs(ay) + fabs(by)) * (fabs(az) + fabs(bz))));
  if (!(det > bd || -det > bd)) return 2;
  return ((det > 0.0) == pos) ? 1 : 0;
}

// Triangle tests as the walks use them.  Filtered only (EXACT false): a
// test the float64 filter cannot decide sets unsure and counts as no
// contact / no crossing for now (the pair is then redone).  EXACT: such a
// test is decided with exact predicates (counted in nex).  local_ok is
// cleared when a fan triangle's crossing at p + e(q-p) stays unknown; the
// pair then needs its midpoint ray.
template <bool EXACT>
__device__ __forceinline__ int contact_test(const Seg &s, const Tri &t, int &xpar, int &ninc,
                                            bool &local_ok, bool &unsure, unsigned &nex) {
  const int r = contact_fast(s, t, xpar, ninc);
  if (r <= 1) return r;
  if (r == 5) {
    local_ok = false;
    return 0;
  }
  if (r == 4) local_ok = false;
  if (!EXACT) {
    unsure = true;
    return 0;
  }
  ++nex;
  const float a3[3] = {t.a.x, t.a.y, t.a.z}, b3[3] = {t.b.x, t.b.y, t.b.z},
              c3[3] = {t.c.x, t.c.y, t.c.z};
  return contact_exact(s.p, s.q, a3, b3, c3) ? 1 : 0;
}

template <bool EXACT>
__device__ __forceinline__ int cross_test(const Seg &s, const Tri &t, bool &unsure,
                                          unsigned &nex) {
  const int r = cross_fast(s, t);
  if (r != 2) return r;
  if (!EXACT) {
    unsure = true;
    return 0;
  }
  ++nex;
  const float a3[3] = {t.a.x, t.a.y, t.a.z}, b3[3] = {t.b.x, t.b.y, t.b.z},
              c3[3] = {t.c.x, t.c.y, t.c.z};
  return ray_cross_exact(s.p, s.q, a3, b3, c3) ? 1 : 0;
}

__device__ __forceinline__ void load_node(const Node *__restrict__ nodes, int idx, float bx[12],
                                          int &c0, int &c1) {
  const Node nd = nodes[idx];
  bx[0] = nd.f0.x, bx[1] = nd.f0.y, bx[2] = nd.f0.z, bx[3] = nd.f0.w;
  bx[4] = nd.f1.x, bx[5] = nd.f1.y, bx[6] = nd.f1.z, bx[7] = nd.f1.w;
  bx[8] = nd.f2.x, bx[9] = nd.f2.y, bx[10] = nd.f2.z, bx[11] = nd.f2.w;
  c0 = __float_as_int(nd.f3.x);
  c1 = __float_as_int(nd.f3.y);
}

// load_node plus the planar group codes of the children (0: none)
__device__ __forceinline__ void load_node_g(const Node *__restrict__ nodes, int idx, float bx[12],
                                            int &c0, int &c1, int &g0, int &g1) {
  const Node nd = nodes[idx];
  bx[0] = nd.f0.x, bx[1] = nd.f0.y, bx[2] = nd.f0.z, bx[3] = nd.f0.w;
  bx[4] = nd.f1.x, bx[5] = nd.f1.y, bx[6] = nd.f1.z, bx[7] = nd.f1.w;
  bx[8] = nd.f2.x, bx[9] = nd.f2.y, bx[10] = nd.f2.z, bx[11] = nd.f2.w;
  c0 = __float_as_int(nd.f3.x);
  c1 = __float_as_int(nd.f3.y);
  g0 = __float_as_int(nd.f3.z);
  g1 = __float_as_int(nd.f3.w);
}

// A walk split over a group of 2^lv lanes: lane g starts at the node reached
// from the root by the lv bits of g (most significant first), entering a
// child only if its box passes the query; a leaf met above level lv belongs
// to the lane whose remaining bits are zero.  Every leaf whose box chain
// passes is thus reached by exactly one lane.  Returns an internal node id,
// a leaf (~triangle) or kNoRoot.
constexpr int kNoRoot = -2147483647 - 1;

template <bool RAY>
__device__ __forceinline__ int split_root(const Node *__restrict__ nodes, int g, int lv,
                                          const Seg &s) {
  int node = 0;
  for (int l = lv - 1; l >= 0; --l) {
    if (node < 0) return (g & ((2 << l) - 1)) ? kNoRoot : node;
    float bx[12];
    int c0, c1;
    load_node(nodes, node, bx, c0, c1);
    const int o = ((g >> l) & 1) ? 6 : 0;
    const bool h = RAY ? ray_box(s, bx[o], bx[o + 1], bx[o + 2], bx[o + 3], bx[o + 4], bx[o + 5])
                       : seg_box(s, bx[o], bx[o + 1], bx[o + 2], bx[o + 3], bx[o + 4], bx[o + 5]);
    if (!h) return kNoRoot;
    node = o ? c1 : c0;
  }
  return node;
}

// Any contact along the segment: 1 contact, 0 none, 2 none but the local
// parity unknown (see contact_test), 4 none certain yet (filtered walk
// only), 3 malformed tree.  stk: this lane's LDS stack, stride 64 entries.
// root: where the walk starts (a split walk's share); stop: a group flag in
// LDS, set on a contact and polled so the group's other lanes end their
// share early.  grp: planar group table of a contact tree (nullable); a child
// of a group is tested against the group's part of the segment only.
template <int STACK, bool COUNT, bool EXACT, bool GROUPS>
__device__ __forceinline__ int walk_any_t(const Node *__restrict__ nodes, const Tri *__restrict__ tris,
                                          const Seg &s, int *stk, int cap, unsigned &nv,
                                          unsigned &nt, unsigned &nex, int &xpar, int &ninc,
                                          int root, volatile int *stop,
                                          const float4 *__restrict__ grp) {
  bool local_ok = true, unsure = false;
  if (root < 0) {
    if (root == kNoRoot) return 0;
    if (COUNT) ++nt;
    if (contact_test<EXACT>(s, tris[~root], xpar, ninc, local_ok, unsure, nex)) {
      if (stop) *stop = 1;
      return 1;
    }
    return unsure ? 4 : (local_ok ? 0 : 2);
  }
  int node = root, sp = 0;
  int gc = 0;  // group whose span [gt0, gt1] was computed last
  float gt0 = 0.0f, gt1 = 1.0f;
  bool gin = true;
  for (int it = 0;; ++it) {
    if (it > cap) return 3;  // visited more nodes than the tree has: malformed tree
    if (stop && *stop) return 0;  // another lane of the group found a contact
    float bx[12];
    int c0, c1, g0 = 0, g1 = 0;
    if (GROUPS)
      load_node_g(nodes, node, bx, c0, c1, g0, g1);
    else
      load_node(nodes, node, bx, c0, c1);
    if (COUNT) ++nv;
    bool h0, h1;
    if (GROUPS && g0) {
      if (g0 != gc) gc = g0, gin = group_span(s, grp, g0, gt0, gt1);
      h0 = gin && seg_box_span(s, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5], gt0, gt1);
    } else {
      h0 = seg_box(s, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5]);
    }
    if (GROUPS && g1) {
      if (g1 != gc) gc = g1, gin = group_span(s, grp, g1, gt0, gt1);
      h1 = gin && seg_box_span(s, bx[6], bx[7], bx[8], bx[9], bx[10], bx[11], gt0, gt1);
    } else {
      h1 = seg_box(s, bx[6], bx[7], bx[8], bx[9], bx[10], bx[11]);
    }
    if (COUNT) nt += (h0 && c0 < 0) + (h1 && c1 < 0);
    if (h0 && c0 < 0) {
      if (contact_test<EXACT>(s, tris[~c0], xpar, ninc, local_ok, unsure, nex)) {
        if (stop) *stop = 1;
        return 1;
      }
      h0 = false;
    }
    if (h1 && c1 < 0) {
      if (contact_test<EXACT>(s, tris[~c1], xpar, ninc, local_ok, unsure, nex)) {
        if (stop) *stop = 1;
        return 1;
      }
      h1 = false;
    }
    if (h0 && h1) {
      stk[(sp++) * 64] = c1;
      node = c0;
    } else if (h0) {
      node = c0;
    } else if (h1) {
      node = c1;
    } else {
      if (sp == 0) break;
      node = stk[(--sp) * 64];
    }
  }
  return unsure ? 4 : (local_ok ? 0 : 2);
}

// walk_any_t with the planar-group pruning compiled in only when the tree
// has groups (a uniform branch outside the walk loop)
template <int STACK, bool COUNT, bool EXACT>
__device__ __forceinline__ int walk_any(const Node *__restrict__ nodes, const Tri *__restrict__ tris,
                                        const Seg &s, int *stk, int cap, unsigned &nv,
                                        unsigned &nt, unsigned &nex, int &xpar, int &ninc,
                                        int root = 0, volatile int *stop = nullptr,
                                        const float4 *__restrict__ grp = nullptr) {
  if (grp)
    return walk_any_t<STACK, COUNT, EXACT, true>(nodes, tris, s, stk, cap, nv, nt, nex, xpar, ninc,
                                                 root, stop, grp);
  return walk_any_t<STACK, COUNT, EXACT, false>(nodes, tris, s, stk, cap, nv, nt, nex, xpar, ninc,
                                                root, stop, nullptr);
}

// Crossing parity of the midpoint ray: 0 even, 1 odd, 2 undecided (filtered
// walk only), 3 malformed tree.  root as for walk_any.
template <int STACK, bool COUNT, bool EXACT>
__device__ __forceinline__ int walk_parity(const Node *__restrict__ nodes,
                                           const Tri *__restrict__ tris, const Seg &s, int *stk,
                                           int cap, unsigned &nv, unsigned &nt, unsigned &nex,
                                           int root = 0) {
  bool unsure = false;
  if (root < 0) {
    if (root == kNoRoot) return 0;
    if (COUNT) ++nt;
    const int r = cross_test<EXACT>(s, tris[~root], unsure, nex);
    return unsure ? 2 : r;
  }
  int node = root, sp = 0, cnt = 0;
  for (int it = 0;; ++it) {
    if (it > cap) return 3;
    float bx[12];
    int c0, c1;
    load_node(nodes, node, bx, c0, c1);
    if (COUNT) ++nv;
    bool h0 = ray_box(s, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5]);
    bool h1 = ray_box(s, bx[6], bx[7], bx[8], bx[9], bx[10], bx[11]);
    if (COUNT) nt += (h0 && c0 < 0) + (h1 && c1 < 0);
    if (h0 && c0 < 0) {
      cnt += cross_test<EXACT>(s, tris[~c0], unsure, nex);
      h0 = false;
    }
    if (h1 && c1 < 0) {
      cnt += cross_test<EXACT>(s, tris[~c1], unsure, nex);
      h1 = false;
    }
    if (h0 && h1) {
      stk[(sp++) * 64] = c1;
      node = c0;
    } else if (h0) {
      node = c0;
    } else if (h1) {
      node = c1;
    } else {
      if (sp == 0) break;
      node = stk[(--sp) * 64];
    }
  }
  return unsure ? 2 : (cnt & 1);
}

__device__ __forceinline__ float lane_of(const float4 &v, int c) {
  return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
}

// What a pair needs to know about its part (one mesh and its point set).
constexpr int kFanMax = 16;  // triangles kept per point fan

// A fan triangle (p, v1, v2) of point p (rotated to start at p), prepared
// once per point so that its test against a pair (p, q) is linear in
// d = q - p: orient3d(p, v1, v2, q) = -d . N with N = e1 x e2, e1 = v1 - p,
// e2 = v2 - p; the (y, z) orientations of the local crossing test are
// orient2d(p, v1, q) = e1y dz - e1z dy, orient2d(v2, p, q) = e2z dy - e2y dz
// and orient2d(p, v1, v2) = N.x (its sign kept, decided as before).
// Filters: |error| of -d . N is below 9 u sum |d_k| A_k (A_k: the absolute
// products of N_k; u = 2^-53, counting the rounding of e1, e2, d, N and the
// dot product), of the 2-D forms below 5 u (|e1y dz| + |e1z dy|); kFilter
// (1e-14 = 90 u) keeps a tenfold margin.  Undecided tests are left to the
// walks, as before.
struct alignas(16) FanRec {
  double nx, ny, nz, ax, ay, az;
  double e1y, e1z, e2y, e2z;
  float v1[3], v2[3];
  int o0;  // sign of orient2d(p, v1, v2) in (y, z): 1, -1, 0 undecided
  int grp;  // planar group code of the triangle: 0 none, g + 1 for group g
};
static_assert(sizeof(FanRec) == 112, "fan record is 112 bytes");

// Two cones of directions at a point with a complete fan, one around an
// axis a and one around -a, each free of the fan's triangles: the point
// p + e u (small e > 0) then lies on one side of the surface for every u in
// the cone, the side of the axis, so the fan test's crossing parity for
// such a u is the axis's (xp for a, xm for -a).  A unit direction u is in
// the first cone when u . a > tp, in the second when u . a < -tm; tp and tm
// are the cosines of the cones' half-angles (the angular distance of +-a
// from the fan's triangles) shrunk by 1e-3 rad, plus 1e-5: far more than
// the float32 error of u . a (below 1e-6).  tp = tm = 2 disables a cone.
// A pair whose direction falls in a cone skips its fan tests at that end.
struct alignas(16) FanCone {
  float ax, ay, az, tp;
  float tm;
  int xp, xm, pad;
};

// Crossing parity of the fan test for direction d from the point, from its
// cones; -1 when d is in neither.
__device__ __forceinline__ int cone_parity(const FanCone &c, float dx, float dy, float dz) {
  const float dd = dx * dx + dy * dy + dz * dz;
  if (!(dd > 1e-30f && dd < 1e30f)) return -1;
  const float cs = (dx * c.ax + dy * c.ay + dz * c.az) * rsqrtf(dd);
  if (cs > c.tp) return c.xp;
  if (cs < -c.tm) return c.xm;
  return -1;
}

// Cosine of the angle between unit u and the cone {l1 e1 + l2 e2 : l1, l2 >= 0}
// (the nearer edge ray's, or the plane's when u projects inside the cone).
__device__ __forceinline__ double sector_cos(const double u[3], const double e1[3],
                                             const double e2[3]) {
  const double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
  const double l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
  const double c1 = (u[0] * e1[0] + u[1] * e1[1] + u[2] * e1[2]) / l1;
  const double c2 = (u[0] * e2[0] + u[1] * e2[1] + u[2] * e2[2]) / l2;
  double best = fmin(1.0, fmax(-1.0, fmax(c1, c2)));  // nearer edge ray
  const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                       e1[0] * e2[1] - e1[1] * e2[0]};
  const double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (ln > 1e-12 * l1 * l2) {
    const double un = (u[0] * n[0] + u[1] * n[1] + u[2] * n[2]) / ln;
    const double w[3] = {u[0] - un * n[0] / ln, u[1] - un * n[1] / ln, u[2] - un * n[2] / ln};
    // w inside the sector: on e2's side of e1 and on e1's side of e2
    const double s1 = (e1[1] * w[2] - e1[2] * w[1]) * n[0] + (e1[2] * w[0] - e1[0] * w[2]) * n[1] +
                      (e1[0] * w[1] - e1[1] * w[0]) * n[2];
    const double s2 = (w[1] * e2[2] - w[2] * e2[1]) * n[0] + (w[2] * e2[0] - w[0] * e2[2]) * n[1] +
                      (w[0] * e2[1] - w[1] * e2[0]) * n[2];
    if (s1 >= 0.0 && s2 >= 0.0) best = fmax(best, sqrt(fmax(0.0, 1.0 - un * un)));
  }
  return best;
}

// cos(theta - 1e-3) + 1e-5 for a cone whose half-angle theta has cosine cm
// (2: no cone when theta <= 2e-3 or theta >= 3.1).
__device__ __forceinline__ float cone_threshold(double cm) {
  if (!(cm < 0.999998 && cm > -0.99913)) return 2.0f;  // cos(2e-3), cos(3.1)
  const double sm = sqrt(fmax(0.0, 1.0 - cm * cm));
  return (float)(cm * 0.9999995000000417 + sm * 9.999998333333417e-4 + 1e-5);  // cos, sin of 1e-3
}

// contact_fast for a fan triangle of the pair's first point, from its
// record: 1 the segment is an edge of the triangle, 0 no contact (xpar
// toggled by a crossing at p + e d, ninc counted), 4 contact undecided,
// 5 crossing undecided.  other: the pair's other point; d = other - p.
__device__ __forceinline__ int fan_test(const FanRec &r, const float other[3], const double d[3],
                                        int &xpar, int &ninc) {
  const float w1[3] = {r.v1[0], r.v1[1], r.v1[2]}, w2[3] = {r.v2[0], r.v2[1], r.v2[2]};
  if (same3(other, w1) || same3(other, w2)) return 1;
  const double det = -(d[0] * r.nx + d[1] * r.ny + d[2] * r.nz);
  const double bd = kFilter * (fabs(d[0]) * r.ax + fabs(d[1]) * r.ay + fabs(d[2]) * r.az);
  if (!(det > bd || -det > bd)) return 4;
  ++ninc;
  const double o1 = r.e1y * d[2] - r.e1z * d[1];
  const double b1 = kFilter * (fabs(r.e1y * d[2]) + fabs(r.e1z * d[1]));
  const double o2 = r.e2z * d[1] - r.e2y * d[2];
  const double b2 = kFilter * (fabs(r.e2z * d[1]) + fabs(r.e2y * d[2]));
  if (r.o0 == 0 || !(o1 > b1 || -o1 > b1) || !(o2 > b2 || -o2 > b2)) return 5;
  const bool s0 = r.o0 > 0;
  if ((o1 > 0.0) == s0 && (o2 > 0.0) == s0 && (det > 0.0) == s0) xpar ^= 1;
  return 0;
}

struct Part {
  const float *pts;
  const uint8_t *mask;
  const uint8_t *far;  // per point: crossing parity of its +x ray with non-incident
                       // triangles (0/1, 2 = undecided); NULL unless the mesh is closed
  const int *fan_n;    // per point: triangles having the point as a vertex (-1: more
                       // than kFanMax); NULL together with far
  const int *fan_idx;  // per point: kFanMax leaf-ordered triangle ids
  const FanRec *fan_rec;  // per point: kFanMax prepared fan triangles
  const FanCone *fan_cone;  // per point: the fan's cones
  const float *clear;  // per point: a lower bound of its distance to every triangle
                       // that does not have it as a vertex (NULL: not computed)
  float clear_r;       // radius of the clearance search (clearances are at most this)
  const Node *snodes;
  const Node4 *snodes4;
  const Tri *stris;
  const Node *rnodes;  // self contact tree (slices of long triangles; else = snodes)
  const Tri *rtris;
  const float4 *rgroups;  // planar groups of the contact tree (NULL: none)
  const Node *cnodes;  // cage (NULL: parity mode)
  const Tri *ctris;
  uint32_t group_axis;  // per planar group of the self mesh, its projection axis (vacd_bvh)
  int64_t n, S;
  int64_t word_begin;   // first ballot word of the part in the batch
  int64_t point_begin;  // first point of the part in the batch
  int scap, ccap, rcap;  // node-visit caps (2 * primitives + 64)
};

struct VisArgs {
  Part one;            // the part of a single-part call
  const Part *parts;   // batch: parts sorted by word_begin
  int nparts;
  int64_t nwords, npoints;
  uint64_t *bits;
  unsigned long long *stats;  // [0] masked, [1] blocked, [2] pairs redone exactly, [3] edges,
                              // [4] malformed, [5] node visits, [6] triangle tests,
                              // [7] walk-queue length, [8] deferred-list length,
                              // [9] exact triangle tests; work counters of
                              // the fan pass (COUNT): pairs left undecided
                              // with [10] no usable fan at p, [11] p's fan
                              // test undecided, [12] no usable fan at q,
                              // [13] q's undecided, [14] both ends leaving
                              // outwards
  int64_t *queue;             // global pair ids (word * 64 + bit) left to walk
  unsigned long long *queue_len;
  uint64_t *qbits;            // per ballot word: its pairs left to walk (compacted into
                              // queue by a scan, in k order: no atomics in the fan pass)
  uint64_t *cbits;            // per ballot word: queued pairs whose fans show both ends
                              // leaving outwards without touching either fan (their
                              // queue entries carry kClip)
  int64_t *dlist;             // global pair ids whose walks met an undecided test
  unsigned long long *dlist_len;
  int64_t front_max;          // walk queues up to this long go to k_walk_front
  // which kernel redoes the first kExactSplit deferred entries a wave per
  // pair: true, k_walk_front (breadth first; it marks what it settles);
  // false, k_walk_split with LV = 6 (depth first).  The lane-per-pair
  // instance of k_walk_split takes the rest, or every unsettled entry.
  bool exact_front = false;
  bool rotated_rays = false;  // undecided +x midpoint rays tried along +y, +z
  int fan_chunk = 1;          // consecutive words a wave of the fan pass takes at a time
};

// Part owning global ballot word w (binary search over word_begin).
template <bool BATCH>
__device__ __forceinline__ Part part_of_word(const VisArgs &a, int64_t w) {
  if (!BATCH) return a.one;
  int lo = 0, hi = a.nparts - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.parts[mid].word_begin <= w) lo = mid;
    else hi = mid - 1;
  }
  return a.parts[lo];
}

template <int STACK, bool COUNT, bool EXACT>
__device__ __forceinline__ int any_hit(const Node *n2, const Tri *tris, const Seg &s, int *stk,
                                       int cap, unsigned &nv, unsigned &nt, unsigned &nex,
                                       int &xpar, int &ninc, const float4 *grp = nullptr) {
  return walk_any<STACK, COUNT, EXACT>(n2, tris, s, stk, cap, nv, nt, nex, xpar, ninc, 0, nullptr,
                                       grp);
}

// Walks for one pair the fan could not settle: any contact with the self
// mesh, then the cage walk (cage mode) or the inside parity (parity mode):
// local when p's fan crossings are all known, else the midpoint ray.  A
// filtered walk (EXACT false) that meets an undecided test sets defer.
template <int STACK, bool CAGE, bool COUNT, bool EXACT>
__device__ __forceinline__ void walk_pair(const Part &P, int64_t i, Seg &s, int *stk, bool &edge,
                                          bool &defer, unsigned &n_blocked, unsigned &n_bad,
                                          unsigned &nv, unsigned &nt, unsigned &nex) {
  make_seg(s);
  int xpar = 0, ninc = 0;
  // contacts: the contact tree
  const int hs = any_hit<STACK, COUNT, EXACT>(P.rnodes, P.rtris, s, stk, P.rcap, nv, nt, nex, xpar,
                                              ninc, P.rgroups);
  if (hs == 1) {
    ++n_blocked;
  } else if (hs == 3) {
    ++n_bad;
  } else if (hs == 4) {
    defer = true;
  } else if (CAGE && P.ctris) {
    int cx = 0, cn = 0;
    const int hc = any_hit<STACK, COUNT, EXACT>(P.cnodes, P.ctris, s, stk, P.ccap, nv, nt, nex, cx,
                                                cn);
    edge = hc == 1;
    defer = hc == 4;
    n_bad += hc == 3;
  } else if (hs == 0 && P.far && ninc > 0 && P.far[i] <= 1) {
    // closed mesh, p on it: inside parity = fan crossings ^ far crossings of p
    edge = ((xpar ^ P.far[i]) & 1) == 0;
  } else {  // no local parity for this pair: the midpoint ray itself
    const int par = walk_parity<STACK, COUNT, EXACT>(P.snodes, P.stris, s, stk, P.scap, nv, nt, nex);
    edge = par == 0;
    defer = par == 2;
    n_bad += par == 3;
  }
}

// Mask of a point: bit 0 new (lib/ptx/ray_segments.cu:61-75: a pair is
// skipped when neither point is new); bit 1 + g: the point lies on the
// plane of planar group g of the self mesh.
__device__ __forceinline__ bool mask_skips(uint8_t a, uint8_t b) { return !((a | b) & 1); }

// Whether direction q - p enters, at p, the closed 2-D angle of a triangle
// of p's fan that belongs to one of the planar groups in `groups` (bit g:
// group g), in the group plane's projection (exact orientations).  Both
// points lie on that plane, so the segment then runs into the group's
// triangles (the cap a cut left) and touches the surface.  Only the exact
// redo calls it (out of line: it would cost the fan pass registers).
__device__ __noinline__ bool enters_cap(const Part &P, int64_t i, const float p[3],
                                        const float q[3], unsigned groups) {
  const int fn = P.fan_n[i];
  const FanRec *fr = P.fan_rec + (size_t)i * kFanMax;
  for (int f = 0; f < fn; ++f) {
    const int code = fr[f].grp;
    if (!code || !((groups >> (code - 1)) & 1u)) continue;
    const int k = (int)((P.group_axis >> (2 * (code - 1))) & 3u);
    if (k == 3) continue;
    const float v1[3] = {fr[f].v1[0], fr[f].v1[1], fr[f].v1[2]};
    const float v2[3] = {fr[f].v2[0], fr[f].v2[1], fr[f].v2[2]};
    const int o = orient2d_exact(p, v1, v2, k);
    if (o == 0) continue;
    const int o1 = orient2d_exact(p, v1, q, k), o2 = orient2d_exact(p, q, v2, k);
    if ((o1 != 0 || o2 != 0) && o * o1 >= 0 && o * o2 >= 0) return true;
  }
  return false;
}

// The same test with float64 filters only, for the fan pass: 1 enters a
// cap triangle (decided), 0 enters none (every sign decided), -1 a sign the
// filter leaves open and no triangle decided "enters" (the exact redo then
// decides the pair with enters_cap).  A zero error bound means both
// products are exactly zero, so the sign (zero) is exact.
__device__ __forceinline__ int enters_cap_filtered(const Part &P, int64_t i, const float p[3],
                                                   const float q[3], unsigned groups) {
  const int fn = P.fan_n[i];
  const FanRec *fr = P.fan_rec + (size_t)i * kFanMax;
  int r = 0;
  for (int f = 0; f < fn; ++f) {
    const int code = fr[f].grp;
    if (!code || !((groups >> (code - 1)) & 1u)) continue;
    const int k = (int)((P.group_axis >> (2 * (code - 1))) & 3u);
    if (k == 3) continue;
    const float v1[3] = {fr[f].v1[0], fr[f].v1[1], fr[f].v1[2]};
    const float v2[3] = {fr[f].v2[0], fr[f].v2[1], fr[f].v2[2]};
    double b0, b1, b2;
    const double d0 = o2_fast(p, v1, v2, k, b0);
    const double d1 = o2_fast(p, v1, q, k, b1);
    const double d2 = o2_fast(p, q, v2, k, b2);
    const bool sure = (fabs(d0) > b0 || b0 == 0.0) && (fabs(d1) > b1 || b1 == 0.0) &&
                      (fabs(d2) > b2 || b2 == 0.0);
    if (!sure) {
      r = -1;
      continue;
    }
    const int o = d0 > 0.0 ? 1 : (d0 < 0.0 ? -1 : 0);
    if (o == 0) continue;
    const int o1 = d1 > 0.0 ? 1 : (d1 < 0.0 ? -1 : 0), o2 = d2 > 0.0 ? 1 : (d2 < 0.0 ? -1 : 0);
    if ((o1 != 0 || o2 != 0) && o * o1 >= 0 && o * o2 >= 0) return 1;
  }
  return r;
}

// Appends `value` of every lane with `mine` set to a list, in lane order,
// with one atomic per wave.
__device__ __forceinline__ void append_lanes(bool mine, int64_t value, int64_t *list,
                                             unsigned long long *len) {
  const uint64_t m = __ballot(mine);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(len, (unsigned long long)__popcll(m));
  base = __shfl(base, 0);
  if (mine) list[base + __popcll(m & (lane ? (~0ull >> (64 - lane)) : 0ull))] = value;
}

// CAP: the mask may carry cap-rule bits (its filtered 2-D tests are
// compiled only where needed: they cost the fan pass registers).
template <int STACK, bool CAGE, bool COUNT, bool BATCH, bool QUEUE, bool CAP = false>
__global__ __launch_bounds__(kBlock) void k_visibility(VisArgs a) {
  __shared__ int stk_all[kWaves][STACK][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int *stk = &stk_all[wave][0][lane];
  unsigned n_masked = 0, n_blocked = 0, n_exact = 0, n_bad = 0, nv = 0, nt = 0, n_walk = 0;
  (void)n_exact;  // filtered walks make no exact tests
  // work counters (COUNT): undecided pairs by the state of each end's fan
  // test
  unsigned n_pno = 0, n_pund = 0, n_qno = 0, n_qund = 0, n_out = 0;
  const int64_t total_waves = (int64_t)gridDim.x * kWaves;
  // Each wave takes chunks of a.fan_chunk consecutive words, the chunks strided
  // over the grid.  A lane's pair in the next word of the same part is its
  // pair 64 places further along the rows, found by stepping (a few integer
  // operations) instead of the closed form (float64 square root and row
  // corrections: measured at 0.5 ms of the 16,384-point pass when done once
  // more per word).
  const int chunk_w = a.fan_chunk;  // chosen with the grid on the host
  int64_t ci = -1, cj = 0, cbeg = -1;  // this lane's last pair and that part's first word
  for (int64_t chunk = (int64_t)blockIdx.x * kWaves + wave; chunk * chunk_w < a.nwords;
       chunk += total_waves) {
  for (int cw = 0; cw < chunk_w; ++cw) {
    const int64_t w = chunk * chunk_w + cw;
    if (w >= a.nwords) break;  // the last chunk's tail
    const Part P = part_of_word<BATCH>(a, w);  // wave-uniform
    const int64_t k = (w - P.word_begin) * 64 + lane;
    bool edge = false, defer = false, queued = false;
    bool pout = false, qout = false;  // the end leaves outwards, its fan untouched (certain)
    if (k < P.S) {
      int64_t i, j;
      if (cw != 0 && cbeg == P.word_begin && ci >= 0) {
        // k - 64 was (ci, cj): row i holds j = i + 1 .. n - 1, then row i + 1
        // starts at j = i + 2
        i = ci;
        j = cj + 64;
        while (j >= P.n) {
          ++i;
          j = j - P.n + i + 1;
        }
      } else {
        pair_of(k, P.n, i, j);
      }
      ci = i, cj = j, cbeg = P.word_begin;
      if (P.mask && mask_skips(P.mask[i], P.mask[j])) {
        ++n_masked;
      } else {
        Seg s;
#pragma unroll
        for (int c = 0; c < 3; ++c) s.p[c] = P.pts[3 * i + c], s.q[c] = P.pts[3 * j + c];
        bool decided = same3(s.p, s.q);  // zero length: not an edge
        int pstate = 2, qstate = 2;  // work counters: 1 outwards, 2 no fan, 3 undecided
        if (CAP && !decided && P.fan_n && !(CAGE && P.ctris) && P.mask) {
          // both points on one planar group's plane (a cut's cap): a pair
          // that runs into the cap at either end touches it
          const unsigned shared = (unsigned)(P.mask[i] & P.mask[j]) >> 1;
          if (shared) {
            int c = enters_cap_filtered(P, i, s.p, s.q, shared);
            if (c != 1) {
              const int cq = enters_cap_filtered(P, j, s.q, s.p, shared);
              c = cq == 1 ? 1 : ((c | cq) < 0 ? -1 : 0);
            }
            if (c == 1) {
              ++n_blocked;
              decided = true;
            } else if (c < 0) {  // an open sign: the exact redo decides the pair
              defer = true;
              decided = true;
            }
          }
        }
        if (!decided && P.fan_n && !(CAGE && P.ctris)) {
          // A pair that leaves p into the solid is blocked or wholly inside:
          // never an edge.  Decide that from p's fan alone (every lane of a
          // row shares p, so the fan loads are broadcasts).
          const int fn = P.fan_n[i];
          if (fn > 0) {
            int xp = cone_parity(P.fan_cone[i], s.q[0] - s.p[0], s.q[1] - s.p[1],
                                 s.q[2] - s.p[2]),
                ni = 0, r = 0;
            if (xp < 0) {  // the direction is near the tangent plane: p's fan tests
              xp = 0;
              const double dq[3] = {(double)s.q[0] - s.p[0], (double)s.q[1] - s.p[1],
                                    (double)s.q[2] - s.p[2]};
              const FanRec *fr = P.fan_rec + (size_t)i * kFanMax;
              for (int f = 0; f < fn && r == 0; ++f) r = fan_test(fr[f], s.q, dq, xp, ni);
            }
            if (r == 1) {  // the segment is an edge of the mesh
              ++n_blocked;
              decided = true;
            } else if (r == 0 && P.far[i] <= 1 && ((xp ^ P.far[i]) & 1)) {
              decided = true;  // starts inside
            }
            pout = r == 0 && P.far[i] <= 1;
            if (COUNT) pstate = pout ? 1 : 3;
          }
          // The same from q's end: a pair that enters the solid at q (the
          // segment seen from q towards p starts inside) is blocked or
          // wholly inside too.  Most pairs that leave p outwards and are not
          // edges pass through the solid and are settled here, so only
          // pairs that leave both ends outwards are walked.
          const int fq = decided ? 0 : P.fan_n[j];
          if (fq > 0) {
            int xq = cone_parity(P.fan_cone[j], s.p[0] - s.q[0], s.p[1] - s.q[1],
                                 s.p[2] - s.q[2]),
                nq = 0, r = 0;
            if (xq < 0) {
              xq = 0;
              const double dp[3] = {(double)s.p[0] - s.q[0], (double)s.p[1] - s.q[1],
                                    (double)s.p[2] - s.q[2]};
              const FanRec *fr = P.fan_rec + (size_t)j * kFanMax;
              for (int f = 0; f < fq && r == 0; ++f) r = fan_test(fr[f], s.p, dp, xq, nq);
            }
            if (r == 1) {
              ++n_blocked;
              decided = true;
            } else if (r == 0 && P.far[j] <= 1 && ((xq ^ P.far[j]) & 1)) {
              decided = true;  // enters the solid at q
            }
            qout = r == 0 && P.far[j] <= 1;
            if (COUNT) qstate = qout ? 1 : 3;
          }
        }
        if (COUNT && !decided) {
          n_pno += pstate == 2, n_pund += pstate == 3, n_qno += qstate == 2, n_qund += qstate == 3;
          n_out += pstate == 1 && qstate == 1;
        }
        if (!decided) {
          if constexpr (QUEUE) {
            queued = true;
          } else {
            ++n_walk;
            walk_pair<STACK, CAGE, COUNT, false>(P, i, s, stk, edge, defer, n_blocked, n_bad,
                                                       nv, nt, n_exact);
          }
        }
      }
    }
    const uint64_t eb = __ballot(edge);
    if (lane == 0) a.bits[w] = eb;
    append_lanes(defer, w * 64 + lane, a.dlist, a.dlist_len);
    if (QUEUE) {  // this wave's undecided pairs, for the walk queue
      const uint64_t qb = __ballot(queued);
      // queued pairs that leave both ends outwards (the fan tests of both
      // ends decided, no contact): their walks may skip the clearance balls
      const uint64_t cb = __ballot(queued && pout && qout && P.clear != nullptr);
      if (lane == 0) a.qbits[w] = qb, a.cbits[w] = cb;
    }
  }
  }
  // per-block totals, one atomic per block and counter (atomics on one
  // address from every wave of the grid serialize)
  for (int off = 32; off > 0; off >>= 1) {
    n_masked += __shfl_down(n_masked, off);
    n_blocked += __shfl_down(n_blocked, off);
    n_bad += __shfl_down(n_bad, off);
    n_walk += __shfl_down(n_walk, off);
  }
  __shared__ unsigned wave_tot[kWaves][4];
  if (lane == 0) {
    wave_tot[wave][0] = n_masked, wave_tot[wave][1] = n_blocked;
    wave_tot[wave][2] = n_bad, wave_tot[wave][3] = n_walk;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long t = 0;
    for (int v = 0; v < kWaves; ++v) t += wave_tot[v][threadIdx.x];
    // stats slots: masked, blocked, malformed, walked (queue mode: queue_len)
    const int slot = threadIdx.x == 0 ? 0 : (threadIdx.x == 1 ? 1 : (threadIdx.x == 2 ? 4 : 7));
    if (t) atomicAdd(&a.stats[slot], t);
  }
  if (COUNT) {
    unsigned long long v = nv, t = nt;
    unsigned c[5] = {n_pno, n_pund, n_qno, n_qund, n_out};
    for (int off = 32; off > 0; off >>= 1) {
      v += __shfl_down(v, off);
      t += __shfl_down(t, off);
#pragma unroll
      for (int k = 0; k < 5; ++k) c[k] += __shfl_down(c[k], off);
    }
    if (lane == 0) {
      atomicAdd(&a.stats[5], v);
      atomicAdd(&a.stats[6], t);
#pragma unroll
      for (int k = 0; k < 5; ++k)
        if (c[k]) atomicAdd(&a.stats[10 + k], (unsigned long long)c[k]);
    }
  }
}

// Dense walks over a list of pairs, 2^LV lanes per pair: each lane walks its
// share of the tree (split_root), the group combines its lanes' results
// through ballots (any contact: OR; fan crossings and ray crossings: XOR),
// and the group's first lane OR-s the outcome into the ballot words (zero
// for these bits until then).  More lanes per pair shorten the longest walk
// of a wave at the price of box tests near the root done by several lanes.
// EXACT false: the walk queue the fan pass left, filtered tests; a pair
// meeting an undecided test goes to the deferred list.  EXACT true: the
// deferred list, every test decided (exact predicates where the filter
// cannot decide); with LV = 6 (a wave per pair, for latency) the first
// kExactSplit entries, with LV = 0 (a lane per pair) the rest.  Only the
// EXACT instances carry the exact code's register cost.
__device__ __forceinline__ uint64_t group_mask(int lane, int G) {
  return G == 64 ? ~0ull : (((1ull << G) - 1ull) << (lane & ~(G - 1)));
}

// Deferred pairs are few (tens on one part, ~10^4 in a batched round of
// cut parts) and their walks long: a lane per pair leaves most SIMDs idle
// (batched rounds of 16,384-point parts: 9.5 ms per redo), a wave per pair
// fills the chip.  So the wave-per-pair kernels take up to kExactSplit
// entries, looping over them with a grid of kExactGrid blocks.
constexpr int64_t kExactSplit = int64_t(1) << 20;
constexpr unsigned kExactGrid = 512;
// Deferred-list entries the breadth-first exact walk settled carry this bit
// (the depth-first exact kernels skip them; the list still names every pair
// redone exactly).
constexpr int64_t kSettled = int64_t(1) << 62;
// Walk-queue entry flag: both ends of the pair leave outwards and neither
// fan touches the segment, so the pair is an edge iff no other triangle
// touches it, and those lie outside both endpoints' clearance balls.
// waves per SIMD of the breadth-first walk of the queue (its register budget:
// 512 / kWalkWaves VGPRs); the walk is bound by the latency of dependent node
// loads, so more resident waves hide more of it
#ifndef VACD_WALK_WAVES
#define VACD_WALK_WAVES 4
#endif
constexpr int kWalkWaves = VACD_WALK_WAVES;
// longest run of consecutive words a wave of the fan pass takes at a time
#ifndef VACD_FAN_CHUNK_MAX
#define VACD_FAN_CHUNK_MAX 8
#endif
constexpr int64_t kClip = int64_t(1) << 61;

template <int STACK, bool COUNT, bool BATCH, int LV, bool EXACT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(EXACT ? 1 : 4))) void k_walk_split(VisArgs a) {
  constexpr int G = 1 << LV;
  constexpr int kGroups = kBlock / G;
  __shared__ int stk_all[kWaves][STACK][64];
  __shared__ int stop_all[kGroups];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int *stk = &stk_all[wave][0][lane];
  const int g = lane & (G - 1);
  const int slot = threadIdx.x / G;
  const uint64_t gm = group_mask(lane, G);
  unsigned n_blocked = 0, n_exact = 0, n_bad = 0, nv = 0, nt = 0;
  const int64_t *list = EXACT ? a.dlist : a.queue;
  const int64_t len = (int64_t)(EXACT ? *a.dlist_len : *a.queue_len);
  if (!EXACT && len <= a.front_max) return;  // k_walk_front took this queue
  // with the breadth-first exact walk (a.exact_front) the lane-per-pair
  // instance takes every entry it left (kSettled clear), the wave-per-pair
  // one is not launched
  const int64_t e_lo = (EXACT && LV == 0 && !a.exact_front) ? kExactSplit : 0;
  const int64_t e_hi = (EXACT && LV == 6) ? (len < kExactSplit ? len : kExactSplit) : len;
  for (int64_t e = e_lo + (int64_t)blockIdx.x * kGroups + slot; e < e_hi;
       e += (int64_t)gridDim.x * kGroups) {
    // (a queue entry's kClip flag is not used here: these walks cover the
    // whole segment)
    const int64_t kg = EXACT ? list[e] : (list[e] & ~kClip);
    if (EXACT && (kg & kSettled)) continue;  // settled by the breadth-first exact walk
    const int64_t w = kg >> 6;
    const int b = (int)(kg & 63);
    const Part P = part_of_word<BATCH>(a, w);
    int64_t i, j;
    pair_of((w - P.word_begin) * 64 + b, P.n, i, j);
    Seg s;
#pragma unroll
    for (int c = 0; c < 3; ++c) s.p[c] = P.pts[3 * i + c], s.q[c] = P.pts[3 * j + c];
    make_seg(s);
    volatile int *stop = &stop_all[slot];
    if (g == 0) *stop = 0;
    int xpar = 0, ninc = 0;
    const int root = split_root<false>(P.rnodes, g, LV, s);
    const int hs = walk_any<STACK, COUNT, EXACT>(P.rnodes, P.rtris, s, stk, P.rcap, nv, nt,
                                                 n_exact, xpar, ninc, root,
                                                 G > 1 ? stop : nullptr, P.rgroups);
    const uint64_t m3 = __ballot(hs == 3) & gm, m1 = __ballot(hs == 1) & gm,
                   m2 = __ballot(hs == 2) & gm, m4 = __ballot(hs == 4) & gm,
                   mx = __ballot(xpar & 1) & gm, mn = __ballot(ninc > 0) & gm;
    bool edge = false, blocked = false, bad = false, defer = false;
    if (m3) {
      bad = true;
    } else if (m1) {
      blocked = true;
    } else if (m4) {
      defer = true;
    } else if (P.ctris) {  // cage mode (deferred cage pairs only): any contact with the cage
      if (g == 0) *stop = 0;
      int cx = 0, cn = 0;
      const int rc = split_root<false>(P.cnodes, g, LV, s);
      const int hc = walk_any<STACK, COUNT, EXACT>(P.cnodes, P.ctris, s, stk, P.ccap, nv, nt,
                                                   n_exact, cx, cn, rc, G > 1 ? stop : nullptr);
      const uint64_t c3 = __ballot(hc == 3) & gm, c1 = __ballot(hc == 1) & gm,
                     c4 = __ballot(hc == 4) & gm;
      if (c3)
        bad = true;
      else if (c1)
        edge = true;
      else if (c4)
        defer = true;
    } else if (!m2 && P.far && mn && P.far[i] <= 1) {
      // closed mesh, p on it: inside parity = fan crossings ^ far crossings of p
      edge = (((__popcll(mx) & 1) ^ P.far[i]) & 1) == 0;
    } else {  // group-uniform branch: the midpoint ray, split the same way
      const int rr = split_root<true>(P.snodes, g, LV, s);
      const int par =
          walk_parity<STACK, COUNT, EXACT>(P.snodes, P.stris, s, stk, P.scap, nv, nt, n_exact, rr);
      const uint64_t p3 = __ballot(par == 3) & gm, p2 = __ballot(par == 2) & gm,
                     p1 = __ballot(par == 1) & gm;
      if (p3)
        bad = true;
      else if (p2)
        defer = true;
      else
        edge = (__popcll(p1) & 1) == 0;
    }
    if (g == 0) {
      if (edge) atomicOr(reinterpret_cast<unsigned long long *>(&a.bits[w]), 1ull << b);
      if (!EXACT && defer) a.dlist[atomicAdd(a.dlist_len, 1ull)] = kg;
      n_blocked += blocked;
      n_bad += bad;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    n_blocked += __shfl_down(n_blocked, off);
    n_bad += __shfl_down(n_bad, off);
    n_exact += __shfl_down(n_exact, off);
    nv += __shfl_down(nv, off);
    nt += __shfl_down(nt, off);
  }
  if (lane == 0) {
    if (n_blocked) atomicAdd(&a.stats[1], (unsigned long long)n_blocked);
    if (n_bad) atomicAdd(&a.stats[4], (unsigned long long)n_bad);
    if (EXACT && n_exact) atomicAdd(&a.stats[9], (unsigned long long)n_exact);
    if (COUNT) {
      atomicAdd(&a.stats[5], (unsigned long long)nv);
      atomicAdd(&a.stats[6], (unsigned long long)nt);
    }
  }
}

// Breadth-first walks of a short walk queue: G lanes per pair (a group),
// the pair's frontier of tree nodes in LDS.  At each level the group's lanes
// take the frontier's nodes, test both children, test the triangles of the
// leaf children, and append the internal children that pass to the next
// frontier (ballot offsets).  A walk then takes about as many dependent
// steps as the tree is deep, instead of one per node visited: a short queue
// does not fill the chip, and its time is the latency of its longest walks
// (edges, which cannot stop early, visit a few hundred nodes).  Outcome as
// in walk_pair: the any-contact walk, then the local parity, or the
// midpoint ray (breadth first as well) when the local parity is unknown; a
// pair whose tests were undecided or whose frontier would overflow goes to
// the deferred list (the exact redo decides it).  Only queues of at most
// a.front_max pairs are taken; k_walk_split walks longer ones.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The cyclic change of coordinates (x, y, z) -> (y, z, x): a rotation, so
// orientations keep their signs, and the +x ray of the rotated problem is
// the +y ray of the original one.
__device__ __forceinline__ Tri rotate_tri(const Tri &t) {
  Tri r;
  r.a = make_float4(t.a.y, t.a.z, t.a.x, t.a.w);
  r.b = make_float4(t.b.y, t.b.z, t.b.x, t.b.w);
  r.c = make_float4(t.c.y, t.c.z, t.c.x, t.c.w);
  return r;
}

__device__ __forceinline__ Seg rotate_seg(const Seg &s) {
  Seg r;
  r.p[0] = s.p[1], r.p[1] = s.p[2], r.p[2] = s.p[0];
  r.q[0] = s.q[1], r.q[1] = s.q[2], r.q[2] = s.q[0];
  make_seg(r);
  return r;
}

// One breadth-first walk by a group of G lanes (cur/nxt: the group's LDS
// frontier buffers of KF entries).  RAY false: any contact of the segment
// with the tree's triangles (the contact tree, planar groups pruned), with
// p's fan crossings met on the way in xpar/ninc/local_ok as in walk_any.
// RAY true: crossings of the +x ray from the midpoint (plain tree), parity
// in xpar.  Per-lane results; returns (group-uniform) 0 done, 1 contact,
// 3 more visits than the tree has nodes, 5 frontier overflow.  EXACT: tests
// the float64 filter cannot decide are decided exactly (counted in nex)
// instead of setting unsure.  [u0, u1]: the segment's parameter range whose
// boxes are tested (RAY false; a triangle met is still tested whole).
template <int G, int KF, bool RAY, bool COUNT, bool EXACT = false, int ROT = 0>
__device__ __forceinline__ int front_walk(const Node *__restrict__ nodes,
                                          const Tri *__restrict__ tris,
                                          const float4 *__restrict__ grp, int cap, const Seg &s,
                                          int *cur, int *nxt, int g, uint64_t gm, uint64_t below,
                                          int &xpar, int &ninc, bool &local_ok, bool &unsure,
                                          unsigned &nv, unsigned &nt, unsigned &nex,
                                          float u0 = 0.0f, float u1 = 1.0f) {
  if (g == 0) cur[0] = 0;  // the root
  wave_lds_sync();
  int n = 1;
  int64_t visits = 0;
  while (n > 0) {
    visits += n;
    if (visits > cap) return 3;
    int nn = 0;
    for (int base = 0; base < n; base += G) {
      const int idx = base + g;
      bool w0 = false, w1 = false, hit = false;
      int c0 = 0, c1 = 0;
      if (idx < n) {
        float bx[12];
        int g0 = 0, g1 = 0;
        if (!RAY && grp) load_node_g(nodes, cur[idx], bx, c0, c1, g0, g1);
        else load_node(nodes, cur[idx], bx, c0, c1);
        if (COUNT) ++nv;
        bool h0, h1;
        if (RAY && ROT == 1) {  // boxes in the rotated frame (x, y, z) -> (y, z, x)
          h0 = ray_box(s, bx[1], bx[2], bx[0], bx[4], bx[5], bx[3]);
          h1 = ray_box(s, bx[7], bx[8], bx[6], bx[10], bx[11], bx[9]);
        } else if (RAY && ROT == 2) {  // (x, y, z) -> (z, x, y)
          h0 = ray_box(s, bx[2], bx[0], bx[1], bx[5], bx[3], bx[4]);
          h1 = ray_box(s, bx[8], bx[6], bx[7], bx[11], bx[9], bx[10]);
        } else if (RAY) {
          h0 = ray_box(s, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5]);
          h1 = ray_box(s, bx[6], bx[7], bx[8], bx[9], bx[10], bx[11]);
        } else {
          float t0, t1;
          h0 = g0 ? (group_span(s, grp, g0, t0, t1) &&
                     seg_box_span(s, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5], fmaxf(t0, u0),
                                  fminf(t1, u1)))
                  : seg_box_span(s, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5], u0, u1);
          h1 = g1 ? (group_span(s, grp, g1, t0, t1) &&
                     seg_box_span(s, bx[6], bx[7], bx[8], bx[9], bx[10], bx[11], fmaxf(t0, u0),
                                  fminf(t1, u1)))
                  : seg_box_span(s, bx[6], bx[7], bx[8], bx[9], bx[10], bx[11], u0, u1);
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const int ch = c ? c1 : c0;
          if (!(c ? h1 : h0) || ch >= 0) continue;
          if (COUNT) ++nt;
          if (RAY && ROT == 1) xpar ^= cross_test<EXACT>(s, rotate_tri(tris[~ch]), unsure, nex);
          else if (RAY && ROT == 2)
            xpar ^= cross_test<EXACT>(s, rotate_tri(rotate_tri(tris[~ch])), unsure, nex);
          else if (RAY) xpar ^= cross_test<EXACT>(s, tris[~ch], unsure, nex);
          else hit |= contact_test<EXACT>(s, tris[~ch], xpar, ninc, local_ok, unsure, nex) == 1;
        }
        w0 = h0 && c0 >= 0;
        w1 = h1 && c1 >= 0;
      }
      const uint64_t m0 = __ballot(w0) & gm, m1 = __ballot(w1) & gm;
      if (!RAY && (__ballot(hit) & gm)) return 1;
      const int cnt0 = __popcll(m0), tot = cnt0 + __popcll(m1);
      if (nn + tot > KF) return 5;
      if (w0) nxt[nn + __popcll(m0 & below)] = c0;
      if (w1) nxt[nn + cnt0 + __popcll(m1 & below)] = c1;
      nn += tot;
    }
    wave_lds_sync();
    int *t = cur;
    cur = nxt;
    nxt = t;
    n = nn;
  }
  return 0;
}

// EXACT false: the walk queue (only when it holds at most a.front_max
// pairs).  EXACT true: the first kExactSplit entries of the deferred list,
// every test decided; an entry this kernel settles gets the kSettled bit
// (k_walk_split skips it), one whose frontier overflows or that has a cage
// is left to k_walk_split.
template <bool COUNT, bool BATCH, int G, bool EXACT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(EXACT ? 1 : kWalkWaves))) void k_walk_front(VisArgs a) {
  constexpr int kGroups = kBlock / G;
  constexpr int kFront = G >= 8 ? 16 * G : 128;  // frontier entries per group and buffer
  __shared__ int front_all[kGroups][2][kFront];
  const int lane = threadIdx.x & 63;
  const int slot = threadIdx.x / G;
  const int g = lane & (G - 1);
  const uint64_t gm = group_mask(lane, G);
  const uint64_t below = gm & ((1ull << lane) - 1ull);  // lanes of the group before this one
  unsigned n_blocked = 0, n_bad = 0, nv = 0, nt = 0, n_exact = 0, n_clip = 0, n_turned = 0;
  int64_t *list = EXACT ? a.dlist : a.queue;
  int64_t len = (int64_t)(EXACT ? *a.dlist_len : *a.queue_len);
  if (EXACT && len > kExactSplit) len = kExactSplit;
  if (!EXACT && len > a.front_max) return;
  int *fa = front_all[slot][0], *fb = front_all[slot][1];
  for (int64_t e = (int64_t)blockIdx.x * kGroups + slot; e < len;
       e += (int64_t)gridDim.x * kGroups) {
    const int64_t raw = list[e];
    const bool clip = !EXACT && (raw & kClip);
    const int64_t kg = EXACT ? raw : (raw & ~kClip);
    if (EXACT && (kg & kSettled)) continue;  // settled by the exact cap rule
    const int64_t w = kg >> 6;
    const int b = (int)(kg & 63);
    const Part P = part_of_word<BATCH>(a, w);
    if (EXACT && P.ctris) continue;  // cage pairs: k_walk_split
    int64_t i, j;
    pair_of((w - P.word_begin) * 64 + b, P.n, i, j);
    Seg s;
#pragma unroll
    for (int c = 0; c < 3; ++c) s.p[c] = P.pts[3 * i + c], s.q[c] = P.pts[3 * j + c];
    make_seg(s);
    // A pair whose fans showed both ends leaving outwards and neither fan
    // touching the segment is an edge iff no other triangle touches it; such
    // a triangle keeps clear[i] from p and clear[j] from q, so only the
    // parameters [clear_i / len, 1 - clear_j / len] need box tests (0.1 %
    // and 1e-6 of slack cover the float32 rounding of these bounds).  When
    // the two balls cover the segment, no triangle can touch it.
    float u0 = 0.0f, u1 = 1.0f;
    if (clip) {
      const float dx = s.q[0] - s.p[0], dy = s.q[1] - s.p[1], dz = s.q[2] - s.p[2];
      const float len = sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
      u0 = fmaf(0.999f, P.clear[i] / len, -1e-6f);
      u1 = 1.0f - fmaf(0.999f, P.clear[j] / len, -1e-6f);
      if (g == 0) ++n_clip;
    }
    int xpar = 0, ninc = 0;
    bool local_ok = true, unsure = false;
    const int hs = (clip && !(u0 < u1))
                       ? 0
                       : front_walk<G, kFront, false, COUNT, EXACT>(
                             P.rnodes, P.rtris, P.rgroups, P.rcap, s, fa, fb, g, gm, below, xpar,
                             ninc, local_ok, unsure, nv, nt, n_exact, u0, u1);
    uint64_t mu = __ballot(unsure) & gm;
    const uint64_t ml = __ballot(!local_ok) & gm, mx = __ballot(xpar & 1) & gm,
                   mn = __ballot(ninc > 0) & gm;
    bool edge = false, defer = false, left = false;
    if (hs == 3) {
      if (g == 0) ++n_bad;
    } else if (hs == 1) {
      if (g == 0) ++n_blocked;
    } else if (hs == 5 || mu) {
      // undecided tests or a frontier overflow: the exact redo (EXACT: the
      // depth-first one)
      if (EXACT) left = true;
      else defer = true;
    } else if (clip) {
      edge = true;  // outside at both ends, untouched
    } else if (!ml && P.far && mn && P.far[i] <= 1) {
      // closed mesh, p on it: inside parity = fan crossings ^ far crossings of p
      edge = (((__popcll(mx) & 1) ^ P.far[i]) & 1) == 0;
    } else {  // no local parity for this pair: the midpoint ray, breadth first too
      wave_lds_sync();
      int cross = 0, unused = 0;
      bool ok = true;
      unsure = false;
      const int hr = front_walk<G, kFront, true, COUNT, EXACT>(
          P.snodes, P.stris, nullptr, P.scap, s, fa, fb, g, gm, below, cross, unused, ok, unsure,
          nv, nt, n_exact);
      mu = __ballot(unsure) & gm;
      uint64_t mc = __ballot(cross & 1) & gm;
      int hr2 = hr;
      if (!EXACT && a.rotated_rays && P.far && hr != 3 && (hr == 5 || mu)) {
        // The +x ray met a case its filters cannot decide.  The segment
        // touches no triangle (decided above), so its midpoint is off the
        // surface, and on a closed mesh its crossing parity is the same
        // along every ray: try the +y ray (the rotated problem's +x ray)
        // and then the +z ray before handing the pair to the exact redo.
        wave_lds_sync();
        const Seg sr = rotate_seg(s);
        cross = 0;
        unsure = false;
        hr2 = front_walk<G, kFront, true, COUNT, false, 1>(
            P.snodes, P.stris, nullptr, P.scap, sr, fa, fb, g, gm, below, cross, unused, ok,
            unsure, nv, nt, n_exact);
        mu = __ballot(unsure) & gm;
        mc = __ballot(cross & 1) & gm;
        if (hr2 != 3 && (hr2 == 5 || mu)) {  // and the +z ray
          wave_lds_sync();
          const Seg sz = rotate_seg(sr);
          cross = 0;
          unsure = false;
          hr2 = front_walk<G, kFront, true, COUNT, false, 2>(
              P.snodes, P.stris, nullptr, P.scap, sz, fa, fb, g, gm, below, cross, unused, ok,
              unsure, nv, nt, n_exact);
          mu = __ballot(unsure) & gm;
          mc = __ballot(cross & 1) & gm;
        }
        if (g == 0 && hr2 != 5 && !mu) ++n_turned;
      }
      if (hr2 == 3) {
        if (g == 0) ++n_bad;
      } else if (hr2 == 5 || mu) {
        if (EXACT) left = true;
        else defer = true;
      } else {
        edge = (__popcll(mc) & 1) == 0;
      }
    }
    if (g == 0) {
      if (edge) atomicOr(reinterpret_cast<unsigned long long *>(&a.bits[w]), 1ull << b);
      if (defer) a.dlist[atomicAdd(a.dlist_len, 1ull)] = kg;
      if (EXACT && !left) list[e] = kg | kSettled;  // settled here
    }
    // the next pair of this group reuses the frontier buffers: keep its
    // writes after every lane's last read of this pair's frontier
    wave_lds_sync();
  }
  for (int off = 32; off > 0; off >>= 1) {
    n_blocked += __shfl_down(n_blocked, off);
    n_bad += __shfl_down(n_bad, off);
    n_exact += __shfl_down(n_exact, off);
    n_clip += __shfl_down(n_clip, off);
    n_turned += __shfl_down(n_turned, off);
    nv += __shfl_down(nv, off);
    nt += __shfl_down(nt, off);
  }
  if (lane == 0) {
    if (n_blocked) atomicAdd(&a.stats[1], (unsigned long long)n_blocked);
    if (n_bad) atomicAdd(&a.stats[4], (unsigned long long)n_bad);
    if (EXACT && n_exact) atomicAdd(&a.stats[9], (unsigned long long)n_exact);
    if (!EXACT && n_clip) atomicAdd(&a.stats[15], (unsigned long long)n_clip);
    if (!EXACT && n_turned) atomicAdd(&a.stats[2], (unsigned long long)n_turned);
    if (COUNT) {
      atomicAdd(&a.stats[5], (unsigned long long)nv);
      atomicAdd(&a.stats[6], (unsigned long long)nt);
    }
  }
}

// ------------------------------------------------------------ exact redo
// The cap rule decided exactly for the deferred pairs whose points share a
// planar group (the fan pass defers those whose filtered 2-D signs stay
// open), a lane per entry, before the exact walks: a pair that enters a
// cap at either end is blocked and gets the kSettled bit (the exact walks
// skip it); for every other pair the rule cannot hold (the filter decided
// it or the exact test says no) and the walks decide it.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_cap_redo(VisArgs a) {
  const int64_t len = (int64_t)*a.dlist_len;
  unsigned n_blocked = 0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < len;
       e += (int64_t)gridDim.x * 256) {
    const int64_t kg = a.dlist[e];
    const int64_t w = kg >> 6;
    const Part P = part_of_word<BATCH>(a, w);
    if (!P.mask || !P.fan_n || P.ctris) continue;
    int64_t i, j;
    pair_of((w - P.word_begin) * 64 + (kg & 63), P.n, i, j);
    const unsigned shared = (unsigned)(P.mask[i] & P.mask[j]) >> 1;
    if (!shared) continue;
    float p[3], q[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = P.pts[3 * i + c], q[c] = P.pts[3 * j + c];
    if (enters_cap(P, i, p, q, shared) || enters_cap(P, j, q, p, shared)) {
      a.dlist[e] = kg | kSettled;
      ++n_blocked;
    }
  }
  for (int off = 32; off > 0; off >>= 1) n_blocked += __shfl_down(n_blocked, off);
  if ((threadIdx.x & 63) == 0 && n_blocked) atomicAdd(&a.stats[1], (unsigned long long)n_blocked);
}

// ------------------------------------------------------------ far parity
// Crossings (mod 2) of the +x ray from the segment's midpoint with the
// triangles of this lane's share.  Midpoint mode (skip NULL): the ray is
// perturbed by (0, e, e^2), the definition of midpoint parity.  Point mode
// (skip = the point, s.p == s.q): triangles having the point as a vertex are
// left out and the ray is not perturbed; a test whose answer would depend
// on a perturbation sets degenerate instead (for pairs starting at the
// point, the pair's own direction would have to break that tie).
__device__ int exact_ray_parity(const Node *__restrict__ nodes, const Tri *__restrict__ tris,
                                const Seg &s, int root, int *stk, int cap, const float *skip,
                                bool &bad, bool &degenerate) {
  auto leaf = [&](int id) {
    const Tri tr = tris[id];
    const float a3[3] = {tr.a.x, tr.a.y, tr.a.z}, b3[3] = {tr.b.x, tr.b.y, tr.b.z},
                c3[3] = {tr.c.x, tr.c.y, tr.c.z};
    if (skip && (same3(skip, a3) || same3(skip, b3) || same3(skip, c3))) return 0;
    int r = cross_fast(s, tr);
    if (r == 2)
      r = skip ? ray_cross_point_exact(s.p, a3, b3, c3, degenerate)
               : (ray_cross_exact(s.p, s.q, a3, b3, c3) ? 1 : 0);
    return r;
  };
  if (root == kNoRoot) return 0;
  if (root < 0) return leaf(~root);
  int node = root, sp = 0, cnt = 0;
  for (int it = 0;; ++it) {
    if (it > cap) {
      bad = true;
      return 0;
    }
    float bx[12];
    int c0, c1;
    load_node(nodes, node, bx, c0, c1);
    int ch[2] = {c0, c1};
    bool h[2] = {ray_box(s, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5]),
                 ray_box(s, bx[6], bx[7], bx[8], bx[9], bx[10], bx[11])};
    for (int t = 0; t < 2; ++t) {
      if (h[t] && ch[t] < 0) {
        cnt += leaf(~ch[t]);
        h[t] = false;
      }
    }
    if (h[0] && h[1]) {
      stk[(sp++) * 64] = ch[1];
      node = ch[0];
    } else if (h[0]) {
      node = ch[0];
    } else if (h[1]) {
      node = ch[1];
    } else {
      if (sp == 0) return cnt & 1;
      node = stk[(--sp) * 64];
    }
  }
}

// Squared distance from p to the closed segment ab (float64).
__device__ __forceinline__ double seg_dist2(const double p[3], const double a[3],
                                            const double b[3]) {
  const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
  const double l2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
  double t = l2 > 0.0 ? (ap[0] * ab[0] + ap[1] * ab[1] + ap[2] * ab[2]) / l2 : 0.0;
  t = fmin(fmax(t, 0.0), 1.0);
  const double d[3] = {ap[0] - t * ab[0], ap[1] - t * ab[1], ap[2] - t * ab[2]};
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// Squared distance from p to the closed triangle abc (float64): the nearest
// of its three edges, or its plane when p projects inside it.  Degenerate
// triangles are their edges.
__device__ __forceinline__ double tri_dist2(const double p[3], const double a[3],
                                            const double b[3], const double c[3]) {
  double d2 = fmin(seg_dist2(p, a, b), fmin(seg_dist2(p, b, c), seg_dist2(p, c, a)));
  const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                       e1[0] * e2[1] - e1[1] * e2[0]};
  const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (nn > 0.0) {
    const double *v[3] = {a, b, c};
    bool inside = true;
    for (int k = 0; k < 3 && inside; ++k) {
      const double *u = v[k], *w = v[(k + 1) % 3];
      const double e[3] = {w[0] - u[0], w[1] - u[1], w[2] - u[2]};
      const double r[3] = {p[0] - u[0], p[1] - u[1], p[2] - u[2]};
      const double x[3] = {e[1] * r[2] - e[2] * r[1], e[2] * r[0] - e[0] * r[2],
                           e[0] * r[1] - e[1] * r[0]};
      inside = x[0] * n[0] + x[1] * n[1] + x[2] * n[2] >= 0.0;
    }
    if (inside) {
      const double h = (p[0] - a[0]) * n[0] + (p[1] - a[1]) * n[1] + (p[2] - a[2]) * n[2];
      d2 = fmin(d2, h * h / nn);
    }
  }
  return d2;
}

// Minimum of a double over the lanes of a group of G consecutive lanes.
template <int G>
__device__ __forceinline__ double group_min(double v) {
#pragma unroll
  for (int off = 1; off < G; off <<= 1) v = fmin(v, __shfl_xor(v, off));
  return v;
}

// The records and cones of point g's fan (cnt triangle ids at out; p the
// point).
__device__ __forceinline__ void fan_finish(const Part &P, int64_t g, const float p[3],
                                           const int *out, int cnt,
                                           FanRec *__restrict__ fan_rec,
                                           FanCone *__restrict__ fan_cone) {
  // the fan's records (see FanRec)
  for (int f = 0; f < cnt; ++f) {
    const Tri t = P.stris[out[f]];
    const float a3[3] = {t.a.x, t.a.y, t.a.z}, b3[3] = {t.b.x, t.b.y, t.b.z},
                c3[3] = {t.c.x, t.c.y, t.c.z};
    const bool pa = same3(p, a3), pb = same3(p, b3);
    const float *v1 = pa ? b3 : (pb ? c3 : a3);
    const float *v2 = pa ? c3 : (pb ? a3 : b3);
    FanRec r;
    const double e1x = (double)v1[0] - p[0], e1y = (double)v1[1] - p[1], e1z = (double)v1[2] - p[2];
    const double e2x = (double)v2[0] - p[0], e2y = (double)v2[1] - p[1], e2z = (double)v2[2] - p[2];
    r.nx = e1y * e2z - e1z * e2y;
    r.ny = e1z * e2x - e1x * e2z;
    r.nz = e1x * e2y - e1y * e2x;
    r.ax = fabs(e1y * e2z) + fabs(e1z * e2y);
    r.ay = fabs(e1z * e2x) + fabs(e1x * e2z);
    r.az = fabs(e1x * e2y) + fabs(e1y * e2x);
    r.e1y = e1y, r.e1z = e1z, r.e2y = e2y, r.e2z = e2z;
    for (int k = 0; k < 3; ++k) r.v1[k] = v1[k], r.v2[k] = v2[k];
    double b0;
    const double o0 = o2yz(p, v1, v2, b0);
    r.o0 = o0 > b0 ? 1 : (-o0 > b0 ? -1 : 0);
    r.grp = (int)__float_as_uint(t.c.w);
    fan_rec[(size_t)g * kFanMax + f] = r;
  }
  // the cones: axis = sum of the unit normals, each turned to the first's side
  double ax = 0.0, ay = 0.0, az = 0.0, n0[3] = {0.0, 0.0, 0.0};
  for (int f = 0; f < cnt; ++f) {
    const FanRec &r = fan_rec[(size_t)g * kFanMax + f];
    const double ln = sqrt(r.nx * r.nx + r.ny * r.ny + r.nz * r.nz);
    if (!(ln > 0.0)) continue;
    double u[3] = {r.nx / ln, r.ny / ln, r.nz / ln};
    if (n0[0] == 0.0 && n0[1] == 0.0 && n0[2] == 0.0) n0[0] = u[0], n0[1] = u[1], n0[2] = u[2];
    if (u[0] * n0[0] + u[1] * n0[1] + u[2] * n0[2] < 0.0) u[0] = -u[0], u[1] = -u[1], u[2] = -u[2];
    ax += u[0], ay += u[1], az += u[2];
  }
  FanCone cone;
  cone.tp = cone.tm = 2.0f;
  cone.xp = cone.xm = 0;
  cone.pad = 0;
  const double la = sqrt(ax * ax + ay * ay + az * az);
  const float fa[3] = {(float)(ax / la), (float)(ay / la), (float)(az / la)};
  cone.ax = fa[0], cone.ay = fa[1], cone.az = fa[2];
  if (la > 0.0 && cnt > 0) {
    // the float axis as used by the pairs, unit in double
    const double lf = sqrt((double)fa[0] * fa[0] + (double)fa[1] * fa[1] + (double)fa[2] * fa[2]);
    const double up[3] = {fa[0] / lf, fa[1] / lf, fa[2] / lf};
    const double um[3] = {-up[0], -up[1], -up[2]};
    double cp = -1.0, cm = -1.0;  // cosines of the cones' half-angles
    int xpp = 0, xpm = 0, np = 0, nm = 0, rp = 0, rm = 0;
    for (int f = 0; f < cnt; ++f) {
      const FanRec &r = fan_rec[(size_t)g * kFanMax + f];
      const double e1[3] = {(double)r.v1[0] - p[0], (double)r.v1[1] - p[1], (double)r.v1[2] - p[2]};
      const double e2[3] = {(double)r.v2[0] - p[0], (double)r.v2[1] - p[1], (double)r.v2[2] - p[2]};
      cp = fmax(cp, sector_cos(up, e1, e2));
      cm = fmax(cm, sector_cos(um, e1, e2));
      // the axes' own crossing parities (p itself stands for the other
      // point: it is no vertex of the fan but its apex)
      if (rp == 0) rp = fan_test(r, p, up, xpp, np);
      if (rm == 0) rm = fan_test(r, p, um, xpm, nm);
    }
    if (rp == 0) {
      cone.tp = cone_threshold(cp);
      cone.xp = xpp & 1;
    }
    if (rm == 0) {
      cone.tm = cone_threshold(cm);
      cone.xm = xpm & 1;
    }
  }
  fan_cone[g] = cone;
}

// Fan of each point: the leaf-ordered ids of the triangles that have the
// point as a vertex (bitwise-equal coordinates), found by a point query on
// the 4-wide BVH; -1 when there are more than kFanMax.  One point g per
// lane; stk: the lane's walk stack in LDS, 64 entries at a stride of 64 ints
// (lane-interleaved: a private array would live in scratch memory).
template <bool BATCH>
__device__ __forceinline__ void fan_point(const VisArgs &a, int64_t g, int *stk,
                                          int *__restrict__ fan_n, int *__restrict__ fan_idx,
                                          FanRec *__restrict__ fan_rec,
                                          FanCone *__restrict__ fan_cone) {
  Part P = a.one;
  if (BATCH) {
    int lo = 0, hi = a.nparts - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.parts[mid].point_begin <= g) lo = mid;
      else hi = mid - 1;
    }
    P = a.parts[lo];
  }
  if (!P.fan_n) return;
  const int64_t i = g - P.point_begin;
  const float p[3] = {P.pts[3 * i], P.pts[3 * i + 1], P.pts[3 * i + 2]};
  int node = 0, sp = 0, cnt = 0;
  bool done = false;
  int *out = fan_idx + (size_t)g * kFanMax;
  for (int it = 0; it <= P.scap && !done; ++it) {
    const Node4 nd = P.snodes4[node];
    int next = -1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const bool in = p[0] >= lane_of(nd.f[0], c) && p[0] <= lane_of(nd.f[3], c) &&
                      p[1] >= lane_of(nd.f[1], c) && p[1] <= lane_of(nd.f[4], c) &&
                      p[2] >= lane_of(nd.f[2], c) && p[2] <= lane_of(nd.f[5], c);
      if (!in) continue;
      const int id = __float_as_int(lane_of(nd.f[6], c));
      if (id < 0) {
        const Tri t = P.stris[~id];
        const float a3[3] = {t.a.x, t.a.y, t.a.z}, b3[3] = {t.b.x, t.b.y, t.b.z},
                    c3[3] = {t.c.x, t.c.y, t.c.z};
        if (same3(p, a3) || same3(p, b3) || same3(p, c3)) {
          if (cnt < kFanMax) out[cnt] = ~id;
          ++cnt;
        }
      } else if (next < 0) {
        next = id;
      } else if (sp < 64) {
        stk[(sp++) * 64] = id;
      } else {
        cnt = kFanMax + 1;  // stack exhausted: give up on this point's fan
      }
    }
    if (next >= 0) {
      node = next;
    } else if (sp == 0) {
      done = true;
    } else {
      node = stk[(--sp) * 64];
    }
  }
  // an unfinished walk or an overfull fan leaves the point to the full walk
  fan_n[g] = (!done || cnt > kFanMax) ? -1 : cnt;
  if (!done || cnt > kFanMax) return;
  fan_finish(P, g, p, out, cnt, fan_rec, fan_cone);
}

// The fan query breadth first: G lanes per point (the group), the point's
// frontier of tree nodes (binary tree: boxes containing the point) in LDS as
// in the far walks, fan triangles appended to the point's list in ballot
// order (in LDS) and then sorted; a group per point of block `block` (NT
// threads, lds: (NT / G) * (2 * 16 * G + kFanMax) ints).  About as many
// dependent steps as the tree is deep instead of one per visited node.
template <bool BATCH, int G, int NT>
__device__ __forceinline__ void fan_front_block(const VisArgs &a, int64_t block, int *lds,
                                                int *__restrict__ fan_n,
                                                int *__restrict__ fan_idx,
                                                FanRec *__restrict__ fan_rec,
                                                FanCone *__restrict__ fan_cone,
                                                float *__restrict__ clear) {
  constexpr int kGroups = NT / G;
  constexpr int kF = 16 * G;
  int(*front_all)[2][kF] = reinterpret_cast<int(*)[2][kF]>(lds);
  const int lane = threadIdx.x & 63;
  const int slot = threadIdx.x / G;
  const int g = lane & (G - 1);
  const uint64_t gm = group_mask(lane, G);
  const uint64_t below = gm & ((1ull << lane) - 1ull);
  const int64_t pt = block * kGroups + slot;
  if (pt >= a.npoints) return;  // group-uniform; no block-wide barrier below
  Part P = a.one;
  if (BATCH) {
    int lo = 0, hi = a.nparts - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.parts[mid].point_begin <= pt) lo = mid;
      else hi = mid - 1;
    }
    P = a.parts[lo];
  }
  if (!P.fan_n) return;
  const int64_t i = pt - P.point_begin;
  const float p[3] = {P.pts[3 * i], P.pts[3 * i + 1], P.pts[3 * i + 2]};
  // the clearance (P.clear): the same walk also takes boxes within the best
  // distance so far (at first clear_r) of a triangle not at p, and measures
  // those triangles (bound as in clear_point)
  const bool want_clear = P.clear != nullptr;
  const double pd[3] = {p[0], p[1], p[2]};
  const double R = (double)P.clear_r;
  double best = R * R;  // squared, this lane's; the group's minimum prunes
  int *out = fan_idx + (size_t)pt * kFanMax;
  int *fl = lds + kGroups * 2 * kF + slot * kFanMax;  // the group's fan list
  int *cur = front_all[slot][0], *nxt = front_all[slot][1];
  if (g == 0) cur[0] = 0;
  wave_lds_sync();
  int n = 1, cnt = 0;
  int64_t visits = 0;
  bool fail = false;
  while (n > 0 && !fail) {
    visits += n;
    if (visits > P.scap) {
      fail = true;
      break;
    }
    const double bound = want_clear ? group_min<G>(best) : 0.0;
    int nn = 0;
    for (int base = 0; base < n; base += G) {
      const int idx = base + g;
      bool w0 = false, w1 = false, f0 = false, f1 = false;
      int c0 = 0, c1 = 0;
      if (idx < n) {
        float bx[12];
        load_node(P.snodes, cur[idx], bx, c0, c1);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const float *b = bx + 6 * c;
          const bool in = p[0] >= b[0] && p[0] <= b[3] && p[1] >= b[1] && p[1] <= b[4] &&
                          p[2] >= b[2] && p[2] <= b[5];
          bool near = false;  // within the clearance bound (boxes holding p aside)
          if (want_clear && !in) {
            double d2 = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const double o = fmax(fmax((double)b[k] - pd[k], pd[k] - (double)b[3 + k]), 0.0);
              d2 += o * o;
            }
            near = d2 < bound;
          }
          if (!in && !near) continue;
          const int ch = c ? c1 : c0;
          bool in_fan = false;
          if (ch < 0) {
            const Tri t = P.stris[~ch];
            const float a3[3] = {t.a.x, t.a.y, t.a.z}, b3[3] = {t.b.x, t.b.y, t.b.z},
                        c3[3] = {t.c.x, t.c.y, t.c.z};
            in_fan = in && (same3(p, a3) || same3(p, b3) || same3(p, c3));
            if (want_clear && !in_fan) {
              const double A[3] = {a3[0], a3[1], a3[2]}, B[3] = {b3[0], b3[1], b3[2]},
                           C[3] = {c3[0], c3[1], c3[2]};
              best = fmin(best, tri_dist2(pd, A, B, C));
            }
          }
          if (c) f1 = in_fan, w1 = ch >= 0;
          else f0 = in_fan, w0 = ch >= 0;
        }
      }
      const uint64_t mf0 = __ballot(f0) & gm, mf1 = __ballot(f1) & gm;
      const int nf0 = __popcll(mf0);
      if (f0 && cnt + __popcll(mf0 & below) < kFanMax) fl[cnt + __popcll(mf0 & below)] = ~c0;
      if (f1 && cnt + nf0 + __popcll(mf1 & below) < kFanMax)
        fl[cnt + nf0 + __popcll(mf1 & below)] = ~c1;
      cnt += nf0 + __popcll(mf1);
      const uint64_t m0 = __ballot(w0) & gm, m1 = __ballot(w1) & gm;
      const int cnt0 = __popcll(m0), tot = cnt0 + __popcll(m1);
      if (nn + tot > kF) {
        fail = true;  // frontier overflow: k_fan_redo queries the point depth first
        break;
      }
      if (w0) nxt[nn + __popcll(m0 & below)] = c0;
      if (w1) nxt[nn + cnt0 + __popcll(m1 & below)] = c1;
      nn += tot;
    }
    wave_lds_sync();
    int *t = cur;
    cur = nxt;
    nxt = t;
    n = nn;
  }
  wave_lds_sync();  // the group's frontier buffers are free after this
  if (want_clear) {
    best = group_min<G>(best);
    // a failed walk leaves no bound (k_fan_redo writes none either)
    if (g == 0) clear[pt] = fail ? 0.0f : __double2float_rz(sqrt(best) * (1.0 - 1e-6));
  }
  if (g != 0) return;
  if (fail || cnt > kFanMax) {
    fan_n[pt] = fail ? -2 : -1;  // -2: k_fan_redo
    return;
  }
  // canonical order: ascending triangle id
  for (int u = 1; u < cnt; ++u) {
    const int v = fl[u];
    int w = u - 1;
    while (w >= 0 && fl[w] > v) {
      fl[w + 1] = fl[w];
      --w;
    }
    fl[w + 1] = v;
  }
  for (int u = 0; u < cnt; ++u) out[u] = fl[u];
  fan_n[pt] = cnt;
  fan_finish(P, pt, p, fl, cnt, fan_rec, fan_cone);
}


// Points whose breadth-first fan query overflowed its frontier (fan_n -2),
// queried again depth first, a lane per point.
// Clearance of point g (one lane): a lower bound of its distance to every
// triangle that does not have it as a vertex (bitwise-equal coordinates),
// capped at the part's search radius clear_r: a nearest-triangle walk of the
// 4-wide tree, nearest child first, that skips boxes no nearer than the best
// distance so far (at first clear_r), so it visits only the boxes within
// clear_r of the point.  Distances are float64 from the float32 coordinates;
// the result is lowered by a relative 1e-6 (far above their rounding) and
// rounded down to float32.  0 when the walk's stack (KS entries: ids in stk,
// their box distances in dst, both at a stride of 64) or its visit cap runs
// out.  clear: the whole batch's buffer, indexed by global point.
template <bool BATCH, int KS>
__device__ __forceinline__ void clear_point(const VisArgs &a, int64_t g, int *stk, float *dst,
                                            float *__restrict__ clear) {
  Part P = a.one;
  if (BATCH) {
    int lo = 0, hi = a.nparts - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.parts[mid].point_begin <= g) lo = mid;
      else hi = mid - 1;
    }
    P = a.parts[lo];
  }
  if (!P.clear) return;
  const int64_t i = g - P.point_begin;
  const float pf[3] = {P.pts[3 * i], P.pts[3 * i + 1], P.pts[3 * i + 2]};
  const double p[3] = {pf[0], pf[1], pf[2]};
  const double R = (double)P.clear_r;
  double best = R * R;  // squared
  int node = 0, sp = 0;
  bool ok = false;
  for (int it = 0; it <= P.scap; ++it) {
    const Node4 nd = P.snodes4[node];
    int next = -1;
    double nd2 = 0.0;
    bool full = false;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (lane_of(nd.f[0], c) == kEmpty) continue;
      double d2 = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double lo = lane_of(nd.f[k], c), hi = lane_of(nd.f[3 + k], c);
        const double o = fmax(fmax(lo - p[k], p[k] - hi), 0.0);
        d2 += o * o;
      }
      if (!(d2 < best)) continue;
      const int id = __float_as_int(lane_of(nd.f[6], c));
      if (id < 0) {
        const Tri t = P.stris[~id];
        const float a3[3] = {t.a.x, t.a.y, t.a.z}, b3[3] = {t.b.x, t.b.y, t.b.z},
                    c3[3] = {t.c.x, t.c.y, t.c.z};
        if (same3(pf, a3) || same3(pf, b3) || same3(pf, c3)) continue;
        const double A[3] = {a3[0], a3[1], a3[2]}, B[3] = {b3[0], b3[1], b3[2]},
                     C[3] = {c3[0], c3[1], c3[2]};
        best = fmin(best, tri_dist2(p, A, B, C));
      } else if (next < 0) {
        next = id, nd2 = d2;
      } else {
        // descend into the nearer one, push the other
        int push = id;
        double pd = d2;
        if (d2 < nd2) push = next, pd = nd2, next = id, nd2 = d2;
        if (sp >= KS) {
          full = true;
          break;
        }
        stk[sp * 64] = push;
        dst[sp * 64] = (float)pd * 0.999f;  // lowered: skipping it on a pop stays safe
        ++sp;
      }
    }
    if (full) break;  // no bound for this point
    if (next >= 0 && nd2 < best) {
      node = next;
      continue;
    }
    node = -1;
    while (sp > 0) {
      --sp;
      if ((double)dst[sp * 64] < best) {
        node = stk[sp * 64];
        break;
      }
    }
    if (node < 0) {
      ok = true;
      break;
    }
  }
  clear[g] = ok ? __double2float_rz(sqrt(best) * (1.0 - 1e-6)) : 0.0f;
}


// The clearances of all points on their own (a lane per point), when the far
// walks run without k_far_fan.
template <bool BATCH>
__global__ __launch_bounds__(64) void k_clear(VisArgs a, float *__restrict__ clear) {
  __shared__ int stk_all[64][64];
  __shared__ float dst_all[64][64];
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g < a.npoints) clear_point<BATCH, 64>(a, g, &stk_all[0][threadIdx.x], &dst_all[0][threadIdx.x], clear);
}

template <bool BATCH>
__global__ __launch_bounds__(64) void k_fan_redo(VisArgs a, int *__restrict__ fan_n,
                                                 int *__restrict__ fan_idx,
                                                 FanRec *__restrict__ fan_rec,
                                                 FanCone *__restrict__ fan_cone) {
  __shared__ int stk_all[64][64];
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g < a.npoints && fan_n[g] == -2)
    fan_point<BATCH>(a, g, &stk_all[0][threadIdx.x], fan_n, fan_idx, fan_rec, fan_cone);
}

template <bool BATCH>
__global__ __launch_bounds__(64) void k_fan(VisArgs a, int *__restrict__ fan_n,
                                            int *__restrict__ fan_idx,
                                            FanRec *__restrict__ fan_rec,
                                            FanCone *__restrict__ fan_cone) {
  __shared__ int stk_all[64][64];
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g < a.npoints) fan_point<BATCH>(a, g, &stk_all[0][threadIdx.x], fan_n, fan_idx, fan_rec, fan_cone);
}

// Crossing parity (0/1) of the +x ray from each point with the triangles
// that do not have the point as a vertex; tests the filter leaves open are
// decided exactly.  2 when a test is degenerate (the ray meets a triangle's
// boundary exactly: then the parity seen from p + e(q-p) depends on q, and
// the point's walked pairs walk their midpoint rays instead) or a walk
// exceeds its node budget.  One wave per point, the 64 lanes splitting the walk
// (split_root at level 6).  In a batch, far is indexed by the global point.
template <bool BATCH, bool REDO = false>  // REDO: only the points marked 3
__global__ __launch_bounds__(64) void k_far_parity(VisArgs a, uint8_t *__restrict__ far) {
  // a walk holds at most 32 entries: depth - 1 <= 32, or small-first order
  // (binary_stack)
  __shared__ int stk_all[32][64];
  const int lane = threadIdx.x;
  int *stk = &stk_all[0][lane];
  for (int64_t g = blockIdx.x; g < a.npoints; g += gridDim.x) {
    if (REDO && far[g] != 3) continue;  // (the part is checked below)
    Part P = a.one;
    if (BATCH) {
      int lo = 0, hi = a.nparts - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.parts[mid].point_begin <= g) lo = mid;
        else hi = mid - 1;
      }
      P = a.parts[lo];
    }
    if (!P.far) continue;  // open mesh or cage mode: not needed
    const int64_t i = g - P.point_begin;
    Seg s;
    for (int c = 0; c < 3; ++c) s.p[c] = s.q[c] = P.pts[3 * i + c];
    make_seg(s);
    bool bad = false, degenerate = false;
    const int root = split_root<true>(P.snodes, lane, 6, s);
    const int par =
        exact_ray_parity(P.snodes, P.stris, s, root, stk, P.scap, s.p, bad, degenerate);
    const uint64_t odd = __ballot(par & 1), mb = __ballot(bad || degenerate);
    if (lane == 0) far[g] = mb ? 2 : (uint8_t)(__popcll(odd) & 1);
  }
}

// The same, breadth first: G lanes per point (the group), the point's
// frontier of tree nodes in LDS (as in k_walk_front), so that a wave walks
// 64 / G points at once; a point whose frontier overflows is marked 3 and
// redone depth first by k_far_parity<.., true>.
// The walks of block `block` of `nblocks` blocks of NT threads; lds: the
// block's frontiers, (NT / G) * 2 * 16 * G ints.
template <bool BATCH, int G, bool AGAIN, int NT>
__device__ __forceinline__ void far_front_block(const VisArgs &a, uint8_t *__restrict__ far,
                                                int64_t block, int64_t nblocks, int *lds) {
  constexpr int kGroups = NT / G;
  constexpr int kF = 16 * G;
  int(*front_all)[2][kF] = reinterpret_cast<int(*)[2][kF]>(lds);
  const int lane = threadIdx.x & 63;
  const int slot = threadIdx.x / G;
  const int g = lane & (G - 1);
  const uint64_t gm = group_mask(lane, G);
  const uint64_t below = gm & ((1ull << lane) - 1ull);
  for (int64_t pt = block * kGroups + slot; pt < a.npoints; pt += nblocks * kGroups) {
    if (AGAIN && far[pt] != 3) continue;  // (the part is checked below)
    Part P = a.one;
    if (BATCH) {
      int lo = 0, hi = a.nparts - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.parts[mid].point_begin <= pt) lo = mid;
        else hi = mid - 1;
      }
      P = a.parts[lo];
    }
    if (!P.far) continue;  // open mesh or cage mode: not needed (group-uniform)
    const int64_t i = pt - P.point_begin;
    Seg s;
    for (int c = 0; c < 3; ++c) s.p[c] = s.q[c] = P.pts[3 * i + c];
    make_seg(s);
    int *cur = front_all[slot][0], *nxt = front_all[slot][1];
    if (g == 0) cur[0] = 0;
    wave_lds_sync();
    int n = 1, cnt = 0;
    int64_t visits = 0;
    bool bad = false, degenerate = false, over = false;
    while (n > 0) {
      visits += n;
      if (visits > P.scap) {
        bad = true;
        break;
      }
      int nn = 0;
      for (int base = 0; base < n; base += G) {
        const int idx = base + g;
        bool w0 = false, w1 = false;
        int c0 = 0, c1 = 0;
        if (idx < n) {
          float bx[12];
          load_node(P.snodes, cur[idx], bx, c0, c1);
          const bool h0 = ray_box(s, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5]);
          const bool h1 = ray_box(s, bx[6], bx[7], bx[8], bx[9], bx[10], bx[11]);
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            const int ch = c ? c1 : c0;
            if (!(c ? h1 : h0) || ch >= 0) continue;
            const Tri tr = P.stris[~ch];
            const float a3[3] = {tr.a.x, tr.a.y, tr.a.z}, b3[3] = {tr.b.x, tr.b.y, tr.b.z},
                        c3[3] = {tr.c.x, tr.c.y, tr.c.z};
            if (same3(s.p, a3) || same3(s.p, b3) || same3(s.p, c3)) continue;
            int r = cross_fast(s, tr);
            if (r == 2) r = ray_cross_point_exact(s.p, a3, b3, c3, degenerate);
            cnt += r;
          }
          w0 = h0 && c0 >= 0;
          w1 = h1 && c1 >= 0;
        }
        const uint64_t m0 = __ballot(w0) & gm, m1 = __ballot(w1) & gm;
        const int cnt0 = __popcll(m0), tot = cnt0 + __popcll(m1);
        if (nn + tot > kF) {
          over = true;
          break;
        }
        if (w0) nxt[nn + __popcll(m0 & below)] = c0;
        if (w1) nxt[nn + cnt0 + __popcll(m1 & below)] = c1;
        nn += tot;
      }
      if (over) break;
      wave_lds_sync();
      int *t = cur;
      cur = nxt;
      nxt = t;
      n = nn;
    }
    const uint64_t odd = __ballot(cnt & 1) & gm, mb = __ballot(degenerate) & gm;
    // 3: the frontier overflowed; k_far_parity<.., true> redoes the point
    if (g == 0) far[pt] = over ? 3 : ((bad || mb) ? 2 : (uint8_t)(__popcll(odd) & 1));
    wave_lds_sync();
  }
}

template <bool BATCH, int G, bool AGAIN = false>  // AGAIN: only the points marked 3
__global__ __launch_bounds__(kBlock) void k_far_front(VisArgs a, uint8_t *__restrict__ far) {
  __shared__ int lds[(kBlock / G) * 2 * 16 * G];
  far_front_block<BATCH, G, AGAIN, kBlock>(a, far, blockIdx.x, gridDim.x, lds);
}

// The fan query and the first far-parity walks in one launch (they are
// independent): blocks below fan_blocks take 8 points' fans (breadth
// first, 8 lanes each), the others the far walks of 8 (G = 8) or 4
// (G = 16) points each.  One launch instead of two back to back.
// One launch for two independent per-point jobs: blocks [0, fan_blocks)
// the fan queries (with the clearances), the rest the far walks.
template <bool BATCH, int G>
__global__ __launch_bounds__(64) void k_far_fan(VisArgs a, uint8_t *__restrict__ far,
                                                unsigned fan_blocks, int *__restrict__ fan_n,
                                                int *__restrict__ fan_idx,
                                                FanRec *__restrict__ fan_rec,
                                                FanCone *__restrict__ fan_cone,
                                                float *__restrict__ clear) {
  // the frontiers and fan lists of the fan queries, or the far walks' frontiers
  __shared__ int lds[(64 / 8) * (2 * 16 * 8 + kFanMax)];
  static_assert((64 / G) * 2 * 16 * G <= (64 / 8) * 2 * 16 * 8, "far frontiers fit");
  if (blockIdx.x < fan_blocks) {
    fan_front_block<BATCH, 8, 64>(a, blockIdx.x, lds, fan_n, fan_idx, fan_rec, fan_cone, clear);
  } else {
    far_front_block<BATCH, G, false, 64>(a, far, blockIdx.x - fan_blocks,
                                         gridDim.x - fan_blocks, lds);
  }
}

// ------------------------------------------------------------ compaction
__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long v,
                                                              unsigned long long *sh,
                                                              unsigned long long &total) {
  // 256 threads
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    unsigned long long x = threadIdx.x >= off ? sh[threadIdx.x - off] : 0ull;
    __syncthreads();
    sh[threadIdx.x] += x;
    __syncthreads();
  }
  total = sh[255];
  const unsigned long long incl = sh[threadIdx.x];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(256) void k_word_scan(const uint64_t *__restrict__ bits,
                                                   int64_t nwords,
                                                   unsigned long long *__restrict__ block_tot) {
  __shared__ unsigned long long sh[256];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
  unsigned long long c = 0;
  for (int t = 0; t < 4; ++t)
    if (base + t < nwords) c += __popcll(bits[base + t]);
  unsigned long long total;
  block_excl_scan(c, sh, total);
  if (threadIdx.x == 0) block_tot[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_blocks(unsigned long long *__restrict__ v,
                                                      int64_t m, int64_t *__restrict__ count,
                                                      unsigned long long *stats_edges) {
  // count may be NULL (the walk queue: only its length in stats)
  __shared__ unsigned long long part[1024];
  const int64_t per = (m + 1023) / 1024;
  const int64_t beg = threadIdx.x * per;
  const int64_t end = beg + per < m ? beg + per : m;
  unsigned long long s = 0;
  for (int64_t i = beg; i < end; ++i) s += v[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    unsigned long long x = threadIdx.x >= off ? part[threadIdx.x - off] : 0ull;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  unsigned long long run = threadIdx.x ? part[threadIdx.x - 1] : 0ull;
  for (int64_t i = beg; i < end; ++i) {
    const unsigned long long x = v[i];
    v[i] = run;
    run += x;
  }
  if (threadIdx.x == 1023) {
    if (count) *count = (int64_t)part[1023];
    *stats_edges = part[1023];
  }
}

// The walk queue from the fan pass's per-word bits: global pair ids
// (word * 64 + bit) in k order, at offsets from k_word_scan/k_scan_blocks.
__global__ __launch_bounds__(256) void k_queue_emit(VisArgs a,
                                                    const unsigned long long *__restrict__ block_off) {
  __shared__ unsigned long long sh[256];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
  uint64_t w4[4], c4[4];
  unsigned long long c = 0;
  for (int t = 0; t < 4; ++t) {
    w4[t] = base + t < a.nwords ? a.qbits[base + t] : 0ull;
    c4[t] = base + t < a.nwords ? a.cbits[base + t] : 0ull;
    c += __popcll(w4[t]);
  }
  unsigned long long total;
  unsigned long long out = block_excl_scan(c, sh, total) + block_off[blockIdx.x];
  for (int t = 0; t < 4; ++t) {
    uint64_t m = w4[t];
    while (m) {
      const int b = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      a.queue[out++] = ((base + t) * 64 + b) | (((c4[t] >> b) & 1ull) ? kClip : 0);
    }
  }
}

template <bool BATCH>
__device__ __forceinline__ int part_index(const VisArgs &a, int64_t w) {
  if (!BATCH) return 0;
  int lo = 0, hi = a.nparts - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.parts[mid].word_begin <= w) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Pairs of every set bit in k order.  In a batch the pairs are local to
// their part and part_off[p] receives the index of part p's first edge.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_emit(VisArgs a,
                                              const unsigned long long *__restrict__ block_off,
                                              uint32_t *__restrict__ pairs, int64_t cap,
                                              int64_t *__restrict__ part_off) {
  __shared__ unsigned long long sh[256];
  const uint64_t *__restrict__ bits = a.bits;
  const int64_t nwords = a.nwords;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
  uint64_t w4[4];
  unsigned long long c = 0;
  for (int t = 0; t < 4; ++t) {
    w4[t] = base + t < nwords ? bits[base + t] : 0ull;
    c += __popcll(w4[t]);
  }
  unsigned long long total;
  unsigned long long out = block_excl_scan(c, sh, total) + block_off[blockIdx.x];
  for (int t = 0; t < 4; ++t) {
    if (base + t >= nwords) break;
    const int pi = part_index<BATCH>(a, base + t);
    const Part &P = BATCH ? a.parts[pi] : a.one;
    if (BATCH && P.word_begin == base + t) part_off[pi] = (int64_t)out;
    uint64_t m = w4[t];
    if (!m) continue;
    const int64_t n = P.n;
    int64_t k0 = (base + t - P.word_begin) * 64;
    int64_t i, j;
    pair_of(k0, n, i, j);
    int64_t kc = k0;
    while (m) {
      const int b = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      int64_t delta = k0 + b - kc;
      while (delta > 0) {
        const int64_t room = n - 1 - j;  // entries left in row i after j
        if (delta <= room) {
          j += delta;
          delta = 0;
        } else {
          delta -= room + 1;
          ++i;
          j = i + 1;
        }
      }
      kc = k0 + b;
      if ((int64_t)out < cap) {
        uint2 pr = make_uint2((uint32_t)i, (uint32_t)j);
        *reinterpret_cast<uint2 *>(pairs + 2 * out) = pr;
      }
      ++out;
    }
  }
}

}  // namespace

namespace {

__global__ void k_fix_part_offsets(VisArgs a, const int64_t *__restrict__ total,
                                   int64_t *__restrict__ part_off) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  part_off[a.nparts] = *total;
  for (int p = a.nparts - 1; p >= 0; --p)
    if (a.parts[p].S == 0) part_off[p] = part_off[p + 1];
}

struct Launch {
  unsigned grid, qgrid;
  unsigned long long *qscan;  // block offsets of the queue compaction
  int64_t nscan;              // its blocks
  int lv;     // log2 of the lanes per queued pair
  int front;  // lanes per pair of the breadth-first walks of short queues (0: off)
  bool cap;   // a mask of some part may carry cap-rule bits
  hipStream_t st;
};

template <bool N, bool B>
void launch_front(const Launch &L, const VisArgs &a) {
  constexpr unsigned kFrontGrid = 256 * kWalkWaves;  // every wave slot of the chip
  switch (L.front) {
    case 0: break;
    case 4: k_walk_front<N, B, 4, false><<<kFrontGrid, kBlock, 0, L.st>>>(a); break;
    case 8: k_walk_front<N, B, 8, false><<<kFrontGrid, kBlock, 0, L.st>>>(a); break;
    case 16: k_walk_front<N, B, 16, false><<<kFrontGrid, kBlock, 0, L.st>>>(a); break;
    case 32: k_walk_front<N, B, 32, false><<<kFrontGrid, kBlock, 0, L.st>>>(a); break;
    default: k_walk_front<N, B, 64, false><<<kFrontGrid, kBlock, 0, L.st>>>(a); break;
  }
}

template <int S, bool N, bool B>
void launch_split(const Launch &L, const VisArgs &a) {
  switch (L.lv) {
    case 0: k_walk_split<S, N, B, 0, false><<<L.qgrid, kBlock, 0, L.st>>>(a); break;
    case 1: k_walk_split<S, N, B, 1, false><<<L.qgrid, kBlock, 0, L.st>>>(a); break;
    case 2: k_walk_split<S, N, B, 2, false><<<L.qgrid, kBlock, 0, L.st>>>(a); break;
    case 3: k_walk_split<S, N, B, 3, false><<<L.qgrid, kBlock, 0, L.st>>>(a); break;
    default: k_walk_split<S, N, B, 4, false><<<L.qgrid, kBlock, 0, L.st>>>(a); break;
  }
}

// The fan pass queues what it cannot settle (no walk stack needed there);
// the queue is then walked densely.
template <int S, bool C, bool N, bool B>
void launch_main(bool q, const Launch &L, const VisArgs &a) {
  if (q && !C) {
    if (L.cap)  // no walk, no stack
      k_visibility<1, C, N, B, true, true><<<L.grid, kBlock, 0, L.st>>>(a);
    else
      k_visibility<1, C, N, B, true><<<L.grid, kBlock, 0, L.st>>>(a);
    // compact the per-word bits into the queue
    k_word_scan<<<(unsigned)L.nscan, 256, 0, L.st>>>(a.qbits, a.nwords, L.qscan);
    k_scan_blocks<<<1, 1024, 0, L.st>>>(L.qscan, L.nscan, nullptr, a.queue_len);
    k_queue_emit<<<(unsigned)L.nscan, 256, 0, L.st>>>(a, L.qscan);
    launch_front<N, B>(L, a);  // a short queue: breadth-first walks
    // a long one: depth-first walks (not launched when no queue can exceed
    // the breadth-first bound)
    if (L.front == 0 || a.front_max < a.nwords * 64) launch_split<S, N, B>(L, a);
  } else {  // the rare path (open parts, cages): one instance, cap rule compiled in
    k_visibility<S, C, N, B, false, true><<<L.grid, kBlock, 0, L.st>>>(a);
  }
}
template <int S, bool C, bool N>
void pick_b(bool b, bool q, const Launch &L, const VisArgs &a) {
  if (b) launch_main<S, C, N, true>(q, L, a);
  else launch_main<S, C, N, false>(q, L, a);
}
template <int S, bool C>
void pick_n(bool n, bool b, bool q, const Launch &L, const VisArgs &a) {
  if (n) pick_b<S, C, true>(b, q, L, a);
  else pick_b<S, C, false>(b, q, L, a);
}
template <int S>
void pick_c(bool c, bool n, bool b, bool q, const Launch &L, const VisArgs &a) {
  if (c) pick_n<S, true>(n, b, q, L, a);
  else pick_n<S, false>(n, b, q, L, a);
}

Part make_part(const float *d_pts, int64_t n, const uint8_t *d_mask, const vacd_bvh *self,
               const vacd_bvh *cage) {
  Part P;
  P.pts = d_pts;
  P.mask = d_mask;
  P.far = nullptr;
  P.fan_n = nullptr;
  P.fan_idx = nullptr;
  P.fan_rec = nullptr;
  P.fan_cone = nullptr;
  P.clear = nullptr;
  P.clear_r = 0.0f;
  P.snodes = self->nodes.as<Node>();
  P.snodes4 = self->nodes4.as<Node4>();
  P.stris = self->tris.as<Tri>();
  P.cnodes = cage ? cage->nodes.as<Node>() : nullptr;
  P.ctris = cage ? cage->tris.as<Tri>() : nullptr;
  P.n = n;
  P.S = n >= 2 ? n * (n - 1) / 2 : 0;
  P.word_begin = 0;
  P.point_begin = 0;
  P.scap = (int)std::min<int64_t>(2 * self->n_tris + 64, INT32_MAX);
  P.rgroups = nullptr;
  if (self->n_refs > 0) {
    P.rnodes = self->rnodes.as<Node>();
    P.rtris = self->rtris.as<Tri>();
    P.rcap = (int)std::min<int64_t>(2 * self->n_refs + 64, INT32_MAX);
    if (self->n_groups > 0) P.rgroups = self->groups.as<float4>();
  } else {
    P.rnodes = P.snodes;
    P.rtris = P.stris;
    P.rcap = P.scap;
  }
  P.ccap = cage ? (int)std::min<int64_t>(2 * cage->n_tris + 64, INT32_MAX) : 0;
  P.group_axis = self->group_axis;
  return P;
}

// Stack entries a binary walk of a tree can hold: depth - 1, one pending
// sibling per level.  In a small-first tree child 0 of every node has no more
// leaves than child 1 and the walks push child 1 while descending into child
// 0, so with s entries pushed the current subtree holds at most n / 2^s
// leaves: s <= log2(n) - 1.
int binary_stack(int depth, int64_t leaves, bool small_first) {
  if (leaves <= 0) return 0;
  if (!small_first) return depth - 1;
  int lg = 0;
  while ((int64_t(2) << lg) <= leaves) ++lg;  // floor(log2(leaves))
  return std::min(depth - 1, lg);
}

// Search radius of the points' clearances on a mesh: three times its extent
// over sqrt(triangles), about a mean edge length of a closed surface mesh
// (the nearest triangle not at a vertex is typically 0.8 of one away).  Any
// radius gives valid clearances (capped at it); a larger one finds a little
// more clearance at the cost of a wider search.
float clear_radius(const vacd_bvh *b) {
  const float ext = std::max({b->bounds[3] - b->bounds[0], b->bounds[4] - b->bounds[1],
                              b->bounds[5] - b->bounds[2]});
  return b->n_tris > 0 ? 3.0f * ext / std::sqrt((float)b->n_tris) : 0.0f;
}

// Enqueues far parity, the main kernel, the exact redo and the compaction for
// a prepared VisArgs (a.nwords > 0).
void run(vacd_ctx *ctx, VisArgs &a, bool batch, bool any_cage, bool any_far, bool cap_bits,
         int64_t max_tris, int depth, int stack2,
         uint32_t *d_pairs, int64_t cap, int64_t *d_count, int64_t *d_part_off) {
  hipStream_t st = ctx->stream;
  const int64_t nwords = a.nwords;
  vacd::DevBuf &bits = ctx->buf[SLOT_BITS];
  vacd::DevBuf &dlist = ctx->buf[SLOT_DEFER];
  vacd::DevBuf &scan = ctx->buf[SLOT_SCAN];
  bits.ensure(8 * (size_t)nwords);
  dlist.ensure(8 * (size_t)nwords * 64);
  const int64_t nscan = (nwords + kScanTile - 1) / kScanTile;
  scan.ensure(8 * (size_t)nscan);
  a.bits = bits.as<uint64_t>();
  a.dlist = dlist.as<int64_t>();
  a.dlist_len = a.stats + 8;
  VACD_REQUIRE(depth <= 64, VACD_ERR_LIMIT, "BVH deeper than the traversal stack");
  // stack entries a walk can hold: stack2 (see binary_stack); the fan
  // query's 4-wide walk, up to three per visited level (every other binary
  // level)
  const int need2 = stack2, need4 = 3 * (depth / 2);
  // parity mode with fans: queue the pairs the fan pass leaves open
  const bool queue = any_far && ctx->fan_pairs && !any_cage && need4 <= 64;
  // the points' clearances (parts with fans), for the walks of queued pairs
  // that leave both ends outwards
  const bool want_clear = queue && ctx->clip_min > 0;
  float *clear = ctx->buf[SLOT_CLEAR].as<float>();
  if (any_far) {
    const unsigned fg = (unsigned)((a.npoints + 63) / 64);
    const unsigned pg = (unsigned)std::min<int64_t>(a.npoints, 16384);  // a wave per point
    vacd::DevBuf &farb = ctx->buf[SLOT_FAR];
    vacd::DevBuf &fanb = ctx->buf[SLOT_FAN];
    int *fan_n = fanb.as<int>();
    int *fan_idx = fan_n + a.npoints;
    FanRec *fan_rec = ctx->buf[SLOT_FANREC].as<FanRec>();
    FanCone *fan_cone = reinterpret_cast<FanCone *>(fan_rec + (size_t)a.npoints * kFanMax);
    // far parity breadth first, 8 lanes per point, with the fan query in
    // the same launch (a wave per point, the walk split by subtree, and the
    // fan query on its own, when the breadth-first walks are off).
    // A ray across a mesh of T triangles meets about T^(2/3) boxes at the
    // deeper levels: meshes above 131,072 triangles take 16 lanes per point
    // (a frontier of 256 nodes) so that fewer points overflow to the
    // depth-first redo.
    const int fl = max_tris > 131072 ? 16 : 8;
    const unsigned far_blocks =
        (unsigned)std::min<int64_t>((a.npoints + 64 / fl - 1) / (64 / fl), 16384);
    const unsigned fan_blocks = ctx->fan_pairs ? (unsigned)((a.npoints + 7) / 8) : 0;
    // points whose frontier overflows: again with a wave per point (a
    // frontier of 1024 nodes), then depth first
    const unsigned fpg64 = (unsigned)std::min<int64_t>((a.npoints + kWaves - 1) / kWaves, 4096);
    if (ctx->front_lanes) {
      // the fan queries also measure the clearances (want_clear: the parts'
      // clear pointers are set)
      const unsigned grid = fan_blocks + far_blocks;  // > 0: a closed part has points
      if (batch && fl == 16)
        k_far_fan<true, 16><<<grid, 64, 0, st>>>(a, farb.as<uint8_t>(), fan_blocks, fan_n, fan_idx, fan_rec, fan_cone, clear);
      else if (batch)
        k_far_fan<true, 8><<<grid, 64, 0, st>>>(a, farb.as<uint8_t>(), fan_blocks, fan_n, fan_idx, fan_rec, fan_cone, clear);
      else if (fl == 16)
        k_far_fan<false, 16><<<grid, 64, 0, st>>>(a, farb.as<uint8_t>(), fan_blocks, fan_n, fan_idx, fan_rec, fan_cone, clear);
      else
        k_far_fan<false, 8><<<grid, 64, 0, st>>>(a, farb.as<uint8_t>(), fan_blocks, fan_n, fan_idx, fan_rec, fan_cone, clear);
      if (fan_blocks) {
        if (batch) k_fan_redo<true><<<fg, 64, 0, st>>>(a, fan_n, fan_idx, fan_rec, fan_cone);
        else k_fan_redo<false><<<fg, 64, 0, st>>>(a, fan_n, fan_idx, fan_rec, fan_cone);
      }
      if (batch) {
        k_far_front<true, 64, true><<<fpg64, kBlock, 0, st>>>(a, farb.as<uint8_t>());
        k_far_parity<true, true><<<pg, 64, 0, st>>>(a, farb.as<uint8_t>());
      } else {
        k_far_front<false, 64, true><<<fpg64, kBlock, 0, st>>>(a, farb.as<uint8_t>());
        k_far_parity<false, true><<<pg, 64, 0, st>>>(a, farb.as<uint8_t>());
      }
    } else {
      if (batch) k_far_parity<true><<<pg, 64, 0, st>>>(a, farb.as<uint8_t>());
      else k_far_parity<false><<<pg, 64, 0, st>>>(a, farb.as<uint8_t>());
      if (want_clear) {
        if (batch) k_clear<true><<<fg, 64, 0, st>>>(a, clear);
        else k_clear<false><<<fg, 64, 0, st>>>(a, clear);
      }
    }
    VACD_HIP(hipGetLastError());
    if (ctx->fan_pairs && !ctx->front_lanes) {
      if (batch)
        k_fan<true><<<fg, 64, 0, st>>>(a, fan_n, fan_idx, fan_rec, fan_cone);
      else
        k_fan<false><<<fg, 64, 0, st>>>(a, fan_n, fan_idx, fan_rec, fan_cone);
      VACD_HIP(hipGetLastError());
    }
  }
  // the fan pass: chunks of up to VACD_FAN_CHUNK_MAX consecutive words per
  // wave, as many as keep 16 blocks per CU busy
  a.fan_chunk = (int)std::max<int64_t>(1, std::min<int64_t>(VACD_FAN_CHUNK_MAX, nwords / (256 * 16 * kWaves)));
  const int64_t want_blocks = (nwords + kWaves * a.fan_chunk - 1) / (kWaves * a.fan_chunk);
  Launch L;
  L.grid = (unsigned)std::min<int64_t>(want_blocks, 256 * 16);
  L.lv = ctx->lanes_per_pair_log2;
  L.front = ctx->front_lanes;
  L.cap = cap_bits;
  a.front_max = L.front ? ctx->front_max : -1;
  a.rotated_rays = ctx->rotated_rays;
  // the queue length is known on the device only: enough groups to cover
  // every pair once when all are queued, at most 16 blocks per CU
  L.qgrid = (unsigned)std::min<int64_t>(((nwords * 64) << L.lv) / kBlock + 1, 256 * 16);
  L.st = st;
  a.queue_len = a.stats + 7;
  a.queue = nullptr;
  a.qbits = nullptr;
  L.qscan = scan.as<unsigned long long>();  // free until the edge scan
  L.nscan = nscan;
  a.cbits = nullptr;
  if (queue) {
    vacd::DevBuf &qb = ctx->buf[SLOT_QUEUE];
    qb.ensure(8 * (size_t)nwords * 66);  // the queue, then the two per-word bit arrays
    a.queue = qb.as<int64_t>();
    a.qbits = reinterpret_cast<uint64_t *>(a.queue + nwords * 64);
    a.cbits = a.qbits + nwords;
  }
  timed_begin(ctx, TIMED_VIS);
  if (need2 <= 32)
    pick_c<32>(any_cage, ctx->count_work, batch, queue, L, a);
  else
    pick_c<64>(any_cage, ctx->count_work, batch, queue, L, a);
  VACD_HIP(hipGetLastError());
  timed_end(ctx, TIMED_VIS);
  // the deferred pairs (the list length is read on the device): a wave per
  // pair for the first kExactSplit, a lane per pair for the rest.  One flag,
  // set here where the launches are chosen, tells the lane-per-pair kernel
  // which entries the wave-per-pair kernel took.
  a.exact_front = L.front != 0;
  const unsigned wgrid = kExactGrid;
  const unsigned xgrid = (unsigned)std::min<int64_t>((nwords * 64 + kBlock - 1) / kBlock, 1024);
  if (cap_bits) {  // the exact cap rule before the exact walks
    if (batch) k_cap_redo<true><<<xgrid, 256, 0, st>>>(a);
    else k_cap_redo<false><<<xgrid, 256, 0, st>>>(a);
  }
  if (a.exact_front) {  // breadth first (a wave per pair) before the depth-first kernels
    if (batch) k_walk_front<false, true, 64, true><<<wgrid, kBlock, 0, st>>>(a);
    else k_walk_front<false, false, 64, true><<<wgrid, kBlock, 0, st>>>(a);
  }
  const bool wave = !a.exact_front;  // the depth-first wave-per-pair instance
  if (need2 <= 32) {
    if (batch) {
      if (wave) k_walk_split<32, false, true, 6, true><<<wgrid, kBlock, 0, st>>>(a);
      k_walk_split<32, false, true, 0, true><<<xgrid, kBlock, 0, st>>>(a);
    } else {
      if (wave) k_walk_split<32, false, false, 6, true><<<wgrid, kBlock, 0, st>>>(a);
      k_walk_split<32, false, false, 0, true><<<xgrid, kBlock, 0, st>>>(a);
    }
  } else {
    if (batch) {
      if (wave) k_walk_split<64, false, true, 6, true><<<wgrid, kBlock, 0, st>>>(a);
      k_walk_split<64, false, true, 0, true><<<xgrid, kBlock, 0, st>>>(a);
    } else {
      if (wave) k_walk_split<64, false, false, 6, true><<<wgrid, kBlock, 0, st>>>(a);
      k_walk_split<64, false, false, 0, true><<<xgrid, kBlock, 0, st>>>(a);
    }
  }
  VACD_HIP(hipGetLastError());
  unsigned long long *blk = scan.as<unsigned long long>();
  k_word_scan<<<(unsigned)nscan, 256, 0, st>>>(a.bits, nwords, blk);
  k_scan_blocks<<<1, 1024, 0, st>>>(blk, nscan, d_count, a.stats + 3);
  if (batch) {
    k_emit<true><<<(unsigned)nscan, 256, 0, st>>>(a, blk, d_pairs, cap, d_part_off);
    k_fix_part_offsets<<<1, 64, 0, st>>>(a, d_count, d_part_off);
  } else {
    k_emit<false><<<(unsigned)nscan, 256, 0, st>>>(a, blk, d_pairs, cap, nullptr);
  }
  VACD_HIP(hipGetLastError());
}

void reset_stats(vacd_ctx *ctx) {
  vacd::DevBuf &stats = ctx->buf[SLOT_STATS];
  stats.ensure(16 * sizeof(unsigned long long));
  fill_u64(ctx, stats.ptr, 16, 0ull);  // a kernel: the pass may be captured in a graph
}

}  // namespace

void launch_visibility(vacd_ctx *ctx, const float *d_pts, int64_t n, const uint8_t *d_mask,
                       const vacd_bvh *self, const vacd_bvh *cage, uint32_t *d_pairs, int64_t cap,
                       int64_t *d_count) {
  VACD_REQUIRE(n >= 0 && n <= (int64_t(1) << 26), VACD_ERR_LIMIT, "visibility needs n <= 2^26");
  VACD_REQUIRE(self != nullptr, VACD_ERR_INVALID, "self BVH is required");
  reset_stats(ctx);
  VisArgs a;
  a.one = make_part(d_pts, n, d_mask, self, cage);
  a.parts = nullptr;
  a.nparts = 1;
  a.nwords = (a.one.S + 63) / 64;
  a.npoints = n;
  a.stats = ctx->buf[SLOT_STATS].as<unsigned long long>();
  ctx->last_segments = a.one.S;
  if (a.nwords == 0) {
    fill_u64(ctx, d_count, 1, 0ull);
    return;
  }
  const bool any_far = !cage && self->closed;
  const int depth = std::max(std::max(self->depth, self->depth_refs), cage ? cage->depth : 0);
  if (any_far) {
    vacd::DevBuf &farb = ctx->buf[SLOT_FAR];
    farb.ensure((size_t)n);
    a.one.far = farb.as<uint8_t>();
    if (ctx->fan_pairs && 3 * (depth / 2) <= 64) {  // the fan query walks the 4-wide tree
      vacd::DevBuf &fanb = ctx->buf[SLOT_FAN];
      fanb.ensure(sizeof(int) * (size_t)n * (1 + kFanMax));
      a.one.fan_n = fanb.as<int>();
      a.one.fan_idx = fanb.as<int>() + n;
      vacd::DevBuf &recb = ctx->buf[SLOT_FANREC];
      recb.ensure((sizeof(FanRec) * kFanMax + sizeof(FanCone)) * (size_t)n);  // records, cones
      a.one.fan_rec = recb.as<FanRec>();
      a.one.fan_cone = reinterpret_cast<const FanCone *>(a.one.fan_rec + (size_t)n * kFanMax);
      if (ctx->clip_min > 0 && a.one.S >= ctx->clip_min) {
        vacd::DevBuf &clb = ctx->buf[SLOT_CLEAR];
        clb.ensure(sizeof(float) * (size_t)n);
        a.one.clear = clb.as<float>();
        a.one.clear_r = clear_radius(self);
      }
    }
  }
  const int stack2 = std::max({binary_stack(self->depth, self->n_tris, self->small_first),
                               binary_stack(self->depth_refs, self->n_refs, self->small_first_refs),
                               cage ? binary_stack(cage->depth, cage->n_tris, cage->small_first) : 0});
  // the cap rule needs a mask and planar groups in the self BVH
  const bool cap_bits = d_mask != nullptr && self->group_axis != 0xffffffffu;
  run(ctx, a, false, cage != nullptr, any_far, cap_bits, self->n_tris, depth, stack2, d_pairs, cap, d_count,
      nullptr);
}

void launch_visibility_batch(vacd_ctx *ctx, const vacd_part *parts, int nparts, uint32_t *d_pairs,
                             int64_t cap, int64_t *d_count, int64_t *d_part_off) {
  VACD_REQUIRE(nparts >= 1, VACD_ERR_INVALID, "empty batch");
  reset_stats(ctx);
  std::vector<Part> hp((size_t)nparts);
  int64_t words = 0, points = 0, segs = 0;
  int depth = 0, stack2 = 0;
  int64_t max_tris = 0;
  bool any_cage = false, any_far = false;
  for (int p = 0; p < nparts; ++p) {
    const vacd_part &q = parts[p];
    VACD_REQUIRE(q.self != nullptr, VACD_ERR_INVALID, "part without a self BVH");
    VACD_REQUIRE(q.n >= 0 && q.n <= (int64_t(1) << 26), VACD_ERR_LIMIT,
                 "visibility needs n <= 2^26");
    VACD_REQUIRE(q.n < 2 || q.pts, VACD_ERR_INVALID, "NULL points");
    VACD_REQUIRE(q.self->device == ctx->device && (!q.cage || q.cage->device == ctx->device),
                 VACD_ERR_INVALID, "BVH built on another device");
    Part P = make_part(q.pts, q.n, q.new_mask, q.self, q.cage);
    P.word_begin = words;
    P.point_begin = points;
    words += (P.S + 63) / 64;
    points += q.n;
    segs += P.S;
    depth = std::max({depth, q.self->depth, q.self->depth_refs, q.cage ? q.cage->depth : 0});
    stack2 = std::max({stack2, binary_stack(q.self->depth, q.self->n_tris, q.self->small_first),
                       binary_stack(q.self->depth_refs, q.self->n_refs, q.self->small_first_refs),
                       q.cage ? binary_stack(q.cage->depth, q.cage->n_tris, q.cage->small_first) : 0});
    any_cage |= q.cage != nullptr;
    any_far |= !q.cage && q.self->closed;
    max_tris = std::max(max_tris, q.self->n_tris);
    hp[p] = P;
  }
  ctx->last_segments = segs;
  if (any_far) {
    vacd::DevBuf &farb = ctx->buf[SLOT_FAR];
    farb.ensure((size_t)std::max<int64_t>(points, 1));
    const bool fan = ctx->fan_pairs && 3 * (depth / 2) <= 64;
    vacd::DevBuf &fanb = ctx->buf[SLOT_FAN];
    vacd::DevBuf &recb = ctx->buf[SLOT_FANREC];
    vacd::DevBuf &clb = ctx->buf[SLOT_CLEAR];
    if (fan) {
      fanb.ensure(sizeof(int) * (size_t)std::max<int64_t>(points, 1) * (1 + kFanMax));
      recb.ensure((sizeof(FanRec) * kFanMax + sizeof(FanCone)) *
                  (size_t)std::max<int64_t>(points, 1));
      if (ctx->clip_min > 0) clb.ensure(sizeof(float) * (size_t)std::max<int64_t>(points, 1));
    }
    for (int p = 0; p < nparts; ++p)
      if (!hp[p].ctris && parts[p].self->closed) {
        hp[p].far = farb.as<uint8_t>() + hp[p].point_begin;
        if (fan) {
          hp[p].fan_n = fanb.as<int>() + hp[p].point_begin;
          hp[p].fan_idx = fanb.as<int>() + points + hp[p].point_begin * kFanMax;
          hp[p].fan_rec = recb.as<FanRec>() + hp[p].point_begin * kFanMax;
          hp[p].fan_cone = reinterpret_cast<const FanCone *>(recb.as<FanRec>() +
                                                             (size_t)points * kFanMax) +
                           hp[p].point_begin;
          if (ctx->clip_min > 0 && hp[p].S >= ctx->clip_min) {
            hp[p].clear = clb.as<float>() + hp[p].point_begin;
            hp[p].clear_r = clear_radius(parts[p].self);
          }
        }
      }
  }
  // the part table goes through pinned memory, which a captured copy reads
  // again at every replay
  vacd::DevBuf &pb = ctx->buf[SLOT_PARTS];
  pb.ensure(sizeof(Part) * (size_t)nparts);
  std::memcpy(pinned_stage(ctx, PINNED_VIS, sizeof(Part) * (size_t)nparts), hp.data(),
              sizeof(Part) * (size_t)nparts);
  pinned_upload(ctx, PINNED_VIS, pb.ptr, sizeof(Part) * (size_t)nparts);
  VisArgs a;
  a.one = hp[0];
  a.parts = pb.as<Part>();
  a.nparts = nparts;
  a.nwords = words;
  a.npoints = points;
  a.stats = ctx->buf[SLOT_STATS].as<unsigned long long>();
  if (words == 0) {
    fill_u64(ctx, d_count, 1, 0ull);
    fill_u64(ctx, d_part_off, nparts + 1, 0ull);
    return;
  }
  bool cap_bits = false;
  for (int p = 0; p < nparts; ++p)
    cap_bits |= parts[p].new_mask != nullptr && parts[p].self->group_axis != 0xffffffffu;
  run(ctx, a, true, any_cage, any_far, cap_bits, max_tris, depth, stack2, d_pairs, cap, d_count,
      d_part_off);
}

}  // namespace vacd
