// Internal declarations shared by the HIP translation units of libvacd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/vacd.h"

namespace vacd {

// ---------------------------------------------------------------- layouts
// BVH node (64 B): boxes of both children plus their ids.  A child id >= 0
// is an internal node, a negative id c is leaf ~c (a leaf-ordered triangle).
//   f0 = lo0.x lo0.y lo0.z hi0.x   f1 = hi0.y hi0.z lo1.x lo1.y
//   f2 = lo1.z hi1.x hi1.y hi1.z   f3 = c0 c1 (as int bits) pad pad
struct alignas(16) Node {
  float4 f0, f1, f2, f3;
};
static_assert(sizeof(Node) == 64, "node is 64 bytes");

// 4-wide node (128 B), structure of arrays over its (up to) four children:
//   f[0..2] = lo.x, lo.y, lo.z of children 0..3
//   f[3..5] = hi.x, hi.y, hi.z of children 0..3
//   f[6]    = child ids (int bits; same encoding as Node), f[7] unused
// Node4 x holds the grandchildren of binary node x (a leaf child stays a
// child); empty slots carry an inverted box.
struct alignas(16) Node4 {
  float4 f[8];
};
static_assert(sizeof(Node4) == 128, "node4 is 128 bytes");

// Box coordinate of an empty child slot: the degenerate box at (1e30, 1e30,
// 1e30) lies outside every scene, so both the slab test and the +x ray test
// reject it (an inverted box would not be rejected by the slab test).
constexpr float kEmpty = 1e30f;

// Leaf-ordered triangle (48 B): a.xyz + original id bits, b.xyz, c.xyz +
// planar group code bits (0 none, g + 1: group g; plain tree only).
struct alignas(16) Tri {
  float4 a, b, c;
};
static_assert(sizeof(Tri) == 48, "triangle is 48 bytes");

// ---------------------------------------------------------------- errors
struct Error {
  int code;
  std::string msg;
};

#define VACD_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::vacd::Error{VACD_ERR_HIP, std::string(#call) + ": " +           \
                                            hipGetErrorString(e_)};           \
  } while (0)

#define VACD_REQUIRE(cond, code, msg)                                         \
  do {                                                                        \
    if (!(cond)) throw ::vacd::Error{(code), (msg)};                          \
  } while (0)

// Grow-only device buffer owned by a context or a BVH.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  void ensure(size_t n) {
    if (n <= bytes) return;
    if (ptr) VACD_HIP(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    size_t want = n < 256 ? 256 : n;
    VACD_HIP(hipMalloc(&ptr, want));
    bytes = want;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(ptr);
  }
};

}  // namespace vacd

// Pool slots used by the entry points.
enum VacdSlot {
  SLOT_PTS = 0,
  SLOT_MASK,
  SLOT_PAIRS,
  SLOT_COUNT,
  SLOT_BITS,
  SLOT_DEFER,
  SLOT_SCAN,
  SLOT_STATS,
  SLOT_PLANES,
  SLOT_EDGES,
  SLOT_EREC,
  SLOT_SCORES,
  SLOT_SCORES_F,
  SLOT_ARGMAX,
  SLOT_VERTS,
  SLOT_TRIS,
  SLOT_SORT,
  SLOT_MISC,
  SLOT_FAR,
  SLOT_PARTS,
  SLOT_FAN,
  SLOT_QUEUE,
  SLOT_SLICE,
  SLOT_REFS,
  SLOT_SCORE_TABLES,
  SLOT_MLP_IN,
  SLOT_GROUPS,
  SLOT_DIST,
  SLOT_EXTENT,
  SLOT_GRID,
  SLOT_FANREC,
  SLOT_SCANTMP,
  SLOT_CLEAR,
  SLOT_COUNT_MAX
};

struct vacd_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::string last_error;
  vacd::DevBuf buf[SLOT_COUNT_MAX];
  int64_t last_segments = 0;
  bool timing = false;
  bool count_work = false;
  bool fan_pairs = true;  // decide pairs that start inside from the fan of p
  int64_t score_chunk = 0;  // edges per scoring block (0: sized for ~4096 blocks)
  int lanes_per_pair_log2 = 0;  // queued-pair walks split over 2^k lanes (k <= 4)
  // walk queues of at most front_max pairs are walked breadth-first with
  // front_lanes lanes per pair (0: never); longer ones depth first.  The
  // default walks every queue breadth first: on the 16,384-point parts of a
  // decomposition that measured faster at every queue length seen
  int front_lanes = 8;
  int64_t front_max = int64_t(1) << 40;
  // breadth-first walks of pairs that leave both ends outwards test only the
  // part of the segment outside both endpoints' clearance balls, in parts of
  // at least clip_min segments (0: never).  Clipping pays where the walk
  // queue is long enough to keep the chip busy (the walks' total work
  // counts); on smaller parts the queue's time is the latency of its
  // longest walks and the clearance query would only add to the pass.
  int64_t clip_min = int64_t(1) << 25;
  // an undecided +x midpoint ray of a closed part is tried along +y and +z
  // (the parity of a point off a closed surface is the same along any ray)
  // before the pair goes to the exact redo
  bool rotated_rays = true;
  // recorded (start, stop) event pairs per timed kernel kind
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[3];
  // pinned host staging for descriptor tables copied to the device, one
  // region per table kind; a copy captured into a graph reads its region
  // again at every replay
  void *pinned[3] = {nullptr, nullptr, nullptr};
  size_t pinned_bytes[3] = {0, 0, 0};
  hipEvent_t pinned_free[3] = {nullptr, nullptr, nullptr};  // after the last copy out of it
};

namespace vacd {
enum PinnedKind { PINNED_VIS = 0, PINNED_SCORE = 1, PINNED_DIST = 2 };

inline bool capturing(const vacd_ctx *ctx) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  VACD_HIP(hipStreamIsCapturing(ctx->stream, &st));
  return st != hipStreamCaptureStatusNone;
}

// Pinned staging region `k` of at least `bytes`, once the copies enqueued
// from it earlier have completed.  While the stream is being captured the
// region must already be large enough (a warm-up call sizes it), and it is
// not waited on: the graph's copy reads it at replay time.
inline void *pinned_stage(vacd_ctx *ctx, int k, size_t bytes) {
  const bool cap = capturing(ctx);
  if (!cap && ctx->pinned_free[k]) VACD_HIP(hipEventSynchronize(ctx->pinned_free[k]));
  if (bytes > ctx->pinned_bytes[k]) {
    VACD_REQUIRE(!cap, VACD_ERR_INVALID, "staging buffer too small while capturing: warm up first");
    if (ctx->pinned[k]) VACD_HIP(hipHostFree(ctx->pinned[k]));
    ctx->pinned[k] = nullptr;
    ctx->pinned_bytes[k] = 0;
    const size_t want = bytes < 4096 ? 4096 : bytes;
    VACD_HIP(hipHostMalloc(&ctx->pinned[k], want, hipHostMallocDefault));
    ctx->pinned_bytes[k] = want;
  }
  return ctx->pinned[k];
}
// Copies `bytes` from pinned region k to dst on the context stream.
inline void pinned_upload(vacd_ctx *ctx, int k, void *dst, size_t bytes) {
  VACD_HIP(hipMemcpyAsync(dst, ctx->pinned[k], bytes, hipMemcpyHostToDevice, ctx->stream));
  if (capturing(ctx)) return;
  if (!ctx->pinned_free[k])
    VACD_HIP(hipEventCreateWithFlags(&ctx->pinned_free[k], hipEventDisableTiming));
  VACD_HIP(hipEventRecord(ctx->pinned_free[k], ctx->stream));
}

enum TimedKind { TIMED_VIS = 0, TIMED_SCORE = 1, TIMED_BVH = 2 };
// Records a start event when timing is on; returns the index to close.
inline void timed_begin(vacd_ctx *ctx, int kind) {
  if (!ctx->timing) return;
  hipEvent_t a, b;
  VACD_HIP(hipEventCreate(&a));
  VACD_HIP(hipEventCreate(&b));
  VACD_HIP(hipEventRecord(a, ctx->stream));
  ctx->ev[kind].push_back({a, b});
}
inline void timed_end(vacd_ctx *ctx, int kind) {
  if (!ctx->timing || ctx->ev[kind].empty()) return;
  VACD_HIP(hipEventRecord(ctx->ev[kind].back().second, ctx->stream));
}
}  // namespace vacd

struct vacd_bvh {
  int device = 0;
  int64_t n_tris = 0;
  int32_t depth = 0;
  // trees deeper than 33 levels store the child with fewer leaves first
  bool small_first = false, small_first_refs = false;
  // every undirected edge (by vertex index) has an even number of incident
  // triangles: crossing parity is then the same from any point of an open
  // segment that touches no triangle
  bool closed = false;
  float bounds[6] = {0, 0, 0, 0, 0, 0};
  vacd::DevBuf nodes;   // Node[max(1, n_tris-1)]
  vacd::DevBuf nodes4;  // Node4[max(1, n_tris-1)]
  vacd::DevBuf tris;    // Tri[n_tris]
  // Contact tree: a triangle much longer than the mesh's typical edge is
  // stored as slices across its length, each slice a leaf with its own box
  // referring to the whole triangle (b.w = slice | slices << 8 | apex << 16
  // as int bits), so that a long sliver no longer overlaps every query near
  // it.  Used by the any-contact walks only (a triangle met twice cannot
  // change a contact answer); crossing counts use nodes/tris.  n_refs = 0:
  // no triangle needed slicing, the contact walks use nodes/tris.
  int64_t n_refs = 0;
  int32_t depth_refs = 0;
  vacd::DevBuf rnodes;  // Node[max(1, n_refs-1)]
  vacd::DevBuf rtris;   // Tri[n_refs]
  // Planar groups of the contact tree: group g's triangles lie within tau of
  // the plane n.x + d = 0 (|n| = 1); float4 pairs (n.x, n.y, n.z, d),
  // (tau, 0, 0, 0).  A segment can touch them only where it is within tau of
  // that plane, so the walks test the boxes of a group's subtree against
  // that part of the segment only.
  int n_groups = 0;
  vacd::DevBuf groups;
  // Projection of each planar group's plane for 2-D tests within it: 2 bits
  // per group g at bit 2g, the axis dropped (the largest component of the
  // plane normal: 0 x, 1 y, 2 z); 3 for a group dropped by its tolerance.
  uint32_t group_axis = 0xffffffffu;
  void release() {
    nodes.release();
    nodes4.release();
    tris.release();
    rnodes.release();
    rtris.release();
    groups.release();
  }
};

namespace vacd {

// LBVH build from device vertices/triangles (lbvh.hip).  d_grp (nullable):
// per triangle its planar group code, 0 = none, 1..7 = group 0..6.
void build_lbvh(vacd_ctx *ctx, const float *d_vertices, int64_t n_vertices,
                const uint32_t *d_tris, int64_t n_tris, vacd_bvh *out,
                const uint8_t *d_grp = nullptr);

// Visibility (visibility.hip): enqueue everything; writes pairs and count.
void launch_visibility(vacd_ctx *ctx, const float *d_pts, int64_t n,
                       const uint8_t *d_mask, const vacd_bvh *self,
                       const vacd_bvh *cage, uint32_t *d_pairs, int64_t cap,
                       int64_t *d_count);

void launch_visibility_batch(vacd_ctx *ctx, const vacd_part *parts, int nparts, uint32_t *d_pairs,
                             int64_t cap, int64_t *d_count, int64_t *d_part_off);

// Largest point-to-mesh distance (distance.hip); writes float bits.
void launch_max_distance(vacd_ctx *ctx, const float *d_pts, int64_t m, const vacd_bvh *mesh,
                         unsigned *d_out_bits);
// Same for many (points, mesh) queries at once, every point against every
// triangle of its mesh (no BVH); host arrays in, out[nq] written on return.
void max_distance_batch(vacd_ctx *ctx, const vacd_distance_query *q, int nq, float *out);
// Extents of point sets along unit normals, float64 (scoring.hip).
void plane_extents_batch(vacd_ctx *ctx, const vacd_extent_query *q, int nq, double *lo_hi);

// Scoring (scoring.hip).
void launch_scoring(vacd_ctx *ctx, const float *d_planes, int64_t P,
                    const float *d_pts, int64_t n, const uint32_t *d_edges,
                    int64_t E, int64_t *d_q32, float *d_f, int32_t *d_argmax);
void launch_scoring_batch(vacd_ctx *ctx, const vacd_score_part *parts, int nparts,
                          int64_t *d_q32, int32_t *d_argmax);
// Allocation-free scoring (scoring.hip): noalloc_setup looks up, once per
// device and with that device current, the address of the library's fixed
// scoring scratch array; launch_scoring_noalloc then scores with it.
void noalloc_setup(int device);
void launch_scoring_noalloc(vacd_ctx *ctx, const float *d_planes, int64_t P, const float *d_pts,
                            int64_t n, const uint32_t *d_edges, int64_t E, int64_t *d_q32,
                            float *d_f, int32_t *d_argmax);

// bf16 GEMM on MFMA (mlp.hip): C = op(A) op(B) (+ bias, relu), fp32 and/or
// bf16 output; A(m, k) = ta ? A[k * lda + m] : A[m * lda + k], B(k, n) = tb ?
// B[n * ldb + k] : B[k * ldb + n]; acc adds into the fp32 output.
void gemm_bf16(vacd_ctx *ctx, bool ta, bool tb, int M, int N, int K, const void *A, int lda,
               const void *B, int ldb, const float *bias, bool relu, bool acc, float *Cf,
               void *Cb, int ldc);
void to_bf16(vacd_ctx *ctx, const float *x, int64_t n, void *y);

// Learned plane proposal model (mlp.hip).
int64_t MlpLayoutCount();
void mlp_destroy(vacd_mlp *M);
void mlp_create(vacd_ctx *ctx, uint64_t seed, vacd_mlp **out);
void mlp_set(vacd_ctx *ctx, vacd_mlp *M, const float *host);
void mlp_get(vacd_ctx *ctx, vacd_mlp *M, float *host);
void mlp_get_grads(vacd_ctx *ctx, vacd_mlp *M, float *host);
void mlp_forward(vacd_ctx *ctx, vacd_mlp *M, int B, int E, const float *feat, int P,
                 const float *planes, float *out);
void mlp_forward_host(vacd_ctx *ctx, vacd_mlp *M, int B, int E, const float *feat, int P,
                      const float *planes, float *out);
float mlp_train_step(vacd_ctx *ctx, vacd_mlp *M, int B, int E, const float *feat, int P,
                     const float *planes, const float *target, float lr);

// Exclusive prefix sum of m uint32 values in place (one block; lbvh.hip).
void scan_u32(vacd_ctx *ctx, uint32_t *d_data, int64_t m);

// Signed distances (negative inside) of a triangle mesh at the nodes lo +
// (i, j, k) h of an n[0] x n[1] x n[2] grid (distance.hip), x fastest, into
// host out; *sign_mode 0 by column parity (closed mesh), 1 by winding number.
void sdf_grid(vacd_ctx *ctx, const float *V, int64_t nV, const uint32_t *T, int64_t nT,
              const double lo[3], double h, const int n[3], float *out, int *sign_mode);

// Signed-distance remesh of a triangle mesh (distance.hip): the iso-surface
// at distance level_set outside the mesh, on a grid of spacing 1 / scale,
// extracted as quads split into triangles.  Host arrays in and out.
void sdf_remesh(vacd_ctx *ctx, const float *V, int64_t nV, const uint32_t *T, int64_t nT,
                double scale, double level_set, std::vector<float> &Vout,
                std::vector<uint32_t> &Tout, int *sign_mode);

// Fills n 64-bit words with v on the context stream (a kernel, so that it
// may be captured into a graph).
void fill_u64(vacd_ctx *ctx, void *p, int64_t n, unsigned long long v);

}  // namespace vacd
