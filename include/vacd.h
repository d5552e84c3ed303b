/*
 * vacd.h -- C ABI of the MI355X visibility-edge concavity engine.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (VACD_OK == 0); vacd_last_error() gives the message of the last failure
 * on that context (or of the last context-less failure when ctx is NULL).
 *
 * Host-buffer entry points (no suffix) copy inputs to the device, run, copy
 * results back and return when done.  "_d" entry points take device
 * pointers, enqueue on the context's stream and return without waiting.
 *
 * One context per host thread; calls on one context are serialised by the
 * caller.  Device memory (BVHs aside) is pooled in the context and reused
 * across calls.
 */
#ifndef VACD_H
#define VACD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VACD_OK 0
#define VACD_ERR_INVALID 1  /* bad argument (null pointer, negative size, ...) */
#define VACD_ERR_HIP 2      /* HIP runtime error (launch, copy, allocation)    */
#define VACD_ERR_CAPACITY 3 /* output buffer smaller than the result            */
#define VACD_ERR_LIMIT 4    /* input beyond a documented size limit              */

typedef struct vacd_ctx vacd_ctx;
typedef struct vacd_bvh vacd_bvh;

/* Statistics of the last visibility call (vacd_visibility_stats). */
typedef struct vacd_vis_stats {
  int64_t segments;   /* pairs enumerated, n(n-1)/2                       */
  int64_t masked;     /* skipped by the new-vertex mask (reference state 2) */
  int64_t blocked;    /* touched the self mesh                            */
  int64_t edges;      /* visibility edges                                 */
  int64_t exact_calls;/* pairs redone by the exact-predicate fallback     */
  int64_t node_visits;/* BVH nodes loaded (VACD_OPT_COUNT_WORK only)     */
  int64_t tri_tests;  /* triangle tests (VACD_OPT_COUNT_WORK only)       */
  int64_t walked;     /* pairs that needed a BVH walk (the rest were      */
                      /* masked, zero-length or settled by a fan)         */
} vacd_vis_stats;

/* Work counters of the last visibility call, for diagnosis: out[k] for
 * k < n (n <= 16) = 0 masked, 1 blocked, 2 pairs whose undecided midpoint
 * ray was settled along +y instead, 3 edges, 4 malformed-tree walks,
 * 5 node visits, 6 triangle tests, 7 walked, 8 redone exactly, 9 exact
 * triangle tests; with VACD_OPT_COUNT_WORK the fan pass also counts the
 * pairs it leaves to the walks with 10 no usable fan at p, 11 p's fan test
 * undecided, 12 no usable fan at q, 13 q's undecided, 14 both ends leaving
 * outwards; 15 queued pairs walked over their clipped segment only
 * (VACD_OPT_CLIP_ENDS).  Waits for the stream. */
int vacd_visibility_counters(vacd_ctx *ctx, int64_t *out, int32_t n);

/* Context options (vacd_ctx_set_option). */
#define VACD_OPT_TIMING 1     /* record HIP events around the main kernels  */
#define VACD_OPT_COUNT_WORK 2 /* visibility counts node visits / tri tests  */
#define VACD_OPT_FAN_PAIRS 4  /* parity mode: settle pairs that enter the  */
                              /* solid at p or at q from that end's fan (1) */
#define VACD_OPT_SCORE_CHUNK 5 /* edges per scoring block, 0 = automatic    */
#define VACD_OPT_LANES_PER_PAIR 6 /* lanes sharing one BVH walk of a pair   */
                                  /* the fan leaves open: 1,2,4,8,16 (1)    */
#define VACD_OPT_FRONT_LANES 8 /* lanes per pair of the breadth-first walks */
                               /* of short walk queues: 0, 4, 8, 16, 32, 64 */
                               /* (8; 0 = depth-first walks only)           */
#define VACD_OPT_FRONT_MAX 9   /* longest queue walked breadth-first, longer */
                               /* ones depth first (2^40: every queue)      */
#define VACD_OPT_CLIP_ENDS 10  /* breadth-first walks of pairs whose fans   */
                               /* show both ends leaving outwards test only */
                               /* the segment outside both endpoints'       */
                               /* clearance balls, in parts of at least     */
                               /* this many segments (2^25; 0 = never)      */
#define VACD_OPT_ROTATED_RAYS 11 /* a walked pair on a closed part whose +x  */
                                 /* midpoint ray meets an undecided test     */
                                 /* tries the +y, then the +z ray before the */
                                 /* exact redo (1; 0 = exact redo at once)   */

/* Accumulated device time of the main kernels since the last reset
 * (VACD_OPT_TIMING): the visibility traversal kernel, the plane-scoring
 * kernel and a whole LBVH build. */
typedef struct vacd_timing {
  double vis_ms;
  int64_t vis_launches;
  double score_ms;
  int64_t score_launches;
  double bvh_ms;
  int64_t bvh_launches;
} vacd_timing;

/* ---- context -------------------------------------------------------------
 * Replaces createContext / optixDeviceContextDestroy
 * (lib/include/optixUtils.hpp:80-108, called at lib/src/process.cpp:246,331)
 * and the CUCHK/OCHK error macros (lib/include/optixUtils.hpp:56-71). */
int vacd_ctx_create(int device, vacd_ctx **out);
int vacd_ctx_destroy(vacd_ctx *ctx);
const char *vacd_last_error(const vacd_ctx *ctx);
/* The hipStream_t the context enqueues on (returned as void*). */
int vacd_ctx_stream(vacd_ctx *ctx, void **stream);
/* Enqueue on a caller-owned hipStream_t instead (NULL restores the own one). */
int vacd_ctx_set_stream(vacd_ctx *ctx, void *stream);
int vacd_sync(vacd_ctx *ctx);
int vacd_ctx_set_option(vacd_ctx *ctx, int option, int64_t value);
/* Waits for the context stream, fills *out, and clears the record if reset. */
int vacd_ctx_timing(vacd_ctx *ctx, vacd_timing *out, int reset);

/* ---- BVH -------------------------------------------------------------------
 * Replaces the OptiX GAS build of run_optix (lib/src/intersections.cpp:131-180)
 * with an on-device LBVH (30-bit Morton codes, LDS radix sort, Karras
 * hierarchy, bottom-up AABB refit).  vertices: float32 [n_vertices*3];
 * tris: uint32 [n_tris*3] vertex indices.  vacd_bvh_build copies from host
 * memory; vacd_bvh_build_d reads device memory.  Both wait for the build.
 * Limits: 1 <= n_tris < 2^30. */
int vacd_bvh_build(vacd_ctx *ctx, const float *vertices, int64_t n_vertices,
                   const uint32_t *tris, int64_t n_tris, vacd_bvh **out);
/* Same as vacd_bvh_build, with planar groups: tri_group[t] in [0, n_groups)
 * tags triangle t as lying in the plane planes[4g..4g+3] (a, b, c, d with
 * a x + b y + c z + d = 0; -1: no group), n_groups <= 7.  The contact walks
 * then test a group's boxes only against the part of a segment within the
 * group's measured distance of its plane; results are the same as without
 * groups (a group whose triangles lie farther than 1e-4 of the mesh extent
 * from its plane is ignored).  Meant for the planar caps that clipping
 * leaves, whose long triangles overlap the boxes around every cap vertex.
 * No counterpart in the reference (OptiX builds from the soup only,
 * lib/src/intersections.cpp:131-180). */
int vacd_bvh_build_planar(vacd_ctx *ctx, const float *vertices, int64_t n_vertices,
                          const uint32_t *tris, int64_t n_tris, const int8_t *tri_group,
                          const double *planes, int n_groups, vacd_bvh **out);
int vacd_bvh_build_d(vacd_ctx *ctx, const float *d_vertices, int64_t n_vertices,
                     const uint32_t *d_tris, int64_t n_tris, vacd_bvh **out);
int vacd_bvh_destroy(vacd_bvh *bvh);
/* n_tris, tree depth (root = 1), scene bounds lo[3], hi[3]. */
int vacd_bvh_info(const vacd_bvh *bvh, int64_t *n_tris, int32_t *depth,
                  float *bounds6);
/* Copies the node array (64 B per internal node, n_tris-1 nodes) and the
 * leaf-ordered triangle array (48 B per triangle) to host memory. */
/* Diagnostic: primitives (triangles and slices of long triangles) and depth
 * of the BVH's contact tree; 0 and 0 when no triangle needed slicing.  No
 * counterpart in the reference (its any-hit programs run on OptiX's own
 * acceleration structure). */
int vacd_bvh_contact_info(const vacd_bvh *bvh, int64_t *n_refs, int32_t *depth);

int vacd_bvh_export(const vacd_bvh *bvh, void *nodes, void *tris);

/* ---- visibility --------------------------------------------------------------
 * Replaces run_optix x2 + the combine loop of compute_intersection_matrix
 * (lib/src/intersections.cpp:75-325, 327-385; device programs
 * lib/ptx/ray_segments.cu:49-125).  Tests every pair (i<j) of the n points
 * pts[n*3].  With new_mask non-NULL, new_mask[i] holds bit 0 = new point
 * (the reference's mask: a pair is skipped when neither point has it) and
 * bit 1 + g = the point lies on the plane of planar group g of the self BVH
 * (vacd_bvh_build_planar; the reference has no such bits).  A pair whose
 * points share such a bit is not an edge when, at either end, its
 * direction enters a group-g triangle of that end's fan (exact 2-D
 * orientations in the group plane's projection; closed self mesh, parity
 * mode; ends with more than 16 incident triangles are not tested there):
 * the cap rule.  Otherwise, and with bits 1-7 clear, a pair is a visibility
 * edge iff its segment does not touch the self mesh and (cage != NULL)
 * touches the cage mesh, or
 * (cage == NULL) its midpoint lies outside the self mesh.  Edges are written
 * to pairs[2*e], pairs[2*e+1] = (i, j) in increasing k = i(2n-i-1)/2 + j-i-1
 * order, at most cap pairs; *count receives the total number of edges.
 * Returns VACD_ERR_CAPACITY (after setting *count) if count > cap.
 * Limits: 2 <= n <= 2^26 (k must fit 52 bits); n = 0 or 1 gives no edges. */
int vacd_visibility(vacd_ctx *ctx, const float *pts, int64_t n,
                    const uint8_t *new_mask, const vacd_bvh *self,
                    const vacd_bvh *cage, uint32_t *pairs, int64_t cap,
                    int64_t *count);
/* Device-pointer form: d_count is a device int64; no capacity error is
 * raised (the caller reads *d_count and compares with cap). */
int vacd_visibility_d(vacd_ctx *ctx, const float *d_pts, int64_t n,
                      const uint8_t *d_new_mask, const vacd_bvh *self,
                      const vacd_bvh *cage, uint32_t *d_pairs, int64_t cap,
                      int64_t *d_count);
/* Statistics of the last visibility call on ctx (waits for it). */
int vacd_visibility_stats(vacd_ctx *ctx, vacd_vis_stats *out);

/* Diagnostic: the pair ids k (global across the parts of a batch) that the
 * last visibility call redid with exact predicates, in no particular order;
 * copies min(count, cap) ids to host memory and sets *count.  No counterpart
 * in the reference (its any-hit programs have no exact fallback). */
int vacd_visibility_deferred(vacd_ctx *ctx, int64_t *pair_ids, int64_t cap, int64_t *count);

/* One part of a batched visibility call: device points, optional device
 * new-vertex mask, its own BVH and optional cage BVH. */
typedef struct vacd_part {
  const float *pts;
  int64_t n;
  const uint8_t *new_mask;
  const vacd_bvh *self;
  const vacd_bvh *cage;
} vacd_part;

/* Batched frontier: the visibility of every part in one launch sequence
 * (the reference visits parts one at a time, lib/src/process.cpp:166-178).
 * parts is a host array of nparts descriptors whose pointers are device
 * pointers.  Edges are written part after part, each part's in its own k
 * order with indices local to the part; d_part_off (device int64[nparts+1])
 * receives each part's first edge index and the total; d_count the total.
 * Pairs beyond cap are dropped (compare *d_count with cap). */
int vacd_visibility_batch_d(vacd_ctx *ctx, const vacd_part *parts, int32_t nparts,
                            uint32_t *d_pairs, int64_t cap, int64_t *d_count,
                            int64_t *d_part_off);

/* Host form of vacd_visibility_batch_d: parts[p].pts and parts[p].new_mask
 * are host arrays; pairs (host, at most cap) and part_off (host, nparts + 1
 * entries, part p's edges are pairs[part_off[p] .. part_off[p+1])) are
 * written, *count receives the total; waits.  Returns VACD_ERR_CAPACITY
 * (after setting *count) if count > cap. */
int vacd_visibility_batch(vacd_ctx *ctx, const vacd_part *parts, int32_t nparts,
                          uint32_t *pairs, int64_t cap, int64_t *count, int64_t *part_off);

/* ---- plane scoring ------------------------------------------------------------
 * Replaces classify_and_rate_planes + plane_point_side_and_rate_kernel
 * (lib/cuda/plane_intersections.cu:7-133).  planes: float32 [P*4] (a,b,c,d);
 * pts: float32 [n*3]; edges: uint32 [E*2].  side(v) = sign(fmaf(c,z,
 * fmaf(b,y,a*x)) + d) with dead zone 1e-6; an edge whose endpoints have
 * nonzero opposite sides adds its length sqrtf(fmaf(dz,dz,fmaf(dy,dy,dx*dx)))
 * rounded once to Q32.32.  Outputs (each nullable): score_q32 int64 [P]
 * (exact, order-independent), score_f float32 [P] (= score_q32 * 2^-32),
 * argmax (first maximum of score_q32; -1 when P == 0).
 * Edges with an endpoint index >= n are ignored (as the reference does).
 * Limits: P < 2^31, E < 2^40. */
int vacd_score_planes(vacd_ctx *ctx, const float *planes, int64_t P,
                      const float *pts, int64_t n, const uint32_t *edges,
                      int64_t E, int64_t *score_q32, float *score_f,
                      int32_t *argmax);
int vacd_score_planes_d(vacd_ctx *ctx, const float *d_planes, int64_t P,
                        const float *d_pts, int64_t n, const uint32_t *d_edges,
                        int64_t E, int64_t *d_score_q32, float *d_score_f,
                        int32_t *d_argmax);

/* Allocation-free scoring.  No counterpart in the reference: meant for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * definitions and results as vacd_score_planes_d (device pointers, enqueued
 * on the context stream, returns without waiting), but the plane sums and
 * edge records live in a fixed device array of the library (one per GPU,
 * held until the process ends, about 32.5 MiB) instead of the context's
 * pool, so the call makes no allocation of any kind.  Limits: P <=
 * VACD_NOALLOC_MAX_PLANES, E <= VACD_NOALLOC_MAX_EDGES (VACD_ERR_LIMIT
 * beyond); d_planes must be 16-byte and d_edges 8-byte aligned
 * (VACD_ERR_INVALID otherwise).  Every call on one GPU shares that array:
 * calls that overlap on one GPU (two streams, two contexts, two threads)
 * race and must be serialised by the caller. */
#define VACD_NOALLOC_MAX_PLANES 65536
#define VACD_NOALLOC_MAX_EDGES 1048576
int vacd_score_planes_noalloc_d(vacd_ctx *ctx, const float *d_planes, int64_t P,
                                const float *d_pts, int64_t n, const uint32_t *d_edges,
                                int64_t E, int64_t *d_score_q32, float *d_score_f,
                                int32_t *d_argmax);

/* One part of a batched scoring call: its planes, points and edges (device
 * pointers for the _d form, host pointers for the host form). */
typedef struct vacd_score_part {
  const float *planes;   /* [P][4] */
  int64_t P;
  const float *pts;      /* [n][3] */
  int64_t n;
  const uint32_t *edges; /* [E][2], indices into pts */
  int64_t E;
} vacd_score_part;

/* Batched scoring: every part's planes on its own edges, with the same
 * definitions as vacd_score_planes, in one set of launches (the recursion
 * frontier of process(), lib/src/process.cpp:74-134, one call for all parts
 * split at once).  Scores of part p start at plane_off[p] = sum of the P of
 * parts 0..p-1 in score_q32 (nullable); argmax[p] is part p's first maximum
 * (-1 when it has no planes).  The _d form takes device pointers, writes
 * device outputs and does not wait; the host form copies and waits. */
int vacd_score_planes_batch_d(vacd_ctx *ctx, const vacd_score_part *parts, int nparts,
                              int64_t *d_score_q32, int32_t *d_argmax);
int vacd_score_planes_batch(vacd_ctx *ctx, const vacd_score_part *parts, int nparts,
                            int64_t *score_q32, int32_t *argmax);

/* ---- learned plane proposal ----------------------------------------------------
 * No counterpart in the reference.  vacd_gemm_bf16_d is the product the model
 * is built from, exported for its tests: C[M][N] = op(A) op(B) (+ bias[N],
 * relu), A, B bf16 device arrays, A(m,k) = ta ? A[k*lda+m] : A[m*lda+k],
 * B(k,n) = tb ? B[n*ldb+k] : B[k*ldb+n], fp32 output C (ldc), accumulate
 * into C when acc != 0. */
int vacd_gemm_bf16_d(vacd_ctx *ctx, int ta, int tb, int M, int N, int K, const void *d_A,
                     int lda, const void *d_B, int ldb, const float *d_bias, int relu, int acc,
                     float *d_C, int ldc);

/* The model: a set encoder over E sampled visibility edges of a part (edge
 * features (u, v, |v - u|), 7 floats) with max-pooling, and a head that
 * predicts, for each plane (n, d) of the part, its cut length relative to the
 * part's other planes (trained on cut length / largest cut length among the
 * part's planes; only the order is used).  Weights are FP32 masters with bf16 shadows on
 * MFMA.  Batches hold B parts with E edges and P planes each: d_feat
 * [B][E][7], d_planes [B][P][4], d_out / d_target [B][P] (device).
 * Parameters travel as vacd_mlp_param_count() floats. */
typedef struct vacd_mlp vacd_mlp;
int64_t vacd_mlp_param_count(void);
int vacd_mlp_create(vacd_ctx *ctx, uint64_t seed, vacd_mlp **out);
int vacd_mlp_destroy(vacd_mlp *mlp);
int vacd_mlp_set_params(vacd_ctx *ctx, vacd_mlp *mlp, const float *params);
int vacd_mlp_get_params(vacd_ctx *ctx, vacd_mlp *mlp, float *params);
int vacd_mlp_forward_d(vacd_ctx *ctx, vacd_mlp *mlp, int B, int E, const float *d_feat, int P,
                       const float *d_planes, float *d_out);
/* Same on host buffers (feat [B][E][7], planes [B][P][4], out [B][P]);
 * returns when out is written. */
int vacd_mlp_forward(vacd_ctx *ctx, vacd_mlp *mlp, int B, int E, const float *feat, int P,
                     const float *planes, float *out);
/* The gradient of the loss of the last vacd_mlp_train_step_d with respect to
 * the parameters (vacd_mlp_param_count() floats, laid out as the
 * parameters), copied to host memory; waits.  For the tests of the training
 * kernels (no counterpart in the reference). */
int vacd_mlp_get_grads(vacd_ctx *ctx, vacd_mlp *mlp, float *grads);
/* One Adam step (lr; betas 0.9, 0.999) on the mean squared error; waits and
 * writes the batch loss (before the step) to *loss. */
int vacd_mlp_train_step_d(vacd_ctx *ctx, vacd_mlp *mlp, int B, int E, const float *d_feat,
                          int P, const float *d_planes, const float *d_target, float lr,
                          float *loss);

/* ---- surface distance ---------------------------------------------------------
 * Largest distance from m points to a triangle mesh (its BVH), the one-sided
 * term of the sampled Hausdorff distance in the concavity metric
 * (lib/src/cost.cpp:69-105, lib/include/hausdorff.hpp:214-298; the reference
 * approximates each point's distance from its 10 nearest samples, this
 * measures it against every triangle).  vacd_max_distance copies points
 * from the host and waits; the _d form writes a device float.
 * Limit: the BVH depth must allow a 64-entry walk stack (3 * depth/2 <= 64). */
int vacd_max_distance(vacd_ctx *ctx, const float *pts, int64_t m,
                      const vacd_bvh *mesh, float *out);
int vacd_max_distance_d(vacd_ctx *ctx, const float *d_pts, int64_t m,
                        const vacd_bvh *mesh, float *d_out);

/* Batched form for the concavity metric: for each query, the largest
 * distance from its m points to its triangle mesh (float32 vertices
 * [nV][3], triangles [nT][3] of vertex ids), every point measured against
 * every triangle (no BVH to build; same value as the BVH form).  One launch
 * for all queries; out[nq] is written when the call returns (+inf for a
 * query with points but no triangles, 0 for one without points).  Replaces,
 * for a whole decomposition round at once, the per-part Hausdorff sampling
 * of lib/src/cost.cpp:69-105. */
typedef struct vacd_distance_query {
  const float *pts;
  int64_t m;
  const float *V;
  int64_t nV;
  const uint32_t *T;
  int64_t nT;
} vacd_distance_query;
int vacd_max_distance_batch(vacd_ctx *ctx, const vacd_distance_query *q, int32_t nq,
                            float *out);

/* ---- signed-distance remesh ------------------------------------------------
 * Replaces manifold_preprocess / sdf_manifold (lib/src/preprocess.cpp:18-67,
 * OpenVDB meshToSignedDistanceField + volumeToMesh), which the reference uses
 * to remesh its input into a closed manifold (scale 30 or 20, level_set
 * 0.55 / scale, lib/src/process.cpp:252-257) and to build the dilated cage
 * its visibility rule tests against (scale 40, level_set 0.03 for the input
 * and 0.02 per part, lib/src/process.cpp:170-171, 260-261).  The surface at
 * signed distance level_set outside the mesh (V float32 [nV][3], T [nT][3]),
 * sampled on a grid of nodes at integer multiples of 1 / scale: distances
 * by the mesh's BVH, inside/outside by crossing parity along grid columns
 * (closed meshes) or by the generalised winding number (open ones), the
 * surface by surface nets (a vertex per straddling cell, two triangles per
 * straddling grid edge, facing outwards; the result is closed).  Writes at
 * most v_cap vertices and t_cap triangles, sets *nV_out and *nT_out, and
 * returns VACD_ERR_CAPACITY when they do not fit.  Waits.  Limits: at most
 * 2,048 grid nodes per axis. */
int vacd_sdf_remesh(vacd_ctx *ctx, const float *V, int64_t nV, const uint32_t *T, int64_t nT,
                    double scale, double level_set, float *V_out, int64_t v_cap,
                    uint32_t *T_out, int64_t t_cap, int64_t *nV_out, int64_t *nT_out);
/* The signed distance field underneath (negative inside) at the nodes
 * lo + (i, j, k) * h of an n[0] x n[1] x n[2] grid, x fastest, into the host
 * array out (n[0] n[1] n[2] floats); *sign_mode (nullable) receives 0 when
 * the signs came from column parity (closed mesh) and 1 from winding numbers.
 * Waits.  Limits: at most 2,048 nodes per axis. */
int vacd_sdf_grid(vacd_ctx *ctx, const float *V, int64_t nV, const uint32_t *T, int64_t nT,
                  const double *lo, double h, const int32_t *n, float *out, int32_t *sign_mode);

/* ---- candidate-plane extents ---------------------------------------------
 * For each query, the extent [min, max] of n.p over its points p for each of
 * its P normals n, in float64 with the products and sums rounded as the host
 * evaluates ((n0 p0 + n1 p1) + n2 p2) without fused multiply-add, so the
 * values equal a host loop's bit for bit.  Used to place the benchmark
 * workloads' candidate planes (normals uniform on the sphere, offsets uniform
 * over the part's extent along the normal) for many parts in one launch.
 * lo_hi receives [P][2] doubles per query, queries back to back. */
typedef struct vacd_extent_query {
  const double *normals; /* [P][3] */
  int64_t P;
  const double *pts; /* [n][3] */
  int64_t n;
} vacd_extent_query;
int vacd_plane_extents_batch(vacd_ctx *ctx, const vacd_extent_query *q, int32_t nq,
                             double *lo_hi);

/* Reference-signature shim: same arguments and behaviour as
 * classify_and_rate_planes (lib/cuda/plane_intersections.cu:71-78,
 * declared at lib/src/process.cpp:74-79): host pointers, synchronous,
 * scores[] overwritten.  Uses a per-thread default context on device 0;
 * failures are reported on stderr, as the reference does. */
void classify_and_rate_planes(const float *planes, const float *points,
                              const unsigned int *edges, float *scores,
                              int numPlanes, int numPoints, int numEdges);

#ifdef __cplusplus
}
#endif

#endif /* VACD_H */
