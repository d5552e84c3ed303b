// Largest point-to-mesh distance, for the Hausdorff term of the concavity
// metric (lib/src/cost.cpp:69-105, lib/include/hausdorff.hpp:214-298).
//
// The reference estimates each sample's distance to the other mesh from the
// triangles of its 10 nearest samples (kd-tree); here every sample is
// measured exactly against the whole mesh with a closest-point walk of the
// mesh's 4-wide BVH: children are visited nearest box first and pruned by
// the best squared distance so far.  One lane per sample; the per-sample
// distances are max-reduced per wave and folded with an integer atomicMax
// on the float bits (non-negative floats order like their bit patterns).
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <thread>

namespace vacd {
namespace {

constexpr int kStack = 64;

__device__ __forceinline__ float box_d2(const float p[3], float lx, float ly, float lz, float hx,
                                        float hy, float hz) {
  const float dx = fmaxf(fmaxf(lx - p[0], p[0] - hx), 0.0f);
  const float dy = fmaxf(fmaxf(ly - p[1], p[1] - hy), 0.0f);
  const float dz = fmaxf(fmaxf(lz - p[2], p[2] - hz), 0.0f);
  return dx * dx + dy * dy + dz * dz;
}

// Squared distance from p to triangle abc (closest-point regions; C. Ericson,
// Real-Time Collision Detection, 2005, section 5.1.5).
__device__ float tri_d2(const float p[3], const float4 &A, const float4 &B, const float4 &C) {
  const float ab[3] = {B.x - A.x, B.y - A.y, B.z - A.z}, ac[3] = {C.x - A.x, C.y - A.y, C.z - A.z};
  const float ap[3] = {p[0] - A.x, p[1] - A.y, p[2] - A.z};
  auto dot3 = [](const float *u, const float *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  auto d2to = [&](float x, float y, float z) {
    const float dx = p[0] - x, dy = p[1] - y, dz = p[2] - z;
    return dx * dx + dy * dy + dz * dz;
  };
  const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.f && d2 <= 0.f) return d2to(A.x, A.y, A.z);
  const float bp[3] = {p[0] - B.x, p[1] - B.y, p[2] - B.z};
  const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.f && d4 <= d3) return d2to(B.x, B.y, B.z);
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
    const float v = d1 / (d1 - d3);
    return d2to(A.x + v * ab[0], A.y + v * ab[1], A.z + v * ab[2]);
  }
  const float cp[3] = {p[0] - C.x, p[1] - C.y, p[2] - C.z};
  const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.f && d5 <= d6) return d2to(C.x, C.y, C.z);
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
    const float w = d2 / (d2 - d6);
    return d2to(A.x + w * ac[0], A.y + w * ac[1], A.z + w * ac[2]);
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return d2to(B.x + w * (C.x - B.x), B.y + w * (C.y - B.y), B.z + w * (C.z - B.z));
  }
  const float den = 1.0f / (va + vb + vc);
  const float v = vb * den, w = vc * den;
  return d2to(A.x + ab[0] * v + ac[0] * w, A.y + ab[1] * v + ac[1] * w, A.z + ab[2] * v + ac[2] * w);
}

__device__ __forceinline__ float comp4(const float4 &v, int c) {
  return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
}

__global__ __launch_bounds__(256) void k_max_distance(const float *__restrict__ pts, int64_t m,
                                                      const Node4 *__restrict__ nodes,
                                                      const Tri *__restrict__ tris, int cap,
                                                      unsigned *__restrict__ out_bits) {
  __shared__ int stk_all[4][kStack][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int *stk = &stk_all[wave][0][lane];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  float best = 0.0f;
  if (i < m) {
    const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    float bd = INFINITY;
    int node = 0, sp = 0;
    for (int it = 0; it <= cap; ++it) {
      const Node4 nd = nodes[node];
      float d[4];
      int id[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        d[c] = box_d2(p, comp4(nd.f[0], c), comp4(nd.f[1], c), comp4(nd.f[2], c),
                      comp4(nd.f[3], c), comp4(nd.f[4], c), comp4(nd.f[5], c));
        id[c] = __float_as_int(comp4(nd.f[6], c));
        if (comp4(nd.f[0], c) == kEmpty) d[c] = INFINITY;
      }
      // leaves first, then internal children nearest first
      int next = -1;
      float next_d = INFINITY;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (!(d[c] < bd)) continue;
        if (id[c] < 0) {
          const Tri t = tris[~id[c]];
          bd = fminf(bd, tri_d2(p, t.a, t.b, t.c));
        } else if (d[c] < next_d) {
          if (next >= 0 && sp < kStack) stk[(sp++) * 64] = next;
          next = id[c];
          next_d = d[c];
        } else if (sp < kStack) {
          stk[(sp++) * 64] = id[c];
        }
      }
      if (next >= 0) {
        node = next;
      } else {
        if (sp == 0) break;
        node = stk[(--sp) * 64];
      }
    }
    best = sqrtf(bd);
  }
  for (int off = 32; off > 0; off >>= 1) best = fmaxf(best, __shfl_down(best, off));
  if (lane == 0) atomicMax(out_bits, __float_as_uint(best));
}

}  // namespace

void launch_max_distance(vacd_ctx *ctx, const float *d_pts, int64_t m, const vacd_bvh *mesh,
                         unsigned *d_out_bits) {
  hipStream_t st = ctx->stream;
  VACD_HIP(hipMemsetAsync(d_out_bits, 0, sizeof(unsigned), st));
  if (m == 0) return;
  const int cap = (int)std::min<int64_t>(2 * mesh->n_tris + 64, INT32_MAX);
  k_max_distance<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(
      d_pts, m, mesh->nodes4.as<Node4>(), mesh->tris.as<Tri>(), cap, d_out_bits);
  VACD_HIP(hipGetLastError());
}

}  // namespace vacd

// ------------------------------------------------------------- batched form
// Many (points, mesh) queries in one launch, without a BVH: the point-to-mesh
// distance is needed for a few thousand surface samples against every part
// of a decomposition round and against its hull, and building a BVH per
// mesh (a dozen launches, allocations and a host wait each) costs more than
// this.  Three steps:
//   k_dist_bound: an upper bound per sample, its distance to the nearest of
//     up to kRep vertices spread over the query's triangles (points of the
//     mesh), staged in LDS;
//   k_dist_chunks: a bounding sphere per chunk of kDistChunk triangles (a
//     mesh lists neighbouring triangles near each other);
//   k_dist_near: each sample against every triangle of its nearest chunk
//     (by sphere), which holds its closest triangle or one close to it: the
//     bounds drop near the true distances before the chunk-pruned pass;
//   k_dist_seed: per query, the exact distance of the sample with the
//     largest bound, a lower bound h of the query's maximum (a query's
//     result is the maximum over its samples of their distance to the mesh:
//     a sample whose distance cannot exceed h does not change it);
//   k_dist_batch: a block takes 256 samples and one chunk; lanes whose bound
//     is at most h, or that the chunk's sphere cannot improve, skip it (the
//     block skips it when all do); the rest stage its triangles 256 at a
//     time in LDS with their own spheres and measure exactly those that can
//     come closer than the best so far (relative slack well above float32
//     rounding), until the best falls to h.
// Per-sample minima (squared, as float bits) meet through atomicMin; a last
// kernel takes each query's maximum.  A sample that ends above h holds its
// exact minimum over all triangles (equal to that of the BVH walk, which
// prunes the same way); one that stops at or below h holds a value between
// its minimum and h, and the seed sample holds h itself: so the maximum is
// exactly the maximum of the exact minima.
namespace vacd {
namespace {

constexpr int kDistChunk = 1024;  // triangles per block
constexpr int kRep = 512;         // vertices of the upper-bound pass

struct DistQ {
  int64_t pt0, m;      // first sample, sample count
  int64_t v0, t0, nT;  // first vertex, first triangle, triangle count
  int64_t ch0;         // first chunk (kDistChunk triangles)
  int64_t blk0;        // first block of the query
  int64_t tchunks;     // triangle chunks of the query
  int64_t sblk0;       // first block of the bound pass
};

__device__ __forceinline__ int find_query(const DistQ *__restrict__ qs, int nq, int64_t key,
                                          int which) {
  int lo = 0, hi = nq - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t v = which == 0 ? qs[mid].blk0 : (which == 1 ? qs[mid].ch0 : qs[mid].sblk0);
    if (v <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// dmin[s] = squared distance to the nearest of the query's kRep spread
// vertices (+inf without triangles); one block per 256 samples.
__global__ __launch_bounds__(256) void k_dist_bound(const DistQ *__restrict__ qs, int nq,
                                                    const float *__restrict__ pts,
                                                    const float *__restrict__ V,
                                                    const uint32_t *__restrict__ T,
                                                    unsigned *__restrict__ dmin) {
  __shared__ float4 s_rep[kRep];
  const DistQ q = qs[find_query(qs, nq, blockIdx.x, 2)];
  const int64_t s = (blockIdx.x - q.sblk0) * 256 + threadIdx.x;
  const int nr = (int)(q.nT < kRep ? q.nT : kRep);
  for (int r = threadIdx.x; r < nr; r += 256) {
    const int64_t t = (int64_t)r * q.nT / nr;  // spread over the triangle list
    const float *X = V + 3 * (q.v0 + T[3 * (q.t0 + t)]);
    s_rep[r] = make_float4(X[0], X[1], X[2], 0.f);
  }
  __syncthreads();
  if (s >= q.m) return;
  const float p[3] = {pts[3 * (q.pt0 + s)], pts[3 * (q.pt0 + s) + 1], pts[3 * (q.pt0 + s) + 2]};
  float best = INFINITY;
  for (int r = 0; r < nr; ++r) {
    const float4 x = s_rep[r];
    const float dx = p[0] - x.x, dy = p[1] - x.y, dz = p[2] - x.z;
    best = fminf(best, dx * dx + dy * dy + dz * dz);
  }
  dmin[q.pt0 + s] = __float_as_uint(best);
}

// Bounding sphere (box centre, farthest vertex) of each chunk.
__global__ __launch_bounds__(256) void k_dist_chunks(const DistQ *__restrict__ qs, int nq,
                                                     int64_t nch, const float *__restrict__ V,
                                                     const uint32_t *__restrict__ T,
                                                     float4 *__restrict__ sph) {
  __shared__ float red[6][256];
  const int64_t c = blockIdx.x;  // one block per chunk
  const DistQ q = qs[find_query(qs, nq, c, 1)];
  const int64_t t0 = (c - q.ch0) * kDistChunk;
  const int64_t t1 = q.nT < t0 + kDistChunk ? q.nT : t0 + kDistChunk;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t t = t0 + threadIdx.x; t < t1; t += 256)
    for (int v = 0; v < 3; ++v) {
      const float *X = V + 3 * (q.v0 + T[3 * (q.t0 + t) + v]);
      for (int k = 0; k < 3; ++k) lo[k] = fminf(lo[k], X[k]), hi[k] = fmaxf(hi[k], X[k]);
    }
  for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = lo[k], red[3 + k][threadIdx.x] = hi[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + w]);
      }
    __syncthreads();
  }
  const float cx = 0.5f * (red[0][0] + red[3][0]), cy = 0.5f * (red[1][0] + red[4][0]),
              cz = 0.5f * (red[2][0] + red[5][0]);
  __syncthreads();
  float r2 = 0.0f;
  for (int64_t t = t0 + threadIdx.x; t < t1; t += 256)
    for (int v = 0; v < 3; ++v) {
      const float *X = V + 3 * (q.v0 + T[3 * (q.t0 + t) + v]);
      const float dx = X[0] - cx, dy = X[1] - cy, dz = X[2] - cz;
      r2 = fmaxf(r2, dx * dx + dy * dy + dz * dz);
    }
  red[0][threadIdx.x] = r2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[0][threadIdx.x] = fmaxf(red[0][threadIdx.x], red[0][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) sph[c] = make_float4(cx, cy, cz, sqrtf(red[0][0]));
}

// dmin[s] = min(dmin[s], squared distance to every triangle of the chunk
// whose sphere comes nearest to the sample); one lane per sample.
__global__ __launch_bounds__(256) void k_dist_near(const DistQ *__restrict__ qs, int nq,
                                                   const float *__restrict__ pts,
                                                   const float *__restrict__ V,
                                                   const uint32_t *__restrict__ T,
                                                   const float4 *__restrict__ csph,
                                                   unsigned *__restrict__ dmin) {
  const DistQ q = qs[find_query(qs, nq, blockIdx.x, 2)];
  const int64_t s = (blockIdx.x - q.sblk0) * 256 + threadIdx.x;
  if (s >= q.m || q.nT == 0) return;
  const float p[3] = {pts[3 * (q.pt0 + s)], pts[3 * (q.pt0 + s) + 1], pts[3 * (q.pt0 + s) + 2]};
  int64_t near = 0;
  float nd = INFINITY;
  for (int64_t c = 0; c < q.tchunks; ++c) {
    const float4 b = csph[q.ch0 + c];
    const float dx = p[0] - b.x, dy = p[1] - b.y, dz = p[2] - b.z;
    const float d = sqrtf(dx * dx + dy * dy + dz * dz) - b.w;
    if (d < nd) nd = d, near = c;
  }
  float best = __uint_as_float(dmin[q.pt0 + s]);
  const int64_t t0 = near * kDistChunk;
  const int64_t t1 = q.nT < t0 + kDistChunk ? q.nT : t0 + kDistChunk;
  for (int64_t t = t0; t < t1; ++t) {
    const uint32_t *tr = T + 3 * (q.t0 + t);
    const float *A = V + 3 * (q.v0 + tr[0]), *B = V + 3 * (q.v0 + tr[1]), *C = V + 3 * (q.v0 + tr[2]);
    best = fminf(best, tri_d2(p, make_float4(A[0], A[1], A[2], 0.f), make_float4(B[0], B[1], B[2], 0.f),
                              make_float4(C[0], C[1], C[2], 0.f)));
  }
  dmin[q.pt0 + s] = __float_as_uint(best);
}

// Per query (one block): s* = the sample with the largest bound (lowest
// index among equals), h = the minimum of its bound and its squared
// distance to every triangle of the query's mesh (the value k_dist_batch
// would leave), written to dmin[s*] and hlb[q].  Queries without
// samples or triangles get h = 0.
__global__ __launch_bounds__(256) void k_dist_seed(const DistQ *__restrict__ qs,
                                                   const float *__restrict__ pts,
                                                   const float *__restrict__ V,
                                                   const uint32_t *__restrict__ T,
                                                   unsigned *__restrict__ dmin,
                                                   unsigned *__restrict__ hlb) {
  __shared__ unsigned rv[256];
  __shared__ int64_t ri[256];
  const DistQ q = qs[blockIdx.x];
  if (q.m == 0 || q.nT == 0) {
    if (threadIdx.x == 0) hlb[blockIdx.x] = 0u;
    return;
  }
  unsigned bv = 0u;
  int64_t bi = -1;
  for (int64_t s = threadIdx.x; s < q.m; s += 256) {
    const unsigned v = dmin[q.pt0 + s];
    if (bi < 0 || v > bv) bv = v, bi = s;
  }
  rv[threadIdx.x] = bv, ri[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const unsigned v = rv[threadIdx.x + w];
      const int64_t i = ri[threadIdx.x + w];
      if (i >= 0 && (ri[threadIdx.x] < 0 || v > rv[threadIdx.x] ||
                     (v == rv[threadIdx.x] && i < ri[threadIdx.x])))
        rv[threadIdx.x] = v, ri[threadIdx.x] = i;
    }
    __syncthreads();
  }
  const int64_t s = ri[0];
  const unsigned bound = rv[0];
  __syncthreads();
  const float p[3] = {pts[3 * (q.pt0 + s)], pts[3 * (q.pt0 + s) + 1], pts[3 * (q.pt0 + s) + 2]};
  float best = __uint_as_float(bound);  // the minimum includes the bound, as in k_dist_batch
  for (int64_t t = threadIdx.x; t < q.nT; t += 256) {
    const uint32_t *tr = T + 3 * (q.t0 + t);
    const float *A = V + 3 * (q.v0 + tr[0]), *B = V + 3 * (q.v0 + tr[1]), *C = V + 3 * (q.v0 + tr[2]);
    best = fminf(best, tri_d2(p, make_float4(A[0], A[1], A[2], 0.f), make_float4(B[0], B[1], B[2], 0.f),
                              make_float4(C[0], C[1], C[2], 0.f)));
  }
  rv[threadIdx.x] = __float_as_uint(best);  // non-negative floats order as their bits
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) rv[threadIdx.x] = min(rv[threadIdx.x], rv[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    dmin[q.pt0 + s] = rv[0];
    hlb[blockIdx.x] = rv[0];
  }
}

__global__ __launch_bounds__(256) void k_dist_batch(const DistQ *__restrict__ qs, int nq,
                                                    const float *__restrict__ pts,
                                                    const float *__restrict__ V,
                                                    const uint32_t *__restrict__ T,
                                                    const float4 *__restrict__ csph,
                                                    const unsigned *__restrict__ hlb,
                                                    unsigned *__restrict__ dmin) {
  __shared__ float4 sa[256], sb[256], sc[256];
  __shared__ float sr[256];
  const int64_t blk = blockIdx.x;
  const int qi = find_query(qs, nq, blk, 0);
  const DistQ q = qs[qi];  // block-uniform
  const float h = __uint_as_float(hlb[qi]);
  const int64_t local = blk - q.blk0;
  const int64_t st = local / q.tchunks, tc = local % q.tchunks;
  const int64_t s = st * 256 + threadIdx.x;
  bool valid = s < q.m;
  float p[3] = {0.f, 0.f, 0.f};
  float best = 0.f;
  if (valid) {
    p[0] = pts[3 * (q.pt0 + s)], p[1] = pts[3 * (q.pt0 + s) + 1], p[2] = pts[3 * (q.pt0 + s) + 2];
    best = __uint_as_float(dmin[q.pt0 + s]);
    valid = best > h;  // at or below the query's lower bound: cannot change its maximum
  }
  float reach = sqrtf(best);
  {  // the whole chunk first: skipped when no lane can improve on its bound
    const float4 b = csph[q.ch0 + tc];
    const float dx = p[0] - b.x, dy = p[1] - b.y, dz = p[2] - b.z;
    const float lim = (reach + b.w) * 1.0001f + 1e-30f;
    const bool need = valid && dx * dx + dy * dy + dz * dz <= lim * lim;
    if (!__syncthreads_or(need)) return;
  }
  const int64_t t_begin = tc * kDistChunk;
  const int64_t t_end = q.nT < t_begin + kDistChunk ? q.nT : t_begin + kDistChunk;
  for (int64_t t0 = t_begin; t0 < t_end; t0 += 256) {
    // every lane down to the lower bound: the rest of the chunk changes nothing
    if (!__syncthreads_or(valid && best > h)) break;
    const int64_t t = t0 + threadIdx.x;
    if (t < t_end) {
      const uint32_t *tr = T + 3 * (q.t0 + t);
      const float *A = V + 3 * (q.v0 + tr[0]), *B = V + 3 * (q.v0 + tr[1]),
                  *C = V + 3 * (q.v0 + tr[2]);
      const float cx = (A[0] + B[0] + C[0]) * (1.0f / 3.0f), cy = (A[1] + B[1] + C[1]) * (1.0f / 3.0f),
                  cz = (A[2] + B[2] + C[2]) * (1.0f / 3.0f);
      auto d2c = [&](const float *X) {
        const float dx = X[0] - cx, dy = X[1] - cy, dz = X[2] - cz;
        return dx * dx + dy * dy + dz * dz;
      };
      sa[threadIdx.x] = make_float4(A[0], A[1], A[2], cx);
      sb[threadIdx.x] = make_float4(B[0], B[1], B[2], cy);
      sc[threadIdx.x] = make_float4(C[0], C[1], C[2], cz);
      sr[threadIdx.x] = sqrtf(fmaxf(fmaxf(d2c(A), d2c(B)), d2c(C)));
    }
    __syncthreads();
    const int cnt = (int)(t_end - t0 < 256 ? t_end - t0 : 256);
    if (!valid) continue;
    for (int k = 0; k < cnt && best > h; ++k) {
      const float4 a = sa[k], b = sb[k], c = sc[k];
      const float dx = p[0] - a.w, dy = p[1] - b.w, dz = p[2] - c.w;
      const float lim = (reach + sr[k]) * 1.0001f + 1e-30f;
      if (dx * dx + dy * dy + dz * dz > lim * lim) continue;  // sphere beyond the best
      const float d = tri_d2(p, a, b, c);
      if (d < best) best = d, reach = sqrtf(d);
    }
  }
  if (valid) atomicMin(&dmin[q.pt0 + s], __float_as_uint(best));
}

// out[q] = sqrt(max over the query's samples of dmin); 0 without samples
__global__ __launch_bounds__(256) void k_dist_max(const DistQ *__restrict__ qs,
                                                  const unsigned *__restrict__ dmin,
                                                  float *__restrict__ out) {
  __shared__ unsigned red[256];
  const DistQ q = qs[blockIdx.x];
  unsigned m = 0;
  for (int64_t s = threadIdx.x; s < q.m; s += 256) m = max(m, dmin[q.pt0 + s]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sqrtf(__uint_as_float(red[0]));
}

}  // namespace

namespace {
// Copies m samples to dst in the order of their cell in a 16^3 grid over
// their box (Morton order of the cells, counting sort), so that a block's
// 256 samples lie close together and skip the same far chunks.  A query's
// result, a maximum over its samples, does not depend on their order.
void sort_samples(const float *src, int64_t m, float *dst, std::vector<uint32_t> &cell) {
  float lo[3] = {src[0], src[1], src[2]}, hi[3] = {src[0], src[1], src[2]};
  for (int64_t i = 1; i < m; ++i)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], src[3 * i + k]), hi[k] = std::max(hi[k], src[3 * i + k]);
  float inv[3];
  for (int k = 0; k < 3; ++k) inv[k] = hi[k] > lo[k] ? 16.0f / (hi[k] - lo[k]) : 0.0f;
  cell.resize((size_t)m);
  std::vector<uint32_t> start(4097, 0);
  for (int64_t i = 0; i < m; ++i) {
    uint32_t key = 0;
    for (int k = 0; k < 3; ++k) {
      int c = (int)((src[3 * i + k] - lo[k]) * inv[k]);
      c = c < 0 ? 0 : (c > 15 ? 15 : c);
      for (int b = 0; b < 4; ++b) key |= (uint32_t)((c >> b) & 1) << (3 * b + k);
    }
    cell[(size_t)i] = key;
    ++start[key + 1];
  }
  for (int c = 0; c < 4096; ++c) start[c + 1] += start[c];
  for (int64_t i = 0; i < m; ++i) {
    const uint32_t at = start[cell[(size_t)i]]++;
    for (int k = 0; k < 3; ++k) dst[3 * (size_t)at + k] = src[3 * i + k];
  }
}

// Copies nT triangles (vertex ids) to dst in the order of the cell of their
// centroid in a 32^3 grid over the mesh's box (Morton order of the cells,
// counting sort), so that a chunk of kDistChunk consecutive triangles is
// compact and its bounding sphere small.  The distance to the mesh does not
// depend on the triangles' order.
void sort_tris(const float *V, int64_t nV, const uint32_t *T, int64_t nT, uint32_t *dst,
               std::vector<uint32_t> &cell) {
  float lo[3] = {V[0], V[1], V[2]}, hi[3] = {V[0], V[1], V[2]};
  for (int64_t i = 1; i < nV; ++i)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], V[3 * i + k]), hi[k] = std::max(hi[k], V[3 * i + k]);
  float inv[3];
  for (int k = 0; k < 3; ++k) inv[k] = hi[k] > lo[k] ? 32.0f / (hi[k] - lo[k]) : 0.0f;
  cell.resize((size_t)nT);
  std::vector<uint32_t> start(32769, 0);
  for (int64_t t = 0; t < nT; ++t) {
    uint32_t key = 0;
    for (int k = 0; k < 3; ++k) {
      const float c = (V[3 * T[3 * t] + k] + V[3 * T[3 * t + 1] + k] + V[3 * T[3 * t + 2] + k]) *
                      (1.0f / 3.0f);
      int x = (int)((c - lo[k]) * inv[k]);
      x = x < 0 ? 0 : (x > 31 ? 31 : x);
      for (int b = 0; b < 5; ++b) key |= (uint32_t)((x >> b) & 1) << (3 * b + k);
    }
    cell[(size_t)t] = key;
    ++start[key + 1];
  }
  for (int c = 0; c < 32768; ++c) start[c + 1] += start[c];
  for (int64_t t = 0; t < nT; ++t) {
    const uint32_t at = start[cell[(size_t)t]]++;
    for (int k = 0; k < 3; ++k) dst[3 * (size_t)at + k] = T[3 * t + k];
  }
}
}  // namespace

void max_distance_batch(vacd_ctx *ctx, const vacd_distance_query *q, int nq, float *out) {
  hipStream_t st = ctx->stream;
  std::vector<DistQ> qs((size_t)nq);
  int64_t np = 0, nv = 0, nt = 0, nch = 0, nb = 0, nsb = 0;
  for (int i = 0; i < nq; ++i) {
    VACD_REQUIRE(q[i].m >= 0 && q[i].nV >= 0 && q[i].nT >= 0, VACD_ERR_INVALID,
                 "negative size in a distance query");
    VACD_REQUIRE((q[i].m == 0 || q[i].pts) && (q[i].nV == 0 || q[i].V) && (q[i].nT == 0 || q[i].T),
                 VACD_ERR_INVALID, "NULL array in a distance query");
    DistQ &d = qs[(size_t)i];
    d.pt0 = np, d.m = q[i].m, d.v0 = nv, d.t0 = nt, d.nT = q[i].nT, d.blk0 = nb, d.ch0 = nch;
    d.sblk0 = nsb;
    d.tchunks = std::max<int64_t>(1, (q[i].nT + kDistChunk - 1) / kDistChunk);
    if (q[i].m > 0 && q[i].nT > 0) nb += ((q[i].m + 255) / 256) * d.tchunks;
    nsb += (q[i].m + 255) / 256;
    np += q[i].m, nv += q[i].nV, nt += q[i].nT;
    nch += (q[i].nT + kDistChunk - 1) / kDistChunk;
  }
  VACD_REQUIRE(nb < (int64_t(1) << 31) && nch < (int64_t(1) << 31), VACD_ERR_LIMIT,
               "distance batch too large");
  // one pinned region: query table | points | vertices | triangles
  const size_t b_q = ((sizeof(DistQ) * (size_t)nq + 255) & ~size_t(255));
  const size_t b_p = ((12 * (size_t)np + 255) & ~size_t(255));
  const size_t b_v = ((12 * (size_t)nv + 255) & ~size_t(255));
  const size_t b_t = ((12 * (size_t)nt + 255) & ~size_t(255));
  const size_t bytes = b_q + b_p + b_v + b_t;
  char *h = static_cast<char *>(pinned_stage(ctx, PINNED_DIST, bytes));
  std::memcpy(h, qs.data(), sizeof(DistQ) * (size_t)nq);
  char *hp = h + b_q, *hv = hp + b_p, *ht = hv + b_v;
  // the queries' samples and triangles sorted into the staging region on
  // host threads: each query's samples, vertices and triangles are separate
  // tasks (largest first), so a large query is not one thread's serial work;
  // the triangle task also checks the query's vertex ids
  struct Task {
    int q, kind;  // kind 0 samples, 1 vertices, 2 triangles
    int64_t size;
  };
  std::vector<Task> tasks;
  for (int i = 0; i < nq; ++i) {
    if (q[i].m) tasks.push_back({i, 0, q[i].m});
    if (q[i].nV) tasks.push_back({i, 1, q[i].nV / 4});
    if (q[i].nT) tasks.push_back({i, 2, 2 * q[i].nT});
  }
  std::sort(tasks.begin(), tasks.end(), [](const Task &x, const Task &y) {
    return x.size > y.size || (x.size == y.size && (x.q < y.q || (x.q == y.q && x.kind < y.kind)));
  });
  std::atomic<int> next{0};
  std::atomic<bool> bad_ids{false};
  auto work = [&] {
    std::vector<uint32_t> cell;
    for (int k; (k = next.fetch_add(1)) < (int)tasks.size();) {
      const int i = tasks[(size_t)k].q;
      const DistQ &d = qs[(size_t)i];
      switch (tasks[(size_t)k].kind) {
        case 0:
          sort_samples(q[i].pts, d.m, reinterpret_cast<float *>(hp + 12 * d.pt0), cell);
          break;
        case 1:
          std::memcpy(hv + 12 * d.v0, q[i].V, 12 * (size_t)q[i].nV);
          break;
        default: {
          const uint32_t *src = q[i].T;
          bool ok = true;
          for (int64_t t = 0; t < 3 * q[i].nT; ++t) ok &= src[t] < (uint64_t)q[i].nV;
          if (!ok) {
            bad_ids = true;
            break;
          }
          sort_tris(q[i].V, q[i].nV, q[i].T, d.nT, reinterpret_cast<uint32_t *>(ht + 12 * d.t0),
                    cell);
        }
      }
    }
  };
  const int nth = std::min<int>((int)tasks.size(),
                                std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  std::vector<std::thread> pool;
  for (int t = 1; t < nth; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  VACD_REQUIRE(!bad_ids, VACD_ERR_INVALID, "triangle index out of range in a distance query");
  DevBuf &b = ctx->buf[SLOT_DIST];
  const size_t b_min = ((4 * (size_t)np + 255) & ~size_t(255));
  const size_t b_out = ((4 * (size_t)nq + 255) & ~size_t(255));
  const size_t b_ch = 16 * (size_t)std::max<int64_t>(nch, 1);
  const size_t b_h = ((4 * (size_t)nq + 255) & ~size_t(255));
  b.ensure(bytes + b_min + b_out + b_ch + b_h);
  char *d = b.as<char>();
  pinned_upload(ctx, PINNED_DIST, d, bytes);
  const DistQ *dq = reinterpret_cast<const DistQ *>(d);
  const float *dp = reinterpret_cast<const float *>(d + b_q);
  const float *dv = reinterpret_cast<const float *>(d + b_q + b_p);
  const uint32_t *dt = reinterpret_cast<const uint32_t *>(d + b_q + b_p + b_v);
  unsigned *dmin = reinterpret_cast<unsigned *>(d + bytes);
  float *dout = reinterpret_cast<float *>(d + bytes + b_min);
  float4 *csph = reinterpret_cast<float4 *>(d + bytes + b_min + b_out);
  unsigned *hlb = reinterpret_cast<unsigned *>(d + bytes + b_min + b_out + b_ch);
  if (nsb) k_dist_bound<<<(unsigned)nsb, 256, 0, st>>>(dq, nq, dp, dv, dt, dmin);
  if (nch) k_dist_chunks<<<(unsigned)nch, 256, 0, st>>>(dq, nq, nch, dv, dt, csph);
  if (nsb && nch) k_dist_near<<<(unsigned)nsb, 256, 0, st>>>(dq, nq, dp, dv, dt, csph, dmin);
  if (nq) k_dist_seed<<<(unsigned)nq, 256, 0, st>>>(dq, dp, dv, dt, dmin, hlb);
  if (nb) k_dist_batch<<<(unsigned)nb, 256, 0, st>>>(dq, nq, dp, dv, dt, csph, hlb, dmin);
  if (nq) k_dist_max<<<(unsigned)nq, 256, 0, st>>>(dq, dmin, dout);
  VACD_HIP(hipGetLastError());
  VACD_HIP(hipMemcpyAsync(out, dout, 4 * (size_t)nq, hipMemcpyDeviceToHost, st));
  VACD_HIP(hipStreamSynchronize(st));
}

}  // namespace vacd

// ------------------------------------------------------ signed-distance remesh
// Replaces the OpenVDB pipeline of lib/src/preprocess.cpp:18-61 (mesh ->
// narrow-band signed distance field -> iso-surface at level_set, voxel
// 1 / scale), used by the reference to remesh its input into a closed
// manifold (scale 30 or 20, level 0.55 voxel) and to build the dilated
// "cage" it tests segments against (scale 40, level 0.03 / 0.02).  Here:
//   k_grid_dist    unsigned distance of every grid node to the mesh, by the
//                  closest-point walk of the mesh's 4-wide BVH (one lane per
//                  node);
//   k_grid_parity  sign of a closed mesh: per (y, z) column, the crossings of
//                  the +x line with the triangles (triangles staged in LDS,
//                  one lane per column; the column is offset by a small
//                  irrational fraction of a voxel so it misses edges and
//                  vertices), a node inside when an odd number lie below it;
//   k_grid_wind    sign of an open mesh: generalised winding number (sum of
//                  the triangles' solid angles over 4 pi, van Oosterom and
//                  Strackee's formula) above one half;
//   k_sn_cells / k_sn_quads  the iso-surface by surface nets: one vertex per
//                  cell whose corners straddle the level (the mean of its
//                  edge crossings), one quad per grid edge that straddles it,
//                  joining the four cells around that edge and facing from
//                  inside to outside; every dual edge is shared by an even
//                  number of quads, so the output is closed.
// Cell and quad lists are compacted by exclusive scans, so the output order
// is deterministic.
namespace vacd {
namespace {

struct Grid {
  double lo[3];  // node (0, 0, 0)
  double h;      // spacing
  int n[3];      // nodes per axis
};

__device__ __forceinline__ int64_t gidx(const Grid &g, int i, int j, int k) {
  return ((int64_t)k * g.n[1] + j) * g.n[0] + i;
}

__device__ float mesh_d2(const float p[3], const Node4 *__restrict__ nodes,
                         const Tri *__restrict__ tris, int cap, int *stk) {
  float bd = INFINITY;
  int node = 0, sp = 0;
  for (int it = 0; it <= cap; ++it) {
    const Node4 nd = nodes[node];
    float d[4];
    int id[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      d[c] = box_d2(p, comp4(nd.f[0], c), comp4(nd.f[1], c), comp4(nd.f[2], c),
                    comp4(nd.f[3], c), comp4(nd.f[4], c), comp4(nd.f[5], c));
      id[c] = __float_as_int(comp4(nd.f[6], c));
      if (comp4(nd.f[0], c) == kEmpty) d[c] = INFINITY;
    }
    int next = -1;
    float next_d = INFINITY;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (!(d[c] < bd)) continue;
      if (id[c] < 0) {
        const Tri t = tris[~id[c]];
        bd = fminf(bd, tri_d2(p, t.a, t.b, t.c));
      } else if (d[c] < next_d) {
        if (next >= 0 && sp < kStack) stk[(sp++) * 64] = next;
        next = id[c];
        next_d = d[c];
      } else if (sp < kStack) {
        stk[(sp++) * 64] = id[c];
      }
    }
    if (next >= 0) {
      node = next;
    } else {
      if (sp == 0) break;
      node = stk[(--sp) * 64];
    }
  }
  return bd;
}

__global__ __launch_bounds__(256) void k_grid_dist(Grid g, const Node4 *__restrict__ nodes,
                                                   const Tri *__restrict__ tris, int cap,
                                                   float *__restrict__ dist) {
  __shared__ int stk_all[4][kStack][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t N = (int64_t)g.n[0] * g.n[1] * g.n[2];
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= N) return;
  const int i = (int)(v % g.n[0]), j = (int)((v / g.n[0]) % g.n[1]), k = (int)(v / ((int64_t)g.n[0] * g.n[1]));
  const float p[3] = {(float)(g.lo[0] + i * g.h), (float)(g.lo[1] + j * g.h), (float)(g.lo[2] + k * g.h)};
  dist[v] = sqrtf(mesh_d2(p, nodes, tris, cap, &stk_all[wave][0][lane]));
}

constexpr int kMaxCross = 64;    // crossings kept per column
constexpr int kTriStage = 256;   // triangles per LDS stage

// Column (j, k) of the grid: the +x line at y_j + 0.000123 h, z_k + 0.0000789 h.
__global__ __launch_bounds__(64) void k_grid_parity(Grid g, const float *__restrict__ V,
                                                    const uint32_t *__restrict__ T, int64_t nT,
                                                    float *__restrict__ dist,
                                                    unsigned *__restrict__ overflow) {
  __shared__ float sv[kTriStage][9];
  __shared__ float xs[kMaxCross][64];
  const int lane = threadIdx.x;
  const int64_t col = blockIdx.x * (int64_t)64 + lane;
  const int64_t ncol = (int64_t)g.n[1] * g.n[2];
  const bool live = col < ncol;
  const int j = live ? (int)(col % g.n[1]) : 0, k = live ? (int)(col / g.n[1]) : 0;
  const double y = g.lo[1] + (j + 1.23e-4) * g.h, z = g.lo[2] + (k + 7.89e-5) * g.h;
  int cnt = 0;
  for (int64_t t0 = 0; t0 < nT; t0 += kTriStage) {
    const int m = (int)(nT - t0 < kTriStage ? nT - t0 : kTriStage);
    __syncthreads();
    for (int x = lane; x < 9 * m; x += 64) {
      const int t = x / 9, r = x % 9;
      sv[t][r] = V[3 * (size_t)T[3 * (t0 + t) + r / 3] + r % 3];
    }
    __syncthreads();
    if (!live) continue;
    for (int t = 0; t < m; ++t) {
      const double ay = sv[t][1], az = sv[t][2], by = sv[t][4], bz = sv[t][5], cy = sv[t][7], cz = sv[t][8];
      // (y, z) inside the projected triangle: three orientations of one sign
      const double o1 = (by - ay) * (z - az) - (bz - az) * (y - ay);
      const double o2 = (cy - by) * (z - bz) - (cz - bz) * (y - by);
      const double o3 = (ay - cy) * (z - cz) - (az - cz) * (y - cy);
      if (!((o1 > 0 && o2 > 0 && o3 > 0) || (o1 < 0 && o2 < 0 && o3 < 0))) continue;
      const double s = o1 + o2 + o3;  // twice the projected area, signed
      // barycentric weights of a, b, c: o2, o3, o1 over s
      const double x = (o2 * sv[t][0] + o3 * sv[t][3] + o1 * sv[t][6]) / s;
      if (cnt < kMaxCross) xs[cnt][lane] = (float)x;
      ++cnt;
    }
  }
  if (!live) return;
  if (cnt > kMaxCross) {
    atomicOr(overflow, 1u);
    return;
  }
  for (int i = 0; i < g.n[0]; ++i) {
    const float xi = (float)(g.lo[0] + i * g.h);
    int below = 0;
    for (int c = 0; c < cnt; ++c) below += xs[c][lane] < xi;
    const int64_t v = gidx(g, i, j, k);
    if (below & 1) dist[v] = -dist[v];
  }
}

// Generalised winding number of every node (open meshes): inside when the
// solid angle sum exceeds 2 pi.
__global__ __launch_bounds__(256) void k_grid_wind(Grid g, const float *__restrict__ V,
                                                   const uint32_t *__restrict__ T, int64_t nT,
                                                   float *__restrict__ dist) {
  __shared__ float sv[kTriStage][9];
  const int64_t N = (int64_t)g.n[0] * g.n[1] * g.n[2];
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool live = v < N;
  double p[3] = {0, 0, 0};
  if (live) {
    const int i = (int)(v % g.n[0]), j = (int)((v / g.n[0]) % g.n[1]), k = (int)(v / ((int64_t)g.n[0] * g.n[1]));
    p[0] = g.lo[0] + i * g.h, p[1] = g.lo[1] + j * g.h, p[2] = g.lo[2] + k * g.h;
  }
  double w = 0.0;
  for (int64_t t0 = 0; t0 < nT; t0 += kTriStage) {
    const int m = (int)(nT - t0 < kTriStage ? nT - t0 : kTriStage);
    __syncthreads();
    for (int x = threadIdx.x; x < 9 * m; x += 256) {
      const int t = x / 9, r = x % 9;
      sv[t][r] = V[3 * (size_t)T[3 * (t0 + t) + r / 3] + r % 3];
    }
    __syncthreads();
    if (!live) continue;
    for (int t = 0; t < m; ++t) {
      const double a[3] = {sv[t][0] - p[0], sv[t][1] - p[1], sv[t][2] - p[2]};
      const double b[3] = {sv[t][3] - p[0], sv[t][4] - p[1], sv[t][5] - p[2]};
      const double c[3] = {sv[t][6] - p[0], sv[t][7] - p[1], sv[t][8] - p[2]};
      const double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      const double lb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      const double lc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                         a[2] * (b[0] * c[1] - b[1] * c[0]);
      const double den = la * lb * lc + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * lc +
                         (a[0] * c[0] + a[1] * c[1] + a[2] * c[2]) * lb +
                         (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) * la;
      w += 2.0 * atan2(det, den);
    }
  }
  if (live && fabs(w) > 2.0 * 3.14159265358979323846) dist[v] = -dist[v];
}

// Cell (i, j, k) spans nodes i..i+1, j..j+1, k..k+1; corner c = bits (x, y, z).
__global__ __launch_bounds__(256) void k_sn_cells(Grid g, const float *__restrict__ sdf, float iso,
                                                  uint32_t *__restrict__ flag,
                                                  float *__restrict__ cvert) {
  const int cx = g.n[0] - 1, cy = g.n[1] - 1, cz = g.n[2] - 1;
  const int64_t C = (int64_t)cx * cy * cz;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int i = (int)(c % cx), j = (int)((c / cx) % cy), k = (int)(c / ((int64_t)cx * cy));
  float v[8];
  int in = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    v[q] = sdf[gidx(g, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2))] - iso;
    in += v[q] < 0.0f;
  }
  if (in == 0 || in == 8) {
    flag[c] = 0;
    return;
  }
  flag[c] = 1;
  // the 12 cube edges as corner pairs
  const int E[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                        {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  double acc[3] = {0, 0, 0};
  int ne = 0;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const int a = E[e][0], b = E[e][1];
    if ((v[a] < 0.0f) == (v[b] < 0.0f)) continue;
    const double t = (double)v[a] / ((double)v[a] - (double)v[b]);
    const int da[3] = {a & 1, (a >> 1) & 1, a >> 2}, db[3] = {b & 1, (b >> 1) & 1, b >> 2};
    acc[0] += da[0] + t * (db[0] - da[0]);
    acc[1] += da[1] + t * (db[1] - da[1]);
    acc[2] += da[2] + t * (db[2] - da[2]);
    ++ne;
  }
  cvert[3 * c] = (float)(g.lo[0] + (i + acc[0] / ne) * g.h);
  cvert[3 * c + 1] = (float)(g.lo[1] + (j + acc[1] / ne) * g.h);
  cvert[3 * c + 2] = (float)(g.lo[2] + (k + acc[2] / ne) * g.h);
}

// Grid edge e along axis a (0 x, 1 y, 2 z) from node (i, j, k): edge ids
// a * N + node.  A quad when the edge straddles the level and its four cells
// exist; mode 0 flags, mode 1 writes the two triangles (vertex ids through
// the cell scan) at the quad's scanned slot.
__global__ __launch_bounds__(256) void k_sn_quads(Grid g, const float *__restrict__ sdf, float iso,
                                                  int mode, uint32_t *__restrict__ qflag,
                                                  const uint32_t *__restrict__ cell_id,
                                                  uint32_t *__restrict__ tris) {
  const int64_t N = (int64_t)g.n[0] * g.n[1] * g.n[2];
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 3 * N) return;
  const int a = (int)(e / N);
  const int64_t v = e % N;
  int p[3] = {(int)(v % g.n[0]), (int)((v / g.n[0]) % g.n[1]), (int)(v / ((int64_t)g.n[0] * g.n[1]))};
  const int u = (a + 1) % 3, w = (a + 2) % 3;
  bool ok = p[a] + 1 < g.n[a] && p[u] >= 1 && p[w] >= 1 && p[u] <= g.n[u] - 2 && p[w] <= g.n[w] - 2;
  float v0 = 0.f, v1 = 0.f;
  if (ok) {
    v0 = sdf[v] - iso;
    int q[3] = {p[0], p[1], p[2]};
    q[a] += 1;
    v1 = sdf[gidx(g, q[0], q[1], q[2])] - iso;
    ok = (v0 < 0.0f) != (v1 < 0.0f);
  }
  if (mode == 0) {
    qflag[e] = ok ? 1u : 0u;
    return;
  }
  if (!ok) return;
  const int cx = g.n[0] - 1, cy = g.n[1] - 1;
  uint32_t id[4];
  const int du[4] = {-1, 0, 0, -1}, dw[4] = {-1, -1, 0, 0};  // around the edge: +u then +w
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int c[3] = {p[0], p[1], p[2]};
    c[u] += du[r];
    c[w] += dw[r];
    id[r] = cell_id[((int64_t)c[2] * cy + c[1]) * cx + c[0]];
  }
  const uint32_t o = 6 * qflag[e];
  if (v0 < 0.0f) {  // inside at the lower end: the surface faces +a
    tris[o] = id[0], tris[o + 1] = id[1], tris[o + 2] = id[2];
    tris[o + 3] = id[0], tris[o + 4] = id[2], tris[o + 5] = id[3];
  } else {
    tris[o] = id[0], tris[o + 1] = id[2], tris[o + 2] = id[1];
    tris[o + 3] = id[0], tris[o + 4] = id[3], tris[o + 5] = id[2];
  }
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

__global__ __launch_bounds__(256) void k_sn_compact(int64_t C, const uint32_t *__restrict__ flag,
                                                    const uint32_t *__restrict__ cell_id,
                                                    const float *__restrict__ cvert,
                                                    float *__restrict__ out) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= C || !flag[c]) return;
  const uint32_t o = cell_id[c];
  out[3 * (size_t)o] = cvert[3 * c], out[3 * (size_t)o + 1] = cvert[3 * c + 1];
  out[3 * (size_t)o + 2] = cvert[3 * c + 2];
}

}  // namespace

namespace {
// Signed distances (negative inside) of every node of g into sdf (device,
// N floats); dv, dt: the mesh on the device, bvh its BVH.  Returns the sign
// method: 0 column parity (closed mesh), 1 winding number.
int signed_grid(vacd_ctx *ctx, const Grid &g, const vacd_bvh &bvh, const float *dv,
                const uint32_t *dt, int64_t nT, float *sdf, unsigned *over) {
  hipStream_t st = ctx->stream;
  const int64_t N = (int64_t)g.n[0] * g.n[1] * g.n[2];
  const int cap = (int)std::min<int64_t>(2 * nT + 64, INT32_MAX);
  k_grid_dist<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(g, bvh.nodes4.as<Node4>(),
                                                          bvh.tris.as<Tri>(), cap, sdf);
  int mode = bvh.closed ? 0 : 1;
  if (mode == 0) {
    k_zero_u32<<<1, 1, 0, st>>>(over);
    const int64_t ncol = (int64_t)g.n[1] * g.n[2];
    k_grid_parity<<<(unsigned)((ncol + 63) / 64), 64, 0, st>>>(g, dv, dt, nT, sdf, over);
    unsigned h_over = 0;
    VACD_HIP(hipMemcpyAsync(&h_over, over, 4, hipMemcpyDeviceToHost, st));
    VACD_HIP(hipStreamSynchronize(st));
    if (h_over) {  // a column met too many crossings: distances again, signs by winding numbers
      mode = 1;
      k_grid_dist<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(g, bvh.nodes4.as<Node4>(),
                                                              bvh.tris.as<Tri>(), cap, sdf);
    }
  }
  if (mode == 1)
    k_grid_wind<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(g, dv, dt, nT, sdf);
  VACD_HIP(hipGetLastError());
  return mode;
}
}  // namespace

void sdf_grid(vacd_ctx *ctx, const float *V, int64_t nV, const uint32_t *T, int64_t nT,
              const double lo[3], double h, const int n[3], float *out, int *sign_mode) {
  VACD_REQUIRE(nV >= 3 && nT >= 1, VACD_ERR_INVALID, "signed distances need a mesh");
  VACD_REQUIRE(h > 0.0 && std::isfinite(h), VACD_ERR_INVALID, "grid spacing must be positive");
  for (int64_t t = 0; t < 3 * nT; ++t)
    VACD_REQUIRE((int64_t)T[t] < nV, VACD_ERR_INVALID, "triangle index out of range");
  Grid g;
  int64_t N = 1;
  for (int k = 0; k < 3; ++k) {
    VACD_REQUIRE(n[k] >= 1 && n[k] <= 2048, VACD_ERR_LIMIT, "grid needs 1..2048 nodes per axis");
    g.lo[k] = lo[k];
    g.n[k] = n[k];
    N *= n[k];
  }
  g.h = h;
  hipStream_t st = ctx->stream;
  vacd_bvh bvh;
  DevBuf dv, dt, work;
  try {
    dv.ensure(12 * (size_t)nV);
    dt.ensure(12 * (size_t)nT);
    VACD_HIP(hipMemcpyAsync(dv.ptr, V, 12 * (size_t)nV, hipMemcpyHostToDevice, st));
    VACD_HIP(hipMemcpyAsync(dt.ptr, T, 12 * (size_t)nT, hipMemcpyHostToDevice, st));
    build_lbvh(ctx, dv.as<float>(), nV, dt.as<uint32_t>(), nT, &bvh);
    work.ensure(4 * (size_t)N + 256);
    const int mode = signed_grid(ctx, g, bvh, dv.as<float>(), dt.as<uint32_t>(), nT, work.as<float>(),
                                 reinterpret_cast<unsigned *>(work.as<char>() + 4 * (size_t)N));
    if (sign_mode) *sign_mode = mode;
    VACD_HIP(hipMemcpyAsync(out, work.ptr, 4 * (size_t)N, hipMemcpyDeviceToHost, st));
    VACD_HIP(hipStreamSynchronize(st));
  } catch (...) {
    bvh.release(), dv.release(), dt.release(), work.release();
    throw;
  }
  bvh.release(), dv.release(), dt.release(), work.release();
}

void sdf_remesh(vacd_ctx *ctx, const float *V, int64_t nV, const uint32_t *T, int64_t nT,
                double scale, double level_set, std::vector<float> &Vout,
                std::vector<uint32_t> &Tout, int *sign_mode) {
  VACD_REQUIRE(nV >= 3 && nT >= 1, VACD_ERR_INVALID, "remesh needs a mesh");
  VACD_REQUIRE(scale > 0.0 && std::isfinite(scale) && level_set >= 0.0 && std::isfinite(level_set),
               VACD_ERR_INVALID, "remesh needs scale > 0 and level_set >= 0");
  for (int64_t t = 0; t < 3 * nT; ++t)
    VACD_REQUIRE((int64_t)T[t] < nV, VACD_ERR_INVALID, "triangle index out of range");
  hipStream_t st = ctx->stream;
  // grid: nodes at integer multiples of 1 / scale (the reference's index
  // space), covering the mesh's box widened by the level and two voxels
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) lo[k] = INFINITY, hi[k] = -INFINITY;
  for (int64_t v = 0; v < nV; ++v)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], (double)V[3 * v + k]), hi[k] = std::max(hi[k], (double)V[3 * v + k]);
  const int margin = (int)std::ceil(level_set * scale) + 2;
  Grid g;
  g.h = 1.0 / scale;
  int64_t N = 1;
  for (int k = 0; k < 3; ++k) {
    const int64_t i0 = (int64_t)std::floor(lo[k] * scale) - margin;
    const int64_t i1 = (int64_t)std::ceil(hi[k] * scale) + margin;
    VACD_REQUIRE(i1 - i0 + 1 <= 2048, VACD_ERR_LIMIT, "remesh grid above 2048 nodes per axis");
    g.lo[k] = (double)i0 / scale;
    g.n[k] = (int)(i1 - i0 + 1);
    N *= g.n[k];
  }
  VACD_REQUIRE(N < (int64_t(1) << 30), VACD_ERR_LIMIT, "remesh grid too large");
  // the mesh's BVH for the distances (closedness picks the sign method)
  vacd_bvh bvh;
  DevBuf dv, dt;
  dv.ensure(12 * (size_t)nV);
  dt.ensure(12 * (size_t)nT);
  VACD_HIP(hipMemcpyAsync(dv.ptr, V, 12 * (size_t)nV, hipMemcpyHostToDevice, st));
  VACD_HIP(hipMemcpyAsync(dt.ptr, T, 12 * (size_t)nT, hipMemcpyHostToDevice, st));
  try {
    build_lbvh(ctx, dv.as<float>(), nV, dt.as<uint32_t>(), nT, &bvh);
    const int64_t C = (int64_t)(g.n[0] - 1) * (g.n[1] - 1) * (g.n[2] - 1);
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_sdf = 0, o_flag = o_sdf + up(4 * (size_t)N), o_cvert = o_flag + up(4 * (size_t)(C + 1)),
                 o_qflag = o_cvert + up(12 * (size_t)C), o_over = o_qflag + up(4 * (size_t)(3 * N + 1)),
                 bytes = o_over + 256;
    DevBuf work;
    work.ensure(bytes);
    char *base = work.as<char>();
    float *sdf = reinterpret_cast<float *>(base + o_sdf);
    uint32_t *flag = reinterpret_cast<uint32_t *>(base + o_flag);
    float *cvert = reinterpret_cast<float *>(base + o_cvert);
    uint32_t *qflag = reinterpret_cast<uint32_t *>(base + o_qflag);
    unsigned *over = reinterpret_cast<unsigned *>(base + o_over);
    const int mode = signed_grid(ctx, g, bvh, dv.as<float>(), dt.as<uint32_t>(), nT, sdf, over);
    if (sign_mode) *sign_mode = mode;
    const float iso = (float)level_set;
    k_sn_cells<<<(unsigned)((C + 255) / 256), 256, 0, st>>>(g, sdf, iso, flag, cvert);
    k_zero_u32<<<1, 1, 0, st>>>(flag + C);  // the scan's total lands here
    scan_u32(ctx, flag, C + 1);
    k_sn_quads<<<(unsigned)((3 * N + 255) / 256), 256, 0, st>>>(g, sdf, iso, 0, qflag, nullptr, nullptr);
    k_zero_u32<<<1, 1, 0, st>>>(qflag + 3 * N);
    scan_u32(ctx, qflag, 3 * N + 1);
    uint32_t nverts = 0, nquads = 0;
    VACD_HIP(hipMemcpyAsync(&nverts, flag + C, 4, hipMemcpyDeviceToHost, st));
    VACD_HIP(hipMemcpyAsync(&nquads, qflag + 3 * N, 4, hipMemcpyDeviceToHost, st));
    VACD_HIP(hipStreamSynchronize(st));
    // the cell flags were scanned in place: flag[c] is now the cell's vertex
    // id; k_sn_cells runs again to restore the flags for the compaction
    DevBuf outv, outt, ids;
    outv.ensure(12 * (size_t)std::max<uint32_t>(nverts, 1));
    outt.ensure(24 * (size_t)std::max<uint32_t>(nquads, 1));
    ids.ensure(4 * (size_t)(C + 1));
    VACD_HIP(hipMemcpyAsync(ids.ptr, flag, 4 * (size_t)(C + 1), hipMemcpyDeviceToDevice, st));
    k_sn_cells<<<(unsigned)((C + 255) / 256), 256, 0, st>>>(g, sdf, iso, flag, cvert);
    k_sn_compact<<<(unsigned)((C + 255) / 256), 256, 0, st>>>(C, flag, ids.as<uint32_t>(), cvert,
                                                             outv.as<float>());
    k_sn_quads<<<(unsigned)((3 * N + 255) / 256), 256, 0, st>>>(g, sdf, iso, 1, qflag,
                                                               ids.as<uint32_t>(), outt.as<uint32_t>());
    VACD_HIP(hipGetLastError());
    Vout.resize(3 * (size_t)nverts);
    Tout.resize(6 * (size_t)nquads);
    if (nverts)
      VACD_HIP(hipMemcpyAsync(Vout.data(), outv.ptr, 12 * (size_t)nverts, hipMemcpyDeviceToHost, st));
    if (nquads)
      VACD_HIP(hipMemcpyAsync(Tout.data(), outt.ptr, 24 * (size_t)nquads, hipMemcpyDeviceToHost, st));
    VACD_HIP(hipStreamSynchronize(st));
    work.release();
    outv.release();
    outt.release();
    ids.release();
  } catch (...) {
    bvh.release();
    dv.release();
    dt.release();
    throw;
  }
  bvh.release();
  dv.release();
  dt.release();
}

}  // namespace vacd
