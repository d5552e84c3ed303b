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
#include <cstring>

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
//   k_dist_batch: a block takes 256 samples and one chunk; lanes whose bound
//     the chunk's sphere cannot beat skip it (the block skips it when all
//     do); the rest stage its triangles 256 at a time in LDS with their own
//     spheres and measure exactly those that can come closer than the best
//     so far (relative slack well above float32 rounding).
// Per-sample minima (squared, as float bits) meet through atomicMin; a last
// kernel takes each query's maximum.  The minimum over all triangles equals
// that of the BVH walk, which prunes the same way.
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

__global__ __launch_bounds__(256) void k_dist_batch(const DistQ *__restrict__ qs, int nq,
                                                    const float *__restrict__ pts,
                                                    const float *__restrict__ V,
                                                    const uint32_t *__restrict__ T,
                                                    const float4 *__restrict__ csph,
                                                    unsigned *__restrict__ dmin) {
  __shared__ float4 sa[256], sb[256], sc[256];
  __shared__ float sr[256];
  const int64_t blk = blockIdx.x;
  const DistQ q = qs[find_query(qs, nq, blk, 0)];  // block-uniform
  const int64_t local = blk - q.blk0;
  const int64_t st = local / q.tchunks, tc = local % q.tchunks;
  const int64_t s = st * 256 + threadIdx.x;
  const bool valid = s < q.m;
  float p[3] = {0.f, 0.f, 0.f};
  float best = 0.f;
  if (valid) {
    p[0] = pts[3 * (q.pt0 + s)], p[1] = pts[3 * (q.pt0 + s) + 1], p[2] = pts[3 * (q.pt0 + s) + 2];
    best = __uint_as_float(dmin[q.pt0 + s]);
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
    __syncthreads();
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
    for (int k = 0; k < cnt; ++k) {
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
}  // namespace

void max_distance_batch(vacd_ctx *ctx, const vacd_distance_query *q, int nq, float *out) {
  hipStream_t st = ctx->stream;
  std::vector<uint32_t> cell;
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
  for (int i = 0; i < nq; ++i) {
    const DistQ &d = qs[(size_t)i];
    if (d.m) sort_samples(q[i].pts, d.m, reinterpret_cast<float *>(hp + 12 * d.pt0), cell);
    if (q[i].nV) std::memcpy(hv + 12 * d.v0, q[i].V, 12 * (size_t)q[i].nV);
    if (d.nT) {
      const uint32_t *src = q[i].T;
      for (int64_t t = 0; t < 3 * d.nT; ++t)
        VACD_REQUIRE(src[t] < (uint64_t)q[i].nV, VACD_ERR_INVALID,
                     "triangle index out of range in a distance query");
      std::memcpy(ht + 12 * d.t0, src, 12 * (size_t)d.nT);
    }
  }
  DevBuf &b = ctx->buf[SLOT_DIST];
  const size_t b_min = ((4 * (size_t)np + 255) & ~size_t(255));
  const size_t b_out = ((4 * (size_t)nq + 255) & ~size_t(255));
  const size_t b_ch = 16 * (size_t)std::max<int64_t>(nch, 1);
  b.ensure(bytes + b_min + b_out + b_ch);
  char *d = b.as<char>();
  pinned_upload(ctx, PINNED_DIST, d, bytes);
  const DistQ *dq = reinterpret_cast<const DistQ *>(d);
  const float *dp = reinterpret_cast<const float *>(d + b_q);
  const float *dv = reinterpret_cast<const float *>(d + b_q + b_p);
  const uint32_t *dt = reinterpret_cast<const uint32_t *>(d + b_q + b_p + b_v);
  unsigned *dmin = reinterpret_cast<unsigned *>(d + bytes);
  float *dout = reinterpret_cast<float *>(d + bytes + b_min);
  float4 *csph = reinterpret_cast<float4 *>(d + bytes + b_min + b_out);
  if (nsb) k_dist_bound<<<(unsigned)nsb, 256, 0, st>>>(dq, nq, dp, dv, dt, dmin);
  if (nch) k_dist_chunks<<<(unsigned)nch, 256, 0, st>>>(dq, nq, nch, dv, dt, csph);
  if (nb) k_dist_batch<<<(unsigned)nb, 256, 0, st>>>(dq, nq, dp, dv, dt, csph, dmin);
  if (nq) k_dist_max<<<(unsigned)nq, 256, 0, st>>>(dq, dmin, dout);
  VACD_HIP(hipGetLastError());
  VACD_HIP(hipMemcpyAsync(out, dout, 4 * (size_t)nq, hipMemcpyDeviceToHost, st));
  VACD_HIP(hipStreamSynchronize(st));
}

}  // namespace vacd
