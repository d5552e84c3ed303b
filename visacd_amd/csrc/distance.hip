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
// mesh (a dozen launches, allocations and a host wait each) costs far more
// than measuring every sample against every triangle.  A block takes 256
// samples of one query and a chunk of kDistChunk of its triangles, staged
// 256 at a time in LDS with their bounding spheres; a triangle whose sphere
// lies farther than the best distance so far (with a relative slack well
// above float32 rounding) cannot hold the minimum and is skipped.  Per-sample
// minima (squared, as float bits) meet through atomicMin in a device array;
// a second kernel takes each query's maximum.  The minimum over all
// triangles equals that of the BVH walk, which prunes the same way.
namespace vacd {
namespace {

constexpr int kDistChunk = 1024;  // triangles per block

struct DistQ {
  int64_t pt0, m;     // first sample, sample count
  int64_t v0, t0, nT; // first vertex, first triangle, triangle count
  int64_t blk0;       // first block of the query
  int64_t tchunks;    // triangle chunks of the query
};

__global__ __launch_bounds__(256) void k_dist_fill(unsigned *__restrict__ d, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) d[i] = 0x7f800000u;  // +inf
}

__global__ __launch_bounds__(256) void k_dist_batch(const DistQ *__restrict__ qs, int nq,
                                                    const float *__restrict__ pts,
                                                    const float *__restrict__ V,
                                                    const uint32_t *__restrict__ T,
                                                    unsigned *__restrict__ dmin) {
  __shared__ float4 sa[256], sb[256], sc[256];
  __shared__ float sr[256];
  const int64_t blk = blockIdx.x;
  int lo = 0, hi = nq - 1;
  while (lo < hi) {  // query owning this block (block-uniform)
    const int mid = (lo + hi + 1) >> 1;
    if (qs[mid].blk0 <= blk) lo = mid;
    else hi = mid - 1;
  }
  const DistQ q = qs[lo];
  const int64_t local = blk - q.blk0;
  const int64_t st = local / q.tchunks, tc = local % q.tchunks;
  const int64_t s = st * 256 + threadIdx.x;
  const bool valid = s < q.m;
  float p[3] = {0.f, 0.f, 0.f};
  float best = 0.f;
  if (valid) {
    p[0] = pts[3 * (q.pt0 + s)], p[1] = pts[3 * (q.pt0 + s) + 1], p[2] = pts[3 * (q.pt0 + s) + 2];
    best = __uint_as_float(dmin[q.pt0 + s]);  // what other chunks found already
  }
  float reach = sqrtf(best);
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

void max_distance_batch(vacd_ctx *ctx, const vacd_distance_query *q, int nq, float *out) {
  hipStream_t st = ctx->stream;
  std::vector<DistQ> qs((size_t)nq);
  int64_t np = 0, nv = 0, nt = 0, nb = 0;
  for (int i = 0; i < nq; ++i) {
    VACD_REQUIRE(q[i].m >= 0 && q[i].nV >= 0 && q[i].nT >= 0, VACD_ERR_INVALID,
                 "negative size in a distance query");
    VACD_REQUIRE((q[i].m == 0 || q[i].pts) && (q[i].nV == 0 || q[i].V) && (q[i].nT == 0 || q[i].T),
                 VACD_ERR_INVALID, "NULL array in a distance query");
    DistQ &d = qs[(size_t)i];
    d.pt0 = np, d.m = q[i].m, d.v0 = nv, d.t0 = nt, d.nT = q[i].nT, d.blk0 = nb;
    d.tchunks = std::max<int64_t>(1, (q[i].nT + kDistChunk - 1) / kDistChunk);
    if (q[i].m > 0 && q[i].nT > 0) nb += ((q[i].m + 255) / 256) * d.tchunks;
    np += q[i].m, nv += q[i].nV, nt += q[i].nT;
  }
  VACD_REQUIRE(nb < (int64_t(1) << 31), VACD_ERR_LIMIT, "distance batch too large");
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
    if (d.m) std::memcpy(hp + 12 * d.pt0, q[i].pts, 12 * (size_t)d.m);
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
  b.ensure(bytes + b_min + b_out);
  char *d = b.as<char>();
  pinned_upload(ctx, PINNED_DIST, d, bytes);
  const DistQ *dq = reinterpret_cast<const DistQ *>(d);
  const float *dp = reinterpret_cast<const float *>(d + b_q);
  const float *dv = reinterpret_cast<const float *>(d + b_q + b_p);
  const uint32_t *dt = reinterpret_cast<const uint32_t *>(d + b_q + b_p + b_v);
  unsigned *dmin = reinterpret_cast<unsigned *>(d + bytes);
  float *dout = reinterpret_cast<float *>(d + bytes + b_min);
  if (np) k_dist_fill<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(dmin, np);
  if (nb) k_dist_batch<<<(unsigned)nb, 256, 0, st>>>(dq, nq, dp, dv, dt, dmin);
  if (nq) k_dist_max<<<(unsigned)nq, 256, 0, st>>>(dq, dmin, dout);
  VACD_HIP(hipGetLastError());
  VACD_HIP(hipMemcpyAsync(out, dout, 4 * (size_t)nq, hipMemcpyDeviceToHost, st));
  VACD_HIP(hipStreamSynchronize(st));
}

}  // namespace vacd
