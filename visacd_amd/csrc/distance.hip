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
