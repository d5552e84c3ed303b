// Plane-scoring kernels for gfx950.
//
// Replaces plane_point_side_and_rate_kernel and classify_and_rate_planes
// (lib/cuda/plane_intersections.cu:7-133), which give every (plane, edge)
// pair a thread and float-atomicAdd the edge length into scores[plane].
// Here:
//   k_edge_prep  one 32-byte record per edge: both endpoints and the edge
//                length rounded once to Q32.32 (int64).
//   k_score      a block owns 2048 planes (8 per lane, in registers) and one
//                chunk of edges; every lane reads the same edge record, so the
//                records come through scalar loads; crossings add the Q32.32
//                length to per-lane int64 sums, flushed with one integer
//                atomic per plane per block.  Integer addition makes the
//                result exact and independent of order.
//   k_finalize   float view of the sums and the argmax (first maximum).
// Side rule (lib/cuda/plane_intersections.cu:43-61): val = a*x+b*y+c*z+d,
// evaluated as fmaf(c, z, fmaf(b, y, a*x)) + d; side +1 if val > 1e-6, -1
// if val < -1e-6, else 0; an edge counts iff its sides are +1 and -1.
#include "internal.hpp"

#include <algorithm>

namespace vacd {
namespace {

constexpr int kBlock = 256;
constexpr int kPPT = 8;                     // planes per lane
constexpr int kPlanesPerBlock = kBlock * kPPT;
constexpr int kEdgeQuantum = 32;            // edge chunks are multiples of this
constexpr float kEps = 1e-6f;

struct alignas(16) EdgeRec {
  float4 a;  // x0 y0 z0 x1
  float4 b;  // y1 z1 lenq(lo) lenq(hi) | kSameFirst
};
// Set in lenq(hi) when the edge's first endpoint is the previous edge's (edges
// in k order come in runs of one first vertex); lengths stay below 2^36.
constexpr unsigned kSameFirst = 0x40000000u;

__global__ __launch_bounds__(256) void k_edge_prep(const float *__restrict__ pts, int64_t n,
                                                   const uint32_t *__restrict__ edges, int64_t E,
                                                   EdgeRec *__restrict__ rec) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= E) return;
  const uint2 ij = reinterpret_cast<const uint2 *>(edges)[e];
  EdgeRec r;
  if ((int64_t)ij.x >= n || (int64_t)ij.y >= n) {
    // out-of-range endpoints contribute nothing (lib/cuda/plane_intersections.cu:34-36)
    r.a = make_float4(0.f, 0.f, 0.f, 0.f);
    r.b = make_float4(0.f, 0.f, 0.f, 0.f);
  } else {
    const float x0 = pts[3 * (size_t)ij.x], y0 = pts[3 * (size_t)ij.x + 1],
                z0 = pts[3 * (size_t)ij.x + 2];
    const float x1 = pts[3 * (size_t)ij.y], y1 = pts[3 * (size_t)ij.y + 1],
                z1 = pts[3 * (size_t)ij.y + 2];
    const float dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    const float len = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
    const long long q = __double2ll_rn((double)len * 4294967296.0);
    // the previous record must hold real coordinates for its first endpoint
    const uint2 prev = e > 0 ? reinterpret_cast<const uint2 *>(edges)[e - 1] : make_uint2(~0u, ~0u);
    const bool same_first =
        e > 0 && prev.x == ij.x && (int64_t)prev.x < n && (int64_t)prev.y < n;
    r.a = make_float4(x0, y0, z0, x1);
    r.b = make_float4(y1, z1, __int_as_float((int)(q & 0xffffffffll)),
                      __uint_as_float((unsigned)(q >> 32) | (same_first ? kSameFirst : 0u)));
  }
  rec[e] = r;
}

// Planes held per lane, oriented per run of edges that share a first
// endpoint u: the copy (fa, fb, fc, fd) is the plane itself when u is on its
// + side, the negated plane when u is on its - side, and (0, 0, 0, 1) when u
// is in the dead zone.  Negation is exact in round-to-nearest, so the
// oriented value of the second endpoint v is exactly +-val(v), and the edge
// crosses iff that value is < -1e-6 (with (0, 0, 0, 1) it never is).
template <int PPT>
struct Lanes {
  float pa[PPT], pb[PPT], pc[PPT], pd[PPT];
  float fa[PPT], fb[PPT], fc[PPT], fd[PPT];
  unsigned long long sum[PPT];
};

__device__ __forceinline__ float plane_val(float a, float b, float c, float d, float x, float y,
                                           float z) {
  return fmaf(c, z, fmaf(b, y, a * x)) + d;
}

// Edge records are read with wave-uniform addresses (every lane of a block
// walks the same chunk), so they arrive through scalar loads and feed the
// vector ALU as scalar operands; no LDS staging is needed.
__global__ __launch_bounds__(kBlock) void k_score(const float4 *__restrict__ planes, int64_t P,
                                                  const EdgeRec *__restrict__ rec, int64_t E,
                                                  int64_t chunk, unsigned long long *__restrict__ acc) {
  Lanes<kPPT> L;
#pragma unroll
  for (int r = 0; r < kPPT; ++r) {
    const int64_t p = (int64_t)blockIdx.x * kPlanesPerBlock + r * kBlock + threadIdx.x;
    const float4 pl = p < P ? planes[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    L.pa[r] = pl.x, L.pb[r] = pl.y, L.pc[r] = pl.z, L.pd[r] = pl.w;
    L.fa[r] = 0.f, L.fb[r] = 0.f, L.fc[r] = 0.f, L.fd[r] = 1.f;
    L.sum[r] = 0ull;
  }
  const int64_t e0 = (int64_t)blockIdx.y * chunk;
  const int64_t e1 = e0 + chunk < E ? e0 + chunk : E;
  if (e0 >= e1) return;
  EdgeRec nxt = rec[e0];
  for (int64_t e = e0; e < e1; ++e) {
    const EdgeRec rc = nxt;
    if (e + 1 < e1) nxt = rec[e + 1];  // issued one edge ahead of its use
    const float4 ra = rc.a, rb = rc.b;
    const unsigned hi = __float_as_uint(rb.w);
    const unsigned long long lq =
        ((unsigned long long)(hi & ~kSameFirst) << 32) | (unsigned)__float_as_int(rb.z);
    if (e == e0 || !(hi & kSameFirst)) {  // new first endpoint: re-orient the planes
#pragma unroll
      for (int r = 0; r < kPPT; ++r) {
        const float vu = plane_val(L.pa[r], L.pb[r], L.pc[r], L.pd[r], ra.x, ra.y, ra.z);
        const float sg = vu > kEps ? 1.f : (vu < -kEps ? -1.f : 0.f);
        L.fa[r] = sg * L.pa[r];
        L.fb[r] = sg * L.pb[r];
        L.fc[r] = sg * L.pc[r];
        L.fd[r] = sg != 0.f ? sg * L.pd[r] : 1.f;
      }
    }
#pragma unroll
    for (int r = 0; r < kPPT; ++r) {
      const float vv = plane_val(L.fa[r], L.fb[r], L.fc[r], L.fd[r], ra.w, rb.x, rb.y);
      L.sum[r] += vv < -kEps ? lq : 0ull;
    }
  }
#pragma unroll
  for (int r = 0; r < kPPT; ++r) {
    const int64_t p = (int64_t)blockIdx.x * kPlanesPerBlock + r * kBlock + threadIdx.x;
    if (p < P && L.sum[r]) atomicAdd(&acc[p], L.sum[r]);
  }
}

__global__ __launch_bounds__(1024) void k_finalize(const unsigned long long *__restrict__ acc,
                                                   int64_t P, int64_t *__restrict__ q32,
                                                   float *__restrict__ f, int32_t *__restrict__ amax) {
  __shared__ long long bv[1024];
  __shared__ int bi[1024];
  long long best = -1;
  int besti = -1;
  for (int64_t p = threadIdx.x; p < P; p += 1024) {
    const long long v = (long long)acc[p];
    if (q32) q32[p] = v;
    if (f) f[p] = (float)((double)v * 2.3283064365386963e-10);
    if (v > best) best = v, besti = (int)p;  // p increases: keeps the first max
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = besti;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const long long ov = bv[threadIdx.x + s];
      const int oi = bi[threadIdx.x + s];
      if (ov > bv[threadIdx.x] || (ov == bv[threadIdx.x] && oi >= 0 && (bi[threadIdx.x] < 0 || oi < bi[threadIdx.x]))) {
        bv[threadIdx.x] = ov;
        bi[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && amax) *amax = bi[0];
}

}  // namespace

void launch_scoring(vacd_ctx *ctx, const float *d_planes, int64_t P, const float *d_pts, int64_t n,
                    const uint32_t *d_edges, int64_t E, int64_t *d_q32, float *d_f,
                    int32_t *d_argmax) {
  hipStream_t st = ctx->stream;
  VACD_REQUIRE(P >= 0 && P < (int64_t(1) << 31), VACD_ERR_LIMIT, "scoring needs P < 2^31");
  VACD_REQUIRE(E >= 0 && E < (int64_t(1) << 40), VACD_ERR_LIMIT, "scoring needs E < 2^40");
  if (P == 0) {
    if (d_argmax) VACD_HIP(hipMemsetAsync(d_argmax, 0xff, sizeof(int32_t), st));
    return;
  }
  vacd::DevBuf &accb = ctx->buf[SLOT_SCORES];
  accb.ensure(8 * (size_t)P);
  unsigned long long *acc = accb.as<unsigned long long>();
  VACD_HIP(hipMemsetAsync(acc, 0, 8 * (size_t)P, st));
  if (E > 0) {
    vacd::DevBuf &recb = ctx->buf[SLOT_EREC];
    recb.ensure(sizeof(EdgeRec) * (size_t)E);
    EdgeRec *rec = recb.as<EdgeRec>();
    k_edge_prep<<<(unsigned)((E + 255) / 256), 256, 0, st>>>(d_pts, n, d_edges, E, rec);
    const int64_t ptiles = (P + kPlanesPerBlock - 1) / kPlanesPerBlock;
    // about 4096 blocks in total unless the context fixes the chunk length
    int64_t nchunks = std::max<int64_t>(1, 4096 / ptiles);
    int64_t chunk = ctx->score_chunk > 0 ? ctx->score_chunk : (E + nchunks - 1) / nchunks;
    chunk = std::max<int64_t>(kEdgeQuantum, (chunk + kEdgeQuantum - 1) / kEdgeQuantum * kEdgeQuantum);
    nchunks = (E + chunk - 1) / chunk;
    VACD_REQUIRE(nchunks < 65536, VACD_ERR_LIMIT, "too many edge chunks");
    dim3 grid((unsigned)ptiles, (unsigned)nchunks);
    timed_begin(ctx, TIMED_SCORE);
    k_score<<<grid, kBlock, 0, st>>>(reinterpret_cast<const float4 *>(d_planes), P, rec, E, chunk,
                                     acc);
    timed_end(ctx, TIMED_SCORE);
  }
  k_finalize<<<1, 1024, 0, st>>>(acc, P, d_q32, d_f, d_argmax);
  VACD_HIP(hipGetLastError());
}

}  // namespace vacd
