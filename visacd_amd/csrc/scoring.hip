// Plane-scoring kernels for gfx950.
//
// Replaces plane_point_side_and_rate_kernel and classify_and_rate_planes
// (lib/cuda/plane_intersections.cu:7-133), which give every (plane, edge)
// pair a thread and float-atomicAdd the edge length into scores[plane].
// Here:
//   k_edge_prep  one 32-byte record per edge: both endpoints and the edge
//                length rounded once to Q32.32 (int64).
//   k_score      a block (one wave) owns 512 planes (8 per lane, in
//                registers) and one chunk of edges, staged through LDS 256
//                records at a time; every lane reads the same record (a
//                broadcast); crossings add the Q32.32 length to per-lane
//                int64 sums, flushed with one integer atomic per plane per
//                block.  Integer addition makes the
//                result exact and independent of order.
//   k_finalize   float view of the sums and the argmax (first maximum).
// Side rule (lib/cuda/plane_intersections.cu:43-61): val = a*x+b*y+c*z+d,
// evaluated as fmaf(c, z, fmaf(b, y, a*x)) + d; side +1 if val > 1e-6, -1
// if val < -1e-6, else 0; an edge counts iff its sides are +1 and -1.
#include "internal.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace vacd {
namespace {

constexpr int kBlock = 64;
constexpr int kPPT = 8;                     // planes per lane
constexpr int kPlanesPerBlock = kBlock * kPPT;
constexpr int kEdgeQuantum = 32;            // edge chunks are multiples of this
constexpr int kStage = 256;                 // edge records per LDS stage
// one-wave blocks, each a sequential walk over its edges: about eight per
// SIMD hide the LDS and ALU latencies (measured: 4,096 planes x 48,869 edges
// take 56 us with 64-edge chunks, 6,100 blocks, against 93 us with 256)
constexpr int64_t kTargetBlocks = 8192;
constexpr float kEps = 1e-6f;

struct alignas(16) EdgeRec {
  float4 a;  // x0 y0 z0 x1
  float4 b;  // y1 z1 lenq(lo) lenq(hi) | kSameFirst
};
// Set in lenq(hi) when the edge's first endpoint is the previous edge's (edges
// in k order come in runs of one first vertex); lengths stay below 2^36.
constexpr unsigned kSameFirst = 0x40000000u;

// Record of edge e of one part (pts[n], edges[E]).
__device__ __forceinline__ EdgeRec edge_record(const float *__restrict__ pts, int64_t n,
                                               const uint32_t *__restrict__ edges, int64_t e) {
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
  return r;
}

__global__ __launch_bounds__(256) void k_edge_prep(const float *__restrict__ pts, int64_t n,
                                                   const uint32_t *__restrict__ edges, int64_t E,
                                                   EdgeRec *__restrict__ rec) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= E) return;
  rec[e] = edge_record(pts, n, edges, e);
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

// One block's work: planes [tile * kPlanesPerBlock, +kPlanesPerBlock) of a
// plane set (P planes; its sums at acc[0..P)) against edge records [e0, e1).
// Every lane walks the same records; they are staged in LDS by one coalesced
// load per stage (measured faster than scalar loads issued one record ahead:
// 56 against 70 us for 4,096 planes x 48,869 edges).
__device__ __forceinline__ void score_tile(const float4 *__restrict__ planes, int64_t P,
                                           int64_t tile, const EdgeRec *__restrict__ rec,
                                           int64_t e0, int64_t e1,
                                           unsigned long long *__restrict__ acc) {
  Lanes<kPPT> L;
#pragma unroll
  for (int r = 0; r < kPPT; ++r) {
    const int64_t p = tile * kPlanesPerBlock + r * kBlock + threadIdx.x;
    const float4 pl = p < P ? planes[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    L.pa[r] = pl.x, L.pb[r] = pl.y, L.pc[r] = pl.z, L.pd[r] = pl.w;
    L.fa[r] = 0.f, L.fb[r] = 0.f, L.fc[r] = 0.f, L.fd[r] = 1.f;
    L.sum[r] = 0ull;
  }
  if (e0 >= e1) return;
  // the chunk's records pass through LDS kStage at a time: one coalesced
  // load by the block, then every lane reads each record from LDS (the same
  // address for all lanes: a broadcast)
  __shared__ EdgeRec stage[kStage];
  for (int64_t es = e0; es < e1; es += kStage) {
    const int cnt = (int)(e1 - es < kStage ? e1 - es : kStage);
    __syncthreads();
    for (int x = threadIdx.x; x < 2 * cnt; x += kBlock)
      reinterpret_cast<float4 *>(stage)[x] = reinterpret_cast<const float4 *>(rec + es)[x];
    __syncthreads();
  for (int g = 0; g < cnt; ++g) {
    const int64_t e = es + g;
    const EdgeRec rc = stage[g];
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
  }
#pragma unroll
  for (int r = 0; r < kPPT; ++r) {
    const int64_t p = tile * kPlanesPerBlock + r * kBlock + threadIdx.x;
    if (p < P && L.sum[r]) atomicAdd(&acc[p], L.sum[r]);
  }
}

__global__ __launch_bounds__(kBlock) void k_score(const float4 *__restrict__ planes, int64_t P,
                                                  const EdgeRec *__restrict__ rec, int64_t E,
                                                  int64_t chunk, unsigned long long *__restrict__ acc) {
  const int64_t e0 = (int64_t)blockIdx.y * chunk;
  const int64_t e1 = e0 + chunk < E ? e0 + chunk : E;
  score_tile(planes, P, blockIdx.x, rec, e0, e1, acc);
}

// q32 / float views of P sums and their first maximum (a 1024-thread block).
__device__ __forceinline__ void finalize_range(const unsigned long long *__restrict__ acc,
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

__global__ __launch_bounds__(1024) void k_finalize(const unsigned long long *__restrict__ acc,
                                                   int64_t P, int64_t *__restrict__ q32,
                                                   float *__restrict__ f, int32_t *__restrict__ amax) {
  finalize_range(acc, P, q32, f, amax);
}

// ---------------------------------------------------------------- batch
struct ScorePartDev {
  const float4 *planes;
  int64_t P, plane_base;  // plane_base: index of its first plane in the batch
  const float *pts;
  int64_t n;
  const uint32_t *edges;
  int64_t E, rec_base;    // rec_base: index of its first edge record
};
struct ScoreBlock {
  int part, tile;
  int64_t e0, e1;  // edge records of the batch
};

__global__ __launch_bounds__(256) void k_edge_prep_batch(const ScorePartDev *__restrict__ parts,
                                                         int nparts, int64_t Etot,
                                                         EdgeRec *__restrict__ rec) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= Etot) return;
  int lo = 0, hi = nparts - 1;  // last part with rec_base <= e
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (parts[mid].rec_base <= e) lo = mid;
    else hi = mid - 1;
  }
  const ScorePartDev &P = parts[lo];
  rec[e] = edge_record(P.pts, P.n, P.edges, e - P.rec_base);
}

__global__ __launch_bounds__(kBlock) void k_score_batch(const ScorePartDev *__restrict__ parts,
                                                        const ScoreBlock *__restrict__ blocks,
                                                        const EdgeRec *__restrict__ rec,
                                                        unsigned long long *__restrict__ acc) {
  const ScoreBlock b = blocks[blockIdx.x];
  const ScorePartDev &P = parts[b.part];
  score_tile(P.planes, P.P, b.tile, rec, b.e0, b.e1, acc + P.plane_base);
}

__global__ __launch_bounds__(1024) void k_finalize_batch(const unsigned long long *__restrict__ acc,
                                                         const ScorePartDev *__restrict__ parts,
                                                         int64_t *__restrict__ q32,
                                                         int32_t *__restrict__ amax) {
  const ScorePartDev &P = parts[blockIdx.x];
  finalize_range(acc + P.plane_base, P.P, q32 ? q32 + P.plane_base : nullptr, nullptr,
                 amax + blockIdx.x);
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long *__restrict__ p, int64_t n,
                                                  unsigned long long v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

// The launches of one scoring call, given its scratch: acc (P sums) and
// rec (E edge records).  Nothing here allocates.
static void score_launches(vacd_ctx *ctx, const float *d_planes, int64_t P, const float *d_pts,
                           int64_t n, const uint32_t *d_edges, int64_t E,
                           unsigned long long *acc, EdgeRec *rec, int64_t *d_q32, float *d_f,
                           int32_t *d_argmax, bool timed) {
  hipStream_t st = ctx->stream;
  fill_u64(ctx, acc, P, 0ull);
  if (E > 0) {
    k_edge_prep<<<(unsigned)((E + 255) / 256), 256, 0, st>>>(d_pts, n, d_edges, E, rec);
    const int64_t ptiles = (P + kPlanesPerBlock - 1) / kPlanesPerBlock;
    // about kTargetBlocks blocks in total unless the context fixes the chunk length
    int64_t nchunks = std::max<int64_t>(1, kTargetBlocks / ptiles);
    int64_t chunk = ctx->score_chunk > 0 ? ctx->score_chunk : (E + nchunks - 1) / nchunks;
    chunk = std::max<int64_t>(kEdgeQuantum, (chunk + kEdgeQuantum - 1) / kEdgeQuantum * kEdgeQuantum);
    nchunks = (E + chunk - 1) / chunk;
    VACD_REQUIRE(nchunks < 65536, VACD_ERR_LIMIT, "too many edge chunks");
    dim3 grid((unsigned)ptiles, (unsigned)nchunks);
    if (timed) timed_begin(ctx, TIMED_SCORE);
    k_score<<<grid, kBlock, 0, st>>>(reinterpret_cast<const float4 *>(d_planes), P, rec, E, chunk,
                                     acc);
    if (timed) timed_end(ctx, TIMED_SCORE);
  }
  k_finalize<<<1, 1024, 0, st>>>(acc, P, d_q32, d_f, d_argmax);
  VACD_HIP(hipGetLastError());
}

void launch_scoring(vacd_ctx *ctx, const float *d_planes, int64_t P, const float *d_pts, int64_t n,
                    const uint32_t *d_edges, int64_t E, int64_t *d_q32, float *d_f,
                    int32_t *d_argmax) {
  hipStream_t st = ctx->stream;
  VACD_REQUIRE(P >= 0 && P < (int64_t(1) << 31), VACD_ERR_LIMIT, "scoring needs P < 2^31");
  VACD_REQUIRE(E >= 0 && E < (int64_t(1) << 40), VACD_ERR_LIMIT, "scoring needs E < 2^40");
  if (P == 0) {
    if (d_argmax) {
      // -1: the first 4 bytes of an all-ones word written into the 8-byte slot
      vacd::DevBuf &tmp = ctx->buf[SLOT_ARGMAX];
      tmp.ensure(8);
      fill_u64(ctx, tmp.ptr, 1, ~0ull);
      VACD_HIP(hipMemcpyAsync(d_argmax, tmp.ptr, sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    }
    return;
  }
  vacd::DevBuf &accb = ctx->buf[SLOT_SCORES];
  accb.ensure(8 * (size_t)P);
  EdgeRec *rec = nullptr;
  if (E > 0) {
    vacd::DevBuf &recb = ctx->buf[SLOT_EREC];
    recb.ensure(sizeof(EdgeRec) * (size_t)E);
    rec = recb.as<EdgeRec>();
  }
  score_launches(ctx, d_planes, P, d_pts, n, d_edges, E, accb.as<unsigned long long>(), rec,
                 d_q32, d_f, d_argmax, true);
}

// ------------------------------------------------------- allocation-free form
// Scratch of vacd_score_planes_noalloc_d: the Q32.32 sums of up to
// VACD_NOALLOC_MAX_PLANES planes, then the records of up to
// VACD_NOALLOC_MAX_EDGES edges.  Each GPU holds its own copy from module load
// until the process ends.  Host code never names the array itself (its name
// is a host-side stand-in, not a device address): noalloc_setup() looks up
// the device address once per device and every launch uses that.
namespace {
constexpr size_t kNoAllocAccBytes = 8 * (size_t)VACD_NOALLOC_MAX_PLANES;
constexpr size_t kNoAllocBytes = kNoAllocAccBytes + sizeof(EdgeRec) * (size_t)VACD_NOALLOC_MAX_EDGES;
static_assert(kNoAllocAccBytes % 256 == 0, "edge records start 256-byte aligned");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocDevices] = {};

__global__ void k_set_i32(int32_t *p, int32_t v) { *p = v; }
}  // namespace

void noalloc_setup(int device) {
  VACD_REQUIRE(device >= 0 && device < kNoAllocDevices, VACD_ERR_LIMIT,
               "allocation-free scoring supports device ordinals below 64");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return;
  int cur = -1;
  VACD_HIP(hipGetDevice(&cur));
  VACD_REQUIRE(cur == device, VACD_ERR_INVALID, "noalloc_setup: device is not current");
  void *p = nullptr;
  VACD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  VACD_REQUIRE(p != nullptr, VACD_ERR_HIP, "no device address for the scoring scratch array");
  g_noalloc_ptr[device] = p;
}

void launch_scoring_noalloc(vacd_ctx *ctx, const float *d_planes, int64_t P, const float *d_pts,
                            int64_t n, const uint32_t *d_edges, int64_t E, int64_t *d_q32,
                            float *d_f, int32_t *d_argmax) {
  VACD_REQUIRE(P <= VACD_NOALLOC_MAX_PLANES, VACD_ERR_LIMIT,
               "allocation-free scoring: more planes than VACD_NOALLOC_MAX_PLANES (65,536)");
  VACD_REQUIRE(E <= VACD_NOALLOC_MAX_EDGES, VACD_ERR_LIMIT,
               "allocation-free scoring: more edges than VACD_NOALLOC_MAX_EDGES (1,048,576)");
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    base = g_noalloc_ptr[ctx->device];
  }
  VACD_REQUIRE(base != nullptr, VACD_ERR_INVALID,
               "allocation-free scoring: the context's device has no scratch address");
  if (P == 0) {
    if (d_argmax) k_set_i32<<<1, 1, 0, ctx->stream>>>(d_argmax, -1);
    VACD_HIP(hipGetLastError());
    return;
  }
  unsigned long long *acc = static_cast<unsigned long long *>(base);
  EdgeRec *rec = reinterpret_cast<EdgeRec *>(static_cast<char *>(base) + kNoAllocAccBytes);
  score_launches(ctx, d_planes, P, d_pts, n, d_edges, E, acc, rec, d_q32, d_f, d_argmax, false);
}

void fill_u64(vacd_ctx *ctx, void *p, int64_t n, unsigned long long v) {
  if (n <= 0) return;
  const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
  k_fill_u64<<<grid, 256, 0, ctx->stream>>>(static_cast<unsigned long long *>(p), n, v);
  VACD_HIP(hipGetLastError());
}

void launch_scoring_batch(vacd_ctx *ctx, const vacd_score_part *parts, int nparts,
                          int64_t *d_q32, int32_t *d_argmax) {
  hipStream_t st = ctx->stream;
  VACD_REQUIRE(nparts >= 0, VACD_ERR_INVALID, "negative part count");
  if (nparts == 0) return;
  VACD_REQUIRE(d_argmax, VACD_ERR_INVALID, "NULL argmax");
  std::vector<ScorePartDev> dev(nparts);
  int64_t Ptot = 0, Etot = 0, work = 0;
  for (int p = 0; p < nparts; ++p) {
    const vacd_score_part &q = parts[p];
    VACD_REQUIRE(q.P >= 0 && q.n >= 0 && q.E >= 0, VACD_ERR_INVALID, "negative size");
    VACD_REQUIRE(q.P == 0 || (q.planes && ((uintptr_t)q.planes % 16) == 0), VACD_ERR_INVALID,
                 "planes must be 16-byte aligned");
    VACD_REQUIRE(q.E == 0 || (q.pts && q.edges && ((uintptr_t)q.edges % 8) == 0),
                 VACD_ERR_INVALID, "edges must be 8-byte aligned");
    VACD_REQUIRE(q.P < (int64_t(1) << 31), VACD_ERR_LIMIT, "scoring needs P < 2^31 per part");
    dev[p] = ScorePartDev{reinterpret_cast<const float4 *>(q.planes), q.P, Ptot, q.pts, q.n,
                          q.edges, q.E, Etot};
    Ptot += q.P;
    Etot += q.E;
    work += ((q.P + kPlanesPerBlock - 1) / kPlanesPerBlock) * q.E;
  }
  VACD_REQUIRE(Etot < (int64_t(1) << 40), VACD_ERR_LIMIT, "scoring needs E < 2^40");
  // one edge-chunk length for the whole batch: about kTargetBlocks blocks in total
  int64_t chunk = ctx->score_chunk > 0 ? ctx->score_chunk : (work + kTargetBlocks - 1) / kTargetBlocks;
  chunk = std::max<int64_t>(kEdgeQuantum, (chunk + kEdgeQuantum - 1) / kEdgeQuantum * kEdgeQuantum);
  std::vector<ScoreBlock> blocks;
  for (int p = 0; p < nparts; ++p) {
    const ScorePartDev &q = dev[p];
    const int64_t ptiles = (q.P + kPlanesPerBlock - 1) / kPlanesPerBlock;
    for (int64_t t = 0; t < ptiles; ++t)
      for (int64_t e = 0; e < q.E; e += chunk)
        blocks.push_back(ScoreBlock{p, (int)t, q.rec_base + e, q.rec_base + std::min(e + chunk, q.E)});
  }
  VACD_REQUIRE(blocks.size() < (size_t(1) << 31), VACD_ERR_LIMIT, "too many scoring blocks");
  // tables: parts then blocks, staged through pinned memory
  const size_t parts_bytes = sizeof(ScorePartDev) * (size_t)nparts;
  const size_t blocks_off = (parts_bytes + 255) & ~size_t(255);
  const size_t tables_bytes = blocks_off + sizeof(ScoreBlock) * blocks.size();
  char *host = static_cast<char *>(pinned_stage(ctx, PINNED_SCORE, tables_bytes));
  std::memcpy(host, dev.data(), parts_bytes);
  if (!blocks.empty()) std::memcpy(host + blocks_off, blocks.data(), sizeof(ScoreBlock) * blocks.size());
  vacd::DevBuf &tb = ctx->buf[SLOT_SCORE_TABLES];
  tb.ensure(tables_bytes);
  pinned_upload(ctx, PINNED_SCORE, tb.ptr, tables_bytes);
  const ScorePartDev *d_parts = tb.as<ScorePartDev>();
  const ScoreBlock *d_blocks = reinterpret_cast<const ScoreBlock *>(tb.as<char>() + blocks_off);
  vacd::DevBuf &accb = ctx->buf[SLOT_SCORES];
  accb.ensure(8 * (size_t)std::max<int64_t>(Ptot, 1));
  unsigned long long *acc = accb.as<unsigned long long>();
  fill_u64(ctx, acc, Ptot, 0ull);
  if (Etot > 0) {
    vacd::DevBuf &recb = ctx->buf[SLOT_EREC];
    recb.ensure(sizeof(EdgeRec) * (size_t)Etot);
    EdgeRec *rec = recb.as<EdgeRec>();
    k_edge_prep_batch<<<(unsigned)((Etot + 255) / 256), 256, 0, st>>>(d_parts, nparts, Etot, rec);
    if (!blocks.empty()) {
      timed_begin(ctx, TIMED_SCORE);
      k_score_batch<<<(unsigned)blocks.size(), kBlock, 0, st>>>(d_parts, d_blocks, rec, acc);
      timed_end(ctx, TIMED_SCORE);
    }
  }
  k_finalize_batch<<<(unsigned)nparts, 1024, 0, st>>>(acc, d_parts, d_q32, d_argmax);
  VACD_HIP(hipGetLastError());
}

}  // namespace vacd

// ------------------------------------------------- candidate-plane extents
// Extent of each part's points along each of its candidate normals (the
// sphere-sampled candidate planes put their offset uniformly in it).  A
// thread owns one normal and scans the part's points, staged 1024 at a time
// in LDS; float64 products and sums are rounded one by one (contraction into
// fused multiply-adds is switched off) in the host's order, so the extents
// equal the host loop's.
namespace vacd {
namespace {

struct ExtQ {
  int64_t n0, P;    // first normal, normal count
  int64_t p0, n;    // first point, point count
  int64_t blk0;     // first block
};

constexpr int kExtPts = 1024;

__global__ __launch_bounds__(256) void k_extents(const ExtQ *__restrict__ qs, int nq,
                                                 const double *__restrict__ normals,
                                                 const double *__restrict__ pts,
                                                 double *__restrict__ lo_hi) {
  __shared__ double sp[3 * kExtPts];
  const int64_t blk = blockIdx.x;
  int lo = 0, hi = nq - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (qs[mid].blk0 <= blk) lo = mid;
    else hi = mid - 1;
  }
  const ExtQ q = qs[lo];
  const int64_t i = (blk - q.blk0) * 256 + threadIdx.x;
  const bool valid = i < q.P;
  double nx = 0.0, ny = 0.0, nz = 0.0;
  if (valid) {
    nx = normals[3 * (q.n0 + i)], ny = normals[3 * (q.n0 + i) + 1], nz = normals[3 * (q.n0 + i) + 2];
  }
  double mn = 1e300, mx = -1e300;
  for (int64_t b = 0; b < q.n; b += kExtPts) {
    const int cnt = (int)(q.n - b < kExtPts ? q.n - b : kExtPts);
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * cnt; k += 256) sp[k] = pts[3 * (q.p0 + b) + k];
    __syncthreads();
    if (!valid) continue;
    for (int k = 0; k < cnt; ++k) {
#pragma clang fp contract(off)  // a fused multiply-add would round differently from the host
      const double d = (nx * sp[3 * k] + ny * sp[3 * k + 1]) + nz * sp[3 * k + 2];
      mn = fmin(mn, d), mx = fmax(mx, d);
    }
  }
  if (valid) lo_hi[2 * (q.n0 + i)] = mn, lo_hi[2 * (q.n0 + i) + 1] = mx;
}

}  // namespace

void plane_extents_batch(vacd_ctx *ctx, const vacd_extent_query *q, int nq, double *lo_hi) {
  hipStream_t st = ctx->stream;
  std::vector<ExtQ> qs((size_t)nq);
  int64_t tn = 0, tp = 0, nb = 0;
  for (int i = 0; i < nq; ++i) {
    VACD_REQUIRE(q[i].P >= 0 && q[i].n >= 0, VACD_ERR_INVALID, "negative size in an extent query");
    VACD_REQUIRE((q[i].P == 0 || q[i].normals) && (q[i].n == 0 || q[i].pts), VACD_ERR_INVALID,
                 "NULL array in an extent query");
    qs[(size_t)i] = ExtQ{tn, q[i].P, tp, q[i].n, nb};
    nb += (q[i].P + 255) / 256;
    tn += q[i].P, tp += q[i].n;
  }
  VACD_REQUIRE(nb < (int64_t(1) << 31), VACD_ERR_LIMIT, "extent batch too large");
  const size_t b_q = (sizeof(ExtQ) * (size_t)nq + 255) & ~size_t(255);
  const size_t b_n = (24 * (size_t)tn + 255) & ~size_t(255);
  const size_t b_p = (24 * (size_t)tp + 255) & ~size_t(255);
  const size_t bytes = b_q + b_n + b_p;
  char *h = static_cast<char *>(pinned_stage(ctx, PINNED_DIST, bytes));
  std::memcpy(h, qs.data(), sizeof(ExtQ) * (size_t)nq);
  for (int i = 0; i < nq; ++i) {
    const ExtQ &e = qs[(size_t)i];
    if (e.P) std::memcpy(h + b_q + 24 * e.n0, q[i].normals, 24 * (size_t)e.P);
    if (e.n) std::memcpy(h + b_q + b_n + 24 * e.p0, q[i].pts, 24 * (size_t)e.n);
  }
  DevBuf &b = ctx->buf[SLOT_EXTENT];
  b.ensure(bytes + 16 * (size_t)std::max<int64_t>(tn, 1));
  char *d = b.as<char>();
  pinned_upload(ctx, PINNED_DIST, d, bytes);
  double *dout = reinterpret_cast<double *>(d + bytes);
  if (nb)
    k_extents<<<(unsigned)nb, 256, 0, st>>>(reinterpret_cast<const ExtQ *>(d), nq,
                                            reinterpret_cast<const double *>(d + b_q),
                                            reinterpret_cast<const double *>(d + b_q + b_n), dout);
  VACD_HIP(hipGetLastError());
  if (tn) VACD_HIP(hipMemcpyAsync(lo_hi, dout, 16 * (size_t)tn, hipMemcpyDeviceToHost, st));
  VACD_HIP(hipStreamSynchronize(st));
}

}  // namespace vacd
