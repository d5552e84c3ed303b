// On-device LBVH build for gfx950.
//
// Replaces the OptiX GAS build of run_optix (lib/src/intersections.cpp:131-180).
// Steps, all on the context stream:
//   1. k_bounds   triangle centroid bounds and vertex bounds (block reduce +
//                 ordered-integer atomics).
//   2. k_morton   30-bit Morton code of each centroid quantised to 10 bits per
//                 axis over the centroid bounds.
//   3. radix sort four stable LSD passes of 8 bits: per 256-key tile a digit
//                 histogram, one exclusive scan of the digit-major histogram
//                 table, then a scatter that ranks keys with wave ballots
//                 (match on the 8 digit bits) and per-wave digit counts in LDS.
//   4. k_karras   Karras (HPG 2012) hierarchy: each internal node finds its key
//                 range and split, with the key index breaking ties between
//                 equal codes; the child with fewer leaves is then stored
//                 first (k_small_first).
//   5. k_leaves   leaf-ordered 48-byte triangles.
//   6. boxes      a segment tree over the leaf boxes (k_seg_up, 8 levels per
//                 launch in LDS), then each internal node's two child boxes as
//                 unions over its children's leaf ranges (k_node_boxes).
//   7. k_depth    tree depth (longest leaf-to-root chain) for stack sizing.
// Child boxes are widened by 2^-17 of the scene extent so that the float32
// slab tests of the traversal stay conservative.
// Contact tree (build_contact_tree): triangles whose longest edge exceeds
// twice the mesh's mean longest edge are cut into up to 64 slices parallel
// to their shortest edge; the slices' boxes go through the same steps 1-7
// as primitives of their own, each leaf holding its whole triangle.
// Planar groups: triangles the caller tags as lying in one of up to 7 planes
// (the caps of earlier cuts) get the group code (1..7) as the top 3 bits of
// their 30-bit key (27 Morton bits below), so each group fills whole
// subtrees; k_karras stores each child's group code in the node's spare
// words (0: no single group) for the traversal's slab pruning.
#include "internal.hpp"
#include "predicates.hpp"

#include <algorithm>

namespace vacd {
namespace {

constexpr int kTile = 256;  // keys per sort tile (one per thread)

__device__ __forceinline__ unsigned f2ord(float f) {
  unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// bounds layout (unsigned ordered floats): [0..2] centroid min, [3..5]
// centroid max, [6..8] vertex min, [9..11] vertex max.
__global__ void k_bounds_init(unsigned *b) {
  int t = threadIdx.x;
  if (t < 12) b[t] = ((t / 3) % 2 == 0) ? 0xffffffffu : 0u;
}

__global__ __launch_bounds__(256) void k_bounds(const float *__restrict__ V,
                                                const uint32_t *__restrict__ T, int64_t nT,
                                                unsigned *__restrict__ bounds) {
  float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  float vmin[3] = {INFINITY, INFINITY, INFINITY}, vmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nT;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t i0 = T[3 * t], i1 = T[3 * t + 1], i2 = T[3 * t + 2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float a = V[3 * (size_t)i0 + k], b = V[3 * (size_t)i1 + k], c = V[3 * (size_t)i2 + k];
      const float cc = (a + b + c) * (1.0f / 3.0f);
      cmin[k] = fminf(cmin[k], cc);
      cmax[k] = fmaxf(cmax[k], cc);
      vmin[k] = fminf(vmin[k], fminf(a, fminf(b, c)));
      vmax[k] = fmaxf(vmax[k], fmaxf(a, fmaxf(b, c)));
    }
  }
  __shared__ float red[12][256];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = cmin[k];
    red[3 + k][threadIdx.x] = cmax[k];
    red[6 + k][threadIdx.x] = vmin[k];
    red[9 + k][threadIdx.x] = vmax[k];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int r = 0; r < 12; ++r) {
        const float a = red[r][threadIdx.x], b = red[r][threadIdx.x + s];
        red[r][threadIdx.x] = ((r / 3) % 2 == 0) ? fminf(a, b) : fmaxf(a, b);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 12) {
    const int r = threadIdx.x;
    const unsigned v = f2ord(red[r][0]);
    if ((r / 3) % 2 == 0)
      atomicMin(&bounds[r], v);
    else
      atomicMax(&bounds[r], v);
  }
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ __launch_bounds__(256) void k_morton(const float *__restrict__ V,
                                                const uint32_t *__restrict__ T, int64_t nT,
                                                const unsigned *__restrict__ bounds,
                                                uint32_t *__restrict__ keys,
                                                uint32_t *__restrict__ vals) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nT) return;
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float lo = ord2f(bounds[k]), hi = ord2f(bounds[3 + k]);
    const float a = V[3 * (size_t)T[3 * t] + k], b = V[3 * (size_t)T[3 * t + 1] + k],
                c = V[3 * (size_t)T[3 * t + 2] + k];
    const float cc = (a + b + c) * (1.0f / 3.0f);
    const float ext = hi - lo;
    float u = ext > 0.0f ? (cc - lo) / ext : 0.5f;
    u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
    q[k] = (uint32_t)u;
  }
  keys[t] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  vals[t] = (uint32_t)t;
}

// --- radix sort ---------------------------------------------------------
// hist layout: digit-major, hist[d * ntiles + tile].
__global__ __launch_bounds__(kTile) void k_hist(const uint32_t *__restrict__ keys, int64_t n,
                                                int shift, uint32_t *__restrict__ hist,
                                                int ntiles) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i = blockIdx.x * (int64_t)kTile + threadIdx.x;
  if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xffu], 1u);
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of 4096 consecutive entries data[base ..] (those below m)
// by a block of 1024 threads, 4 per thread: thread sums, a wave scan by
// shuffles, the 16 wave totals scanned in LDS; every entry gets carry plus
// the sum before it.  Returns the chunk's total in every thread.
__device__ __forceinline__ uint32_t scan_chunk(uint32_t *__restrict__ data, int64_t base,
                                               int64_t m, uint32_t carry, uint32_t *wsum) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t i0 = base + 4 * (int64_t)threadIdx.x;
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = i0 + k < m ? data[i0 + k] : 0u;
  const uint32_t tsum = v[0] + v[1] + v[2] + v[3];
  uint32_t inc = tsum;  // inclusive scan over the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off);
    if (lane >= off) inc += y;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < 16; ++w) {
      const uint32_t x = wsum[w];
      wsum[w] = run;
      run += x;
    }
    wsum[16] = run;
  }
  __syncthreads();
  uint32_t run = carry + wsum[wave] + (inc - tsum);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (i0 + k < m) data[i0 + k] = run;
    run += v[k];
  }
  const uint32_t total = wsum[16];
  __syncthreads();  // wsum is reused by the next chunk
  return total;
}

// Exclusive scan of m entries by one block, chunk after chunk (short arrays).
__global__ __launch_bounds__(1024) void k_scan_u32(uint32_t *__restrict__ data, int64_t m) {
  __shared__ uint32_t wsum[17];
  uint32_t carry = 0u;
  for (int64_t base = 0; base < m; base += 4096) carry += scan_chunk(data, base, m, carry, wsum);
}

// Long arrays: the total of every 4096-entry chunk, ...
__global__ __launch_bounds__(1024) void k_scan_sums(const uint32_t *__restrict__ data, int64_t m,
                                                    uint32_t *__restrict__ sums) {
  __shared__ uint32_t wsum[16];
  const int64_t i0 = (int64_t)blockIdx.x * 4096 + 4 * (int64_t)threadIdx.x;
  uint32_t t = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) t += i0 + k < m ? data[i0 + k] : 0u;
  for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < 16; ++w) s += wsum[w];
    sums[blockIdx.x] = s;
  }
}

// ... their exclusive scan (k_scan_u32), then every chunk scanned with its
// chunk's offset as the carry.
__global__ __launch_bounds__(1024) void k_scan_add(uint32_t *__restrict__ data, int64_t m,
                                                   const uint32_t *__restrict__ sums) {
  __shared__ uint32_t wsum[17];
  scan_chunk(data, (int64_t)blockIdx.x * 4096, m, sums[blockIdx.x], wsum);
}

// Exclusive scan of m entries on the context's stream (three launches for
// more than 16,384 entries: one block scanning 327,680 entries serially
// took 190 us).
void scan_launch(vacd_ctx *ctx, uint32_t *data, int64_t m) {
  hipStream_t st = ctx->stream;
  if (m <= 4 * 4096) {
    k_scan_u32<<<1, 1024, 0, st>>>(data, m);
    return;
  }
  const int64_t nblk = (m + 4095) / 4096;
  vacd::DevBuf &tb = ctx->buf[SLOT_SCANTMP];
  tb.ensure(4 * (size_t)nblk + 256);
  uint32_t *sums = tb.as<uint32_t>();
  k_scan_sums<<<(unsigned)nblk, 1024, 0, st>>>(data, m, sums);
  k_scan_u32<<<1, 1024, 0, st>>>(sums, nblk);
  k_scan_add<<<(unsigned)nblk, 1024, 0, st>>>(data, m, sums);
}

__global__ __launch_bounds__(kTile) void k_scatter(const uint32_t *__restrict__ kin,
                                                   const uint32_t *__restrict__ vin, int64_t n,
                                                   int shift, const uint32_t *__restrict__ hist,
                                                   int ntiles, uint32_t *__restrict__ kout,
                                                   uint32_t *__restrict__ vout) {
  constexpr int kWaves = kTile / 64;
  __shared__ uint32_t wcnt[kWaves][256];
  for (int t = threadIdx.x; t < kWaves * 256; t += kTile) (&wcnt[0][0])[t] = 0;
  __syncthreads();
  const int64_t i = blockIdx.x * (int64_t)kTile + threadIdx.x;
  const bool valid = i < n;
  const uint32_t key = valid ? kin[i] : 0u;
  const uint32_t digit = (key >> shift) & 0xffu;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // lanes of this wave holding the same digit
  uint64_t same = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint64_t m = __ballot(valid && ((digit >> b) & 1u));
    same &= ((digit >> b) & 1u) ? m : ~m;
  }
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t rank = (uint32_t)__popcll(same & below);
  if (valid && rank == 0) wcnt[wave][digit] = (uint32_t)__popcll(same);
  __syncthreads();
  if (valid) {
    uint32_t off = hist[(size_t)digit * ntiles + blockIdx.x] + rank;
    for (int w = 0; w < wave; ++w) off += wcnt[w][digit];
    kout[off] = key;
    vout[off] = vin[i];
  }
}

// --- hierarchy ------------------------------------------------------------
__device__ __forceinline__ int delta(const uint32_t *__restrict__ k, int64_t n, int64_t i,
                                     int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = k[i], b = k[j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((uint32_t)(i ^ j));
}

// gshift: keys carry a planar group code in bits gshift.. (0: no groups).
// larger0[i] = 1 when child 0 of node i has more leaves than child 1.
__global__ __launch_bounds__(256) void k_karras(const uint32_t *__restrict__ keys, int64_t n,
                                                Node *__restrict__ nodes,
                                                int32_t *__restrict__ parent_int,
                                                int32_t *__restrict__ parent_leaf,
                                                uint8_t *__restrict__ larger0,
                                                int32_t *__restrict__ rng, int gshift = 0) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int64_t lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int64_t l = 0;
  for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int64_t s = 0;
  int64_t div = 2;
  for (;;) {
    const int64_t t = (l + div - 1) / div;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
    div <<= 1;
  }
  const int64_t g = i + s * d + (d < 0 ? -1 : 0);
  const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
  int32_t c0, c1;
  if (lo == g) {
    c0 = ~(int32_t)g;
    parent_leaf[g] = (int32_t)i;
  } else {
    c0 = (int32_t)g;
    parent_int[g] = (int32_t)i;
  }
  if (hi == g + 1) {
    c1 = ~(int32_t)(g + 1);
    parent_leaf[g + 1] = (int32_t)i;
  } else {
    c1 = (int32_t)(g + 1);
    parent_int[g + 1] = (int32_t)i;
  }
  int g0 = 0, g1 = 0;  // group code of each child's key range, 0 unless one group
  if (gshift > 0) {
    const uint32_t a = keys[lo] >> gshift, b = keys[g] >> gshift, c = keys[g + 1] >> gshift,
                   d = keys[hi] >> gshift;
    g0 = a == b ? (int)a : 0;
    g1 = c == d ? (int)c : 0;
  }
  larger0[i] = (uint8_t)(g - lo + 1 > hi - g);
  rng[3 * i] = (int32_t)lo, rng[3 * i + 1] = (int32_t)g, rng[3 * i + 2] = (int32_t)hi;
  nodes[i].f3 = make_float4(__int_as_float(c0), __int_as_float(c1), __int_as_float(g0),
                            __int_as_float(g1));
  if (i == 0) parent_int[0] = -1;
}

// Child 0 of every node becomes the one with fewer leaves: the walks push
// child 1 and descend into child 0, so a walk's stack never holds more than
// log2(leaves) entries however deep the tree (binary_stack in visibility.hip).
// (Measured on clipped parts, this order also visits fewer nodes.)
__global__ __launch_bounds__(256) void k_small_first(Node *__restrict__ nodes, int64_t n_int,
                                                     const uint8_t *__restrict__ larger0) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n_int || !larger0[i]) return;
  Node nd = nodes[i];
  float *f = reinterpret_cast<float *>(&nd);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const float x = f[k];
    f[k] = f[6 + k], f[6 + k] = x;
  }
  nd.f3 = make_float4(nd.f3.y, nd.f3.x, nd.f3.w, nd.f3.z);
  nodes[i] = nd;
}

// Leaf-ordered triangles; c.w holds the triangle's planar group code (0
// none, g + 1 for group g) as integer bits.
__global__ __launch_bounds__(256) void k_leaves(const float *__restrict__ V,
                                                const uint32_t *__restrict__ T,
                                                const uint32_t *__restrict__ order, int64_t n,
                                                const uint8_t *__restrict__ grp,
                                                Tri *__restrict__ tris) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t t = order[j];
  const float *a = V + 3 * (size_t)T[3 * (size_t)t];
  const float *b = V + 3 * (size_t)T[3 * (size_t)t + 1];
  const float *c = V + 3 * (size_t)T[3 * (size_t)t + 2];
  Tri r;
  r.a = make_float4(a[0], a[1], a[2], __uint_as_float(t));
  r.b = make_float4(b[0], b[1], b[2], 0.0f);
  r.c = make_float4(c[0], c[1], c[2], __uint_as_float(grp ? (unsigned)grp[t] : 0u));
  tris[j] = r;
}

__device__ __forceinline__ void put_box(Node *nd, int slot, const float lo[3], const float hi[3]) {
  float *f = reinterpret_cast<float *>(nd);
  const int o = slot * 6;
  // plain stores; made visible by the release half of the counter update
  f[o + 0] = lo[0];
  f[o + 1] = lo[1];
  f[o + 2] = lo[2];
  f[o + 3] = hi[0];
  f[o + 4] = hi[1];
  f[o + 5] = hi[2];
}

// --- boxes ------------------------------------------------------------------
// Every internal node covers a contiguous range of leaves (Karras): child 0
// leaves lo..g, child 1 leaves g+1..hi.  Its two boxes are unions of leaf
// boxes over those ranges, read from a segment tree over the leaves (level
// L entry k = union of leaves k 2^L .. (k+1) 2^L - 1), built level by level
// in LDS, 8 levels per launch.  A range takes at most two entries per level.
// No atomics or device-wide fences: an atomic bottom-up climb crosses the
// XCDs' L2s at every level and measured ~1.2 ms on 327,680 triangles.
// Unions are exact (min / max), so the boxes equal those of any climb order.
// Level L has ceil(n / 2^L) entries; level L starts at entry sum of the
// sizes below it.  Entries are (lo, hi) float4 pairs.
__device__ __forceinline__ void seg_level(int64_t n, int L, int64_t &off, int64_t &size) {
  off = 0, size = n;
  for (int k = 0; k < L; ++k) off += size, size = (size + 1) >> 1;
}

__device__ __forceinline__ void box_union(float4 &lo, float4 &hi, const float4 &blo,
                                          const float4 &bhi) {
  lo.x = fminf(lo.x, blo.x), lo.y = fminf(lo.y, blo.y), lo.z = fminf(lo.z, blo.z);
  hi.x = fmaxf(hi.x, bhi.x), hi.y = fmaxf(hi.y, bhi.y), hi.z = fmaxf(hi.z, bhi.z);
}

// Levels L0 + 1 .. L0 + m from level L0 (L0 = 0: first the leaf boxes, the
// triangles' boxes or, with blo, the boxes blo/bhi[order[j]], widened by
// 2^-17 of the extent as the walks expect); 256 entries of level L0 per
// block.
__global__ __launch_bounds__(256) void k_seg_up(float4 *__restrict__ seg, int64_t n, int L0,
                                                int m, const Tri *__restrict__ tris,
                                                const unsigned *__restrict__ bounds,
                                                const float *__restrict__ blo,
                                                const float *__restrict__ bhi,
                                                const uint32_t *__restrict__ order) {
  __shared__ float4 slo[256], shi[256];
  const int t = threadIdx.x;
  int64_t off, size;
  seg_level(n, L0, off, size);
  const int64_t i = (int64_t)blockIdx.x * 256 + t;
  float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
  if (i < size) {
    if (L0 == 0) {
      float ext = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        ext = fmaxf(ext, fabsf(ord2f(bounds[6 + k])) + fabsf(ord2f(bounds[9 + k])));
      const float eps = ext * 7.62939453125e-06f + 1e-30f;  // 2^-17 of the extent
      if (blo) {
        const size_t r = order[i];
        lo = make_float4(blo[3 * r] - eps, blo[3 * r + 1] - eps, blo[3 * r + 2] - eps, 0.f);
        hi = make_float4(bhi[3 * r] + eps, bhi[3 * r + 1] + eps, bhi[3 * r + 2] + eps, 0.f);
      } else {
        const Tri tr = tris[i];
        lo = make_float4(fminf(tr.a.x, fminf(tr.b.x, tr.c.x)) - eps,
                         fminf(tr.a.y, fminf(tr.b.y, tr.c.y)) - eps,
                         fminf(tr.a.z, fminf(tr.b.z, tr.c.z)) - eps, 0.f);
        hi = make_float4(fmaxf(tr.a.x, fmaxf(tr.b.x, tr.c.x)) + eps,
                         fmaxf(tr.a.y, fmaxf(tr.b.y, tr.c.y)) + eps,
                         fmaxf(tr.a.z, fmaxf(tr.b.z, tr.c.z)) + eps, 0.f);
      }
      seg[2 * i] = lo, seg[2 * i + 1] = hi;
    } else {
      lo = seg[2 * (off + i)], hi = seg[2 * (off + i) + 1];
    }
  }
  slo[t] = lo, shi[t] = hi;
  for (int k = 1; k <= m; ++k) {
    const int64_t poff = off, psize = size;
    off += size, size = (size + 1) >> 1;
    const int cnt = 256 >> k;
    const int64_t idx = ((int64_t)blockIdx.x * 256 >> k) + t;
    __syncthreads();
    float4 nlo, nhi;
    const bool mine = t < cnt && idx < size;
    if (mine) {
      nlo = slo[2 * t], nhi = shi[2 * t];
      if (2 * idx + 1 < psize) box_union(nlo, nhi, slo[2 * t + 1], shi[2 * t + 1]);
    }
    (void)poff;
    __syncthreads();
    if (mine) {
      slo[t] = nlo, shi[t] = nhi;
      seg[2 * (off + idx)] = nlo, seg[2 * (off + idx) + 1] = nhi;
    }
  }
}

// Union of leaf boxes a..b (inclusive) from the segment tree.
__device__ __forceinline__ void seg_query(const float4 *__restrict__ seg, int64_t n, int64_t a,
                                          int64_t b, float4 &lo, float4 &hi) {
  lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
  hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
  int64_t l = a, r = b + 1, off = 0, size = n;
  while (l < r) {
    if (l & 1) {
      box_union(lo, hi, seg[2 * (off + l)], seg[2 * (off + l) + 1]);
      ++l;
    }
    if (r & 1) {
      --r;
      box_union(lo, hi, seg[2 * (off + r)], seg[2 * (off + r) + 1]);
    }
    l >>= 1, r >>= 1;
    off += size, size = (size + 1) >> 1;
  }
}

// The two child boxes of every internal node, and the root box; a single
// leaf gets node 0 with the leaf in slot 0 and an empty slot 1.
__global__ __launch_bounds__(256) void k_node_boxes(const float4 *__restrict__ seg, int64_t n,
                                                    const int32_t *__restrict__ rng,
                                                    Node *__restrict__ nodes,
                                                    float *__restrict__ root_box) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n == 1) {
    if (i == 0) {
      const float4 lo = seg[0], hi = seg[1];
      float *f = reinterpret_cast<float *>(&nodes[0]);
      f[0] = lo.x, f[1] = lo.y, f[2] = lo.z, f[3] = hi.x, f[4] = hi.y, f[5] = hi.z;
      for (int k = 6; k < 12; ++k) f[k] = kEmpty;
      nodes[0].f3 = make_float4(__int_as_float(~0), __int_as_float(~0), 0.0f, 0.0f);
      root_box[0] = lo.x, root_box[1] = lo.y, root_box[2] = lo.z;
      root_box[3] = hi.x, root_box[4] = hi.y, root_box[5] = hi.z;
    }
    return;
  }
  if (i >= n - 1) return;
  const int64_t lo_i = rng[3 * i], g = rng[3 * i + 1], hi_i = rng[3 * i + 2];
  float4 l0, h0, l1, h1;
  seg_query(seg, n, lo_i, g, l0, h0);
  seg_query(seg, n, g + 1, hi_i, l1, h1);
  float4 *f4 = reinterpret_cast<float4 *>(&nodes[i]);
  f4[0] = make_float4(l0.x, l0.y, l0.z, h0.x);
  f4[1] = make_float4(h0.y, h0.z, l1.x, l1.y);
  f4[2] = make_float4(l1.z, h1.x, h1.y, h1.z);
  if (i == 0) {
    root_box[0] = fminf(l0.x, l1.x), root_box[1] = fminf(l0.y, l1.y), root_box[2] = fminf(l0.z, l1.z);
    root_box[3] = fmaxf(h0.x, h1.x), root_box[4] = fmaxf(h0.y, h1.y), root_box[5] = fmaxf(h0.z, h1.z);
  }
}

// Segment tree and node boxes for n leaves (seg: 2 float4 per entry, at
// least 2 (2n + 40) float4).
void fill_boxes(hipStream_t st, float4 *seg, int64_t n, const int32_t *rng, Node *nodes,
                float *root, const Tri *tris, const unsigned *bounds, const float *blo,
                const float *bhi, const uint32_t *order) {
  int nlev = 1;
  for (int64_t sz = n; sz > 1; sz = (sz + 1) >> 1) ++nlev;
  int64_t size = n;
  int L0 = 0;
  do {  // the leaf level, then 8 levels per launch up to the single top entry
    const int m = std::min(8, nlev - 1 - L0);
    k_seg_up<<<(unsigned)((size + 255) / 256), 256, 0, st>>>(seg, n, L0, m, tris, bounds, blo, bhi,
                                                           order);
    for (int k = 0; k < m; ++k) size = (size + 1) >> 1;
    L0 += m;
  } while (L0 < nlev - 1);
  k_node_boxes<<<(unsigned)((std::max<int64_t>(n - 1, 1) + 255) / 256), 256, 0, st>>>(
      seg, n, rng, nodes, root);
}

// Tree depth (nodes on the longest leaf-to-root path).  The deepest leaf's
// sibling is a leaf too (were it an internal node, its leaves would lie
// deeper), so only nodes with two leaf children climb: about a third of the
// leaves' climbs, and one atomic per block.
__global__ __launch_bounds__(256) void k_depth(int64_t n, const Node *__restrict__ nodes,
                                               const int32_t *__restrict__ parent_int,
                                               int32_t *depth) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int32_t d = 0;
  if (n == 1) {
    d = i == 0 ? 2 : 0;
  } else if (i < n - 1) {
    const float4 f3 = nodes[i].f3;
    if (__float_as_int(f3.x) < 0 && __float_as_int(f3.y) < 0) {
      d = 2;  // the leaf and node i
      int32_t x = parent_int[i];
      while (x >= 0 && d < 1000) {
        ++d;
        x = parent_int[x];
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) d = max(d, __shfl_down(d, off));
  __shared__ int32_t wd[4];
  if ((threadIdx.x & 63) == 0) wd[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    d = max(max(wd[0], wd[1]), max(wd[2], wd[3]));
    if (d) atomicMax(depth, d);
  }
}

// Closedness (mod 2): every undirected edge must belong to an even number
// of triangles.  Each edge key lo<<32|hi is hashed by two independent
// 64-bit mixers (splitmix64 finalizer, D. Lemire / S. Vigna) and all hashes
// are XOR-ed: an edge met an even number of times cancels, so a closed mesh
// gives zero; an open one gives zero only if the hashes of its odd edges
// happen to XOR to zero in both halves (probability 2^-128 for a random
// function).  One pass over the triangles, no table.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_edge_xor(const uint32_t *__restrict__ T, int64_t nT,
                                                  unsigned long long *__restrict__ acc) {
  unsigned long long h1 = 0, h2 = 0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nT;
       t += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const uint32_t u = T[3 * t + e], v = T[3 * t + (e + 1) % 3];
      if (u == v) continue;
      const unsigned long long key =
          ((unsigned long long)(u < v ? u : v) << 32) | (unsigned long long)(u < v ? v : u);
      h1 ^= mix64(key + 0x9E3779B97F4A7C15ull);
      h2 ^= mix64((key ^ 0xD6E8FEB86659FD93ull) * 0xA0761D6478BD642Full + 0x2545F4914F6CDD1Dull);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    h1 ^= __shfl_down(h1, off);
    h2 ^= __shfl_down(h2, off);
  }
  // one atomic pair per block (atomics on one address from every wave of
  // the grid serialize)
  __shared__ unsigned long long wh[2][4];
  if ((threadIdx.x & 63) == 0) wh[0][threadIdx.x >> 6] = h1, wh[1][threadIdx.x >> 6] = h2;
  __syncthreads();
  if (threadIdx.x == 0) {
    h1 = wh[0][0] ^ wh[0][1] ^ wh[0][2] ^ wh[0][3];
    h2 = wh[1][0] ^ wh[1][1] ^ wh[1][2] ^ wh[1][3];
    if (h1 | h2) {
      atomicXor(&acc[0], h1);
      atomicXor(&acc[1], h2);
    }
  }
}

// Node4[x] = the (up to four) grandchildren of binary node x.
__global__ __launch_bounds__(256) void k_collapse4(const Node *__restrict__ nodes, int64_t n_int,
                                                   Node4 *__restrict__ out) {
  const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (x >= n_int) return;
  float lo[3][4], hi[3][4];
  int id[4];
  int k = 0;
  auto take = [&](const Node &nd, int slot) {
    const float *f = reinterpret_cast<const float *>(&nd);
    for (int a = 0; a < 3; ++a) lo[a][k] = f[6 * slot + a], hi[a][k] = f[6 * slot + 3 + a];
    id[k] = __float_as_int(slot == 0 ? nd.f3.x : nd.f3.y);
    ++k;
  };
  const Node nd = nodes[x];
  for (int s = 0; s < 2; ++s) {
    const int c = __float_as_int(s == 0 ? nd.f3.x : nd.f3.y);
    const float *f = reinterpret_cast<const float *>(&nd);
    if (c < 0 || f[6 * s] == kEmpty) {  // leaf or empty slot stays a child
      take(nd, s);
    } else {
      const Node cn = nodes[c];
      take(cn, 0);
      take(cn, 1);
    }
  }
  for (; k < 4; ++k) {
    for (int a = 0; a < 3; ++a) lo[a][k] = kEmpty, hi[a][k] = kEmpty;
    id[k] = ~0;
  }
  Node4 r;
  for (int a = 0; a < 3; ++a) {
    r.f[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
    r.f[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
  }
  r.f[6] = make_float4(__int_as_float(id[0]), __int_as_float(id[1]), __int_as_float(id[2]),
                       __int_as_float(id[3]));
  r.f[7] = make_float4(0.f, 0.f, 0.f, 0.f);
  out[x] = r;
}

// ---------------------------------------------------------- contact tree
constexpr int kMaxSlices = 64;
constexpr double kLenScale = 1048576.0;  // fixed point of the edge-length sum

__device__ __forceinline__ float dist3(const float *a, const float *b) {
  const float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
  return sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
}

__device__ __forceinline__ void load_tri(const float *V, const uint32_t *T, int64_t t,
                                         float p[3][3]) {
#pragma unroll
  for (int v = 0; v < 3; ++v)
#pragma unroll
    for (int k = 0; k < 3; ++k) p[v][k] = V[3 * (size_t)T[3 * t + v] + k];
}

// Sum over triangles of the longest edge, in fixed point so that the slice
// length (and the tree) do not depend on the order of the additions.
__global__ __launch_bounds__(256) void k_edge_sum(const float *__restrict__ V,
                                                  const uint32_t *__restrict__ T, int64_t nT,
                                                  unsigned long long *sum) {
  unsigned long long acc = 0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nT;
       t += (int64_t)gridDim.x * blockDim.x) {
    float p[3][3];
    load_tri(V, T, t, p);
    const float L = fmaxf(dist3(p[0], p[1]), fmaxf(dist3(p[1], p[2]), dist3(p[2], p[0])));
    acc += (unsigned long long)((double)L * kLenScale);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
  __shared__ unsigned long long wacc[4];  // one atomic per block
  if ((threadIdx.x & 63) == 0) wacc[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    acc = wacc[0] + wacc[1] + wacc[2] + wacc[3];
    if (acc) atomicAdd(sum, acc);
  }
}

// Slicing of a triangle: the apex is the vertex opposite its shortest edge,
// slice s is the band of the triangle between the lines parallel to that
// edge at s/k and (s+1)/k of the way to the apex, k = ceil(longest / h).
struct SliceFrame {
  float b0[3], b1[3], ap[3];
  int apex, k;
};

__device__ __forceinline__ SliceFrame slice_frame(const float *V, const uint32_t *T, int64_t t,
                                                  float h, int max_k) {
  float p[3][3];
  load_tri(V, T, t, p);
  const float e0 = dist3(p[0], p[1]), e1 = dist3(p[1], p[2]), e2 = dist3(p[2], p[0]);
  SliceFrame f;
  f.apex = 2;  // e0 = |p0 p1| is opposite p2, e1 opposite p0, e2 opposite p1
  float shortest = e0;
  if (e1 < shortest) shortest = e1, f.apex = 0;
  if (e2 < shortest) shortest = e2, f.apex = 1;
  const float longest = fmaxf(e0, fmaxf(e1, e2));
  int k = h > 0.0f ? (int)ceilf(longest / h) : 1;
  f.k = k < 1 ? 1 : (k > max_k ? max_k : k);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    f.ap[c] = p[f.apex][c];
    f.b0[c] = p[(f.apex + 1) % 3][c];
    f.b1[c] = p[(f.apex + 2) % 3][c];
  }
  return f;
}

__device__ __forceinline__ float slice_h(const unsigned long long *sum, int64_t nT) {
  return (float)(2.0 * ((double)*sum / kLenScale) / (double)nT);
}

// counts[t] = slices of triangle t; counts[nT] = 0 (the scan then ends
// with the total).
__global__ __launch_bounds__(256) void k_slice_count(const float *__restrict__ V,
                                                     const uint32_t *__restrict__ T, int64_t nT,
                                                     const unsigned long long *sum,
                                                     uint32_t *__restrict__ counts, int max_k) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t == 0) counts[nT] = 0;
  if (t >= nT) return;
  counts[t] = (uint32_t)slice_frame(V, T, t, slice_h(sum, nT), max_k).k;
}

__global__ __launch_bounds__(256) void k_slices(const float *__restrict__ V,
                                                const uint32_t *__restrict__ T, int64_t nT,
                                                const unsigned long long *sum,
                                                const uint32_t *__restrict__ offsets,
                                                float *__restrict__ blo, float *__restrict__ bhi,
                                                uint32_t *__restrict__ ref_tri,
                                                uint32_t *__restrict__ ref_meta, int max_k) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nT) return;
  const SliceFrame f = slice_frame(V, T, t, slice_h(sum, nT), max_k);
  const size_t base = offsets[t];
  for (int s = 0; s < f.k; ++s) {
    const float u0 = (float)s / (float)f.k;
    const float u1 = s + 1 == f.k ? 1.0f : (float)(s + 1) / (float)f.k;
    const size_t r = base + s;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float q0 = fmaf(f.ap[c] - f.b0[c], u0, f.b0[c]), q1 = fmaf(f.ap[c] - f.b1[c], u0, f.b1[c]);
      const float q2 = fmaf(f.ap[c] - f.b0[c], u1, f.b0[c]), q3 = fmaf(f.ap[c] - f.b1[c], u1, f.b1[c]);
      blo[3 * r + c] = fminf(fminf(q0, q1), fminf(q2, q3));
      bhi[3 * r + c] = fmaxf(fmaxf(q0, q1), fmaxf(q2, q3));
    }
    ref_tri[r] = (uint32_t)t;
    ref_meta[r] = (uint32_t)s | ((uint32_t)f.k << 8) | ((uint32_t)f.apex << 16);
  }
}

// bounds of box centroids and of the boxes (same layout as k_bounds)
__global__ __launch_bounds__(256) void k_bounds_boxes(const float *__restrict__ blo,
                                                      const float *__restrict__ bhi, int64_t n,
                                                      unsigned *__restrict__ bounds) {
  float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  float vmin[3] = {INFINITY, INFINITY, INFINITY}, vmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float lo = blo[3 * r + k], hi = bhi[3 * r + k], cc = 0.5f * (lo + hi);
      cmin[k] = fminf(cmin[k], cc);
      cmax[k] = fmaxf(cmax[k], cc);
      vmin[k] = fminf(vmin[k], lo);
      vmax[k] = fmaxf(vmax[k], hi);
    }
  }
  __shared__ float red[12][256];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = cmin[k];
    red[3 + k][threadIdx.x] = cmax[k];
    red[6 + k][threadIdx.x] = vmin[k];
    red[9 + k][threadIdx.x] = vmax[k];
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
#pragma unroll
      for (int r = 0; r < 12; ++r) {
        const float a = red[r][threadIdx.x], b = red[r][threadIdx.x + st];
        red[r][threadIdx.x] = ((r / 3) % 2 == 0) ? fminf(a, b) : fmaxf(a, b);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 12) {
    const int r = threadIdx.x;
    const unsigned v = f2ord(red[r][0]);
    if ((r / 3) % 2 == 0)
      atomicMin(&bounds[r], v);
    else
      atomicMax(&bounds[r], v);
  }
}

__global__ __launch_bounds__(256) void k_morton_boxes(const float *__restrict__ blo,
                                                      const float *__restrict__ bhi, int64_t n,
                                                      const unsigned *__restrict__ bounds,
                                                      uint32_t *__restrict__ keys,
                                                      uint32_t *__restrict__ vals,
                                                      const uint8_t *__restrict__ grp,
                                                      const uint32_t *__restrict__ ref_tri) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float lo = ord2f(bounds[k]), hi = ord2f(bounds[3 + k]);
    const float cc = 0.5f * (blo[3 * r + k] + bhi[3 * r + k]);
    const float ext = hi - lo;
    float u = ext > 0.0f ? (cc - lo) / ext : 0.5f;
    u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
    q[k] = (uint32_t)u;
  }
  const uint32_t code = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  keys[r] = grp ? ((uint32_t)grp[ref_tri[r]] << 27) | (code >> 3) : code;
  vals[r] = (uint32_t)r;
}

__global__ __launch_bounds__(256) void k_leaves_refs(const float *__restrict__ V,
                                                     const uint32_t *__restrict__ T,
                                                     const uint32_t *__restrict__ ref_tri,
                                                     const uint32_t *__restrict__ ref_meta,
                                                     const uint32_t *__restrict__ order, int64_t n,
                                                     Tri *__restrict__ tris) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t r = order[j];
  const uint32_t t = ref_tri[r];
  const float *a = V + 3 * (size_t)T[3 * (size_t)t];
  const float *b = V + 3 * (size_t)T[3 * (size_t)t + 1];
  const float *c = V + 3 * (size_t)T[3 * (size_t)t + 2];
  Tri x;
  x.a = make_float4(a[0], a[1], a[2], __uint_as_float(t));
  x.b = make_float4(b[0], b[1], b[2], __uint_as_float(ref_meta[r]));
  x.c = make_float4(c[0], c[1], c[2], 0.0f);
  tris[j] = x;
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }
// grid of the whole-mesh reductions (bounds, edge sums, closedness hash):
// each block ends in a few atomics on a few addresses, which serialize, so
// fewer, fuller blocks (1,024 blocks took ~100 us on 327,680 triangles)
constexpr unsigned kReduceGrid = 128;

// The contact tree of out (see vacd_bvh): slices of the long triangles and
// whole short ones as primitives with their own boxes, then the steps of the
// plain build.  Left empty (n_refs = 0) when no triangle is sliced, when the
// slices would exceed four per triangle, or when the tree is too deep for
// the traversal stacks.
void build_contact_tree(vacd_ctx *ctx, const float *d_V, const uint32_t *d_T, int64_t nT,
                        vacd_bvh *out, const uint8_t *d_grp) {
  out->n_refs = 0;
  out->depth_refs = 0;
  if (nT < 2) return;
  hipStream_t st = ctx->stream;
  vacd::DevBuf &cb = ctx->buf[SLOT_SLICE];
  cb.ensure(256 + 4 * (size_t)(nT + 1));
  unsigned long long *sum = cb.as<unsigned long long>();
  uint32_t *offsets = reinterpret_cast<uint32_t *>(cb.as<char>() + 256);
  VACD_HIP(hipMemsetAsync(sum, 0, 8, st));
  k_edge_sum<<<std::min(grid_for(nT, 256), kReduceGrid), 256, 0, st>>>(d_V, d_T, nT, sum);
  int max_k = kMaxSlices;
  uint32_t total = 0;
  for (;;) {
    k_slice_count<<<grid_for(nT, 256), 256, 0, st>>>(d_V, d_T, nT, sum, offsets, max_k);
    scan_launch(ctx, offsets, nT + 1);
    VACD_HIP(hipGetLastError());
    VACD_HIP(hipMemcpyAsync(&total, offsets + nT, 4, hipMemcpyDeviceToHost, st));
    VACD_HIP(hipStreamSynchronize(st));
    // with planar groups the tree is built even without slices (for the
    // groups' subtrees); too many slices fall back to whole triangles
    if ((int64_t)total > 4 * nT && d_grp && max_k > 1) {
      max_k = 1;
      continue;
    }
    break;
  }
  if ((int64_t)total > 4 * nT || ((int64_t)total <= nT && !d_grp)) return;
  const int64_t R = total;
  vacd::DevBuf &rb = ctx->buf[SLOT_REFS];
  rb.ensure(32 * (size_t)R + 1024);
  float *blo = rb.as<float>();
  float *bhi = blo + 3 * R;
  uint32_t *ref_tri = reinterpret_cast<uint32_t *>(bhi + 3 * R);
  uint32_t *ref_meta = ref_tri + R;
  k_slices<<<grid_for(nT, 256), 256, 0, st>>>(d_V, d_T, nT, sum, offsets, blo, bhi, ref_tri,
                                              ref_meta, max_k);
  const int64_t ntiles = (R + kTile - 1) / kTile;
  const size_t hist_n = (size_t)256 * ntiles;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  const size_t o_bounds = take(16 * 4), o_depth = take(4), o_k0 = take(4 * R), o_k1 = take(4 * R),
               o_v0 = take(4 * R), o_v1 = take(4 * R), o_hist = take(4 * hist_n),
               o_pint = take(4 * (R - 1)), o_pleaf = take(4 * R), o_rng = take(12 * (R - 1)),
               o_root = take(6 * 4), o_larger = take(R), o_seg = take(32 * (2 * (size_t)R + 40));
  vacd::DevBuf &scratch = ctx->buf[SLOT_SORT];
  scratch.ensure(off);
  char *base = scratch.as<char>();
  unsigned *bounds = reinterpret_cast<unsigned *>(base + o_bounds);
  int32_t *depth = reinterpret_cast<int32_t *>(base + o_depth);
  uint32_t *k0 = reinterpret_cast<uint32_t *>(base + o_k0);
  uint32_t *k1 = reinterpret_cast<uint32_t *>(base + o_k1);
  uint32_t *v0 = reinterpret_cast<uint32_t *>(base + o_v0);
  uint32_t *v1 = reinterpret_cast<uint32_t *>(base + o_v1);
  uint32_t *hist = reinterpret_cast<uint32_t *>(base + o_hist);
  int32_t *pint = reinterpret_cast<int32_t *>(base + o_pint);
  int32_t *pleaf = reinterpret_cast<int32_t *>(base + o_pleaf);
  int32_t *rng = reinterpret_cast<int32_t *>(base + o_rng);
  float *root = reinterpret_cast<float *>(base + o_root);
  uint8_t *larger0 = reinterpret_cast<uint8_t *>(base + o_larger);
  float4 *seg = reinterpret_cast<float4 *>(base + o_seg);
  out->rnodes.ensure(sizeof(Node) * (size_t)(R - 1));
  out->rtris.ensure(sizeof(Tri) * (size_t)R);
  Node *nodes = out->rnodes.as<Node>();
  Tri *tris = out->rtris.as<Tri>();
  k_bounds_init<<<1, 64, 0, st>>>(bounds);
  k_bounds_boxes<<<std::min(grid_for(R, 256), kReduceGrid), 256, 0, st>>>(blo, bhi, R, bounds);
  k_morton_boxes<<<grid_for(R, 256), 256, 0, st>>>(blo, bhi, R, bounds, k0, v0, d_grp, ref_tri);
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 8 * pass;
    k_hist<<<(unsigned)ntiles, kTile, 0, st>>>(k0, R, shift, hist, (int)ntiles);
    scan_launch(ctx, hist, (int64_t)hist_n);
    k_scatter<<<(unsigned)ntiles, kTile, 0, st>>>(k0, v0, R, shift, hist, (int)ntiles, k1, v1);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  VACD_HIP(hipMemsetAsync(depth, 0, 4, st));
  VACD_HIP(hipMemsetAsync(nodes, 0, sizeof(Node) * (size_t)(R - 1), st));
  k_karras<<<grid_for(R - 1, 256), 256, 0, st>>>(k0, R, nodes, pint, pleaf, larger0, rng,
                                                 d_grp ? 27 : 0);
  k_leaves_refs<<<grid_for(R, 256), 256, 0, st>>>(d_V, d_T, ref_tri, ref_meta, v0, R, tris);
  fill_boxes(st, seg, R, rng, nodes, root, tris, bounds, blo, bhi, v0);
  k_depth<<<grid_for(R - 1, 256), 256, 0, st>>>(R, nodes, pint, depth);
  k_small_first<<<grid_for(R - 1, 256), 256, 0, st>>>(nodes, R - 1, larger0);
  VACD_HIP(hipGetLastError());
  int32_t h_depth = 0;
  VACD_HIP(hipMemcpyAsync(&h_depth, depth, 4, hipMemcpyDeviceToHost, st));
  VACD_HIP(hipStreamSynchronize(st));
  if (h_depth > 64) {  // the walks' stacks hold 64 entries
    out->rnodes.release();
    out->rtris.release();
    return;
  }
  out->n_refs = R;
  out->depth_refs = h_depth;
  out->small_first_refs = true;
}


}  // namespace

void scan_u32(vacd_ctx *ctx, uint32_t *d_data, int64_t m) {
  if (m <= 0) return;
  scan_launch(ctx, d_data, m);
  VACD_HIP(hipGetLastError());
}

void build_lbvh(vacd_ctx *ctx, const float *d_V, int64_t nV, const uint32_t *d_T, int64_t nT,
                vacd_bvh *out, const uint8_t *d_grp) {
  (void)nV;
  VACD_REQUIRE(nT >= 1 && nT < (int64_t(1) << 30), VACD_ERR_LIMIT,
               "BVH needs 1 <= n_tris < 2^30");
  hipStream_t st = ctx->stream;
  const int64_t ntiles = (nT + kTile - 1) / kTile;
  const int64_t n_int = nT > 1 ? nT - 1 : 1;
  // scratch: bounds(12 u32) | depth | keys x2 | vals x2 | hist | parents | ranges | root box |
  // closedness hash | child order | segment tree
  const size_t hist_n = (size_t)256 * ntiles;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  const size_t o_bounds = take(16 * 4), o_depth = take(4), o_k0 = take(4 * nT),
               o_k1 = take(4 * nT), o_v0 = take(4 * nT), o_v1 = take(4 * nT),
               o_hist = take(4 * hist_n), o_pint = take(4 * n_int), o_pleaf = take(4 * nT),
               o_rng = take(12 * n_int), o_root = take(6 * 4), o_odd = take(16),
               o_larger = take(n_int), o_seg = take(32 * (2 * (size_t)nT + 40));
  vacd::DevBuf &scratch = ctx->buf[SLOT_SORT];
  scratch.ensure(off);
  char *base = scratch.as<char>();
  unsigned *bounds = reinterpret_cast<unsigned *>(base + o_bounds);
  int32_t *depth = reinterpret_cast<int32_t *>(base + o_depth);
  uint32_t *k0 = reinterpret_cast<uint32_t *>(base + o_k0);
  uint32_t *k1 = reinterpret_cast<uint32_t *>(base + o_k1);
  uint32_t *v0 = reinterpret_cast<uint32_t *>(base + o_v0);
  uint32_t *v1 = reinterpret_cast<uint32_t *>(base + o_v1);
  uint32_t *hist = reinterpret_cast<uint32_t *>(base + o_hist);
  int32_t *pint = reinterpret_cast<int32_t *>(base + o_pint);
  int32_t *pleaf = reinterpret_cast<int32_t *>(base + o_pleaf);
  int32_t *rng = reinterpret_cast<int32_t *>(base + o_rng);
  float *root = reinterpret_cast<float *>(base + o_root);
  float4 *seg = reinterpret_cast<float4 *>(base + o_seg);
  int32_t *odd = reinterpret_cast<int32_t *>(base + o_odd);
  uint8_t *larger0 = reinterpret_cast<uint8_t *>(base + o_larger);

  out->n_tris = nT;
  out->nodes.ensure(sizeof(Node) * (size_t)n_int);
  out->tris.ensure(sizeof(Tri) * (size_t)nT);
  Node *nodes = out->nodes.as<Node>();
  Tri *tris = out->tris.as<Tri>();

  timed_begin(ctx, TIMED_BVH);
  k_bounds_init<<<1, 64, 0, st>>>(bounds);
  const unsigned gb = std::min(grid_for(nT, 256), kReduceGrid);
  k_bounds<<<gb, 256, 0, st>>>(d_V, d_T, nT, bounds);
  k_morton<<<grid_for(nT, 256), 256, 0, st>>>(d_V, d_T, nT, bounds, k0, v0);
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 8 * pass;
    k_hist<<<(unsigned)ntiles, kTile, 0, st>>>(k0, nT, shift, hist, (int)ntiles);
    scan_launch(ctx, hist, (int64_t)hist_n);
    k_scatter<<<(unsigned)ntiles, kTile, 0, st>>>(k0, v0, nT, shift, hist, (int)ntiles, k1, v1);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  VACD_HIP(hipGetLastError());
  VACD_HIP(hipMemsetAsync(depth, 0, 4, st));
  if (nT > 1) {
    VACD_HIP(hipMemsetAsync(nodes, 0, sizeof(Node) * (size_t)n_int, st));
    k_karras<<<grid_for(nT - 1, 256), 256, 0, st>>>(k0, nT, nodes, pint, pleaf, larger0, rng);
  }
  k_leaves<<<grid_for(nT, 256), 256, 0, st>>>(d_V, d_T, v0, nT, d_grp, tris);
  fill_boxes(st, seg, nT, rng, nodes, root, tris, bounds, nullptr, nullptr, nullptr);
  k_depth<<<grid_for(std::max<int64_t>(nT - 1, 1), 256), 256, 0, st>>>(nT, nodes, pint, depth);
  if (nT > 1) k_small_first<<<grid_for(nT - 1, 256), 256, 0, st>>>(nodes, nT - 1, larger0);
  out->nodes4.ensure(sizeof(Node4) * (size_t)n_int);
  k_collapse4<<<grid_for(n_int, 256), 256, 0, st>>>(nodes, n_int, out->nodes4.as<Node4>());
  VACD_HIP(hipGetLastError());
  timed_end(ctx, TIMED_BVH);
  VACD_HIP(hipMemsetAsync(odd, 0, 16, st));
  unsigned long long *edge_hash = reinterpret_cast<unsigned long long *>(odd);
  k_edge_xor<<<std::min(grid_for(nT, 256), kReduceGrid), 256, 0, st>>>(d_T, nT, edge_hash);
  VACD_HIP(hipGetLastError());
  int32_t h_depth = 0;
  unsigned long long h_hash[2] = {1, 1};
  float h_root[6];
  VACD_HIP(hipMemcpyAsync(h_hash, edge_hash, 16, hipMemcpyDeviceToHost, st));
  VACD_HIP(hipMemcpyAsync(&h_depth, depth, 4, hipMemcpyDeviceToHost, st));
  VACD_HIP(hipMemcpyAsync(h_root, root, 24, hipMemcpyDeviceToHost, st));
  VACD_HIP(hipStreamSynchronize(st));
  out->depth = h_depth;
  out->small_first = true;
  out->closed = h_hash[0] == 0 && h_hash[1] == 0;
  for (int k = 0; k < 6; ++k) out->bounds[k] = h_root[k];
  build_contact_tree(ctx, d_V, d_T, nT, out, d_grp);
}

}  // namespace vacd
