// Learned plane proposal for gfx950: a set-encoder MLP with BF16 weights on
// MFMA and FP32 accumulation, with hand-written forward, backward and Adam.
//
// Model (per part; D = 256, E = 512 sampled visibility edges, P planes):
//   edge features x_e = (u, v, |v - u|) in R^7, zero-padded to 8;
//   encoder   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3)
//             (shared over the edges), z = max over the edges of h3;
//   head      g1 = relu(Wz z + Wp q + c1) for plane q = (n, d): the pooled part
//             Wz z + c1 is computed once per part, Wp q per plane;
//             g2 = relu(V2 g1 + c2);  y = v3 . g2 + c3, the predicted cut length
//             of the plane relative to the largest one among the part's planes.
// Every layer product runs through k_gemm: C = op(A) op(B) with bf16 operands,
// v_mfma_f32_32x32x8_bf16 tiles and FP32 accumulation; a block computes a
// 64 x 64 tile with four waves (32 x 32 each), K staged through LDS 16 at a
// time.  Activations and gradients that feed products are kept in bf16,
// weights in FP32 master copies with bf16 shadows for the products.
#include "internal.hpp"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace vacd {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  // round to nearest even (NaN is not produced by these layers)
  const unsigned u = __float_as_uint(f);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

constexpr int kTM = 64, kTN = 64, kTK = 16;

// C[m][n] (+)= sum_k A(m, k) B(k, n) (+ bias[n]), optionally relu'd.
// A(m, k) = TA ? A[k * lda + m] : A[m * lda + k]; B(k, n) = TB ? B[n * ldb + k]
// : B[k * ldb + n].  Outputs: Cf (fp32, nullable) and Cb (bf16, nullable);
// ACC adds into Cf instead of storing.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const unsigned short *__restrict__ A,
                                              int lda, const unsigned short *__restrict__ B, int ldb,
                                              const float *__restrict__ bias, int relu, int acc,
                                              float *__restrict__ Cf, unsigned short *__restrict__ Cb,
                                              int ldc) {
  __shared__ unsigned short As[kTM][kTK + 4];  // As[m][k]
  __shared__ unsigned short Bs[kTN][kTK + 4];  // Bs[n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 c;
#pragma unroll
  for (int r = 0; r < 16; ++r) c[r] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += kTK) {
    // 64 x 16 elements of each operand, 4 per thread
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int e = tid * 4 + t;
      const int r = e / kTK, k = e % kTK;
      const int gm = m0 + r, gn = n0 + r, gk = k0 + k;
      unsigned short av = 0, bv = 0;
      if (gm < M && gk < K) av = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      if (gn < N && gk < K) bv = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
      As[r][k] = av;
      Bs[r][k] = bv;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 8) {
      // lane l: row / column l % 32, k = kk + 4 * (l / 32) .. + 3
      const int rc = lane & 31, kb = kk + 4 * (lane >> 5);
      bf16x4 a, b;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = (short)As[wm + rc][kb + i];
        b[i] = (short)Bs[wn + rc][kb + i];
      }
      c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
    }
    __syncthreads();
  }
  // D layout: element r of lane l is row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
  const int gn = n0 + wn + (lane & 31);
  if (gn >= N) return;
  const float bn = bias ? bias[gn] : 0.0f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int gm = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (gm >= M) continue;
    float v = c[r] + bn;
    const size_t o = (size_t)gm * ldc + gn;
    if (acc) v += Cf[o];
    if (relu) v = fmaxf(v, 0.0f);
    if (Cf) Cf[o] = v;
    if (Cb) Cb[o] = f2bf(v);
  }
}

__global__ void k_f2bf(const float *__restrict__ x, int64_t n, unsigned short *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

constexpr int kD = 256;  // hidden width

// edge features (R rows of 7 floats) -> bf16 rows of 8 (zero-padded)
__global__ void k_edge_in(const float *__restrict__ x, int64_t R, unsigned short *__restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= R * 8) return;
  const int64_t r = i >> 3;
  const int c = (int)(i & 7);
  y[i] = c < 7 ? f2bf(x[r * 7 + c]) : (unsigned short)0;
}

// z[b][f] = max over part b's E rows of h[.][f]; arg = that row (first max)
__global__ void k_maxpool(const unsigned short *__restrict__ h, int E, float *__restrict__ z,
                          int *__restrict__ arg) {
  const int b = blockIdx.x, f = threadIdx.x;
  float best = -1e30f;
  int bi = 0;
  for (int e = 0; e < E; ++e) {
    const float v = bf2f(h[((size_t)b * E + e) * kD + f]);
    if (v > best) best = v, bi = e;
  }
  z[b * kD + f] = best;
  arg[b * kD + f] = bi;
}

// g1[p][f] = relu(u[b][f] + Wp[f] . q[p]) for the P planes of each part b
__global__ void k_head1(const float *__restrict__ u, const float *__restrict__ Wp,
                        const float *__restrict__ q, int64_t BP, int P,
                        unsigned short *__restrict__ g1) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= BP * kD) return;
  const int64_t p = i / kD;
  const int f = (int)(i % kD);
  const int64_t b = p / P;
  const float *w = Wp + 4 * f;
  const float *x = q + 4 * p;
  const float v = u[b * kD + f] + w[0] * x[0] + w[1] * x[1] + w[2] * x[2] + w[3] * x[3];
  g1[i] = f2bf(fmaxf(v, 0.0f));
}

// out[p] = v3 . g2[p] + c3, a wave per plane
__global__ void k_head_out(const unsigned short *__restrict__ g2, const float *__restrict__ v3,
                           const float *__restrict__ c3, int64_t BP, float *__restrict__ out) {
  const int64_t p = blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= BP) return;
  float s = 0.0f;
#pragma unroll
  for (int t = 0; t < 4; ++t) s += v3[lane + 64 * t] * bf2f(g2[p * kD + lane + 64 * t]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) out[p] = s + c3[0];
}

// Inference head, fused: out[p] = v3 . bf16(relu(V2 g1[p] + c2)) + c3 with
// g1[p] = bf16(relu(u[b] + Wp q[p])), every product on MFMA and neither g1
// nor g2 in memory (the unfused training path writes both).  A wave takes
// RT tiles of 32 planes at a time; v_mfma_f32_32x32x16_bf16 operands: lane l
// holds row (A) or column (B) l % 32 and k = 8 (l / 32) .. +7; result element
// r of lane l is row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.  The
// products run transposed (hidden units as rows, planes as columns), so that
// what one MFMA leaves in a lane -- 16 hidden units of the lane's plane -- is
// the B operand of the next one for two k-steps, once the hidden units are
// renumbered; the V2 columns and v3 are stored in that order (head_perm).
//   g1: eight MFMAs per tile, one per 32 hidden units: A = (Wp_hi | Wp_lo)
//       in lanes 0-31, (Wp_hi | u_hi u_lo of part c, u_hi u_lo of part c + 1)
//       in lanes 32-63, B = (q_hi | q_hi) and (q_lo | s0 s0 s1 s1), where s0
//       / s1 say whether the plane lies in part c / c + 1 (c: the part of the
//       tile's first plane; with parts of 32 planes or more a tile spans at
//       most two).  So Wp q + u with bf16 splits of q, Wp and u; only the lo x
//       lo term of Wp q (about 2^-16 of it) and u's third bf16 term (2^-16 of
//       u) are left out, far below the bf16 rounding of g1 that follows, as in
//       the unfused path, with relu.  Parts of fewer than 32 planes (SMALLP)
//       add the u of a later part in float32.
//   g2: 16 k-steps per 32 outputs with V2 from LDS (rows padded by 16 bytes
//       against bank conflicts) shared by the RT tiles, + c2, relu, bf16.
//   y:  two MFMAs per 32 outputs with A = (v3_hi, v3_lo) in rows 0 and 1:
//       y = row 0 + row 1 + c3.  Products of bf16 values are exact in the
//       float32 accumulators.
// relu of a bf16 pair is a signed 16-bit max with 0 (relu_bf16x2).  One
// block of NW waves per CU (V2 fills most of its LDS); each wave takes a
// contiguous run of tiles.  64 parts x 16,384 planes: 0.30 ms with g1 and
// the v3 dot on the VALU, 0.134 ms this way (the whole head; one tile per
// step measured 0.136 ms at three waves per SIMD, two tiles 0.134 ms at two).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
constexpr int kHeadFusedWaves = 8;  // the head's
constexpr int kHeadTiles = 2;       // its 32-plane tiles per wave and step
constexpr int kV2Stride = kD + 8;  // bf16 per padded LDS row

// Position of hidden unit c in the order of the head's B operands: for c =
// 32 t + 16 b + 8 j + 4 h + r, the lanes of half h take units 32 t + 16 b +
// 8 h + [0, 8) of k-step 2 t + b, which hold 4 j + r.
__host__ __device__ __forceinline__ int head_perm(int c) {
  return (c & ~31) | (c & 16) | ((c & 4) << 1) | ((c & 8) >> 1) | (c & 3);
}

// bf16 pair (a, b) of relu(x), relu(y): round both (v_cvt_pk_bf16_f32), then
// max with 0 as signed 16-bit integers -- a negative bf16 (or -0) has the sign
// bit set, so this is relu, and relu commutes with the rounding.
typedef short short2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned relu_bf16x2(float x, float y) {
  // one packed conversion (element-wise casts cost two and a permute)
  typedef float f32x2_t __attribute__((ext_vector_type(2)));
  const bf16x2_t v = __builtin_convertvector((f32x2_t){x, y}, bf16x2_t);
  const short2_t m = __builtin_elementwise_max(__builtin_bit_cast(short2_t, v), (short2_t)(0));
  return __builtin_bit_cast(unsigned, m);
}

__device__ __forceinline__ void split2(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <int NW, int RT, bool SMALLP>
__global__ __launch_bounds__(64 * NW) void k_head_fused(
    const float *__restrict__ u, const float *__restrict__ Wp, const float *__restrict__ q,
    int64_t BP, int P, const unsigned short *__restrict__ V2, const float *__restrict__ c2,
    const float *__restrict__ v3, const float *__restrict__ c3, float *__restrict__ out) {
  __shared__ __attribute__((aligned(16))) unsigned short sV[kD * kV2Stride];
  __shared__ __attribute__((aligned(16))) __bf16 sWp[kD * 8];  // per unit: Wp_hi (4), Wp_lo (4)
  __shared__ __attribute__((aligned(16))) __bf16 sV3[2 * kD];
  __shared__ __attribute__((aligned(16))) float sc2[kD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // V2 rows into LDS with permuted columns: 8 consecutive columns (aligned)
  // are two runs of 4 in the permuted order.  Every load is issued before
  // the first store (a load-store loop waits out one memory latency per
  // iteration).
  constexpr int kChunks = kD * kD / 8, kPer = (kChunks + 64 * NW - 1) / (64 * NW);
  {
    uint4 v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int c = tid + j * 64 * NW;
      if (c < kChunks) v[j] = reinterpret_cast<const uint4 *>(V2)[c];
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int c = tid + j * 64 * NW;
      if (c >= kChunks) break;
      const int n = c / (kD / 8), k = 8 * (c % (kD / 8));
      *reinterpret_cast<uint2 *>(&sV[n * kV2Stride + head_perm(k)]) = make_uint2(v[j].x, v[j].y);
      *reinterpret_cast<uint2 *>(&sV[n * kV2Stride + head_perm(k + 4)]) = make_uint2(v[j].z, v[j].w);
    }
  }
  for (int k = tid; k < kD; k += 64 * NW) {
    __bf16 hi, lo;
    split2(v3[k], hi, lo);
    sV3[head_perm(k)] = hi;
    sV3[kD + head_perm(k)] = lo;
    sc2[k] = c2[k];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      split2(Wp[4 * k + c], hi, lo);
      sWp[8 * k + c] = hi;
      sWp[8 * k + 4 + c] = lo;
    }
  }
  __syncthreads();
  const float bias3 = c3[0];
  const int half = lane >> 5, col = lane & 31;
  uint2 au[8];  // per 32 units: u_hi, u_lo of part cur, u_hi, u_lo of part cur + 1
  // the part whose u terms au holds and the first row after it; a wave's
  // tiles ascend, so the part advances by additions, with no division
  int64_t cur = -1, cur_end = 0;
  // each wave takes a contiguous run of tiles, so its part changes rarely
  const int64_t tiles = (BP + 31) / 32;
  const int64_t nw = (int64_t)gridDim.x * NW, wid = (int64_t)blockIdx.x * NW + wave;
  const int64_t run = (tiles + nw - 1) / nw;
  const int64_t t_end = std::min<int64_t>(tiles, (wid + 1) * run);
  for (int64_t tile = wid * run; tile < t_end; tile += RT) {
    // RT tiles at a time share every V2 fragment read from LDS
    int64_t row[RT];
    bool valid[RT];
    // g1 of each tile: the B operands of the main product, k-step 2 t + b in
    // b2[x][2 t + b]
    bf16x8_t b2[RT][16];
#pragma unroll
    for (int x = 0; x < RT; ++x) {
      row[x] = (tile + x) * 32 + col;  // this lane's plane
      valid[x] = tile + x < t_end && row[x] < BP;
      const int64_t r0 = std::min<int64_t>(tile + x, tiles - 1) * 32;
      if (r0 >= cur_end) {  // wave-uniform: the u terms of parts p0 and p0 + 1
        int64_t p0 = cur < 0 ? r0 / P : cur, e = (p0 + 1) * P;
        while (r0 >= e) ++p0, e += P;
        cur = p0;
        cur_end = e;
        const int64_t p1 = std::min<int64_t>(p0 + 1, (BP - 1) / P);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          __bf16 h0, l0, h1, l1;
          split2(u[p0 * kD + 32 * t + col], h0, l0);
          split2(u[p1 * kD + 32 * t + col], h1, l1);
          bf16x2_t lo, hi;
          lo[0] = h0, lo[1] = l0, hi[0] = h1, hi[1] = l1;
          au[t] = make_uint2(__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi));
        }
      }
      // parts of at least 32 planes: a tile's rows lie in parts cur and cur + 1
      const int64_t part =
          !valid[x] ? cur : SMALLP ? row[x] / P : row[x] >= cur_end ? cur + 1 : cur;
      const float4 qv = valid[x] ? *reinterpret_cast<const float4 *>(q + 4 * row[x])
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
      // B operand of g1: (q_hi | q_hi) in lanes 0-31, (q_lo | s0 s0 s1 s1) in
      // lanes 32-63, s0 / s1: the plane lies in part cur / cur + 1 (with
      // parts of at least 32 planes a tile spans at most two)
      bf16x8_t b1;
      {
        const float qa[4] = {qv.x, qv.y, qv.z, qv.w};
        const __bf16 s0 = (__bf16)(part == cur ? 1.0f : 0.0f);
        const __bf16 s1 = (__bf16)(part == cur + 1 ? 1.0f : 0.0f);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          __bf16 hi, lo;
          split2(qa[c], hi, lo);
          b1[c] = half ? lo : hi;
          b1[4 + c] = half ? (c < 2 ? s0 : s1) : hi;
        }
      }
      const bool other = SMALLP && __ballot(part != cur && part != cur + 1) != 0;
      // A: (Wp_hi | Wp_lo) in lanes 0-31, (Wp_hi | u terms) in lanes 32-63
      auto g1_product = [&](int t) {
        uint4 aw = *reinterpret_cast<const uint4 *>(&sWp[8 * (32 * t + col)]);
        if (half) aw.z = au[t].x, aw.w = au[t].y;
        f32x16 d;
#pragma unroll
        for (int r = 0; r < 16; ++r) d[r] = 0.0f;
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, aw), b1, d, 0,
                                                       0, 0);
      };
      auto g1_finish = [&](int t, f32x16 d) {
        if (SMALLP && other && part != cur && part != cur + 1) {
          // parts of fewer than 32 planes: a later part's u added directly
#pragma unroll
          for (int r = 0; r < 16; ++r)
            d[r] += u[part * kD + 32 * t + 8 * (r >> 2) + 4 * half + (r & 3)];
        }
        uint4 g0, g1;
        g0.x = relu_bf16x2(d[0], d[1]), g0.y = relu_bf16x2(d[2], d[3]);
        g0.z = relu_bf16x2(d[4], d[5]), g0.w = relu_bf16x2(d[6], d[7]);
        g1.x = relu_bf16x2(d[8], d[9]), g1.y = relu_bf16x2(d[10], d[11]);
        g1.z = relu_bf16x2(d[12], d[13]), g1.w = relu_bf16x2(d[14], d[15]);
        b2[x][2 * t] = __builtin_bit_cast(bf16x8_t, g0);
        b2[x][2 * t + 1] = __builtin_bit_cast(bf16x8_t, g1);
      };
      // product t + 1 is issued before the VALU finishes product t, so the
      // matrix core runs while the activations are packed; two 16-register
      // results live at a time (the scheduler would otherwise start all
      // eight products first and run out of registers)
      f32x16 dprev = g1_product(0);
#pragma unroll
      for (int t = 1; t < 8; ++t) {
        const f32x16 dcur = g1_product(t);
        g1_finish(t - 1, dprev);
        __builtin_amdgcn_sched_barrier(0);
        dprev = dcur;
      }
      g1_finish(7, dprev);
    }
    f32x16 y[RT];
#pragma unroll
    for (int x = 0; x < RT; ++x)
#pragma unroll
      for (int r = 0; r < 16; ++r) y[x][r] = 0.0f;
#pragma unroll 1
    for (int s = 0; s < 8; ++s) {
      // the bias c2 enters as the products' starting value (element r of
      // lane l: unit 32 s + 8 (r / 4) + 4 half + r % 4)
      f32x16 cf;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const float4 c = *reinterpret_cast<const float4 *>(&sc2[32 * s + 8 * g + 4 * half]);
        cf[4 * g] = c.x, cf[4 * g + 1] = c.y, cf[4 * g + 2] = c.z, cf[4 * g + 3] = c.w;
      }
      f32x16 d[RT];
      const unsigned short *vrow = &sV[(32 * s + col) * kV2Stride + 8 * half];
      // a wave issuing its matrix products runs at raised priority, so the
      // SIMD's arbiter favours whichever wave keeps the matrix core busy
      // while the other packs activations (3 % on 64 parts)
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) {
        const bf16x8_t a = *reinterpret_cast<const bf16x8_t *>(vrow + 16 * kk);
#pragma unroll
        for (int x = 0; x < RT; ++x)
          d[x] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b2[x][kk], kk ? d[x] : cf, 0, 0, 0);
      }
      __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        bf16x8_t a;
        if (col < 2) {
          a = *reinterpret_cast<const bf16x8_t *>(&sV3[col * kD + 32 * s + 16 * b + 8 * half]);
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) a[i] = (__bf16)0.0f;
        }
        const int e = 8 * b;
#pragma unroll
        for (int x = 0; x < RT; ++x) {
          uint4 g;
          g.x = relu_bf16x2(d[x][e], d[x][e + 1]);
          g.y = relu_bf16x2(d[x][e + 2], d[x][e + 3]);
          g.z = relu_bf16x2(d[x][e + 4], d[x][e + 5]);
          g.w = relu_bf16x2(d[x][e + 6], d[x][e + 7]);
          y[x] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8_t, g), y[x],
                                                         0, 0, 0);
        }
      }
    }
    // rows 0 and 1 (v3_hi and v3_lo products) are elements 0 and 1 of lanes 0-31
#pragma unroll
    for (int x = 0; x < RT; ++x)
      if (half == 0 && valid[x]) out[row[x]] = y[x][0] + y[x][1] + bias3;
  }
}

// Inference encoder, fused: per 32 edges of x (one wave), h1 = relu(W1 x +
// b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3) (each rounded to bf16 as
// in the unfused path) and the max-pool of h3 over each part's E edges into
// zbits (float bits of the non-negative pooled values, which order like
// unsigned integers; zbits must start at 0), with no activation in memory.
// Layers 1 and 2 run transposed (units as rows, edges as columns), so each
// result is the next product's B operand (head_perm order, as in
// k_head_fused); layer 3 runs untransposed with h2 as the A operand (the same
// register layout serves both) and W3 as B, so each lane ends with 16 edges
// of one unit and pools them in registers: one atomic per unit and tile.
// The block's waves take NW tiles per round: W2 (columns permuted) is staged
// into LDS for layers 1-2, then W3 (rows permuted) into the same LDS for
// layer 3, with block barriers between.
// u[b][n] = c1[n] + sum_k Wz[n][k] bf16(z[b][k]) for the pooled parts (the
// product of the unfused path in float32 with the same bf16 operands, in
// another summation order): a block per part, a thread per output, its Wz
// row read 16 bytes at a time with every load issued up front.  A 64 x 256 x
// 256 product through k_gemm runs in 4 blocks and waits out 16 load
// latencies (21 us); this takes about 3.
__global__ __launch_bounds__(kD) void k_pool_proj(float *__restrict__ z,
                                                  const unsigned short *__restrict__ Wz,
                                                  const float *__restrict__ c1,
                                                  float *__restrict__ u) {
  __shared__ float zs[kD];
  const int b = blockIdx.x, n = threadIdx.x;
  zs[n] = bf2f(f2bf(z[b * kD + n]));
  z[b * kD + n] = 0.0f;  // the next call's pooling starts from zero
  __syncthreads();
  uint4 w[kD / 8];
#pragma unroll
  for (int j = 0; j < kD / 8; ++j) w[j] = reinterpret_cast<const uint4 *>(Wz + (size_t)n * kD)[j];
  float acc = 0.0f;
#pragma unroll
  for (int j = 0; j < kD / 8; ++j) {
    const unsigned x[4] = {w[j].x, w[j].y, w[j].z, w[j].w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      acc = fmaf(__uint_as_float(x[i] << 16), zs[8 * j + 2 * i], acc);
      acc = fmaf(__uint_as_float(x[i] & 0xffff0000u), zs[8 * j + 2 * i + 1], acc);
    }
  }
  u[b * kD + n] = acc + c1[n];
}

template <int NT>  // threads of the block
__device__ __forceinline__ void stage_w(unsigned short *sW, const unsigned short *__restrict__ W,
                                        int tid) {
  // rows of W with permuted columns; every load issued before the first store
  constexpr int kChunks = kD * kD / 8, kPer = (kChunks + NT - 1) / NT;
  uint4 v[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int c = tid + j * NT;
    if (c < kChunks) v[j] = reinterpret_cast<const uint4 *>(W)[c];
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int c = tid + j * NT;
    if (c >= kChunks) break;
    const int n = c / (kD / 8), k = 8 * (c % (kD / 8));
    *reinterpret_cast<uint2 *>(&sW[n * kV2Stride + head_perm(k)]) = make_uint2(v[j].x, v[j].y);
    *reinterpret_cast<uint2 *>(&sW[n * kV2Stride + head_perm(k + 4)]) = make_uint2(v[j].z, v[j].w);
  }
}

template <int NW>
__global__ __launch_bounds__(64 * NW) void k_encoder(
    const float *__restrict__ feat, int64_t R, int E, const unsigned short *__restrict__ W1,
    const float *__restrict__ b1, const unsigned short *__restrict__ W2,
    const float *__restrict__ b2, const unsigned short *__restrict__ W3,
    const float *__restrict__ b3, unsigned *__restrict__ zbits) {
  __shared__ __attribute__((aligned(16))) unsigned short sW[kD * kV2Stride];
  __shared__ __attribute__((aligned(16))) unsigned short sW1[kD * 8];
  __shared__ __attribute__((aligned(16))) float sb[3][kD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int k = tid; k < kD; k += 64 * NW) {
    *reinterpret_cast<uint4 *>(&sW1[8 * k]) = reinterpret_cast<const uint4 *>(W1)[k];
    sb[0][k] = b1[k], sb[1][k] = b2[k], sb[2][k] = b3[k];
  }
  const int half = lane >> 5, col = lane & 31;
  const int64_t tiles = (R + 31) / 32;
  for (int64_t base = (int64_t)blockIdx.x * NW; base < tiles; base += (int64_t)gridDim.x * NW) {
    const int64_t tile = base + wave;
    const bool live = tile < tiles;  // wave-uniform; every wave joins the barriers
    const int64_t e0 = tile * 32;
    stage_w<64 * NW>(sW, W2, tid);
    __syncthreads();
    bf16x8_t h2[16];
    if (live) {
      // layer 1: B = x^T (lanes 0-31: the 7 features and a zero of edge col;
      // lanes 32-63: zeros, k padded to 16), A = W1 rows (lanes 0-31)
      bf16x8_t bx;
      {
        const int64_t e = e0 + col;
#pragma unroll
        for (int i = 0; i < 8; ++i)
          bx[i] = (__bf16)((half == 0 && i < 7 && e < R) ? feat[e * 7 + i] : 0.0f);
      }
      bf16x8_t h1[16];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        bf16x8_t a = *reinterpret_cast<const bf16x8_t *>(&sW1[8 * (32 * t + col)]);
        if (half) {
#pragma unroll
          for (int i = 0; i < 8; ++i) a[i] = (__bf16)0.0f;
        }
        f32x16 d;
#pragma unroll
        for (int r = 0; r < 16; ++r) d[r] = 0.0f;
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bx, d, 0, 0, 0);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const float4 c0 = *reinterpret_cast<const float4 *>(&sb[0][32 * t + 16 * b + 4 * half]);
          const float4 c1 =
              *reinterpret_cast<const float4 *>(&sb[0][32 * t + 16 * b + 8 + 4 * half]);
          const int e = 8 * b;
          uint4 g;
          g.x = relu_bf16x2(d[e] + c0.x, d[e + 1] + c0.y);
          g.y = relu_bf16x2(d[e + 2] + c0.z, d[e + 3] + c0.w);
          g.z = relu_bf16x2(d[e + 4] + c1.x, d[e + 5] + c1.y);
          g.w = relu_bf16x2(d[e + 6] + c1.z, d[e + 7] + c1.w);
          h1[2 * t + b] = __builtin_bit_cast(bf16x8_t, g);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      // layer 2: 16 k-steps per 32 units, A = W2 rows from LDS
#pragma unroll
      for (int s2 = 0; s2 < 8; ++s2) {
        f32x16 d;
#pragma unroll
        for (int r = 0; r < 16; ++r) d[r] = 0.0f;
        const unsigned short *wrow = &sW[(32 * s2 + col) * kV2Stride + 8 * half];
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
          d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
              *reinterpret_cast<const bf16x8_t *>(wrow + 16 * kk), h1[kk], d, 0, 0, 0);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const float4 c0 = *reinterpret_cast<const float4 *>(&sb[1][32 * s2 + 16 * b + 4 * half]);
          const float4 c1 =
              *reinterpret_cast<const float4 *>(&sb[1][32 * s2 + 16 * b + 8 + 4 * half]);
          const int e = 8 * b;
          uint4 g;
          g.x = relu_bf16x2(d[e] + c0.x, d[e + 1] + c0.y);
          g.y = relu_bf16x2(d[e + 2] + c0.z, d[e + 3] + c0.w);
          g.z = relu_bf16x2(d[e + 4] + c1.x, d[e + 5] + c1.y);
          g.w = relu_bf16x2(d[e + 6] + c1.z, d[e + 7] + c1.w);
          h2[2 * s2 + b] = __builtin_bit_cast(bf16x8_t, g);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();  // every wave is done with W2
    stage_w<64 * NW>(sW, W3, tid);
    __syncthreads();
    if (live) {
      // layer 3 untransposed: rows = edges (A = h2), columns = units (B = W3
      // rows from LDS); element r of lane l: edge e0 + 8 (r / 4) + 4 half +
      // r % 4, unit 32 s3 + col
      const int64_t part0 = e0 / E;
      const bool one_part = (std::min<int64_t>(e0 + 31, R - 1)) / E == part0;
      // a whole tile of one part: rounding to bf16, the relu and adding the
      // bias are monotone, so the pooled value is that of the largest raw
      // sum, found before any of them is applied
      const bool whole = one_part && e0 + 32 <= R;
#pragma unroll 1
      for (int s3 = 0; s3 < 8; ++s3) {
        const int n = 32 * s3 + col;
        const unsigned short *wcol = &sW[n * kV2Stride + 8 * half];
        f32x16 d;
#pragma unroll
        for (int r = 0; r < 16; ++r) d[r] = 0.0f;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
          d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
              h2[kk], *reinterpret_cast<const bf16x8_t *>(wcol + 16 * kk), d, 0, 0, 0);
        const float bn = sb[2][n];
        if (whole) {
          float m = d[0];
#pragma unroll
          for (int r = 1; r < 16; ++r) m = fmaxf(m, d[r]);
          m = fmaxf(m, __shfl_xor(m, 32));  // the other half-wave's edges
          if (half == 0)
            atomicMax(&zbits[part0 * kD + n], __float_as_uint(bf2f(f2bf(fmaxf(m + bn, 0.0f)))));
          continue;
        }
        float mx = 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int64_t e = e0 + 8 * (r >> 2) + 4 * half + (r & 3);
          if (e >= R) continue;
          // the value the unfused path pools: bf16 of the relu output
          const float v = bf2f(f2bf(fmaxf(d[r] + bn, 0.0f)));
          if (one_part) mx = fmaxf(mx, v);
          else atomicMax(&zbits[(e / E) * kD + n], __float_as_uint(v));
        }
        if (one_part) {
          mx = fmaxf(mx, __shfl_xor(mx, 32));  // the other half-wave's edges
          if (half == 0) atomicMax(&zbits[part0 * kD + n], __float_as_uint(mx));
        }
      }
    }
    __syncthreads();  // every wave is done with W3 before the next round's W2
  }
}



// loss gradient and the last layer: dout = 2 (out - t) / BP; dg2 = relu'(g2) dout v3;
// dv3 += dout g2 and dc3 += dout (atomics); loss += (out - t)^2 / BP
__global__ void k_head_back(const float *__restrict__ out, const float *__restrict__ t,
                            const unsigned short *__restrict__ g2, const float *__restrict__ v3,
                            int64_t BP, unsigned short *__restrict__ dg2, float *__restrict__ dv3,
                            float *__restrict__ dc3, float *__restrict__ loss) {
  const int64_t p = blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= BP) return;
  const float r = out[p] - t[p];
  const float d = 2.0f * r / (float)BP;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int f = lane + 64 * k;
    const float g = bf2f(g2[p * kD + f]);
    dg2[p * kD + f] = f2bf(g > 0.0f ? d * v3[f] : 0.0f);
    atomicAdd(&dv3[f], d * g);
  }
  if (lane == 0) {
    atomicAdd(dc3, d);
    atomicAdd(loss, r * r / (float)BP);
  }
}

// out[r][c] = bf16(keep[r][c] > 0 ? x[r][c] : 0) (backward through relu)
__global__ void k_relu_back(const float *__restrict__ x, const unsigned short *__restrict__ keep,
                            int64_t n, unsigned short *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = f2bf(bf2f(keep[i]) > 0.0f ? x[i] : 0.0f);
}

// column sums of R rows of a bf16 [R][256] array into out (atomics), rows
// grouped by `group` (out[(r / group) * 256 + c]) when group > 0
__global__ void k_colsum(const unsigned short *__restrict__ x, int64_t R, int64_t group,
                         float *__restrict__ out) {
  const int c = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int64_t r1 = r0 + 64 < R ? r0 + 64 : R;
  float s = 0.0f;
  int64_t g = group > 0 ? r0 / group : 0;
  for (int64_t r = r0; r < r1; ++r) {
    const int64_t gr = group > 0 ? r / group : 0;
    if (gr != g) {
      atomicAdd(&out[g * kD + c], s);
      s = 0.0f;
      g = gr;
    }
    s += bf2f(x[r * kD + c]);
  }
  atomicAdd(&out[g * kD + c], s);
}

// dWp[f][j] += sum_p dg1[p][f] q[p][j]
__global__ void k_wp_grad(const unsigned short *__restrict__ dg1, const float *__restrict__ q,
                          int64_t BP, float *__restrict__ dWp) {
  const int f = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int64_t r1 = r0 + 64 < BP ? r0 + 64 : BP;
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  for (int64_t p = r0; p < r1; ++p) {
    const float d = bf2f(dg1[p * kD + f]);
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] += d * q[4 * p + j];
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) atomicAdd(&dWp[4 * f + j], s[j]);
}

// gradient through the max-pool: row arg[b][f] of part b gets dz[b][f]
// (when the pooled value is positive), every other row 0
__global__ void k_unpool(const float *__restrict__ dz, const int *__restrict__ arg,
                         const float *__restrict__ z, int B, int E,
                         unsigned short *__restrict__ dh) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * E * kD) return;
  const int64_t row = i / kD;
  const int f = (int)(i % kD);
  const int b = (int)(row / E), e = (int)(row % E);
  const bool hit = arg[b * kD + f] == e && z[b * kD + f] > 0.0f;
  dh[i] = f2bf(hit ? dz[b * kD + f] : 0.0f);
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

}  // namespace

void gemm_bf16(vacd_ctx *ctx, bool ta, bool tb, int M, int N, int K, const void *A, int lda,
               const void *B, int ldb, const float *bias, bool relu, bool acc, float *Cf,
               void *Cb, int ldc) {
  if (M <= 0 || N <= 0) return;
  const dim3 grid((unsigned)((N + kTN - 1) / kTN), (unsigned)((M + kTM - 1) / kTM));
  auto a = static_cast<const unsigned short *>(A);
  auto b = static_cast<const unsigned short *>(B);
  auto cb = static_cast<unsigned short *>(Cb);
  hipStream_t st = ctx->stream;
  if (!ta && !tb)
    k_gemm<false, false><<<grid, 256, 0, st>>>(M, N, K, a, lda, b, ldb, bias, relu, acc, Cf, cb, ldc);
  else if (!ta && tb)
    k_gemm<false, true><<<grid, 256, 0, st>>>(M, N, K, a, lda, b, ldb, bias, relu, acc, Cf, cb, ldc);
  else if (ta && !tb)
    k_gemm<true, false><<<grid, 256, 0, st>>>(M, N, K, a, lda, b, ldb, bias, relu, acc, Cf, cb, ldc);
  else
    k_gemm<true, true><<<grid, 256, 0, st>>>(M, N, K, a, lda, b, ldb, bias, relu, acc, Cf, cb, ldc);
  VACD_HIP(hipGetLastError());
}

void to_bf16(vacd_ctx *ctx, const float *x, int64_t n, void *y) {
  if (n <= 0) return;
  k_f2bf<<<(unsigned)std::min<int64_t>((n + 255) / 256, 4096), 256, 0, ctx->stream>>>(
      x, n, static_cast<unsigned short *>(y));
  VACD_HIP(hipGetLastError());
}

// Parameter layout (fp32, in this order): W1[256][8] b1[256] W2[256][256] b2
// W3[256][256] b3 Wz[256][256] Wp[256][4] c1[256] V2[256][256] c2 v3[256] c3[1].
struct MlpLayout {
  int64_t W1, b1, W2, b2, W3, b3, Wz, Wp, c1, V2, c2, v3, c3, n;
  MlpLayout() {
    int64_t o = 0;
    auto take = [&](int64_t k) {
      const int64_t r = o;
      o += k;
      return r;
    };
    W1 = take(kD * 8), b1 = take(kD), W2 = take(kD * kD), b2 = take(kD), W3 = take(kD * kD);
    b3 = take(kD), Wz = take(kD * kD), Wp = take(kD * 4), c1 = take(kD), V2 = take(kD * kD);
    c2 = take(kD), v3 = take(kD), c3 = take(1), n = o;
  }
};
// bf16 shadows: W1, W2, W3, Wz, V2 at these element offsets
constexpr int64_t kB_W1 = 0, kB_W2 = kD * 8, kB_W3 = kB_W2 + kD * kD, kB_Wz = kB_W3 + kD * kD,
                  kB_V2 = kB_Wz + kD * kD, kB_n = kB_V2 + kD * kD;

}  // namespace vacd

struct vacd_mlp {
  int device = 0;
  vacd::MlpLayout L;
  vacd::DevBuf w, g, m, v, wb, act;
  int64_t steps = 0;
  // pooled maxima of the fused inference path: the projection that reads a
  // row clears it, so a call finds its rows at zero and needs no fill; rows
  // [0, zpool_clean) are known to be zero between calls
  vacd::DevBuf zpool;
  int64_t zpool_clean = 0;
};

namespace vacd {
namespace {

void refresh_shadows(vacd_ctx *ctx, vacd_mlp *M) {
  const float *w = M->w.as<float>();
  unsigned short *wb = M->wb.as<unsigned short>();
  to_bf16(ctx, w + M->L.W1, kD * 8, wb + kB_W1);
  to_bf16(ctx, w + M->L.W2, kD * kD, wb + kB_W2);
  to_bf16(ctx, w + M->L.W3, kD * kD, wb + kB_W3);
  to_bf16(ctx, w + M->L.Wz, kD * kD, wb + kB_Wz);
  to_bf16(ctx, w + M->L.V2, kD * kD, wb + kB_V2);
}

// Activation buffers of one forward pass (bf16 unless noted)
struct Acts {
  unsigned short *x0, *h1, *h2, *h3, *zb, *g1, *g2;
  float *z, *u, *out;  // fp32
  int *arg;
  // backward
  unsigned short *dg2, *dg1, *dub, *dh3, *dh2, *dh1;
  float *tmp, *du, *dz, *loss;
};

// keep_head false: no g1 / g2 (the fused inference head keeps them in registers)
Acts carve(vacd_mlp *M, int64_t R, int64_t B, int64_t BP, bool backward, bool keep_head = true) {
  size_t off = 0;
  std::vector<std::pair<void **, size_t>> want;
  Acts a{};
  auto add = [&](void **p, size_t bytes) { want.push_back({p, bytes}); };
  add((void **)&a.x0, 2 * R * 8);
  add((void **)&a.h1, 2 * R * kD);
  add((void **)&a.h2, 2 * R * kD);
  add((void **)&a.h3, 2 * R * kD);
  add((void **)&a.zb, 2 * B * kD);
  if (keep_head) {
    add((void **)&a.g1, 2 * BP * kD);
    add((void **)&a.g2, 2 * BP * kD);
  }
  add((void **)&a.z, 4 * B * kD);
  add((void **)&a.u, 4 * B * kD);
  add((void **)&a.out, 4 * BP);
  add((void **)&a.arg, 4 * B * kD);
  if (backward) {
    add((void **)&a.dg2, 2 * BP * kD);
    add((void **)&a.dg1, 2 * BP * kD);
    add((void **)&a.dub, 2 * B * kD);
    add((void **)&a.dh3, 2 * R * kD);
    add((void **)&a.dh2, 2 * R * kD);
    add((void **)&a.dh1, 2 * R * kD);
    add((void **)&a.tmp, 4 * std::max(BP, R) * kD);
    add((void **)&a.du, 4 * B * kD);
    add((void **)&a.dz, 4 * B * kD);
    add((void **)&a.loss, 256);
  }
  for (auto &x : want) off += (x.second + 255) & ~size_t(255);
  M->act.ensure(off);
  char *base = M->act.as<char>();
  off = 0;
  for (auto &x : want) {
    *x.first = base + off;
    off += (x.second + 255) & ~size_t(255);
  }
  return a;
}

unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

// fused: the inference head (k_head_fused) in place of g1, g2 and the output
// layer as separate passes (training needs them)
void forward(vacd_ctx *ctx, vacd_mlp *M, const Acts &a, int B, int E, const float *feat, int P,
             const float *planes, bool fused = false) {
  hipStream_t st = ctx->stream;
  const float *w = M->w.as<float>();
  const unsigned short *wb = M->wb.as<unsigned short>();
  const int64_t R = (int64_t)B * E, BP = (int64_t)B * P;
  float *zp = a.z;
  int64_t clean_after = 0;
  if (fused) {  // the whole encoder and its max-pool in one pass
    const size_t zbytes = 4 * (size_t)B * kD;
    if (M->zpool.bytes < zbytes) {
      M->zpool.ensure(zbytes);
      M->zpool_clean = 0;
    }
    zp = M->zpool.as<float>();
    if (M->zpool_clean < B) fill_u64(ctx, zp, (int64_t)B * kD / 2, 0ull);
    clean_after = std::max<int64_t>(M->zpool_clean, B);
    M->zpool_clean = 0;  // in use until this call's projection has cleared its rows
    // one wave per SIMD (a tile of 32 edges per wave): 64 parts of 512 edges
    // make 1024 tiles, one per SIMD, and a wave may hold both layers'
    // activations in registers
    constexpr int kEncWaves = 4;
    const unsigned grid =
        (unsigned)std::min<int64_t>(((R + 31) / 32 + kEncWaves - 1) / kEncWaves, 256);
    k_encoder<kEncWaves><<<grid, 64 * kEncWaves, 0, st>>>(
        feat, R, E, wb + kB_W1, w + M->L.b1, wb + kB_W2, w + M->L.b2, wb + kB_W3, w + M->L.b3,
        reinterpret_cast<unsigned *>(zp));
    VACD_HIP(hipGetLastError());
  } else {
    k_edge_in<<<blocks_for(R * 8), 256, 0, st>>>(feat, R, a.x0);
    gemm_bf16(ctx, false, true, (int)R, kD, 8, a.x0, 8, wb + kB_W1, 8, w + M->L.b1, true, false,
              nullptr, a.h1, kD);
    gemm_bf16(ctx, false, true, (int)R, kD, kD, a.h1, kD, wb + kB_W2, kD, w + M->L.b2, true, false,
              nullptr, a.h2, kD);
    gemm_bf16(ctx, false, true, (int)R, kD, kD, a.h2, kD, wb + kB_W3, kD, w + M->L.b3, true, false,
              nullptr, a.h3, kD);
  }
  if (!fused) {
    k_maxpool<<<(unsigned)B, kD, 0, st>>>(a.h3, E, a.z, a.arg);
    to_bf16(ctx, a.z, (int64_t)B * kD, a.zb);
    gemm_bf16(ctx, false, true, B, kD, kD, a.zb, kD, wb + kB_Wz, kD, w + M->L.c1, false, false,
              a.u, nullptr, kD);
  } else {
    k_pool_proj<<<(unsigned)B, kD, 0, st>>>(zp, wb + kB_Wz, w + M->L.c1, a.u);
    VACD_HIP(hipGetLastError());
    M->zpool_clean = clean_after;
  }
  if (fused) {
    // 32 planes per wave and step, one block of kHeadFusedWaves waves per CU
    constexpr int kWaves = kHeadFusedWaves;
    const int64_t tiles = (BP + 31) / 32;
    const unsigned grid = (unsigned)std::min<int64_t>((tiles + kWaves - 1) / kWaves, 256);
    if (P >= 32)
      k_head_fused<kWaves, kHeadTiles, false><<<grid, 64 * kWaves, 0, st>>>(
          a.u, w + M->L.Wp, planes, BP, P, wb + kB_V2, w + M->L.c2, w + M->L.v3, w + M->L.c3, a.out);
    else
      k_head_fused<kWaves, 1, true><<<grid, 64 * kWaves, 0, st>>>(
          a.u, w + M->L.Wp, planes, BP, P, wb + kB_V2, w + M->L.c2, w + M->L.v3, w + M->L.c3, a.out);
    VACD_HIP(hipGetLastError());
    return;
  }
  k_head1<<<blocks_for(BP * kD), 256, 0, st>>>(a.u, w + M->L.Wp, planes, BP, P, a.g1);
  gemm_bf16(ctx, false, true, (int)BP, kD, kD, a.g1, kD, wb + kB_V2, kD, w + M->L.c2, true, false,
            nullptr, a.g2, kD);
  k_head_out<<<blocks_for(BP, 4), 256, 0, st>>>(a.g2, w + M->L.v3, w + M->L.c3, BP, a.out);
  VACD_HIP(hipGetLastError());
}

void check_dims(int B, int E, int P) {
  VACD_REQUIRE(B >= 1 && E >= 1 && P >= 1, VACD_ERR_INVALID, "empty MLP batch");
  VACD_REQUIRE((int64_t)B * E < (int64_t(1) << 26) && (int64_t)B * P < (int64_t(1) << 26),
               VACD_ERR_LIMIT, "MLP batch too large");
}

}  // namespace

void mlp_create(vacd_ctx *ctx, uint64_t seed, vacd_mlp **out) {
  vacd_mlp *M = new vacd_mlp();
  try {
    M->device = ctx->device;
    const int64_t n = M->L.n;
    // padded to whole 8-byte words: the buffers are cleared with 64-bit fills
    M->w.ensure(8 * ((n + 1) / 2));
    M->g.ensure(8 * ((n + 1) / 2));
    M->m.ensure(8 * ((n + 1) / 2));
    M->v.ensure(8 * ((n + 1) / 2));
    M->wb.ensure(2 * kB_n);
    // He-uniform weights from a seeded generator, zero biases
    std::vector<float> h(n, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int64_t o, int64_t count, int fan_in) {
      const double lim = std::sqrt(6.0 / fan_in);
      std::uniform_real_distribution<double> u(-lim, lim);
      for (int64_t i = 0; i < count; ++i) h[o + i] = (float)u(rng);
    };
    fill(M->L.W1, kD * 8, 7);
    fill(M->L.W2, kD * kD, kD);
    fill(M->L.W3, kD * kD, kD);
    fill(M->L.Wz, kD * kD, kD);
    fill(M->L.Wp, kD * 4, 4);
    fill(M->L.V2, kD * kD, kD);
    fill(M->L.v3, kD, kD);
    VACD_HIP(hipMemcpyAsync(M->w.ptr, h.data(), 4 * n, hipMemcpyHostToDevice, ctx->stream));
    fill_u64(ctx, M->m.ptr, (n + 1) / 2, 0ull);
    fill_u64(ctx, M->v.ptr, (n + 1) / 2, 0ull);
    refresh_shadows(ctx, M);
    VACD_HIP(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    M->w.release(), M->g.release(), M->m.release(), M->v.release(), M->wb.release();
    delete M;
    throw;
  }
  *out = M;
}

int64_t MlpLayoutCount() { return MlpLayout().n; }

void mlp_destroy(vacd_mlp *M) {
  if (!M) return;
  (void)hipSetDevice(M->device);
  M->w.release(), M->g.release(), M->m.release(), M->v.release(), M->wb.release();
  M->act.release();
  M->zpool.release();
  delete M;
}

void mlp_set(vacd_ctx *ctx, vacd_mlp *M, const float *host) {
  VACD_HIP(hipMemcpyAsync(M->w.ptr, host, 4 * M->L.n, hipMemcpyHostToDevice, ctx->stream));
  refresh_shadows(ctx, M);
  VACD_HIP(hipStreamSynchronize(ctx->stream));
}

void mlp_get(vacd_ctx *ctx, vacd_mlp *M, float *host) {
  VACD_HIP(hipMemcpyAsync(host, M->w.ptr, 4 * M->L.n, hipMemcpyDeviceToHost, ctx->stream));
  VACD_HIP(hipStreamSynchronize(ctx->stream));
}

void mlp_get_grads(vacd_ctx *ctx, vacd_mlp *M, float *host) {
  VACD_HIP(hipMemcpyAsync(host, M->g.ptr, 4 * M->L.n, hipMemcpyDeviceToHost, ctx->stream));
  VACD_HIP(hipStreamSynchronize(ctx->stream));
}

void mlp_forward(vacd_ctx *ctx, vacd_mlp *M, int B, int E, const float *feat, int P,
                 const float *planes, float *out) {
  check_dims(B, E, P);
  Acts a = carve(M, (int64_t)B * E, B, (int64_t)B * P, false, false);
  a.out = out;  // the fused head writes the predictions straight to the caller's buffer
  forward(ctx, M, a, B, E, feat, P, planes, true);
}

void mlp_forward_host(vacd_ctx *ctx, vacd_mlp *M, int B, int E, const float *feat, int P,
                      const float *planes, float *out) {
  check_dims(B, E, P);
  const int64_t R = (int64_t)B * E, BP = (int64_t)B * P;
  DevBuf &din = ctx->buf[SLOT_MLP_IN];
  din.ensure(4 * (size_t)(7 * R + 4 * BP));
  float *dfeat = din.as<float>(), *dplanes = dfeat + 7 * R;
  VACD_HIP(hipMemcpyAsync(dfeat, feat, 28 * (size_t)R, hipMemcpyHostToDevice, ctx->stream));
  VACD_HIP(hipMemcpyAsync(dplanes, planes, 16 * (size_t)BP, hipMemcpyHostToDevice, ctx->stream));
  const Acts a = carve(M, R, B, BP, false, false);
  forward(ctx, M, a, B, E, dfeat, P, dplanes, true);
  VACD_HIP(hipMemcpyAsync(out, a.out, 4 * (size_t)BP, hipMemcpyDeviceToHost, ctx->stream));
  VACD_HIP(hipStreamSynchronize(ctx->stream));
}

float mlp_train_step(vacd_ctx *ctx, vacd_mlp *M, int B, int E, const float *feat, int P,
                     const float *planes, const float *target, float lr) {
  check_dims(B, E, P);
  hipStream_t st = ctx->stream;
  const int64_t R = (int64_t)B * E, BP = (int64_t)B * P;
  const Acts a = carve(M, R, B, BP, true);
  forward(ctx, M, a, B, E, feat, P, planes);
  const MlpLayout &L = M->L;
  float *w = M->w.as<float>(), *g = M->g.as<float>();
  const unsigned short *wb = M->wb.as<unsigned short>();
  fill_u64(ctx, g, (L.n + 1) / 2, 0ull);
  fill_u64(ctx, a.loss, 1, 0ull);
  fill_u64(ctx, a.du, (int64_t)B * kD / 2, 0ull);
  // head
  k_head_back<<<blocks_for(BP, 4), 256, 0, st>>>(a.out, target, a.g2, w + L.v3, BP, a.dg2,
                                                  g + L.v3, g + L.c3, a.loss);
  gemm_bf16(ctx, true, false, kD, kD, (int)BP, a.dg2, kD, a.g1, kD, nullptr, false, false,
            g + L.V2, nullptr, kD);                                   // dV2 = dg2^T g1
  k_colsum<<<blocks_for(BP, 64), kD, 0, st>>>(a.dg2, BP, 0, g + L.c2);  // dc2
  gemm_bf16(ctx, false, false, (int)BP, kD, kD, a.dg2, kD, wb + kB_V2, kD, nullptr, false, false,
            a.tmp, nullptr, kD);                                      // dg1 = dg2 V2
  k_relu_back<<<blocks_for(BP * kD), 256, 0, st>>>(a.tmp, a.g1, BP * kD, a.dg1);
  k_wp_grad<<<blocks_for(BP, 64), kD, 0, st>>>(a.dg1, planes, BP, g + L.Wp);
  k_colsum<<<blocks_for(BP, 64), kD, 0, st>>>(a.dg1, BP, P, a.du);    // du per part
  k_colsum<<<blocks_for(BP, 64), kD, 0, st>>>(a.dg1, BP, 0, g + L.c1);
  to_bf16(ctx, a.du, (int64_t)B * kD, a.dub);
  gemm_bf16(ctx, true, false, kD, kD, B, a.dub, kD, a.zb, kD, nullptr, false, false, g + L.Wz,
            nullptr, kD);                                             // dWz = du^T z
  gemm_bf16(ctx, false, false, B, kD, kD, a.dub, kD, wb + kB_Wz, kD, nullptr, false, false, a.dz,
            nullptr, kD);                                             // dz = du Wz
  // encoder
  k_unpool<<<blocks_for(R * kD), 256, 0, st>>>(a.dz, a.arg, a.z, B, E, a.dh3);
  gemm_bf16(ctx, true, false, kD, kD, (int)R, a.dh3, kD, a.h2, kD, nullptr, false, false,
            g + L.W3, nullptr, kD);
  k_colsum<<<blocks_for(R, 64), kD, 0, st>>>(a.dh3, R, 0, g + L.b3);
  gemm_bf16(ctx, false, false, (int)R, kD, kD, a.dh3, kD, wb + kB_W3, kD, nullptr, false, false,
            a.tmp, nullptr, kD);
  k_relu_back<<<blocks_for(R * kD), 256, 0, st>>>(a.tmp, a.h2, R * kD, a.dh2);
  gemm_bf16(ctx, true, false, kD, kD, (int)R, a.dh2, kD, a.h1, kD, nullptr, false, false,
            g + L.W2, nullptr, kD);
  k_colsum<<<blocks_for(R, 64), kD, 0, st>>>(a.dh2, R, 0, g + L.b2);
  gemm_bf16(ctx, false, false, (int)R, kD, kD, a.dh2, kD, wb + kB_W2, kD, nullptr, false, false,
            a.tmp, nullptr, kD);
  k_relu_back<<<blocks_for(R * kD), 256, 0, st>>>(a.tmp, a.h1, R * kD, a.dh1);
  gemm_bf16(ctx, true, false, kD, 8, (int)R, a.dh1, kD, a.x0, 8, nullptr, false, false, g + L.W1,
            nullptr, 8);
  k_colsum<<<blocks_for(R, 64), kD, 0, st>>>(a.dh1, R, 0, g + L.b1);
  // Adam
  M->steps += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)M->steps), c2 = 1.0f - std::pow(b2, (float)M->steps);
  k_adam<<<blocks_for(L.n), 256, 0, st>>>(w, g, M->m.as<float>(), M->v.as<float>(), L.n, lr, b1, b2,
                                          1e-8f, c1, c2);
  refresh_shadows(ctx, M);
  float loss = 0.0f;
  VACD_HIP(hipMemcpyAsync(&loss, a.loss, 4, hipMemcpyDeviceToHost, st));
  VACD_HIP(hipStreamSynchronize(st));
  return loss;
}

}  // namespace vacd
