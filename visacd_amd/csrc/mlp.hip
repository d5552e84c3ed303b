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
// nor g2 in memory (the unfused training path writes both).  A wave takes 32
// planes at a time; v_mfma_f32_32x32x16_bf16 operands: lane l holds row (A)
// or column (B) l % 32 and k = 8 (l / 32) .. +7; result element r of lane l
// is row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.  The products run
// transposed (hidden units as rows, planes as columns), so that what one
// MFMA leaves in a lane -- 16 hidden units of the lane's plane -- is the B
// operand of the next one for two k-steps, once the hidden units are
// renumbered; the V2 columns and v3 are stored in that order (head_perm).
//   g1: eight MFMAs, one per 32 hidden units: A = (Wp_hi | Wp_lo) in lanes
//       0-31, (Wp_hi | u_hi u_mid u_lo 0) in lanes 32-63, B = (q_hi | q_hi)
//       and (q_lo | 1 1 1 0): Wp q + u with bf16 splits of q and Wp (only the
//       lo x lo term, about 2^-16 of the product, is left out) and u exact as
//       three bf16 terms; then relu and the bf16 rounding of the unfused
//       path.  u belongs to the part of the tile's first plane; lanes whose
//       plane lies in the next part (tiles spanning parts) are corrected by
//       the difference of the two parts' u in float32.
//   g2: 16 k-steps per 32 outputs with V2 from LDS (rows padded by 16 bytes
//       against bank conflicts), + c2, relu, bf16.
//   y:  two MFMAs per 32 outputs with A = (v3_hi, v3_lo) in rows 0 and 1:
//       y = row 0 + row 1 + c3.  Products of bf16 values are exact in the
//       float32 accumulators.
// One block per CU (V2 fills most of its LDS), NW waves.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
constexpr int kHeadWaves = 8;
constexpr int kV2Stride = kD + 8;  // bf16 per padded LDS row

// Position of hidden unit c in the order of the head's B operands: for c =
// 32 t + 16 b + 8 j + 4 h + r, the lanes of half h take units 32 t + 16 b +
// 8 h + [0, 8) of k-step 2 t + b, which hold 4 j + r.
__host__ __device__ __forceinline__ int head_perm(int c) {
  return (c & ~31) | (c & 16) | ((c & 4) << 1) | ((c & 8) >> 1) | (c & 3);
}

__device__ __forceinline__ void split2(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <int NW>
__global__ __launch_bounds__(64 * NW) void k_head_fused(
    const float *__restrict__ u, const float *__restrict__ Wp, const float *__restrict__ q,
    int64_t BP, int P, const unsigned short *__restrict__ V2, const float *__restrict__ c2,
    const float *__restrict__ v3, const float *__restrict__ c3, float *__restrict__ out) {
  __shared__ __attribute__((aligned(16))) unsigned short sV[kD * kV2Stride];
  __shared__ __attribute__((aligned(16))) __bf16 sV3[2 * kD];
  __shared__ float sc2[kD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // V2 rows into LDS with permuted columns: 8 consecutive columns (aligned)
  // are two runs of 4 in the permuted order
  for (int c = tid; c < kD * kD / 8; c += 64 * NW) {
    const int n = c / (kD / 8), k = 8 * (c % (kD / 8));
    const uint4 v = reinterpret_cast<const uint4 *>(V2)[c];
    *reinterpret_cast<uint2 *>(&sV[n * kV2Stride + head_perm(k)]) = make_uint2(v.x, v.y);
    *reinterpret_cast<uint2 *>(&sV[n * kV2Stride + head_perm(k + 4)]) = make_uint2(v.z, v.w);
  }
  for (int k = tid; k < kD; k += 64 * NW) {
    __bf16 hi, lo;
    split2(v3[k], hi, lo);
    sV3[head_perm(k)] = hi;
    sV3[kD + head_perm(k)] = lo;
    sc2[k] = c2[k];
  }
  __syncthreads();
  const float bias3 = c3[0];
  const int half = lane >> 5, col = lane & 31;
  // the g1 A operands: Wp splits (both halves), u terms (half 1, per part)
  bf16x8_t a1[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int m = 32 * t + col;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      __bf16 hi, lo;
      split2(Wp[4 * m + c], hi, lo);
      a1[t][c] = hi;
      a1[t][4 + c] = half ? (__bf16)0.0f : lo;
    }
  }
  int64_t cur = -1;  // the part whose u the a1 operands hold
  const int64_t tiles = (BP + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * NW + wave; tile < tiles;
       tile += (int64_t)gridDim.x * NW) {
    const int64_t row = tile * 32 + col;  // this lane's plane
    const bool valid = row < BP;
    const int64_t part0 = tile * 32 / P;
    const int64_t part = valid ? row / P : part0;
    const bool spans = (std::min<int64_t>(tile * 32 + 31, BP - 1)) / P != part0;
    if (part0 != cur) {  // wave-uniform
      cur = part0;
      if (half) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const float x = u[part0 * kD + 32 * t + col];
          const __bf16 h0 = (__bf16)x;
          const float r1 = x - (float)h0;
          const __bf16 h1 = (__bf16)r1;
          a1[t][4] = h0;
          a1[t][5] = h1;
          a1[t][6] = (__bf16)(r1 - (float)h1);
        }
      }
    }
    const float4 qv = valid ? *reinterpret_cast<const float4 *>(q + 4 * row)
                            : make_float4(0.f, 0.f, 0.f, 0.f);
    bf16x8_t b1;
    {
      const float qa[4] = {qv.x, qv.y, qv.z, qv.w};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        __bf16 hi, lo;
        split2(qa[c], hi, lo);
        b1[c] = half ? lo : hi;
        b1[4 + c] = half ? (__bf16)(c < 3 ? 1.0f : 0.0f) : hi;
      }
    }
    // g1: the B operands of the main product, k-step 2 t + b in b2[2 t + b]
    bf16x8_t b2[16];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      f32x16 d;
#pragma unroll
      for (int r = 0; r < 16; ++r) d[r] = 0.0f;
      d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[t], b1, d, 0, 0, 0);
      if (spans && part != part0) {  // the lane's plane is in the next part
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int m = 32 * t + 8 * (r >> 2) + 4 * half + (r & 3);
          d[r] += u[part * kD + m] - u[part0 * kD + m];
        }
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) b2[2 * t + (r >> 3)][r & 7] = (__bf16)fmaxf(d[r], 0.0f);
    }
    f32x16 y;
#pragma unroll
    for (int r = 0; r < 16; ++r) y[r] = 0.0f;
#pragma unroll 1
    for (int s = 0; s < 8; ++s) {
      f32x16 d;
#pragma unroll
      for (int r = 0; r < 16; ++r) d[r] = 0.0f;
      const unsigned short *vrow = &sV[(32 * s + col) * kV2Stride + 8 * half];
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) {
        const bf16x8_t a = *reinterpret_cast<const bf16x8_t *>(vrow + 16 * kk);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b2[kk], d, 0, 0, 0);
      }
      bf16x8_t b3[2];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int o = 32 * s + 8 * (r >> 2) + 4 * half + (r & 3);
        b3[r >> 3][r & 7] = (__bf16)fmaxf(d[r] + sc2[o], 0.0f);
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        bf16x8_t a;
        if (col < 2) {
          a = *reinterpret_cast<const bf16x8_t *>(&sV3[col * kD + 32 * s + 16 * b + 8 * half]);
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) a[i] = (__bf16)0.0f;
        }
        y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b3[b], y, 0, 0, 0);
      }
    }
    // rows 0 and 1 (v3_hi and v3_lo products) are elements 0 and 1 of lanes 0-31
    if (half == 0 && valid) out[row] = y[0] + y[1] + bias3;
  }
}

// C[m][:] = bf16(relu(W A[m][:] + bias)) for a 256-wide layer (A bf16
// [M][256], W bf16 [256][256] as [out][in]): the inference form of the
// encoder's 256 -> 256 layers, laid out like k_head_fused (W in padded LDS
// rows, eight waves, 32 rows per wave at a time) with the A operand read
// straight from memory (lane l: row l % 32, 16 bytes at k = 8 (l / 32)).
// POOL: instead of storing C, max-pool it over each part's E rows into
// zbits[b][n] (float bits of the non-negative relu outputs, which order like
// unsigned integers; zbits must start at 0).
template <bool POOL>
__global__ __launch_bounds__(64 * kHeadWaves) void k_layer256(
    const unsigned short *__restrict__ A, int64_t M, const unsigned short *__restrict__ W,
    const float *__restrict__ bias, unsigned short *__restrict__ C, int E,
    unsigned *__restrict__ zbits) {
  __shared__ __attribute__((aligned(16))) unsigned short sV[kD * kV2Stride];
  __shared__ float sb[kD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int c = tid; c < kD * kD / 8; c += 64 * kHeadWaves) {
    const int n = c / (kD / 8), k8 = c % (kD / 8);
    *reinterpret_cast<uint4 *>(&sV[n * kV2Stride + 8 * k8]) = reinterpret_cast<const uint4 *>(W)[c];
  }
  for (int k = tid; k < kD; k += 64 * kHeadWaves) sb[k] = bias[k];
  __syncthreads();
  const int half = lane >> 5, col = lane & 31;
  const int64_t tiles = (M + 31) / 32;
  for (int64_t t = (int64_t)blockIdx.x * kHeadWaves + wave; t < tiles;
       t += (int64_t)gridDim.x * kHeadWaves) {
    const int64_t row = t * 32 + col;
    const bool valid = row < M;
    f32x16 acc[8];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[nt][r] = 0.0f;
    for (int ks = 0; ks < kD / 16; ++ks) {
      const int k0 = ks * 16 + 8 * half;
      bf16x8_t a;
      if (valid) {
        a = *reinterpret_cast<const bf16x8_t *>(A + row * kD + k0);
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = (__bf16)0.0f;
      }
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) {
        const bf16x8_t b =
            *reinterpret_cast<const bf16x8_t *>(&sV[(nt * 32 + col) * kV2Stride + k0]);
        acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[nt], 0, 0, 0);
      }
    }
    if (!POOL) {
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) {
        const int n = nt * 32 + col;
        const float bn = sb[n];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int64_t orow = t * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
          if (orow < M) C[orow * kD + n] = f2bf(fmaxf(acc[nt][r] + bn, 0.0f));
        }
      }
    } else {
      const int64_t last = (t * 32 + 31 < M ? t * 32 + 31 : M - 1);
      const bool one_part = (t * 32) / E == last / E;
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) {
        const int n = nt * 32 + col;
        const float bn = sb[n];
        float mx = 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int64_t orow = t * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
          if (orow >= M) continue;
          // the value the unfused path pools: bf16 of the relu output
          const float v = bf2f(f2bf(fmaxf(acc[nt][r] + bn, 0.0f)));
          if (one_part) mx = fmaxf(mx, v);
          else atomicMax(&zbits[(orow / E) * kD + n], __float_as_uint(v));
        }
        if (one_part) {
          mx = fmaxf(mx, __shfl_xor(mx, 32));  // the other half-wave's rows
          if (half == 0) atomicMax(&zbits[((t * 32) / E) * kD + n], __float_as_uint(mx));
        }
      }
    }
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
  k_edge_in<<<blocks_for(R * 8), 256, 0, st>>>(feat, R, a.x0);
  gemm_bf16(ctx, false, true, (int)R, kD, 8, a.x0, 8, wb + kB_W1, 8, w + M->L.b1, true, false,
            nullptr, a.h1, kD);
  if (fused) {  // the 256 -> 256 encoder layers with the weights in LDS,
                // the last one max-pooled in its epilogue
    const unsigned grid =
        (unsigned)std::min<int64_t>(((R + 31) / 32 + kHeadWaves - 1) / kHeadWaves, 256);
    k_layer256<false><<<grid, 64 * kHeadWaves, 0, st>>>(a.h1, R, wb + kB_W2, w + M->L.b2, a.h2, E,
                                                        nullptr);
    fill_u64(ctx, a.z, (int64_t)B * kD / 2, 0ull);
    k_layer256<true><<<grid, 64 * kHeadWaves, 0, st>>>(a.h2, R, wb + kB_W3, w + M->L.b3, nullptr,
                                                       E, reinterpret_cast<unsigned *>(a.z));
    VACD_HIP(hipGetLastError());
  } else {
    gemm_bf16(ctx, false, true, (int)R, kD, kD, a.h1, kD, wb + kB_W2, kD, w + M->L.b2, true, false,
              nullptr, a.h2, kD);
    gemm_bf16(ctx, false, true, (int)R, kD, kD, a.h2, kD, wb + kB_W3, kD, w + M->L.b3, true, false,
              nullptr, a.h3, kD);
  }
  if (!fused) k_maxpool<<<(unsigned)B, kD, 0, st>>>(a.h3, E, a.z, a.arg);
  to_bf16(ctx, a.z, (int64_t)B * kD, a.zb);
  gemm_bf16(ctx, false, true, B, kD, kD, a.zb, kD, wb + kB_Wz, kD, w + M->L.c1, false, false, a.u,
            nullptr, kD);
  if (fused) {
    // 32 planes per wave and step, one block of kHeadWaves waves per CU
    constexpr int kWaves = kHeadWaves;
    const int64_t tiles = (BP + 31) / 32;
    const unsigned grid = (unsigned)std::min<int64_t>((tiles + kWaves - 1) / kWaves, 256);
    k_head_fused<kWaves><<<grid, 64 * kWaves, 0, st>>>(a.u, w + M->L.Wp, planes, BP, P, wb + kB_V2,
                                                       w + M->L.c2, w + M->L.v3, w + M->L.c3, a.out);
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
  const Acts a = carve(M, (int64_t)B * E, B, (int64_t)B * P, false, false);
  forward(ctx, M, a, B, E, feat, P, planes, true);
  VACD_HIP(hipMemcpyAsync(out, a.out, 4 * (size_t)B * P, hipMemcpyDeviceToDevice, ctx->stream));
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
