// Device orientation predicates on float32 coordinates.
//
// Fast forms evaluate the determinants in float64 (differences of float32
// values, products and sums) and return a sign only when the value exceeds
// 1e-14 times its permanent, which bounds the float64 evaluation error
// (J. R. Shewchuk, DCG 18, 1997, bound A of orient2d/orient3d is below
// 8e-16 times the same permanent); otherwise they report "uncertain".
//
// Exact forms expand each determinant over the raw coordinates (rows
// x, y, z, 1): every term is a product of two or three float32 values,
// represented exactly by one double (two factors) or by a double pair
// (three factors, remainder via fma).  The terms are summed exactly into a
// nonoverlapping expansion (Grow-Expansion with zero elimination, Shewchuk
// section 2.4); the sign of its largest component is the exact sign.
//
// Before the expansions, the exact forms try integer arithmetic: float32
// values are integers m * 2^e (|m| < 2^24); when the nonzero inputs of one
// predicate have exponents e within kGridWindow bits of each other, all of
// them are integers on the grid 2^emin below 2^39, their differences below
// 2^40, and a 3x3 determinant of differences (or the sum of two) is below
// 2^124: computed exactly in signed 128-bit integers in registers.  Only
// inputs of very different magnitude take the expansion path.
//
// Only compiler builtins are used (no device intrinsics), so the same code
// also builds for the host (tests/predicates_host.cpp checks it there
// against the oracle).
#pragma once

#include <hip/hip_runtime.h>

namespace vacd {

constexpr double kFilter = 1e-14;

struct Exp {  // nonoverlapping expansion, increasing magnitude
  double c[96];
  int n;
};

__device__ __forceinline__ void exp_add(Exp &e, double b) {
  if (b == 0.0) return;
  double q = b;
  int m = 0;
  for (int t = 0; t < e.n; ++t) {
    const double ct = e.c[t];
    const double s = q + ct;
    const double bb = s - q;
    const double r = (q - (s - bb)) + (ct - bb);
    if (r != 0.0) e.c[m++] = r;
    q = s;
  }
  if (q != 0.0) e.c[m++] = q;
  e.n = m;
}

__device__ __forceinline__ int exp_sign(const Exp &e) {
  return e.n == 0 ? 0 : (e.c[e.n - 1] > 0.0 ? 1 : -1);
}

// x*y*z for float32-valued doubles, added exactly.
__device__ __forceinline__ void exp_add3(Exp &e, double x, double y, double z) {
  const double xy = x * y;  // 48 significant bits: exact
  const double hi = xy * z;
  exp_add(e, hi);
  exp_add(e, __builtin_fma(xy, z, -hi));  // exact remainder (fused)
}

// sign * det4 | r0 1 ; r1 1 ; r2 1 ; r3 1 | accumulated into e.
// Terms: for every ordered choice (rx, ry, rz) of distinct rows, the sign of
// the permutation (rx, ry, rz, remaining row) times rx.x * ry.y * rz.z.
__device__ __forceinline__ void exp_det4(Exp &e, const float *r0, const float *r1,
                                         const float *r2, const float *r3, int sign) {
  const float *rows[4] = {r0, r1, r2, r3};
#pragma unroll 1
  for (int rx = 0; rx < 4; ++rx) {
#pragma unroll 1
    for (int ry = 0; ry < 4; ++ry) {
      if (ry == rx) continue;
#pragma unroll 1
      for (int rz = 0; rz < 4; ++rz) {
        if (rz == rx || rz == ry) continue;
        const int rw = 6 - rx - ry - rz;
        // parity of (rx, ry, rz, rw) by counting inversions
        int inv = (rx > ry) + (rx > rz) + (rx > rw) + (ry > rz) + (ry > rw) + (rz > rw);
        const int s = ((inv & 1) ? -1 : 1) * sign;
        const double x = (double)rows[rx][0];
        exp_add3(e, s > 0 ? x : -x, (double)rows[ry][1], (double)rows[rz][2]);
      }
    }
  }
}

// sign * det3 | a.u a.v 1 ; b.u b.v 1 ; c.u c.v 1 | (u, v) = coordinates
// (k==0: y,z  k==1: z,x  k==2: x,y) accumulated into e.
__device__ __forceinline__ void exp_det3(Exp &e, const float *a, const float *b,
                                         const float *c, int k, int sign) {
  const int u = k == 0 ? 1 : (k == 1 ? 2 : 0);
  const int v = k == 0 ? 2 : (k == 1 ? 0 : 1);
  const float *rows[3] = {a, b, c};
#pragma unroll 1
  for (int ru = 0; ru < 3; ++ru) {
#pragma unroll 1
    for (int rv = 0; rv < 3; ++rv) {
      if (rv == ru) continue;
      const int rw = 3 - ru - rv;
      const int inv = (ru > rv) + (ru > rw) + (rv > rw);
      const int s = ((inv & 1) ? -1 : 1) * sign;
      const double x = (double)rows[ru][u];
      exp_add(e, (s > 0 ? x : -x) * (double)rows[rv][v]);  // exact product
    }
  }
}

// ------------------------------------------------------------ integer path
constexpr int kGridWindow = 15;
using i128 = __int128;

// f[i] = m[i] * 2^emin with one emin for all; false when the nonzero values'
// exponents span more than kGridWindow.
template <int N>
__device__ __forceinline__ bool int_grid(const float (&f)[N], int64_t (&m)[N]) {
  int e[N];
  int lo = 1 << 20, hi = -(1 << 20);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const unsigned b = __builtin_bit_cast(unsigned, f[i]);
    const int be = (int)((b >> 23) & 0xffu);
    const int mm = (int)(b & 0x7fffffu) | (be ? 0x800000 : 0);
    m[i] = (b >> 31) ? -(int64_t)mm : (int64_t)mm;
    e[i] = be ? be - 150 : -149;
    if (mm) {
      lo = e[i] < lo ? e[i] : lo;
      hi = e[i] > hi ? e[i] : hi;
    }
  }
  if (hi - lo > kGridWindow) return false;
#pragma unroll
  for (int i = 0; i < N; ++i) m[i] *= (int64_t)1 << (m[i] ? e[i] - lo : 0);
  return true;
}

__device__ __forceinline__ int sign128(i128 v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }

// det[a-d; b-d; c-d] on grid integers (a, b, c, d at m[0..2], m[3..5], ...).
__device__ __forceinline__ i128 det3_int(const int64_t *m) {
  const int64_t ax = m[0] - m[9], ay = m[1] - m[10], az = m[2] - m[11];
  const int64_t bx = m[3] - m[9], by = m[4] - m[10], bz = m[5] - m[11];
  const int64_t cx = m[6] - m[9], cy = m[7] - m[10], cz = m[8] - m[11];
  return (i128)ax * ((i128)by * cz - (i128)bz * cy) + (i128)bx * ((i128)cy * az - (i128)cz * ay) +
         (i128)cx * ((i128)ay * bz - (i128)az * by);
}

// (a_u - c_u)(b_v - c_v) - (a_v - c_v)(b_u - c_u) on grid integers
// (a, b, c at m[0..1], m[2..3], m[4..5]).
__device__ __forceinline__ i128 det2_int(const int64_t *m) {
  return (i128)(m[0] - m[4]) * (m[3] - m[5]) - (i128)(m[1] - m[5]) * (m[2] - m[4]);
}

__device__ __forceinline__ int proj_u(int k) { return k == 0 ? 1 : (k == 1 ? 2 : 0); }
__device__ __forceinline__ int proj_v(int k) { return k == 0 ? 2 : (k == 1 ? 0 : 1); }

// ------------------------------------------------------------ float64 filters
// det[a-d, b-d, c-d] (orient3d), float64, with its error bound.
__device__ __forceinline__ double o3_fast(const float *a, const float *b, const float *c,
                                          const float *d, double &bound) {
  const double adx = (double)a[0] - d[0], ady = (double)a[1] - d[1], adz = (double)a[2] - d[2];
  const double bdx = (double)b[0] - d[0], bdy = (double)b[1] - d[1], bdz = (double)b[2] - d[2];
  const double cdx = (double)c[0] - d[0], cdy = (double)c[1] - d[1], cdz = (double)c[2] - d[2];
  const double det = adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) +
                     cdx * (ady * bdz - adz * bdy);
  const double perm = fabs(adx) * (fabs(bdy * cdz) + fabs(bdz * cdy)) +
                      fabs(bdx) * (fabs(cdy * adz) + fabs(cdz * ady)) +
                      fabs(cdx) * (fabs(ady * bdz) + fabs(adz * bdy));
  bound = kFilter * perm;
  return det;
}

// orient2d(a, b, c) in projection k, float64, with its error bound.
__device__ __forceinline__ double o2_fast(const float *a, const float *b, const float *c, int k,
                                          double &bound) {
  const int u = k == 0 ? 1 : (k == 1 ? 2 : 0);
  const int v = k == 0 ? 2 : (k == 1 ? 0 : 1);
  const double acu = (double)a[u] - c[u], acv = (double)a[v] - c[v];
  const double bcu = (double)b[u] - c[u], bcv = (double)b[v] - c[v];
  bound = kFilter * (fabs(acu * bcv) + fabs(acv * bcu));
  return acu * bcv - acv * bcu;
}

__device__ __forceinline__ int sgn_or_zero(double v, double bound, bool &certain) {
  certain = v > bound || -v > bound;
  return v > 0.0 ? 1 : (v < 0.0 ? -1 : 0);
}

// ------------------------------------------------------------ exact signs
// The expansion forms (rare: inputs of very different magnitude) are
// separate, out-of-line functions so that the filters and the integer path
// inline into their callers.
__device__ __noinline__ int orient3d_expansion(const float *a, const float *b, const float *c,
                                               const float *d) {
  Exp e;
  e.n = 0;
  exp_det4(e, a, b, c, d, 1);
  return exp_sign(e);
}

__device__ __noinline__ int orient3d_mid_expansion(const float *a, const float *b, const float *c,
                                                   const float *p, const float *q) {
  Exp e;
  e.n = 0;
  exp_det4(e, a, b, c, p, 1);
  exp_det4(e, a, b, c, q, 1);
  return exp_sign(e);
}

__device__ __noinline__ int orient2d_expansion(const float *a, const float *b, const float *c,
                                               int k) {
  Exp e;
  e.n = 0;
  exp_det3(e, a, b, c, k, 1);
  return exp_sign(e);
}

__device__ __noinline__ int orient2d_mid_expansion(const float *a, const float *b, const float *p,
                                                   const float *q, int k) {
  Exp e;
  e.n = 0;
  exp_det3(e, a, b, p, k, 1);
  exp_det3(e, a, b, q, k, 1);
  return exp_sign(e);
}

// Integer paths: false when the inputs' exponents are too far apart for
// the grid; else the exact sign in *sign.
__device__ __forceinline__ bool orient3d_int(const float *a, const float *b, const float *c,
                                             const float *d, int *sign) {
  const float f[12] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], d[0], d[1], d[2]};
  int64_t m[12];
  if (!int_grid(f, m)) return false;
  *sign = sign128(det3_int(m));
  return true;
}

__device__ __forceinline__ bool orient3d_mid_int(const float *a, const float *b, const float *c,
                                                 const float *p, const float *q, int *sign) {
  const float f[15] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2],
                       p[0], p[1], p[2], q[0], q[1], q[2]};
  int64_t m[15];
  if (!int_grid(f, m)) return false;
  const i128 dp = det3_int(m);
  m[9] = m[12], m[10] = m[13], m[11] = m[14];
  *sign = sign128(dp + det3_int(m));
  return true;
}

__device__ __forceinline__ bool orient2d_int(const float *a, const float *b, const float *c, int k,
                                             int *sign) {
  const int u = proj_u(k), v = proj_v(k);
  const float f[6] = {a[u], a[v], b[u], b[v], c[u], c[v]};
  int64_t m[6];
  if (!int_grid(f, m)) return false;
  *sign = sign128(det2_int(m));
  return true;
}

__device__ __forceinline__ bool orient2d_mid_int(const float *a, const float *b, const float *p,
                                                 const float *q, int k, int *sign) {
  const int u = proj_u(k), v = proj_v(k);
  const float f[8] = {a[u], a[v], b[u], b[v], p[u], p[v], q[u], q[v]};
  int64_t m[8];
  if (!int_grid(f, m)) return false;
  const i128 dp = det2_int(m);
  m[4] = m[6], m[5] = m[7];
  *sign = sign128(dp + det2_int(m));
  return true;
}

__device__ __forceinline__ int orient3d_exact(const float *a, const float *b, const float *c,
                                              const float *d) {
  double bound;
  const double det = o3_fast(a, b, c, d, bound);
  if (det > bound) return 1;
  if (-det > bound) return -1;
  int s;
  if (orient3d_int(a, b, c, d, &s)) return s;
  return orient3d_expansion(a, b, c, d);
}

// sign of orient3d(a, b, c, (p+q)/2) = sign(orient3d(.., p) + orient3d(.., q)).
__device__ __forceinline__ int orient3d_mid_exact(const float *a, const float *b, const float *c,
                                                  const float *p, const float *q) {
  double b1, b2;
  const double d1 = o3_fast(a, b, c, p, b1);
  const double d2 = o3_fast(a, b, c, q, b2);
  const double s = d1 + d2;
  const double bound = (b1 + b2) * 1.0000001 + 1e-15 * fabs(s);
  if (s > bound) return 1;
  if (-s > bound) return -1;
  int r;
  if (orient3d_mid_int(a, b, c, p, q, &r)) return r;
  return orient3d_mid_expansion(a, b, c, p, q);
}

__device__ __forceinline__ int orient2d_exact(const float *a, const float *b, const float *c,
                                              int k) {
  double bound;
  const double det = o2_fast(a, b, c, k, bound);
  if (det > bound) return 1;
  if (-det > bound) return -1;
  int s;
  if (orient2d_int(a, b, c, k, &s)) return s;
  return orient2d_expansion(a, b, c, k);
}

// sign of orient2d(a, b, (p+q)/2) in projection k.
__device__ __forceinline__ int orient2d_mid_exact(const float *a, const float *b, const float *p,
                                                  const float *q, int k) {
  double b1, b2;
  const double d1 = o2_fast(a, b, p, k, b1);
  const double d2 = o2_fast(a, b, q, k, b2);
  const double s = d1 + d2;
  const double bound = (b1 + b2) * 1.0000001 + 1e-15 * fabs(s);
  if (s > bound) return 1;
  if (-s > bound) return -1;
  int r;
  if (orient2d_mid_int(a, b, p, q, k, &r)) return r;
  return orient2d_mid_expansion(a, b, p, q, k);
}

// ------------------------------------------------- exact contact / crossing
__device__ __forceinline__ bool same3(const float *a, const float *b) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// A projection with exact nonzero orient2d(a, b, c); -1 for collinear a, b, c.
__device__ __noinline__ int proj_axis_exact(const float *a, const float *b, const float *c) {
  const double e1x = (double)b[0] - a[0], e1y = (double)b[1] - a[1], e1z = (double)b[2] - a[2];
  const double e2x = (double)c[0] - a[0], e2y = (double)c[1] - a[1], e2z = (double)c[2] - a[2];
  const double n0 = fabs(e1y * e2z - e1z * e2y);
  const double n1 = fabs(e1z * e2x - e1x * e2z);
  const double n2 = fabs(e1x * e2y - e1y * e2x);
  int first = 0;
  if (n1 > n0 && n1 >= n2) first = 1;
  else if (n2 > n0 && n2 > n1) first = 2;
  for (int t = 0; t < 3; ++t) {
    const int k = (first + t) % 3;
    if (orient2d_exact(a, b, c, k) != 0) return k;
  }
  return -1;
}

__device__ __forceinline__ bool in_tri_closed(const float *a, const float *b, const float *c,
                                              const float *x, int k) {
  const int s = orient2d_exact(a, b, c, k);
  return s * orient2d_exact(a, b, x, k) >= 0 && s * orient2d_exact(b, c, x, k) >= 0 &&
         s * orient2d_exact(c, a, x, k) >= 0;
}

__device__ __forceinline__ bool between_collinear(const float *p, const float *q, const float *x,
                                                  int k) {
  const int u = k == 0 ? 1 : (k == 1 ? 2 : 0);
  const int v = k == 0 ? 2 : (k == 1 ? 0 : 1);
  return fminf(p[u], q[u]) <= x[u] && x[u] <= fmaxf(p[u], q[u]) && fminf(p[v], q[v]) <= x[v] &&
         x[v] <= fmaxf(p[v], q[v]);
}

__device__ __noinline__ bool segs_touch(const float *p, const float *q, const float *a,
                                        const float *b, int k) {
  const int o1 = orient2d_exact(p, q, a, k), o2 = orient2d_exact(p, q, b, k);
  const int o3 = orient2d_exact(a, b, p, k), o4 = orient2d_exact(a, b, q, k);
  if (o1 * o2 < 0 && o3 * o4 < 0) return true;
  return (o1 == 0 && between_collinear(p, q, a, k)) || (o2 == 0 && between_collinear(p, q, b, k)) ||
         (o3 == 0 && between_collinear(a, b, p, k)) || (o4 == 0 && between_collinear(a, b, q, k));
}

// Closed segment pq and closed triangle (t0, t1, t2) share a point other than
// an endpoint that is one of the triangle's vertices (segment == triangle
// edge counts as contact; collinear triangles are never touched).
__device__ __noinline__ bool contact_exact(const float *p, const float *q, const float *t0,
                                           const float *t1, const float *t2) {
  const float *tv[3] = {t0, t1, t2};
  int ip = -1, iq = -1;
  for (int i = 0; i < 3; ++i) {
    if (ip < 0 && same3(p, tv[i])) ip = i;
    if (iq < 0 && same3(q, tv[i])) iq = i;
  }
  if (ip >= 0 && iq >= 0) return true;
  if (ip >= 0 || iq >= 0) {
    const int iv = ip >= 0 ? ip : iq;
    const float *o = ip >= 0 ? q : p;
    if (orient3d_exact(t0, t1, t2, o) != 0) return false;
    const int k = proj_axis_exact(t0, t1, t2);
    if (k < 0) return false;
    const float *v0 = tv[iv], *v1 = tv[(iv + 1) % 3], *v2 = tv[(iv + 2) % 3];
    const int s = orient2d_exact(v0, v1, v2, k);
    return s * orient2d_exact(v0, v1, o, k) >= 0 && s * orient2d_exact(v0, v2, o, k) <= 0;
  }
  const int op = orient3d_exact(t0, t1, t2, p), oq = orient3d_exact(t0, t1, t2, q);
  if (op != 0 && op == oq) return false;
  if (op != 0 && oq != 0) {
    const int s1 = orient3d_exact(p, q, t0, t1);
    const int s2 = orient3d_exact(p, q, t1, t2);
    const int s3 = orient3d_exact(p, q, t2, t0);
    const bool pos = s1 > 0 || s2 > 0 || s3 > 0, neg = s1 < 0 || s2 < 0 || s3 < 0;
    return !(pos && neg);
  }
  const int k = proj_axis_exact(t0, t1, t2);
  if (k < 0) return false;
  if (op == 0 && oq != 0) return in_tri_closed(t0, t1, t2, p, k);
  if (oq == 0 && op != 0) return in_tri_closed(t0, t1, t2, q, k);
  return in_tri_closed(t0, t1, t2, p, k) || in_tri_closed(t0, t1, t2, q, k) ||
         segs_touch(p, q, t0, t1, k) || segs_touch(p, q, t1, t2, k) || segs_touch(p, q, t2, t0, k);
}

// orient2d in (y, z) of (a, b, m + (0, e, e^2)), m = (p+q)/2.
__device__ __forceinline__ int sos_yz(const float *a, const float *b, const float *p,
                                      const float *q) {
  const int s = orient2d_mid_exact(a, b, p, q, 0);
  if (s != 0) return s;
  if (a[2] != b[2]) return a[2] > b[2] ? 1 : -1;
  if (a[1] != b[1]) return b[1] > a[1] ? 1 : -1;
  return 0;
}

// Ray (p+q)/2 + t(1,0,0), t > 0, perturbed by (0, e, e^2), crosses triangle.
__device__ __noinline__ bool ray_cross_exact(const float *p, const float *q, const float *t0,
                                             const float *t1, const float *t2) {
  const int s1 = sos_yz(t0, t1, p, q);
  if (s1 == 0) return false;
  if (sos_yz(t1, t2, p, q) != s1) return false;
  if (sos_yz(t2, t0, p, q) != s1) return false;
  return orient3d_mid_exact(t0, t1, t2, p, q) == s1;
}

// Unperturbed +x ray from p against a triangle, exactly: 1 crossing, 0 none.
// Sets degenerate (and returns 0) when the answer would depend on how the
// ray is perturbed: p projects onto the triangle's boundary in (y, z) while
// the other edge tests do not exclude it, or p lies in the triangle's plane
// while projecting inside it.
__device__ __noinline__ int ray_cross_point_exact(const float *p, const float *t0,
                                                  const float *t1, const float *t2,
                                                  bool &degenerate) {
  const int s1 = orient2d_mid_exact(t0, t1, p, p, 0);
  const int s2 = orient2d_mid_exact(t1, t2, p, p, 0);
  const int s3 = orient2d_mid_exact(t2, t0, p, p, 0);
  const bool pos = s1 > 0 || s2 > 0 || s3 > 0, neg = s1 < 0 || s2 < 0 || s3 < 0;
  if (pos && neg) return 0;  // outside in projection under any perturbation
  if (s1 == 0 || s2 == 0 || s3 == 0) {
    degenerate = true;
    return 0;
  }
  const int o = orient3d_mid_exact(t0, t1, t2, p, p);
  if (o == 0) {
    degenerate = true;
    return 0;
  }
  return o == s1 ? 1 : 0;
}

}  // namespace vacd
