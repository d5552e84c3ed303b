// C ABI of libvacd (declared in include/vacd.h).
//
// Error convention: every entry point catches vacd::Error, stores its message
// on the context (or in a thread-local slot when there is none) and returns
// its code; the pybind layer turns a non-zero code into RuntimeError, where
// the reference threw std::runtime_error from CUCHK/OCHK
// (lib/include/optixUtils.hpp:56-71).
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

using vacd::Error;

namespace {

thread_local std::string g_last_error;

int fail(vacd_ctx *ctx, const Error &e) {
  if (ctx) ctx->last_error = e.msg;
  g_last_error = e.msg;
  return e.code;
}

template <class F>
int guarded(vacd_ctx *ctx, F &&f) {
  try {
    f();
    return VACD_OK;
  } catch (const Error &e) {
    return fail(ctx, e);
  } catch (const std::bad_alloc &) {
    return fail(ctx, Error{VACD_ERR_HIP, "host allocation failed"});
  } catch (...) {
    return fail(ctx, Error{VACD_ERR_HIP, "unknown failure"});
  }
}

void set_device(const vacd_ctx *ctx) { VACD_HIP(hipSetDevice(ctx->device)); }

template <class T>
T *stage(vacd_ctx *ctx, int slot, const T *host, size_t count) {
  vacd::DevBuf &b = ctx->buf[slot];
  b.ensure(sizeof(T) * (count ? count : 1));
  if (count)
    VACD_HIP(hipMemcpyAsync(b.ptr, host, sizeof(T) * count, hipMemcpyHostToDevice, ctx->stream));
  return b.as<T>();
}

bool aligned(const void *p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

}  // namespace

extern "C" {

int vacd_ctx_create(int device, vacd_ctx **out) {
  if (!out) return fail(nullptr, Error{VACD_ERR_INVALID, "out is NULL"});
  *out = nullptr;
  vacd_ctx *ctx = nullptr;
  int rc = guarded(nullptr, [&] {
    int ndev = 0;
    VACD_HIP(hipGetDeviceCount(&ndev));
    VACD_REQUIRE(device >= 0 && device < ndev, VACD_ERR_INVALID, "no such HIP device");
    ctx = new vacd_ctx();
    ctx->device = device;
    VACD_HIP(hipSetDevice(device));
    VACD_HIP(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
    ctx->stream = ctx->own_stream;
    // once per device, with it current: the scratch address of the
    // allocation-free scoring entry point
    vacd::noalloc_setup(device);
  });
  if (rc != VACD_OK) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return VACD_OK;
}

int vacd_ctx_destroy(vacd_ctx *ctx) {
  if (!ctx) return VACD_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &b : ctx->buf) b.release();
  for (auto &v : ctx->ev)
    for (auto &pr : v) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  for (int k = 0; k < 3; ++k) {
    if (ctx->pinned[k]) (void)hipHostFree(ctx->pinned[k]);
    if (ctx->pinned_free[k]) (void)hipEventDestroy(ctx->pinned_free[k]);
  }
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return VACD_OK;
}

const char *vacd_last_error(const vacd_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : g_last_error.c_str();
}

int vacd_ctx_stream(vacd_ctx *ctx, void **stream) {
  if (!ctx || !stream) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  *stream = (void *)ctx->stream;
  return VACD_OK;
}

int vacd_ctx_set_stream(vacd_ctx *ctx, void *stream) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
  return VACD_OK;
}

int vacd_sync(vacd_ctx *ctx) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    set_device(ctx);
    VACD_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int vacd_ctx_set_option(vacd_ctx *ctx, int option, int64_t value) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  if (option == VACD_OPT_TIMING)
    ctx->timing = value != 0;
  else if (option == VACD_OPT_COUNT_WORK)
    ctx->count_work = value != 0;
  else if (option == VACD_OPT_FRONT_LANES && (value == 0 || value == 4 || value == 8 || value == 16 || value == 32 || value == 64))
    ctx->front_lanes = (int)value;
  else if (option == VACD_OPT_FRONT_MAX && value >= 0)
    ctx->front_max = value;
  else if (option == VACD_OPT_FAN_PAIRS)
    ctx->fan_pairs = value != 0;
  else if (option == VACD_OPT_CLIP_ENDS && value >= 0)
    ctx->clip_min = value;
  else if (option == VACD_OPT_ROTATED_RAYS)
    ctx->rotated_rays = value != 0;
  else if (option == VACD_OPT_SCORE_CHUNK && value >= 0 && value <= (int64_t(1) << 30))
    ctx->score_chunk = value;
  else if (option == VACD_OPT_LANES_PER_PAIR &&
           (value == 1 || value == 2 || value == 4 || value == 8 || value == 16))
    ctx->lanes_per_pair_log2 = value == 1 ? 0 : value == 2 ? 1 : value == 4 ? 2 : value == 8 ? 3 : 4;
  else
    return fail(ctx, Error{VACD_ERR_INVALID, "unknown option"});
  return VACD_OK;
}

int vacd_ctx_timing(vacd_ctx *ctx, vacd_timing *out, int reset) {
  if (!ctx || !out) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    set_device(ctx);
    VACD_HIP(hipStreamSynchronize(ctx->stream));
    double ms[3] = {0, 0, 0};
    int64_t cnt[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
      for (auto &pr : ctx->ev[k]) {
        float t = 0.0f;
        VACD_HIP(hipEventElapsedTime(&t, pr.first, pr.second));
        ms[k] += t;
        ++cnt[k];
      }
      if (reset) {
        for (auto &pr : ctx->ev[k]) {
          (void)hipEventDestroy(pr.first);
          (void)hipEventDestroy(pr.second);
        }
        ctx->ev[k].clear();
      }
    }
    out->vis_ms = ms[0];
    out->vis_launches = cnt[0];
    out->score_ms = ms[1];
    out->score_launches = cnt[1];
    out->bvh_ms = ms[2];
    out->bvh_launches = cnt[2];
  });
}

// ------------------------------------------------------------------ BVH
// Planar group codes of the triangles (0 none, g + 1) and the group table
// (float4 pairs (n, d), (tau, 0, 0, 0)) from host arrays; groups whose
// triangles do not lie within 1e-4 of the extent of their plane are dropped.
// Returns the number of triangles left in a group.
// Axis dropped by the 2-D projection of a plane with normal (a, b, c): the
// largest |component|, ties to the lower axis.
static int plane_axis(const double *p) {
  const double a = std::fabs(p[0]), b = std::fabs(p[1]), c = std::fabs(p[2]);
  int k = b > a ? 1 : 0;
  if (c > std::max(a, b)) k = 2;
  return k;
}

static int64_t planar_groups(const float *V, int64_t nV, const uint32_t *T, int64_t nT,
                             const int8_t *tri_group, const double *planes, int n_groups,
                             std::vector<uint8_t> &code, std::vector<float> &table,
                             uint32_t &axis) {
  VACD_REQUIRE(n_groups >= 0 && n_groups <= 7, VACD_ERR_INVALID, "at most 7 planar groups");
  VACD_REQUIRE(n_groups == 0 || (tri_group && planes), VACD_ERR_INVALID, "NULL group arrays");
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t v = 0; v < nV; ++v)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], V[3 * v + k]), hi[k] = std::max(hi[k], V[3 * v + k]);
  double ext = 0.0;
  for (int k = 0; k < 3; ++k) ext = std::max(ext, (double)hi[k] - (double)lo[k]);
  std::vector<float> nf(4 * (size_t)n_groups);
  std::vector<double> tau((size_t)n_groups, 0.0);
  for (int g = 0; g < n_groups; ++g) {
    const double *p = planes + 4 * g;
    const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    VACD_REQUIRE(l > 0.0 && std::isfinite(l) && std::isfinite(p[3]), VACD_ERR_INVALID,
                 "planar group plane has no normal");
    for (int k = 0; k < 4; ++k) nf[4 * g + k] = (float)(p[k] / l);
  }
  for (int64_t t = 0; t < nT; ++t) {
    const int g = tri_group[t];
    VACD_REQUIRE(g >= -1 && g < n_groups, VACD_ERR_INVALID, "triangle group out of range");
    if (g < 0) continue;
    const float *n = &nf[4 * g];
    for (int c = 0; c < 3; ++c) {
      const float *x = V + 3 * (size_t)T[3 * t + c];
      const double r = (double)n[0] * x[0] + (double)n[1] * x[1] + (double)n[2] * x[2] + (double)n[3];
      tau[g] = std::max(tau[g], std::fabs(r));
    }
  }
  code.assign((size_t)nT, 0);
  table.assign(8 * (size_t)n_groups, 0.0f);
  int64_t kept = 0;
  for (int64_t t = 0; t < nT; ++t) {
    const int g = tri_group[t];
    if (g < 0 || tau[g] > 1e-4 * ext) continue;
    code[t] = (uint8_t)(g + 1);
    ++kept;
  }
  axis = 0xffffffffu;
  for (int g = 0; g < n_groups; ++g) {
    if (tau[g] <= 1e-4 * ext) {
      axis &= ~(3u << (2 * g));
      axis |= (uint32_t)plane_axis(planes + 4 * g) << (2 * g);
    }
    for (int k = 0; k < 4; ++k) table[8 * g + k] = nf[4 * g + k];
    // rounded up: the slab must hold every vertex of the group
    table[8 * g + 4] = std::nextafter((float)(tau[g] * (1.0 + 1e-6)), INFINITY);
  }
  return kept;
}

static int bvh_build_impl(vacd_ctx *ctx, const float *V, int64_t nV, const uint32_t *T, int64_t nT,
                          bool device_ptrs, vacd_bvh **out, const int8_t *tri_group = nullptr,
                          const double *planes = nullptr, int n_groups = 0) {
  if (!ctx || !out) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  *out = nullptr;
  vacd_bvh *b = nullptr;
  int rc = guarded(ctx, [&] {
    VACD_REQUIRE(V && T, VACD_ERR_INVALID, "NULL mesh arrays");
    VACD_REQUIRE(nV >= 1 && nT >= 1, VACD_ERR_INVALID, "empty mesh");
    VACD_REQUIRE(nT < (int64_t(1) << 30) && nV < (int64_t(1) << 32), VACD_ERR_LIMIT,
                 "mesh too large");
    set_device(ctx);
    const float *dV = V;
    const uint32_t *dT = T;
    if (!device_ptrs) {
      // indices are validated on the host copy
      for (int64_t t = 0; t < 3 * nT; ++t)
        VACD_REQUIRE((int64_t)T[t] < nV, VACD_ERR_INVALID, "triangle index out of range");
      dV = stage(ctx, SLOT_VERTS, V, (size_t)(3 * nV));
      dT = stage(ctx, SLOT_TRIS, T, (size_t)(3 * nT));
    }
    std::vector<uint8_t> code;
    std::vector<float> table;
    uint32_t axis = 0xffffffffu;
    const int64_t grouped =
        n_groups > 0 ? planar_groups(V, nV, T, nT, tri_group, planes, n_groups, code, table, axis)
                     : 0;
    b = new vacd_bvh();
    b->device = ctx->device;
    const uint8_t *dgrp = grouped > 0 ? stage(ctx, SLOT_GROUPS, code.data(), code.size()) : nullptr;
    vacd::build_lbvh(ctx, dV, nV, dT, nT, b, dgrp);
    if (dgrp) b->group_axis = axis;
    if (dgrp && b->n_refs > 0) {
      b->groups.ensure(4 * table.size());
      VACD_HIP(hipMemcpyAsync(b->groups.ptr, table.data(), 4 * table.size(), hipMemcpyHostToDevice,
                              ctx->stream));
      VACD_HIP(hipStreamSynchronize(ctx->stream));
      b->n_groups = n_groups;
    }
  });
  if (rc != VACD_OK) {
    if (b) {
      b->release();
      delete b;
    }
    return rc;
  }
  *out = b;
  return VACD_OK;
}

int vacd_bvh_build(vacd_ctx *ctx, const float *vertices, int64_t n_vertices, const uint32_t *tris,
                   int64_t n_tris, vacd_bvh **out) {
  return bvh_build_impl(ctx, vertices, n_vertices, tris, n_tris, false, out);
}

int vacd_bvh_build_planar(vacd_ctx *ctx, const float *vertices, int64_t n_vertices,
                          const uint32_t *tris, int64_t n_tris, const int8_t *tri_group,
                          const double *planes, int n_groups, vacd_bvh **out) {
  return bvh_build_impl(ctx, vertices, n_vertices, tris, n_tris, false, out, tri_group, planes,
                        n_groups);
}

int vacd_bvh_build_d(vacd_ctx *ctx, const float *d_vertices, int64_t n_vertices,
                     const uint32_t *d_tris, int64_t n_tris, vacd_bvh **out) {
  return bvh_build_impl(ctx, d_vertices, n_vertices, d_tris, n_tris, true, out);
}

int vacd_bvh_destroy(vacd_bvh *bvh) {
  if (!bvh) return VACD_OK;
  (void)hipSetDevice(bvh->device);
  bvh->release();
  delete bvh;
  return VACD_OK;
}

int vacd_bvh_info(const vacd_bvh *bvh, int64_t *n_tris, int32_t *depth, float *bounds6) {
  if (!bvh) return fail(nullptr, Error{VACD_ERR_INVALID, "NULL BVH"});
  if (n_tris) *n_tris = bvh->n_tris;
  if (depth) *depth = bvh->depth;
  if (bounds6)
    for (int k = 0; k < 6; ++k) bounds6[k] = bvh->bounds[k];
  return VACD_OK;
}

int vacd_bvh_contact_info(const vacd_bvh *bvh, int64_t *n_refs, int32_t *depth) {
  if (!bvh) return fail(nullptr, Error{VACD_ERR_INVALID, "NULL BVH"});
  if (n_refs) *n_refs = bvh->n_refs;
  if (depth) *depth = bvh->depth_refs;
  return VACD_OK;
}

int vacd_bvh_export(const vacd_bvh *bvh, void *nodes, void *tris) {
  if (!bvh) return fail(nullptr, Error{VACD_ERR_INVALID, "NULL BVH"});
  return guarded(nullptr, [&] {
    VACD_HIP(hipSetDevice(bvh->device));
    const int64_t n_int = bvh->n_tris > 1 ? bvh->n_tris - 1 : 1;
    if (nodes)
      VACD_HIP(hipMemcpy(nodes, bvh->nodes.ptr, sizeof(vacd::Node) * n_int, hipMemcpyDeviceToHost));
    if (tris)
      VACD_HIP(hipMemcpy(tris, bvh->tris.ptr, sizeof(vacd::Tri) * bvh->n_tris,
                         hipMemcpyDeviceToHost));
  });
}

// ------------------------------------------------------------ visibility
int vacd_visibility_d(vacd_ctx *ctx, const float *d_pts, int64_t n, const uint8_t *d_new_mask,
                      const vacd_bvh *self, const vacd_bvh *cage, uint32_t *d_pairs, int64_t cap,
                      int64_t *d_count) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(self && d_count, VACD_ERR_INVALID, "NULL argument");
    VACD_REQUIRE(n >= 0 && cap >= 0, VACD_ERR_INVALID, "negative size");
    VACD_REQUIRE(n < 2 || d_pts, VACD_ERR_INVALID, "NULL points");
    VACD_REQUIRE(cap == 0 || d_pairs, VACD_ERR_INVALID, "NULL pair buffer");
    VACD_REQUIRE(self->device == ctx->device && (!cage || cage->device == ctx->device),
                 VACD_ERR_INVALID, "BVH built on another device");
    set_device(ctx);
    uint32_t *pairs = d_pairs;
    if (cap > 0 && !aligned(d_pairs, 8)) {
      vacd::DevBuf &pb = ctx->buf[SLOT_PAIRS];
      pb.ensure(8 * (size_t)cap);
      pairs = pb.as<uint32_t>();
    }
    vacd::launch_visibility(ctx, d_pts, n, d_new_mask, self, cage, pairs, cap, d_count);
    if (pairs != d_pairs)
      VACD_HIP(hipMemcpyAsync(d_pairs, pairs, 8 * (size_t)cap, hipMemcpyDeviceToDevice,
                              ctx->stream));
  });
}

int vacd_visibility_batch_d(vacd_ctx *ctx, const vacd_part *parts, int32_t nparts,
                            uint32_t *d_pairs, int64_t cap, int64_t *d_count,
                            int64_t *d_part_off) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(parts && nparts > 0 && d_count && d_part_off, VACD_ERR_INVALID,
                 "NULL argument");
    VACD_REQUIRE(cap >= 0 && (cap == 0 || d_pairs), VACD_ERR_INVALID, "bad pair buffer");
    VACD_REQUIRE(cap == 0 || aligned(d_pairs, 8), VACD_ERR_INVALID,
                 "pair buffer must be 8-byte aligned");
    set_device(ctx);
    vacd::launch_visibility_batch(ctx, parts, nparts, d_pairs, cap, d_count, d_part_off);
  });
}

int vacd_visibility_batch(vacd_ctx *ctx, const vacd_part *parts, int32_t nparts,
                          uint32_t *pairs, int64_t cap, int64_t *count, int64_t *part_off) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  int64_t total = 0;
  int rc = guarded(ctx, [&] {
    VACD_REQUIRE(parts && nparts > 0 && count && part_off, VACD_ERR_INVALID, "NULL argument");
    VACD_REQUIRE(cap >= 0 && (cap == 0 || pairs), VACD_ERR_INVALID, "bad pair buffer");
    set_device(ctx);
    // points and masks of every part in one device buffer
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t bytes = 0;
    for (int p = 0; p < nparts; ++p) {
      VACD_REQUIRE(parts[p].n >= 0 && parts[p].n <= (int64_t(1) << 26), VACD_ERR_LIMIT,
                   "visibility needs n <= 2^26");
      VACD_REQUIRE(parts[p].n < 2 || parts[p].pts, VACD_ERR_INVALID, "NULL points");
      bytes += up(12 * (size_t)parts[p].n) + (parts[p].new_mask ? up((size_t)parts[p].n) : 0);
    }
    vacd::DevBuf &sb = ctx->buf[SLOT_PTS];
    sb.ensure(bytes + 256);
    std::vector<vacd_part> dev(parts, parts + nparts);
    size_t off = 0;
    for (int p = 0; p < nparts; ++p) {
      const vacd_part &q = parts[p];
      char *dp = sb.as<char>() + off;
      off += up(12 * (size_t)q.n);
      if (q.n) VACD_HIP(hipMemcpyAsync(dp, q.pts, 12 * (size_t)q.n, hipMemcpyHostToDevice, ctx->stream));
      dev[p].pts = reinterpret_cast<const float *>(dp);
      if (q.new_mask) {
        char *dm = sb.as<char>() + off;
        off += up((size_t)q.n);
        if (q.n) VACD_HIP(hipMemcpyAsync(dm, q.new_mask, (size_t)q.n, hipMemcpyHostToDevice, ctx->stream));
        dev[p].new_mask = reinterpret_cast<const uint8_t *>(dm);
      }
    }
    vacd::DevBuf &pb = ctx->buf[SLOT_PAIRS];
    pb.ensure(8 * (size_t)(cap > 0 ? cap : 1));
    vacd::DevBuf &cb = ctx->buf[SLOT_COUNT];
    cb.ensure(8 * (size_t)(nparts + 2));
    int64_t *dcount = cb.as<int64_t>(), *doff = dcount + 1;
    vacd::launch_visibility_batch(ctx, dev.data(), nparts, pb.as<uint32_t>(), cap, dcount, doff);
    unsigned long long bad = 0;
    VACD_HIP(hipMemcpyAsync(&total, dcount, 8, hipMemcpyDeviceToHost, ctx->stream));
    VACD_HIP(hipMemcpyAsync(part_off, doff, 8 * (size_t)(nparts + 1), hipMemcpyDeviceToHost,
                            ctx->stream));
    VACD_HIP(hipMemcpyAsync(&bad, ctx->buf[SLOT_STATS].as<unsigned long long>() + 4, sizeof(bad),
                            hipMemcpyDeviceToHost, ctx->stream));
    VACD_HIP(hipStreamSynchronize(ctx->stream));
    VACD_REQUIRE(bad == 0, VACD_ERR_HIP, "visibility: BVH walk exceeded its node budget");
    const int64_t ncopy = total < cap ? total : cap;
    if (ncopy > 0)
      VACD_HIP(hipMemcpy(pairs, pb.ptr, 8 * (size_t)ncopy, hipMemcpyDeviceToHost));
  });
  if (rc != VACD_OK) return rc;
  *count = total;
  if (total > cap)
    return fail(ctx, Error{VACD_ERR_CAPACITY, "visibility: more edges than the pair buffer holds"});
  return VACD_OK;
}

int vacd_visibility(vacd_ctx *ctx, const float *pts, int64_t n, const uint8_t *new_mask,
                    const vacd_bvh *self, const vacd_bvh *cage, uint32_t *pairs, int64_t cap,
                    int64_t *count) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  int64_t total = 0;
  int rc = guarded(ctx, [&] {
    VACD_REQUIRE(self && count, VACD_ERR_INVALID, "NULL argument");
    VACD_REQUIRE(n >= 0 && cap >= 0, VACD_ERR_INVALID, "negative size");
    VACD_REQUIRE(n < 2 || pts, VACD_ERR_INVALID, "NULL points");
    VACD_REQUIRE(cap == 0 || pairs, VACD_ERR_INVALID, "NULL pair buffer");
    VACD_REQUIRE(n <= (int64_t(1) << 26), VACD_ERR_LIMIT, "visibility needs n <= 2^26");
    set_device(ctx);
    const float *dp = stage(ctx, SLOT_PTS, pts, (size_t)(3 * n));
    const uint8_t *dm = new_mask ? stage(ctx, SLOT_MASK, new_mask, (size_t)n) : nullptr;
    vacd::DevBuf &pb = ctx->buf[SLOT_PAIRS];
    pb.ensure(8 * (size_t)(cap > 0 ? cap : 1));
    vacd::DevBuf &cb = ctx->buf[SLOT_COUNT];
    cb.ensure(sizeof(int64_t));
    vacd::launch_visibility(ctx, dp, n, dm, self, cage, pb.as<uint32_t>(), cap, cb.as<int64_t>());
    VACD_HIP(hipMemcpyAsync(&total, cb.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    unsigned long long bad = 0;
    if (n >= 2)
      VACD_HIP(hipMemcpyAsync(&bad, ctx->buf[SLOT_STATS].as<unsigned long long>() + 4,
                              sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
    VACD_HIP(hipStreamSynchronize(ctx->stream));
    VACD_REQUIRE(bad == 0, VACD_ERR_HIP, "visibility: BVH walk exceeded its node budget");
    const int64_t ncopy = total < cap ? total : cap;
    if (ncopy > 0)
      VACD_HIP(hipMemcpy(pairs, pb.ptr, 8 * (size_t)ncopy, hipMemcpyDeviceToHost));
  });
  if (rc != VACD_OK) return rc;
  *count = total;
  if (total > cap)
    return fail(ctx, Error{VACD_ERR_CAPACITY, "visibility: more edges than the pair buffer holds"});
  return VACD_OK;
}

int vacd_visibility_stats(vacd_ctx *ctx, vacd_vis_stats *out) {
  if (!ctx || !out) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    set_device(ctx);
    unsigned long long s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    VACD_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->buf[SLOT_STATS].ptr)
      VACD_HIP(hipMemcpy(s, ctx->buf[SLOT_STATS].ptr, sizeof(s), hipMemcpyDeviceToHost));
    out->segments = ctx->last_segments;
    out->masked = (int64_t)s[0];
    out->blocked = (int64_t)s[1];
    out->exact_calls = (int64_t)s[8];
    out->edges = (int64_t)s[3];
    out->node_visits = (int64_t)s[5];
    out->tri_tests = (int64_t)s[6];
    out->walked = (int64_t)s[7];
  });
}

int vacd_visibility_counters(vacd_ctx *ctx, int64_t *out, int32_t n) {
  if (!ctx || (n > 0 && !out)) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(n >= 0 && n <= 16, VACD_ERR_INVALID, "at most 16 counters");
    set_device(ctx);
    unsigned long long s[16] = {};
    VACD_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->buf[SLOT_STATS].ptr)
      VACD_HIP(hipMemcpy(s, ctx->buf[SLOT_STATS].ptr, sizeof(s), hipMemcpyDeviceToHost));
    for (int k = 0; k < n; ++k) out[k] = (int64_t)s[k];
  });
}

int vacd_visibility_deferred(vacd_ctx *ctx, int64_t *pair_ids, int64_t cap, int64_t *count) {
  if (!ctx || !count) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(cap >= 0 && (cap == 0 || pair_ids), VACD_ERR_INVALID, "bad id buffer");
    set_device(ctx);
    unsigned long long len = 0;
    VACD_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->buf[SLOT_STATS].ptr)
      VACD_HIP(hipMemcpy(&len, ctx->buf[SLOT_STATS].as<unsigned long long>() + 8, sizeof(len),
                         hipMemcpyDeviceToHost));
    *count = (int64_t)len;
    const int64_t n = (int64_t)len < cap ? (int64_t)len : cap;
    if (n > 0)
      VACD_HIP(hipMemcpy(pair_ids, ctx->buf[SLOT_DEFER].ptr, 8 * (size_t)n, hipMemcpyDeviceToHost));
    // bit 62 marks entries the breadth-first exact walk settled
    for (int64_t t = 0; t < n; ++t) pair_ids[t] &= ~(int64_t(1) << 62);
  });
}

// ------------------------------------------------------------ scoring
int vacd_score_planes_d(vacd_ctx *ctx, const float *d_planes, int64_t P, const float *d_pts,
                        int64_t n, const uint32_t *d_edges, int64_t E, int64_t *d_score_q32,
                        float *d_score_f, int32_t *d_argmax) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(P >= 0 && n >= 0 && E >= 0, VACD_ERR_INVALID, "negative size");
    VACD_REQUIRE(P == 0 || d_planes, VACD_ERR_INVALID, "NULL planes");
    VACD_REQUIRE(E == 0 || (d_edges && d_pts), VACD_ERR_INVALID, "NULL edges or points");
    set_device(ctx);
    const float *pl = d_planes;
    if (P > 0 && !aligned(d_planes, 16)) {
      vacd::DevBuf &b = ctx->buf[SLOT_PLANES];
      b.ensure(16 * (size_t)P);
      VACD_HIP(hipMemcpyAsync(b.ptr, d_planes, 16 * (size_t)P, hipMemcpyDeviceToDevice, ctx->stream));
      pl = b.as<float>();
    }
    const uint32_t *ed = d_edges;
    if (E > 0 && !aligned(d_edges, 8)) {
      vacd::DevBuf &b = ctx->buf[SLOT_EDGES];
      b.ensure(8 * (size_t)E);
      VACD_HIP(hipMemcpyAsync(b.ptr, d_edges, 8 * (size_t)E, hipMemcpyDeviceToDevice, ctx->stream));
      ed = b.as<uint32_t>();
    }
    vacd::launch_scoring(ctx, pl, P, d_pts, n, ed, E, d_score_q32, d_score_f, d_argmax);
  });
}

int vacd_score_planes_noalloc_d(vacd_ctx *ctx, const float *d_planes, int64_t P,
                                const float *d_pts, int64_t n, const uint32_t *d_edges, int64_t E,
                                int64_t *d_score_q32, float *d_score_f, int32_t *d_argmax) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(P >= 0 && n >= 0 && E >= 0, VACD_ERR_INVALID, "negative size");
    VACD_REQUIRE(P == 0 || d_planes, VACD_ERR_INVALID, "NULL planes");
    VACD_REQUIRE(E == 0 || (d_edges && d_pts), VACD_ERR_INVALID, "NULL edges or points");
    VACD_REQUIRE(P == 0 || aligned(d_planes, 16), VACD_ERR_INVALID,
                 "allocation-free scoring needs 16-byte aligned planes");
    VACD_REQUIRE(E == 0 || aligned(d_edges, 8), VACD_ERR_INVALID,
                 "allocation-free scoring needs 8-byte aligned edges");
    set_device(ctx);
    vacd::launch_scoring_noalloc(ctx, d_planes, P, d_pts, n, d_edges, E, d_score_q32, d_score_f,
                                 d_argmax);
  });
}

int vacd_score_planes_batch_d(vacd_ctx *ctx, const vacd_score_part *parts, int nparts,
                              int64_t *d_score_q32, int32_t *d_argmax) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(nparts == 0 || parts, VACD_ERR_INVALID, "NULL parts");
    set_device(ctx);
    vacd::launch_scoring_batch(ctx, parts, nparts, d_score_q32, d_argmax);
  });
}

int vacd_score_planes_batch(vacd_ctx *ctx, const vacd_score_part *parts, int nparts,
                            int64_t *score_q32, int32_t *argmax) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(nparts >= 0 && (nparts == 0 || (parts && argmax)), VACD_ERR_INVALID,
                 "NULL argument");
    if (nparts == 0) return;
    set_device(ctx);
    // one device buffer: planes | points | edges of every part, each aligned
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t bytes = 0;
    int64_t Ptot = 0;
    for (int p = 0; p < nparts; ++p) {
      const vacd_score_part &q = parts[p];
      VACD_REQUIRE(q.P >= 0 && q.n >= 0 && q.E >= 0, VACD_ERR_INVALID, "negative size");
      VACD_REQUIRE((q.P == 0 || q.planes) && (q.E == 0 || (q.pts && q.edges)), VACD_ERR_INVALID,
                   "NULL part arrays");
      for (int64_t e = 0; e < 2 * q.E; ++e)
        VACD_REQUIRE((int64_t)q.edges[e] < q.n, VACD_ERR_INVALID, "edge index out of range");
      bytes += up(16 * (size_t)q.P) + up(12 * (size_t)q.n) + up(8 * (size_t)q.E);
      Ptot += q.P;
    }
    vacd::DevBuf &b = ctx->buf[SLOT_EDGES];
    b.ensure(bytes + 256);
    char *base = b.as<char>();
    std::vector<vacd_score_part> dev(nparts);
    size_t off = 0;
    for (int p = 0; p < nparts; ++p) {
      const vacd_score_part &q = parts[p];
      dev[p] = q;
      char *pl = base + off;
      off += up(16 * (size_t)q.P);
      char *pt = base + off;
      off += up(12 * (size_t)q.n);
      char *ed = base + off;
      off += up(8 * (size_t)q.E);
      if (q.P) VACD_HIP(hipMemcpyAsync(pl, q.planes, 16 * (size_t)q.P, hipMemcpyHostToDevice, ctx->stream));
      if (q.n) VACD_HIP(hipMemcpyAsync(pt, q.pts, 12 * (size_t)q.n, hipMemcpyHostToDevice, ctx->stream));
      if (q.E) VACD_HIP(hipMemcpyAsync(ed, q.edges, 8 * (size_t)q.E, hipMemcpyHostToDevice, ctx->stream));
      dev[p].planes = reinterpret_cast<const float *>(pl);
      dev[p].pts = reinterpret_cast<const float *>(pt);
      dev[p].edges = reinterpret_cast<const uint32_t *>(ed);
    }
    vacd::DevBuf &qb = ctx->buf[SLOT_SCORES_F];
    qb.ensure(8 * (size_t)std::max<int64_t>(Ptot, 1) + 4 * (size_t)nparts + 256);
    int64_t *dq = qb.as<int64_t>();
    int32_t *da = reinterpret_cast<int32_t *>(qb.as<char>() + ((8 * (size_t)std::max<int64_t>(Ptot, 1) + 255) & ~size_t(255)));
    vacd::launch_scoring_batch(ctx, dev.data(), nparts, dq, da);
    if (score_q32 && Ptot)
      VACD_HIP(hipMemcpyAsync(score_q32, dq, 8 * (size_t)Ptot, hipMemcpyDeviceToHost, ctx->stream));
    VACD_HIP(hipMemcpyAsync(argmax, da, 4 * (size_t)nparts, hipMemcpyDeviceToHost, ctx->stream));
    VACD_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int vacd_gemm_bf16_d(vacd_ctx *ctx, int ta, int tb, int M, int N, int K, const void *d_A,
                     int lda, const void *d_B, int ldb, const float *d_bias, int relu, int acc,
                     float *d_C, int ldc) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(M >= 0 && N >= 0 && K >= 0 && d_A && d_B && d_C, VACD_ERR_INVALID,
                 "bad GEMM arguments");
    set_device(ctx);
    vacd::gemm_bf16(ctx, ta != 0, tb != 0, M, N, K, d_A, lda, d_B, ldb, d_bias, relu != 0,
                    acc != 0, d_C, nullptr, ldc);
  });
}

int64_t vacd_mlp_param_count(void) { return vacd::MlpLayoutCount(); }

int vacd_mlp_create(vacd_ctx *ctx, uint64_t seed, vacd_mlp **out) {
  if (!ctx || !out) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  *out = nullptr;
  return guarded(ctx, [&] {
    set_device(ctx);
    vacd::mlp_create(ctx, seed, out);
  });
}

int vacd_mlp_destroy(vacd_mlp *mlp) {
  vacd::mlp_destroy(mlp);
  return VACD_OK;
}

int vacd_mlp_set_params(vacd_ctx *ctx, vacd_mlp *mlp, const float *params) {
  if (!ctx || !mlp || !params) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    set_device(ctx);
    vacd::mlp_set(ctx, mlp, params);
  });
}

int vacd_mlp_get_params(vacd_ctx *ctx, vacd_mlp *mlp, float *params) {
  if (!ctx || !mlp || !params) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    set_device(ctx);
    vacd::mlp_get(ctx, mlp, params);
  });
}

int vacd_mlp_get_grads(vacd_ctx *ctx, vacd_mlp *mlp, float *grads) {
  if (!ctx || !mlp || !grads) return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    set_device(ctx);
    vacd::mlp_get_grads(ctx, mlp, grads);
  });
}

int vacd_mlp_forward_d(vacd_ctx *ctx, vacd_mlp *mlp, int B, int E, const float *d_feat, int P,
                       const float *d_planes, float *d_out) {
  if (!ctx || !mlp || !d_feat || !d_planes || !d_out)
    return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    set_device(ctx);
    vacd::mlp_forward(ctx, mlp, B, E, d_feat, P, d_planes, d_out);
  });
}

int vacd_mlp_forward(vacd_ctx *ctx, vacd_mlp *mlp, int B, int E, const float *feat, int P,
                     const float *planes, float *out) {
  if (!ctx || !mlp || !feat || !planes || !out)
    return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    set_device(ctx);
    vacd::mlp_forward_host(ctx, mlp, B, E, feat, P, planes, out);
  });
}

int vacd_mlp_train_step_d(vacd_ctx *ctx, vacd_mlp *mlp, int B, int E, const float *d_feat,
                          int P, const float *d_planes, const float *d_target, float lr,
                          float *loss) {
  if (!ctx || !mlp || !d_feat || !d_planes || !d_target || !loss)
    return fail(ctx, Error{VACD_ERR_INVALID, "NULL argument"});
  return guarded(ctx, [&] {
    set_device(ctx);
    *loss = vacd::mlp_train_step(ctx, mlp, B, E, d_feat, P, d_planes, d_target, lr);
  });
}

int vacd_score_planes(vacd_ctx *ctx, const float *planes, int64_t P, const float *pts, int64_t n,
                      const uint32_t *edges, int64_t E, int64_t *score_q32, float *score_f,
                      int32_t *argmax) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(P >= 0 && n >= 0 && E >= 0, VACD_ERR_INVALID, "negative size");
    VACD_REQUIRE(P == 0 || planes, VACD_ERR_INVALID, "NULL planes");
    VACD_REQUIRE(E == 0 || (edges && pts), VACD_ERR_INVALID, "NULL edges or points");
    set_device(ctx);
    const float *dpl = stage(ctx, SLOT_PLANES, planes, (size_t)(4 * P));
    const float *dpt = stage(ctx, SLOT_PTS, pts, (size_t)(3 * n));
    const uint32_t *ded = stage(ctx, SLOT_EDGES, edges, (size_t)(2 * E));
    vacd::DevBuf &q = ctx->buf[SLOT_SCORES_F];  // q32 then f32 view
    q.ensure(12 * (size_t)(P > 0 ? P : 1));
    int64_t *dq = q.as<int64_t>();
    float *df = reinterpret_cast<float *>(q.as<char>() + 8 * (size_t)(P > 0 ? P : 1));
    vacd::DevBuf &am = ctx->buf[SLOT_ARGMAX];
    am.ensure(sizeof(int32_t));
    vacd::launch_scoring(ctx, dpl, P, dpt, n, ded, E, dq, df, am.as<int32_t>());
    if (score_q32 && P)
      VACD_HIP(hipMemcpyAsync(score_q32, dq, 8 * (size_t)P, hipMemcpyDeviceToHost, ctx->stream));
    if (score_f && P)
      VACD_HIP(hipMemcpyAsync(score_f, df, 4 * (size_t)P, hipMemcpyDeviceToHost, ctx->stream));
    int32_t h_am = -1;
    if (P) VACD_HIP(hipMemcpyAsync(&h_am, am.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
    VACD_HIP(hipStreamSynchronize(ctx->stream));
    if (argmax) *argmax = h_am;
  });
}

static void check_distance_bvh(const vacd_ctx *ctx, const vacd_bvh *mesh) {
  VACD_REQUIRE(mesh != nullptr, VACD_ERR_INVALID, "NULL BVH");
  VACD_REQUIRE(mesh->device == ctx->device, VACD_ERR_INVALID, "BVH built on another device");
  VACD_REQUIRE(3 * (mesh->depth / 2) <= 64, VACD_ERR_LIMIT, "BVH too deep for the distance walk");
}

int vacd_max_distance_d(vacd_ctx *ctx, const float *d_pts, int64_t m, const vacd_bvh *mesh,
                        float *d_out) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(m >= 0 && (m == 0 || d_pts) && d_out, VACD_ERR_INVALID, "bad arguments");
    check_distance_bvh(ctx, mesh);
    set_device(ctx);
    vacd::launch_max_distance(ctx, d_pts, m, mesh, reinterpret_cast<unsigned *>(d_out));
  });
}

int vacd_max_distance(vacd_ctx *ctx, const float *pts, int64_t m, const vacd_bvh *mesh, float *out) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(m >= 0 && (m == 0 || pts) && out, VACD_ERR_INVALID, "bad arguments");
    check_distance_bvh(ctx, mesh);
    set_device(ctx);
    const float *dp = stage(ctx, SLOT_PTS, pts, (size_t)(3 * m));
    vacd::DevBuf &o = ctx->buf[SLOT_COUNT];
    o.ensure(sizeof(int64_t));
    vacd::launch_max_distance(ctx, dp, m, mesh, o.as<unsigned>());
    VACD_HIP(hipMemcpyAsync(out, o.ptr, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    VACD_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int vacd_max_distance_batch(vacd_ctx *ctx, const vacd_distance_query *q, int32_t nq, float *out) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(nq >= 0 && (nq == 0 || (q && out)), VACD_ERR_INVALID, "bad arguments");
    if (nq == 0) return;
    set_device(ctx);
    vacd::max_distance_batch(ctx, q, nq, out);
  });
}

int vacd_plane_extents_batch(vacd_ctx *ctx, const vacd_extent_query *q, int32_t nq, double *lo_hi) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(nq >= 0 && (nq == 0 || (q && lo_hi)), VACD_ERR_INVALID, "bad arguments");
    if (nq == 0) return;
    set_device(ctx);
    vacd::plane_extents_batch(ctx, q, nq, lo_hi);
  });
}

int vacd_sdf_remesh(vacd_ctx *ctx, const float *V, int64_t nV, const uint32_t *T, int64_t nT,
                    double scale, double level_set, float *V_out, int64_t v_cap, uint32_t *T_out,
                    int64_t t_cap, int64_t *nV_out, int64_t *nT_out) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  int64_t nv = 0, nt = 0;
  int rc = guarded(ctx, [&] {
    VACD_REQUIRE(V && T && nV_out && nT_out, VACD_ERR_INVALID, "NULL argument");
    VACD_REQUIRE(v_cap >= 0 && t_cap >= 0 && (v_cap == 0 || V_out) && (t_cap == 0 || T_out),
                 VACD_ERR_INVALID, "bad output buffers");
    set_device(ctx);
    std::vector<float> vo;
    std::vector<uint32_t> to;
    vacd::sdf_remesh(ctx, V, nV, T, nT, scale, level_set, vo, to, nullptr);
    nv = (int64_t)vo.size() / 3;
    nt = (int64_t)to.size() / 3;
    if (nv <= v_cap && nv) std::memcpy(V_out, vo.data(), 4 * vo.size());
    if (nt <= t_cap && nt) std::memcpy(T_out, to.data(), 4 * to.size());
  });
  if (rc != VACD_OK) return rc;
  *nV_out = nv;
  *nT_out = nt;
  if (nv > v_cap || nt > t_cap)
    return fail(ctx, Error{VACD_ERR_CAPACITY, "remesh: more vertices or triangles than the buffers hold"});
  return VACD_OK;
}

int vacd_sdf_grid(vacd_ctx *ctx, const float *V, int64_t nV, const uint32_t *T, int64_t nT,
                  const double *lo, double h, const int32_t *n, float *out, int32_t *sign_mode) {
  if (!ctx) return fail(ctx, Error{VACD_ERR_INVALID, "NULL context"});
  return guarded(ctx, [&] {
    VACD_REQUIRE(V && T && lo && n && out, VACD_ERR_INVALID, "NULL argument");
    set_device(ctx);
    const int nn[3] = {n[0], n[1], n[2]};
    int mode = 0;
    vacd::sdf_grid(ctx, V, nV, T, nT, lo, h, nn, out, &mode);
    if (sign_mode) *sign_mode = mode;
  });
}

void classify_and_rate_planes(const float *planes, const float *points, const unsigned int *edges,
                              float *scores, int numPlanes, int numPoints, int numEdges) {
  thread_local vacd_ctx *tctx = nullptr;
  if (!tctx && vacd_ctx_create(0, &tctx) != VACD_OK) {
    std::fprintf(stderr, "classify_and_rate_planes: %s\n", vacd_last_error(nullptr));
    return;
  }
  int rc = vacd_score_planes(tctx, planes, numPlanes, points, numPoints, edges, numEdges, nullptr,
                             scores, nullptr);
  if (rc != VACD_OK) std::fprintf(stderr, "classify_and_rate_planes: %s\n", vacd_last_error(tctx));
}

}  // extern "C"
