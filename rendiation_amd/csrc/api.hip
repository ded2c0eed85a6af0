// api.hip -- the C-ABI of librdrt (include/rdrt.h).
//
// Mirrors the reference provider (shader/ray-tracing/src/backend/
// wavefront_compute/geometry/naive/mod.rs:497-609): sources are copied at
// creation, handles are never reused, any change invalidates the built
// scene, and the build happens on rdrt_build / rdrt_commit instead of lazily
// on the next create_comp_instance.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"
#include "attrib.h"
#include "device_build.h"
#include "refit.h"

using namespace rdrt;

namespace {
thread_local std::string g_last_error;

TreeOptions tree_opt(uint32_t bin, uint32_t depth, uint32_t bins, uint32_t def_bin) {
  return TreeOptions{bin ? bin : def_bin, depth ? depth : 50u, bins ? bins : 4u};
}
}  // namespace

namespace rdrt {

int32_t fail(int32_t code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

void invalidate(rdrt_ctx *c) {
  c->built_valid = false;
  c->uploaded = false;
}

void free_device_scene(DeviceScene &d) {
  release(d.tri_pairs);
  release(d.tlas_pairs);
  release(d.tri_quads);
  release(d.tlas_quads);
  release(d.tri_qquads);
  release(d.tlas_qquads);
  release(d.tlas_quad_roots);
  release(d.inst_lookup);
  release(d.tlas_inst_first);
  release(d.uvs);
  release(d.tris);
  release(d.redirect);
  release(d.tri_ids);
  release(d.indices);
  release(d.binding);
  release(d.tri_material);
  release(d.anyhit);
  release(d.vertices);
  release(d.normals);
  release(d.geos);
  release(d.insts);
  release(d.inst_fwd);
  release(d.tlas_roots);
  release(d.tri_forest);
  release(d.tlas_forest);
  d.bytes = 0;
}

void release_refit(rdrt_ctx *c) {
  release(c->refit_forest);
  release(c->refit_pairs);
  release(c->refit_quads);
  c->refit_boxes.free();
  c->refit_ready = false;
}

// Shading attributes and / or hit context of closest hits (one kernel).
int32_t fetch_hit_data(rdrt_ctx *c, const rdrt_ray *rays, const rdrt_ray_meta *ray_meta, const rdrt_hit *hits,
                       const rdrt_hit_meta *hit_meta, uint32_t n, uint32_t tlas_idx, uint32_t sbt_offset,
                       uint32_t sbt_stride, rdrt_hit_attributes *out, rdrt_hit_transform *xf) {
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->uploaded) return fail(RDRT_ERR_NOT_BUILT, "scene is not committed (call rdrt_commit)");
  AttribArgs a = scene_attrib_args(c);
  a.rays = rays;
  a.meta = ray_meta;
  a.hits = hits;
  a.hit_meta = hit_meta;
  a.n = n;
  a.tlas_idx = tlas_idx;
  a.sbt_offset = sbt_offset;
  a.sbt_stride = sbt_stride;
  a.out = out;
  a.xf_out = xf;
  RDRT_HIP(enqueue_hit_attributes(a, c->stream));
  return RDRT_OK;
}

AttribArgs scene_attrib_args(const rdrt_ctx *c) {
  const DeviceScene &d = c->dev;
  AttribArgs a{};
  a.binding = d.binding.ptr;
  a.nbinding = (uint32_t)d.binding.count;
  a.ntlas = (uint32_t)d.tlas_inst_first.count;
  a.ninst = (uint32_t)d.inst_lookup.count;
  a.tlas_inst_first = d.tlas_inst_first.ptr;
  a.inst_lookup = d.inst_lookup.ptr;
  a.insts = d.insts.ptr;
  a.inst_fwd = d.inst_fwd.ptr;
  a.geos = d.geos.ptr;
  a.indices = d.indices.ptr;
  a.vertices = d.vertices.ptr;
  a.normals = d.normals.count ? d.normals.ptr : nullptr;
  a.uvs = d.uvs.count ? d.uvs.ptr : nullptr;
  return a;
}

int32_t need_device(rdrt_ctx *c) {
  if (!c->has_device) return fail(RDRT_ERR_NO_DEVICE, "context was created without a device (device = -1)");
  RDRT_HIP(hipSetDevice(c->cfg.device));
  return RDRT_OK;
}

TraceParams scene_trace_params(const rdrt_ctx *c) {
  const DeviceScene &d = c->dev;
  TraceParams p{};
  p.status = c->status;
  p.tri_pairs = d.tri_pairs.ptr;
  p.tris = d.tris.ptr;
  p.redirect = d.redirect.ptr;
  p.tri_ids = d.tri_ids.ptr;
  p.geos = d.geos.ptr;
  p.tlas_pairs = d.tlas_pairs.ptr;
  p.tri_quads = d.tri_quads.ptr;
  p.tlas_quads = d.tlas_quads.ptr;
  p.tri_qquads = d.tri_qquads.ptr;
  p.tlas_qquads = d.tlas_qquads.ptr;
  p.tlas_quad_roots = d.tlas_quad_roots.ptr;
  p.insts = d.insts.ptr;
  p.tlas_roots = d.tlas_roots.ptr;
  p.binding = d.binding.ptr;
  p.nbinding = (uint32_t)d.binding.count;
  p.ntlas = (uint32_t)d.tlas_roots.count;
  p.tri_forest = d.tri_forest.ptr;
  p.tlas_forest = d.tlas_forest.ptr;
  p.indices = d.indices.ptr;
  p.vertices = d.vertices.ptr;
  p.anyhit = d.anyhit.count ? d.anyhit.ptr : nullptr;
  return p;
}

uint32_t push_depth(const rdrt_ctx *c, const TraceTune &t) {
  return t.wide ? c->built.blas_stack4 + c->built.tlas_stack4 : c->built.blas_stack + c->built.tlas_stack;
}

namespace {

// Replaces a device array with a new host array, keeping the byte count.
template <class T>
hipError_t reupload(DevArray<T> &a, const std::vector<T> &h, uint64_t &bytes) {
  const uint64_t old = (uint64_t)(a.count ? a.count : 1) * sizeof(T);
  if (a.ptr) bytes -= std::min(bytes, old);
  release(a);
  return upload(a, h, bytes);
}

// Captured frames and batched calls hold the scene's device pointers.
void drop_captures(rdrt_ctx *c) {
  PtState &s = c->pt;
  if (s.exec) (void)hipGraphExecDestroy(s.exec);
  if (s.graph) (void)hipGraphDestroy(s.graph);
  s.exec = nullptr;
  s.graph = nullptr;
  s.have_cached = false;
  s.last_recorded = false;
  drop_ao_capture(c->ao);
  BatchState &bs = c->batch;
  if (bs.exec) (void)hipGraphExecDestroy(bs.exec);
  if (bs.graph) (void)hipGraphDestroy(bs.graph);
  bs.exec = nullptr;
  bs.graph = nullptr;
  bs.cached.clear();
}

}  // namespace

}  // namespace rdrt

extern "C" {

const char *rdrt_last_error(void) { return g_last_error.c_str(); }
int32_t rdrt_abi_version(void) { return RDRT_ABI_VERSION; }

int32_t rdrt_config_default(rdrt_config *out) {
  if (!out) return fail(RDRT_ERR_INVALID_ARG, "null config");
  std::memset(out, 0, sizeof(*out));
  out->device = 0;
  out->build = rdrt_build_opts{2, 50, 10, 50, 4};
  out->traversal = RDRT_TRAVERSAL_FAST;
  out->block_size = 0;
  return RDRT_OK;
}

int32_t rdrt_ctx_create(const rdrt_config *cfg, rdrt_ctx **out) {
  if (!cfg || !out) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (cfg->traversal > RDRT_TRAVERSAL_REFERENCE) return fail(RDRT_ERR_INVALID_ARG, "unknown traversal");
  if (cfg->block_size != 0 && cfg->block_size != 64 && cfg->block_size != 128 && cfg->block_size != 256)
    return fail(RDRT_ERR_INVALID_ARG, "block_size must be 0, 64, 128 or 256");
  rdrt_ctx *c = new rdrt_ctx();
  c->cfg = *cfg;
  c->tune.block = cfg->block_size ? cfg->block_size : 256;
  const rdrt_build_opts &b = cfg->build;
  c->src.blas_opt = tree_opt(b.blas_bin_size, b.blas_max_depth, b.sah_bins, 2);
  c->src.tlas_opt = tree_opt(b.tlas_bin_size, b.tlas_max_depth, b.sah_bins, 10);
  if (c->src.blas_opt.max_depth > 64 || c->src.tlas_opt.max_depth > 64) {
    delete c;
    return fail(RDRT_ERR_INVALID_ARG, "max_depth above 64 is not supported");
  }
  if (cfg->device >= 0) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || cfg->device >= n) {
      delete c;
      return fail(RDRT_ERR_NO_DEVICE, std::string("no HIP device ") + std::to_string(cfg->device) +
                                          (e != hipSuccess ? std::string(": ") + hipGetErrorString(e) : ""));
    }
    c->has_device = true;
    c->stream = (hipStream_t)cfg->stream;
    if (hipSetDevice(cfg->device) != hipSuccess || hipEventCreate(&c->ev0) != hipSuccess ||
        hipEventCreate(&c->ev1) != hipSuccess || hipMalloc(&c->status, sizeof(uint32_t)) != hipSuccess ||
        hipMemset(c->status, 0, sizeof(uint32_t)) != hipSuccess) {
      delete c;
      return fail(RDRT_ERR_HIP, "device set-up failed");
    }
    static_scratch_init(cfg->device);
    // Graph capture needs a real stream: the user's, or a blocking stream of
    // our own (ordered with the null stream) at the default priority.
    c->work_stream = c->stream;
    if (!c->work_stream) {
      if (hipStreamCreate(&c->work_stream) != hipSuccess) {
        delete c;
        return fail(RDRT_ERR_HIP, "stream creation failed");
      }
      c->owns_work_stream = true;
    }
  }
  *out = c;
  return RDRT_OK;
}

int32_t rdrt_ctx_destroy(rdrt_ctx *c) {
  if (!c) return RDRT_OK;
  if (c->has_device) {
    (void)hipSetDevice(c->cfg.device);
    (void)hipStreamSynchronize(c->stream);
    if (c->work_stream) (void)hipStreamSynchronize(c->work_stream);
    free_device_scene(c->dev);
    release_refit(c);
    c->pt.free_all();
    c->batch.free_all();
    free_ao(c->ao);
    c->compact_scratch.free();
    c->trace_work.free();
    c->trace_cull.free();
    if (c->owns_work_stream) (void)hipStreamDestroy(c->work_stream);
    if (c->status) (void)hipFree(c->status);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
  }
  delete c;
  return RDRT_OK;
}

int32_t rdrt_sync(rdrt_ctx *c) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  int32_t rc = need_device(c);
  if (rc) return rc;
  RDRT_HIP(hipStreamSynchronize(c->stream));
  if (c->work_stream != c->stream) RDRT_HIP(hipStreamSynchronize(c->work_stream));
  uint32_t st = 0;
  RDRT_HIP(hipMemcpy(&st, c->status, sizeof(st), hipMemcpyDeviceToHost));
  if (st) {
    RDRT_HIP(hipMemset(c->status, 0, sizeof(uint32_t)));
    return fail(RDRT_ERR_CAPACITY, "a traversal stack overflowed; results of the last launch are incomplete");
  }
  return RDRT_OK;
}

int32_t rdrt_blas_begin(rdrt_ctx *c) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  c->pending.clear();
  c->in_blas = true;
  return RDRT_OK;
}

int32_t rdrt_blas_add_triangles(rdrt_ctx *c, const float *pos, uint32_t nverts, const uint32_t *idx, uint32_t ntris,
                                uint32_t flags) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  if (!c->in_blas) return fail(RDRT_ERR_STATE, "rdrt_blas_add_triangles outside rdrt_blas_begin/end");
  // A rejected geometry aborts the whole pending BLAS.
  auto abort_blas = [c](const char *msg) {
    c->pending.clear();
    c->in_blas = false;
    return fail(RDRT_ERR_INVALID_ARG, msg);
  };
  if (nverts && !pos) return abort_blas("null positions");
  if (idx == nullptr && ntris != 0 && ntris != nverts / 3)
    return abort_blas("non-indexed geometry: ntris must be 0 or nverts/3");
  if (idx)
    for (size_t i = 0; i < 3 * (size_t)ntris; i++)
      if (idx[i] >= nverts) return abort_blas("triangle index out of range");
  GeometrySrc g;
  g.pos.assign(pos, pos + 3 * (size_t)nverts);
  g.indexed = idx != nullptr;
  if (idx) g.idx.assign(idx, idx + 3 * (size_t)ntris);
  g.ntris = idx ? ntris : nverts / 3;
  g.flags = flags;
  c->pending.push_back(std::move(g));
  return RDRT_OK;
}

int32_t rdrt_blas_add_triangles_u16(rdrt_ctx *c, const float *pos, uint32_t nverts, const uint16_t *idx,
                                    uint32_t ntris, uint32_t flags) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  if (!idx) return rdrt_blas_add_triangles(c, pos, nverts, nullptr, ntris, flags);
  std::vector<uint32_t> wide(idx, idx + 3 * (size_t)ntris);
  return rdrt_blas_add_triangles(c, pos, nverts, wide.data(), ntris, flags);
}

int32_t rdrt_blas_end(rdrt_ctx *c, uint32_t *out) {
  if (!c || !out) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (!c->in_blas) return fail(RDRT_ERR_STATE, "rdrt_blas_end without rdrt_blas_begin");
  c->src.blas.push_back(std::move(c->pending));
  c->src.blas_live.push_back(1);
  c->pending.clear();
  c->in_blas = false;
  invalidate(c);
  *out = (uint32_t)(c->src.blas.size() - 1);
  return RDRT_OK;
}

int32_t rdrt_blas_destroy(rdrt_ctx *c, uint32_t h) {
  if (!c || h >= c->src.blas.size()) return fail(RDRT_ERR_INVALID_ARG, "unknown BLAS handle");
  c->src.blas_live[h] = 0;
  c->src.blas[h].clear();
  invalidate(c);
  return RDRT_OK;
}

int32_t rdrt_tlas_create(rdrt_ctx *c, const rdrt_instance *inst, uint32_t n, uint32_t *out) {
  if (!c || !out || (n && !inst)) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  c->src.tlas.emplace_back(inst, inst + n);
  c->src.tlas_live.push_back(1);
  invalidate(c);
  *out = (uint32_t)(c->src.tlas.size() - 1);
  return RDRT_OK;
}

int32_t rdrt_tlas_destroy(rdrt_ctx *c, uint32_t h) {
  if (!c || h >= c->src.tlas.size()) return fail(RDRT_ERR_INVALID_ARG, "unknown TLAS handle");
  c->src.tlas_live[h] = 0;
  c->src.tlas[h].clear();
  invalidate(c);
  return RDRT_OK;
}

int32_t rdrt_bind_tlas(rdrt_ctx *c, const uint32_t *h, uint32_t n) {
  if (!c || (n && !h)) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  std::vector<uint32_t> b(h, h + n);
  if (b != c->src.binding) {  // naive/mod.rs:539-545
    c->src.binding = b;
    invalidate(c);
  }
  return RDRT_OK;
}

uint32_t rdrt_bind_tlas_max_len(const rdrt_ctx *) { return 0xffffffffu; }

int32_t rdrt_build(rdrt_ctx *c) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  if (c->built_valid) return RDRT_OK;
  std::string err;
  int32_t rc;
  if (c->device_build) {
    rc = need_device(c);
    if (rc) return rc;
    std::unique_ptr<DeviceTreeBuilder> trees = make_device_tree_builder(c->stream);
    rc = build_scene(c->src, c->built, err, trees.get(), true);
  } else {
    // with a device, the commit computes the triangle records and quantized
    // quads there instead of uploading host copies
    rc = build_scene(c->src, c->built, err, nullptr, c->has_device);
  }
  if (rc) return fail(rc, err);
  c->built_valid = true;
  c->uploaded = false;
  return RDRT_OK;
}

int32_t rdrt_pick(rdrt_ctx *c, uint32_t blas, uint32_t geometry, const float *rays, uint32_t n, uint32_t face_side,
                  rdrt_pick_hit *out) {
  if (!c || (n && (!rays || !out))) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = rdrt_build(c);
  if (rc) return rc;
  const BuiltScene &b = c->built;
  if (blas >= b.blas_geo_begin.size() || geometry >= b.blas_geo_end[blas] - b.blas_geo_begin[blas])
    return fail(RDRT_ERR_INVALID_ARG, "unknown BLAS or geometry");
  std::string err;
  rc = pick_nearest(b, b.blas_geo_begin[blas] + geometry, rays, n, face_side, out, err);
  if (rc) return fail(rc, err);
  return RDRT_OK;
}

int32_t rdrt_pick_all(rdrt_ctx *c, uint32_t blas, uint32_t geometry, const float *ray, uint32_t face_side,
                      rdrt_pick_hit *out, uint32_t capacity, uint32_t *count) {
  if (!c || !ray || !count || (capacity && !out)) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = rdrt_build(c);
  if (rc) return rc;
  const BuiltScene &b = c->built;
  if (blas >= b.blas_geo_begin.size() || geometry >= b.blas_geo_end[blas] - b.blas_geo_begin[blas])
    return fail(RDRT_ERR_INVALID_ARG, "unknown BLAS or geometry");
  std::string err;
  rc = pick_all(b, b.blas_geo_begin[blas] + geometry, ray, face_side, out, capacity, count, err);
  if (rc) return fail(rc, err);
  return RDRT_OK;
}

int32_t rdrt_set_device_build(rdrt_ctx *c, uint32_t enable) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  if (enable > 1) return fail(RDRT_ERR_INVALID_ARG, "enable is 0 or 1");
  if (enable && !c->has_device) return fail(RDRT_ERR_NO_DEVICE, "a device build needs a device context");
  if (c->device_build != (enable != 0)) invalidate(c);
  c->device_build = enable != 0;
  return RDRT_OK;
}

int32_t rdrt_commit(rdrt_ctx *c) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  int32_t rc = need_device(c);
  if (rc) return rc;
  rc = rdrt_build(c);
  if (rc) return rc;
  if (c->uploaded) return RDRT_OK;
  RDRT_HIP(hipStreamSynchronize(c->stream));
  free_device_scene(c->dev);
  release_refit(c);
  const BuiltScene &b = c->built;
  DeviceScene &d = c->dev;
  RDRT_HIP(upload(d.tri_pairs, b.tri_pairs, d.bytes));
  RDRT_HIP(upload(d.tlas_pairs, b.tlas_pairs, d.bytes));
  RDRT_HIP(upload(d.tri_quads, b.tri_quads, d.bytes));
  RDRT_HIP(upload(d.tlas_quads, b.tlas_quads, d.bytes));
  if (!b.device_records) RDRT_HIP(upload(d.tri_qquads, b.tri_qquads, d.bytes));
  RDRT_HIP(upload(d.tlas_qquads, b.tlas_qquads, d.bytes));
  RDRT_HIP(upload(d.tlas_quad_roots, b.tlas_quad_roots, d.bytes));
  if (!b.device_records) RDRT_HIP(upload(d.tris, b.tris, d.bytes));
  RDRT_HIP(upload(d.redirect, b.redirect, d.bytes));
  RDRT_HIP(upload(d.tri_ids, b.tri_ids, d.bytes));
  RDRT_HIP(upload(d.indices, b.indices, d.bytes));
  RDRT_HIP(upload(d.binding, b.binding, d.bytes));
  RDRT_HIP(upload(d.vertices, b.vertices, d.bytes));
  RDRT_HIP(upload(d.geos, b.geos, d.bytes));
  RDRT_HIP(upload(d.insts, b.insts, d.bytes));
  RDRT_HIP(upload(d.inst_fwd, b.inst_fwd, d.bytes));
  RDRT_HIP(upload(d.tlas_roots, b.tlas_roots, d.bytes));
  RDRT_HIP(upload(d.tri_forest, b.tri_forest, d.bytes));
  RDRT_HIP(upload(d.tlas_forest, b.tlas_forest, d.bytes));
  RDRT_HIP(upload(d.normals, b.normals, d.bytes));
  RDRT_HIP(upload(d.uvs, b.uvs, d.bytes));
  RDRT_HIP(upload(d.inst_lookup, b.inst_lookup, d.bytes));
  RDRT_HIP(upload(d.tlas_inst_first, b.tlas_inst_first, d.bytes));
  RDRT_HIP(upload(d.tri_material, b.tri_material, d.bytes));
  RDRT_HIP(upload(d.anyhit, b.anyhit, d.bytes));
  if (b.device_records) {
    // computed here from the arrays just uploaded (the same arithmetic as
    // the host builder's), instead of on the host and copied
    const size_t nslots = b.tri_ids.size(), nquads = b.tri_quads.size();
    RDRT_HIP(hipMalloc(&d.tris.ptr, std::max<size_t>(nslots, 1) * sizeof(TriPre)));
    d.tris.count = nslots;
    d.bytes += std::max<size_t>(nslots, 1) * sizeof(TriPre);
    RDRT_HIP(hipMalloc(&d.tri_qquads.ptr, std::max<size_t>(nquads, 1) * sizeof(QuadNodeQ)));
    d.tri_qquads.count = nquads;
    d.bytes += std::max<size_t>(nquads, 1) * sizeof(QuadNodeQ);
    RDRT_HIP(enqueue_scene_records(d.vertices.ptr, d.indices.ptr, d.tri_ids.ptr, (uint32_t)nslots, d.tris.ptr,
                                   d.tri_quads.ptr, (uint32_t)nquads, d.tri_qquads.ptr, c->stream));
    RDRT_HIP(hipStreamSynchronize(c->stream));
  }
  c->uploaded = true;
  return RDRT_OK;
}

int32_t rdrt_refit(rdrt_ctx *c, const float *positions, const float *normals) {
  if (!c || !positions) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->uploaded) return fail(RDRT_ERR_NOT_BUILT, "scene is not committed (call rdrt_commit)");
  BuiltScene &b = c->built;
  DeviceScene &d = c->dev;
  if (normals && d.normals.count == 0) return fail(RDRT_ERR_STATE, "the scene has no vertex normals to replace");
  hipStream_t st = c->stream;
  if (!c->refit_ready) {
    const RefitLevels L = refit_levels(b);
    uint64_t scratch_bytes = 0;
    RDRT_HIP(upload(c->refit_forest, L.forest, scratch_bytes));
    RDRT_HIP(upload(c->refit_pairs, L.pairs, scratch_bytes));
    RDRT_HIP(upload(c->refit_quads, L.quads, scratch_bytes));
    c->refit_forest_off = L.forest_off;
    c->refit_pairs_off = L.pairs_off;
    c->refit_quads_off = L.quads_off;
    c->refit_ready = true;
  }
  const uint32_t ntris = (uint32_t)d.tri_ids.count;
  RDRT_HIP(c->refit_boxes.reserve((size_t)std::max<uint32_t>(ntris, 1u) * 6 * sizeof(float)));
  if (d.vertices.count)
    RDRT_HIP(hipMemcpyAsync(d.vertices.ptr, positions, d.vertices.count * sizeof(float), hipMemcpyDeviceToDevice, st));
  if (normals)
    RDRT_HIP(hipMemcpyAsync(d.normals.ptr, normals, d.normals.count * sizeof(float), hipMemcpyDeviceToDevice, st));
  RefitArgs a{};
  a.vertices = d.vertices.ptr;
  a.indices = d.indices.ptr;
  a.tri_ids = d.tri_ids.ptr;
  a.redirect = d.redirect.ptr;
  a.nslots = ntris;
  a.tri_box = (float *)c->refit_boxes.ptr;
  a.tris = d.tris.ptr;
  a.forest = d.tri_forest.ptr;
  a.pairs = d.tri_pairs.ptr;
  a.quads = d.tri_quads.ptr;
  a.qquads = d.tri_qquads.ptr;
  a.geos = d.geos.ptr;
  a.ngeos = (uint32_t)d.geos.count;
  a.forest_nodes = c->refit_forest.ptr;
  a.pair_nodes = c->refit_pairs.ptr;
  a.quad_nodes = c->refit_quads.ptr;
  auto levels = [](const std::vector<uint32_t> &off) { return off.empty() ? 0u : (uint32_t)off.size() - 1; };
  RDRT_HIP(enqueue_refit(a, c->refit_forest_off.data(), levels(c->refit_forest_off), c->refit_pairs_off.data(),
                         levels(c->refit_pairs_off), c->refit_quads_off.data(), levels(c->refit_quads_off), st));
  // The TLAS layer is small: rebuilt on the host from the refitted BLAS boxes.
  if (!b.geos.empty())
    RDRT_HIP(hipMemcpyAsync(b.geos.data(), d.geos.ptr, b.geos.size() * sizeof(GeoRec), hipMemcpyDeviceToHost, st));
  RDRT_HIP(hipStreamSynchronize(st));
  for (size_t k = 0; k < b.blas_box.size(); k++) {
    Box bx = box_empty();
    for (uint32_t g = b.blas_geo_begin[k]; g < b.blas_geo_end[k]; g++)
      for (int j = 0; j < 3; j++) {
        bx.mn[j] = std::fmin(bx.mn[j], b.geos[g].mn[j]);
        bx.mx[j] = std::fmax(bx.mx[j], b.geos[g].mx[j]);
      }
    b.blas_box[k] = bx;
  }
  std::string err;
  rc = build_tlas_layer(c->src, b, err);
  if (rc != RDRT_OK) return fail(rc, err);
  drop_captures(c);
  RDRT_HIP(reupload(d.insts, b.insts, d.bytes));
  RDRT_HIP(reupload(d.inst_fwd, b.inst_fwd, d.bytes));
  RDRT_HIP(reupload(d.tlas_forest, b.tlas_forest, d.bytes));
  RDRT_HIP(reupload(d.tlas_pairs, b.tlas_pairs, d.bytes));
  RDRT_HIP(reupload(d.tlas_quads, b.tlas_quads, d.bytes));
  RDRT_HIP(reupload(d.tlas_qquads, b.tlas_qquads, d.bytes));
  RDRT_HIP(reupload(d.tlas_roots, b.tlas_roots, d.bytes));
  RDRT_HIP(reupload(d.tlas_quad_roots, b.tlas_quad_roots, d.bytes));
  RDRT_HIP(reupload(d.inst_lookup, b.inst_lookup, d.bytes));
  RDRT_HIP(reupload(d.tlas_inst_first, b.tlas_inst_first, d.bytes));
  return RDRT_OK;
}

int32_t rdrt_get_scene_info(const rdrt_ctx *c, rdrt_scene_info *o) {
  if (!c || !o) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (!c->built_valid) return fail(RDRT_ERR_NOT_BUILT, "scene is not built");
  const BuiltScene &b = c->built;
  std::memset(o, 0, sizeof(*o));
  o->tri_nodes = (uint32_t)b.tri_forest.size();
  o->tlas_nodes = (uint32_t)b.tlas_forest.size();
  o->tris = (uint32_t)b.redirect.size();
  o->verts = (uint32_t)(b.vertices.size() / 3);
  o->instances = (uint32_t)b.insts.size();
  o->geometries = (uint32_t)b.geos.size();
  o->blas = (uint32_t)c->src.blas.size();
  o->tlas = (uint32_t)b.tlas_roots.size();
  o->max_blas_depth = b.max_blas_depth;
  o->max_tlas_depth = b.max_tlas_depth;
  o->stack_entries = stack_entries_for(b.blas_stack + b.tlas_stack);
  o->quad_stack_entries = stack_entries_for(b.blas_stack4 + b.tlas_stack4);
  o->device_bytes = c->dev.bytes;
  o->build_ms = b.build_ms;
  for (int k = 0; k < 4; k++) o->build_phase_ms[k] = b.phase_ms[k];
  o->tri_quads = (uint32_t)b.tri_quads.size();
  o->tlas_quads = (uint32_t)b.tlas_quads.size();
  return RDRT_OK;
}

int32_t rdrt_copy_built(const rdrt_ctx *c, uint32_t which, void *out, size_t bytes) {
  if (!c || !out) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (!c->built_valid) return fail(RDRT_ERR_NOT_BUILT, "scene is not built");
  const BuiltScene &b = c->built;
  const void *src = nullptr;
  size_t need = 0;
  std::vector<uint32_t> order;
  std::vector<QuadNodeQ> qq;
  switch (which) {
    case 0:
      src = b.tri_forest.data();
      need = b.tri_forest.size() * sizeof(RefNode);
      break;
    case 1:
      src = b.tlas_forest.data();
      need = b.tlas_forest.size() * sizeof(RefNode);
      break;
    case 2:
      src = b.redirect.data();
      need = b.redirect.size() * 4;
      break;
    case 3:
      for (const InstRec &r : b.insts) order.push_back(r.order);
      src = order.data();
      need = order.size() * 4;
      break;
    case 4:
      src = b.tri_quads.data();
      need = b.tri_quads.size() * sizeof(QuadNode);
      break;
    case 5:
      if (b.device_records) {  // the device holds them: the host twin of the same arithmetic
        qq.resize(b.tri_quads.size());
        for (size_t k = 0; k < qq.size(); k++) qq[k] = quantize_quad(b.tri_quads[k]);
        src = qq.data();
        need = qq.size() * sizeof(QuadNodeQ);
      } else {
        src = b.tri_qquads.data();
        need = b.tri_qquads.size() * sizeof(QuadNodeQ);
      }
      break;
    default:
      return fail(RDRT_ERR_INVALID_ARG, "unknown array");
  }
  if (bytes < need) return fail(RDRT_ERR_INVALID_ARG, "output buffer too small");
  if (need) std::memcpy(out, src, need);
  return RDRT_OK;
}

}  // extern "C"

namespace rdrt {

// Carves the cull pass's scratch out of one grow-only buffer of the context.
int32_t cull_scratch(rdrt_ctx *c, uint32_t n, CullScratch &cs) {
  RDRT_HIP(c->trace_cull.reserve(cull_scratch_words(n) * 4));
  const size_t nb = ((size_t)n + kCullBlock - 1) / kCullBlock;
  uint32_t *w = (uint32_t *)c->trace_cull.ptr;
  cs.tmp = w;
  cs.blk = w + nb * kCullBlock;
  cs.queue = cs.blk + 2 * nb;
  cs.count = cs.queue + n;
  cs.later = nullptr;
  cs.later_rounds = 0;
  cs.stride = n;
  return RDRT_OK;
}

}  // namespace rdrt

extern "C" {

int32_t rdrt_trace(rdrt_ctx *c, const rdrt_trace_args *a) {
  if (!c || !a) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->uploaded) return fail(RDRT_ERR_NOT_BUILT, "scene is not committed (call rdrt_commit)");
  if (a->n && (!a->rays || !a->hits)) return fail(RDRT_ERR_INVALID_ARG, "null rays or hits");
  if (a->traversal > RDRT_TRAVERSAL_REFERENCE) return fail(RDRT_ERR_INVALID_ARG, "unknown traversal");
  TraceParams p = scene_trace_params(c);
  p.rays = a->rays;
  p.meta = a->ray_meta;
  p.hits = a->hits;
  p.hit_meta = a->hit_meta;
  p.counters = a->counters;
  p.n = a->n;
  p.flags = a->flags;
  p.cull_mask = a->cull_mask;
  p.tlas_idx = a->tlas_idx;
  uint32_t push = push_depth(c, c->tune);
  if (a->traversal == RDRT_TRAVERSAL_FAST && stack_entries_for(push) == 0)
    return fail(RDRT_ERR_CAPACITY, "BVH too deep for the fast traversal stack (use RDRT_TRAVERSAL_REFERENCE)");
  if (c->tune.persist) {
    RDRT_HIP(c->trace_work.reserve(4));
    p.work = (uint32_t *)c->trace_work.ptr;
  }
  const bool cull = c->tune.cull && a->traversal == RDRT_TRAVERSAL_FAST;
  CullScratch cs{};
  if (cull) {
    rc = cull_scratch(c, a->n, cs);
    if (rc) return rc;
  }
  RDRT_HIP(hipEventRecord(c->ev0, c->stream));
  if (cull) RDRT_HIP(launch_trace_culled(p, push, c->tune, cs, c->stream));
  else RDRT_HIP(launch_trace(p, a->traversal, push, c->tune, c->stream));
  RDRT_HIP(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  c->call_timed = true;
  return RDRT_OK;
}

int32_t rdrt_fetch_hit_attributes(rdrt_ctx *c, const rdrt_ray *rays, const rdrt_ray_meta *ray_meta,
                                  const rdrt_hit *hits, const rdrt_hit_meta *hit_meta, uint32_t n, uint32_t tlas_idx,
                                  uint32_t sbt_offset, uint32_t sbt_stride, rdrt_hit_attributes *out) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  if (n && (!rays || !hits || !hit_meta || !out)) return fail(RDRT_ERR_INVALID_ARG, "null rays, hits, metadata or output");
  return fetch_hit_data(c, rays, ray_meta, hits, hit_meta, n, tlas_idx, sbt_offset, sbt_stride, out, nullptr);
}

int32_t rdrt_fetch_hit_transforms(rdrt_ctx *c, const rdrt_ray *rays, const rdrt_ray_meta *ray_meta,
                                  const rdrt_hit *hits, const rdrt_hit_meta *hit_meta, uint32_t n, uint32_t tlas_idx,
                                  rdrt_hit_transform *out) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  if (n && (!rays || !hits || !hit_meta || !out)) return fail(RDRT_ERR_INVALID_ARG, "null rays, hits, metadata or output");
  return fetch_hit_data(c, rays, ray_meta, hits, hit_meta, n, tlas_idx, 0, 0, nullptr, out);
}

int32_t rdrt_set_tune(rdrt_ctx *c, const rdrt_tune *t) {
  if (!c || !t) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if ((t->block_size != 64 && t->block_size != 128 && t->block_size != 256) ||
      (t->lds_entries != 8 && t->lds_entries != 16) || (t->waves != 0 && t->waves != 6 && t->waves != 8) ||
      t->persistent > 1 ||
      t->cull > 1 || t->wide > 2 || t->coop > (1u << 22) || t->tail > kPtMaxBounces)
    return fail(RDRT_ERR_INVALID_ARG,
                "block_size 64/128/256, lds_entries 8/16, waves 0/6/8, persistent 0/1, cull 0/1, wide 0/1/2, "
                "coop <= 2^22, tail <= 16");
  TraceTune n;
  n.block = t->block_size;
  n.lds = t->lds_entries;
  n.waves = t->waves;
  n.persist = t->persistent;
  n.cull = t->cull;
  n.wide = t->wide;
  n.coop = t->coop;
  n.tail = t->tail;
  if (c->built_valid && stack_entries_for(push_depth(c, n)) == 0)
    return fail(RDRT_ERR_CAPACITY, "BVH too deep for the fast traversal stack of this layout");
  c->tune = n;
  return RDRT_OK;
}

int32_t rdrt_get_tune(rdrt_ctx *c, rdrt_tune *t) {
  if (!c || !t) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  std::memset(t, 0, sizeof(*t));
  t->block_size = c->tune.block;
  t->lds_entries = c->tune.lds;
  t->waves = c->tune.waves;
  t->persistent = c->tune.persist;
  t->cull = c->tune.cull;
  t->wide = c->tune.wide;
  t->ms = c->tune_ms;
  t->coop = c->tune.coop;
  t->tail = c->tune.tail;
  return RDRT_OK;
}

int32_t rdrt_autotune(rdrt_ctx *c, const rdrt_trace_args *a, uint32_t reps, rdrt_tune *best) {
  if (!c || !a) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->uploaded) return fail(RDRT_ERR_NOT_BUILT, "scene is not committed (call rdrt_commit)");
  if (a->n == 0 || !a->rays || !a->hits) return fail(RDRT_ERR_INVALID_ARG, "the sample needs rays and hits");
  if (stack_entries_for(push_depth(c, TraceTune{})) == 0)
    return fail(RDRT_ERR_CAPACITY, "BVH too deep for the fast traversal stack");
  if (reps == 0) reps = 3;
  TraceParams p = scene_trace_params(c);
  p.rays = a->rays;
  p.meta = a->ray_meta;
  p.hits = a->hits;
  p.n = a->n;
  p.flags = a->flags;
  p.cull_mask = a->cull_mask;
  p.tlas_idx = a->tlas_idx;
  RDRT_HIP(c->trace_work.reserve(4));
  p.work = (uint32_t *)c->trace_work.ptr;
  CullScratch cs{};
  rc = cull_scratch(c, a->n, cs);
  if (rc) return rc;
  hipStream_t st = c->stream;
  TraceTune chosen = c->tune;
  float chosen_ms = -1.0f;
  for (uint32_t block : {64u, 128u, 256u})
    for (uint32_t lds : {8u, 16u})
      for (uint32_t waves : {0u, 6u, 8u})
        for (uint32_t shape = 0; shape < 12; shape++) {
          TraceTune t;
          t.block = block;
          t.lds = lds;
          t.waves = waves;
          t.persist = shape & 1u;
          t.cull = (shape >> 1) & 1u;
          t.wide = shape >> 2;
          const uint32_t push = push_depth(c, t);
          if (stack_entries_for(push) == 0) continue;  // layout too deep for the stack
          auto run = [&]() -> hipError_t {
            return t.cull ? launch_trace_culled(p, push, t, cs, st) : launch_trace(p, RDRT_TRAVERSAL_FAST, push, t, st);
          };
          RDRT_HIP(run());  // warm-up
          RDRT_HIP(hipEventRecord(c->ev0, st));
          for (uint32_t r = 0; r < reps; r++) RDRT_HIP(run());
          RDRT_HIP(hipEventRecord(c->ev1, st));
          RDRT_HIP(hipEventSynchronize(c->ev1));
          float ms = 0.0f;
          RDRT_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
          ms /= (float)reps;
          if (chosen_ms < 0.0f || ms < chosen_ms) {
            chosen_ms = ms;
            chosen = t;
          }
        }
  c->tune = chosen;
  c->tune_ms = chosen_ms;
  c->timed = false;
  c->call_timed = false;
  if (best) return rdrt_get_tune(c, best);
  return RDRT_OK;
}

int32_t rdrt_trace_closest(rdrt_ctx *c, const rdrt_ray *rays, uint32_t n, rdrt_hit *hits) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  rdrt_trace_args a;
  std::memset(&a, 0, sizeof(a));
  a.rays = rays;
  a.hits = hits;
  a.n = n;
  a.cull_mask = 0xffffffffu;
  a.traversal = c->cfg.traversal;
  return rdrt_trace(c, &a);
}

int32_t rdrt_blas_create(rdrt_ctx *c, const float *pos, uint32_t nverts, const uint32_t *idx, uint32_t ntris,
                         uint32_t geometry_flags, const rdrt_build_opts *opts, uint32_t *out) {
  if (!c || !out) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (opts) {
    const TreeOptions o = tree_opt(opts->blas_bin_size, opts->blas_max_depth, opts->sah_bins, 2);
    const TreeOptions &h = c->src.blas_opt;
    if (o.bin_size != h.bin_size || o.max_depth != h.max_depth || o.bins != h.bins)
      return fail(RDRT_ERR_UNSUPPORTED, "per-BLAS build options must equal the context's");
  }
  int32_t rc = rdrt_blas_begin(c);
  if (rc) return rc;
  rc = rdrt_blas_add_triangles(c, pos, nverts, idx, ntris, geometry_flags);
  if (rc) {
    c->pending.clear();
    c->in_blas = false;
    return rc;
  }
  return rdrt_blas_end(c, out);
}

int32_t rdrt_sobol(uint32_t n, uint32_t s0, uint32_t s1, float out2[2]) {
  if (!out2) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  // 24-bit fixed point to [0, 1 - 2^-23] (sobol.rs:47-50,64-67)
  auto to_float = [](uint32_t x) {
    const float f = (float)((x >> 8) & 0xffffffu) / 16777216.0f;
    return f < 1.0f - 1.1920928955078125e-7f ? f : 1.0f - 1.1920928955078125e-7f;
  };
  uint32_t r = 0;  // radical inverse: the bits of n reversed
  for (int k = 0; k < 32; k++) r |= ((n >> k) & 1u) << (31 - k);
  out2[0] = to_float(r ^ s0);
  uint32_t v = 1u << 31, x = s1;  // dimension 1: v_{k+1} = v_k ^ (v_k >> 1)
  for (uint32_t m = n; m; m >>= 1, v ^= v >> 1)
    if (m & 1u) x ^= v;
  out2[1] = to_float(x);
  return RDRT_OK;
}

int32_t rdrt_get_metrics(rdrt_ctx *c, rdrt_metrics *o) {
  if (!c || !o) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  std::memset(o, 0, sizeof(*o));
  o->last_call_ms = -1.0f;
  o->tune_ms = c->tune_ms;
  o->device_bytes = c->dev.bytes;
  o->build_ms = c->built_valid ? c->built.build_ms : 0.0;
  if (!c->has_device) return RDRT_OK;
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (c->call_timed) {
    RDRT_HIP(hipEventSynchronize(c->ev1));
    RDRT_HIP(hipEventElapsedTime(&o->last_call_ms, c->ev0, c->ev1));
  }
  RDRT_HIP(hipStreamSynchronize(c->work_stream));
  RDRT_HIP(hipMemcpy(&o->status, c->status, 4, hipMemcpyDeviceToHost));
  return RDRT_OK;
}

int32_t rdrt_last_trace_ms(rdrt_ctx *c, float *ms) {
  if (!c || !ms) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (!c->timed) return fail(RDRT_ERR_STATE, "no trace recorded");
  int32_t rc = need_device(c);
  if (rc) return rc;
  RDRT_HIP(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return RDRT_OK;
}

}  // extern "C"
