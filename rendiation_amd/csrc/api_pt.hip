// api_pt.hip -- C-ABI of the path tracer, shading attributes and the device
// primitives (compaction, Sobol) of include/rdrt.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "context.h"

using namespace rdrt;

namespace rdrt {

void PtState::free_all() {
  for (Scratch *s : {&rays, &shadow_rays, &hits, &shadow_hits, &hitx, &beta, &L, &nee, &qa, &qb, &qs, &counts,
                     &tmp_a, &tmp_b, &blk, &materials, &diag_counters, &diag_totals, &record, &views, &work, &tflags, &det_acc,
                     &det_flags, &det_hdr, &grec, &loss_part})
    s->free();
  host_views.clear();
  materials_dirty = true;
  last_recorded = false;
  for (hipEvent_t e : ev_begin) (void)hipEventDestroy(e);
  for (hipEvent_t e : ev_end) (void)hipEventDestroy(e);
  ev_begin.clear();
  ev_end.clear();
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  have_cached = false;
  capacity = 0;
}

}  // namespace rdrt

namespace {

// Device scratch for n reproducible sums.
int32_t reserve_det(PtState &s, size_t n, DetBufs &db) {
  RDRT_HIP(s.det_acc.reserve(std::max<size_t>(n, 1) * sizeof(det::u64)));
  RDRT_HIP(s.det_flags.reserve(std::max<size_t>(n, 1) * sizeof(uint32_t)));
  RDRT_HIP(s.det_hdr.reserve(sizeof(det::Header)));
  db.acc = (det::u64 *)s.det_acc.ptr;
  db.flags = (uint32_t *)s.det_flags.ptr;
  db.h = (det::Header *)s.det_hdr.ptr;
  return RDRT_OK;
}

int32_t reserve_pt(rdrt_ctx *c, uint32_t npaths) {
  PtState &s = c->pt;
  if (npaths <= s.capacity) return RDRT_OK;
  size_t n = npaths;
  RDRT_HIP(s.rays.reserve(n * sizeof(rdrt_ray)));
  RDRT_HIP(s.shadow_rays.reserve(n * sizeof(rdrt_ray)));
  RDRT_HIP(s.hits.reserve(n * sizeof(rdrt_hit)));
  RDRT_HIP(s.shadow_hits.reserve(n * sizeof(rdrt_hit)));
  RDRT_HIP(s.hitx.reserve(n * sizeof(uint2)));
  RDRT_HIP(s.beta.reserve(n * 12));
  RDRT_HIP(s.L.reserve(n * 12));
  RDRT_HIP(s.nee.reserve(n * 12));
  RDRT_HIP(s.qa.reserve(n * 4));
  RDRT_HIP(s.qb.reserve(n * 4));
  RDRT_HIP(s.qs.reserve(n * 4));
  size_t stride = (n + kPtBlock - 1) / kPtBlock;
  RDRT_HIP(s.counts.reserve((2 * kPtMaxBounces + 3) * 4));
  RDRT_HIP(s.tmp_a.reserve(stride * kPtBlock * 4));
  RDRT_HIP(s.tmp_b.reserve(stride * kPtBlock * 4));
  RDRT_HIP(s.blk.reserve(4 * stride * 4));
  s.capacity = npaths;
  s.have_cached = false;
  return RDRT_OK;
}

GeometrySrc *geometry_of(rdrt_ctx *c, uint32_t blas, uint32_t geometry) {
  if (blas >= c->src.blas.size() || !c->src.blas_live[blas] || geometry >= c->src.blas[blas].size()) return nullptr;
  return &c->src.blas[blas][geometry];
}

}  // namespace

extern "C" {

int32_t rdrt_geometry_vertex_range(rdrt_ctx *c, uint32_t blas, uint32_t geometry, uint32_t *first,
                                   uint32_t *count) {
  if (!c || !first || !count) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (!geometry_of(c, blas, geometry)) return fail(RDRT_ERR_INVALID_ARG, "unknown BLAS or geometry");
  // vertices are concatenated over live BLAS in creation order, geometries in order
  uint64_t at = 0;
  for (uint32_t b = 0; b < c->src.blas.size(); b++) {
    if (!c->src.blas_live[b]) continue;
    for (uint32_t gi = 0; gi < c->src.blas[b].size(); gi++) {
      const uint64_t nv = c->src.blas[b][gi].pos.size() / 3;
      if (b == blas && gi == geometry) {
        *first = (uint32_t)at;
        *count = (uint32_t)nv;
        return RDRT_OK;
      }
      at += nv;
    }
  }
  return fail(RDRT_ERR_INVALID_ARG, "unknown BLAS or geometry");
}

int32_t rdrt_blas_set_vertex_normals(rdrt_ctx *c, uint32_t blas, uint32_t geometry, const float *normals,
                                     uint32_t nverts) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  GeometrySrc *g = geometry_of(c, blas, geometry);
  if (!g) return fail(RDRT_ERR_INVALID_ARG, "unknown BLAS or geometry");
  if (nverts != g->pos.size() / 3 || (nverts && !normals))
    return fail(RDRT_ERR_INVALID_ARG, "normals must give one xyz triple per vertex");
  g->normals.assign(normals, normals + 3 * (size_t)nverts);
  invalidate(c);
  return RDRT_OK;
}

int32_t rdrt_blas_set_vertex_uvs(rdrt_ctx *c, uint32_t blas, uint32_t geometry, const float *uvs, uint32_t nverts) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  GeometrySrc *g = geometry_of(c, blas, geometry);
  if (!g) return fail(RDRT_ERR_INVALID_ARG, "unknown BLAS or geometry");
  if (nverts != g->pos.size() / 3 || (nverts && !uvs))
    return fail(RDRT_ERR_INVALID_ARG, "uvs must give one pair per vertex");
  g->uvs.assign(uvs, uvs + 2 * (size_t)nverts);
  invalidate(c);
  return RDRT_OK;
}

int32_t rdrt_blas_set_triangle_materials(rdrt_ctx *c, uint32_t blas, uint32_t geometry, const uint32_t *offsets,
                                         uint32_t ntris) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  GeometrySrc *g = geometry_of(c, blas, geometry);
  if (!g) return fail(RDRT_ERR_INVALID_ARG, "unknown BLAS or geometry");
  uint32_t nt = g->indexed ? (uint32_t)(g->idx.size() / 3) : (uint32_t)(g->pos.size() / 9);
  if (ntris != nt || (ntris && !offsets)) return fail(RDRT_ERR_INVALID_ARG, "one material offset per triangle");
  g->tri_material.assign(offsets, offsets + ntris);
  invalidate(c);
  return RDRT_OK;
}

int32_t rdrt_blas_set_anyhit(rdrt_ctx *c, uint32_t blas, uint32_t geometry, const uint8_t *behavior, uint32_t ntris) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  GeometrySrc *g = geometry_of(c, blas, geometry);
  if (!g) return fail(RDRT_ERR_INVALID_ARG, "unknown BLAS or geometry");
  uint32_t nt = g->indexed ? (uint32_t)(g->idx.size() / 3) : (uint32_t)(g->pos.size() / 9);
  if (ntris != nt || (ntris && !behavior)) return fail(RDRT_ERR_INVALID_ARG, "one any-hit behaviour per triangle");
  for (uint32_t t = 0; t < ntris; t++)
    if (behavior[t] & ~(RDRT_ANYHIT_ACCEPT | RDRT_ANYHIT_END_SEARCH))
      return fail(RDRT_ERR_INVALID_ARG, "unknown any-hit behaviour bits");
  g->anyhit.assign(behavior, behavior + ntris);
  invalidate(c);
  return RDRT_OK;
}

int32_t rdrt_set_materials(rdrt_ctx *c, const rdrt_material *m, uint32_t n) {
  if (!c || (n && !m)) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  c->materials.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    Material &d = c->materials[i];
    std::memset(&d, 0, sizeof(d));
    for (int k = 0; k < 3; k++) {
      d.albedo[k] = m[i].albedo[k];
      d.emission[k] = m[i].emission[k];
    }
  }
  c->pt.materials_dirty = true;
  return RDRT_OK;
}

int32_t rdrt_pt_render(rdrt_ctx *c, const rdrt_pt_params *prm, float *radiance) {
  if (!c || !prm || !radiance) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->uploaded) return fail(RDRT_ERR_NOT_BUILT, "scene is not committed (call rdrt_commit)");
  if (prm->width == 0 || prm->height == 0 || prm->spp == 0) return fail(RDRT_ERR_INVALID_ARG, "empty frame");
  if (prm->max_bounces > kPtMaxBounces) return fail(RDRT_ERR_INVALID_ARG, "max_bounces above 16");
  const uint32_t nviews = prm->views > 1 ? prm->views : 1u;
  if (prm->views > 1 && !prm->view_params) return fail(RDRT_ERR_INVALID_ARG, "views > 1 needs view_params");
  uint64_t npaths64 = (uint64_t)nviews * prm->width * prm->height * prm->spp;
  if (npaths64 > (1ull << 30)) return fail(RDRT_ERR_CAPACITY, "more than 2^30 paths per frame");
  if (c->materials.empty()) return fail(RDRT_ERR_STATE, "no materials set (rdrt_set_materials)");
  // every instance must address a material: custom index + largest triangle offset
  const BuiltScene &b = c->built;
  for (const InstRec &in : b.insts) {
    uint32_t mx = 0;
    for (uint32_t g = in.geo_begin; g < in.geo_end; g++) mx = std::max(mx, b.geo_max_material[g]);
    if ((uint64_t)in.custom_index + mx >= c->materials.size())
      return fail(RDRT_ERR_INVALID_ARG, "an instance addresses a material beyond rdrt_set_materials");
  }
  if (stack_entries_for(push_depth(c, c->tune)) == 0) return fail(RDRT_ERR_CAPACITY, "BVH too deep for the fast traversal");
  const uint32_t npaths = (uint32_t)npaths64;
  rc = reserve_pt(c, npaths);
  if (rc) return rc;
  PtState &s = c->pt;
  // camera list: view_params, or the single camera of the parameters
  std::vector<float> views(20 * (size_t)nviews, 0.0f);
  for (uint32_t v = 0; v < nviews; v++) {
    if (prm->view_params) {
      std::memcpy(&views[20 * (size_t)v], prm->view_params + 20 * (size_t)v, 19 * sizeof(float));
    } else {
      std::memcpy(&views[0], prm->view_projection_inv, 16 * sizeof(float));
      std::memcpy(&views[16], prm->camera_position, 3 * sizeof(float));
    }
  }
  // synchronous uploads wait for the frames still reading the old contents
  if (views != s.host_views) {
    RDRT_HIP(hipStreamSynchronize(c->work_stream));
    RDRT_HIP(s.views.reserve(views.size() * sizeof(float)));
    RDRT_HIP(hipMemcpy(s.views.ptr, views.data(), views.size() * sizeof(float), hipMemcpyHostToDevice));
    s.host_views = views;
  }
  if (s.materials_dirty) {
    RDRT_HIP(hipStreamSynchronize(c->work_stream));
    RDRT_HIP(s.materials.reserve(c->materials.size() * sizeof(Material)));
    RDRT_HIP(hipMemcpy(s.materials.ptr, c->materials.data(), c->materials.size() * sizeof(Material),
                       hipMemcpyHostToDevice));
    s.nmaterials_uploaded = (uint32_t)c->materials.size();
    s.materials_dirty = false;
  }

  PtLaunch L;
  std::memset(&L, 0, sizeof(L));
  L.views = (const float *)s.views.ptr;
  L.nviews = nviews;
  L.width = prm->width;
  L.height = prm->height;
  L.spp = prm->spp;
  L.npaths = npaths;
  L.max_bounces = prm->max_bounces;
  L.rr_start = prm->rr_start;
  L.seed = prm->seed;
  L.sample_base = prm->sample_index_base;
  std::memcpy(L.sky_up, prm->sky_up, 12);
  std::memcpy(L.sky_down, prm->sky_down, 12);
  std::memcpy(L.sun_dir, prm->sun_direction, 12);
  std::memcpy(L.sun_radiance, prm->sun_radiance, 12);
  const DeviceScene &d = c->dev;
  L.insts = d.insts.ptr;
  L.geos = d.geos.ptr;
  L.indices = d.indices.ptr;
  L.vertices = d.vertices.ptr;
  L.normals = d.normals.count ? d.normals.ptr : nullptr;
  L.tri_material = d.tri_material.count ? d.tri_material.ptr : nullptr;
  L.materials = (const Material *)s.materials.ptr;
  L.nmaterials = s.nmaterials_uploaded;
  L.rays = (rdrt_ray *)s.rays.ptr;
  L.shadow_rays = (rdrt_ray *)s.shadow_rays.ptr;
  L.hits = (rdrt_hit *)s.hits.ptr;
  L.shadow_hits = (rdrt_hit *)s.shadow_hits.ptr;
  L.hitx = (uint2 *)s.hitx.ptr;
  L.beta = (float *)s.beta.ptr;
  L.L = (float *)s.L.ptr;
  L.nee = (float *)s.nee.ptr;
  L.qa = (uint32_t *)s.qa.ptr;
  L.qb = (uint32_t *)s.qb.ptr;
  L.qs = (uint32_t *)s.qs.ptr;
  L.counts = (uint32_t *)s.counts.ptr;
  L.tmp_a = (uint32_t *)s.tmp_a.ptr;
  L.tmp_b = (uint32_t *)s.tmp_b.ptr;
  L.blk = (uint32_t *)s.blk.ptr;
  L.blk_stride = (npaths + kPtBlock - 1) / kPtBlock;
  L.status = c->status;
  RDRT_HIP(s.work.reserve(8));
  L.work = (uint32_t *)s.work.ptr;
  L.out = radiance;
  if (prm->record) {
    RDRT_HIP(s.record.reserve((size_t)(prm->max_bounces + 2) * npaths * sizeof(uint2)));
    L.record = (uint2 *)s.record.ptr;
  }

  const TraceParams tp = scene_trace_params(c);
  hipStream_t st = c->work_stream;
  if (prm->diagnostics & ~(RDRT_PT_DIAG_TIME_TRACES | RDRT_PT_DIAG_COUNT_VISITS | RDRT_PT_DIAG_COUNT_FAST))
    return fail(RDRT_ERR_INVALID_ARG, "unknown diagnostics bits");
  const uint32_t counting = prm->diagnostics & (RDRT_PT_DIAG_COUNT_VISITS | RDRT_PT_DIAG_COUNT_FAST);
  if (counting == (RDRT_PT_DIAG_COUNT_VISITS | RDRT_PT_DIAG_COUNT_FAST))
    return fail(RDRT_ERR_INVALID_ARG, "count visits in one order per frame");
  PtDiag diag;
  if (prm->diagnostics) {
    const size_t launches = 2 * (kPtMaxBounces + 1);
    if (prm->diagnostics & RDRT_PT_DIAG_TIME_TRACES) {
      while (s.ev_begin.size() < launches) {
        hipEvent_t a, b;
        RDRT_HIP(hipEventCreate(&a));
        RDRT_HIP(hipEventCreate(&b));
        s.ev_begin.push_back(a);
        s.ev_end.push_back(b);
      }
      diag.ev_begin = s.ev_begin.data();
      diag.ev_end = s.ev_end.data();
    }
    if (counting) {
      RDRT_HIP(s.diag_counters.reserve((size_t)npaths * 12));
      RDRT_HIP(s.diag_totals.reserve(3 * 8 + (2 * kPtMaxBounces + 1) * 4));
      diag.counters = (uint32_t *)s.diag_counters.ptr;
      diag.totals = (unsigned long long *)s.diag_totals.ptr;
      diag.max_visits = (uint32_t *)((char *)s.diag_totals.ptr + 3 * 8);
      diag.fast_order = counting == RDRT_PT_DIAG_COUNT_FAST;
    }
  }
  s.last_diag = prm->diagnostics;
  // the tail kernel (not with visit counting: one ray per lane there)
  if (c->tune.tail && c->tune.tail < prm->max_bounces && !counting) {
    RDRT_HIP(s.tflags.reserve((size_t)npaths * 4));
    L.tail_b0 = c->tune.tail;
    L.tflags = (uint32_t *)s.tflags.ptr;
  }
  // shadow rays of bounce b share a launch with extension rays of b + 1,
  // except in visit-counting frames; the tail kernel is one launch
  s.last_launches = counting ? 2 * prm->max_bounces + 1 : L.tail_b0 ? L.tail_b0 + 2 : prm->max_bounces + 1;
  RDRT_HIP(hipEventRecord(c->ev0, st));
  if (prm->diagnostics) {
    RDRT_HIP(enqueue_pt_frame(L, tp, push_depth(c, c->tune), c->tune, st, &diag));
  } else if (prm->use_graph) {
    if (!s.exec || !s.have_cached || std::memcmp(&s.cached, &L, sizeof(L)) != 0 ||
        std::memcmp(&s.cached_tune, &c->tune, sizeof(TraceTune)) != 0) {
      if (s.exec) (void)hipGraphExecDestroy(s.exec);
      if (s.graph) (void)hipGraphDestroy(s.graph);
      s.exec = nullptr;
      s.graph = nullptr;
      RDRT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      hipError_t e = enqueue_pt_frame(L, tp, push_depth(c, c->tune), c->tune, st);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(st, &g);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(RDRT_ERR_HIP, std::string("frame capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
      }
      s.graph = g;
      RDRT_HIP(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
      s.cached = L;
      s.cached_tune = c->tune;
      s.have_cached = true;
    }
    RDRT_HIP(hipGraphLaunch(s.exec, st));
  } else {
    RDRT_HIP(enqueue_pt_frame(L, tp, push_depth(c, c->tune), c->tune, st));
  }
  RDRT_HIP(hipEventRecord(c->ev1, st));
  c->call_timed = true;
  s.last_bounces = prm->max_bounces;
  s.last_frame = L;
  s.last_recorded = prm->record != 0;
  return RDRT_OK;
}

int32_t rdrt_pt_forward(rdrt_ctx *c, const rdrt_pt_params *prm, float *radiance) {
  return rdrt_pt_render(c, prm, radiance);
}

int32_t rdrt_pt_autotune(rdrt_ctx *c, const rdrt_pt_params *prm, float *radiance, uint32_t reps, rdrt_tune *best) {
  if (!c || !prm || !radiance) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (prm->record || prm->diagnostics) return fail(RDRT_ERR_INVALID_ARG, "tune on plain frames (no record, no diagnostics)");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (reps == 0) reps = 3;
  rdrt_pt_params q = *prm;
  q.use_graph = 0;  // eager frames: no capture per shape, the same launches
  hipEvent_t e0 = nullptr, e1 = nullptr;
  RDRT_HIP(hipEventCreate(&e0));
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(RDRT_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(he));
  }
  const TraceTune before = c->tune;
  TraceTune chosen = before;
  float chosen_ms = -1.0f;
  rc = RDRT_OK;
  std::vector<std::pair<float, TraceTune>> timed;  // every shape's time, for the refinement below
  // average frame time of shape t over n frames
  auto frame_ms = [&](const TraceTune &t, uint32_t n, float &ms) -> int32_t {
    c->tune = t;
    int32_t r2 = rdrt_pt_render(c, &q, radiance);  // warm-up
    if (r2 == RDRT_OK && hipEventRecord(e0, c->work_stream) != hipSuccess) r2 = fail(RDRT_ERR_HIP, "event");
    for (uint32_t r = 0; r < n && r2 == RDRT_OK; r++) r2 = rdrt_pt_render(c, &q, radiance);
    if (r2 != RDRT_OK) return r2;
    if (hipEventRecord(e1, c->work_stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      return fail(RDRT_ERR_HIP, "frame timing failed");
    ms /= (float)n;
    return RDRT_OK;
  };
  // the shape's time, kept when it is the fastest so far
  auto time_shape = [&](const TraceTune &t) -> int32_t {
    float ms = 0.0f;
    const int32_t r2 = frame_ms(t, reps, ms);
    if (r2 != RDRT_OK) return r2;
    timed.push_back({ms, t});
    if (chosen_ms < 0.0f || ms < chosen_ms) {
      chosen_ms = ms;
      chosen = t;
    }
    return RDRT_OK;
  };
  for (uint32_t block : {64u, 128u, 256u})
    for (uint32_t lds : {8u, 16u})
      for (uint32_t waves : {0u, 6u, 8u})
        for (uint32_t shape = 0; shape < 18 && rc == RDRT_OK; shape++) {
          TraceTune t;
          t.block = block;
          t.lds = lds;
          t.waves = waves;
          t.persist = shape & 1u;
          t.wide = (shape >> 1) % 3u;
          t.coop = kCoopChoices[shape / 6u];
          if (stack_entries_for(push_depth(c, t)) == 0) continue;
          rc = time_shape(t);
        }
  // then where the tail kernel takes over, with the fastest shape
  const TraceTune base = chosen;
  for (uint32_t tail = 1; tail < prm->max_bounces && tail <= 3 && rc == RDRT_OK; tail++) {
    TraceTune t = base;
    t.tail = tail;
    rc = time_shape(t);
  }
  // the shapes within a few percent of each other are told apart by their
  // timing noise: the fastest few are timed again on 3 x reps frames, twice,
  // in turn, and the lowest mean is kept
  if (rc == RDRT_OK && timed.size() > 1) {
    std::sort(timed.begin(), timed.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    const size_t nk = std::min<size_t>(4, timed.size());
    std::vector<float> sum(nk, 0.0f);
    for (uint32_t pass = 0; pass < 2 && rc == RDRT_OK; pass++)
      for (size_t k = 0; k < nk && rc == RDRT_OK; k++) {
        float ms = 0.0f;
        rc = frame_ms(timed[k].second, 3 * reps, ms);
        sum[k] += ms;
      }
    if (rc == RDRT_OK) {
      size_t bk = 0;
      for (size_t k = 1; k < nk; k++)
        if (sum[k] < sum[bk]) bk = k;
      chosen = timed[bk].second;
      chosen_ms = sum[bk] / 2.0f;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != RDRT_OK) {
    c->tune = before;
    return rc;
  }
  // the cull pass is not part of a frame; keep the context's choice for rdrt_trace
  chosen.cull = before.cull;
  c->tune = chosen;
  c->tune_ms = chosen_ms;
  if (best) return rdrt_get_tune(c, best);
  return RDRT_OK;
}

int32_t rdrt_pt_set_material_params(rdrt_ctx *c, const float *albedo, const float *emission, uint32_t n) {
  if (!c || !albedo || !emission) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (n == 0 || n >= kVertSkyUp) return fail(RDRT_ERR_INVALID_ARG, "material count out of range");
  int32_t rc = need_device(c);
  if (rc) return rc;
  PtState &s = c->pt;
  RDRT_HIP(s.materials.reserve((size_t)n * sizeof(Material)));
  RDRT_HIP(enqueue_pack_materials(albedo, emission, n, (Material *)s.materials.ptr, c->work_stream));
  // the device copy is now the authoritative one; the host list keeps the count
  c->materials.assign(n, Material{});
  s.nmaterials_uploaded = n;
  s.materials_dirty = false;
  return RDRT_OK;
}

int32_t rdrt_pt_backward(rdrt_ctx *c, const float *adj, float *grad_albedo, float *grad_emission) {
  if (!c || !adj || !grad_albedo || !grad_emission) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  PtState &s = c->pt;
  if (!s.last_recorded || !s.last_frame.record)
    return fail(RDRT_ERR_STATE, "no recorded frame (render with record = 1 first)");
  if (s.last_frame.rr_start < s.last_frame.max_bounces)
    return fail(RDRT_ERR_UNSUPPORTED, "the adjoint needs a frame without Russian roulette (rr_start >= max_bounces)");
  PtLaunch L = s.last_frame;
  // materials as they are now (set_material_params may have moved them)
  L.materials = (const Material *)s.materials.ptr;
  if (s.nmaterials_uploaded != L.nmaterials)
    return fail(RDRT_ERR_STATE, "the material count changed since the recorded frame");
  DetBufs db;
  rc = reserve_det(s, 6 * (size_t)L.nmaterials, db);
  if (rc) return rc;
  RDRT_HIP(enqueue_pt_backward(L, adj, grad_albedo, grad_emission, db, c->work_stream));
  return RDRT_OK;
}

int32_t rdrt_pt_backward_geometry(rdrt_ctx *c, const float *adj, float *grad_pos, float *grad_nrm) {
  if (!c || !adj || !grad_pos) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  PtState &s = c->pt;
  if (!s.last_recorded || !s.last_frame.record)
    return fail(RDRT_ERR_STATE, "no recorded frame (render with record = 1 first)");
  if (s.last_frame.rr_start < s.last_frame.max_bounces)
    return fail(RDRT_ERR_UNSUPPORTED, "the adjoint needs a frame without Russian roulette (rr_start >= max_bounces)");
  if (s.nmaterials_uploaded != s.last_frame.nmaterials)
    return fail(RDRT_ERR_STATE, "the material count changed since the recorded frame");
  if (c->dev.normals.count && !grad_nrm) return fail(RDRT_ERR_INVALID_ARG, "the scene has vertex normals: grad_nrm needed");
  PtLaunch L = s.last_frame;
  L.materials = (const Material *)s.materials.ptr;
  L.adj = adj;
  L.g_pos = grad_pos;
  L.g_nrm = grad_nrm;  // zeroed by the replay; stays zero without vertex normals
  L.nvertices = (uint32_t)(c->dev.vertices.count / 3);
  DetBufs db;
  rc = reserve_det(s, 6 * (size_t)L.nvertices, db);
  if (rc) return rc;
  RDRT_HIP(s.grec.reserve(std::max<size_t>((size_t)L.max_bounces * L.npaths, 1) * 3 * sizeof(uint4)));
  L.grec = (uint4 *)s.grec.ptr;
  L.det_h = db.h;
  const TraceParams tp = scene_trace_params(c);
  RDRT_HIP(enqueue_pt_replay(L, tp, push_depth(c, c->tune), c->tune, db, c->work_stream));
  return RDRT_OK;
}

int32_t rdrt_l2_adjoint(rdrt_ctx *c, const float *a, const float *b, uint32_t n, float *adj, double *loss) {
  if (!c || !loss || (n && (!a || !b || !adj))) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  RDRT_HIP(c->pt.loss_part.reserve(l2_adjoint_parts(n) * sizeof(double)));
  RDRT_HIP(enqueue_l2_adjoint(a, b, n, adj, loss, (double *)c->pt.loss_part.ptr, c->work_stream));
  return RDRT_OK;
}

int32_t rdrt_pt_get_stats(rdrt_ctx *c, rdrt_pt_stats *o) {
  if (!c || !o) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->pt.counts.ptr) return fail(RDRT_ERR_STATE, "no frame rendered");
  RDRT_HIP(hipStreamSynchronize(c->work_stream));
  std::memset(o, 0, sizeof(*o));
  uint32_t mb = c->pt.last_bounces;
  uint32_t counts[2 * kPtMaxBounces + 3];
  RDRT_HIP(hipMemcpy(counts, c->pt.counts.ptr, (2 * mb + 3) * 4, hipMemcpyDeviceToHost));
  o->bounces = mb + 1;
  for (uint32_t b = 0; b <= mb; b++) {
    o->extension_rays[b] = counts[b];
    o->shadow_rays[b] = b < mb ? counts[mb + 2 + b] : 0u;
  }
  RDRT_HIP(hipEventElapsedTime(&o->frame_ms, c->ev0, c->ev1));
  o->trace_launches = c->pt.last_launches;
  if (c->pt.last_diag & RDRT_PT_DIAG_TIME_TRACES) {
    for (uint32_t i = 0; i < c->pt.last_launches; i++) {
      float ms = 0.0f;
      RDRT_HIP(hipEventElapsedTime(&ms, c->pt.ev_begin[i], c->pt.ev_end[i]));
      o->trace_ms += ms;
      if (i < 33) o->launch_ms[i] = ms;
    }
  }
  if (c->pt.last_diag & (RDRT_PT_DIAG_COUNT_VISITS | RDRT_PT_DIAG_COUNT_FAST)) {
    RDRT_HIP(hipMemcpy(o->visits, c->pt.diag_totals.ptr, 24, hipMemcpyDeviceToHost));
    RDRT_HIP(hipMemcpy(o->max_ray_visits, (char *)c->pt.diag_totals.ptr + 24, c->pt.last_launches * 4,
                       hipMemcpyDeviceToHost));
  }
  return RDRT_OK;
}

int32_t rdrt_compact(rdrt_ctx *c, const uint32_t *mask, uint32_t n, uint32_t *out_idx, uint32_t *count) {
  if (!c || !count || (n && (!mask || !out_idx))) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (n > (1u << 31) - 1) return fail(RDRT_ERR_CAPACITY, "compaction supports fewer than 2^31 elements");
  int32_t rc = need_device(c);
  if (rc) return rc;
  RDRT_HIP(c->compact_scratch.reserve(compact_scratch_words(n) * 4));
  RDRT_HIP(enqueue_compact(mask, n, out_idx, count, (uint32_t *)c->compact_scratch.ptr, c->stream));
  return RDRT_OK;
}

int32_t rdrt_sample_sobol(rdrt_ctx *c, const uint32_t *index, const uint32_t *scr, uint32_t n, float *out) {
  if (!c || (n && (!index || !scr || !out))) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  RDRT_HIP(enqueue_sobol(index, scr, n, out, c->stream));
  return RDRT_OK;
}

}  // extern "C"
