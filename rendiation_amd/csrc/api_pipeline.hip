// api_pipeline.hip -- C-ABI of the programmable ray-tracing pipeline
// (include/rdrt.h): GPURaytracingPipelineAndBindingSource and its executor
// (shader/ray-tracing/src/api/pipeline.rs:75-130; wavefront_compute/
// pipeline.rs:26-277) and RayTracingEncoderProvider::trace_ray
// (api/backend.rs:48-56; wavefront_compute/mod.rs:111-190) over the device
// task graph.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "attrib.h"
#include "context.h"
#include "pipeline.h"

using namespace rdrt;

struct rdrt_pipeline {
  rdrt_ctx *ctx = nullptr;
  rdrt_pipeline_desc desc{};
  uint32_t ngroups = 0, capacity = 0;  // tile entries the buffers hold
  const void *stages[tg::kMaxGroups] = {};
  tg::GraphMemory mem;
  tg::Graph last{};  // the graph as the last call left it
  PipeLaunch base{};
  std::vector<Scratch> bufs;
  Scratch work;
  // captured call (use_graph): the launches of one (launch size, user
  // pointer, scene, buffers, launch shape), replayed while they stay the same
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> key;
  tg::Graph graph_last{};
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
  }
  ~rdrt_pipeline() {
    drop();
    mem.release();
    for (Scratch &s : bufs) s.free();
    work.free();
  }
};

namespace {

constexpr uint32_t kDefaultTile = 1u << 22;

template <class T>
T *carve(rdrt_pipeline *p, size_t count, hipError_t &e) {
  p->bufs.emplace_back();
  Scratch &s = p->bufs.back();
  if (e == hipSuccess) e = s.reserve(std::max<size_t>(count, 1) * sizeof(T));
  return static_cast<T *>(s.ptr);
}

// Device memory for tiles of `cap` launch entries.
int32_t reserve_pipeline(rdrt_pipeline *p, uint32_t cap) {
  if (cap <= p->capacity) return RDRT_OK;
  RDRT_HIP(hipStreamSynchronize(p->ctx->work_stream));
  p->drop();
  p->mem.release();
  for (Scratch &s : p->bufs) s.free();
  p->bufs.clear();
  p->bufs.reserve(64);
  const rdrt_pipeline_desc &d = p->desc;
  const uint32_t inflight = d.max_in_flight ? d.max_in_flight : 2u;
  const uint64_t ntrace64 = (uint64_t)inflight * cap;
  if (ntrace64 >= (1ull << 28)) return fail(RDRT_ERR_CAPACITY, "tile * max_in_flight above 2^28 tasks");
  const uint32_t ntrace = (uint32_t)ntrace64;
  std::vector<uint32_t> sizes(p->ngroups), words(p->ngroups);
  for (uint32_t g = 0; g < p->ngroups; g++) {
    sizes[g] = g == kPipeRayGen ? cap : ntrace;
    words[g] = g == kPipeRayGen ? 1u : 0u;  // ray gen: the spawn index of a root
  }
  if (p->mem.alloc(sizes, words) != 0) return fail(RDRT_ERR_OUT_OF_MEMORY, "task graph allocation failed");
  PipeLaunch &L = p->base;
  L = PipeLaunch{};
  L.gr = p->mem.dev;
  L.words = d.payload_words;
  L.nclosest = d.nclosest;
  L.nmiss = d.nmiss;
  for (uint32_t k = 0; k < 4; k++) {
    L.closest_group[k] = k < d.nclosest && d.closest_hit[k] ? 2u + k : tg::kNone;
    L.miss_group[k] = k < d.nmiss && d.miss[k] ? 2u + d.nclosest + k : tg::kNone;
  }
  hipError_t e = hipSuccess;
  const size_t nt = (size_t)ntrace + 1, P = d.payload_words;
  L.rays = carve<rdrt_ray>(p, nt, e);
  L.meta = carve<rdrt_ray_meta>(p, nt, e);
  L.sbt = carve<uint4>(p, nt, e);
  L.hits = carve<rdrt_hit>(p, nt, e);
  L.hit_meta = carve<rdrt_hit_meta>(p, nt, e);
  L.hitx = carve<uint2>(p, nt, e);
  L.attrs = d.hit_context & RDRT_PIPELINE_HIT_ATTRIBUTES ? carve<rdrt_hit_attributes>(p, nt, e) : nullptr;
  L.xforms = d.hit_context & RDRT_PIPELINE_HIT_TRANSFORMS ? carve<rdrt_hit_transform>(p, nt, e) : nullptr;
  L.tpay = carve<uint32_t>(p, nt * P, e);
  L.tlaunch = carve<uint32_t>(p, nt, e);
  L.tawait = carve<uint32_t>(p, nt, e);
  L.tq = carve<uint32_t>(p, ntrace, e);
  L.tcount = carve<uint32_t>(p, 4, e);
  for (uint32_t g = kPipeRayGen; g < p->ngroups; g++) {
    const size_t n = (size_t)sizes[g] + 1;
    StageArrays &S = L.st[g];
    S.payload = carve<uint32_t>(p, n * P, e);
    S.result = carve<uint32_t>(p, n * P, e);
    S.action = carve<uint32_t>(p, n, e);
    S.call_payload = carve<uint32_t>(p, n * P, e);
    S.launch = carve<uint32_t>(p, n, e);
    S.trace_of = carve<uint32_t>(p, n, e);
    S.call = carve<rdrt_trace_call>(p, n, e);
  }
  if (e != hipSuccess) return fail(RDRT_ERR_OUT_OF_MEMORY, std::string("pipeline buffers: ") + hipGetErrorString(e));
  RDRT_HIP(p->work.reserve(4));
  p->capacity = cap;
  return RDRT_OK;
}

}  // namespace

extern "C" {

int32_t rdrt_pipeline_create(rdrt_ctx *c, const rdrt_pipeline_desc *d, rdrt_pipeline **out) {
  if (!c || !d || !out) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (!d->ray_gen) return fail(RDRT_ERR_INVALID_ARG, "a ray-gen stage is required");
  if (d->nclosest > 4 || d->nmiss > 4 || d->nclosest + d->nmiss > tg::kMaxGroups - 2)
    return fail(RDRT_ERR_INVALID_ARG, "at most 4 closest-hit and 4 miss stages, 6 together");
  if (d->payload_words == 0 || d->payload_words > RDRT_PIPELINE_MAX_PAYLOAD)
    return fail(RDRT_ERR_INVALID_ARG, "payload_words must be 1..16");
  if (d->max_in_flight > 64) return fail(RDRT_ERR_INVALID_ARG, "max_in_flight above 64");
  if (d->tile_pixels > (1u << 24)) return fail(RDRT_ERR_INVALID_ARG, "tile_pixels above 2^24");
  if (d->hit_context & ~(RDRT_PIPELINE_HIT_ATTRIBUTES | RDRT_PIPELINE_HIT_TRANSFORMS))
    return fail(RDRT_ERR_INVALID_ARG, "unknown hit_context bits");
  if (d->use_graph > 1) return fail(RDRT_ERR_INVALID_ARG, "use_graph is 0 or 1");
  int32_t rc = need_device(c);
  if (rc) return rc;
  rdrt_pipeline *p = new rdrt_pipeline();
  p->ctx = c;
  p->desc = *d;
  p->ngroups = 2 + d->nclosest + d->nmiss;
  p->stages[kPipeRayGen] = d->ray_gen;
  for (uint32_t k = 0; k < d->nclosest; k++) p->stages[2 + k] = d->closest_hit[k];
  for (uint32_t k = 0; k < d->nmiss; k++) p->stages[2 + d->nclosest + k] = d->miss[k];
  *out = p;
  return RDRT_OK;
}

int32_t rdrt_pipeline_destroy(rdrt_pipeline *p) {
  if (!p) return RDRT_OK;
  (void)hipSetDevice(p->ctx->cfg.device);
  (void)hipStreamSynchronize(p->ctx->work_stream);
  delete p;
  return RDRT_OK;
}

int32_t rdrt_pipeline_trace_rays(rdrt_pipeline *p, uint32_t width, uint32_t height, void *user) {
  if (!p) return fail(RDRT_ERR_INVALID_ARG, "null pipeline");
  rdrt_ctx *c = p->ctx;
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->uploaded) return fail(RDRT_ERR_NOT_BUILT, "scene is not committed (call rdrt_commit)");
  if (width == 0 || height == 0) return fail(RDRT_ERR_INVALID_ARG, "empty launch");
  if ((uint64_t)width * height > (1ull << 28)) return fail(RDRT_ERR_CAPACITY, "more than 2^28 launch entries");
  if (stack_entries_for(push_depth(c, c->tune)) == 0)
    return fail(RDRT_ERR_CAPACITY, "BVH too deep for the fast traversal");
  // the reference executes 512 x 512 entries at a time (wavefront_compute/
  // mod.rs:134) to bound its buffers; here the default is 2^22 entries (a
  // 1080p launch in one tile): a tile's launches are latency-bound, so fewer,
  // larger tiles run the same work in fewer rounds of launches
  const uint32_t tile = p->desc.tile_pixels ? p->desc.tile_pixels : kDefaultTile;
  uint32_t rows = std::max<uint32_t>(1u, tile / width);
  rows = std::min(rows, height);
  rc = reserve_pipeline(p, rows * width);
  if (rc) return rc;
  const uint32_t rounds = p->desc.round_hint ? p->desc.round_hint : 8u;
  PipeLaunch L = p->base;
  L.gr = p->mem.dev;
  L.width = width;
  L.height = height;
  L.tile_x = 0;
  L.tile_w = width;
  L.insts = c->dev.insts.ptr;
  const TraceParams scene = scene_trace_params(c);
  const AttribArgs attrib = scene_attrib_args(c);
  TraceTune tune = c->tune;
  tune.cull = 0;  // the queue is the trace group's own
  const uint32_t push = push_depth(c, tune);
  hipStream_t st = c->work_stream;
  tg::Graph last{};
  auto enqueue_all = [&]() -> hipError_t {
    bool first = true;
    for (uint32_t y = 0; y < height; y += rows) {
      const uint32_t h = std::min(rows, height - y);
      L.tile_y = y;
      hipError_t e = enqueue_pipeline_tile(L, p->stages, h, rounds, scene, attrib, push, tune,
                                           (uint32_t *)p->work.ptr, user, first, st, &last);
      if (e != hipSuccess) return e;
      first = false;
    }
    return hipSuccess;
  };
  if (!p->desc.use_graph) {
    RDRT_HIP(enqueue_all());
    p->last = last;
    return RDRT_OK;
  }
  std::vector<uint8_t> key(sizeof(L) + sizeof(scene) + sizeof(attrib) + sizeof(TraceTune) + 4 * 4 + sizeof(void *));
  uint8_t *k = key.data();
  std::memcpy(k, &L, sizeof(L));
  k += sizeof(L);
  std::memcpy(k, &scene, sizeof(scene));
  k += sizeof(scene);
  std::memcpy(k, &attrib, sizeof(attrib));
  k += sizeof(attrib);
  std::memcpy(k, &tune, sizeof(TraceTune));
  k += sizeof(TraceTune);
  const uint32_t extra[4] = {height, rows, rounds, push};
  std::memcpy(k, extra, sizeof(extra));
  k += sizeof(extra);
  std::memcpy(k, &user, sizeof(void *));
  if (!p->exec || key != p->key) {
    p->drop();
    RDRT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_all();
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(RDRT_ERR_HIP, std::string("pipeline capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    p->graph = g;
    RDRT_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    p->key = key;
    p->graph_last = last;
  }
  RDRT_HIP(hipGraphLaunch(p->exec, st));
  p->last = p->graph_last;
  return RDRT_OK;
}

int32_t rdrt_pipeline_states(rdrt_pipeline *p, uint32_t *wake, uint32_t *sof, uint32_t *empty, uint32_t *errors) {
  if (!p) return fail(RDRT_ERR_INVALID_ARG, "null pipeline");
  int32_t rc = need_device(p->ctx);
  if (rc) return rc;
  if (!p->capacity) return fail(RDRT_ERR_STATE, "no call traced yet");
  return read_task_states(p->last, p->ctx->work_stream, wake, sof, empty, errors);
}

int32_t rdrt_pipeline_ray_count(rdrt_pipeline *p, uint64_t *out) {
  if (!p || !out) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(p->ctx);
  if (rc) return rc;
  if (!p->capacity) return fail(RDRT_ERR_STATE, "no call traced yet");
  uint32_t v = 0;
  RDRT_HIP(hipMemcpyAsync(&v, p->base.tcount + 2, 4, hipMemcpyDeviceToHost, p->ctx->work_stream));
  RDRT_HIP(hipStreamSynchronize(p->ctx->work_stream));
  *out = v;
  return RDRT_OK;
}

}  // extern "C"
