// api_ao.hip -- C-ABI of the device task graph and of ambient occlusion on
// it (include/rdrt.h).

#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

#include "ao.h"
#include "context.h"

using namespace rdrt;

namespace rdrt {

struct AoState {
  tg::GraphMemory mem;
  tg::Graph call_end{};   // the graph as a call's launches leave it
  tg::Graph last{};       // call_end of the last call, advanced by rdrt_ao_states
  uint32_t capacity = 0;  // pixels per tile the buffers hold
  Scratch rays, hits, hitx, tq, counts, frame, work;
  float *host_frame = nullptr;  // pinned staging of the frame block
  hipEvent_t frame_copied = nullptr;
  bool frame_pending = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> key;  // launch key of the captured graph
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
  }
};

void drop_ao_capture(AoState *s) {
  if (s) s->drop();
}

void free_ao(AoState *&s) {
  if (!s) return;
  s->drop();
  s->mem.release();
  for (Scratch *x : {&s->rays, &s->hits, &s->hitx, &s->tq, &s->counts, &s->frame, &s->work}) x->free();
  if (s->host_frame) (void)hipHostFree(s->host_frame);
  if (s->frame_copied) (void)hipEventDestroy(s->frame_copied);
  delete s;
  s = nullptr;
}

// Counters of every group after a prepare and a compaction of each
// (read_back_execution_states, mod.rs:305-370).
int32_t read_task_states(tg::Graph &gr, hipStream_t st, uint32_t *wake, uint32_t *sof, uint32_t *empty,
                         uint32_t *errors) {
  for (uint32_t k = 0; k < gr.n; k++) {
    RDRT_HIP(tg::enqueue_prepare(gr, k, st));
    RDRT_HIP(tg::enqueue_compact(gr, k, st));
  }
  std::vector<uint32_t> ctr((size_t)gr.n * tg::kCtrWords);
  for (uint32_t k = 0; k < gr.n; k++)
    RDRT_HIP(hipMemcpyAsync(ctr.data() + (size_t)k * tg::kCtrWords, gr.g[k].ctr, tg::kCtrWords * 4,
                            hipMemcpyDeviceToHost, st));
  RDRT_HIP(hipStreamSynchronize(st));
  for (uint32_t k = 0; k < gr.n; k++) {
    const uint32_t *c = ctr.data() + (size_t)k * tg::kCtrWords;
    const uint32_t w = c[tg::kActive0 + gr.g[k].cur], e = c[tg::kEmpty0 + gr.g[k].pp];
    if (wake) wake[k] = w;
    if (empty) empty[k] = e;
    if (sof) sof[k] = gr.g[k].size - (e + w <= gr.g[k].size ? e + w : gr.g[k].size);
    if (errors) errors[k] = c[tg::kError];
  }
  return RDRT_OK;
}

}  // namespace rdrt

namespace {

void hemisphere_sample(uint32_t s, float out[3]) {
  // hammersley_2d(s, 256) (sampling.rs:33-50)
  const float ux = (float)s / (float)kAoMaxSample;
  uint32_t b = s;
  b = (b << 16) | (b >> 16);
  b = ((b & 0x55555555u) << 1) | ((b & 0xAAAAAAAAu) >> 1);
  b = ((b & 0x33333333u) << 2) | ((b & 0xCCCCCCCCu) >> 2);
  b = ((b & 0x0F0F0F0Fu) << 4) | ((b & 0xF0F0F0F0u) >> 4);
  b = ((b & 0x00FF00FFu) << 8) | ((b & 0xFF00FF00u) >> 8);
  const float uy = (float)b * 2.3283064e-10f;
  // sample_hemisphere_cos (sampling.rs:61-68); 2 pi rounded to f32
  const float phi = 6.2831855f * uy;
  const float cos_theta = std::sqrt(1.0f - ux);
  const float sin_theta = std::sqrt(1.0f - cos_theta * cos_theta);
  out[0] = (float)std::cos((double)phi) * sin_theta;
  out[1] = (float)std::sin((double)phi) * sin_theta;
  out[2] = cos_theta;
}

int32_t reserve_ao(rdrt_ctx *c, uint32_t cap) {
  if (!c->ao) c->ao = new AoState();
  AoState &s = *c->ao;
  if (!s.host_frame) {
    RDRT_HIP(hipHostMalloc((void **)&s.host_frame, kAoFrameWords * sizeof(float), hipHostMallocDefault));
    RDRT_HIP(hipEventCreateWithFlags(&s.frame_copied, hipEventDisableTiming));
  }
  if (cap <= s.capacity) return RDRT_OK;
  RDRT_HIP(hipStreamSynchronize(c->work_stream));
  s.drop();
  const uint32_t ntrace = 2 * cap;  // two traces in flight per pixel (ao.rs:297)
  if (s.mem.alloc({ntrace, cap, cap, cap, cap}, {3, 1, 1, 1, 1}) != 0)
    return fail(RDRT_ERR_HIP, "task graph allocation failed");
  RDRT_HIP(s.rays.reserve((size_t)(ntrace + 1) * sizeof(rdrt_ray)));
  RDRT_HIP(s.hits.reserve((size_t)(ntrace + 1) * sizeof(rdrt_hit)));
  RDRT_HIP(s.hitx.reserve((size_t)(ntrace + 1) * sizeof(uint2)));
  RDRT_HIP(s.tq.reserve((size_t)2 * ntrace * 4));
  RDRT_HIP(s.counts.reserve(8 * 4));
  RDRT_HIP(s.frame.reserve(kAoFrameWords * sizeof(float)));
  RDRT_HIP(s.work.reserve(4));
  s.capacity = cap;
  return RDRT_OK;
}

}  // namespace

struct rdrt_task_graph {
  rdrt_ctx *ctx = nullptr;
  tg::GraphMemory mem;
  std::vector<tg::KindDesc> kinds;
};

extern "C" {

int32_t rdrt_task_graph_create(rdrt_ctx *c, const rdrt_task_group_desc *groups, uint32_t n, uint32_t capacity,
                               rdrt_task_graph **out) {
  if (!c || !groups || !out) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (n == 0 || n > tg::kMaxGroups) return fail(RDRT_ERR_INVALID_ARG, "1 to 8 task groups");
  std::vector<uint32_t> sizes, words;
  std::vector<tg::KindDesc> kinds;
  for (uint32_t k = 0; k < n; k++) {
    const rdrt_task_group_desc &d = groups[k];
    if (d.kind > RDRT_TASK_SPAWN_THEN) return fail(RDRT_ERR_INVALID_ARG, "unknown task kind");
    if (d.kind == RDRT_TASK_SPAWN_THEN && d.child_group >= n) return fail(RDRT_ERR_INVALID_ARG, "child group out of range");
    const uint64_t size = (uint64_t)d.max_in_flight * capacity;
    if (size == 0 || size > (1u << 26)) return fail(RDRT_ERR_INVALID_ARG, "group slots must be 1 .. 2^26");
    sizes.push_back((uint32_t)size);
    words.push_back(1);
    kinds.push_back(tg::KindDesc{d.kind, d.child_group});
  }
  rdrt_task_graph *g = new rdrt_task_graph();
  g->ctx = c;
  g->kinds = kinds;
  if (g->mem.alloc(sizes, words) != 0) {
    delete g;
    return fail(RDRT_ERR_HIP, "task graph allocation failed");
  }
  hipError_t e = tg::enqueue_reset(g->mem.dev, c->work_stream);
  if (e != hipSuccess) {
    g->mem.release();
    delete g;
    return fail(RDRT_ERR_HIP, hipGetErrorString(e));
  }
  *out = g;
  return RDRT_OK;
}

int32_t rdrt_task_graph_destroy(rdrt_task_graph *g) {
  if (!g) return RDRT_OK;
  (void)hipSetDevice(g->ctx->cfg.device);
  (void)hipStreamSynchronize(g->ctx->work_stream);
  g->mem.release();
  delete g;
  return RDRT_OK;
}

int32_t rdrt_task_graph_spawn(rdrt_task_graph *g, uint32_t group, uint32_t count) {
  if (!g) return fail(RDRT_ERR_INVALID_ARG, "null graph");
  if (group >= g->mem.dev.n) return fail(RDRT_ERR_INVALID_ARG, "group out of range");
  int32_t rc = need_device(g->ctx);
  if (rc) return rc;
  RDRT_HIP(tg::enqueue_spawn_roots(g->mem.dev, group, count, g->ctx->work_stream));
  return RDRT_OK;
}

int32_t rdrt_task_graph_execute(rdrt_task_graph *g, uint32_t rounds) {
  if (!g) return fail(RDRT_ERR_INVALID_ARG, "null graph");
  int32_t rc = need_device(g->ctx);
  if (rc) return rc;
  hipStream_t st = g->ctx->work_stream;
  tg::Graph &gr = g->mem.dev;
  const tg::KindDesc *kinds = g->kinds.data();
  RDRT_HIP(tg::enqueue_rounds(gr, rounds, [&](uint32_t gid) { return tg::enqueue_generic_poll(gr, gid, kinds, st); },
                              st));
  return RDRT_OK;
}

int32_t rdrt_task_graph_states(rdrt_task_graph *g, uint32_t *wake, uint32_t *sof, uint32_t *empty, uint32_t *errors) {
  if (!g) return fail(RDRT_ERR_INVALID_ARG, "null graph");
  int32_t rc = need_device(g->ctx);
  if (rc) return rc;
  return read_task_states(g->mem.dev, g->ctx->work_stream, wake, sof, empty, errors);
}

int32_t rdrt_ao_hemisphere_sample(uint32_t s, float out[3]) {
  if (!out) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  hemisphere_sample(s, out);
  return RDRT_OK;
}

int32_t rdrt_ao_render(rdrt_ctx *c, const rdrt_ao_params *prm, void *accum) {
  if (!c || !prm || !accum) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->uploaded) return fail(RDRT_ERR_NOT_BUILT, "scene is not committed (call rdrt_commit)");
  if (prm->width == 0 || prm->height == 0) return fail(RDRT_ERR_INVALID_ARG, "empty frame");
  if (prm->format > RDRT_AO_FORMAT_RGBA8) return fail(RDRT_ERR_INVALID_ARG, "unknown AO format");
  if ((uint64_t)prm->width * prm->height > (1ull << 28)) return fail(RDRT_ERR_CAPACITY, "more than 2^28 pixels");
  if (stack_entries_for(push_depth(c, c->tune)) == 0)
    return fail(RDRT_ERR_CAPACITY, "BVH too deep for the fast traversal");
  const uint32_t tile_pixels = prm->tile_pixels ? prm->tile_pixels : 512u * 512u;
  if (tile_pixels > (1u << 24)) return fail(RDRT_ERR_INVALID_ARG, "tile_pixels above 2^24");
  uint32_t rows = tile_pixels / prm->width;
  if (rows == 0) rows = 1;
  if (rows > prm->height) rows = prm->height;
  const uint32_t rounds_hint = prm->rounds ? prm->rounds : 8u;  // ao.rs:300
  const uint32_t rounds = rounds_hint < kAoChainRounds ? rounds_hint : kAoChainRounds;
  rc = reserve_ao(c, rows * prm->width);
  if (rc) return rc;
  AoState &s = *c->ao;
  hipStream_t st = c->work_stream;

  // the frame block: camera, hemisphere direction of this sample, index
  if (s.frame_pending) RDRT_HIP(hipEventSynchronize(s.frame_copied));
  float *f = s.host_frame;
  std::memcpy(f, prm->view_projection_inv, 16 * sizeof(float));
  std::memcpy(f + 16, prm->camera_position, 3 * sizeof(float));
  f[19] = 0.0f;
  hemisphere_sample(prm->sample_index, f + 20);
  std::memcpy(f + 23, &prm->sample_index, 4);
  RDRT_HIP(hipMemcpyAsync(s.frame.ptr, f, kAoFrameWords * sizeof(float), hipMemcpyHostToDevice, st));
  RDRT_HIP(hipEventRecord(s.frame_copied, st));
  s.frame_pending = true;

  const DeviceScene &d = c->dev;
  AoLaunch L;
  std::memset(&L, 0, sizeof(L));
  L.gr = s.mem.dev;
  L.rays = (rdrt_ray *)s.rays.ptr;
  L.hits = (rdrt_hit *)s.hits.ptr;
  L.hitx = (uint2 *)s.hitx.ptr;
  L.tq0 = (uint32_t *)s.tq.ptr;
  L.tq1 = L.tq0 + 2 * (size_t)s.capacity;
  L.tcount = (uint32_t *)s.counts.ptr;
  L.frame = (const float *)s.frame.ptr;
  L.width = prm->width;
  L.height = prm->height;
  L.tile_w = prm->width;
  L.accum = accum;
  L.format = prm->format;
  L.insts = d.insts.ptr;
  L.geos = d.geos.ptr;
  L.indices = d.indices.ptr;
  L.vertices = d.vertices.ptr;
  L.normals = d.normals.count ? d.normals.ptr : nullptr;
  const TraceParams tp = scene_trace_params(c);
  const uint32_t depth = push_depth(c, c->tune);
  uint32_t *work = (uint32_t *)s.work.ptr;
  auto enqueue_all = [&]() -> hipError_t {
    for (uint32_t y0 = 0; y0 < prm->height; y0 += rows) {
      AoLaunch T = L;
      T.tile_y = y0;
      const uint32_t h = prm->height - y0 < rows ? prm->height - y0 : rows;
      hipError_t e = enqueue_ao_tile(T, h, rounds, tp, depth, c->tune, work, y0 == 0, st, &s.call_end);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  if (!prm->use_graph) {
    RDRT_HIP(enqueue_all());
    s.last = s.call_end;
    return RDRT_OK;
  }
  std::vector<uint8_t> key(sizeof(L) + sizeof(tp) + sizeof(TraceTune) + 3 * 4);
  uint8_t *k = key.data();
  std::memcpy(k, &L, sizeof(L));
  std::memcpy(k + sizeof(L), &tp, sizeof(tp));
  std::memcpy(k + sizeof(L) + sizeof(tp), &c->tune, sizeof(TraceTune));
  const uint32_t extra[3] = {rows, rounds, depth};
  std::memcpy(k + sizeof(L) + sizeof(tp) + sizeof(TraceTune), extra, sizeof(extra));
  if (!s.exec || key != s.key) {
    s.drop();
    RDRT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_all();
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(RDRT_ERR_HIP, std::string("AO capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    s.graph = g;
    RDRT_HIP(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
    s.key = key;
  }
  RDRT_HIP(hipGraphLaunch(s.exec, st));
  s.last = s.call_end;
  return RDRT_OK;
}

int32_t rdrt_ao_autotune(rdrt_ctx *c, const rdrt_ao_params *prm, void *accum, uint32_t reps, struct rdrt_tune *best) {
  if (!c || !prm || !accum) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (reps == 0) reps = 3;
  rdrt_ao_params q = *prm;
  q.use_graph = 0;  // eager calls: no capture per shape, the same launches
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
  // average call time of shape t over n calls
  auto call_ms = [&](const TraceTune &t, uint32_t n, float &ms) -> int32_t {
    c->tune = t;
    int32_t r2 = rdrt_ao_render(c, &q, accum);  // warm-up
    if (r2 == RDRT_OK && hipEventRecord(e0, c->work_stream) != hipSuccess) r2 = fail(RDRT_ERR_HIP, "event");
    for (uint32_t r = 0; r < n && r2 == RDRT_OK; r++) r2 = rdrt_ao_render(c, &q, accum);
    if (r2 != RDRT_OK) return r2;
    if (hipEventRecord(e1, c->work_stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      return fail(RDRT_ERR_HIP, "AO timing failed");
    ms /= (float)n;
    return RDRT_OK;
  };
  std::vector<std::pair<float, TraceTune>> timed;
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
          float ms = 0.0f;
          rc = call_ms(t, reps, ms);
          if (rc != RDRT_OK) break;
          timed.push_back({ms, t});
          if (chosen_ms < 0.0f || ms < chosen_ms) {
            chosen_ms = ms;
            chosen = t;
          }
        }
  // the fastest few, timed again on 3 x reps calls, twice, in turn (shapes
  // within a few percent of each other are otherwise told apart by noise)
  if (rc == RDRT_OK && timed.size() > 1) {
    std::sort(timed.begin(), timed.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    const size_t nk = std::min<size_t>(4, timed.size());
    std::vector<float> sum(nk, 0.0f);
    for (uint32_t pass = 0; pass < 2 && rc == RDRT_OK; pass++)
      for (size_t k = 0; k < nk && rc == RDRT_OK; k++) {
        float ms = 0.0f;
        rc = call_ms(timed[k].second, 3 * reps, ms);
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
  c->tune = chosen;
  c->tune_ms = chosen_ms;
  if (best) {
    best->block_size = chosen.block;
    best->lds_entries = chosen.lds;
    best->waves = chosen.waves;
    best->persistent = chosen.persist;
    best->cull = chosen.cull;
    best->wide = chosen.wide;
    best->ms = chosen_ms;
    best->coop = chosen.coop;
  }
  return RDRT_OK;
}

int32_t rdrt_ao_ray_counts(rdrt_ctx *c, uint32_t out[2]) {
  if (!c || !out) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->ao || !c->ao->capacity) return fail(RDRT_ERR_STATE, "no AO frame rendered yet");
  uint32_t h[6];
  RDRT_HIP(hipMemcpyAsync(h, c->ao->counts.ptr, sizeof(h), hipMemcpyDeviceToHost, c->work_stream));
  RDRT_HIP(hipStreamSynchronize(c->work_stream));
  out[0] = h[4];
  out[1] = h[5];
  return RDRT_OK;
}

int32_t rdrt_ao_states(rdrt_ctx *c, uint32_t *wake, uint32_t *sof, uint32_t *empty, uint32_t *errors) {
  if (!c) return fail(RDRT_ERR_INVALID_ARG, "null context");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->ao || !c->ao->capacity) return fail(RDRT_ERR_STATE, "no AO frame rendered yet");
  return read_task_states(c->ao->last, c->work_stream, wake, sof, empty, errors);
}

}  // extern "C"
