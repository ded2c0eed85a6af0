// api_batch.hip -- C-ABI of batched multi-scene tracing (include/rdrt.h).
//
// Persistent mode: per round, one persistent traversal launch over the
// queue of all scenes, then one spawn + compaction launch; the whole call is
// a captured hipGraph replayed while its parameters are unchanged.
// Sequential mode (the baseline): the same kernels scene by scene and round
// by round, with plain (one-ray-per-thread) traversal launches.

#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "batch.h"
#include "context.h"

using namespace rdrt;

namespace rdrt {

void BatchState::free_all() {
  for (Scratch *s : {&wrays, &hitx, &qa, &qb, &counts, &tmp, &blk, &work}) s->free();
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  cached.clear();
}

}  // namespace rdrt

namespace {

struct Plan {
  uint32_t rounds, nscenes;
  std::vector<uint32_t> off;  // scene ray offsets
};

Plan make_plan(const rdrt_batch_params *a) {
  Plan p;
  p.rounds = a->bounces + 1;
  if (a->mode == RDRT_BATCH_SEQUENTIAL) {
    p.nscenes = a->nscenes;
    p.off.assign(a->scene_offsets, a->scene_offsets + a->nscenes + 1);
  } else {
    p.nscenes = 1;
    p.off = {0u, a->n};
  }
  return p;
}

// Enqueue all rounds of scene range k (the whole batch in persistent mode).
// prefilled: every later round already holds miss records (the cull pass
// wrote them), so paths that end need not write theirs.
hipError_t enqueue_scene(rdrt_ctx *c, const rdrt_batch_params *a, const Plan &pl, uint32_t k, bool persist,
                         bool prefilled, uint32_t &launches) {
  BatchState &s = c->batch;
  const uint32_t first = pl.off[k], nk = pl.off[k + 1] - first;
  if (nk == 0) return hipSuccess;
  const uint32_t R = pl.rounds;
  uint32_t *counts = (uint32_t *)s.counts.ptr + (size_t)k * (R + 1);
  uint32_t *queues[2] = {(uint32_t *)s.qa.ptr + first, (uint32_t *)s.qb.ptr + first};
  TraceParams base = scene_trace_params(c);
  base.meta = a->ray_meta;
  base.cull_mask = 0xffffffffu;
  base.n = a->n;
  base.slots = nk;
  base.hitx = (uint2 *)s.hitx.ptr;
  base.work = (uint32_t *)s.work.ptr;
  TraceTune tune = c->tune;
  tune.persist = persist ? 1u : 0u;
  for (uint32_t r = 0; r < R; r++) {
    TraceParams tp = base;
    tp.rays = r == 0 ? a->rays : (const rdrt_ray *)s.wrays.ptr;
    tp.hits = a->hits + (size_t)r * a->n;
    tp.queue = queues[r & 1];
    tp.count = counts + r;
    hipError_t e = launch_trace(tp, RDRT_TRAVERSAL_FAST, push_depth(c, c->tune), tune, c->work_stream);
    if (e != hipSuccess) return e;
    launches += persist ? 2 : 1;
    if (r + 1 == R) break;
    SpawnArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.qin = queues[r & 1];
    sa.count_in = counts + r;
    sa.slots = nk;
    sa.qout = queues[(r + 1) & 1];
    sa.count_out = counts + r + 1;
    // compaction scratch, shared by every spawn (they run in stream order)
    sa.tmp = (uint32_t *)s.tmp.ptr;
    sa.blk = (uint32_t *)s.blk.ptr;
    sa.rays = tp.rays;
    sa.out_rays = (rdrt_ray *)s.wrays.ptr;
    sa.record = a->spawned ? a->spawned + (size_t)r * a->n : nullptr;
    sa.hits = tp.hits;
    sa.hits_later = tp.hits + a->n;
    sa.later_rounds = prefilled ? 0u : R - 1 - r;
    sa.hitx = base.hitx;
    sa.n = a->n;
    sa.round = r + 1;
    sa.seed = a->seed;
    sa.insts = c->dev.insts.ptr;
    sa.geos = c->dev.geos.ptr;
    sa.indices = c->dev.indices.ptr;
    sa.vertices = c->dev.vertices.ptr;
    sa.status = c->status;
    e = enqueue_spawn(sa, c->work_stream);
    if (e != hipSuccess) return e;
    launches++;
  }
  return hipSuccess;
}

}  // namespace

extern "C" {

int32_t rdrt_trace_batched(rdrt_ctx *c, const rdrt_batch_params *a) {
  if (!c || !a) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (!c->uploaded) return fail(RDRT_ERR_NOT_BUILT, "scene is not committed (call rdrt_commit)");
  if (a->n == 0) return RDRT_OK;
  if (!a->rays || !a->ray_meta || !a->hits) return fail(RDRT_ERR_INVALID_ARG, "null rays, metadata or hits");
  if (a->bounces > 16) return fail(RDRT_ERR_INVALID_ARG, "bounces above 16");
  if (a->mode > RDRT_BATCH_SEQUENTIAL) return fail(RDRT_ERR_INVALID_ARG, "unknown batch mode");
  if (a->n > (1u << 30)) return fail(RDRT_ERR_CAPACITY, "more than 2^30 rays per batch");
  if (a->mode == RDRT_BATCH_SEQUENTIAL) {
    if (!a->scene_offsets || a->nscenes == 0) return fail(RDRT_ERR_INVALID_ARG, "sequential mode needs scene offsets");
    if (a->scene_offsets[0] != 0 || a->scene_offsets[a->nscenes] != a->n)
      return fail(RDRT_ERR_INVALID_ARG, "scene offsets must start at 0 and end at n");
    for (uint32_t k = 0; k < a->nscenes; k++)
      if (a->scene_offsets[k] > a->scene_offsets[k + 1]) return fail(RDRT_ERR_INVALID_ARG, "offsets must increase");
  }
  if (stack_entries_for(push_depth(c, c->tune)) == 0) return fail(RDRT_ERR_CAPACITY, "BVH too deep for the fast traversal");
  const Plan pl = make_plan(a);
  BatchState &s = c->batch;
  const size_t n = a->n;
  RDRT_HIP(s.wrays.reserve(n * sizeof(rdrt_ray)));
  RDRT_HIP(s.hitx.reserve(n * 8));
  RDRT_HIP(s.qa.reserve(n * 4));
  RDRT_HIP(s.qb.reserve(n * 4));
  RDRT_HIP(s.counts.reserve((size_t)pl.nscenes * (pl.rounds + 1) * 4));
  const size_t nblk = (n + kSpawnBlock - 1) / kSpawnBlock;
  RDRT_HIP(s.tmp.reserve(nblk * kSpawnBlock * 4));
  RDRT_HIP(s.blk.reserve(2 * nblk * 4));
  RDRT_HIP(s.work.reserve(4));
  hipStream_t st = c->work_stream;
  s.rounds = pl.rounds;
  s.nscenes = pl.nscenes;
  s.mode = a->mode;
  if (a->mode == RDRT_BATCH_SEQUENTIAL) {
    // per-scene queue sizes start at the scene's ray count; uploaded (and
    // waited for) ahead of the timed region so the host array may be reused
    s.host_counts.assign((size_t)pl.nscenes * (pl.rounds + 1), 0u);
    for (uint32_t k = 0; k < pl.nscenes; k++) s.host_counts[(size_t)k * (pl.rounds + 1)] = pl.off[k + 1] - pl.off[k];
    RDRT_HIP(hipMemcpyAsync(s.counts.ptr, s.host_counts.data(), s.host_counts.size() * 4, hipMemcpyHostToDevice, st));
    RDRT_HIP(hipStreamSynchronize(st));
  }
  RDRT_HIP(hipEventRecord(c->ev0, st));
  if (a->mode == RDRT_BATCH_PERSISTENT) {
    // graph key: every parameter and buffer the captured launches use
    std::vector<uint8_t> key(sizeof(*a) + 8 * sizeof(void *) + sizeof(TraceTune));
    std::memcpy(key.data(), a, sizeof(*a));
    const void *ptrs[8] = {s.wrays.ptr, s.hitx.ptr, s.qa.ptr, s.qb.ptr, s.counts.ptr, s.tmp.ptr, s.blk.ptr,
                           s.work.ptr};
    std::memcpy(key.data() + sizeof(*a), ptrs, sizeof(ptrs));
    std::memcpy(key.data() + sizeof(*a) + sizeof(ptrs), &c->tune, sizeof(TraceTune));  // kernels captured
    if (!s.exec || key != s.cached) {
      if (s.exec) (void)hipGraphExecDestroy(s.exec);
      if (s.graph) (void)hipGraphDestroy(s.graph);
      s.exec = nullptr;
      s.graph = nullptr;
      uint32_t launches = 1;
      RDRT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      hipError_t e;
      if (c->tune.cull) {
        // round 0's queue from the cull pass: rays that miss their scene's
        // root box get miss records for every round and are not queued
        e = enqueue_batch_init((uint32_t *)s.qa.ptr, 0u, 0u, (uint32_t *)s.work.ptr, (uint32_t *)s.counts.ptr + 1,
                               pl.rounds, st);
        TraceParams tp = scene_trace_params(c);
        tp.rays = a->rays;
        tp.meta = a->ray_meta;
        tp.hits = a->hits;
        tp.hitx = (uint2 *)s.hitx.ptr;
        tp.n = a->n;
        tp.cull_mask = 0xffffffffu;
        CullScratch cs;
        cs.tmp = (uint32_t *)s.tmp.ptr;
        cs.blk = (uint32_t *)s.blk.ptr;
        cs.queue = (uint32_t *)s.qa.ptr;
        cs.count = (uint32_t *)s.counts.ptr;
        cs.later = a->hits;
        cs.later_rounds = pl.rounds - 1;
        cs.stride = a->n;
        if (e == hipSuccess) e = enqueue_cull(tp, cs, st);
      } else {
        e = enqueue_batch_init((uint32_t *)s.qa.ptr, 0u, a->n, (uint32_t *)s.counts.ptr, (uint32_t *)s.counts.ptr + 1,
                               pl.rounds, st);
      }
      if (e == hipSuccess) e = enqueue_scene(c, a, pl, 0, true, c->tune.cull != 0, launches);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(st, &g);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(RDRT_ERR_HIP, std::string("batch capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
      }
      s.graph = g;
      RDRT_HIP(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
      s.cached = key;
      s.launches = launches;
    }
    RDRT_HIP(hipGraphLaunch(s.exec, st));
  } else {
    uint32_t launches = 1;
    // identity queue over all rays (scene k owns [off[k], off[k+1])); the count
    // slot written is the unused work word
    RDRT_HIP(enqueue_batch_init((uint32_t *)s.qa.ptr, 0u, a->n, (uint32_t *)s.work.ptr, nullptr, 0u, st));
    for (uint32_t k = 0; k < pl.nscenes; k++) RDRT_HIP(enqueue_scene(c, a, pl, k, false, false, launches));
    s.launches = launches;
  }
  RDRT_HIP(hipEventRecord(c->ev1, st));
  c->call_timed = true;
  return RDRT_OK;
}

int32_t rdrt_batch_autotune(rdrt_ctx *c, const rdrt_batch_params *a, uint32_t reps, rdrt_tune *best) {
  if (!c || !a) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (a->mode != RDRT_BATCH_PERSISTENT) return fail(RDRT_ERR_INVALID_ARG, "the tuner times persistent batched calls");
  if (a->n == 0) return fail(RDRT_ERR_INVALID_ARG, "the sample needs rays");
  int32_t rc = need_device(c);
  if (rc) return rc;
  if (reps == 0) reps = 2;
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
  for (uint32_t block : {64u, 128u, 256u})
    for (uint32_t lds : {8u, 16u})
      for (uint32_t waves : {0u, 6u, 8u})
        for (uint32_t shape = 0; shape < 6 && rc == RDRT_OK; shape++) {
          TraceTune t;
          t.block = block;
          t.lds = lds;
          t.waves = waves;
          t.persist = before.persist;  // batched rounds always run persistent
          t.cull = shape & 1u;
          t.wide = shape >> 1;
          if (stack_entries_for(push_depth(c, t)) == 0) continue;
          c->tune = t;
          rc = rdrt_trace_batched(c, a);  // capture and first run
          if (rc == RDRT_OK && hipEventRecord(e0, c->work_stream) != hipSuccess) rc = fail(RDRT_ERR_HIP, "event");
          for (uint32_t r = 0; r < reps && rc == RDRT_OK; r++) rc = rdrt_trace_batched(c, a);
          if (rc != RDRT_OK) break;
          float ms = 0.0f;
          if (hipEventRecord(e1, c->work_stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(RDRT_ERR_HIP, "batch timing failed");
            break;
          }
          ms /= (float)reps;
          if (chosen_ms < 0.0f || ms < chosen_ms) {
            chosen_ms = ms;
            chosen = t;
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
  if (best) return rdrt_get_tune(c, best);
  return RDRT_OK;
}

int32_t rdrt_batch_get_stats(rdrt_ctx *c, rdrt_batch_stats *o) {
  if (!c || !o) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  int32_t rc = need_device(c);
  if (rc) return rc;
  BatchState &s = c->batch;
  if (!s.counts.ptr || s.rounds == 0) return fail(RDRT_ERR_STATE, "no batch traced");
  RDRT_HIP(hipStreamSynchronize(c->work_stream));
  std::memset(o, 0, sizeof(*o));
  std::vector<uint32_t> counts((size_t)s.nscenes * (s.rounds + 1));
  RDRT_HIP(hipMemcpy(counts.data(), s.counts.ptr, counts.size() * 4, hipMemcpyDeviceToHost));
  o->rounds = s.rounds;
  for (uint32_t k = 0; k < s.nscenes; k++)
    for (uint32_t r = 0; r < s.rounds; r++) o->rays[r] += counts[(size_t)k * (s.rounds + 1) + r];
  o->launches = s.launches;
  RDRT_HIP(hipEventElapsedTime(&o->ms, c->ev0, c->ev1));
  return RDRT_OK;
}

}  // extern "C"
