// pipeline_kernels.hip -- the library's side of a programmable pipeline on
// the device task graph: the trace task (traversal, then the closest-hit or
// miss task of the ray, then its payload back to the caller;
// wavefront_compute/trace_task.rs:149-493) and the pre / post kernels around
// each caller stage (task_group.rs:118-150: a polled task resolves or goes
// to sleep on the trace it spawned; future/then.rs: a finished child is read
// back and freed).
#include "pipeline.h"

namespace rdrt {
namespace {

using tg::Graph;
using tg::Group;

// 1024-thread blocks: each pass takes one counter atomic per block and list,
// and same-address atomics serialize, so fewer, larger blocks finish sooner.
// At most 512 of them (about as many as the chip holds at once: the passes
// loop over larger lists): a pass over an empty group, as in the rounds
// after a pipeline has drained, then costs a few microseconds instead of
// the dispatch of a grid sized for the whole pool (1.86 -> 1.59 ms for the
// 1080p pipeline test, whose last five of eight rounds find no task)
constexpr uint32_t kBlock = 1024;
constexpr uint32_t kMaxGrid = 512;
constexpr uint32_t kStageBlock = 256;  // the caller's stage kernels

uint32_t grid_for(uint32_t n) {
  const uint32_t g = (n + kBlock - 1) / kBlock;
  return g < 1u ? 1u : (g > kMaxGrid ? kMaxGrid : g);
}

__device__ __forceinline__ void copy_words(uint32_t *dst, const uint32_t *src, uint32_t n) {
  for (uint32_t k = 0; k < n; k++) dst[k] = src[k];
}

// ---- trace group ----
// New trace tasks go to the traversal queue; a task woken by its finished
// closest-hit / miss task takes that payload, frees it and resolves.
__global__ void k_pipe_trace_pre(PipeLaunch L) {
  const Group &T = L.gr.g[kPipeTrace];
  uint32_t *cnt = L.tcount + L.qpar;
  // the other parity's count (used by the previous round, all of whose
  // launches have finished) starts the next round at zero, and so does the
  // work counter of the traversal that follows
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    L.tcount[L.qpar ^ 1u] = 0u;
    *L.work = 0u;
  }
  tg::begin_phase(T);
  tg::refill_empty(T);  // trace slots are taken and freed only by stage polls
  tg::for_each_active(T, [&](uint32_t, bool valid, uint32_t idx) {
    const bool fresh = valid && T.stage[idx] == 0u;
    const uint32_t pos = tg::block_add(cnt, fresh);
    if (fresh) L.tq[pos] = idx;
    const bool waiting = valid && !fresh;
    const uint32_t cg = waiting ? L.tawait[idx] : tg::kNone;
    const uint32_t c = waiting ? T.child[idx] : tg::kNone;
    const bool done = waiting && cg < L.gr.n && c != tg::kNone && L.gr.g[cg].state[c] == tg::kFinished;
    if (done) copy_words(L.tpay + (size_t)idx * L.words, L.st[cg].payload + (size_t)c * L.words, L.words);
    tg::cleanup_in(L.gr, done, cg, c);
    tg::end_poll(L.gr, kPipeTrace, waiting, idx, done, true);
  });
}

// After the traversal: each traced task spawns the closest-hit task of its
// hit group (offset + stride * geometry + instance offset, ctx.rs:53-55) or
// the miss task of its miss index, hands it the payload and sleeps; a ray
// whose group or miss index has no stage resolves with its payload as is.
__global__ void k_pipe_trace_post(PipeLaunch L) {
  const Group &T = L.gr.g[kPipeTrace];
  const uint32_t n = L.tcount[L.qpar];
  if (blockIdx.x == 0 && threadIdx.x == 0) L.tcount[2] += n;  // nothing else writes it during the launch
  for (uint32_t j0 = blockIdx.x * blockDim.x; j0 < n; j0 += gridDim.x * blockDim.x) {
    const uint32_t j = j0 + threadIdx.x;
    const bool valid = j < n;
    const uint32_t idx = valid ? L.tq[j] : 0u;
    uint32_t cg = tg::kNone;
    if (valid) {
      const uint4 sb = L.sbt[idx];
      if (L.hits[idx].prim_id != RDRT_MISS) {
        const uint32_t hg = sb.x + sb.y * L.hit_meta[idx].geometry_id + L.insts[L.hitx[idx].y].sbt_offset;
        cg = hg < L.nclosest ? L.closest_group[hg] : tg::kNone;
      } else {
        cg = sb.z < L.nmiss ? L.miss_group[sb.z] : tg::kNone;
      }
    }
    uint32_t c = tg::kNone;
    for (uint32_t g = kPipeRayGen + 1; g < L.gr.n; g++) {
      const bool mine = valid && cg == g;
      if (!__syncthreads_or(mine)) continue;
      const uint32_t cc = tg::spawn(L.gr.g[g], mine, idx, kPipeTrace);
      if (mine) c = cc;
    }
    if (c != tg::kNone) {
      const StageArrays &S = L.st[cg];
      copy_words(S.payload + (size_t)c * L.words, L.tpay + (size_t)idx * L.words, L.words);
      S.launch[c] = L.tlaunch[idx];
      S.trace_of[c] = idx;
    }
    if (valid) {
      L.tawait[idx] = cg;
      T.child[idx] = c;
      T.stage[idx] = 1u;
    }
    // no stage for the ray, or no slot left for its task (error bit set): resolve now
    tg::end_poll(L.gr, kPipeTrace, valid, idx, valid && c == tg::kNone, true);
  }
}

// ---- stage groups ----
// Before the caller's kernel: a ray-gen root at its first poll turns its
// spawn index into the launch index and zeroes its payload; a task whose
// trace has finished takes the trace's payload into trace_result, frees it
// and counts the poll; every polled task's action is preset to resolve.
__device__ __forceinline__ void stage_pre(const PipeLaunch &L, uint32_t gid) {
  const Group &G = L.gr.g[gid];
  const Group &T = L.gr.g[kPipeTrace];
  const StageArrays &S = L.st[gid];
  tg::for_each_active(G, [&](uint32_t, bool valid, uint32_t idx) {
    const uint32_t c = valid ? G.child[idx] : tg::kNone;
    if (gid == kPipeRayGen && valid && G.stage[idx] == 0u && c == tg::kNone) {
      const uint32_t k = G.payload[idx];  // spawn index within the tile
      S.launch[idx] = (L.tile_y + k / L.tile_w) * L.width + L.tile_x + k % L.tile_w;
      for (uint32_t w = 0; w < L.words; w++) S.payload[(size_t)idx * L.words + w] = 0u;
    }
    const bool done = c != tg::kNone && T.state[c] == tg::kFinished;
    if (done) {
      copy_words(S.result + (size_t)idx * L.words, L.tpay + (size_t)c * L.words, L.words);
      G.stage[idx] += 1u;
      G.child[idx] = tg::kNone;
    }
    tg::cleanup(T, done, c);
    if (valid) S.action[idx] = RDRT_STAGE_RESOLVE;
  });
}

// After the caller's kernel: a task that asked for a trace spawns the trace
// task (its ray, flags, hit-group config and payload) and sleeps on it;
// every other polled task resolves (its parent trace task wakes).
__device__ __forceinline__ void stage_post(const PipeLaunch &L, uint32_t gid) {
  const Group &G = L.gr.g[gid];
  const Group &T = L.gr.g[kPipeTrace];
  const StageArrays &S = L.st[gid];
  tg::for_each_active(G, [&](uint32_t, bool valid, uint32_t idx) {
    const bool tr = valid && S.action[idx] == RDRT_STAGE_TRACE;
    const uint32_t c = tg::spawn(T, tr, idx, gid);
    if (c != tg::kNone) {
      const rdrt_trace_call cl = S.call[idx];
      float4 *r = reinterpret_cast<float4 *>(L.rays + c);
      r[0] = make_float4(cl.origin[0], cl.origin[1], cl.origin[2], cl.tmin);
      r[1] = make_float4(cl.direction[0], cl.direction[1], cl.direction[2], cl.tmax);
      L.meta[c] = rdrt_ray_meta{cl.flags, cl.cull_mask, cl.tlas_idx, 0u};
      L.sbt[c] = make_uint4(cl.sbt_offset, cl.sbt_stride, cl.miss_index, 0u);
      copy_words(L.tpay + (size_t)c * L.words, S.call_payload + (size_t)idx * L.words, L.words);
      L.tlaunch[c] = S.launch[idx];
      L.tawait[c] = tg::kNone;
    }
    if (tr) G.child[idx] = c;
    // a trace that found no free slot (error bit set) resolves instead of
    // sleeping for ever; a stage task's parent is a trace task or none
    tg::end_poll_in(L.gr, gid, valid, idx, valid && (!tr || c == tg::kNone), kPipeTrace, true);
  });
}

// The pre passes of every stage group in one launch, and their post passes
// in another, with the caller's stage kernels between them in group order.
// Polled one group at a time, pre(g), stage(g), post(g) would run before
// pre(g + 1); the result is the same: stage tasks spawn, wake and read back
// only trace tasks, and no stage pass finishes a trace task or returns a
// freed trace slot to the empty stack, so what one stage group's passes do
// is invisible to another's in the same round.  The caller's kernels still
// run in group order.  The pre launch also fixes each group's snapshot
// (the count the caller's kernel reads) and returns the freed slots of the
// closest-hit and miss groups to their stacks: those slots are taken and
// freed only by trace polls.  Ray-gen roots are spawned once per tile.
__global__ void k_stages_pre(PipeLaunch L) {
  for (uint32_t g = kPipeRayGen; g < L.gr.n; g++) {
    const Group &G = L.gr.g[g];
    tg::begin_phase(G);
    if (g > kPipeRayGen) tg::refill_empty(G);
    if (blockIdx.x == 0 && threadIdx.x == 0) G.ctr[tg::kSnapshot] = min(*tg::active_count(G), G.size);
  }
  for (uint32_t g = kPipeRayGen; g < L.gr.n; g++) stage_pre(L, g);
}
__global__ void k_stages_post(PipeLaunch L) {
  for (uint32_t g = kPipeRayGen; g < L.gr.n; g++) stage_post(L, g);
}

}  // namespace

hipError_t enqueue_pipeline_tile(const PipeLaunch &L0, const void *const *stages, uint32_t tile_h, uint32_t rounds,
                                 const TraceParams &scene, const AttribArgs &attrib, uint32_t push_depth,
                                 const TraceTune &tune, uint32_t *work, void *user, bool first_tile, hipStream_t st,
                                 tg::Graph *after) {
  PipeLaunch L = L0;  // its graph's parities advance as launches are enqueued
  Graph &gr = L.gr;
  // empty graph, the tile's ray-gen roots, the trace counts at zero (and
  // the call's traced-ray total on its first tile)
  hipError_t e = tg::enqueue_reset_spawn(gr, kPipeRayGen, L.tile_w * tile_h, L.tcount, first_tile ? 3u : 2u, st);
  if (e != hipSuccess) return e;
  L.qpar = 0u;
  const uint32_t ntrace = gr.g[kPipeTrace].size;
  TraceParams tp = scene;
  tp.rays = L.rays;
  tp.meta = L.meta;
  tp.hits = L.hits;
  tp.hit_meta = L.hit_meta;
  tp.hitx = L.hitx;
  tp.counters = nullptr;
  tp.n = ntrace + 1;
  tp.queue = L.tq;
  tp.slots = ntrace;
  tp.work = work;
  tp.work_zeroed = 1u;  // by k_pipe_trace_pre, right before each traversal
  L.work = work;
  auto poll_trace = [&]() -> hipError_t {
    hipLaunchKernelGGL(k_pipe_trace_pre, dim3(grid_for(ntrace)), dim3(kBlock), 0, st, L);
    TraceParams q = tp;
    q.count = L.tcount + L.qpar;
    hipError_t e2 = launch_trace(q, RDRT_TRAVERSAL_FAST, push_depth, tune, st);
    if (e2 != hipSuccess) return e2;
    if (L.attrs || L.xforms) {
      AttribArgs a = attrib;
      a.rays = L.rays;
      a.meta = L.meta;
      a.hits = L.hits;
      a.hit_meta = L.hit_meta;
      a.sbt = L.sbt;
      a.queue = L.tq;
      a.count = q.count;
      a.n = ntrace;
      a.out = L.attrs;
      a.xf_out = L.xforms;
      e2 = enqueue_hit_attributes(a, st);
      if (e2 != hipSuccess) return e2;
    }
    hipLaunchKernelGGL(k_pipe_trace_post, dim3(grid_for(ntrace)), dim3(kBlock), 0, st, L);
    L.qpar ^= 1u;  // the next round's trace poll uses the other count
    return hipGetLastError();
  };
  auto poll_stages = [&]() -> hipError_t {
    uint32_t grid = 1;
    for (uint32_t g = kPipeRayGen; g < gr.n; g++) grid = std::max(grid, grid_for(gr.g[g].size));
    hipLaunchKernelGGL(k_stages_pre, dim3(grid), dim3(kBlock), 0, st, L);
    for (uint32_t g = kPipeRayGen; g < gr.n; g++) {
      const Group &G = gr.g[g];
      const StageArrays &S = L.st[g];
      rdrt_stage_args a;
      a.count = G.ctr + tg::kSnapshot;
      a.slots = G.cur ? G.list1 : G.list0;
      a.payload = S.payload;
      a.trace_result = S.result;
      a.polls = G.stage;
      a.launch = S.launch;
      a.trace_of = S.trace_of;
      a.rays = L.rays;
      a.ray_meta = L.meta;
      a.hits = L.hits;
      a.hit_meta = L.hit_meta;
      a.attributes = L.attrs;
      a.transforms = L.xforms;
      a.action = S.action;
      a.call = S.call;
      a.call_payload = S.call_payload;
      a.user = user;
      a.launch_size[0] = L.width;
      a.launch_size[1] = L.height;
      a.payload_words = L.words;
      a.capacity = G.size;
      void *args[] = {&a};
      hipError_t e2 =
          hipLaunchKernel(stages[g], dim3((G.size + kStageBlock - 1) / kStageBlock), dim3(kStageBlock), args, 0, st);
      if (e2 != hipSuccess) return e2;
    }
    hipLaunchKernelGGL(k_stages_post, dim3(grid), dim3(kBlock), 0, st, L);
    return hipGetLastError();
  };
  // Each round polls the trace group, then the stage groups (see
  // k_stages_pre), as the executor's round does (mod.rs:396-460).  The two
  // are the phases of task_graph.h, so no list launch runs between them:
  // the trace poll spawns, wakes and frees only stage tasks, and the stage
  // polls only trace tasks (and ray-gen roots, freed once).  The first
  // round of a tile skips the trace group: right after the reset and the
  // root spawns it holds no task, so its poll would do nothing.
  const uint32_t trace_mask = 1u << kPipeTrace;
  const uint32_t stage_mask = ((gr.n >= 32 ? 0u : 1u << gr.n) - 1u) & ~trace_mask;
  const uint32_t refill_mask = stage_mask & ~(1u << kPipeRayGen);
  for (uint32_t r = 0; r < rounds; r++) {
    if (r > 0) {
      e = poll_trace();
      if (e != hipSuccess) return e;
      tg::end_phase(gr, trace_mask);
      tg::flip_refilled(gr, trace_mask);
    }
    e = poll_stages();
    if (e != hipSuccess) return e;
    tg::end_phase(gr, stage_mask);
    tg::flip_refilled(gr, refill_mask);
  }
  if (after) *after = gr;
  return hipSuccess;
}

}  // namespace rdrt
