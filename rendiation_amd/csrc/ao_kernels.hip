// ao_kernels.hip -- the poll kernels of the ambient-occlusion pipeline on the
// device task graph (scene/rendering/gpu-ray-tracing/src/feature/
// ao.rs:192-304), and the trace task's poll around the traversal launch
// (shader/ray-tracing/src/backend/wavefront_compute/trace_task.rs).
#include "ao.h"
#include "pt_device.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace {

using pt::V3;
using pt::v3;
using tg::Graph;
using tg::Group;

// 1024-thread blocks: a pass takes one counter atomic per block and list,
// and same-address atomics serialize
constexpr uint32_t kBlock = 1024;
constexpr uint32_t kMaxGrid = 1024;
constexpr float kFltEps = 1.1920928955078125e-7f;  // ShaderRayRange::default (api/ty.rs:49-56)
constexpr float kFltMax = 3.402823466e+38f;

uint32_t grid_for(uint32_t n) {
  const uint32_t g = (n + kBlock - 1) / kBlock;
  return g < 1u ? 1u : (g > kMaxGrid ? kMaxGrid : g);
}

__device__ __forceinline__ void write_ray(rdrt_ray *r, V3 o, float tmin, V3 d, float tmax) {
  float4 *q = reinterpret_cast<float4 *>(r);
  q[0] = make_float4(o.x, o.y, o.z, tmin);
  q[1] = make_float4(d.x, d.y, d.z, tmax);
}
__device__ __forceinline__ V3 load3(const float *p) { return v3(p[0], p[1], p[2]); }

// Payload words.  Trace: [0] ray type (0 primary, 1 occlusion), [1] payload
// value, [2] the group of the awaited closest-hit / miss task.  Ray gen:
// [0] tile pixel.  Closest / miss: [0] payload value.
__device__ __forceinline__ uint32_t *pw(const Group &G, uint32_t idx) { return G.payload + (size_t)idx * G.words; }

// Starts a trace task (then_trace) for the lanes where `pred` holds: the
// ray goes to the trace group's ray array at the new task's slot, and the
// slot to the traversal queue of its ray type for the next trace poll (a
// new trace task's first poll is its traversal; every task spawned now is in
// that poll, so queueing it here is what that poll would do).  ray_type is
// the same in every lane.
__device__ __forceinline__ uint32_t spawn_trace(const AoLaunch &L, bool pred, uint32_t parent, uint32_t ptype,
                                                uint32_t ray_type, float value, V3 o, float tmin, V3 d, float tmax) {
  const Group &T = L.gr.g[kAoTrace];
  // tg::spawn, with its push to the active list and the push to the
  // traversal queue issued together
  const uint32_t c = tg::pop_empty(T, pred);
  const bool ok = c != tg::kNone;
  uint32_t pos[2];
  tg::block_add_n<2>({tg::active_count(T), L.tcount + 2 * L.qpar + ray_type}, {ok, ok}, pos);
  if (ok) {
    T.state[c] = tg::kWaken;
    T.parent[c] = parent;
    T.ptype[c] = ptype;
    T.stage[c] = 0u;
    T.child[c] = tg::kNone;
    if (pos[0] < T.size) tg::active_list(T, T.cur)[pos[0]] = c;
    else atomicOr(T.ctr + tg::kError, 2u);
    uint32_t *w = pw(T, c);
    w[0] = ray_type;
    w[1] = __float_as_uint(value);
    w[2] = tg::kNone;
    write_ray(L.rays + c, o, tmin, d, tmax);
    if (pos[1] < L.qcap) (ray_type ? L.tq1 : L.tq0)[pos[1]] = c;
    else atomicOr(T.ctr + tg::kError, 2u);  // more rays than the queue holds (not reached: one per pixel)
  }
  return c;
}

// ---- group 0: trace ----
// The poll after the traversal of the rays queued at spawn: each task just
// traced spawns the closest-hit task of its hit group (ao.rs:20-47: every
// instance maps ray type 0 to the occlusion closest hit, ray type 1 to the
// one that writes 0) or the miss task, passes it the payload and sleeps; a
// task woken by its finished closest-hit / miss task takes that payload,
// frees it and resolves.
__global__ void k_trace_poll(AoLaunch L) {
  const Group &T = L.gr.g[kAoTrace];
  if (blockIdx.x == 0 && threadIdx.x == 0) {  // nothing else writes these during the launch
    L.tcount[4] += L.tcount[2 * L.qpar];
    L.tcount[5] += L.tcount[2 * L.qpar + 1];
  }
  // the other parity's counts (last used by the previous round, whose
  // launches have all finished) start the queues of the next round at zero
  if (blockIdx.x == 0 && threadIdx.x < 2) L.tcount[2 * (L.qpar ^ 1u) + threadIdx.x] = 0u;
  tg::begin_phase(T);
  tg::for_each_active(T, [&](uint32_t, bool valid, uint32_t idx) {
    uint32_t *w = pw(T, idx);
    // every load that no block operation orders, issued up front: a poll is
    // a chain of dependent round trips, and each barrier ends a batch
    const uint32_t stage = valid ? T.stage[idx] : 1u;
    const uint32_t w0 = valid ? w[0] : 0u, w1 = valid ? w[1] : 0u, w2 = valid ? w[2] : tg::kNone;
    const uint32_t child = valid ? T.child[idx] : tg::kNone;
    const uint32_t parent = valid ? T.parent[idx] : tg::kNone, ptype = valid ? T.ptype[idx] : tg::kNone;
    const uint32_t prim = valid ? L.hits[idx].prim_id : RDRT_MISS;
    const bool fresh = valid && stage == 0u;
    const bool waiting = valid && !fresh;
    // the awaited closest-hit / miss task (payload word 0 = its value)
    const bool cvalid = waiting && w2 >= kAoClosest && w2 <= kAoMiss && child != tg::kNone;
    const uint32_t cstate = cvalid ? L.gr.g[w2].state[child] : tg::kNotExist;
    const uint32_t cval = cvalid ? pw(L.gr.g[w2], child)[0] : 0u;
    // a task just traced: the closest-hit task of its hit group or the miss task
    const bool hit = fresh && prim != RDRT_MISS;
    const uint32_t cg = !fresh ? tg::kNone : hit ? (w0 == 0u ? kAoClosest : kAoClosestTest) : kAoMiss;
    uint32_t c = tg::kNone;
    for (uint32_t g = kAoClosest; g <= kAoMiss; g++) {
      const bool mine = cg == g;
      if (!__syncthreads_or(mine)) continue;
      const uint32_t cc = tg::spawn(L.gr.g[g], mine, idx, kAoTrace);
      if (mine) c = cc;
    }
    if (fresh) {
      if (c != tg::kNone) pw(L.gr.g[cg], c)[0] = w1;
      w[2] = cg;
      T.child[idx] = c;
      T.stage[idx] = 1u;
    }
    // a waiting task whose child finished: takes its value, frees it
    // (TaskFuture poll, task_group.rs:464-488), resolves and wakes its
    // parent (a ray-gen or closest-hit task, asleep in the other phase);
    // the others sleep
    const bool done = cvalid && cstate == tg::kFinished;
    if (done) w[1] = cval;
    if (valid) T.state[idx] = done ? tg::kFinished : tg::kSleep;
    const Group &C2 = L.gr.g[kAoClosest], &C3 = L.gr.g[kAoClosestTest], &C4 = L.gr.g[kAoMiss];
    const Group &P1 = L.gr.g[kAoRayGen], &P2 = L.gr.g[kAoClosest];
    uint32_t pos[5];
    tg::block_add_n<5>({tg::removed_count(C2), tg::removed_count(C3), tg::removed_count(C4), tg::active_count(P1),
                        tg::active_count(P2)},
                       {done && w2 == kAoClosest, done && w2 == kAoClosestTest, done && w2 == kAoMiss,
                        done && ptype == kAoRayGen, done && ptype == kAoClosest},
                       pos);
    tg::free_at(C2, done && w2 == kAoClosest, child, pos[0]);
    tg::free_at(C3, done && w2 == kAoClosestTest, child, pos[1]);
    tg::free_at(C4, done && w2 == kAoMiss, child, pos[2]);
    tg::wake_at(P1, done && ptype == kAoRayGen, parent, pos[3]);
    tg::wake_at(P2, done && ptype == kAoClosest, parent, pos[4]);
  });
}

// ---- group 1: ray generation (ao.rs:192-245) ----
// PCGRandomSampler (scene/rendering/gpu-ray-tracing/src/sampler.rs:11-16,
// 48-71): seeded by xxhash32(pixel, sample index); next() maps the top 23
// bits to [0, 1).
__device__ __forceinline__ float pcg_next(uint32_t &state) {
  const uint32_t prev = state * 747796405u + 2891336453u;
  const uint32_t word = ((prev >> ((prev >> 28u) + 4u)) ^ prev) * 277803737u;
  state = prev;
  const uint32_t r = (word >> 22u) ^ word;
  return __uint_as_float(0x3f800000u | (r >> 9u)) - 1.0f;
}

__device__ __forceinline__ void accumulate(const AoLaunch &L, uint32_t x, uint32_t y, uint32_t s, float payload) {
  const size_t px = (size_t)y * L.width + x;
  const float prev_count = (float)s;
  const float all = prev_count + 1.0f;
  if (L.format == RDRT_AO_FORMAT_F32) {
    float *a = static_cast<float *>(L.accum) + px;
    const float prev = *a;
    if (s < kAoMaxSample) *a = (prev * prev_count + payload) / all;
    return;
  }
  // RGBA8 unorm like the reference's texture: the stored value is read back
  // as k / 255 and written rounded to the nearest of those
  uchar4 *a = static_cast<uchar4 *>(L.accum) + px;
  const uchar4 old = *a;
  const float prev = (float)old.x / 255.0f;
  float v = s < kAoMaxSample ? (prev * prev_count + payload) / all : prev;
  v = fminf(fmaxf(v, 0.0f), 1.0f);
  const unsigned char q = (unsigned char)rintf(v * 255.0f);
  *a = make_uchar4(q, q, q, 255);
}

__device__ __forceinline__ void poll_raygen(const AoLaunch &L) {
  const Group &G = L.gr.g[kAoRayGen];
  const uint32_t s = __float_as_uint(L.frame[23]);
  const Group &T = L.gr.g[kAoTrace];
  tg::for_each_active(G, [&](uint32_t, bool valid, uint32_t idx) {
    const uint32_t tp = valid ? pw(G, idx)[0] : 0u;
    const uint32_t x = L.tile_x + tp % L.tile_w, y = L.tile_y + tp / L.tile_w;
    const uint32_t stage = valid ? G.stage[idx] : 1u, child = valid ? G.child[idx] : tg::kNone;
    const bool first = valid && stage == 0u;
    // the awaited trace (loaded before the block operations below)
    const bool tvalid = valid && !first && child != tg::kNone;
    const uint32_t tstate = tvalid ? T.state[child] : tg::kNotExist;
    const uint32_t tval = tvalid ? pw(T, child)[1] : 0u;
    V3 cam = v3(0.0f, 0.0f, 0.0f), d = v3(0.0f, 0.0f, 1.0f);
    if (first) {
      uint32_t state = pt::xxhash32(x, y, s);
      const float r0 = pcg_next(state);
      const float r1 = pcg_next(state);
      // camera.rs:66-99
      const float u = (float)x / (float)L.width + r0 / (float)L.width;
      const float v = (float)y / (float)L.height + r1 / (float)L.height;
      const V3 target = pt::unproject(L.frame, u, v);
      cam = load3(L.frame + 16);
      d = pt::normalize3(target - cam);
    }
    // payload 0: occluded unless the miss shader writes 1 (ao.rs:216-219)
    const uint32_t c = spawn_trace(L, first, idx, kAoRayGen, 0u, 0.0f, cam, kFltEps, d, kFltMax);
    if (first) {
      G.child[idx] = c;
      G.stage[idx] = 1u;
    }
    // a finished trace: its value goes into the running mean, the trace and
    // this root are freed (a root resolves without a parent)
    const bool done = tvalid && tstate == tg::kFinished;
    if (done) accumulate(L, x, y, s, __uint_as_float(tval));
    if (valid && !done) G.state[idx] = tg::kSleep;
    uint32_t pos[2];
    tg::block_add_n<2>({tg::removed_count(T), tg::removed_count(G)}, {done, done}, pos);
    tg::free_at(T, done, child, pos[0]);
    tg::free_at(G, done, idx, pos[1]);
  });
}

// ---- group 2: closest hit of primary rays (ao.rs:252-291) ----
// World position o + t d, shading normal from the bindless mesh
// (bindless_mesh_bridge.rs:86-123), the occlusion direction tbn(n) * the
// frame's cosine-weighted hemisphere sample, range [0.01, 100], first hit.
__device__ __forceinline__ void occlusion_ray(const AoLaunch &L, uint32_t p, V3 &pos, V3 &dir) {
  const rdrt_hit h = L.hits[p];  // the parent trace task's ray and hit
  const uint2 hx = L.hitx[p];
  const float4 *rq = reinterpret_cast<const float4 *>(L.rays + p);
  const float4 q0 = rq[0], q1 = rq[1];
  const V3 o = v3(q0.x, q0.y, q0.z), d = v3(q1.x, q1.y, q1.z);
  const GeoRec &g = L.geos[hx.x];
  const InstRec &in = L.insts[hx.y];
  const size_t tri = (size_t)g.primitive_start + h.prim_id;
  const uint32_t ia = L.indices[3 * tri], ib = L.indices[3 * tri + 1], ic = L.indices[3 * tri + 2];
  pos = o + d * h.t;  // api/ctx.rs:209-211
  const V3 pa = load3(L.vertices + 3 * (size_t)ia), pb = load3(L.vertices + 3 * (size_t)ib),
           pc = load3(L.vertices + 3 * (size_t)ic);
  const float *m = in.inv;  // normal matrix: transpose of the world-to-object 3x3
  auto nmat = [m](V3 n) {
    return v3(m[0] * n.x + m[1] * n.y + m[2] * n.z, m[4] * n.x + m[5] * n.y + m[6] * n.z,
              m[8] * n.x + m[9] * n.y + m[10] * n.z);
  };
  V3 gn = pt::normalize3(nmat(pt::cross3(pa - pb, pa - pc)));
  if (pt::dot3(o - pos, gn) < 0.0f) gn = gn * -1.0f;
  V3 sn = gn;
  if (L.normals) {
    const float w = 1.0f - h.u - h.v;
    const V3 n = load3(L.normals + 3 * (size_t)ia) * w + load3(L.normals + 3 * (size_t)ib) * h.u +
                 load3(L.normals + 3 * (size_t)ic) * h.v;
    if (pt::dot3(n, n) > 0.0f) {
      sn = pt::normalize3(nmat(n));
      if (pt::dot3(gn, sn) < 0.0f) sn = sn * -1.0f;
    }
  }
  dir = pt::tbn_apply(sn, load3(L.frame + 20));
}

__device__ __forceinline__ void poll_closest(const AoLaunch &L) {
  const Group &G = L.gr.g[kAoClosest];
  const Group &T = L.gr.g[kAoTrace];
  tg::for_each_active(G, [&](uint32_t, bool valid, uint32_t idx) {
    const uint32_t stage = valid ? G.stage[idx] : 1u, child = valid ? G.child[idx] : tg::kNone;
    const uint32_t parent = valid ? G.parent[idx] : tg::kNone;
    const bool first = valid && stage == 0u;
    const bool tvalid = valid && !first && child != tg::kNone;
    const uint32_t tstate = tvalid ? T.state[child] : tg::kNotExist;
    const uint32_t tval = tvalid ? pw(T, child)[1] : 0u;
    V3 pos = v3(0.0f, 0.0f, 0.0f), dir = v3(0.0f, 0.0f, 1.0f);
    if (first) occlusion_ray(L, parent, pos, dir);
    // payload 1: unoccluded unless the occlusion closest hit writes 0
    const uint32_t c = spawn_trace(L, first, idx, kAoClosest, 1u, 1.0f, pos, 0.01f, dir, 100.0f);
    if (first) {
      G.child[idx] = c;
      G.stage[idx] = 1u;
    }
    // a finished occlusion trace: its value becomes this task's payload,
    // the trace is freed, this task resolves and wakes its camera trace
    const bool done = tvalid && tstate == tg::kFinished;
    if (done) pw(G, idx)[0] = tval;
    if (valid) G.state[idx] = done ? tg::kFinished : tg::kSleep;
    uint32_t p2[2];
    tg::block_add_n<2>({tg::removed_count(T), tg::active_count(T)}, {done, done}, p2);
    tg::free_at(T, done, child, p2[0]);
    tg::wake_at(T, done, parent, p2[1]);
  });
}

// ---- groups 3 and 4: store a constant and resolve (ao.rs:299-314) ----
__device__ __forceinline__ void poll_store(const AoLaunch &L, uint32_t gid, float value) {
  const Group &G = L.gr.g[gid];
  const Group &T = L.gr.g[kAoTrace];
  tg::for_each_active(G, [&](uint32_t, bool valid, uint32_t idx) {
    const uint32_t parent = valid ? G.parent[idx] : tg::kNone;
    if (valid) {
      pw(G, idx)[0] = __float_as_uint(value);
      G.state[idx] = tg::kFinished;
    }
    // resolves and wakes its trace (asleep in the other phase)
    const bool wakes = valid && parent != tg::kNone;  // always: store tasks are spawned by traces
    uint32_t pos[1];
    tg::block_add_n<1>({tg::active_count(T)}, {wakes}, pos);
    tg::wake_at(T, wakes, parent, pos[0]);
  });
}

// Groups 1-4 of one round in one launch.  Polled in order they would run
// one launch each; here every block polls its share of each group in turn.
// The result is the same: these groups spawn only trace tasks (group 0),
// wake only trace tasks, and read back only finished trace tasks, which no
// poll of groups 1-4 finishes; none of them spawns, wakes or frees a task of
// another of them, so no poll sees what another one does in this round.
__global__ void k_poll_stages(AoLaunch L) {
  for (uint32_t g = kAoRayGen; g <= kAoMiss; g++) tg::begin_phase(L.gr.g[g]);
  if (blockIdx.x == 0 && threadIdx.x == 0) *L.work = 0u;  // for the next round's traversal
  poll_raygen(L);
  poll_closest(L);
  poll_store(L, kAoClosestTest, 0.0f);
  poll_store(L, kAoMiss, 1.0f);
}

// The reset and round 1 of a tile in one launch.  After the reset (every
// pool empty; root k of the tile's n ray-gen roots in slot k + 1 and active
// entry k, payload k) round 1 polls only those roots, each for the first
// time: it spawns the camera trace, queues its ray and sleeps.  With the
// trace pool empty its slots come off the empty stack in order, so root k's
// trace takes slot k + 1, active entry k and queue entry k; this launch
// writes that state directly (what tg::enqueue_reset_spawn and the first
// k_poll_stages leave, in one pass).
__global__ void k_ao_begin(AoLaunch L, uint32_t n, uint32_t nzero) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nzero) L.tcount[t] = t == 0 ? n : 0u;  // the camera rays queued for round 2
  if (t == 0) *L.work = 0u;                        // for round 2's traversal
  const uint32_t s = __float_as_uint(L.frame[23]);
  for (uint32_t g = 0; g < L.gr.n; g++) {
    const Group &G = L.gr.g[g];
    const uint32_t kk = (g == kAoRayGen || g == kAoTrace) ? n : 0u;
    for (uint32_t i = t; i <= G.size; i += stride) {
      if (i >= 1u && i <= kk) {
        const uint32_t tp = i - 1u;  // the root's tile pixel
        if (g == kAoRayGen) {        // asleep on its camera trace
          G.state[i] = tg::kSleep;
          G.parent[i] = tg::kNone;
          G.ptype[i] = tg::kNone;
          G.stage[i] = 1u;
          G.child[i] = i;
          pw(G, i)[0] = tp;
        } else {  // the camera trace, awake and queued (spawn_trace)
          const uint32_t x = L.tile_x + tp % L.tile_w, y = L.tile_y + tp / L.tile_w;
          uint32_t state = pt::xxhash32(x, y, s);
          const float r0 = pcg_next(state);
          const float r1 = pcg_next(state);
          // camera.rs:66-99
          const float u = (float)x / (float)L.width + r0 / (float)L.width;
          const float v = (float)y / (float)L.height + r1 / (float)L.height;
          const V3 target = pt::unproject(L.frame, u, v);
          const V3 cam = load3(L.frame + 16);
          const V3 d = pt::normalize3(target - cam);
          G.state[i] = tg::kWaken;
          G.parent[i] = i;
          G.ptype[i] = kAoRayGen;
          G.stage[i] = 0u;
          G.child[i] = tg::kNone;
          uint32_t *w = pw(G, i);
          w[0] = 0u;  // primary ray; payload 0 unless the miss writes 1 (ao.rs:216-219)
          w[1] = __float_as_uint(0.0f);
          w[2] = tg::kNone;
          write_ray(L.rays + i, cam, kFltEps, d, kFltMax);
          L.tq0[tp] = i;
        }
        G.list0[tp] = i;
      } else {
        G.state[i] = tg::kNotExist;
      }
      if (i < G.size) G.empty[i] = G.size - i;
    }
    if (blockIdx.x == 0 && threadIdx.x < tg::kCtrWords)
      G.ctr[threadIdx.x] = threadIdx.x == tg::kEmpty0 ? G.size - kk : threadIdx.x == tg::kActive0 ? kk : 0u;
  }
}

// Rounds 4 and 5 of a tile in one launch, for executions of five rounds.
// When round 4 starts no task will be spawned any more (the ray gen spawns
// its camera trace in round 1, the closest hit its occlusion trace in round
// 2, both traced by round 3), and the only awake tasks are the occlusion
// traces of the pixels whose camera ray hit, woken in round 3 by their
// finished closest-hit-test / miss task.  Each such pixel's chain takes
// exactly the four polls left, one a phase: the occlusion trace takes its
// child's value and resolves (round 4, trace phase), the closest hit it
// wakes takes the value and resolves (round 4, stage phase), the camera
// trace it wakes does the same (round 5, trace phase), and the ray gen adds
// the value to the running mean (round 5, stage phase).  A chain touches only
// its own five tasks, so one thread runs the four polls of one chain in
// turn, loading the whole chain up front, and frees its five tasks as those
// polls would: pixel values and task states come out as after the two
// rounds, and every active list ends empty.  A task that does not fit this
// shape is not polled and sets error bit 3 of the trace group.
__global__ void k_finish_chains(AoLaunch L) {
  const Group &T = L.gr.g[kAoTrace], &G = L.gr.g[kAoRayGen], &C2 = L.gr.g[kAoClosest];
  const Group &C3 = L.gr.g[kAoClosestTest], &C4 = L.gr.g[kAoMiss];
  for (uint32_t g = kAoTrace; g <= kAoMiss; g++) tg::begin_phase(L.gr.g[g]);
  const uint32_t s = __float_as_uint(L.frame[23]);
  tg::for_each_active(T, [&](uint32_t, bool valid, uint32_t idx) {
    // the occlusion trace: its finished child and its parent (a closest hit)
    const uint32_t *w = pw(T, idx);
    const uint32_t stage = valid ? T.stage[idx] : 0u, w0 = valid ? w[0] : 0u, w2 = valid ? w[2] : tg::kNone;
    const uint32_t child = valid ? T.child[idx] : tg::kNone;
    const uint32_t p = valid ? T.parent[idx] : tg::kNone, ptype = valid ? T.ptype[idx] : tg::kNone;
    const Group &CG = w2 == kAoMiss ? C4 : C3;
    bool ok = valid && stage == 1u && w0 == 1u && (w2 == kAoClosestTest || w2 == kAoMiss) && child >= 1u &&
              child <= CG.size && ptype == kAoClosest && p >= 1u && p <= C2.size;
    const uint32_t cstate = ok ? CG.state[child] : tg::kNotExist, cval = ok ? pw(CG, child)[0] : 0u;
    const uint32_t pstate = ok ? C2.state[p] : tg::kNotExist, pstage = ok ? C2.stage[p] : 0u;
    const uint32_t pchild = ok ? C2.child[p] : tg::kNone;
    const uint32_t q = ok ? C2.parent[p] : tg::kNone, qtype = ok ? C2.ptype[p] : tg::kNone;
    ok = ok && cstate == tg::kFinished && pstate == tg::kSleep && pstage == 1u && pchild == idx &&
         qtype == kAoTrace && q >= 1u && q <= T.size;
    // the camera trace, asleep on the closest hit, and its ray gen
    const uint32_t *wq = pw(T, ok ? q : 0u);
    const uint32_t qstate = ok ? T.state[q] : tg::kNotExist, qstage = ok ? T.stage[q] : 0u;
    const uint32_t qchild = ok ? T.child[q] : tg::kNone, qw0 = ok ? wq[0] : 1u, qw2 = ok ? wq[2] : tg::kNone;
    const uint32_t r = ok ? T.parent[q] : tg::kNone, rtype = ok ? T.ptype[q] : tg::kNone;
    ok = ok && qstate == tg::kSleep && qstage == 1u && qchild == p && qw0 == 0u && qw2 == kAoClosest &&
         rtype == kAoRayGen && r >= 1u && r <= G.size;
    const uint32_t rstate = ok ? G.state[r] : tg::kNotExist, rstage = ok ? G.stage[r] : 0u;
    const uint32_t rchild = ok ? G.child[r] : tg::kNone, rparent = ok ? G.parent[r] : 0u;
    const uint32_t tp = ok ? pw(G, r)[0] : 0u;
    ok = ok && rstate == tg::kSleep && rstage == 1u && rchild == q && rparent == tg::kNone;
    if (valid && !ok) atomicOr(T.ctr + tg::kError, 8u);
    // the occlusion value (0 from the closest hit test, 1 from the miss)
    // passes up the chain into the running mean
    if (ok) accumulate(L, L.tile_x + tp % L.tile_w, L.tile_y + tp / L.tile_w, s, __uint_as_float(cval));
    uint32_t pos[6];
    tg::block_add_n<6>({tg::removed_count(C3), tg::removed_count(C4), tg::removed_count(T), tg::removed_count(T),
                        tg::removed_count(C2), tg::removed_count(G)},
                       {ok && w2 == kAoClosestTest, ok && w2 == kAoMiss, ok, ok, ok, ok}, pos);
    tg::free_at(C3, ok && w2 == kAoClosestTest, child, pos[0]);
    tg::free_at(C4, ok && w2 == kAoMiss, child, pos[1]);
    tg::free_at(T, ok, idx, pos[2]);
    tg::free_at(T, ok, q, pos[3]);
    tg::free_at(C2, ok, p, pos[4]);
    tg::free_at(G, ok, r, pos[5]);
  });
}

}  // namespace

hipError_t enqueue_ao_tile(const AoLaunch &L0, uint32_t tile_h, uint32_t rounds, const TraceParams &scene,
                           uint32_t push_depth, const TraceTune &tune, uint32_t *work, bool first_tile,
                           hipStream_t st, tg::Graph *after) {
  AoLaunch L = L0;  // its graph's parities advance as launches are enqueued
  Graph &gr = L.gr;
  const uint32_t nroots = L.tile_w * tile_h;
  hipError_t e = hipSuccess;
  L.qpar = 0u;
  const uint32_t ntrace = gr.g[kAoTrace].size;
  TraceParams tp = scene;
  tp.rays = L.rays;
  tp.hits = L.hits;
  tp.hitx = L.hitx;
  tp.meta = nullptr;
  tp.hit_meta = nullptr;
  tp.counters = nullptr;
  tp.n = ntrace + 1;
  tp.queue = L.tq0;
  tp.count = L.tcount;  // per round below
  // each queue holds at most one ray per pixel of the tile, and so do both
  // together: a pixel's chain has one trace task at a time in traversal
  const uint32_t per_queue = L.tile_w * tile_h;
  L.qcap = per_queue;
  tp.slots = per_queue;
  tp.live_max = per_queue;
  tp.flags = RDRT_RAY_FLAG_CULL_BACK_FACING_TRIANGLES;  // ao.rs:209
  tp.cull_mask = 0xffffffffu;
  tp.tlas_idx = 0;
  tp.work = work;
  tp.work_zeroed = 1u;  // by k_poll_stages, the launch before each traversal
  L.work = work;
  tp.rays2 = L.rays;
  tp.hits2 = L.hits;
  tp.hitx2 = L.hitx;
  tp.queue2 = L.tq1;
  tp.count2 = L.tcount + 1;  // per round below
  tp.slots2 = per_queue;
  tp.flags2 = RDRT_RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH;  // ao.rs:276
  // Each round polls the trace group (the traversal of the queued rays,
  // then k_trace_poll), then groups 1-4 (one launch, see k_poll_stages), as
  // the executor's round does (mod.rs:396-460).  The two are the phases of
  // task_graph.h, so no list launch runs between them: the trace poll
  // spawns, wakes and frees only tasks of groups 1-4, and their polls only
  // trace tasks; each pool holds every task of a tile (a pixel's chain has
  // one task of each group but the trace group, which holds two per pixel),
  // so no freed slot is taken again before the next tile's reset.  The
  // first round of a tile skips the trace group: right after the reset and
  // the root spawns it holds no task.  Only rounds 2 and 3 trace: the camera
  // rays that the ray gen spawns in round 1, the occlusion rays that the
  // closest hit spawns in round 2; no later poll spawns a trace, so later
  // trace polls run without the traversal.
  constexpr uint32_t kStages = (1u << kAoRayGen) | (1u << kAoClosest) | (1u << kAoClosestTest) | (1u << kAoMiss);
  constexpr uint32_t kTrace = 1u << kAoTrace;
  uint32_t stage_grid = 1;
  for (uint32_t g = kAoRayGen; g <= kAoMiss; g++) stage_grid = std::max(stage_grid, grid_for(gr.g[g].size));
  uint32_t r0 = 0;
  if (rounds >= 1) {
    // empty graph, the tile's ray-gen roots, the queue counts at zero (and
    // the call's ray totals on its first tile), and round 1 (k_ao_begin)
    if (nroots > gr.g[kAoRayGen].size || nroots > gr.g[kAoTrace].size || nroots > L.qcap) return hipErrorInvalidValue;
    for (uint32_t k = 0; k < gr.n; k++) gr.g[k].cur = gr.g[k].pp = 0u;
    hipLaunchKernelGGL(k_ao_begin, dim3(grid_for(gr.g[kAoTrace].size + 1)), dim3(kBlock), 0, st, L, nroots,
                       first_tile ? 6u : 4u);
    tg::end_phase(gr, kStages);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    r0 = 1;
  } else {
    e = tg::enqueue_reset_spawn(gr, kAoRayGen, nroots, L.tcount, first_tile ? 6u : 4u, st);
    if (e != hipSuccess) return e;
  }
  for (uint32_t r = r0; r < rounds; r++) {
    if (r == 3 && rounds == kAoChainRounds) {  // rounds 4 and 5 (k_finish_chains)
      hipLaunchKernelGGL(k_finish_chains, dim3(grid_for(ntrace)), dim3(kBlock), 0, st, L);
      tg::end_phase(gr, kTrace | kStages);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
      break;
    }
    if (r > 0) {
      if (r <= 2) {
        TraceParams q = tp;
        q.count = L.tcount + 2 * L.qpar;
        q.count2 = L.tcount + 2 * L.qpar + 1;
        e = launch_trace(q, RDRT_TRAVERSAL_FAST, push_depth, tune, st);
        if (e != hipSuccess) return e;
      }
      hipLaunchKernelGGL(k_trace_poll, dim3(grid_for(ntrace)), dim3(kBlock), 0, st, L);
      tg::end_phase(gr, kTrace);
      L.qpar ^= 1u;  // spawns from here on queue for the next trace poll
    }
    hipLaunchKernelGGL(k_poll_stages, dim3(stage_grid), dim3(kBlock), 0, st, L);
    tg::end_phase(gr, kStages);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (after) *after = gr;
  return hipSuccess;
}

}  // namespace rdrt
