// task_graph.h -- device task graph: groups of tasks that spawn, sleep on,
// wake and finish each other, polled round by round on one stream
// (shader/task-graph/src/runtime/{mod.rs,task_group.rs,task_pool.rs}).
//
// A group owns a pool of task slots (slot 0 is never used, slots 1..size
// hold tasks), the per-slot state, parent reference, poll stage, awaited
// child and payload words, an active list of the slots to poll (two
// buffers; compaction writes the other one and flips a device word), a
// stack of empty slots and a list of slots freed since the group last ran.
// Every count lives in device memory, so a whole execution is enqueued
// without host synchronization and can be captured in a graph.
//
// One round runs, for each group in order: prepare (freed slots go back to
// the empty stack; the number of active entries to poll is fixed), poll
// (one thread per active entry), compact (tasks that are awake stay active;
// tasks that went to sleep leave the list).
#ifndef RDRT_TASK_GRAPH_H
#define RDRT_TASK_GRAPH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace rdrt {
namespace tg {

// Task states (task_pool.rs:85-107).
constexpr uint32_t kNotExist = 0;  // free slot
constexpr uint32_t kWaken = 1;     // spawned or woken: in the active list, to be polled
constexpr uint32_t kGoToSleep = 2;  // polled, pending, still in the active list until compaction
constexpr uint32_t kSleep = 3;      // pending, out of the active list
constexpr uint32_t kFinished = 4;   // resolved, waiting for its parent to read it back
constexpr uint32_t kNone = 0xffffffffu;  // no parent / no task

// Counter words of a group.  The active, empty and removed counts are
// double-buffered: a prepare or a compaction reads the current word and
// writes the other one, and the host flips the group's parities as it
// enqueues, so no launch needs a cross-block handshake to commit a count.
enum Ctr : uint32_t {
  kActive0 = 0,   // entries in active buffer 0 / 1
  kEmpty0 = 2,    // entries in the empty stack (parity pp)
  kRemoved0 = 4,  // slots freed since the group last ran (parity pp)
  kSnapshot = 6,  // active entries the current poll covers
  kError = 7,     // bit 0: empty stack exhausted; bit 1: active list full; bit 2: removed list full
  kDone = 8,      // blocks finished (last-block commit of the root spawn)
  kCtrWords = 12,
};

constexpr uint32_t kMaxGroups = 8;

// Device view of one group.
struct Group {
  uint32_t *state, *parent, *ptype, *stage, *child;  // per slot (size + 1)
  uint32_t *payload;                                 // words per slot
  uint32_t *list0, *list1;                           // active buffers (size each)
  uint32_t *empty, *removed;                         // size each
  uint32_t *ctr;
  uint32_t size, words;
  uint32_t cur, pp;  // host-kept parities: current active buffer; empty / removed counts
};

struct Graph {
  Group g[kMaxGroups];
  uint32_t n;
};

__device__ __forceinline__ uint32_t *active_count(const Group &G) { return G.ctr + kActive0 + G.cur; }
__device__ __forceinline__ uint32_t *empty_count(const Group &G) { return G.ctr + kEmpty0 + G.pp; }
__device__ __forceinline__ uint32_t *removed_count(const Group &G) { return G.ctr + kRemoved0 + G.pp; }

__device__ __forceinline__ uint32_t *active_list(const Group &G, uint32_t cur) { return cur ? G.list1 : G.list0; }

// ---- block-level list operations ----
// Every function below is called by all threads of a block together (in
// block-uniform control flow), each thread with its own predicate; a list's
// counter then takes one atomic per block instead of one per thread.  Blocks
// have at most kMaxBlock threads.

constexpr uint32_t kMaxBlock = 1024;

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ uint32_t rank_in(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// This thread's rank among the block's threads where `pred` holds, and
// their number (`total`, the same in every thread).
__device__ __forceinline__ uint32_t block_rank(bool pred, uint32_t &total) {
  __shared__ uint32_t wave_off[kMaxBlock / 64 + 1];
  const uint64_t m = __ballot(pred);
  const uint32_t wid = threadIdx.x >> 6;
  const uint32_t nw = (blockDim.x + 63) >> 6;
  if (lane_id() == 0) wave_off[wid] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < nw; w++) {
      const uint32_t c = wave_off[w];
      wave_off[w] = t;
      t += c;
    }
    wave_off[nw] = t;
  }
  __syncthreads();
  const uint32_t r = wave_off[wid] + rank_in(m);
  total = wave_off[nw];
  __syncthreads();  // the array is reused by the next call
  return r;
}

// N independent block_add calls in one pass: pos[k] = old value of ctr[k]
// + this thread's rank among the threads where pred[k] holds.  The N
// counter atomics are issued together (lane k of wave 0 takes counter k),
// so the block waits for one round trip instead of N.  A null counter must
// have pred false in every thread.
template <int N>
__device__ __forceinline__ void block_add_n(uint32_t *const (&ctr)[N], const bool (&pred)[N], uint32_t (&pos)[N]) {
  __shared__ uint32_t off[N][kMaxBlock / 64 + 1];
  __shared__ uint32_t base[N];
  const uint32_t wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  uint64_t m[N];
#pragma unroll
  for (int k = 0; k < N; k++) m[k] = __ballot(pred[k]);
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < N; k++) off[k][wid] = (uint32_t)__popcll(m[k]);
  }
  __syncthreads();
  if (threadIdx.x < (uint32_t)N) {
    const uint32_t k = threadIdx.x;
    uint32_t *c = ctr[0];
#pragma unroll
    for (int j = 1; j < N; j++)
      if (k == (uint32_t)j) c = ctr[j];
    uint32_t t = 0;
    for (uint32_t w = 0; w < nw; w++) {
      const uint32_t v = off[k][w];
      off[k][w] = t;
      t += v;
    }
    base[k] = t ? atomicAdd(c, t) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) pos[k] = base[k] + off[k][wid] + rank_in(m[k]);
  __syncthreads();  // the arrays are reused by the next call
}

// Old counter value + this thread's rank among the threads where `pred`
// holds (meaningful for those threads only).
__device__ __forceinline__ uint32_t block_add(uint32_t *ctr, bool pred) {
  __shared__ uint32_t base;
  uint32_t total;
  const uint32_t r = block_rank(pred, total);
  if (total == 0) return 0u;
  if (threadIdx.x == 0) base = atomicAdd(ctr, total);
  __syncthreads();
  const uint32_t b = base;
  __syncthreads();
  return b + r;
}

__device__ __forceinline__ void push_active(const Group &G, bool pred, uint32_t idx) {
  const uint32_t pos = block_add(active_count(G), pred);
  if (pred) {
    if (pos < G.size) active_list(G, G.cur)[pos] = idx;
    else atomicOr(G.ctr + kError, 2u);
  }
}

// Pops one empty slot per thread where `pred` holds; kNone where the stack
// ran out (error bit 0).  Pops never race with pushes: freed slots return to
// the stack only in the group's prepare.
__device__ __forceinline__ uint32_t pop_empty(const Group &G, bool pred) {
  __shared__ int old_sh;
  uint32_t total;
  const int r = (int)block_rank(pred, total);
  if (total == 0) return kNone;
  const int k = (int)total;
  if (threadIdx.x == 0) {
    const int old = (int)atomicSub(empty_count(G), (uint32_t)k);
    const int got = old < 0 ? 0 : (old < k ? old : k);
    if (got < k) {
      atomicAdd(empty_count(G), (uint32_t)(k - got));
      atomicOr(G.ctr + kError, 1u);
    }
    old_sh = old;
  }
  __syncthreads();
  const int old = old_sh;
  __syncthreads();
  if (!pred || r >= old) return kNone;
  return G.empty[old - 1 - r];
}

// spawn_new_task (task_group.rs:428-446, task_pool.rs:143-172): an empty
// slot becomes a woken task with the given parent and zeroed poll state, at
// the end of the active list.  Returns the slot or kNone; the caller writes
// the payload.
__device__ __forceinline__ uint32_t spawn(const Group &G, bool pred, uint32_t parent, uint32_t ptype) {
  const uint32_t idx = pop_empty(G, pred);
  const bool ok = idx != kNone;
  if (ok) {
    G.state[idx] = kWaken;
    G.parent[idx] = parent;
    G.ptype[idx] = ptype;
    G.stage[idx] = 0u;
    G.child[idx] = kNone;
  }
  push_active(G, ok, idx);
  return idx;
}

// wake_task_dyn (task_group.rs:448-462) of task pidx of group pg (per lane):
// a task still in its active list (going to sleep in this very round) is
// only marked awake.
__device__ __forceinline__ void wake(const Graph &gr, bool pred, uint32_t pg, uint32_t pidx) {
  bool push = false;
  if (pred && pg < gr.n) push = atomicExch(gr.g[pg].state + pidx, kWaken) != kGoToSleep;
  for (uint32_t g = 0; g < gr.n; g++) {
    const bool mine = push && pg == g;
    if (__syncthreads_or(mine)) push_active(gr.g[g], mine, pidx);
  }
}

// cleanup_finished_task_state_and_payload (task_group.rs:490-498).
__device__ __forceinline__ void cleanup(const Group &G, bool pred, uint32_t idx) {
  const uint32_t pos = block_add(removed_count(G), pred);
  if (pred) {
    if (pos < G.size) G.removed[pos] = idx;
    else atomicOr(G.ctr + kError, 4u);
    G.state[idx] = kNotExist;
  }
}
// The same for a per-lane group.
__device__ __forceinline__ void cleanup_in(const Graph &gr, bool pred, uint32_t g_of_lane, uint32_t idx) {
  for (uint32_t g = 0; g < gr.n; g++) {
    const bool mine = pred && g_of_lane == g;
    if (__syncthreads_or(mine)) cleanup(gr.g[g], mine, idx);
  }
}

// The end of a poll (task_group.rs:118-150): a resolved task finishes and
// wakes its parent (or frees itself when it has none); a pending one goes
// to sleep.  `valid` is false for lanes past the end of the work; `phased`
// as for end_poll_in.
__device__ __forceinline__ void end_poll(const Graph &gr, uint32_t gid, bool valid, uint32_t idx, bool resolved,
                                         bool phased = false) {
  const Group &G = gr.g[gid];
  uint32_t p = kNone, pt = kNone;
  if (valid) {
    if (resolved) {
      G.state[idx] = kFinished;
      p = G.parent[idx];
      pt = G.ptype[idx];
    } else {
      G.state[idx] = phased ? kSleep : kGoToSleep;
    }
  }
  const bool fin = valid && resolved;
  cleanup(G, fin && p == kNone, idx);
  wake(gr, fin && p != kNone, pt, p);
}

// end_poll for a group whose tasks' parents all belong to group pg (kNone:
// tasks without parents), which saves the per-group passes of wake.
// `phased`: the poll runs in a phase (below), so a pending task sleeps at
// once.
__device__ __forceinline__ void end_poll_in(const Graph &gr, uint32_t gid, bool valid, uint32_t idx, bool resolved,
                                            uint32_t pg, bool phased = false) {
  const Group &G = gr.g[gid];
  uint32_t p = kNone;
  if (valid) {
    if (resolved) {
      G.state[idx] = kFinished;
      p = G.parent[idx];
    } else {
      G.state[idx] = phased ? kSleep : kGoToSleep;
    }
  }
  const bool fin = valid && resolved;
  cleanup(G, fin && p == kNone, idx);
  if (pg < gr.n) {
    const bool push = fin && p != kNone && atomicExch(gr.g[pg].state + p, kWaken) != kGoToSleep;
    push_active(gr.g[pg], push, p);
  }
}

// Block-uniform loop over the entries a poll covers: body(i, valid, idx).
template <class F>
__device__ __forceinline__ void for_each_polled(const Group &G, F body) {
  const uint32_t snap = G.ctr[kSnapshot];
  const uint32_t *list = active_list(G, G.cur);
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < snap; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    const bool valid = i < snap;
    body(i, valid, valid ? list[i] : 0u);
  }
}

// ---- phases: rounds without list launches ----
// Some executions split their groups into phases such that, while a phase
// polls its groups, none of their tasks is spawned or woken and none of
// their empty stacks is popped (the trace group and the stage groups of a
// ray-tracing pipeline whose stages spawn and wake only trace tasks, and
// whose pools hold every task of a tile, so no freed slot is needed again
// before the tile ends).  Then a phase needs no prepare and no compaction:
// its active lists do not change while it runs, so every block reads the
// entries to poll from the active count; no polled task is awake when the
// phase ends, so a pending task sleeps at once (end_poll `phased`) and the
// compacted list is empty: the phase zeroes the other buffer's count, and
// the host flips the groups' buffers after the launch (end_phase).  Freed
// slots collect on the removed lists until the next prepare.

// Block-uniform loop over a phase's entries of group G: body(i, valid, idx).
template <class F>
__device__ __forceinline__ void for_each_active(const Group &G, F body) {
  const uint32_t n = min(*active_count(G), G.size);
  const uint32_t *list = active_list(G, G.cur);
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    const bool valid = i < n;
    body(i, valid, valid ? list[i] : 0u);
  }
}
// Called by every thread of a phase launch for each polled group: the
// buffer the group's next pushes go to starts empty.
__device__ __forceinline__ void begin_phase(const Group &G) {
  if (blockIdx.x == 0 && threadIdx.x == 0) G.ctr[kActive0 + (G.cur ^ 1u)] = 0u;
}
// Phased polls of executions in which every task awaits at most one child
// at a time: the parent of a resolving task is asleep in a group of the
// other phase and nothing else wakes it in this phase, so the wake is a
// store and a push (no exchange to learn the previous state).  The caller
// issues the push through block_add_n with its other counters and passes
// the position here.
__device__ __forceinline__ void wake_at(const Group &PG, bool pred, uint32_t pidx, uint32_t pos) {
  if (!pred) return;
  PG.state[pidx] = kWaken;
  if (pos < PG.size) active_list(PG, PG.cur)[pos] = pidx;
  else atomicOr(PG.ctr + kError, 2u);
}
// Frees slot idx of G (cleanup) at removed-list position pos from
// block_add_n on removed_count(G).
__device__ __forceinline__ void free_at(const Group &G, bool pred, uint32_t idx, uint32_t pos) {
  if (!pred) return;
  if (pos < G.size) G.removed[pos] = idx;
  else atomicOr(G.ctr + kError, 4u);
  G.state[idx] = kNotExist;
}

// Host: after a phase launch over the groups in mask.
inline void end_phase(Graph &gr, uint32_t mask) {
  for (uint32_t k = 0; k < gr.n; k++)
    if (mask >> k & 1u) gr.g[k].cur ^= 1u;
}
// Freed slots of G back on its empty stack during a launch in which no
// slot of G is popped or freed (a phase of the other groups, when G's tasks
// are spawned and freed only by those groups' polls): reads the counts of
// parity pp, writes those of parity 1 - pp; the host flips pp after the
// launch (flip_refilled).  Every thread of the launch calls it.
__device__ __forceinline__ void refill_empty(const Group &G) {
  const uint32_t ne = *empty_count(G), nr = *removed_count(G);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x)
    if (ne + i < G.size) G.empty[ne + i] = G.removed[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    G.ctr[kEmpty0 + (G.pp ^ 1u)] = ne + nr < G.size ? ne + nr : G.size;
    G.ctr[kRemoved0 + (G.pp ^ 1u)] = 0u;
  }
}
inline void flip_refilled(Graph &gr, uint32_t mask) {
  for (uint32_t k = 0; k < gr.n; k++)
    if (mask >> k & 1u) gr.g[k].pp ^= 1u;
}

// ---- host side ----

// Enqueue helpers (task_graph.hip).  Those that take a mutable Graph flip
// the parities of the groups they prepare or compact; later launches must
// be given the updated Graph.
hipError_t enqueue_reset(Graph &gr, hipStream_t st);
hipError_t enqueue_prepare(Graph &gr, uint32_t gid, hipStream_t st);
hipError_t enqueue_compact(Graph &gr, uint32_t gid, hipStream_t st);
// The compaction of group gc followed by the prepare of group gp (either
// kNone), in one launch when they are different groups.
hipError_t enqueue_list(Graph &gr, uint32_t gc, uint32_t gp, hipStream_t st);
// The compactions of the groups whose bits are set in cmask, then the
// prepares of those in pmask; one launch when no group is in both.
hipError_t enqueue_lists(Graph &gr, uint32_t cmask, uint32_t pmask, hipStream_t st);

// `rounds` rounds of execute (mod.rs:396-460): per group, prepare, then
// poll(gid) (which enqueues the group's poll launches), then compact; each
// compaction shares its launch with the next group's prepare.
template <class PollFn>
hipError_t enqueue_rounds(Graph &gr, uint32_t rounds, PollFn poll, hipStream_t st) {
  uint32_t pending = kNone;  // the group whose compaction is still due
  for (uint32_t r = 0; r < rounds; r++)
    for (uint32_t gid = 0; gid < gr.n; gid++) {
      hipError_t e = enqueue_list(gr, pending, gid, st);
      if (e == hipSuccess) e = poll(gid);
      if (e != hipSuccess) return e;
      pending = gid;
    }
  return enqueue_list(gr, pending, kNone, st);
}
// dispatch_allocate_init_task (mod.rs:223-303) for the generic task kinds:
// `count` root tasks whose payload word 0 is their spawn index.
hipError_t enqueue_spawn_roots(Graph &gr, uint32_t gid, uint32_t count, hipStream_t st);
// The same right after enqueue_reset (every slot empty): no prepare, no
// atomics, counts written by one thread.  count <= the group's slots.
hipError_t enqueue_spawn_fresh(Graph &gr, uint32_t gid, uint32_t count, hipStream_t st);

// enqueue_reset then enqueue_spawn_fresh(gid, count) in one launch, which
// also sets the caller's nzero words at `zero` (at most 1024) to zero.
hipError_t enqueue_reset_spawn(Graph &gr, uint32_t gid, uint32_t count, uint32_t *zero, uint32_t nzero,
                               hipStream_t st);

// Generic task kinds (the C-ABI task graph): kind 0 resolves on its first
// poll; kind 1 spawns one task of group `child_group` with payload
// (own payload + 1) and itself as parent, sleeps, and once that child has
// finished reads its payload back, frees it and resolves.
struct KindDesc {
  uint32_t kind, child_group;
};
hipError_t enqueue_generic_poll(const Graph &gr, uint32_t gid, const KindDesc *kinds, hipStream_t st);

// Device memory of a graph: one allocation carved into the groups' arrays.
struct GraphMemory {
  Graph dev{};
  void *base = nullptr;
  size_t bytes = 0;
  int alloc(const std::vector<uint32_t> &sizes, const std::vector<uint32_t> &words);
  void release();
};

}  // namespace tg
}  // namespace rdrt

#endif
