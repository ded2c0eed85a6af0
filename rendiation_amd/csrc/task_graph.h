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

// Counter words of a group.
enum Ctr : uint32_t {
  kActive = 0,    // entries in the current active list
  kEmpty = 1,     // entries in the empty stack
  kRemoved = 2,   // slots freed since the group last ran
  kSnapshot = 3,  // active entries the current poll covers
  kCur = 4,       // which active buffer is current (0 / 1)
  kDone = 5,      // blocks finished (last-block commit of prepare / compact)
  kNext = 6,      // entries written to the other active buffer by compaction
  kError = 7,     // bit 0: empty stack exhausted; bit 1: active list full; bit 2: removed list full
  kCtrWords = 8,
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
};

struct Graph {
  Group g[kMaxGroups];
  uint32_t n;
};

__device__ __forceinline__ uint32_t *active_list(const Group &G, uint32_t cur) { return cur ? G.list1 : G.list0; }

__device__ __forceinline__ void push_active(const Group &G, uint32_t idx) {
  const uint32_t cur = __hip_atomic_load(G.ctr + kCur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t pos = atomicAdd(G.ctr + kActive, 1u);
  if (pos < G.size) active_list(G, cur)[pos] = idx;
  else atomicOr(G.ctr + kError, 2u);
}

// spawn_new_task (task_group.rs:428-446, task_pool.rs:143-172): take an
// empty slot, make it a woken task with the given parent and zeroed poll
// state, and append it to the active list.  Returns the slot, or kNone when
// the pool is full (error bit 0).  The caller writes the payload.
__device__ __forceinline__ uint32_t spawn(const Group &G, uint32_t parent, uint32_t ptype) {
  const uint32_t old = atomicSub(G.ctr + kEmpty, 1u);
  if (old == 0u || old > G.size) {
    atomicAdd(G.ctr + kEmpty, 1u);
    atomicOr(G.ctr + kError, 1u);
    return kNone;
  }
  const uint32_t idx = G.empty[old - 1u];
  G.state[idx] = kWaken;
  G.parent[idx] = parent;
  G.ptype[idx] = ptype;
  G.stage[idx] = 0u;
  G.child[idx] = kNone;
  push_active(G, idx);
  return idx;
}

// wake_task_dyn (task_group.rs:448-462): a task still in the active list
// (going to sleep in this very round) is only marked awake.
__device__ __forceinline__ void wake(const Group &G, uint32_t idx) {
  const uint32_t was = atomicExch(G.state + idx, kWaken);
  if (was != kGoToSleep) push_active(G, idx);
}

// cleanup_finished_task_state_and_payload (task_group.rs:490-498).
__device__ __forceinline__ void cleanup(const Group &G, uint32_t idx) {
  const uint32_t pos = atomicAdd(G.ctr + kRemoved, 1u);
  if (pos < G.size) G.removed[pos] = idx;
  else atomicOr(G.ctr + kError, 4u);
  G.state[idx] = kNotExist;
}

// The end of a poll (task_group.rs:118-150): a resolved task finishes and
// wakes its parent (or frees itself when it has none); a pending one goes
// to sleep.
__device__ __forceinline__ void end_poll(const Graph &gr, uint32_t gid, uint32_t idx, bool resolved) {
  const Group &G = gr.g[gid];
  if (!resolved) {
    G.state[idx] = kGoToSleep;
    return;
  }
  G.state[idx] = kFinished;
  const uint32_t p = G.parent[idx], pt = G.ptype[idx];
  if (p == kNone) cleanup(G, idx);
  else if (pt < gr.n) wake(gr.g[pt], p);
}

// The poll of one active entry: thread i < snapshot polls list[i].
__device__ __forceinline__ bool poll_entry(const Group &G, uint32_t i, uint32_t &idx) {
  const uint32_t snap = G.ctr[kSnapshot];
  if (i >= snap) return false;
  idx = active_list(G, G.ctr[kCur])[i];
  return true;
}

// ---- host side ----

// Enqueue helpers (task_graph.hip).  `grid_cap` bounds the blocks of the
// list kernels.
hipError_t enqueue_reset(const Graph &gr, hipStream_t st);
hipError_t enqueue_prepare(const Graph &gr, uint32_t gid, hipStream_t st);
hipError_t enqueue_compact(const Graph &gr, uint32_t gid, hipStream_t st);
// dispatch_allocate_init_task (mod.rs:223-303) for the generic task kinds:
// `count` root tasks whose payload word 0 is their spawn index.
hipError_t enqueue_spawn_roots(const Graph &gr, uint32_t gid, uint32_t count, hipStream_t st);

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
