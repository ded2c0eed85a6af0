// task_graph.hip -- list kernels of the device task graph (reset, prepare,
// compact, root spawns) and the generic task kinds of the C-ABI task graph
// (shader/task-graph/src/runtime/task_group.rs:104-305, mod.rs:223-303).
#include "task_graph.h"

namespace rdrt {
namespace tg {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kMaxGrid = 1024;

uint32_t grid_for(uint32_t n) {
  const uint32_t g = (n + kBlock - 1) / kBlock;
  return g < 1u ? 1u : (g > kMaxGrid ? kMaxGrid : g);
}

// true in thread 0 of the block that finished last (after a fence), so it
// may commit counters every block has read.
__device__ __forceinline__ bool last_block(uint32_t *done) {
  __shared__ bool last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(done, 1u) == gridDim.x - 1u;
  }
  __syncthreads();
  return last && threadIdx.x == 0;
}

// Every group empty: slots 1..size on the empty stack (popped in ascending
// order), all states free, counters zero.
__global__ void k_reset(Graph gr) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = 0; k < gr.n; k++) {
    const Group &G = gr.g[k];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= G.size; i += stride) {
      G.state[i] = kNotExist;
      if (i < G.size) G.empty[i] = G.size - i;
    }
    if (blockIdx.x == 0 && threadIdx.x < kCtrWords) G.ctr[threadIdx.x] = threadIdx.x == kEmpty ? G.size : 0u;
  }
}

// prepare_execution (task_group.rs:213-233): slots freed since the group
// last ran return to the empty stack; the entries to poll are fixed.
__global__ void k_prepare(Graph gr, uint32_t gid) {
  const Group &G = gr.g[gid];
  const uint32_t ne = G.ctr[kEmpty], nr = G.ctr[kRemoved];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x)
    if (ne + i < G.size) G.empty[ne + i] = G.removed[i];
  if (last_block(G.ctr + kDone)) {
    G.ctr[kEmpty] = ne + nr < G.size ? ne + nr : G.size;
    G.ctr[kRemoved] = 0u;
    G.ctr[kSnapshot] = G.ctr[kActive] < G.size ? G.ctr[kActive] : G.size;
    G.ctr[kDone] = 0u;
  }
}

// use_compact_alive_tasks (task_group.rs:168-210, dispatch_compact.rs:60-98):
// awake tasks move to the other buffer, tasks going to sleep leave it.
__global__ void k_compact(Graph gr, uint32_t gid) {
  const Group &G = gr.g[gid];
  const uint32_t cur = G.ctr[kCur];
  const uint32_t n = G.ctr[kActive] < G.size ? G.ctr[kActive] : G.size;
  const uint32_t *src = active_list(G, cur);
  uint32_t *dst = active_list(G, cur ^ 1u);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t idx = src[i];
    const uint32_t s = G.state[idx];
    if (s == kGoToSleep) G.state[idx] = kSleep;
    if (s == kWaken) dst[atomicAdd(G.ctr + kNext, 1u)] = idx;
  }
  if (last_block(G.ctr + kDone)) {
    G.ctr[kActive] = G.ctr[kNext];
    G.ctr[kNext] = 0u;
    G.ctr[kCur] = cur ^ 1u;
    G.ctr[kDone] = 0u;
  }
}

__global__ void k_spawn_roots(Graph gr, uint32_t gid, uint32_t count) {
  const Group &G = gr.g[gid];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const uint32_t idx = spawn(G, kNone, kNone);
    if (idx != kNone && G.words) G.payload[(size_t)idx * G.words] = i;
  }
}

// The generic kinds: 0 = BaseShaderFuture (ready at once); 1 = then(spawn a
// child of `child_group`, await it) (future/then.rs semantics).
__global__ void k_generic_poll(Graph gr, uint32_t gid, uint32_t kind, uint32_t child_group) {
  const Group &G = gr.g[gid];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;; i += stride) {
    uint32_t idx;
    if (!poll_entry(G, i, idx)) break;
    if (kind == 0u) {
      end_poll(gr, gid, idx, true);
      continue;
    }
    const Group &C = gr.g[child_group];
    if (G.stage[idx] == 0u) {
      const uint32_t c = spawn(C, idx, gid);
      if (c != kNone && C.words) C.payload[(size_t)c * C.words] = G.payload[(size_t)idx * G.words] + 1u;
      G.child[idx] = c;
      G.stage[idx] = 1u;
      end_poll(gr, gid, idx, false);  // the child was just spawned: not finished
      continue;
    }
    const uint32_t c = G.child[idx];
    if (c != kNone && C.state[c] == kFinished) {
      if (C.words) G.payload[(size_t)idx * G.words] = C.payload[(size_t)c * C.words];
      cleanup(C, c);
      end_poll(gr, gid, idx, true);
    } else {
      end_poll(gr, gid, idx, false);
    }
  }
}

uint32_t max_size(const Graph &gr) {
  uint32_t m = 1;
  for (uint32_t k = 0; k < gr.n; k++) m = gr.g[k].size + 1 > m ? gr.g[k].size + 1 : m;
  return m;
}

}  // namespace

hipError_t enqueue_reset(const Graph &gr, hipStream_t st) {
  hipLaunchKernelGGL(k_reset, dim3(grid_for(max_size(gr))), dim3(kBlock), 0, st, gr);
  return hipGetLastError();
}

hipError_t enqueue_prepare(const Graph &gr, uint32_t gid, hipStream_t st) {
  hipLaunchKernelGGL(k_prepare, dim3(grid_for(gr.g[gid].size)), dim3(kBlock), 0, st, gr, gid);
  return hipGetLastError();
}

hipError_t enqueue_compact(const Graph &gr, uint32_t gid, hipStream_t st) {
  hipLaunchKernelGGL(k_compact, dim3(grid_for(gr.g[gid].size)), dim3(kBlock), 0, st, gr, gid);
  return hipGetLastError();
}

hipError_t enqueue_spawn_roots(const Graph &gr, uint32_t gid, uint32_t count, hipStream_t st) {
  hipError_t e = enqueue_prepare(gr, gid, st);
  if (e != hipSuccess || count == 0) return e;
  hipLaunchKernelGGL(k_spawn_roots, dim3(grid_for(count)), dim3(kBlock), 0, st, gr, gid, count);
  return hipGetLastError();
}

hipError_t enqueue_generic_poll(const Graph &gr, uint32_t gid, const KindDesc *kinds, hipStream_t st) {
  hipLaunchKernelGGL(k_generic_poll, dim3(grid_for(gr.g[gid].size)), dim3(kBlock), 0, st, gr, gid, kinds[gid].kind,
                     kinds[gid].child_group);
  return hipGetLastError();
}

int GraphMemory::alloc(const std::vector<uint32_t> &sizes, const std::vector<uint32_t> &words) {
  release();
  if (sizes.size() > kMaxGroups || sizes.size() != words.size()) return -1;
  auto up = [](size_t w) { return (w + 63) & ~(size_t)63; };  // 256-byte aligned arrays
  size_t total = 0;
  for (size_t k = 0; k < sizes.size(); k++) {
    const size_t s1 = (size_t)sizes[k] + 1;
    total += 5 * up(s1) + up(s1 * words[k]) + 4 * up(sizes[k]) + up(kCtrWords);
  }
  if (hipMalloc(&base, total * 4) != hipSuccess) {
    base = nullptr;
    return -2;
  }
  bytes = total * 4;
  uint32_t *p = static_cast<uint32_t *>(base);
  auto take = [&](size_t w) {
    uint32_t *r = p;
    p += up(w);
    return r;
  };
  dev.n = (uint32_t)sizes.size();
  for (size_t k = 0; k < sizes.size(); k++) {
    Group &G = dev.g[k];
    const size_t s1 = (size_t)sizes[k] + 1;
    G.size = sizes[k];
    G.words = words[k];
    G.state = take(s1);
    G.parent = take(s1);
    G.ptype = take(s1);
    G.stage = take(s1);
    G.child = take(s1);
    G.payload = take(s1 * words[k]);
    G.list0 = take(sizes[k]);
    G.list1 = take(sizes[k]);
    G.empty = take(sizes[k]);
    G.removed = take(sizes[k]);
    G.ctr = take(kCtrWords);
  }
  return 0;
}

void GraphMemory::release() {
  if (base) (void)hipFree(base);
  base = nullptr;
  bytes = 0;
  dev = Graph{};
}

}  // namespace tg
}  // namespace rdrt
