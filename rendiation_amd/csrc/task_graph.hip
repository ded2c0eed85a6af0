// task_graph.hip -- list kernels of the device task graph (reset, prepare,
// compact, root spawns) and the generic task kinds of the C-ABI task graph
// (shader/task-graph/src/runtime/task_group.rs:104-305, mod.rs:223-303).
#include "task_graph.h"

namespace rdrt {
namespace tg {
namespace {

// 1024-thread blocks: a pass takes one counter atomic per block and list,
// and same-address atomics serialize
constexpr uint32_t kBlock = 1024;
constexpr uint32_t kMaxGrid = 1024;
constexpr uint32_t kListGrid = 1024;

uint32_t grid_for(uint32_t n, uint32_t cap = kMaxGrid) {
  const uint32_t g = (n + kBlock - 1) / kBlock;
  return g < 1u ? 1u : (g > cap ? cap : g);
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
    if (blockIdx.x == 0 && threadIdx.x < kCtrWords) G.ctr[threadIdx.x] = threadIdx.x == kEmpty0 ? G.size : 0u;
  }
}

// prepare_execution (task_group.rs:213-233): slots freed since the group
// last ran return to the empty stack; the entries to poll are fixed.  Reads
// the counts of parity pp, writes those of parity 1 - pp (one writer).
__device__ __forceinline__ void prepare_part(const Group &G) {
  const uint32_t ne = *empty_count(G), nr = *removed_count(G);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x)
    if (ne + i < G.size) G.empty[ne + i] = G.removed[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const uint32_t na = *active_count(G);
    G.ctr[kEmpty0 + (G.pp ^ 1u)] = ne + nr < G.size ? ne + nr : G.size;
    G.ctr[kRemoved0 + (G.pp ^ 1u)] = 0u;
    G.ctr[kSnapshot] = na < G.size ? na : G.size;
    G.ctr[kActive0 + (G.cur ^ 1u)] = 0u;  // the compaction's target
  }
}

// use_compact_alive_tasks (task_group.rs:168-210, dispatch_compact.rs:60-98):
// awake tasks move to the other buffer, tasks going to sleep leave it.
__device__ __forceinline__ void compact_part(const Group &G) {
  const uint32_t n = *active_count(G) < G.size ? *active_count(G) : G.size;
  const uint32_t *src = active_list(G, G.cur);
  uint32_t *dst = active_list(G, G.cur ^ 1u);
  uint32_t *dst_count = G.ctr + kActive0 + (G.cur ^ 1u);
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    const bool valid = i < n;
    const uint32_t idx = valid ? src[i] : 0u;
    const uint32_t s = valid ? G.state[idx] : kNotExist;
    if (s == kGoToSleep) G.state[idx] = kSleep;
    const bool keep = s == kWaken;
    const uint32_t pos = block_add(dst_count, keep);
    if (keep) dst[pos] = idx;
  }
}

// The compactions of the groups in cmask and the prepares of the groups in
// pmask in one launch (disjoint masks): nothing of one group's pass touches
// another group.
__global__ void k_list(Graph gr, uint32_t cmask, uint32_t pmask) {
  for (uint32_t k = 0; k < gr.n; k++)
    if (cmask >> k & 1u) compact_part(gr.g[k]);
  for (uint32_t k = 0; k < gr.n; k++)
    if (pmask >> k & 1u) prepare_part(gr.g[k]);
}

// dispatch_allocate_init_task (mod.rs:223-303) right after a prepare:
// roots take the top `count` empty slots in order and the next active
// entries, without atomics.
__global__ void k_spawn_roots(Graph gr, uint32_t gid, uint32_t count) {
  const Group &G = gr.g[gid];
  const uint32_t ne = *empty_count(G), na = *active_count(G) < G.size ? *active_count(G) : G.size;
  const uint32_t room = G.size - na;
  uint32_t k = count < ne ? count : ne;
  k = k < room ? k : room;
  uint32_t *list = active_list(G, G.cur);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
    const uint32_t idx = G.empty[ne - 1 - i];
    G.state[idx] = kWaken;
    G.parent[idx] = kNone;
    G.ptype[idx] = kNone;
    G.stage[idx] = 0u;
    G.child[idx] = kNone;
    if (G.words) G.payload[(size_t)idx * G.words] = i;
    list[na + i] = idx;
  }
  if (last_block(G.ctr + kDone)) {
    *empty_count(G) = ne - k;
    *active_count(G) = na + k;
    if (count > ne) atomicOr(G.ctr + kError, 1u);
    else if (count > room) atomicOr(G.ctr + kError, 2u);
    G.ctr[kDone] = 0u;
  }
}

// Root spawns right after a reset (every slot empty, nothing active): root
// i takes slot i + 1 and active entry i, and the counts are known on the
// host, so one thread writes them.
__global__ void k_spawn_fresh(Graph gr, uint32_t gid, uint32_t k) {
  const Group &G = gr.g[gid];
  uint32_t *list = active_list(G, G.cur);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
    const uint32_t idx = i + 1u;  // empty[size - 1 - i] after a reset
    G.state[idx] = kWaken;
    G.parent[idx] = kNone;
    G.ptype[idx] = kNone;
    G.stage[idx] = 0u;
    G.child[idx] = kNone;
    if (G.words) G.payload[(size_t)idx * G.words] = i;
    list[i] = idx;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *empty_count(G) = G.size - k;
    *active_count(G) = k;
  }
}

// k_reset followed by k_spawn_fresh(gid, k) in one pass (every write of
// the two depends on its slot only), plus nzero caller words set to zero.
__global__ void k_reset_spawn(Graph gr, uint32_t gid, uint32_t k, uint32_t *zero, uint32_t nzero) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nzero) zero[t] = 0u;
  for (uint32_t g = 0; g < gr.n; g++) {
    const Group &G = gr.g[g];
    const uint32_t kk = g == gid ? k : 0u;
    for (uint32_t i = t; i <= G.size; i += stride) {
      if (i >= 1u && i <= kk) {  // root i - 1 takes slot i and active entry i - 1
        G.state[i] = kWaken;
        G.parent[i] = kNone;
        G.ptype[i] = kNone;
        G.stage[i] = 0u;
        G.child[i] = kNone;
        if (G.words) G.payload[(size_t)i * G.words] = i - 1u;
        G.list0[i - 1u] = i;
      } else {
        G.state[i] = kNotExist;
      }
      if (i < G.size) G.empty[i] = G.size - i;
    }
    if (blockIdx.x == 0 && threadIdx.x < kCtrWords)
      G.ctr[threadIdx.x] = threadIdx.x == kEmpty0 ? G.size - kk : threadIdx.x == kActive0 ? kk : 0u;
  }
}

// The generic kinds: 0 = BaseShaderFuture (ready at once); 1 = then(spawn a
// child of `child_group`, await it) (future/then.rs semantics).
__global__ void k_generic_poll(Graph gr, uint32_t gid, uint32_t kind, uint32_t child_group) {
  const Group &G = gr.g[gid];
  const Group &C = gr.g[child_group];
  for_each_polled(G, [&](uint32_t, bool valid, uint32_t idx) {
    if (kind == 0u) {
      end_poll(gr, gid, valid, idx, true);
      return;
    }
    const bool first = valid && G.stage[idx] == 0u;
    // first poll: spawn the child with payload + 1; it has not finished yet
    const uint32_t c = spawn(C, first, idx, gid);
    if (first) {
      if (c != kNone && C.words) C.payload[(size_t)c * C.words] = G.payload[(size_t)idx * G.words] + 1u;
      G.child[idx] = c;
      G.stage[idx] = 1u;
    }
    // later polls: a finished child is read back and freed
    const uint32_t w = (valid && !first) ? G.child[idx] : kNone;
    const bool done = w != kNone && C.state[w] == kFinished;
    if (done && C.words) G.payload[(size_t)idx * G.words] = C.payload[(size_t)w * C.words];
    cleanup(C, done, w);
    end_poll(gr, gid, valid, idx, done);
  });
}

uint32_t max_size(const Graph &gr) {
  uint32_t m = 1;
  for (uint32_t k = 0; k < gr.n; k++) m = gr.g[k].size + 1 > m ? gr.g[k].size + 1 : m;
  return m;
}

}  // namespace

hipError_t enqueue_reset(Graph &gr, hipStream_t st) {
  for (uint32_t k = 0; k < gr.n; k++) gr.g[k].cur = gr.g[k].pp = 0u;
  hipLaunchKernelGGL(k_reset, dim3(grid_for(max_size(gr))), dim3(kBlock), 0, st, gr);
  return hipGetLastError();
}

hipError_t enqueue_lists(Graph &gr, uint32_t cmask, uint32_t pmask, hipStream_t st) {
  const uint32_t all = gr.n >= 32 ? ~0u : (1u << gr.n) - 1u;
  cmask &= all;
  pmask &= all;
  if (!cmask && !pmask) return hipSuccess;
  if (cmask & pmask) {  // a group in both: its compaction first, then its prepare
    hipError_t e = enqueue_lists(gr, cmask, 0u, st);
    return e != hipSuccess ? e : enqueue_lists(gr, 0u, pmask, st);
  }
  uint32_t n = 0;
  for (uint32_t k = 0; k < gr.n; k++)
    if (((cmask | pmask) >> k & 1u) && gr.g[k].size > n) n = gr.g[k].size;
  hipLaunchKernelGGL(k_list, dim3(grid_for(n, kListGrid)), dim3(kBlock), 0, st, gr, cmask, pmask);
  for (uint32_t k = 0; k < gr.n; k++) {
    if (cmask >> k & 1u) gr.g[k].cur ^= 1u;
    if (pmask >> k & 1u) gr.g[k].pp ^= 1u;
  }
  return hipGetLastError();
}

hipError_t enqueue_list(Graph &gr, uint32_t gc, uint32_t gp, hipStream_t st) {
  return enqueue_lists(gr, gc < gr.n ? 1u << gc : 0u, gp < gr.n ? 1u << gp : 0u, st);
}

hipError_t enqueue_prepare(Graph &gr, uint32_t gid, hipStream_t st) { return enqueue_list(gr, kNone, gid, st); }

hipError_t enqueue_compact(Graph &gr, uint32_t gid, hipStream_t st) { return enqueue_list(gr, gid, kNone, st); }

hipError_t enqueue_spawn_roots(Graph &gr, uint32_t gid, uint32_t count, hipStream_t st) {
  hipError_t e = enqueue_prepare(gr, gid, st);
  if (e != hipSuccess || count == 0) return e;
  hipLaunchKernelGGL(k_spawn_roots, dim3(grid_for(count, kListGrid)), dim3(kBlock), 0, st, gr, gid, count);
  return hipGetLastError();
}

hipError_t enqueue_spawn_fresh(Graph &gr, uint32_t gid, uint32_t count, hipStream_t st) {
  if (count > gr.g[gid].size) return hipErrorInvalidValue;
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_spawn_fresh, dim3(grid_for(count, kListGrid)), dim3(kBlock), 0, st, gr, gid, count);
  return hipGetLastError();
}

hipError_t enqueue_reset_spawn(Graph &gr, uint32_t gid, uint32_t count, uint32_t *zero, uint32_t nzero,
                               hipStream_t st) {
  if (gid >= gr.n || count > gr.g[gid].size || nzero > kBlock) return hipErrorInvalidValue;
  for (uint32_t k = 0; k < gr.n; k++) gr.g[k].cur = gr.g[k].pp = 0u;
  hipLaunchKernelGGL(k_reset_spawn, dim3(grid_for(max_size(gr))), dim3(kBlock), 0, st, gr, gid, count, zero, nzero);
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
