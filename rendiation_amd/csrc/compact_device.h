// compact_device.h -- stable stream compaction for gfx950: wave ballot ranks
// (mbcnt), block sums in LDS, then a scan of the block counts and a scatter.
// No block ever waits on another (a single-pass decoupled look-back chains
// every block to its predecessors and serialized large queues).  The output
// equals a stable copy_if (shader/parallel-compute/src/stream_compaction.rs:
// 3-45).  Included by HIP sources only.
#ifndef RDRT_COMPACT_DEVICE_H
#define RDRT_COMPACT_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rdrt {
namespace compact {

using u64 = unsigned long long;

__device__ __forceinline__ uint32_t lane_rank(u64 m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// ---- three-pass compaction (no block waits on another) ----
// 1. block_compact2 (inside the producing kernel): each block compacts its
//    own items stably into tmp[bid * B + local rank] and writes its counts.
// 2. k_block_scan: one workgroup turns the block counts into exclusive
//    offsets and the totals.
// 3. k_scatter2: each block copies its compacted run to out[offset + j].
// Block i holds the input slots [i * B, (i + 1) * B), so the output is a
// stable copy_if whatever order the blocks run in.

// Stable block-local compaction of two flag sets.  Every thread of the
// block calls it.
template <int B>
__device__ __forceinline__ void block_compact2(bool ka, bool kb, uint32_t item, uint32_t bid, uint32_t *tmp_a,
                                               uint32_t *tmp_b, uint32_t *cnt_a, uint32_t *cnt_b) {
  __shared__ uint32_t s_wa[B / 64], s_wb[B / 64];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const u64 ma = __ballot(ka), mb = __ballot(kb);
  if (lane == 0) {
    s_wa[wave] = (uint32_t)__popcll(ma);
    s_wb[wave] = (uint32_t)__popcll(mb);
  }
  __syncthreads();
  uint32_t pa = 0, pb = 0, ta = 0, tb = 0;
  for (uint32_t w = 0; w < B / 64; w++) {
    if (w == wave) {
      pa = ta;
      pb = tb;
    }
    ta += s_wa[w];
    tb += s_wb[w];
  }
  const size_t base = (size_t)bid * B;
  if (ka) tmp_a[base + pa + lane_rank(ma)] = item;
  if (kb && tmp_b) tmp_b[base + pb + lane_rank(mb)] = item;
  if (tid == 0) {
    cnt_a[bid] = ta;
    if (cnt_b) cnt_b[bid] = tb;
  }
}

constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kScanPer = 8;                         // counts per thread per tile
constexpr uint32_t kScanTile = kScanThreads * kScanPer;  // counts per tile

namespace {

// Exclusive scan of the block counts of the first ceil(min(*count_in,
// slots) / B) blocks (one workgroup of kScanThreads): offsets and totals.
// cnt_b / off_b / total_b may be null.  The counts stream through LDS in
// tiles of kScanTile with coalesced loads and stores; each thread scans
// kScanPer consecutive counts of a tile, the workgroup scans the thread sums,
// and a running carry links the tiles.
__global__ __launch_bounds__(kScanThreads) void k_block_scan(const uint32_t *count_in, uint32_t slots, uint32_t B,
                                                             const uint32_t *cnt_a, const uint32_t *cnt_b,
                                                             uint32_t *off_a, uint32_t *off_b, uint32_t *total_a,
                                                             uint32_t *total_b) {
  __shared__ uint32_t t_a[kScanTile], t_b[kScanTile];
  __shared__ uint32_t s_a[kScanThreads / 64], s_b[kScanThreads / 64];
  const uint32_t n = count_in ? min(*count_in, slots) : slots;
  const uint32_t nb = (n + B - 1) / B;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const bool two = cnt_b != nullptr;
  uint32_t carry_a = 0, carry_b = 0;
  for (uint32_t base = 0; base < nb; base += kScanTile) {
    for (uint32_t k = 0; k < kScanPer; k++) {
      const uint32_t i = base + k * kScanThreads + tid;
      t_a[k * kScanThreads + tid] = i < nb ? cnt_a[i] : 0u;
      if (two) t_b[k * kScanThreads + tid] = i < nb ? cnt_b[i] : 0u;
    }
    __syncthreads();
    uint32_t sa = 0, sb = 0;
    for (uint32_t k = 0; k < kScanPer; k++) {
      sa += t_a[tid * kScanPer + k];
      if (two) sb += t_b[tid * kScanPer + k];
    }
    // inclusive wave scans, then wave totals through LDS
    uint32_t ia = sa, ib = sb;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t xa = __shfl_up(ia, o, 64), xb = __shfl_up(ib, o, 64);
      if (lane >= (uint32_t)o) {
        ia += xa;
        ib += xb;
      }
    }
    if (lane == 63) {
      s_a[wave] = ia;
      s_b[wave] = ib;
    }
    __syncthreads();
    uint32_t wa = 0, wb = 0, ta = 0, tb = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; w++) {
      if (w == wave) {
        wa = ta;
        wb = tb;
      }
      ta += s_a[w];
      tb += s_b[w];
    }
    // exclusive prefixes of this thread's counts, written back in place
    uint32_t ea = carry_a + wa + ia - sa, eb = carry_b + wb + ib - sb;
    for (uint32_t k = 0; k < kScanPer; k++) {
      const uint32_t va = t_a[tid * kScanPer + k];
      t_a[tid * kScanPer + k] = ea;
      ea += va;
      if (two) {
        const uint32_t vb = t_b[tid * kScanPer + k];
        t_b[tid * kScanPer + k] = eb;
        eb += vb;
      }
    }
    __syncthreads();
    for (uint32_t k = 0; k < kScanPer; k++) {
      const uint32_t i = base + k * kScanThreads + tid;
      if (i < nb) {
        off_a[i] = t_a[k * kScanThreads + tid];
        if (two) off_b[i] = t_b[k * kScanThreads + tid];
      }
    }
    carry_a += ta;
    carry_b += tb;
    __syncthreads();  // the tile's LDS is reused by the next tile
  }
  if (tid == 0) {
    if (total_a) *total_a = carry_a;
    if (total_b) *total_b = carry_b;
  }
}

// Copies each block's compacted run to its offset (grid: one block of B
// threads per possible input block; blocks past the live count return).
template <int B>
__global__ __launch_bounds__(B) void k_scatter2(const uint32_t *count_in, uint32_t slots, const uint32_t *tmp_a,
                                                const uint32_t *tmp_b, const uint32_t *cnt_a, const uint32_t *cnt_b,
                                                const uint32_t *off_a, const uint32_t *off_b, uint32_t *out_a,
                                                uint32_t *out_b, uint32_t capacity) {
  const uint32_t n = count_in ? min(*count_in, slots) : slots;
  const uint32_t nb = (n + B - 1) / B, bid = blockIdx.x, j = threadIdx.x;
  if (bid >= nb) return;
  const size_t base = (size_t)bid * B;
  if (j < cnt_a[bid] && off_a[bid] + j < capacity) out_a[off_a[bid] + j] = tmp_a[base + j];
  if (cnt_b && j < cnt_b[bid] && off_b[bid] + j < capacity) out_b[off_b[bid] + j] = tmp_b[base + j];
}

}  // namespace

}  // namespace compact
}  // namespace rdrt

#endif
