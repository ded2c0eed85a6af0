// compact_device.h -- single-pass stable stream compaction for gfx950:
// wave ballot ranks, block sums in LDS, wave-parallel decoupled look-back.
// The output equals a stable copy_if (shader/parallel-compute/src/
// stream_compaction.rs:3-45).  Included by HIP sources only.
#ifndef RDRT_COMPACT_DEVICE_H
#define RDRT_COMPACT_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rdrt {
namespace compact {

using u64 = unsigned long long;
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint32_t lane_rank(u64 m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

struct Pair {
  uint32_t a, b;
};

__device__ __forceinline__ u64 pack(u64 flag, Pair v) { return (flag << 62) | ((u64)v.a << 31) | (u64)v.b; }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Decoupled look-back for block `bid`, run by the 64 lanes of one wave:
// publishes the block's aggregate, then inspects 64 predecessors per step
// (lane l reads block base - l), sums the aggregates up to the nearest
// inclusive prefix, publishes its own inclusive prefix and returns the
// exclusive one to every lane.  Status word: bits 63-62 flag (0 empty,
// 1 aggregate, 2 inclusive), 61-31 count a, 30-0 count b.
__device__ Pair lookback_wave(u64 *status, uint32_t bid, Pair total, uint32_t *err) {
  const uint32_t lane = threadIdx.x & 63u;
  Pair excl{0u, 0u};
  if (bid == 0) {
    if (lane == 0) __hip_atomic_store(status, pack(2, total), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return excl;
  }
  if (lane == 0) __hip_atomic_store(status + bid, pack(1, total), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  int32_t base = (int32_t)bid - 1;
  uint32_t spins = 0;
  for (;;) {
    const int32_t j = base - (int32_t)lane;
    // before block 0 reads as an inclusive zero
    const u64 s = j >= 0 ? __hip_atomic_load(status + j, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : pack(2, Pair{0, 0});
    const uint32_t flag = (uint32_t)(s >> 62);
    const u64 incl = __ballot(flag == 2);
    const u64 empty = __ballot(flag == 0);
    const u64 low = incl & (~incl + 1ull);  // nearest inclusive predecessor
    const u64 upto = incl ? (low | (low - 1ull)) : ~0ull;
    if (empty & upto) {
      // bounded wait: every predecessor already holds its ticket and publishes
      // its aggregate without waiting; the limit only guards against a bug
      if (++spins > kSpinLimit) {
        if (lane == 0) atomicOr(err, 2u);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    const bool mine = (upto >> lane) & 1ull;
    excl.a += wave_sum(mine ? (uint32_t)((s >> 31) & 0x7fffffffu) : 0u);
    excl.b += wave_sum(mine ? (uint32_t)(s & 0x7fffffffu) : 0u);
    if (incl) break;
    base -= 64;
  }
  if (lane == 0)
    __hip_atomic_store(status + bid, pack(2, Pair{excl.a + total.a, excl.b + total.b}), __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// Block-wide stable compaction of two flag sets: writes `item` to
// out_a[offset] / out_b[offset] in thread order across the whole grid.
// Must be called by every thread of the block.  Returns true on the block
// that holds the last item (nblocks - 1) with the grid totals.
template <int B>
__device__ __forceinline__ void compact2(bool ka, bool kb, uint32_t item, uint32_t *out_a, uint32_t *out_b,
                                         u64 *status, uint32_t bid, uint32_t nblocks, uint32_t *total_a,
                                         uint32_t *total_b, uint32_t *err, uint32_t capacity) {
  __shared__ uint32_t s_wa[B / 64], s_wb[B / 64];
  __shared__ uint32_t s_ea, s_eb;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const u64 ma = __ballot(ka), mb = __ballot(kb);
  const uint32_t ra = lane_rank(ma), rb = lane_rank(mb);
  if (lane == 0) {
    s_wa[wave] = (uint32_t)__popcll(ma);
    s_wb[wave] = (uint32_t)__popcll(mb);
  }
  __syncthreads();
  // every thread derives its wave's offset and the block totals
  uint32_t pa = 0, pb = 0, ta = 0, tb = 0;
  for (uint32_t w = 0; w < B / 64; w++) {
    if (w == wave) {
      pa = ta;
      pb = tb;
    }
    ta += s_wa[w];
    tb += s_wb[w];
  }
  if (wave == 0) {
    Pair e = lookback_wave(status, bid, Pair{ta, tb}, err);
    if (lane == 0) {
      s_ea = e.a;
      s_eb = e.b;
      if (bid == nblocks - 1) {
        if (total_a) *total_a = e.a + ta;
        if (total_b) *total_b = e.b + tb;
      }
    }
  }
  __syncthreads();
  // the bound only matters if a look-back gave up (error flag set)
  const uint32_t ia = s_ea + pa + ra, ib = s_eb + pb + rb;
  if (ka && ia < capacity) out_a[ia] = item;
  if (kb && ib < capacity) out_b[ib] = item;
}

}  // namespace compact
}  // namespace rdrt

#endif
