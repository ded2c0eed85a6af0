// pt_tail.h -- the path tracer's tail kernel: from bounce b0 on, one
// persistent launch finishes every remaining bounce of each path instead of
// one set of wavefront launches per bounce.
//
// The tail bounces of a frame carry a few percent of its rays but, traced
// one launch per bounce, each launch lasts as long as its slowest ray: the
// frame pays the sum over bounces of the slowest ray.  Here a group of
// kCoopLanes lanes takes one path (a work item pulled from a counter) and
// runs its bounces back to back -- the shadow ray of the current bounce,
// then the extension ray, then the closest-hit / miss stage -- so the frame
// pays the slowest path instead.  Each path does exactly what the wavefront
// kernels do for it, in the same order (shadow light of bounce b is added
// before the emission of bounce b + 1), with the same per-path functions:
// the radiance is the same; only the visit order of exact-t ties can
// differ (the cooperative traversal's tie rule).
//
// Per group: every lane walks the ray together (cooperative traversal);
// the lane that wrote the hit record shades the path; lane 0 adds the
// unoccluded light.  Between phases, block-scope fences order the lanes'
// stores and loads of the path's records (all lanes of a group are in one
// wave of one work-group).
#ifndef RDRT_PT_TAIL_H
#define RDRT_PT_TAIL_H

#include "pt_shade.h"
#include "trace_device.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace {

constexpr uint32_t kTailBlock = 128;

template <int K, int S, int LAYOUT>
__global__ __launch_bounds__(kTailBlock) void k_pt_tail(PtLaunch P, TraceParams pe, TraceParams ps) {
  constexpr int B = (int)kTailBlock;
  __shared__ uint32_t lds[K * B];
  const uint32_t b0 = P.tail_b0;
  const uint32_t glane = threadIdx.x & (kCoopLanes - 1u);
  const uint32_t leader = __lane_id() & ~(kCoopLanes - 1u);
  const uint32_t nitems = min(P.counts[b0], P.npaths);
  const uint32_t *qin = (b0 & 1u) ? P.qb : P.qa;  // the queue shade(b0) read
  uint32_t *stk = lds + threadIdx.x;
  uint32_t *next = P.work + 1;
  for (;;) {
    uint32_t item = 0;
    if (glane == 0) item = atomicAdd(next, 1u);
    item = (uint32_t)__shfl((int)item, (int)leader, 64);
    if (item >= nitems) break;  // group-uniform
    const uint32_t path = qin[item];
    if (path >= P.npaths) continue;
    uint32_t fl = P.tflags[path];  // bit 0: a shadow ray of bounce b; bit 1: an extension ray of b + 1
    uint32_t b = b0;
    for (;;) {
      if (fl & 1u) {
        const uint32_t r = trace_coop_one<K, S, B, LAYOUT>(ps, path, stk, glane);
        if (glane == 0) {
          if (b > b0) atomicAdd(P.counts + P.max_bounces + 2 + b, 1u);
          if (r & kCoopHit) {
            if (P.record) P.record[(size_t)b * P.npaths + path].y = 0u;
          } else {
            pts::store3(P.L + 3 * (size_t)path,
                        pts::load3(P.L + 3 * (size_t)path) + pts::load3(P.nee + 3 * (size_t)path));
          }
        }
        __threadfence_block();
      }
      if (!(fl & 2u)) break;
      b++;
      const uint32_t r = trace_coop_one<K, S, B, LAYOUT>(pe, path, stk, glane);
      if (glane == 0 && b > b0 + 1) atomicAdd(P.counts + b, 1u);
      __threadfence_block();
      const uint32_t win = r & (kCoopLanes - 1u);
      bool keep = false, shadow = false;
      if (glane == win) {
        pts::GeoTerm gt;
        pts::shade_one<false>(P, b, path, keep, shadow, gt);
      }
      __threadfence_block();
      const uint32_t f = (shadow ? 1u : 0u) | (keep ? 2u : 0u);
      fl = (uint32_t)__shfl((int)f, (int)(leader + win), 64);
    }
  }
}

template <int K, int S, int LAYOUT>
hipError_t launch_tail_ks(const PtLaunch &P, const TraceParams &pe, const TraceParams &ps, hipStream_t st) {
  auto k = k_pt_tail<K, S, LAYOUT>;
  const uint32_t g = std::min(resident_grid((const void *)k, (int)kTailBlock),
                              (P.npaths * kCoopLanes + kTailBlock - 1) / kTailBlock);
  hipLaunchKernelGGL(k, dim3(g), dim3(kTailBlock), 0, st, P, pe, ps);
  return hipGetLastError();
}

template <int S, int LAYOUT>
hipError_t launch_tail_s(const PtLaunch &P, const TraceParams &pe, const TraceParams &ps, const TraceTune &t,
                         hipStream_t st) {
  constexpr int K8 = S < 8 ? S : 8;
  constexpr int K16 = S < 16 ? S : 16;
  if (t.lds <= 8) return launch_tail_ks<K8, S, LAYOUT>(P, pe, ps, st);
  return launch_tail_ks<K16, S, LAYOUT>(P, pe, ps, st);
}

template <int LAYOUT>
hipError_t launch_tail_layout(const PtLaunch &P, const TraceParams &pe, const TraceParams &ps, uint32_t s,
                              const TraceTune &t, hipStream_t st) {
  switch (s) {
    case 8: return launch_tail_s<8, LAYOUT>(P, pe, ps, t, st);
    case 16: return launch_tail_s<16, LAYOUT>(P, pe, ps, t, st);
    case 24: return launch_tail_s<24, LAYOUT>(P, pe, ps, t, st);
    case 32: return launch_tail_s<32, LAYOUT>(P, pe, ps, t, st);
    case 48: return launch_tail_s<48, LAYOUT>(P, pe, ps, t, st);
    case 64: return launch_tail_s<64, LAYOUT>(P, pe, ps, t, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace
}  // namespace rdrt

#endif
