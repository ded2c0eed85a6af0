// trace_kernels.h -- launch interface of the closest-hit kernels.
#ifndef RDRT_TRACE_KERNELS_H
#define RDRT_TRACE_KERNELS_H

#include <hip/hip_runtime.h>

#include "../../include/rdrt.h"
#include "rdrt_layout.h"

namespace rdrt {

// Device view of a committed scene plus one launch's I/O.
struct TraceParams {
  const rdrt_ray *rays;
  const rdrt_ray_meta *meta;
  rdrt_hit *hits;
  rdrt_hit_meta *hit_meta;
  uint32_t *counters;
  uint32_t *status;  // device word; bit 0 = a lane's traversal stack would have overflowed
  uint32_t n, flags, cull_mask, tlas_idx;
  // Optional work queue: thread i traces ray queue[i] for i < *count (both
  // device pointers; count null = n).  Rays and outputs are indexed by the
  // queue entry.
  const uint32_t *queue;
  const uint32_t *count;
  uint32_t slots;  // queue slots the grid covers (0 = n); ray ids stay < n
  uint2 *hitx;     // optional: (global geometry index, forest instance index) per hit
  uint32_t *work;  // device work counter of the persistent kernel
  // Optional second segment traced by the same launch (fast traversal
  // only): queue2[0 .. min(slots2, *count2)) with its own rays, hits, hit
  // extras and flags; ray ids index rays2 / hits2 and stay < n.  The
  // scene, mask, TLAS index and counters are shared.  Used to trace two
  // independent queues (shadow rays of one bounce and extension rays of
  // the next) in one launch instead of two.
  const rdrt_ray *rays2;
  rdrt_hit *hits2;
  uint2 *hitx2;
  const uint32_t *queue2;
  const uint32_t *count2;  // null = no second segment
  uint32_t slots2, flags2;
  // Cooperative traversal (fast layouts): when both segments together hold
  // at most coop_max live rays, the cooperative kernel traces the launch and
  // the one-ray-per-lane kernel returns at once; 0 = never.
  uint32_t coop_max;
  // A bound the caller knows on the live rays of both segments together
  // (0 = their slots): when coop_max reaches it, only the cooperative kernel
  // is launched.
  uint32_t live_max;
  // 1: the caller's previous kernel on the stream left *work at zero, so a
  // persistent launch needs no clearing launch first
  uint32_t work_zeroed;

  const PairNode *tri_pairs;
  const TriPre *tris;
  const uint32_t *tri_ids;   // TriPre slot -> global triangle (fast traversal)
  const uint32_t *redirect;  // forest slot -> global triangle (reference traversal)
  const GeoRec *geos;
  const PairNode *tlas_pairs;
  const InstRec *insts;
  const TlasRoot *tlas_roots;
  const QuadNode *tri_quads;         // child-quad layout of the same trees
  const QuadNode *tlas_quads;
  const QuadNodeQ *tri_qquads;       // its quantized twin
  const QuadNodeQ *tlas_qquads;
  const uint32_t *tlas_quad_roots;   // per TLAS
  const uint32_t *binding;
  uint32_t nbinding, ntlas;

  const RefNode *tri_forest;
  const RefNode *tlas_forest;
  const uint32_t *indices;
  const float *vertices;
  // any-hit behaviour (RDRT_ANYHIT_*) per global triangle of hits on
  // geometry that is not opaque for the ray, or null (accept every hit)
  const uint8_t *anyhit;
};

__host__ __device__ inline uint32_t slots_of(const TraceParams &p) { return p.slots ? p.slots : p.n; }
// Queue slots a launch's grid covers: both segments.
__host__ __device__ inline uint32_t grid_slots(const TraceParams &p, uint32_t block) {
  const uint32_t a = (slots_of(p) + block - 1) / block, b = p.count2 ? (p.slots2 + block - 1) / block : 0u;
  return (a + b) * block;
}

// Smallest stack (entries per lane) that a fast-traversal kernel is built
// with for the given push depth; 0 if none is large enough.
uint32_t stack_entries_for(uint32_t push_depth);

// Launch shape of the fast traversal (the autotuner's search space).
struct TraceTune {
  uint32_t block = 256;  // threads per block: 64, 128 or 256
  uint32_t lds = 16;     // stack entries per lane kept in LDS: 8 or 16 (deeper ones in private memory)
  uint32_t waves = 0;    // 8 / 6 = ask the register allocator for 8 / 6 waves per SIMD; 0 = its own choice
  uint32_t persist = 0;  // 1 = resident-size grid whose waves pull 64-ray chunks from a counter
  uint32_t cull = 0;     // 1 = a first pass finishes the rays that miss their scene's root box and
                         //     queues the others densely for the traversal
  uint32_t wide = 0;     // node layout: 0 = child pairs, 1 = child quads (four boxes per node),
                         //     2 = quantized child quads (64-byte nodes)
  uint32_t coop = 0;     // launches with at most this many live rays use the cooperative kernel
                         //     (8 lanes per ray); 0 = never
  uint32_t tail = 0;     // path-tracer frames: from this bounce on one tail launch finishes every path
                         //     (8 lanes per path); 0 = per-bounce launches throughout
};

// Cooperative-kernel thresholds the frame and AO autotuners try.
constexpr uint32_t kCoopChoices[3] = {0u, 1u << 15, 1u << 18};

// Scratch of a culled launch: tmp (slots words), blk (2 * ceil(slots /
// kCullBlock) words), queue (slots words), count (1 word).
constexpr uint32_t kCullBlock = 256;
struct CullScratch {
  uint32_t *tmp, *blk, *queue, *count;
  // rays that miss also get miss records in `later_rounds` further hit
  // arrays spaced `stride` records apart (batched rounds); 0 = none
  rdrt_hit *later;
  uint32_t later_rounds, stride;
};
size_t cull_scratch_words(uint32_t slots);

// The cull pass and the scan and scatter of its queue into cs.queue /
// cs.count (rays 0..p.n-1; p must have no queue of its own).
hipError_t enqueue_cull(const TraceParams &p, const CullScratch &cs, hipStream_t stream);
// enqueue_cull, then the fast traversal over the queue.
hipError_t launch_trace_culled(const TraceParams &p, uint32_t push_depth, const TraceTune &tune,
                               const CullScratch &cs, hipStream_t stream);

// Enqueues one traversal launch on `stream`.  mode = RDRT_TRAVERSAL_*.  A
// persistent launch needs p.work (zeroed by a kernel launched first, so the
// pair is capturable).  push_depth: the deepest push count of the layout
// tune.wide selects.
hipError_t launch_trace(const TraceParams &p, uint32_t mode, uint32_t push_depth, const TraceTune &tune,
                        hipStream_t stream);

// Resident blocks of `kernel` at `block` threads on the current device.
uint32_t resident_grid(const void *kernel, int block);
// Per-block-size instantiations (trace_fast_<B>.hip); `stack` from stack_entries_for.
hipError_t launch_fast_64(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st);
hipError_t launch_fast_128(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st);
hipError_t launch_fast_256(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st);
// The same over the child-quad layout (trace_quad_<B>.hip) and its
// quantized twin (trace_qquad_<B>.hip).
hipError_t launch_quad_64(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st);
hipError_t launch_quad_128(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st);
hipError_t launch_quad_256(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st);
hipError_t launch_qquad_64(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st);
hipError_t launch_qquad_128(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st);
hipError_t launch_qquad_256(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st);

}  // namespace rdrt

#endif
