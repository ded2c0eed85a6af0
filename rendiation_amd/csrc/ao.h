// ao.h -- ambient occlusion (scene/rendering/gpu-ray-tracing/src/feature/
// ao.rs:176-315) as a ray-tracing pipeline executed on the device task
// graph, the way the reference's wavefront executor runs it
// (shader/ray-tracing/src/backend/wavefront_compute/{mod.rs:113-190,
// pipeline.rs:95-207}).
//
// Task groups, in the reference's order (pipeline.rs:160-206):
//   0 trace      -- one traversal per task; spawns the closest-hit task of
//                   the ray's hit group or the miss task, waits for it and
//                   hands its payload back to the caller (2 in flight per
//                   pixel, ao.rs:297)
//   1 ray gen    -- camera ray, then_trace, then the running average
//   2 closest    -- primary hits: cosine-sampled occlusion ray, then_trace
//   3 closest    -- occlusion-ray hits: payload 0
//   4 miss       -- payload 1
#ifndef RDRT_AO_H
#define RDRT_AO_H

#include <hip/hip_runtime.h>

#include "../../include/rdrt.h"
#include "rdrt_layout.h"
#include "task_graph.h"
#include "trace_kernels.h"

namespace rdrt {

constexpr uint32_t kAoTrace = 0, kAoRayGen = 1, kAoClosest = 2, kAoClosestTest = 3, kAoMiss = 4, kAoGroups = 5;
constexpr uint32_t kAoMaxSample = 256;  // ao.rs:5
// Rounds after which every AO task has finished, whatever the scene.  With
// the groups polled in the order trace, ray gen, closest, closest (occlusion),
// miss: round 1 ray gen spawns the camera trace; round 2 traces it and its
// closest-hit or miss child runs (a miss resolves and wakes the trace);
// round 3 resolves a missed camera trace and its ray gen, and traces the
// occlusion ray whose child resolves; round 4 resolves the occlusion trace
// and the closest hit; round 5 resolves the camera trace and the ray gen.
// Later rounds find every group empty, so running min(rounds, 5) gives the
// same buffer and the same drained graph.
constexpr uint32_t kAoChainRounds = 5;
// Words of the per-call frame block (device): view_projection_inv[16],
// camera position[3], pad, local occlusion direction[3], sample index.
constexpr uint32_t kAoFrameWords = 24;

struct AoLaunch {
  tg::Graph gr;
  // per trace-task slot: the ray, its closest hit and (geometry, instance)
  rdrt_ray *rays;
  rdrt_hit *hits;
  uint2 *hitx;
  uint32_t *tq0, *tq1;  // the next traversal's primary / occlusion rays (queued at spawn), qcap each
  uint32_t *tcount;     // [2q], [2q + 1]: their counts in rounds of parity q; [4], [5]: primary /
                        // occlusion rays traced by the call
  uint32_t qpar;        // parity of the queues being filled / traced (host-kept)
  uint32_t qcap;
  uint32_t *work;       // the persistent traversal's work counter (zeroed by the stage poll)
  const float *frame;   // kAoFrameWords
  uint32_t width, height, tile_x, tile_y, tile_w;
  void *accum;  // f32 per pixel, or RGBA8 per pixel
  uint32_t format;
  // scene
  const InstRec *insts;
  const GeoRec *geos;
  const uint32_t *indices;
  const float *vertices, *normals;
};

// One tile: reset the graph, spawn tile_w * tile_h ray-gen tasks, run
// `rounds` rounds (each group: prepare, poll, compact).  `after` receives
// the graph with the parities the enqueued launches leave.
hipError_t enqueue_ao_tile(const AoLaunch &L, uint32_t tile_h, uint32_t rounds, const TraceParams &scene,
                           uint32_t push_depth, const TraceTune &tune, uint32_t *work, bool first_tile,
                           hipStream_t st, tg::Graph *after = nullptr);

}  // namespace rdrt

#endif
