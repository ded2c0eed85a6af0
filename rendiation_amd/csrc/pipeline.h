// pipeline.h -- programmable ray-tracing pipelines on the device task graph:
// the reference's stage pipelines (shader/ray-tracing/src/api/pipeline.rs:
// 75-130, operator/mod.rs:34-69) run as its wavefront executor runs them
// (wavefront_compute/{mod.rs:113-190, pipeline.rs:95-207, trace_task.rs}),
// with the caller's HIP kernels as the ray-gen, closest-hit and miss stages.
//
// Groups: 0 = trace (the library's), 1 = ray gen, then the closest-hit
// stages, then the miss stages.  Each round polls them in that order.  A
// trace poll is pre (queue new rays, hand back finished payloads), the
// traversal, post (spawn the hit group's closest-hit task or the miss task).
// The stage groups of a round take one pre launch for all of them (take
// returned traces), the caller's stage kernels in group order, and one post
// launch (spawn the traces they asked for, resolve the rest).
#ifndef RDRT_PIPELINE_H
#define RDRT_PIPELINE_H

#include <hip/hip_runtime.h>

#include "../../include/rdrt.h"
#include "attrib.h"
#include "rdrt_layout.h"
#include "task_graph.h"
#include "trace_kernels.h"

namespace rdrt {

constexpr uint32_t kPipeTrace = 0, kPipeRayGen = 1;

// Per-slot arrays of a stage group (the caller's view, rdrt_stage_args).
struct StageArrays {
  uint32_t *payload, *result, *action, *call_payload, *launch, *trace_of;
  rdrt_trace_call *call;
};

struct PipeLaunch {
  tg::Graph gr;
  uint32_t words;                    // payload words
  uint32_t nclosest, nmiss;
  uint32_t closest_group[4], miss_group[4];  // group of each hit group / miss index (tg::kNone: no stage)
  // trace group, per slot
  rdrt_ray *rays;
  rdrt_ray_meta *meta;
  uint4 *sbt;       // (sbt offset, sbt stride, miss index, -)
  rdrt_hit *hits;
  rdrt_hit_meta *hit_meta;
  uint2 *hitx;
  rdrt_hit_attributes *attrs;  // or null
  rdrt_hit_transform *xforms;  // or null
  uint32_t *tpay;   // payload words per slot
  uint32_t *tlaunch, *tawait;  // launch index; group of the awaited closest-hit / miss task
  uint32_t *tq;     // this round's rays to traverse
  uint32_t *tcount;  // [q]: their count in rounds of parity q; [2]: rays traversed by the call
  uint32_t qpar;
  uint32_t *work;    // the persistent traversal's work counter (zeroed by the trace pre-pass)
  StageArrays st[tg::kMaxGroups];    // groups 1..n-1
  // launch geometry of the tile
  uint32_t width, height, tile_x, tile_y, tile_w;
  const InstRec *insts;  // instance hit-group offsets
};

// One tile: reset the graph, spawn tile_w * tile_h ray-gen tasks, run
// `rounds` rounds.  stages[g] is the caller's kernel of group g (g >= 1);
// `after` receives the graph with the parities the launches leave.  With
// L.attrs or L.xforms, `attrib` holds the scene part of the hit-attribute
// fetch that follows each traversal.
hipError_t enqueue_pipeline_tile(const PipeLaunch &L, const void *const *stages, uint32_t tile_h, uint32_t rounds,
                                 const TraceParams &scene, const AttribArgs &attrib, uint32_t push_depth,
                                 const TraceTune &tune, uint32_t *work, void *user, bool first_tile, hipStream_t st,
                                 tg::Graph *after);

}  // namespace rdrt

#endif
