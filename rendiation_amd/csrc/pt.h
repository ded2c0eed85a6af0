// pt.h -- wavefront path tracer: launch description and entry points.
#ifndef RDRT_PT_H
#define RDRT_PT_H

#include <hip/hip_runtime.h>

#include "../../include/rdrt.h"
#include "det_sum.h"
#include "rdrt_layout.h"
#include "trace_kernels.h"

namespace rdrt {

constexpr uint32_t kPtBlock = 256;       // threads per block of the per-path kernels
constexpr uint32_t kPtMaxBounces = 16;   // upper bound on max_bounces

// Everything one frame's kernels read, passed by value (so a captured graph
// holds it).  Device arrays are indexed by path id unless noted.
struct PtLaunch {
  const float *views;  // nviews * 20: view_projection_inv[16], camera_position[3], pad (device)
  uint32_t nviews;
  uint32_t width, height, spp, npaths;
  uint32_t max_bounces, rr_start, seed, sample_base;
  float sky_up[3], sky_down[3];
  float sun_dir[3], sun_radiance[3];

  // scene
  const InstRec *insts;
  const GeoRec *geos;
  const uint32_t *indices;
  const float *vertices;
  const float *normals;          // per-vertex normals or null (geometric normal)
  const uint32_t *tri_material;  // per-triangle material offset or null
  const Material *materials;
  uint32_t nmaterials;

  // path state
  rdrt_ray *rays, *shadow_rays;
  rdrt_hit *hits, *shadow_hits;
  uint2 *hitx;  // (global geometry, forest instance) of each hit
  float *beta, *L, *nee;
  uint32_t *qa, *qb, *qs;  // extension queues (ping-pong) and the shadow queue
  uint32_t *counts;        // [0..max_bounces+1]: extension queue sizes; [max_bounces+2+b]: shadow queue sizes
  // compaction scratch: block-local runs of the two output queues, and per
  // block [count a | count b | offset a | offset b], blk_stride words each
  uint32_t *tmp_a, *tmp_b, *blk;
  uint32_t blk_stride;
  uint32_t *status;
  uint32_t *work;  // [0]: work counter of persistent traversal launches; [1]: of the tail kernel
  // tail: from bounce tail_b0 (>= 1) on, the tail kernel finishes every path
  // (0 = per-bounce launches throughout); tflags[path]: what shade(tail_b0)
  // left the path (bit 0: a shadow ray, bit 1: an extension ray)
  uint32_t tail_b0;
  uint32_t *tflags;
  float *out;  // width * height * 3

  // path vertices for the adjoint, or null: slot k of path p at
  // record[k * npaths + p], max_bounces + 2 slots; (material, cos_l / pi of
  // unoccluded light or 0), or (kVertSkyUp / kVertSkyDown / kVertEnd, 0)
  uint2 *record;

  // geometry adjoint (replay frames): pixel adjoint (views * width * height
  // * 3), gradients w.r.t. vertex positions and normals (nvertices * 3),
  // per-(bounce, queue slot) gradient terms (max_bounces * npaths * 3
  // uint4) and the header of their reproducible sum
  const float *adj;
  float *g_pos, *g_nrm;
  uint32_t nvertices;
  uint4 *grec;
  det::Header *det_h;
};

// Device scratch of one reproducible accumulation (det_sum.h): n sums and
// flags and the header.
struct DetBufs {
  det::u64 *acc;
  uint32_t *flags;
  det::Header *h;
};

constexpr uint32_t kVertSkyUp = 0xFFFFFFF0u, kVertSkyDown = 0xFFFFFFF1u, kVertEnd = 0xFFFFFFF2u;

// Optional per-frame diagnostics (eager frames only).
struct PtDiag {
  hipEvent_t *ev_begin = nullptr, *ev_end = nullptr;  // one pair per traversal launch, or null
  uint32_t *counters = nullptr;                       // 3 u32 per path: visit counters, or null
  unsigned long long *totals = nullptr;               // 3 u64 sums of the counters
  uint32_t *max_visits = nullptr;                     // per launch: largest visit sum of one ray
  bool fast_order = false;                            // count the fast traversal's visits
};

// The tail kernel of a frame (pt_tail_<layout>.hip): every path of the
// queue shade(L.tail_b0) read, from its shadow ray of that bounce to its
// end.  pe / ps: the extension / shadow traversal parameters (no queues).
hipError_t launch_pt_tail(const PtLaunch &L, const TraceParams &pe, const TraceParams &ps, uint32_t push_depth,
                          const TraceTune &tune, hipStream_t stream);
hipError_t launch_pt_tail_0(const PtLaunch &L, const TraceParams &pe, const TraceParams &ps, uint32_t stack,
                            const TraceTune &tune, hipStream_t stream);
hipError_t launch_pt_tail_1(const PtLaunch &L, const TraceParams &pe, const TraceParams &ps, uint32_t stack,
                            const TraceTune &tune, hipStream_t stream);
hipError_t launch_pt_tail_2(const PtLaunch &L, const TraceParams &pe, const TraceParams &ps, uint32_t stack,
                            const TraceTune &tune, hipStream_t stream);

// Enqueues one frame on `stream` (capturable when diag is null: no memset
// nodes, no host round trips).
hipError_t enqueue_pt_frame(const PtLaunch &L, const TraceParams &scene_tp, uint32_t push_depth,
                            const TraceTune &tune, hipStream_t stream, const PtDiag *diag = nullptr);

// Gradient of sum(adj * frame) w.r.t. material albedo and emission from the
// recorded vertices of frame L, written to grad_albedo / grad_emission (3 *
// nmaterials each).  Reproducible: db holds 6 * nmaterials sums.
hipError_t enqueue_pt_backward(const PtLaunch &L, const float *adj, float *grad_albedo, float *grad_emission,
                               const DetBufs &db, hipStream_t stream);
// Replays the recorded frame L (same kernels, no shadow rays), records the
// gradient terms of every bounce in L.grec and writes the gradient w.r.t.
// vertex positions and normals to L.g_pos / L.g_nrm (g_nrm may be null).
// Reproducible: db holds 6 * nvertices sums.
hipError_t enqueue_pt_replay(const PtLaunch &L, const TraceParams &scene_tp, uint32_t push_depth,
                             const TraceTune &tune, const DetBufs &db, hipStream_t stream);
// Packs device albedo / emission arrays (n * 3 each) into Material records.
hipError_t enqueue_pack_materials(const float *albedo, const float *emission, uint32_t n, Material *out,
                                  hipStream_t stream);
// adj[i] = 2 (a[i] - b[i]); *loss = sum (a[i] - b[i])^2 in FP64, summed in
// a fixed order (per-block partials in `part`, l2_adjoint_parts(n) doubles,
// then added in block order).
uint32_t l2_adjoint_parts(uint32_t n);
hipError_t enqueue_l2_adjoint(const float *a, const float *b, uint32_t n, float *adj, double *loss, double *part,
                              hipStream_t stream);

// Stable stream compaction of indices whose mask is non-zero (block-local
// runs, a scan of the block counts, a scatter); out_count receives the kept
// count.  `scratch` holds compact_scratch_words(n) u32.
size_t compact_scratch_words(uint32_t n);
hipError_t enqueue_compact(const uint32_t *mask, uint32_t n, uint32_t *out_idx, uint32_t *out_count,
                           uint32_t *scratch, hipStream_t stream);

// Sobol / van der Corput samples: out[2i] = vdc(index[i], scr[2i]),
// out[2i+1] = sobol1(index[i], scr[2i+1]).
hipError_t enqueue_sobol(const uint32_t *index, const uint32_t *scr, uint32_t n, float *out, hipStream_t stream);

}  // namespace rdrt

#endif
