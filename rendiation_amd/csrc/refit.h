// refit.h -- device refit of the BLAS layers after a vertex update: same
// trees, boxes and triangle records recomputed from the new positions.
#ifndef RDRT_REFIT_H
#define RDRT_REFIT_H

#include <hip/hip_runtime.h>

#include "rdrt_layout.h"

namespace rdrt {

struct RefitArgs {
  const float *vertices;    // 3 per vertex, scene order
  const uint32_t *indices;  // 3 per triangle, scene vertex ids
  const uint32_t *tri_ids;  // fast slot -> triangle
  const uint32_t *redirect; // reference slot -> triangle
  uint32_t nslots;          // triangles
  float *tri_box;           // scratch: 6 per triangle (min xyz, max xyz)
  TriPre *tris;             // fast slot records (rewritten)
  RefNode *forest;
  PairNode *pairs;
  QuadNode *quads;
  QuadNodeQ *qquads;
  GeoRec *geos;
  uint32_t ngeos;
  // node indices per layout, deepest level first, and level boundaries
  // (host arrays of offsets, device arrays of indices)
  const uint32_t *forest_nodes, *pair_nodes, *quad_nodes;
};

// Enqueues the whole BLAS refit on `stream`: triangle records and boxes, then
// the reference forest, the child pairs and the child quads (with their
// quantized twins) level by level, then the geometry boxes.
hipError_t enqueue_refit(const RefitArgs &a, const uint32_t *forest_off, uint32_t forest_levels,
                         const uint32_t *pair_off, uint32_t pair_levels, const uint32_t *quad_off,
                         uint32_t quad_levels, hipStream_t stream);

// The BLAS triangle records (TriPre of triangle tri_ids[k] in slot k) and
// the quantized twins of the child quads, computed on the device with the
// host builder's arithmetic (a commit whose build left them to the device).
hipError_t enqueue_scene_records(const float *vertices, const uint32_t *indices, const uint32_t *tri_ids,
                                 uint32_t nslots, TriPre *tris, const QuadNode *quads, uint32_t nquads,
                                 QuadNodeQ *qquads, hipStream_t stream);

}  // namespace rdrt

#endif
