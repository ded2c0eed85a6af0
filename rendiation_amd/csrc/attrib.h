// attrib.h -- launch interface of the hit-attribute kernel.
#ifndef RDRT_ATTRIB_H
#define RDRT_ATTRIB_H

#include <hip/hip_runtime.h>

#include "../../include/rdrt.h"
#include "rdrt_layout.h"

namespace rdrt {

struct AttribArgs {
  const rdrt_ray *rays;
  const rdrt_ray_meta *meta;  // or null: tlas_idx
  const rdrt_hit *hits;
  const rdrt_hit_meta *hit_meta;
  uint32_t n, tlas_idx, sbt_offset, sbt_stride;
  // optional: per-ray (sbt offset, sbt stride, -, -) in place of the two above
  const uint4 *sbt;
  // optional work queue: thread i fetches hit queue[i] for i < *count (outputs
  // indexed by the queue entry; n = the grid's threads)
  const uint32_t *queue, *count;
  rdrt_hit_attributes *out;  // or null
  rdrt_hit_transform *xf_out;  // or null: hit context (transforms, object-space ray)
  // scene
  const uint32_t *binding;
  uint32_t nbinding, ntlas, ninst;
  const uint32_t *tlas_inst_first, *inst_lookup;
  const InstRec *insts;
  const float *inst_fwd;  // 16 per instance record: object-to-world
  const GeoRec *geos;
  const uint32_t *indices;
  const float *vertices, *normals, *uvs;  // normals / uvs null when the scene has none
};

hipError_t enqueue_hit_attributes(const AttribArgs &a, hipStream_t stream);

// The committed scene's part of AttribArgs (api.hip).
AttribArgs scene_attrib_args(const rdrt_ctx *c);

}  // namespace rdrt

#endif
