// batch.h -- batched multi-scene tracing with per-bounce compaction.
#ifndef RDRT_BATCH_H
#define RDRT_BATCH_H

#include <hip/hip_runtime.h>

#include "../../include/rdrt.h"
#include "rdrt_layout.h"
#include "trace_kernels.h"

namespace rdrt {

// One bounce's spawn step: for every queued path whose ray hit, a new ray
// from the offset hit point in a cosine-weighted direction around the
// geometric normal facing the ray; survivors are compacted (stable) into
// qout.  Passed by value.
struct SpawnArgs {
  const uint32_t *qin;   // queue of this round (path ids)
  const uint32_t *count_in;
  uint32_t slots;        // grid size bound (max queue length)
  uint32_t *qout;
  uint32_t *count_out;
  uint32_t *tmp;  // block-local compacted runs (slots words)
  uint32_t *blk;  // per block: count, then offset (2 * ceil(slots / 256) words)
  const rdrt_ray *rays;  // rays of this round (by path id)
  rdrt_ray *out_rays;    // rays of the next round (by path id)
  rdrt_ray *record;      // optional second copy of the spawned rays (by path id)
  const rdrt_hit *hits;  // hits of this round (by path id)
  rdrt_hit *hits_later;  // hits of the later rounds (stride n): a path that ends gets misses there
  uint32_t later_rounds;
  const uint2 *hitx;
  uint32_t n;            // path-id bound
  uint32_t round, seed;
  const InstRec *insts;
  const GeoRec *geos;
  const uint32_t *indices;
  const float *vertices;
  uint32_t *status;
};

// Spawn kernel, then the scan of its block counts and the scatter into qout.
hipError_t enqueue_spawn(const SpawnArgs &a, hipStream_t st);
constexpr uint32_t kSpawnBlock = 256;
// queue[i] = first + i for i < n; *count = n; zero `ncounts` u32 words of zero_counts.
hipError_t enqueue_batch_init(uint32_t *queue, uint32_t first, uint32_t n, uint32_t *count, uint32_t *zero_counts,
                              uint32_t ncounts, hipStream_t st);

}  // namespace rdrt

#endif
