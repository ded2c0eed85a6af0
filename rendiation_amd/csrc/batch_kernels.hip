// batch_kernels.hip -- spawn and queue-initialisation kernels of batched
// multi-scene tracing.  Rays of many scenes share one queue; each ray's
// rdrt_ray_meta.tlas_idx selects its scene (the reference's per-ray TLAS
// selection, traverse_cpu.rs:63, traverse_gpu.rs:200-201).  Each bounce
// spawns a cosine-weighted ray at every hit and compacts the survivors.

#include "batch.h"
#include "compact_device.h"
#include "pt_device.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace {

using pt::V3;
using pt::v3;
constexpr uint32_t kBlock = kSpawnBlock;
constexpr float kFltEps = 1.1920928955078125e-7f;
constexpr float kFltMax = 3.402823466e+38f;

__device__ __forceinline__ V3 load3(const float *p) { return v3(p[0], p[1], p[2]); }

// Spawn sample of path `path` at bounce `round`: the Sobol (0,2)-sequence of
// the path, indexed by the bounce, scrambled by xxhash32(path, seed, 2 / 3).
__device__ bool spawn_one(const SpawnArgs &a, uint32_t path) {
  const uint4 h = *reinterpret_cast<const uint4 *>(a.hits + path);
  if (h.x == 0xffffffffu) return false;
  const float4 *rq = reinterpret_cast<const float4 *>(a.rays + path);
  const float4 r0 = rq[0], r1 = rq[1];
  const V3 o = v3(r0.x, r0.y, r0.z), d = v3(r1.x, r1.y, r1.z);
  const float t = __uint_as_float(h.y);
  const uint2 hx = a.hitx[path];
  const GeoRec &g = a.geos[hx.x];
  const InstRec &in = a.insts[hx.y];
  const uint32_t tri = g.primitive_start + h.x;
  const V3 p0 = load3(a.vertices + 3 * (size_t)a.indices[3 * (size_t)tri]);
  const V3 p1 = load3(a.vertices + 3 * (size_t)a.indices[3 * (size_t)tri + 1]);
  const V3 p2 = load3(a.vertices + 3 * (size_t)a.indices[3 * (size_t)tri + 2]);
  const V3 c = pt::cross3(p1 - p0, p2 - p0);
  const float *m = in.inv;
  V3 ng = pt::normalize3(v3(m[0] * c.x + m[1] * c.y + m[2] * c.z, m[4] * c.x + m[5] * c.y + m[6] * c.z,
                            m[8] * c.x + m[9] * c.y + m[10] * c.z));
  if (pt::dot3(ng, d) > 0.0f) ng = ng * -1.0f;
  const V3 pos = o + d * t;
  const float s0 = pt::van_der_corput(a.round, pt::xxhash32(path, a.seed, 2u));
  const float s1 = pt::sobol1(a.round, pt::xxhash32(path, a.seed, 3u));
  const V3 wo = pt::sample_cosine(ng, s0, s1);
  const V3 origin = pt::offset_ray_hit(pos, ng);
  const float4 w0 = make_float4(origin.x, origin.y, origin.z, kFltEps), w1 = make_float4(wo.x, wo.y, wo.z, kFltMax);
  float4 *q = reinterpret_cast<float4 *>(a.out_rays + path);
  q[0] = w0;
  q[1] = w1;
  if (a.record) {
    float4 *rr = reinterpret_cast<float4 *>(a.record + path);
    rr[0] = w0;
    rr[1] = w1;
  }
  return true;
}

__global__ __launch_bounds__(kBlock) void k_spawn(SpawnArgs a) {
  const uint32_t bid = blockIdx.x;
  const uint32_t count = min(*a.count_in, a.slots);
  const uint32_t nblocks = (count + kBlock - 1) / kBlock;
  if (bid >= nblocks) return;  // whole block
  const uint32_t i = bid * kBlock + threadIdx.x;
  bool keep = false;
  uint32_t path = 0;
  if (i < count) {
    path = a.qin[i];
    if (path < a.n) {
      keep = spawn_one(a, path);
      if (!keep)
        for (uint32_t r = 0; r < a.later_rounds; r++)
          *reinterpret_cast<uint4 *>(a.hits_later + (size_t)r * a.n + path) = make_uint4(RDRT_MISS, 0u, 0u, 0u);
    }
  }
  compact::block_compact2<kBlock>(keep, false, path, bid, a.tmp, nullptr, a.blk, nullptr);
}

__global__ void k_batch_init(uint32_t *queue, uint32_t first, uint32_t n, uint32_t *count, uint32_t *zero_counts,
                             uint32_t ncounts) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) queue[i] = first + i;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ncounts; i += stride) zero_counts[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = n;
}

}  // namespace

hipError_t enqueue_spawn(const SpawnArgs &a, hipStream_t st) {
  if (a.slots == 0) return hipSuccess;
  const uint32_t grid = (a.slots + kBlock - 1) / kBlock;
  uint32_t *cnt = a.blk, *off = a.blk + grid;
  hipLaunchKernelGGL(k_spawn, dim3(grid), dim3(kBlock), 0, st, a);
  hipLaunchKernelGGL(compact::k_block_scan, dim3(1), dim3(compact::kScanThreads), 0, st, a.count_in, a.slots, kBlock,
                     (const uint32_t *)cnt, (const uint32_t *)nullptr, off, (uint32_t *)nullptr, a.count_out,
                     (uint32_t *)nullptr);
  hipLaunchKernelGGL((compact::k_scatter2<kBlock>), dim3(grid), dim3(kBlock), 0, st, a.count_in, a.slots,
                     (const uint32_t *)a.tmp, (const uint32_t *)nullptr, (const uint32_t *)cnt,
                     (const uint32_t *)nullptr, (const uint32_t *)off, (const uint32_t *)nullptr, a.qout,
                     (uint32_t *)nullptr, a.slots);
  return hipGetLastError();
}

hipError_t enqueue_batch_init(uint32_t *queue, uint32_t first, uint32_t n, uint32_t *count, uint32_t *zero_counts,
                              uint32_t ncounts, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_init, dim3(1024), dim3(256), 0, st, queue, first, n, count, zero_counts, ncounts);
  return hipGetLastError();
}

}  // namespace rdrt
