// pipeline_stages.hip -- stage kernels of the pipeline tests, built into
// tests/stages/libpipeline_stages.so (build_stages.py).  They are the
// caller's side of rdrt_pipeline_*: __global__ functions taking
// rdrt_stage_args, looked up through test_stage().
//
// "debug": the pipeline of shader/ray-tracing/src/test.rs:33-141.  The ray
// gen writes (0, 0, 50/255, 1) to an RGBA8 image, traces a pinhole camera
// ray (origin 0, direction normalize((x + .5) / w * 2 - 1, 1 - (y + .5) / h
// * 2, -1), range [0.1, 100], back faces culled) and, once the trace
// returns, adds 100/255 to blue; the closest hit writes the barycentrics to
// red and green; the miss does nothing.  Writes and reads go through the
// 8-bit UNORM format of the reference's target (round to nearest, ties to
// even).  User buffer: n hit records (the closest hit / miss record of each
// pixel's ray), then the n RGBA8 pixels.
//
// "chain": nested traces.  The ray gen traces the camera ray (origin
// (0.0137, -0.0219, 0), no culling)
// with payload (0, 0, 0); a closest hit adds prim_id + 1 to word 0 and 1 to
// word 1, and at depth (word 2) 0 or 1 traces the mirror ray
// (position + 1e-3 * geometric normal, reflect(d, n), range [0, 100]) with
// word 2 + 1, taking that trace's payload as its own when it returns; a miss
// adds 0x10000 to word 0.  The ray gen stores the returned payload.  User
// buffer: n * 3 rays, then n * 3 hits (per pixel and depth, as the stages
// saw them), then n * 3 hit meta records, then n * 3 payload words.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rdrt.h"

namespace {

__device__ __forceinline__ uint8_t unorm8(float x) {
  x = fminf(fmaxf(x, 0.0f), 1.0f);
  return (uint8_t)rintf(x * 255.0f);
}

__device__ __forceinline__ bool polled(const rdrt_stage_args &a, uint32_t &slot) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *a.count) return false;
  slot = a.slots[i];
  return true;
}

__device__ __forceinline__ void camera_dir(uint32_t id, const rdrt_stage_args &a, float d[3]) {
  const uint32_t w = a.launch_size[0], h = a.launch_size[1];
  const uint32_t x = id % w, y = id / w;
  const float fx = ((float)x + 0.5f) / (float)w * 2.0f - 1.0f;
  const float fy = 1.0f - ((float)y + 0.5f) / (float)h * 2.0f;
  const float fz = -1.0f;
  const float len = sqrtf(fx * fx + fy * fy + fz * fz);
  d[0] = fx / len;
  d[1] = fy / len;
  d[2] = fz / len;
}

__device__ __forceinline__ void set_call(rdrt_trace_call &c, const float o[3], const float d[3], float tmin, float tmax,
                                         uint32_t flags) {
  for (int k = 0; k < 3; k++) {
    c.origin[k] = o[k];
    c.direction[k] = d[k];
  }
  c.tmin = tmin;
  c.tmax = tmax;
  c.flags = flags;
  c.cull_mask = 0xffffffffu;
  c.tlas_idx = 0;
  c.miss_index = 0;
  c.sbt_offset = 0;
  c.sbt_stride = 0;
  c.pad[0] = c.pad[1] = 0;
}

// ---- debug ----

__global__ void debug_ray_gen(rdrt_stage_args a) {
  uint32_t s;
  if (!polled(a, s)) return;
  const uint32_t id = a.launch[s];
  const size_t n = (size_t)a.launch_size[0] * a.launch_size[1];
  uint8_t *px = reinterpret_cast<uint8_t *>(a.user) + n * sizeof(rdrt_hit) + 4 * (size_t)id;
  if (a.polls[s] == 0) {
    px[0] = unorm8(0.0f);
    px[1] = unorm8(0.0f);
    px[2] = unorm8(50.0f / 255.0f);
    px[3] = unorm8(1.0f);
    const float o[3] = {0.0f, 0.0f, 0.0f};
    float d[3];
    camera_dir(id, a, d);
    set_call(a.call[s], o, d, 0.1f, 100.0f, RDRT_RAY_FLAG_CULL_BACK_FACING_TRIANGLES);
    for (uint32_t w = 0; w < a.payload_words; w++) a.call_payload[(size_t)s * a.payload_words + w] = 0u;
    a.action[s] = RDRT_STAGE_TRACE;
  } else {
    const float r = (float)px[0] / 255.0f, g = (float)px[1] / 255.0f, b = (float)px[2] / 255.0f;
    px[0] = unorm8(r);
    px[1] = unorm8(g);
    px[2] = unorm8(b + 100.0f / 255.0f);
    px[3] = unorm8(1.0f);
  }
}

__global__ void debug_closest_hit(rdrt_stage_args a) {
  uint32_t s;
  if (!polled(a, s)) return;
  const uint32_t id = a.launch[s];
  const size_t n = (size_t)a.launch_size[0] * a.launch_size[1];
  const rdrt_hit h = a.hits[a.trace_of[s]];
  reinterpret_cast<rdrt_hit *>(a.user)[id] = h;
  uint8_t *px = reinterpret_cast<uint8_t *>(a.user) + n * sizeof(rdrt_hit) + 4 * (size_t)id;
  const float b = (float)px[2] / 255.0f;
  px[0] = unorm8(h.u);
  px[1] = unorm8(h.v);
  px[2] = unorm8(b);
  px[3] = unorm8(1.0f);
}

__global__ void debug_miss(rdrt_stage_args a) {
  uint32_t s;
  if (!polled(a, s)) return;
  reinterpret_cast<rdrt_hit *>(a.user)[a.launch[s]] = a.hits[a.trace_of[s]];
}

// ---- chain ----

constexpr uint32_t kChainDepth = 3;

__global__ void chain_ray_gen(rdrt_stage_args a) {
  uint32_t s;
  if (!polled(a, s)) return;
  const uint32_t id = a.launch[s];
  const size_t n = (size_t)a.launch_size[0] * a.launch_size[1];
  if (a.polls[s] == 0) {
    // off the axis, so that no camera ray meets a cube edge of the test
    // scenes exactly (a mirror ray from an edge starts on another face's
    // plane, where a hit at t = tmin depends on the box layout)
    const float o[3] = {0.0137f, -0.0219f, 0.0f};
    float d[3];
    camera_dir(id, a, d);
    set_call(a.call[s], o, d, 0.0f, 100.0f, 0u);
    for (uint32_t w = 0; w < 3; w++) a.call_payload[(size_t)s * a.payload_words + w] = 0u;
    a.action[s] = RDRT_STAGE_TRACE;
  } else {
    uint32_t *out = reinterpret_cast<uint32_t *>(
        reinterpret_cast<uint8_t *>(a.user) + n * kChainDepth * (sizeof(rdrt_ray) + sizeof(rdrt_hit) + sizeof(rdrt_hit_meta)));
    for (uint32_t w = 0; w < 3; w++) out[(size_t)id * 3 + w] = a.trace_result[(size_t)s * a.payload_words + w];
  }
}

__device__ __forceinline__ void chain_record(const rdrt_stage_args &a, uint32_t s, uint32_t id, uint32_t depth) {
  const size_t n = (size_t)a.launch_size[0] * a.launch_size[1];
  if (depth >= kChainDepth) return;
  const uint32_t tr = a.trace_of[s];
  rdrt_ray *rays = reinterpret_cast<rdrt_ray *>(a.user);
  rdrt_hit *hits = reinterpret_cast<rdrt_hit *>(rays + n * kChainDepth);
  rdrt_hit_meta *meta = reinterpret_cast<rdrt_hit_meta *>(hits + n * kChainDepth);
  rays[(size_t)id * kChainDepth + depth] = a.rays[tr];
  hits[(size_t)id * kChainDepth + depth] = a.hits[tr];
  meta[(size_t)id * kChainDepth + depth] = a.hit_meta[tr];
}

__global__ void chain_closest_hit(rdrt_stage_args a) {
  uint32_t s;
  if (!polled(a, s)) return;
  uint32_t *p = a.payload + (size_t)s * a.payload_words;
  if (a.polls[s] == 0) {
    const uint32_t id = a.launch[s], tr = a.trace_of[s];
    const uint32_t depth = p[2];
    chain_record(a, s, id, depth);
    p[0] += a.hits[tr].prim_id + 1u;
    p[1] += 1u;
    if (depth < 2u) {
      const rdrt_hit_attributes at = a.attributes[tr];
      const rdrt_ray r = a.rays[tr];
      const float d[3] = {r.dx, r.dy, r.dz};
      const float *nn = at.geometric_normal;
      const float dn = d[0] * nn[0] + d[1] * nn[1] + d[2] * nn[2];
      float o[3], m[3];
      for (int k = 0; k < 3; k++) {
        o[k] = at.position[k] + 1e-3f * nn[k];
        m[k] = d[k] - 2.0f * dn * nn[k];
      }
      set_call(a.call[s], o, m, 0.0f, 100.0f, 0u);
      uint32_t *cp = a.call_payload + (size_t)s * a.payload_words;
      cp[0] = p[0];
      cp[1] = p[1];
      cp[2] = depth + 1u;
      a.action[s] = RDRT_STAGE_TRACE;
    }
  } else {
    for (uint32_t w = 0; w < 3; w++) p[w] = a.trace_result[(size_t)s * a.payload_words + w];
  }
}

__global__ void chain_miss(rdrt_stage_args a) {
  uint32_t s;
  if (!polled(a, s)) return;
  uint32_t *p = a.payload + (size_t)s * a.payload_words;
  chain_record(a, s, a.launch[s], p[2]);
  p[0] += 0x10000u;
}

// ---- sbt ----
// Hit-group and miss dispatch.  The ray gen traces the chain's camera ray
// with sbt offset id % 2, stride (id / 2) % 2 and miss index (id / 4) % 3,
// payload 0xFFFF, and stores the returned payload; closest-hit stage K sets
// the payload to K, miss stage K to 100 + K, and both record the ray, hit,
// hit meta and (closest hit) hit context of their pixel.  User buffer: n
// payload words, n rays, n hits, n hit metas, n hit contexts.

struct SbtOut {
  uint32_t *pay;
  rdrt_ray *rays;
  rdrt_hit *hits;
  rdrt_hit_meta *meta;
  rdrt_hit_transform *xf;
};

__device__ __forceinline__ SbtOut sbt_out(const rdrt_stage_args &a) {
  const size_t n = (size_t)a.launch_size[0] * a.launch_size[1];
  SbtOut o;
  o.pay = reinterpret_cast<uint32_t *>(a.user);
  o.rays = reinterpret_cast<rdrt_ray *>(o.pay + n);
  o.hits = reinterpret_cast<rdrt_hit *>(o.rays + n);
  o.meta = reinterpret_cast<rdrt_hit_meta *>(o.hits + n);
  o.xf = reinterpret_cast<rdrt_hit_transform *>(o.meta + n);
  return o;
}

__global__ void sbt_ray_gen(rdrt_stage_args a) {
  uint32_t s;
  if (!polled(a, s)) return;
  const uint32_t id = a.launch[s];
  if (a.polls[s] == 0) {
    const float o[3] = {0.0137f, -0.0219f, 0.0f};
    float d[3];
    camera_dir(id, a, d);
    set_call(a.call[s], o, d, 0.0f, 100.0f, 0u);
    a.call[s].sbt_offset = id % 2u;
    a.call[s].sbt_stride = (id / 2u) % 2u;
    a.call[s].miss_index = (id / 4u) % 3u;
    a.call_payload[(size_t)s * a.payload_words] = 0xFFFFu;
    a.action[s] = RDRT_STAGE_TRACE;
  } else {
    sbt_out(a).pay[id] = a.trace_result[(size_t)s * a.payload_words];
  }
}

template <uint32_t K, bool HIT>
__global__ void sbt_stage(rdrt_stage_args a) {
  uint32_t s;
  if (!polled(a, s)) return;
  const uint32_t id = a.launch[s], tr = a.trace_of[s];
  const SbtOut o = sbt_out(a);
  o.rays[id] = a.rays[tr];
  o.hits[id] = a.hits[tr];
  o.meta[id] = a.hit_meta[tr];
  if (HIT) o.xf[id] = a.transforms[tr];
  a.payload[(size_t)s * a.payload_words] = HIT ? K : 100u + K;
}

}  // namespace

// Host-side handle of stage k: 0-2 debug ray gen / closest hit / miss,
// 3-5 chain ray gen / closest hit / miss, 6 sbt ray gen, 7-10 sbt closest
// hit 0-3, 11-12 sbt miss 0-1; null past the end.
extern "C" const void *test_stage(uint32_t k) {
  switch (k) {
    case 6: return reinterpret_cast<const void *>(&sbt_ray_gen);
    case 7: return reinterpret_cast<const void *>(&sbt_stage<0, true>);
    case 8: return reinterpret_cast<const void *>(&sbt_stage<1, true>);
    case 9: return reinterpret_cast<const void *>(&sbt_stage<2, true>);
    case 10: return reinterpret_cast<const void *>(&sbt_stage<3, true>);
    case 11: return reinterpret_cast<const void *>(&sbt_stage<0, false>);
    case 12: return reinterpret_cast<const void *>(&sbt_stage<1, false>);
    case 0: return reinterpret_cast<const void *>(&debug_ray_gen);
    case 1: return reinterpret_cast<const void *>(&debug_closest_hit);
    case 2: return reinterpret_cast<const void *>(&debug_miss);
    case 3: return reinterpret_cast<const void *>(&chain_ray_gen);
    case 4: return reinterpret_cast<const void *>(&chain_closest_hit);
    case 5: return reinterpret_cast<const void *>(&chain_miss);
    default: return nullptr;
  }
}

// ---- an exportable allocation for the external-buffer test ----
// Device memory created with HIP's virtual memory API, exportable as a
// POSIX file descriptor (what a Vulkan device hands out for an exported
// buffer), mapped at *ptr for the test to fill and read.  Returns 0, 1 when
// the device has no virtual memory management, or -1 on an error.
extern "C" int test_exportable_create(size_t bytes, int *fd, void **ptr, size_t *mapped, unsigned long long *handle) {
  int vmm = 0;
  if (hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, 0) != hipSuccess || !vmm)
    return 1;
  hipMemAllocationProp prop = {};
  prop.type = hipMemAllocationTypePinned;
  prop.requestedHandleType = hipMemHandleTypePosixFileDescriptor;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = 0;
  size_t gran = 0;
  if (hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum) != hipSuccess || !gran)
    return -1;
  const size_t size = (bytes + gran - 1) / gran * gran;
  hipMemGenericAllocationHandle_t h;
  if (hipMemCreate(&h, size, &prop, 0) != hipSuccess) return -1;
  int f = -1;
  if (hipMemExportToShareableHandle(&f, h, hipMemHandleTypePosixFileDescriptor, 0) != hipSuccess) {
    (void)hipMemRelease(h);
    return -1;
  }
  void *p = nullptr;
  if (hipMemAddressReserve(&p, size, 0, nullptr, 0) != hipSuccess) {
    (void)hipMemRelease(h);
    return -1;
  }
  if (hipMemMap(p, size, 0, h, 0) != hipSuccess) {
    (void)hipMemAddressFree(p, size);
    (void)hipMemRelease(h);
    return -1;
  }
  hipMemAccessDesc acc = {};
  acc.location.type = hipMemLocationTypeDevice;
  acc.location.id = 0;
  acc.flags = hipMemAccessFlagsProtReadWrite;
  if (hipMemSetAccess(p, size, &acc, 1) != hipSuccess) {
    (void)hipMemUnmap(p, size);
    (void)hipMemAddressFree(p, size);
    (void)hipMemRelease(h);
    return -1;
  }
  *fd = f;
  *ptr = p;
  *mapped = size;
  *handle = reinterpret_cast<unsigned long long>(h);
  return 0;
}

extern "C" int test_exportable_destroy(void *ptr, size_t size, unsigned long long handle) {
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemUnmap(ptr, size);
  if (e == hipSuccess) e = hipMemAddressFree(ptr, size);
  if (e == hipSuccess) e = hipMemRelease(reinterpret_cast<hipMemGenericAllocationHandle_t>(handle));
  return e == hipSuccess ? 0 : -1;
}

extern "C" int test_copy(void *dst, const void *src, size_t bytes, int kind) {
  return hipMemcpy(dst, src, bytes, kind ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}
