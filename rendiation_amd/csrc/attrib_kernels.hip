// attrib_kernels.hip -- shading attributes of closest hits: the reference's
// bindless mesh fetch (scene/rendering/gpu-ray-tracing/src/
// bindless_mesh_bridge.rs:43-123) and its shader-binding-table hit group
// (shader/ray-tracing/src/api/ctx.rs:53-55), and the rest of the closest-hit
// context (HitCtxInfo, api/ctx.rs:24-48, as traverse_gpu.rs:575-582,661-683
// fills it): object-to-world and world-to-object transforms and the
// object-space ray; one thread per hit.
#include "attrib.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace {

struct A3 {
  float x, y, z;
};
__device__ __forceinline__ A3 a3(const float *p) { return A3{p[0], p[1], p[2]}; }
__device__ __forceinline__ A3 sub3(A3 a, A3 b) { return A3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ A3 scale3(A3 a, float s) { return A3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ A3 add3(A3 a, A3 b) { return A3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ float dot3(A3 a, A3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ A3 cross3(A3 a, A3 b) {
  return A3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ A3 normalize3(A3 a) {  // vec/dimension.rs:98-105 (zero stays zero)
  const float m = dot3(a, a);
  if (m > 0.0f) return scale3(a, 1.0f / sqrtf(m));
  return a;
}
// The normal matrix: the transpose of the instance's world-to-object 3x3
// (bindless_mesh_bridge.rs:95), applied to an object-space vector.
__device__ __forceinline__ A3 normal_mat(const float *inv, A3 n) {
  return A3{inv[0] * n.x + inv[1] * n.y + inv[2] * n.z, inv[4] * n.x + inv[5] * n.y + inv[6] * n.z,
            inv[8] * n.x + inv[9] * n.y + inv[10] * n.z};
}

__global__ void k_hit_attributes(AttribArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n || (a.queue && t >= *a.count)) return;
  const uint32_t i = a.queue ? a.queue[t] : t;
  rdrt_hit_attributes o;
  for (int k = 0; k < 3; k++) o.position[k] = o.shading_normal[k] = o.geometric_normal[k] = 0.0f;
  o.uv[0] = o.uv[1] = 0.0f;
  o.sbt_record = RDRT_MISS;
  const rdrt_hit h = a.hits[i];
  const rdrt_hit_meta hm = a.hit_meta[i];
  const uint32_t tlas_idx = a.meta ? a.meta[i].tlas_idx : a.tlas_idx;
  bool ok = h.prim_id != RDRT_MISS && tlas_idx < a.nbinding;
  uint32_t inst = 0, geo = 0;
  if (ok) {
    const uint32_t th = a.binding[tlas_idx];
    ok = th < a.ntlas;
    if (ok) {
      const uint32_t slot = a.tlas_inst_first[th] + hm.instance_id;
      ok = slot < a.ninst;
      if (ok) inst = a.inst_lookup[slot];
    }
  }
  if (ok) {
    geo = a.insts[inst].geo_begin + hm.geometry_id;
    ok = geo < a.insts[inst].geo_end;
  }
  if (!ok) {
    if (a.out) a.out[i] = o;
    if (a.xf_out) {
      rdrt_hit_transform z;
      float *zw = reinterpret_cast<float *>(&z);
      for (int k = 0; k < 40; k++) zw[k] = 0.0f;
      a.xf_out[i] = z;
    }
    return;
  }
  const InstRec &in = a.insts[inst];
  const GeoRec &g = a.geos[geo];
  const size_t tri = (size_t)g.primitive_start + h.prim_id;
  const uint32_t ia = a.indices[3 * tri], ib = a.indices[3 * tri + 1], ic = a.indices[3 * tri + 2];
  const float u = h.u, v = h.v;
  const float w = 1.0f - u - v;  // barycentric of the first vertex (bindless_mesh_bridge.rs:52-57)
  const rdrt_ray r = a.rays[i];
  const A3 org = A3{r.ox, r.oy, r.oz}, dir = A3{r.dx, r.dy, r.dz};
  const A3 pos = add3(org, scale3(dir, h.t));
  // geometric normal, facing the ray origin (bindless_mesh_bridge.rs:104-114)
  const A3 pa = a3(a.vertices + 3 * (size_t)ia), pb = a3(a.vertices + 3 * (size_t)ib),
           pc = a3(a.vertices + 3 * (size_t)ic);
  A3 gn = normalize3(normal_mat(in.inv, cross3(sub3(pa, pb), sub3(pa, pc))));
  if (dot3(sub3(org, pos), gn) < 0.0f) gn = scale3(gn, -1.0f);
  // shading normal from the vertex normals, on the geometric side (:93-120)
  A3 sn = gn;
  if (a.normals) {
    const A3 na = a3(a.normals + 3 * (size_t)ia), nb = a3(a.normals + 3 * (size_t)ib),
             nc = a3(a.normals + 3 * (size_t)ic);
    const A3 n = add3(add3(scale3(na, w), scale3(nb, u)), scale3(nc, v));
    if (dot3(n, n) > 0.0f) {  // zero vertex normals: none given for this geometry
      sn = normalize3(normal_mat(in.inv, n));
      if (dot3(gn, sn) < 0.0f) sn = scale3(sn, -1.0f);
    }
  }
  if (a.uvs) {
    const float *ua = a.uvs + 2 * (size_t)ia, *ub = a.uvs + 2 * (size_t)ib, *uc = a.uvs + 2 * (size_t)ic;
    for (int k = 0; k < 2; k++) o.uv[k] = ua[k] * w + ub[k] * u + uc[k] * v;
  }
  o.position[0] = pos.x;
  o.position[1] = pos.y;
  o.position[2] = pos.z;
  o.shading_normal[0] = sn.x;
  o.shading_normal[1] = sn.y;
  o.shading_normal[2] = sn.z;
  o.geometric_normal[0] = gn.x;
  o.geometric_normal[1] = gn.y;
  o.geometric_normal[2] = gn.z;
  const uint32_t sbt_offset = a.sbt ? a.sbt[i].x : a.sbt_offset, sbt_stride = a.sbt ? a.sbt[i].y : a.sbt_stride;
  o.sbt_record = sbt_offset + sbt_stride * hm.geometry_id + in.sbt_offset;  // ctx.rs:53-55
  if (a.out) a.out[i] = o;
  if (a.xf_out) {
    // the object-space ray the traversal used for this instance
    // (traverse_gpu.rs:427-431): inverse * (o, 1) with the w divide, and the
    // inverse's 3x3 times d, normalized
    rdrt_hit_transform x;
    const float *m = in.inv, *f = a.inst_fwd + 16 * (size_t)inst;
    for (int k = 0; k < 16; k++) {
      x.object_to_world[k] = f[k];
      x.world_to_object[k] = m[k];
    }
    const float ox = org.x * m[0] + org.y * m[4] + org.z * m[8] + 1.0f * m[12];
    const float oy = org.x * m[1] + org.y * m[5] + org.z * m[9] + 1.0f * m[13];
    const float oz = org.x * m[2] + org.y * m[6] + org.z * m[10] + 1.0f * m[14];
    const float ow = org.x * m[3] + org.y * m[7] + org.z * m[11] + 1.0f * m[15];
    const A3 bd = normalize3(A3{dir.x * m[0] + dir.y * m[4] + dir.z * m[8], dir.x * m[1] + dir.y * m[5] + dir.z * m[9],
                                dir.x * m[2] + dir.y * m[6] + dir.z * m[10]});
    x.object_ray_origin[0] = ox / ow;
    x.object_ray_origin[1] = oy / ow;
    x.object_ray_origin[2] = oz / ow;
    x.object_ray_direction[0] = bd.x;
    x.object_ray_direction[1] = bd.y;
    x.object_ray_direction[2] = bd.z;
    x.instance_sbt_offset = in.sbt_offset;
    x.pad = 0u;
    a.xf_out[i] = x;
  }
}

}  // namespace

hipError_t enqueue_hit_attributes(const AttribArgs &a, hipStream_t st) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hit_attributes, dim3((a.n + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace rdrt
