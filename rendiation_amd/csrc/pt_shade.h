// pt_shade.h -- per-path device code of the path tracer shared by the
// wavefront kernels (pt_kernels.hip) and the tail kernels (pt_tail_*.hip):
// camera-to-pixel mapping, ray records, and the closest-hit / miss stage of
// one path (ray_hit.rs:6-104, ray_miss.rs:136-158).
#ifndef RDRT_PT_SHADE_H
#define RDRT_PT_SHADE_H

#include "det_sum.h"
#include "pt.h"
#include "pt_device.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace pts {

using pt::V3;
using pt::v3;

constexpr uint32_t kMiss = 0xffffffffu;
constexpr float kFltEps = 1.1920928955078125e-7f;
constexpr float kFltMax = 3.402823466e+38f;
constexpr float kInvPi = 0.318309886183790671538f;

__device__ __forceinline__ V3 load3(const float *p) { return v3(p[0], p[1], p[2]); }
__device__ __forceinline__ void store3(float *p, V3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

// world normal = transpose(inverse 3x3) * n, from the column-major inverse
__device__ __forceinline__ V3 normal_to_world(const float *inv, V3 n) {
  return v3(inv[0] * n.x + inv[1] * n.y + inv[2] * n.z, inv[4] * n.x + inv[5] * n.y + inv[6] * n.z,
            inv[8] * n.x + inv[9] * n.y + inv[10] * n.z);
}

// path -> view, pixel (px, py) within the view, sample index n, and the
// view's sample seed (seed + view)
// Image pixel (views one after the other, each row-major) of pixel slot
// gpix = path / spp.  When both sides are multiples of 8, consecutive slots
// walk 8 x 8 tiles in row-major tile order, so the paths of one wave start
// in a square of the image instead of a 64-pixel strip of one row (more
// coherent rays); otherwise slots are row-major.
__device__ __forceinline__ uint32_t image_pixel(const PtLaunch &P, uint32_t gpix, uint32_t &px, uint32_t &py,
                                                uint32_t &view) {
  const uint32_t npix = P.width * P.height;
  view = gpix / npix;
  const uint32_t s = gpix - view * npix;
  if ((P.width & 7u) == 0u && (P.height & 7u) == 0u) {
    const uint32_t tile = s >> 6, w = s & 63u, ntx = P.width >> 3;
    px = (tile % ntx) * 8u + (w & 7u);
    py = (tile / ntx) * 8u + (w >> 3);
  } else {
    px = s % P.width;
    py = s / P.width;
  }
  return view * npix + py * P.width + px;
}
__device__ __forceinline__ uint32_t image_pixel(const PtLaunch &P, uint32_t gpix) {
  uint32_t px, py, view;
  return image_pixel(P, gpix, px, py, view);
}

__device__ __forceinline__ void pixel_of(const PtLaunch &P, uint32_t path, uint32_t &px, uint32_t &py, uint32_t &n,
                                         uint32_t &view, uint32_t &seed) {
  image_pixel(P, path / P.spp, px, py, view);
  n = P.sample_base + path % P.spp;
  seed = P.seed + view;
}

__device__ __forceinline__ void write_ray(rdrt_ray *r, V3 o, float tmin, V3 d, float tmax) {
  float4 *q = reinterpret_cast<float4 *>(r);
  q[0] = make_float4(o.x, o.y, o.z, tmin);
  q[1] = make_float4(d.x, d.y, d.z, tmax);
}

// Radiance beyond vertex `first` of a path, divided by the throughput at
// that vertex, from its recorded vertices: the sky or nothing at the end,
// e + a (w sun + rest) at each material vertex.
__device__ inline V3 recorded_suffix(const PtLaunch &P, uint32_t path, uint32_t first) {
  const V3 sun = load3(P.sun_radiance);
  V3 T = v3(0.0f, 0.0f, 0.0f), prod = v3(1.0f, 1.0f, 1.0f);
  for (uint32_t j = first; j < P.max_bounces + 2; j++) {
    const uint2 r = P.record[(size_t)j * P.npaths + path];
    if (r.x >= kVertSkyUp) {
      if (r.x != kVertEnd) T = T + prod * load3(r.x == kVertSkyUp ? P.sky_up : P.sky_down);
      break;
    }
    const Material &m = P.materials[r.x < P.nmaterials ? r.x : P.nmaterials - 1];
    const V3 a = v3(m.albedo[0], m.albedo[1], m.albedo[2]);
    T = T + prod * (v3(m.emission[0], m.emission[1], m.emission[2]) + a * sun * __uint_as_float(r.y));
    prod = prod * a;
  }
  return T;
}

// Gradient terms of one replayed (path, bounce) w.r.t. the three vertices
// of the hit triangle.  kind 1 (interpolated normals): normal terms A * w_k
// and position terms B * -w_k for barycentric weights w = (1 - u - v, u,
// v); kind 2 (geometric normal n = e1 x e2): position terms -(A + B), A, B;
// kind 0: none.
struct GeoTerm {
  uint32_t i0, i1, i2, kind;
  V3 A, B;
  float bu, bv;
};

// The per-corner vectors of a term: pos[0..2] (and nrm[0..2] for kind 1).
__device__ __forceinline__ void geo_corners(const GeoTerm &t, V3 pos[3], V3 nrm[3]) {
  if (t.kind == 1u) {
    const float bw = 1.0f - t.bu - t.bv;
    nrm[0] = t.A * bw;
    nrm[1] = t.A * t.bu;
    nrm[2] = t.A * t.bv;
    pos[0] = t.B * -bw;
    pos[1] = t.B * -t.bu;
    pos[2] = t.B * -t.bv;
  } else {
    pos[0] = (t.A + t.B) * -1.0f;
    pos[1] = t.A;
    pos[2] = t.B;
    nrm[0] = nrm[1] = nrm[2] = v3(0.0f, 0.0f, 0.0f);
  }
}

__device__ __forceinline__ uint32_t max_bits3(V3 v) {
  return max(det::abs_bits(v.x), max(det::abs_bits(v.y), det::abs_bits(v.z)));
}

// closest hit + miss for one path (ray_hit.rs:6-104, ray_miss.rs:136-158).
// REPLAY: the same path again after a recorded frame (bit-identical rays,
// hits and samples), adding the gradient of sum(adj * frame) w.r.t. vertex
// positions and normals at this vertex instead of radiance and shadow rays.
// With the path held fixed, the shading normal ns = sigma normalize(R n)
// (R = normal_to_world, n the interpolated object normal or e1 x e2) enters
// the light term beta a sun (ns . wi) / pi and the continuation factor
// a (ns . wo) / cos_wo; (u, v) are where the fixed object ray meets the
// triangle.
template <bool REPLAY>
__device__ inline void shade_one(const PtLaunch &P, uint32_t b, uint32_t path, bool &keep, bool &shadow, GeoTerm &gt) {
  const uint4 h = *reinterpret_cast<const uint4 *>(P.hits + path);
  const float4 *rq = reinterpret_cast<const float4 *>(P.rays + path);
  const float4 r0 = rq[0], r1 = rq[1];
  const V3 o = v3(r0.x, r0.y, r0.z), d = v3(r1.x, r1.y, r1.z);
  V3 beta = load3(P.beta + 3 * (size_t)path);
  V3 L = REPLAY ? v3(0.0f, 0.0f, 0.0f) : load3(P.L + 3 * (size_t)path);
  uint2 *rec = (!REPLAY && P.record) ? P.record + (size_t)b * P.npaths + path : nullptr;
  if (h.x == kMiss) {
    if (REPLAY) return;
    const float *sky = d.y > 0.0f ? P.sky_up : P.sky_down;
    L = L + beta * load3(sky);
    store3(P.L + 3 * (size_t)path, L);
    if (rec) *rec = make_uint2(d.y > 0.0f ? kVertSkyUp : kVertSkyDown, 0u);
    return;
  }
  const float t = __uint_as_float(h.y), bu = __uint_as_float(h.z), bv = __uint_as_float(h.w);
  const uint2 hx = P.hitx[path];
  const GeoRec &g = P.geos[hx.x];
  const InstRec &in = P.insts[hx.y];
  const uint32_t tri = g.primitive_start + h.x;
  const uint32_t i0 = P.indices[3 * (size_t)tri], i1 = P.indices[3 * (size_t)tri + 1],
                 i2 = P.indices[3 * (size_t)tri + 2];
  const V3 p0 = load3(P.vertices + 3 * (size_t)i0), p1 = load3(P.vertices + 3 * (size_t)i1),
           p2 = load3(P.vertices + 3 * (size_t)i2);
  const V3 e1 = p1 - p0, e2 = p2 - p0;
  const V3 m_geo = normal_to_world(in.inv, pt::cross3(e1, e2));
  V3 ng = pt::normalize3(m_geo);
  V3 ns = ng, m_ns = m_geo;
  bool interp = false;
  if (P.normals) {
    const float bw = 1.0f - bu - bv;
    V3 n_obj = load3(P.normals + 3 * (size_t)i0) * bw + load3(P.normals + 3 * (size_t)i1) * bu +
               load3(P.normals + 3 * (size_t)i2) * bv;
    if (pt::dot3(n_obj, n_obj) > 0.0f) {
      m_ns = normal_to_world(in.inv, n_obj);
      ns = pt::normalize3(m_ns);
      interp = true;
    }
  }
  // face the incoming ray (bindless_mesh_bridge.rs:86-123)
  if (pt::dot3(ng, d) > 0.0f) ng = ng * -1.0f;
  float sigma = 1.0f;
  if (pt::dot3(ns, ng) < 0.0f) {
    ns = ns * -1.0f;
    sigma = -1.0f;
  }
  const V3 pos = o + d * t;
  uint32_t mi = in.custom_index + (P.tri_material ? P.tri_material[tri] : 0u);
  if (mi >= P.nmaterials) mi = P.nmaterials - 1;
  const Material &m = P.materials[mi];
  const V3 albedo = v3(m.albedo[0], m.albedo[1], m.albedo[2]);
  L = L + beta * v3(m.emission[0], m.emission[1], m.emission[2]);
  float w_light = 0.0f;
  if (b < P.max_bounces) {
    uint32_t px, py, n, view, seed;
    pixel_of(P, path, px, py, n, view, seed);
    const V3 origin = pt::offset_ray_hit(pos, ng);
    // next-event estimation toward the directional light
    const V3 wi = v3(-P.sun_dir[0], -P.sun_dir[1], -P.sun_dir[2]);
    const float cos_l = pt::dot3(ns, wi);
    if (!REPLAY && cos_l > 0.0f && pt::dot3(ng, wi) > 0.0f &&
        (P.sun_radiance[0] > 0.0f || P.sun_radiance[1] > 0.0f || P.sun_radiance[2] > 0.0f)) {
      const V3 c = beta * albedo * (cos_l * kInvPi) * load3(P.sun_radiance);
      store3(P.nee + 3 * (size_t)path, c);
      write_ray(P.shadow_rays + path, origin, kFltEps, wi, kFltMax);
      shadow = true;
      w_light = cos_l * kInvPi;
    }
    // Lambertian importance sample: f cos / pdf = albedo
    const float s0 = pt::van_der_corput(n, pt::scramble(px, py, seed, 2 + 2 * b));
    const float s1 = pt::sobol1(n, pt::scramble(px, py, seed, 3 + 2 * b));
    const V3 wo = pt::sample_cosine(ns, s0, s1);
    bool alive = pt::dot3(wo, ng) > 0.0f;
    V3 nb = beta * albedo;
    if (alive && b >= P.rr_start) {
      const float q = fminf(fmaxf(fmaxf(nb.x, nb.y), nb.z), 0.95f);
      const float u = pt::van_der_corput(n, pt::scramble(px, py, seed, 64 + b));
      if (u >= q) alive = false;
      else nb = v3(nb.x / q, nb.y / q, nb.z / q);
    }
    if (alive) {
      write_ray(P.rays + path, origin, kFltEps, wo, kFltMax);
      store3(P.beta + 3 * (size_t)path, nb);
      keep = true;
    }
    if (REPLAY) {
      const float w = __uint_as_float(P.record[(size_t)b * P.npaths + path].y);
      const V3 a_pix = load3(P.adj + 3 * (size_t)image_pixel(P, path / P.spp)) * (1.0f / (float)P.spp);
      V3 g_ns = v3(0.0f, 0.0f, 0.0f);
      if (w != 0.0f) g_ns = wi * (pt::dot3(a_pix, beta * albedo * load3(P.sun_radiance)) * kInvPi);
      const float cos_wo = pt::dot3(ns, wo);
      if (alive && cos_wo > 1e-6f) {
        const V3 rem = nb * recorded_suffix(P, path, b + 1);
        g_ns = g_ns + wo * (pt::dot3(a_pix, rem) / cos_wo);
      }
      if (g_ns.x != 0.0f || g_ns.y != 0.0f || g_ns.z != 0.0f) {
        const V3 mv = interp ? m_ns : m_geo;
        const float inv_len = 1.0f / sqrtf(pt::dot3(mv, mv));
        const V3 g_m = (g_ns - ns * pt::dot3(ns, g_ns)) * (sigma * inv_len);
        const float *R = in.inv;
        const V3 g_n = v3(R[0] * g_m.x + R[4] * g_m.y + R[8] * g_m.z, R[1] * g_m.x + R[5] * g_m.y + R[9] * g_m.z,
                          R[2] * g_m.x + R[6] * g_m.y + R[10] * g_m.z);
        gt.i0 = i0;
        gt.i1 = i1;
        gt.i2 = i2;
        if (interp) {
          gt.kind = 1u;
          gt.A = g_n;
          gt.bu = bu;
          gt.bv = bv;
          gt.B = v3(0.0f, 0.0f, 0.0f);
          const V3 n0 = load3(P.normals + 3 * (size_t)i0);
          const float g_u = pt::dot3(g_n, load3(P.normals + 3 * (size_t)i1) - n0);
          const float g_v = pt::dot3(g_n, load3(P.normals + 3 * (size_t)i2) - n0);
          // (t, u, v) solves [-bd e1 e2] x = bo - p0 for the object-space ray
          const V3 c0 = v3(-(R[0] * d.x + R[4] * d.y + R[8] * d.z), -(R[1] * d.x + R[5] * d.y + R[9] * d.z),
                           -(R[2] * d.x + R[6] * d.y + R[10] * d.z));
          const float det = pt::dot3(c0, pt::cross3(e1, e2));
          if (det != 0.0f) gt.B = (pt::cross3(e2, c0) * g_u + pt::cross3(c0, e1) * g_v) * (1.0f / det);
        } else {
          // n = e1 x e2
          gt.kind = 2u;
          gt.A = pt::cross3(e2, g_n);
          gt.B = pt::cross3(g_n, e1);
          gt.bu = 0.0f;
          gt.bv = 0.0f;
        }
      }
    }
  }
  if (REPLAY) return;
  store3(P.L + 3 * (size_t)path, L);
  if (rec) {
    // the light weight stands until the shadow ray finds an occluder
    *rec = make_uint2(mi, __float_as_uint(w_light));
    if (!keep) rec[P.npaths] = make_uint2(kVertEnd, 0u);
  }
}

}  // namespace pts
}  // namespace rdrt

#endif
