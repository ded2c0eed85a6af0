// trace_device.h -- device code of the closest-hit traversal kernels for
// gfx950 (wave64), shared by the translation units that instantiate them.
//
// Two kernels share one result definition, that of the reference CPU
// traversal (shader/ray-tracing/src/backend/wavefront_compute/geometry/naive/
// traverse_cpu.rs:52-245): the last accepted triangle wins, a triangle is
// accepted when its t lies in [t_min, far] of the ray range in object space
// (wavefront_compute/geometry/mod.rs:104-155), and far shrinks to each accepted
// hit.
//
//  * k_trace_ref walks the reference's 48-byte forest by its stackless
//    hit/miss links (traverse_cpu.rs:283-319) in the reference order, with the
//    reference's per-test arithmetic; its hits and visit counters equal the
//    oracle's bit for bit.
//  * k_trace_fast walks the 64-byte child-pair layout front to back with a
//    per-lane stack in LDS (column layout [entry][lane]: each lane owns one
//    bank at every depth, so stack traffic is conflict-free) and tests
//    triangles from 64-byte pre-pass records.  The box and triangle
//    arithmetic is the reference's, so t, u, v are bit-identical; only the
//    visit order differs, which can change the winner among hits with equal t.
//
// Compiled with -ffp-contract=off and IEEE division/sqrt: one rounding per
// operation, as in the reference.
#ifndef RDRT_TRACE_DEVICE_H
#define RDRT_TRACE_DEVICE_H

#include <algorithm>

#include "trace_kernels.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace {

constexpr uint32_t F_FORCE_OPAQUE = 0x01, F_FORCE_NON_OPAQUE = 0x02, F_END_SEARCH = 0x04, F_CULL_BACK = 0x10,
                   F_CULL_FRONT = 0x20, F_CULL_OPAQUE = 0x40, F_CULL_NON_OPAQUE = 0x80,
                   F_SKIP_TRIANGLES = 0x100, F_FLIP = 0x400;
constexpr uint32_t GI_CULL_DISABLE = 0x1, GI_FLIP = 0x2, GI_FORCE_OPAQUE = 0x4, GI_FORCE_NO_OPAQUE = 0x8;

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ F3 normalize(F3 a) {
  float m = dot(a, a);
  if (m > 0.0f) {
    float inv = 1.0f / sqrtf(m);
    return f3(a.x * inv, a.y * inv, a.z * inv);
  }
  return a;
}
__device__ __forceinline__ F3 recip(F3 a) { return f3(1.0f / a.x, 1.0f / a.y, 1.0f / a.z); }

// Ray/AABB slab test (wavefront_compute/geometry/mod.rs:44-64).  inv_d is
// 1/dir, hoisted out of the node loop; the per-node values are unchanged.
__device__ __forceinline__ bool slab(F3 o, F3 inv_d, float tmin, float tmax, float mnx, float mny, float mnz,
                                     float mxx, float mxy, float mxz, float &tnear) {
  float t0x = (mnx - o.x) * inv_d.x, t0y = (mny - o.y) * inv_d.y, t0z = (mnz - o.z) * inv_d.z;
  float t1x = (mxx - o.x) * inv_d.x, t1y = (mxy - o.y) * inv_d.y, t1z = (mxz - o.z) * inv_d.z;
  float nx = fminf(t0x, t1x), ny = fminf(t0y, t1y), nz = fminf(t0z, t1z);
  float fx = fmaxf(t0x, t1x), fy = fmaxf(t0y, t1y), fz = fmaxf(t0z, t1z);
  float tn = fmaxf(fmaxf(nx, ny), nz);
  float tf = fminf(fminf(fx, fy), fz);
  tnear = tn;
  return tn <= tf && tmin < tf && tn < tmax;
}

// Triangle test given the vertex-only quantities (geometry/mod.rs:118-154).
// b == 0 is a miss.
__device__ __forceinline__ bool tri_core(F3 o, F3 d, float rn, float rf, F3 v0, F3 e1, F3 e2, F3 n, float uu,
                                         float uv, float vv, float inverse_d, bool cull_enable, bool cull_back,
                                         float &t, float &u, float &v, bool &back) {
  float b = dot(n, d);
  if (b == 0.0f) return false;
  if (cull_enable && cull_back == (b < 0.0f)) return false;
  F3 w0 = sub(o, v0);
  float a = -dot(n, w0);
  float tt = a / b;
  if (tt < rn || tt > rf) return false;
  F3 p = f3(o.x + tt * d.x, o.y + tt * d.y, o.z + tt * d.z);
  F3 w = sub(p, v0);
  float wu = dot(w, e1);
  float wv = dot(w, e2);
  float uu_ = (uv * wv - vv * wu) * inverse_d;
  if (uu_ < 0.0f || uu_ > 1.0f) return false;
  float vv_ = (uv * wu - uu * wv) * inverse_d;
  if (vv_ < 0.0f || (uu_ + vv_) > 1.0f) return false;
  t = tt;
  u = uu_;
  v = vv_;
  back = b < 0.0f;
  return true;
}

// flag.rs:30-75
__device__ __forceinline__ uint32_t merge_flags(uint32_t f, uint32_t gi) {
  if (gi & GI_CULL_DISABLE) f &= ~(F_CULL_BACK | F_CULL_FRONT);
  if (gi & GI_FLIP) f ^= F_FLIP;
  if (gi & GI_FORCE_OPAQUE) f |= F_FORCE_OPAQUE;
  if (gi & GI_FORCE_NO_OPAQUE) f |= F_FORCE_NON_OPAQUE;
  return f;
}
// flag.rs:78-91.  Opaque for this ray: the geometry's flag, forced either way by the merged
// ray / instance flags (flag.rs cull_geometry).
__device__ __forceinline__ bool geometry_opaque(uint32_t f, uint32_t gflags) {
  return ((gflags & RDRT_GEOMETRY_FLAG_OPAQUE) || (f & F_FORCE_OPAQUE)) && !(f & F_FORCE_NON_OPAQUE);
}
// Any-hit behaviour of a hit on global triangle `tri` (traverse_cpu.rs:
// 164-185): an opaque hit commits; a non-opaque one takes the caller's
// per-triangle behaviour, or is accepted without a table.
__device__ __forceinline__ uint32_t anyhit_behavior(const TraceParams &p, bool opaque, uint32_t tri) {
  return (opaque || !p.anyhit) ? RDRT_ANYHIT_ACCEPT : (uint32_t)p.anyhit[tri];
}
__device__ __forceinline__ bool geometry_passes(uint32_t f, uint32_t gflags) {
  bool opaque = geometry_opaque(f, gflags);
  return (opaque && !(f & F_CULL_OPAQUE)) || (!opaque && !(f & F_CULL_NON_OPAQUE));
}
// flag.rs:94-105
__device__ __forceinline__ void cull_triangle(uint32_t f, bool &enable, bool &back) {
  bool flip = f & F_FLIP, cf = f & F_CULL_FRONT, cb = f & F_CULL_BACK;
  enable = cf || cb;
  back = (flip && cb) || (!flip && cf);
}

// Object-space ray of an instance (traverse_cpu.rs:99-104): inverse transform
// of the origin with the w divide, 3x3 part on the direction, scaling = its
// length, direction normalized.
__device__ __forceinline__ void object_ray(const float *m, F3 o, F3 d, F3 &bo, F3 &bd, float &scaling) {
  float x = o.x * m[0] + o.y * m[4] + o.z * m[8] + 1.0f * m[12];
  float y = o.x * m[1] + o.y * m[5] + o.z * m[9] + 1.0f * m[13];
  float z = o.x * m[2] + o.y * m[6] + o.z * m[10] + 1.0f * m[14];
  float w = o.x * m[3] + o.y * m[7] + o.z * m[11] + 1.0f * m[15];
  bo = f3(x / w, y / w, z / w);
  F3 dd = f3(d.x * m[0] + d.y * m[4] + d.z * m[8], d.x * m[1] + d.y * m[5] + d.z * m[9],
             d.x * m[2] + d.y * m[6] + d.z * m[10]);
  scaling = sqrtf(dot(dd, dd));
  bd = normalize(dd);
}

struct RayIn {
  F3 o, d;
  float tmin, tmax;
  uint32_t flags, mask, tlas_idx;
};

__device__ __forceinline__ RayIn load_ray(const TraceParams &p, uint32_t rid) {
  const float4 *rp = reinterpret_cast<const float4 *>(p.rays + rid);
  float4 a = rp[0], b = rp[1];
  RayIn r;
  r.o = f3(a.x, a.y, a.z);
  r.tmin = a.w;
  r.d = f3(b.x, b.y, b.z);
  r.tmax = b.w;
  r.flags = p.flags;
  r.mask = p.cull_mask;
  r.tlas_idx = p.tlas_idx;
  if (p.meta) {
    uint4 m = *reinterpret_cast<const uint4 *>(p.meta + rid);
    r.flags = m.x;
    r.mask = m.y;
    r.tlas_idx = m.z;
  }
  return r;
}

// Root of the TLAS bound at tlas_idx, or null when unbound / deleted.
__device__ __forceinline__ const TlasRoot *bound_root(const TraceParams &p, uint32_t tlas_idx) {
  if (tlas_idx >= p.nbinding) return nullptr;
  uint32_t th = p.binding[tlas_idx];
  if (th >= p.ntlas) return nullptr;
  const TlasRoot *r = p.tlas_roots + th;
  return r->ref_root == kInvalid ? nullptr : r;
}

struct Best {
  uint32_t k = kInvalid;  // triangle slot (redirect index)
  uint32_t geo = 0, inst = 0;
  float t = 0.0f, u = 0.0f, v = 0.0f;
  bool back = false;
};

// ids maps best.k to the global triangle: p.tri_ids (fast) or p.redirect (reference).
__device__ __forceinline__ void write_result(const TraceParams &p, const uint32_t *ids, uint32_t rid,
                                             const Best &best, uint32_t cn, uint32_t ct, uint32_t ci) {
  uint4 h = make_uint4(RDRT_MISS, 0u, 0u, 0u);
  uint4 hm = make_uint4(RDRT_MISS, RDRT_MISS, 0u, RDRT_MISS);
  if (best.k != kInvalid) {
    const GeoRec &g = p.geos[best.geo];
    h.x = ids[best.k] - g.primitive_start;
    h.y = __float_as_uint(best.t);
    h.z = __float_as_uint(best.u);
    h.w = __float_as_uint(best.v);
    if (p.hit_meta) {
      const InstRec &in = p.insts[best.inst];
      hm = make_uint4(g.geometry_idx, in.order, best.back ? RDRT_HIT_KIND_BACK_FACING_TRIANGLE
                                                          : RDRT_HIT_KIND_FRONT_FACING_TRIANGLE,
                      in.custom_index);
    }
  }
  *reinterpret_cast<uint4 *>(p.hits + rid) = h;
  if (p.hit_meta) *reinterpret_cast<uint4 *>(p.hit_meta + rid) = hm;
  if (p.hitx) p.hitx[rid] = make_uint2(best.geo, best.inst);
  if (p.counters) {
    p.counters[3 * rid + 0] = cn;
    p.counters[3 * rid + 1] = ct;
    p.counters[3 * rid + 2] = ci;
  }
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap the first `used` physical blocks so each XCD works on one contiguous
// range of rays: neighbouring (coherent) rays then share an L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t i, uint32_t used) {
  const uint32_t xcd = i % 8u, local = i / 8u, per = used / 8u, rem = used % 8u;
  return xcd < rem ? xcd * (per + 1u) + local : rem * (per + 1u) + (xcd - rem) * per + local;
}

// Inner-node steps of the fast traversal.  Each tests the children of node
// `ref`; when some are hit it sets ref to the nearest, pushes the others
// (farthest first, so they pop near to far) and returns true.
template <class Push>
__device__ __forceinline__ bool step_pair(const PairNode *nodes, uint32_t &ref, F3 o, F3 inv_d, float tmin, float tmax,
                                          Push &&push) {
  const float4 *q = reinterpret_cast<const float4 *>(nodes + ref);
  const float4 a = q[0], b = q[1], c = q[2], d = q[3];
  float tl, tr;
  const bool hl = slab(o, inv_d, tmin, tmax, a.x, a.y, a.z, b.x, b.y, b.z, tl);
  const bool hr = slab(o, inv_d, tmin, tmax, c.x, c.y, c.z, d.x, d.y, d.z, tr);
  const uint32_t lref = __float_as_uint(a.w), rref = __float_as_uint(c.w);
  if (hl && hr) {
    const bool rfirst = tr < tl;
    push(rfirst ? lref : rref);
    ref = rfirst ? rref : lref;
    return true;
  }
  if (hl || hr) {
    ref = hl ? lref : rref;
    return true;
  }
  return false;
}

// Sort key of one quad child: its entry distance when hit (never NaN, always
// below +inf: a hit's entry lies below tmax), +inf when missed or unused.
__device__ __forceinline__ float quad_key(bool hit, float tnear) {
  return hit ? fmaxf(tnear, -3.402823466e38f) : __builtin_inff();
}
__device__ __forceinline__ void order2(float &ka, uint32_t &ra, float &kb, uint32_t &rb) {
  const bool sw = kb < ka;
  const float k = sw ? kb : ka;
  const uint32_t r = sw ? rb : ra;
  kb = sw ? ka : kb;
  rb = sw ? ra : rb;
  ka = k;
  ra = r;
}

template <class Push>
__device__ __forceinline__ bool step_quad(const QuadNode *nodes, uint32_t &ref, F3 o, F3 inv_d, float tmin, float tmax,
                                          Push &&push) {
  const float4 *q = reinterpret_cast<const float4 *>(nodes + ref);
  const float4 x0 = q[0], x1 = q[1], y0 = q[2], y1 = q[3], z0 = q[4], z1 = q[5];
  const uint4 rf = *reinterpret_cast<const uint4 *>(q + 6);
  float t0, t1, t2, t3;
  const bool h0 = slab(o, inv_d, tmin, tmax, x0.x, y0.x, z0.x, x1.x, y1.x, z1.x, t0);
  const bool h1 = slab(o, inv_d, tmin, tmax, x0.y, y0.y, z0.y, x1.y, y1.y, z1.y, t1);  // slots 0, 1 always used
  const bool h2 = slab(o, inv_d, tmin, tmax, x0.z, y0.z, z0.z, x1.z, y1.z, z1.z, t2) && rf.z != kInvalid;
  const bool h3 = slab(o, inv_d, tmin, tmax, x0.w, y0.w, z0.w, x1.w, y1.w, z1.w, t3) && rf.w != kInvalid;
  const uint32_t n = (uint32_t)h0 + (uint32_t)h1 + (uint32_t)h2 + (uint32_t)h3;
  if (n == 0) return false;
  float k0 = quad_key(h0, t0), k1 = quad_key(h1, t1), k2 = quad_key(h2, t2), k3 = quad_key(h3, t3);
  uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
  // sorting network of four: hits ascend by entry distance, misses last
  order2(k0, r0, k1, r1);
  order2(k2, r2, k3, r3);
  order2(k0, r0, k2, r2);
  order2(k1, r1, k3, r3);
  order2(k1, r1, k2, r2);
  if (n > 3) push(r3);
  if (n > 2) push(r2);
  if (n > 1) push(r1);
  ref = r0;
  return true;
}

// Byte c of a quantized node word, decoded: corner + q * step (q * step is
// exact, the sum rounds once, as on the host that chose q).
__device__ __forceinline__ float dequant(float corner, uint32_t word, int c, float step) {
  return corner + (float)((word >> (8 * c)) & 0xffu) * step;
}

template <class Push>
__device__ __forceinline__ bool step_qquad(const QuadNodeQ *nodes, uint32_t &ref, F3 o, F3 inv_d, float tmin,
                                           float tmax, Push &&push) {
  const float4 *q = reinterpret_cast<const float4 *>(nodes + ref);
  const float4 a = q[0];                                  // corner, x step
  const uint4 rf = *reinterpret_cast<const uint4 *>(q + 1);
  const uint4 xy = *reinterpret_cast<const uint4 *>(q + 2);  // lox, hix, loy, hiy
  const float4 zs = q[3];                                 // loz, hiz (bits), y step, z step
  const uint32_t loz = __float_as_uint(zs.x), hiz = __float_as_uint(zs.y);
  float t[4];
  bool h[4];
#pragma unroll
  for (int c = 0; c < 4; c++)
    h[c] = slab(o, inv_d, tmin, tmax, dequant(a.x, xy.x, c, a.w), dequant(a.y, xy.z, c, zs.z),
                dequant(a.z, loz, c, zs.w), dequant(a.x, xy.y, c, a.w), dequant(a.y, xy.w, c, zs.z),
                dequant(a.z, hiz, c, zs.w), t[c]);
  h[2] = h[2] && rf.z != kInvalid;  // slots 0, 1 always used
  h[3] = h[3] && rf.w != kInvalid;
  const uint32_t n = (uint32_t)h[0] + (uint32_t)h[1] + (uint32_t)h[2] + (uint32_t)h[3];
  if (n == 0) return false;
  float k0 = quad_key(h[0], t[0]), k1 = quad_key(h[1], t[1]), k2 = quad_key(h[2], t[2]), k3 = quad_key(h[3], t[3]);
  uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
  order2(k0, r0, k1, r1);
  order2(k2, r2, k3, r3);
  order2(k0, r0, k2, r2);
  order2(k1, r1, k3, r3);
  order2(k1, r1, k2, r2);
  if (n > 3) push(r3);
  if (n > 2) push(r2);
  if (n > 1) push(r1);
  ref = r0;
  return true;
}

// ------------------------------------------------------------------------
// Fast kernel: front-to-back child-pair traversal.  Stack: the first K
// entries of each lane live in LDS (column layout), deeper ones (rare) in
// the lane's private memory; S bounds the total.
// ------------------------------------------------------------------------
template <int K, int S, int B, bool CNT, int LAYOUT>
__device__ __forceinline__ void trace_fast_one(const TraceParams &p, uint32_t rid, uint32_t *stk) {
  uint32_t spill[S > K ? S - K : 1];
  auto push = [&](uint32_t &sp, uint32_t v) {
    if (sp < (uint32_t)K) stk[sp * B] = v;
    else if (sp < (uint32_t)S) spill[sp - K] = v;
    else {
      atomicOr(p.status, 1u);  // unreachable when S covers the built depth
      return;
    }
    sp++;
  };
  auto pop = [&](uint32_t &sp) -> uint32_t {
    sp--;
    return sp < (uint32_t)K ? stk[sp * B] : spill[sp - K];
  };

  RayIn r = load_ray(p, rid);
  float far_ = r.tmax;
  Best best;
  uint32_t cn = 0, ct = 0, ci = 0;
  bool done = false;

  const TlasRoot *root = bound_root(p, r.tlas_idx);
  F3 winv = recip(r.d);
  float tn;
  if (root && slab(r.o, winv, r.tmin, far_, root->mn[0], root->mn[1], root->mn[2], root->mx[0], root->mx[1],
                   root->mx[2], tn)) {
    uint32_t sp = 0;
    uint32_t ref = LAYOUT ? p.tlas_quad_roots[root - p.tlas_roots] : root->root_ref;
    auto push_tlas = [&](uint32_t v) { push(sp, v); };
    for (;;) {
      if (!(ref & kLeafBit)) {
        if constexpr (LAYOUT == 2) {
          if (step_qquad(p.tlas_qquads, ref, r.o, winv, r.tmin, far_, push_tlas)) continue;
        } else if constexpr (LAYOUT == 1) {
          if (step_quad(p.tlas_quads, ref, r.o, winv, r.tmin, far_, push_tlas)) continue;
        } else {
          if (step_pair(p.tlas_pairs, ref, r.o, winv, r.tmin, far_, push_tlas)) continue;
        }
      } else {
        const uint32_t i0 = (ref & ~kLeafBit) >> kLeafCountBits, i1 = i0 + (ref & kLeafMaxCount);
        for (uint32_t i = i0; i < i1 && !done; i++) {
          if constexpr (CNT) ci++;
          const InstRec &in = p.insts[i];
          float ti;
          if (!slab(r.o, winv, r.tmin, far_, in.mn[0], in.mn[1], in.mn[2], in.mx[0], in.mx[1], in.mx[2], ti))
            continue;
          if ((r.mask & in.mask) == 0) continue;
          const uint32_t fl = merge_flags(r.flags, in.flags);
          if (fl & F_SKIP_TRIANGLES) continue;
          F3 bo, bd;
          float s;
          object_ray(in.inv, r.o, r.d, bo, bd, s);
          const F3 binv = recip(bd);
          const float near_s = r.tmin * s;
          bool ce, cb;
          cull_triangle(fl, ce, cb);
          for (uint32_t g = in.geo_begin; g < in.geo_end && !done; g++) {
            const GeoRec &geo = p.geos[g];
            if (!geometry_passes(fl, geo.flags)) continue;
            const bool opq = geometry_opaque(fl, geo.flags);
            float tg;
            if (!slab(bo, binv, near_s, far_ * s, geo.mn[0], geo.mn[1], geo.mn[2], geo.mx[0], geo.mx[1],
                      geo.mx[2], tg))
              continue;
            // BLAS traversal; its stack entries sit above the TLAS ones.
            uint32_t bsp = sp;
            uint32_t bref = LAYOUT ? geo.quad_root : geo.root_ref;
            auto push_blas = [&](uint32_t v) { push(bsp, v); };
            for (;;) {
              if (!(bref & kLeafBit)) {
                if constexpr (CNT) cn++;
                const float fs = far_ * s;
                if constexpr (LAYOUT == 2) {
                  if (step_qquad(p.tri_qquads, bref, bo, binv, near_s, fs, push_blas)) continue;
                } else if constexpr (LAYOUT == 1) {
                  if (step_quad(p.tri_quads, bref, bo, binv, near_s, fs, push_blas)) continue;
                } else {
                  if (step_pair(p.tri_pairs, bref, bo, binv, near_s, fs, push_blas)) continue;
                }
              } else {
                const uint32_t k0 = (bref & ~kLeafBit) >> kLeafCountBits, k1 = k0 + (bref & kLeafMaxCount);
                for (uint32_t k = k0; k < k1; k++) {
                  if constexpr (CNT) ct++;
                  const float4 *q = reinterpret_cast<const float4 *>(p.tris + k);
                  float4 a = q[0], b = q[1], c = q[2], d = q[3];
                  float t, u, v;
                  bool back;
                  if (tri_core(bo, bd, near_s, far_ * s, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), f3(c.x, c.y, c.z),
                               f3(d.x, d.y, d.z), a.w, b.w, c.w, d.w, ce, cb, t, u, v, back)) {
                    uint32_t beh = opq ? RDRT_ANYHIT_ACCEPT : anyhit_behavior(p, false, p.tri_ids[k]);
                    if (beh & RDRT_ANYHIT_ACCEPT) {
                      far_ = t / s;
                      best.k = k;
                      best.geo = g;
                      best.inst = i;
                      best.u = u;
                      best.v = v;
                      best.back = back;
                      if (fl & F_END_SEARCH) beh |= RDRT_ANYHIT_END_SEARCH;
                    }
                    if (beh & RDRT_ANYHIT_END_SEARCH) {
                      done = true;
                      break;
                    }
                  }
                }
                if (done) break;
              }
              if (bsp == sp) break;
              bref = pop(bsp);
            }
          }
          // re-read the world ray instead of keeping it live through the BLAS loop
          r = load_ray(p, rid);
          winv = recip(r.d);
        }
        if (done) break;
      }
      if (sp == 0) break;
      ref = pop(sp);
    }
  }
  best.t = far_;  // far shrinks to each accepted hit
  write_result(p, p.tri_ids, rid, best, cn, ct, ci);
}

// Live slots of the first and the second segment of a launch.
__device__ __forceinline__ uint32_t live_first(const TraceParams &p) {
  return p.count ? min(slots_of(p), *p.count) : slots_of(p);
}
__device__ __forceinline__ uint32_t live_second(const TraceParams &p) {
  return p.count2 ? min(p.slots2, *p.count2) : 0u;
}
// The launch parameters of the second segment.
__device__ __forceinline__ void select_second(TraceParams &p) {
  p.rays = p.rays2;
  p.hits = p.hits2;
  p.hitx = p.hitx2;
  p.queue = p.queue2;
  p.flags = p.flags2;
  p.counters = nullptr;
}

// ------------------------------------------------------------------------
// Cooperative kernel: kCoopLanes consecutive lanes trace one ray.  Every lane
// of the group walks the TLAS in lockstep (same nodes, same far bound: the
// group's nearest hit so far); the instances it accepts are dealt to the
// lanes in batches of kCoopLanes, each lane walks its instance's BLAS alone,
// and after each batch the group takes the nearest far bound.  A ray's chain
// of dependent node loads becomes the TLAS walk plus the longest BLAS walk of
// each batch, which is what bounds small launches (the tail bounces of a
// path tracer): they last as long as their slowest ray.  The result is the
// nearest accepted hit, as in the one-ray-per-lane kernel; among hits with
// equal t the lowest lane of the last batch wins (a visit-order tie, as
// between the layouts).  Launches of at most kCoopWideMax rays give each ray
// 16 lanes: their time is their slowest rays' chains, and a batch of 16
// instances takes a ray over twice as many instances per BLAS round (the
// last bounces of the 1080p path-tracing frame: 2.88 -> 2.83 ms with every
// launch of at most 32k rays cooperative); larger launches keep 8, where
// the extra lanes cost more occupancy than they save.
// ------------------------------------------------------------------------
constexpr uint32_t kCoopLanes = 8;
constexpr uint32_t kCoopWideLanes = 16;
constexpr uint32_t kCoopWideMax = 32768;

template <int G>
__device__ __forceinline__ float group_min(float v) {
  for (int m = G / 2; m > 0; m >>= 1) v = fminf(v, __shfl_xor(v, m, G));
  return v;
}
template <int G>
__device__ __forceinline__ bool group_any(bool b) {
  int v = b ? 1 : 0;
  for (int m = G / 2; m > 0; m >>= 1) v |= __shfl_xor(v, m, G);
  return v != 0;
}

// Returns (group-uniform) the lane of the group that wrote the record, plus
// kCoopHit when the ray hit.
constexpr uint32_t kCoopHit = 0x100u;

template <int K, int S, int B, int LAYOUT, int G = (int)kCoopLanes>
__device__ __forceinline__ uint32_t trace_coop_one(const TraceParams &p, uint32_t rid, uint32_t *stk,
                                                   uint32_t glane) {
  // the group's hit mask below is a 32-bit (1 << G) - 1
  static_assert(G == 8 || G == 16, "cooperative groups of 8 or 16 lanes");
  uint32_t spill[S > K ? S - K : 1];
  auto push = [&](uint32_t &sp, uint32_t v) {
    if (sp < (uint32_t)K) stk[sp * B] = v;
    else if (sp < (uint32_t)S) spill[sp - K] = v;
    else {
      atomicOr(p.status, 1u);  // unreachable when S covers the built depth
      return;
    }
    sp++;
  };
  auto pop = [&](uint32_t &sp) -> uint32_t {
    sp--;
    return sp < (uint32_t)K ? stk[sp * B] : spill[sp - K];
  };

  RayIn r = load_ray(p, rid);
  float far_ = r.tmax;  // group-uniform outside a batch's BLAS walks
  Best best;
  float best_t = 0.0f;  // this lane's own nearest accepted hit
  bool done = false;
  uint32_t cand = kInvalid, cnt = 0, sp = 0;

  // this lane's instance of the batch (already accepted by the TLAS walk)
  auto run_instance = [&](uint32_t i) {
    const InstRec &in = p.insts[i];
    const uint32_t fl = merge_flags(r.flags, in.flags);
    F3 bo, bd;
    float s;
    object_ray(in.inv, r.o, r.d, bo, bd, s);
    const F3 binv = recip(bd);
    const float near_s = r.tmin * s;
    bool ce, cb;
    cull_triangle(fl, ce, cb);
    for (uint32_t g = in.geo_begin; g < in.geo_end && !done; g++) {
      const GeoRec &geo = p.geos[g];
      if (!geometry_passes(fl, geo.flags)) continue;
      const bool opq = geometry_opaque(fl, geo.flags);
      float tg;
      if (!slab(bo, binv, near_s, far_ * s, geo.mn[0], geo.mn[1], geo.mn[2], geo.mx[0], geo.mx[1], geo.mx[2], tg))
        continue;
      uint32_t bsp = sp;  // above the TLAS entries
      uint32_t bref = LAYOUT ? geo.quad_root : geo.root_ref;
      auto push_blas = [&](uint32_t v) { push(bsp, v); };
      for (;;) {
        if (!(bref & kLeafBit)) {
          const float fs = far_ * s;
          if constexpr (LAYOUT == 2) {
            if (step_qquad(p.tri_qquads, bref, bo, binv, near_s, fs, push_blas)) continue;
          } else if constexpr (LAYOUT == 1) {
            if (step_quad(p.tri_quads, bref, bo, binv, near_s, fs, push_blas)) continue;
          } else {
            if (step_pair(p.tri_pairs, bref, bo, binv, near_s, fs, push_blas)) continue;
          }
        } else {
          const uint32_t k0 = (bref & ~kLeafBit) >> kLeafCountBits, k1 = k0 + (bref & kLeafMaxCount);
          for (uint32_t k = k0; k < k1; k++) {
            const float4 *q = reinterpret_cast<const float4 *>(p.tris + k);
            float4 a = q[0], b = q[1], c = q[2], d = q[3];
            float t, u, v;
            bool back;
            if (tri_core(bo, bd, near_s, far_ * s, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), f3(c.x, c.y, c.z),
                         f3(d.x, d.y, d.z), a.w, b.w, c.w, d.w, ce, cb, t, u, v, back)) {
              uint32_t beh = opq ? RDRT_ANYHIT_ACCEPT : anyhit_behavior(p, false, p.tri_ids[k]);
              if (beh & RDRT_ANYHIT_ACCEPT) {
                far_ = t / s;
                best_t = far_;
                best.k = k;
                best.geo = g;
                best.inst = i;
                best.u = u;
                best.v = v;
                best.back = back;
                if (fl & F_END_SEARCH) beh |= RDRT_ANYHIT_END_SEARCH;
              }
              if (beh & RDRT_ANYHIT_END_SEARCH) {
                done = true;
                break;
              }
            }
          }
          if (done) break;
        }
        if (bsp == sp) break;
        bref = pop(bsp);
      }
    }
  };
  // every lane of the group: walk the batch, then share the far bound
  auto flush = [&]() {
    if (cand != kInvalid) run_instance(cand);
    cand = kInvalid;
    cnt = 0;
    far_ = group_min<G>(far_);
    done = group_any<G>(done);
    r = load_ray(p, rid);
  };

  const TlasRoot *root = bound_root(p, r.tlas_idx);
  F3 winv = recip(r.d);
  float tn;
  if (root && slab(r.o, winv, r.tmin, far_, root->mn[0], root->mn[1], root->mn[2], root->mx[0], root->mx[1],
                   root->mx[2], tn)) {
    uint32_t ref = LAYOUT ? p.tlas_quad_roots[root - p.tlas_roots] : root->root_ref;
    auto push_tlas = [&](uint32_t v) { push(sp, v); };
    for (;;) {
      if (!(ref & kLeafBit)) {
        if constexpr (LAYOUT == 2) {
          if (step_qquad(p.tlas_qquads, ref, r.o, winv, r.tmin, far_, push_tlas)) continue;
        } else if constexpr (LAYOUT == 1) {
          if (step_quad(p.tlas_quads, ref, r.o, winv, r.tmin, far_, push_tlas)) continue;
        } else {
          if (step_pair(p.tlas_pairs, ref, r.o, winv, r.tmin, far_, push_tlas)) continue;
        }
      } else {
        const uint32_t i0 = (ref & ~kLeafBit) >> kLeafCountBits, i1 = i0 + (ref & kLeafMaxCount);
        for (uint32_t i = i0; i < i1 && !done; i++) {
          const InstRec &in = p.insts[i];
          float ti;
          if (!slab(r.o, winv, r.tmin, far_, in.mn[0], in.mn[1], in.mn[2], in.mx[0], in.mx[1], in.mx[2], ti))
            continue;
          if ((r.mask & in.mask) == 0) continue;
          if (merge_flags(r.flags, in.flags) & F_SKIP_TRIANGLES) continue;
          if (glane == cnt) cand = i;
          if (++cnt == (uint32_t)G) {
            flush();
            winv = recip(r.d);
          }
        }
        if (done) break;
      }
      if (sp == 0) break;
      ref = pop(sp);
    }
    if (cnt) flush();
  }
  // the lowest lane whose own hit is the group's nearest reports it
  const bool mine = best.k != kInvalid && best_t == far_;
  const uint64_t m = __ballot(mine);
  const uint32_t lane = __lane_id();
  const uint32_t gm = (uint32_t)(m >> (lane & ~((uint32_t)G - 1u))) & ((1u << G) - 1u);
  const uint32_t win = gm ? (uint32_t)__ffs(gm) - 1u : 0u;
  if (glane == win) {
    if (!gm) best.k = kInvalid;
    best.t = far_;
    write_result(p, p.tri_ids, rid, best, 0u, 0u, 0u);
  }
  return win | (gm ? kCoopHit : 0u);
}

// G lanes per ray: group gi of the grid takes the gi-th live ray.
template <int K, int S, int B, int LAYOUT, int G>
__device__ __forceinline__ void trace_coop_group(TraceParams &p, uint32_t n0, uint32_t n1, uint32_t *lds) {
  const uint32_t gi = (blockIdx.x * B + threadIdx.x) / (uint32_t)G;
  uint32_t slot = gi, n = n0;
  if (gi >= n0) {
    slot = gi - n0;
    n = n1;
    select_second(p);
  }
  if (slot >= n) return;  // whole groups: the lanes of a group share gi
  const uint32_t rid = p.queue ? p.queue[slot] : slot;
  if (rid >= p.n) return;
  trace_coop_one<K, S, B, LAYOUT, G>(p, rid, lds + threadIdx.x, threadIdx.x & ((uint32_t)G - 1u));
}

template <int K, int S, int B, int W, int LAYOUT>
__global__ __launch_bounds__(B, W) void k_trace_coop(TraceParams p) {
  __shared__ uint32_t lds[K * B];
  const uint32_t n0 = live_first(p), n1 = live_second(p);
  if (n0 + n1 > p.coop_max) return;  // the one-ray-per-lane kernel traces launches this large
  // the group width is uniform over the launch (the live count is)
  if (n0 + n1 <= kCoopWideMax) trace_coop_group<K, S, B, LAYOUT, (int)kCoopWideLanes>(p, n0, n1, lds);
  else trace_coop_group<K, S, B, LAYOUT, (int)kCoopLanes>(p, n0, n1, lds);
}

// One ray per thread over a grid sized for both segments' slots; the live
// counts may be smaller (device queue sizes).  The blocks of the second
// segment follow those of the first, so a block's segment is uniform.
template <int K, int S, int B, int W, bool CNT, int LAYOUT>
__global__ __launch_bounds__(B, W) void k_trace_fast(TraceParams p) {
  __shared__ uint32_t lds[K * B];
  const uint32_t n0 = live_first(p), n1 = live_second(p);
  if (p.coop_max && n0 + n1 <= p.coop_max) return;  // the cooperative kernel traces it
  const uint32_t b0 = (n0 + B - 1) / B;
  const uint32_t used = b0 + (n1 + B - 1) / B;
  if (blockIdx.x >= used) return;
  uint32_t blk = xcd_block(blockIdx.x, used);
  uint32_t n = n0;
  if (blk >= b0) {
    blk -= b0;
    n = n1;
    select_second(p);
  }
  const uint32_t slot = blk * B + threadIdx.x;
  if (slot >= n) return;
  const uint32_t rid = p.queue ? p.queue[slot] : slot;
  if (rid >= p.n) return;  // a queue entry always names one of the n rays
  trace_fast_one<K, S, B, CNT, LAYOUT>(p, rid, lds + threadIdx.x);
}

// Persistent variant: a grid sized to the resident capacity; each wave
// takes chunks of 64 consecutive queue slots from the device counter
// p.work until the queue is exhausted, so uneven rays and scenes balance
// across the chip.  Every wave exits once a fetch lands past the count.
// The second segment's slots start at the first 64-aligned position after
// the first segment's, so every chunk lies in one segment.
template <int K, int S, int B, int W, bool CNT, int LAYOUT>
__global__ __launch_bounds__(B, W) void k_trace_persist(TraceParams p) {
  __shared__ uint32_t lds[K * B];
  const uint32_t n0 = live_first(p), n1 = live_second(p);
  if (p.coop_max && n0 + n1 <= p.coop_max) return;  // the cooperative kernel traces it
  const uint32_t start1 = (n0 + 63u) & ~63u;
  const uint32_t total = n1 ? start1 + n1 : n0;
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(p.work, 64u);
    base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));  // wave-uniform: keeps q in scalar registers
    if (base >= total) break;
    TraceParams q = p;
    uint32_t n = n0;
    if (base >= start1) {
      base -= start1;
      n = n1;
      select_second(q);
    }
    const uint32_t slot = base + lane;
    if (slot < n) {
      const uint32_t rid = q.queue ? q.queue[slot] : slot;
      if (rid < q.n) trace_fast_one<K, S, B, CNT, LAYOUT>(q, rid, lds + threadIdx.x);
    }
  }
}

[[maybe_unused]] __global__ void k_zero_work(uint32_t *w) { *w = 0u; }

// One fast-traversal launch with block size B, total stack S and tuning t.
// K = stack entries per lane in LDS, W = minimum waves per SIMD asked of the
// register allocator (1 = its own choice), LAYOUT = TraceTune::wide.
template <int K, int S, int B, int W, int LAYOUT>
hipError_t launch_fast_kw(const TraceParams &p0, bool persist, uint32_t coop, hipStream_t st) {
  TraceParams p = p0;
  p.coop_max = p.counters ? 0u : coop;  // visit counting stays one ray per lane
  if (p.coop_max) {
    // enough groups for every live ray when the launch qualifies
    uint32_t live = slots_of(p) + (p.count2 ? p.slots2 : 0u);
    if (p.live_max && p.live_max < live) live = p.live_max;
    const uint32_t groups = std::min(p.coop_max, live);
    // 8 lanes per ray, or 16 when at most kCoopWideMax rays are live
    const uint32_t threads = std::max(groups * kCoopLanes, std::min(groups, kCoopWideMax) * kCoopWideLanes);
    const uint32_t grid = (threads + B - 1) / B;
    if (grid) hipLaunchKernelGGL((k_trace_coop<K, S, B, W, LAYOUT>), dim3(grid), dim3(B), 0, st, p);
    if (live <= p.coop_max) return hipGetLastError();  // the cooperative kernel always qualifies
  }
  if (persist) {
    if (!p.work) return hipErrorInvalidValue;
    if (!p.work_zeroed) hipLaunchKernelGGL(k_zero_work, dim3(1), dim3(1), 0, st, p.work);
    if (p.counters) {
      auto k = k_trace_persist<K, S, B, W, true, LAYOUT>;
      uint32_t g = std::min(resident_grid((const void *)k, B), grid_slots(p, B) / B);
      hipLaunchKernelGGL(k, dim3(g), dim3(B), 0, st, p);
    } else {
      auto k = k_trace_persist<K, S, B, W, false, LAYOUT>;
      uint32_t g = std::min(resident_grid((const void *)k, B), grid_slots(p, B) / B);
      hipLaunchKernelGGL(k, dim3(g), dim3(B), 0, st, p);
    }
    return hipGetLastError();
  }
  dim3 grid(grid_slots(p, B) / B);
  if (p.counters) hipLaunchKernelGGL((k_trace_fast<K, S, B, W, true, LAYOUT>), grid, dim3(B), 0, st, p);
  else hipLaunchKernelGGL((k_trace_fast<K, S, B, W, false, LAYOUT>), grid, dim3(B), 0, st, p);
  return hipGetLastError();
}

template <int S, int B, int LAYOUT>
hipError_t launch_fast_s(const TraceParams &p, const TraceTune &t, hipStream_t st) {
  constexpr int K8 = S < 8 ? S : 8;
  constexpr int K16 = S < 16 ? S : 16;
  const bool persist = t.persist != 0;
  if (t.waves >= 8) {
    if (t.lds <= 8) return launch_fast_kw<K8, S, B, 8, LAYOUT>(p, persist, t.coop, st);
    return launch_fast_kw<K16, S, B, 8, LAYOUT>(p, persist, t.coop, st);
  }
  if (t.waves >= 6) {
    if (t.lds <= 8) return launch_fast_kw<K8, S, B, 6, LAYOUT>(p, persist, t.coop, st);
    return launch_fast_kw<K16, S, B, 6, LAYOUT>(p, persist, t.coop, st);
  }
  if (t.lds <= 8) return launch_fast_kw<K8, S, B, 1, LAYOUT>(p, persist, t.coop, st);
  return launch_fast_kw<K16, S, B, 1, LAYOUT>(p, persist, t.coop, st);
}

template <int B, int LAYOUT>
hipError_t launch_fast_block(const TraceParams &p, uint32_t s, const TraceTune &t, hipStream_t st) {
  switch (s) {
    case 8: return launch_fast_s<8, B, LAYOUT>(p, t, st);
    case 16: return launch_fast_s<16, B, LAYOUT>(p, t, st);
    case 24: return launch_fast_s<24, B, LAYOUT>(p, t, st);
    case 32: return launch_fast_s<32, B, LAYOUT>(p, t, st);
    case 48: return launch_fast_s<48, B, LAYOUT>(p, t, st);
    case 64: return launch_fast_s<64, B, LAYOUT>(p, t, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace
}  // namespace rdrt

#endif
