// refit_kernels.hip -- device refit of the BLAS layers (see refit.h).
//
// Boxes are unions (fminf / fmaxf, exact and order-free), so a refit to the
// positions a scene was built with reproduces the built boxes bit for bit;
// the triangle records follow the host pre-pass operation for operation
// (single-rounding FP32, IEEE division and square root), so traversal t, u, v
// stay bit-identical to the oracle's on the new positions.
#include <cfloat>

#include "refit.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace {

constexpr uint32_t kRefitBlock = 256;

struct BoxF {
  float mn[3], mx[3];
};

__device__ __forceinline__ BoxF empty_box() {
  const float inf = __builtin_inff();
  return BoxF{{inf, inf, inf}, {-inf, -inf, -inf}};
}
__device__ __forceinline__ void grow(BoxF &b, const float *mn, const float *mx) {
  for (int k = 0; k < 3; k++) {
    b.mn[k] = fminf(b.mn[k], mn[k]);
    b.mx[k] = fmaxf(b.mx[k], mx[k]);
  }
}
__device__ __forceinline__ void grow_tri(BoxF &b, const RefitArgs &a, uint32_t t) {
  const float *q = a.tri_box + 6 * (size_t)t;
  grow(b, q, q + 3);
}
// union of the triangles of fast slots [k0, k1)
__device__ __forceinline__ void grow_slots(BoxF &b, const RefitArgs &a, uint32_t k0, uint32_t k1) {
  for (uint32_t k = k0; k < k1; k++) grow_tri(b, a, a.tri_ids[k]);
}
__device__ __forceinline__ void grow_ref(BoxF &b, const RefitArgs &a, uint32_t ref, bool quads) {
  if (ref & kLeafBit) {
    const uint32_t k0 = (ref & ~kLeafBit) >> kLeafCountBits;
    grow_slots(b, a, k0, k0 + (ref & kLeafMaxCount));
  } else if (quads) {
    const QuadNode &q = a.quads[ref];
    for (int c = 0; c < 4; c++)
      if (q.ref[c] != kInvalid) {
        const float mn[3] = {q.mnx[c], q.mny[c], q.mnz[c]}, mx[3] = {q.mxx[c], q.mxy[c], q.mxz[c]};
        grow(b, mn, mx);
      }
  } else {
    const PairNode &p = a.pairs[ref];
    grow(b, p.lmn, p.lmx);
    grow(b, p.rmn, p.rmx);
  }
}

// The host pre-pass (vertex-only part of the reference triangle test).
__device__ TriPre pre_triangle(const float *v0, const float *v1, const float *v2) {
  TriPre t;
  const float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  const float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
  float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const float mag_sq = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  if (mag_sq > 0.0f) {  // vec/dimension.rs:98-105
    const float inv = 1.0f / sqrtf(mag_sq);
    c[0] = c[0] * inv;
    c[1] = c[1] * inv;
    c[2] = c[2] * inv;
  }
  for (int k = 0; k < 3; k++) {
    t.v0[k] = v0[k];
    t.e1[k] = e1[k];
    t.e2[k] = e2[k];
    t.n[k] = c[k];
  }
  t.uu = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
  t.uv = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
  t.vv = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
  t.inverse_d = 1.0f / (t.uv * t.uv - t.uu * t.vv);
  return t;
}

// The host quantizer of a child-quad node: per axis the smallest
// power-of-two step whose outward-rounded 8-bit bounds fit, each bound
// checked with the decoding arithmetic (corner + q * step).
__device__ QuadNodeQ quantize(const QuadNode &q) {
  QuadNodeQ out;
  int n = 0;
  while (n < 4 && q.ref[n] != kInvalid) n++;
  const float *mn[3] = {q.mnx, q.mny, q.mnz}, *mx[3] = {q.mxx, q.mxy, q.mxz};
  float corner[3], step[3];
  uint32_t lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    float l = mn[a][0], h = mx[a][0];
    for (int k = 1; k < n; k++) {
      l = fminf(l, mn[a][k]);
      h = fmaxf(h, mx[a][k]);
    }
    corner[a] = l;
    const double ext = (double)h - (double)l;
    const double min_step = ldexp(1.0, -126), max_step = ldexp(1.0, 127);
    double s = min_step;
    if (ext > 0.0 && isfinite(ext)) s = fmax(min_step, ldexp(1.0, (int)ceil(log2(ext / 255.0))));
    if (!isfinite(ext)) s = 2.0 * max_step;
    for (;;) {
      if (s > max_step) {
        step[a] = FLT_MAX;
        lo[a] = 0u;
        hi[a] = 0u;
        for (int k = 0; k < n; k++) hi[a] |= 255u << (8 * k);
        break;
      }
      const float sf = (float)s;
      bool ok = true;
      uint32_t wl = 0, wh = 0;
      for (int k = 0; k < n && ok; k++) {
        double ql = floor(((double)mn[a][k] - (double)l) / s), qh = ceil(((double)mx[a][k] - (double)l) / s);
        ql = fmin(fmax(ql, 0.0), 255.0);
        qh = fmax(qh, 0.0);
        uint32_t il = (uint32_t)ql, ih = (uint32_t)fmin(qh, 256.0);
        while (il > 0 && l + (float)il * sf > mn[a][k]) il--;
        while (ih <= 255 && l + (float)ih * sf < mx[a][k]) ih++;
        if (ih > 255) {
          ok = false;
          break;
        }
        wl |= il << (8 * k);
        wh |= ih << (8 * k);
      }
      if (ok) {
        step[a] = sf;
        lo[a] = wl;
        hi[a] = wh;
        break;
      }
      s *= 2.0;
    }
  }
  out.ox = corner[0];
  out.oy = corner[1];
  out.oz = corner[2];
  out.sx = step[0];
  out.sy = step[1];
  out.sz = step[2];
  out.lox = lo[0];
  out.hix = hi[0];
  out.loy = lo[1];
  out.hiy = hi[1];
  out.loz = lo[2];
  out.hiz = hi[2];
  for (int k = 0; k < 4; k++) out.ref[k] = q.ref[k];
  return out;
}

__global__ __launch_bounds__(kRefitBlock) void k_refit_tris(RefitArgs a) {
  const uint32_t k = blockIdx.x * kRefitBlock + threadIdx.x;
  if (k >= a.nslots) return;
  const uint32_t t = a.tri_ids[k];
  const float *v[3];
  for (int j = 0; j < 3; j++) v[j] = a.vertices + 3 * (size_t)a.indices[3 * (size_t)t + j];
  a.tris[k] = pre_triangle(v[0], v[1], v[2]);
  BoxF b = empty_box();
  for (int j = 0; j < 3; j++) grow(b, v[j], v[j]);
  float *q = a.tri_box + 6 * (size_t)t;
  for (int j = 0; j < 3; j++) {
    q[j] = b.mn[j];
    q[3 + j] = b.mx[j];
  }
}

// Reference forest: a leaf (hit == miss) is the union of its triangles, an
// inner node the union of its two children (left = hit, right = the left's
// miss link).
__global__ __launch_bounds__(kRefitBlock) void k_refit_forest(RefitArgs a, uint32_t first, uint32_t count) {
  const uint32_t j = blockIdx.x * kRefitBlock + threadIdx.x;
  if (j >= count) return;
  const uint32_t i = a.forest_nodes[first + j];
  const RefNode &n = a.forest[i];
  BoxF b = empty_box();
  if (n.hit_next == n.miss_next) {
    for (uint32_t r = n.begin; r < n.end; r++) grow_tri(b, a, a.redirect[r]);
  } else {
    const RefNode &l = a.forest[n.hit_next];
    const RefNode &r = a.forest[l.miss_next];
    grow(b, l.mn, l.mx);
    grow(b, r.mn, r.mx);
  }
  RefNode &o = a.forest[i];
  for (int k = 0; k < 3; k++) {
    o.mn[k] = b.mn[k];
    o.mx[k] = b.mx[k];
  }
}

__global__ __launch_bounds__(kRefitBlock) void k_refit_pairs(RefitArgs a, uint32_t first, uint32_t count) {
  const uint32_t j = blockIdx.x * kRefitBlock + threadIdx.x;
  if (j >= count) return;
  const uint32_t i = a.pair_nodes[first + j];
  BoxF l = empty_box(), r = empty_box();
  grow_ref(l, a, a.pairs[i].lref, false);
  grow_ref(r, a, a.pairs[i].rref, false);
  PairNode &p = a.pairs[i];
  for (int k = 0; k < 3; k++) {
    p.lmn[k] = l.mn[k];
    p.lmx[k] = l.mx[k];
    p.rmn[k] = r.mn[k];
    p.rmx[k] = r.mx[k];
  }
}

__global__ __launch_bounds__(kRefitBlock) void k_refit_quads(RefitArgs a, uint32_t first, uint32_t count) {
  const uint32_t j = blockIdx.x * kRefitBlock + threadIdx.x;
  if (j >= count) return;
  const uint32_t i = a.quad_nodes[first + j];
  QuadNode q = a.quads[i];
  for (int c = 0; c < 4; c++) {
    if (q.ref[c] == kInvalid) continue;
    BoxF b = empty_box();
    grow_ref(b, a, q.ref[c], true);
    q.mnx[c] = b.mn[0];
    q.mny[c] = b.mn[1];
    q.mnz[c] = b.mn[2];
    q.mxx[c] = b.mx[0];
    q.mxy[c] = b.mx[1];
    q.mxz[c] = b.mx[2];
  }
  a.quads[i] = q;
  a.qquads[i] = quantize(q);
}

__global__ void k_refit_geos(RefitArgs a) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.ngeos) return;
  const RefNode &root = a.forest[a.geos[g].ref_root];
  for (int k = 0; k < 3; k++) {
    a.geos[g].mn[k] = root.mn[k];
    a.geos[g].mx[k] = root.mx[k];
  }
}

uint32_t blocks(uint32_t n) { return (n + kRefitBlock - 1) / kRefitBlock; }

// The records a commit computes on the device: the triangle record of every
// fast slot, the quantized twin of every child quad.
__global__ __launch_bounds__(kRefitBlock) void k_scene_records(const float *vertices, const uint32_t *indices,
                                                               const uint32_t *tri_ids, uint32_t nslots, TriPre *tris,
                                                               const QuadNode *quads, uint32_t nquads,
                                                               QuadNodeQ *qquads) {
  const uint32_t k = blockIdx.x * kRefitBlock + threadIdx.x;
  if (k < nslots) {
    const uint32_t t = tri_ids[k];
    const float *v[3];
    for (int j = 0; j < 3; j++) v[j] = vertices + 3 * (size_t)indices[3 * (size_t)t + j];
    tris[k] = pre_triangle(v[0], v[1], v[2]);
  }
  if (k < nquads) qquads[k] = quantize(quads[k]);
}

}  // namespace

hipError_t enqueue_scene_records(const float *vertices, const uint32_t *indices, const uint32_t *tri_ids,
                                 uint32_t nslots, TriPre *tris, const QuadNode *quads, uint32_t nquads,
                                 QuadNodeQ *qquads, hipStream_t st) {
  const uint32_t n = nslots > nquads ? nslots : nquads;
  if (n) hipLaunchKernelGGL(k_scene_records, dim3(blocks(n)), dim3(kRefitBlock), 0, st, vertices, indices, tri_ids,
                            nslots, tris, quads, nquads, qquads);
  return hipGetLastError();
}

hipError_t enqueue_refit(const RefitArgs &a, const uint32_t *forest_off, uint32_t forest_levels,
                         const uint32_t *pair_off, uint32_t pair_levels, const uint32_t *quad_off,
                         uint32_t quad_levels, hipStream_t st) {
  if (a.nslots) hipLaunchKernelGGL(k_refit_tris, dim3(blocks(a.nslots)), dim3(kRefitBlock), 0, st, a);
  for (uint32_t l = 0; l < forest_levels; l++) {
    const uint32_t n = forest_off[l + 1] - forest_off[l];
    if (n) hipLaunchKernelGGL(k_refit_forest, dim3(blocks(n)), dim3(kRefitBlock), 0, st, a, forest_off[l], n);
  }
  for (uint32_t l = 0; l < pair_levels; l++) {
    const uint32_t n = pair_off[l + 1] - pair_off[l];
    if (n) hipLaunchKernelGGL(k_refit_pairs, dim3(blocks(n)), dim3(kRefitBlock), 0, st, a, pair_off[l], n);
  }
  for (uint32_t l = 0; l < quad_levels; l++) {
    const uint32_t n = quad_off[l + 1] - quad_off[l];
    if (n) hipLaunchKernelGGL(k_refit_quads, dim3(blocks(n)), dim3(kRefitBlock), 0, st, a, quad_off[l], n);
  }
  if (a.ngeos) hipLaunchKernelGGL(k_refit_geos, dim3((a.ngeos + 63) / 64), dim3(64), 0, st, a);
  return hipGetLastError();
}

}  // namespace rdrt
