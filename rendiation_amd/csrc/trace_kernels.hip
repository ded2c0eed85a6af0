// trace_kernels.hip -- the reference-order traversal kernel and the launch
// dispatch of the closest-hit kernels (device code in trace_device.h).

#include <map>
#include <mutex>

#include "compact_device.h"
#include "trace_device.h"

namespace rdrt {
namespace {

// ------------------------------------------------------------------------
// Reference-order kernel: traverse_cpu.rs:52-245 on the 48-byte forest.
// ------------------------------------------------------------------------
__device__ __forceinline__ bool ref_box(F3 o, F3 inv_d, float tmin, float tmax, const RefNode &n) {
  float tn;
  return slab(o, inv_d, tmin, tmax, n.mn[0], n.mn[1], n.mn[2], n.mx[0], n.mx[1], n.mx[2], tn);
}

__device__ __forceinline__ F3 vertex(const TraceParams &p, uint32_t i) {
  const float *v = p.vertices + 3 * (size_t)i;
  return f3(v[0], v[1], v[2]);
}

template <int B>
__global__ __launch_bounds__(B) void k_trace_ref(TraceParams p) {
  const uint32_t slot = blockIdx.x * B + threadIdx.x;
  if (slot >= slots_of(p) || (p.count && slot >= *p.count)) return;
  const uint32_t rid = p.queue ? p.queue[slot] : slot;
  if (rid >= p.n) return;  // a queue entry always names one of the n rays
  RayIn r = load_ray(p, rid);
  float far_ = r.tmax;
  Best best;
  uint32_t cn = 0, ct = 0, ci = 0;
  bool done = false;
  const TlasRoot *root = bound_root(p, r.tlas_idx);
  const F3 winv = recip(r.d);
  uint32_t cur = root ? root->ref_root : kInvalid;
  while (cur != kInvalid && !done) {
    cn++;
    const RefNode n = p.tlas_forest[cur];
    if (!ref_box(r.o, winv, r.tmin * 1.0f, far_ * 1.0f, n)) {
      cur = n.miss_next;
      continue;
    }
    cur = n.hit_next;
    if (n.hit_next != n.miss_next) continue;
    for (uint32_t i = n.begin; i < n.end && !done; i++) {
      ci++;
      const InstRec &in = p.insts[i];
      // instance bounds against the ray's original range (traverse_cpu.rs:80-86)
      float ti;
      if (!slab(r.o, winv, r.tmin, r.tmax, in.mn[0], in.mn[1], in.mn[2], in.mx[0], in.mx[1], in.mx[2], ti))
        continue;
      if ((r.mask & in.mask) == 0) continue;
      const uint32_t fl = merge_flags(r.flags, in.flags);
      F3 bo, bd;
      float s;
      object_ray(in.inv, r.o, r.d, bo, bd, s);
      if (fl & F_SKIP_TRIANGLES) continue;
      const F3 binv = recip(bd);
      bool ce, cb;
      cull_triangle(fl, ce, cb);
      for (uint32_t g = in.geo_begin; g < in.geo_end && !done; g++) {
        const GeoRec &geo = p.geos[g];
        if (!geometry_passes(fl, geo.flags)) continue;
        const bool opq = geometry_opaque(fl, geo.flags);
        uint32_t bc = geo.ref_root;
        while (bc != kInvalid && !done) {
          cn++;
          const RefNode bn = p.tri_forest[bc];
          if (!ref_box(bo, binv, r.tmin * s, far_ * s, bn)) {
            bc = bn.miss_next;
            continue;
          }
          bc = bn.hit_next;
          if (bn.hit_next != bn.miss_next) continue;
          for (uint32_t k = bn.begin; k < bn.end; k++) {
            const uint32_t tri = p.redirect[k];
            const F3 v0 = vertex(p, p.indices[3 * (size_t)tri]);
            const F3 v1 = vertex(p, p.indices[3 * (size_t)tri + 1]);
            const F3 v2 = vertex(p, p.indices[3 * (size_t)tri + 2]);
            ct++;
            const F3 e1 = sub(v1, v0), e2 = sub(v2, v0);
            const F3 nn = normalize(cross(e1, e2));
            const float uu = dot(e1, e1), uv = dot(e1, e2), vv = dot(e2, e2);
            const float inverse_d = 1.0f / (uv * uv - uu * vv);
            float t, u, v;
            bool back;
            if (tri_core(bo, bd, r.tmin * s, far_ * s, v0, e1, e2, nn, uu, uv, vv, inverse_d, ce, cb, t, u, v,
                         back)) {
              uint32_t beh = anyhit_behavior(p, opq, tri);
              if (beh & RDRT_ANYHIT_ACCEPT) {
                far_ = t / s;
                best.k = k;
                best.geo = g;
                best.inst = i;
                best.t = far_;
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
        }
      }
    }
  }
  write_result(p, p.redirect, rid, best, cn, ct, ci);
}

constexpr uint32_t kStackSizes[] = {8, 16, 24, 32, 48, 64};

// Cull pass: a ray whose range misses the root box of its TLAS (or whose
// TLAS is unbound) cannot hit anything; it gets its miss record here (what
// the traversal would write: no node visited), and the others are compacted
// stably into a dense queue, so traversal waves carry no idle lanes from
// the start.
__global__ __launch_bounds__(kCullBlock) void k_trace_cull(TraceParams p, CullScratch cs) {
  const uint32_t bid = blockIdx.x, i = bid * kCullBlock + threadIdx.x;
  bool keep = false;
  if (i < p.n) {
    const RayIn r = load_ray(p, i);
    const TlasRoot *root = bound_root(p, r.tlas_idx);
    float tn;
    keep = root && slab(r.o, recip(r.d), r.tmin, r.tmax, root->mn[0], root->mn[1], root->mn[2], root->mx[0],
                        root->mx[1], root->mx[2], tn);
    // Every ray gets miss records, kept ones too (their traversal, later in
    // stream order, overwrites them): whole lines are written instead of a
    // scatter of partial ones.
    write_result(p, p.tri_ids, i, Best{}, 0u, 0u, 0u);
    for (uint32_t k = 0; k < cs.later_rounds; k++)
      *reinterpret_cast<uint4 *>(cs.later + (size_t)(k + 1) * cs.stride + i) = make_uint4(RDRT_MISS, 0u, 0u, 0u);
  }
  compact::block_compact2<kCullBlock>(keep, false, i, bid, cs.tmp, nullptr, cs.blk, nullptr);
}

}  // namespace

size_t cull_scratch_words(uint32_t slots) {
  const size_t nb = ((size_t)slots + kCullBlock - 1) / kCullBlock;
  return nb * kCullBlock + 2 * nb + (size_t)slots + 1;
}

hipError_t enqueue_cull(const TraceParams &p, const CullScratch &cs, hipStream_t st) {
  if (p.n == 0) return hipSuccess;
  if (p.queue || p.count) return hipErrorInvalidValue;
  const uint32_t nb = (p.n + kCullBlock - 1) / kCullBlock;
  hipLaunchKernelGGL(k_trace_cull, dim3(nb), dim3(kCullBlock), 0, st, p, cs);
  hipLaunchKernelGGL(compact::k_block_scan, dim3(1), dim3(compact::kScanThreads), 0, st, (const uint32_t *)nullptr,
                     p.n, kCullBlock, (const uint32_t *)cs.blk, (const uint32_t *)nullptr, cs.blk + nb,
                     (uint32_t *)nullptr, cs.count, (uint32_t *)nullptr);
  hipLaunchKernelGGL((compact::k_scatter2<kCullBlock>), dim3(nb), dim3(kCullBlock), 0, st, (const uint32_t *)nullptr,
                     p.n, (const uint32_t *)cs.tmp, (const uint32_t *)nullptr, (const uint32_t *)cs.blk,
                     (const uint32_t *)nullptr, (const uint32_t *)(cs.blk + nb), (const uint32_t *)nullptr, cs.queue,
                     (uint32_t *)nullptr, p.n);
  return hipGetLastError();
}

hipError_t launch_trace_culled(const TraceParams &p, uint32_t push_depth, const TraceTune &t, const CullScratch &cs,
                               hipStream_t st) {
  if (p.n == 0) return hipSuccess;
  hipError_t e = enqueue_cull(p, cs, st);
  if (e != hipSuccess) return e;
  TraceParams q = p;
  q.queue = cs.queue;
  q.count = cs.count;
  q.slots = p.n;
  return launch_trace(q, RDRT_TRAVERSAL_FAST, push_depth, t, st);
}

// Resident blocks of a kernel on the current device (cached per device and kernel).
uint32_t resident_grid(const void *kernel, int block) {
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, uint32_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({dev, kernel});
  if (it != cache.end()) return it->second;
  int cus = 0, per = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0) != hipSuccess || per < 1 || cus < 1)
    return 1024;
  uint32_t g = (uint32_t)(cus * per);
  cache[{dev, kernel}] = g;
  return g;
}

uint32_t stack_entries_for(uint32_t push_depth) {
  for (uint32_t s : kStackSizes)
    if (push_depth <= s) return s;
  return 0;
}

hipError_t launch_trace(const TraceParams &p, uint32_t mode, uint32_t push_depth, const TraceTune &t,
                        hipStream_t stream) {
  if (p.n == 0 || grid_slots(p, 64) == 0) return hipSuccess;
  if (mode == RDRT_TRAVERSAL_REFERENCE) {
    if (p.count2) return hipErrorInvalidValue;  // one segment per reference launch
    dim3 grid((slots_of(p) + 255) / 256);
    hipLaunchKernelGGL((k_trace_ref<256>), grid, dim3(256), 0, stream, p);
    return hipGetLastError();
  }
  const uint32_t s = stack_entries_for(push_depth);
  if (s == 0) return hipErrorInvalidValue;
  if (t.wide == 2) {
    if (!p.tri_qquads || !p.tlas_qquads || !p.tlas_quad_roots) return hipErrorInvalidValue;
    switch (t.block) {
      case 64: return launch_qquad_64(p, s, t, stream);
      case 128: return launch_qquad_128(p, s, t, stream);
      default: return launch_qquad_256(p, s, t, stream);
    }
  }
  if (t.wide == 1) {
    if (!p.tri_quads || !p.tlas_quads || !p.tlas_quad_roots) return hipErrorInvalidValue;
    switch (t.block) {
      case 64: return launch_quad_64(p, s, t, stream);
      case 128: return launch_quad_128(p, s, t, stream);
      default: return launch_quad_256(p, s, t, stream);
    }
  }
  switch (t.block) {
    case 64: return launch_fast_64(p, s, t, stream);
    case 128: return launch_fast_128(p, s, t, stream);
    default: return launch_fast_256(p, s, t, stream);
  }
}

}  // namespace rdrt
