// pt_kernels.hip -- wavefront path tracer for gfx950.
//
// Stage structure of the reference path tracer (scene/rendering/gpu-ray-
// tracing/src/feature/path_tracing/{ray_gen.rs, ray_hit.rs, ray_miss.rs}):
// ray generation, closest hit (emission, next-event estimation toward the
// directional light with an any-hit shadow ray, Lambertian BSDF sample),
// miss (sky).  Instead of the reference's task graph, each bounce is four
// launches over device-resident queues:
//   trace extension rays -> shade (+ stable compaction of the survivors and
//   of the shadow rays) -> trace shadow rays -> add unoccluded light.
// Queue sizes stay on the device, so a whole frame is one capturable
// sequence with no host round trip.
//
// Compaction: one pass fused into the shade kernel.  Each wave ballots its
// keep flags (rank = mbcnt), each block sums its waves in LDS and gets its
// exclusive offset by decoupled look-back over per-block status words; block
// order comes from an atomic ticket, so every predecessor a block waits on
// is already running.  The output equals a stable copy_if
// (shader/parallel-compute/src/stream_compaction.rs:3-45).

#include <algorithm>

#include "compact_device.h"
#include "det_sum.h"
#include "pt.h"
#include "pt_device.h"
#include "pt_shade.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace {

using u64 = unsigned long long;
using pt::V3;
using pt::v3;
using compact::block_compact2;
using compact::k_block_scan;
using compact::k_scatter2;
using compact::kScanThreads;
using namespace pts;

__global__ void k_pt_begin(PtLaunch P) {
  const uint32_t nc = 2 * P.max_bounces + 3;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x)
    P.counts[i] = i == 0 ? P.npaths : 0u;
  // the work counters of the persistent traversal and of the tail kernel
  // start at zero (so their launches need no clearing launch)
  if (blockIdx.x == 0 && threadIdx.x < 2) P.work[threadIdx.x] = 0u;
}

// ray generation (ray_gen.rs:89-107 + camera.rs:66-99)
__global__ __launch_bounds__(kPtBlock) void k_pt_camera(PtLaunch P) {
  const uint32_t path = blockIdx.x * kPtBlock + threadIdx.x;
  if (path >= P.npaths) return;
  uint32_t px, py, n, view, seed;
  pixel_of(P, path, px, py, n, view, seed);
  const float *vw = P.views + 20 * (size_t)view;
  float jx = pt::van_der_corput(n, pt::scramble(px, py, seed, 0));
  float jy = pt::sobol1(n, pt::scramble(px, py, seed, 1));
  float u = (float)px / (float)P.width + jx / (float)P.width;
  float v = (float)py / (float)P.height + jy / (float)P.height;
  V3 target = pt::unproject(vw, u, v);
  V3 cam = v3(vw[16], vw[17], vw[18]);
  V3 d = pt::normalize3(target - cam);
  write_ray(P.rays + path, cam, kFltEps, d, kFltMax);
  store3(P.beta + 3 * (size_t)path, v3(1.0f, 1.0f, 1.0f));
  store3(P.L + 3 * (size_t)path, v3(0.0f, 0.0f, 0.0f));
  P.qa[path] = path;
}

// Shades queue slots [bid * B, bid * B + B) and compacts the survivors and
// the shadow rays block-locally (the scan and scatter launches follow).
// REPLAY: every shaded slot i < count of bounce b < max_bounces writes its
// gradient term to grec[b * npaths + i] (3 x 16 bytes) and notes the
// largest |corner value| for the reproducible sum that follows.
template <int B, bool REPLAY>
__global__ __launch_bounds__(B) void k_pt_shade(PtLaunch P, uint32_t b) {
  const uint32_t bid = blockIdx.x;
  // between the traversal of bounce b and the next one: the work counters of
  // the persistent traversal and the tail kernel start that one at zero
  if (bid == 0 && threadIdx.x < 2) P.work[threadIdx.x] = 0u;
  const uint32_t count_in = min(P.counts[b], P.npaths);
  const uint32_t nblocks = (count_in + B - 1) / B;
  if (bid >= nblocks) return;  // whole block
  const uint32_t *qin = (b & 1) ? P.qb : P.qa;
  const uint32_t i = bid * B + threadIdx.x;
  bool keep = false, shadow = false;
  uint32_t path = 0;
  GeoTerm gt;
  gt.kind = 0u;
  if (i < count_in) {
    path = qin[i];
    if (path < P.npaths) {
      shade_one<REPLAY>(P, b, path, keep, shadow, gt);
      if (!REPLAY && P.tflags && b == P.tail_b0) P.tflags[path] = (shadow ? 1u : 0u) | (keep ? 2u : 0u);
    }
  }
  if (REPLAY && b < P.max_bounces) {
    uint32_t bits = 0u;
    if (i < count_in) {
      uint4 *r = P.grec + 3 * ((size_t)b * P.npaths + i);
      r[0] = make_uint4(gt.i0, gt.i1, gt.i2, gt.kind);
      r[1] = make_uint4(__float_as_uint(gt.A.x), __float_as_uint(gt.A.y), __float_as_uint(gt.A.z),
                        __float_as_uint(gt.bu));
      r[2] = make_uint4(__float_as_uint(gt.B.x), __float_as_uint(gt.B.y), __float_as_uint(gt.B.z),
                        __float_as_uint(gt.bv));
      if (gt.kind) {
        V3 pc[3], nc[3];
        geo_corners(gt, pc, nc);
        for (int k = 0; k < 3; k++) bits = max(bits, max(max_bits3(pc[k]), max_bits3(nc[k])));
      }
    }
    det::det_note_max(P.det_h, bits);
  }
  block_compact2<B>(keep, shadow, path, bid, P.tmp_a, P.tmp_b, P.blk, P.blk + P.blk_stride);
}

// Adds the recorded gradient terms of every replayed bounce into the
// position sums g_acc[0, 3 nv) and the normal sums g_acc[3 nv, 6 nv).
__global__ __launch_bounds__(kPtBlock) void k_geo_scatter(PtLaunch P, det::u64 *acc, uint32_t *flags) {
  const size_t j = (size_t)blockIdx.x * kPtBlock + threadIdx.x;
  const uint32_t b = (uint32_t)(j / P.npaths), i = (uint32_t)(j - (size_t)b * P.npaths);
  if (b >= P.max_bounces || i >= min(P.counts[b], P.npaths)) return;
  const uint4 *r = P.grec + 3 * j;
  const uint4 r0 = r[0];
  if (r0.w == 0u) return;
  const uint4 r1 = r[1], r2 = r[2];
  GeoTerm gt;
  gt.i0 = r0.x;
  gt.i1 = r0.y;
  gt.i2 = r0.z;
  gt.kind = r0.w;
  gt.A = v3(__uint_as_float(r1.x), __uint_as_float(r1.y), __uint_as_float(r1.z));
  gt.bu = __uint_as_float(r1.w);
  gt.B = v3(__uint_as_float(r2.x), __uint_as_float(r2.y), __uint_as_float(r2.z));
  gt.bv = __uint_as_float(r2.w);
  V3 pc[3], nc[3];
  geo_corners(gt, pc, nc);
  const int32_t e = P.det_h->e;
  const size_t nv3 = 3 * (size_t)P.nvertices;
  const uint32_t vi[3] = {gt.i0, gt.i1, gt.i2};
  for (int k = 0; k < 3; k++) {
    const size_t o = 3 * (size_t)vi[k];
    det::det_add(acc, flags, o, pc[k].x, e);
    det::det_add(acc, flags, o + 1, pc[k].y, e);
    det::det_add(acc, flags, o + 2, pc[k].z, e);
    if (gt.kind == 1u) {
      det::det_add(acc, flags, nv3 + o, nc[k].x, e);
      det::det_add(acc, flags, nv3 + o + 1, nc[k].y, e);
      det::det_add(acc, flags, nv3 + o + 2, nc[k].z, e);
    }
  }
}

// add the light of unoccluded shadow rays
__global__ __launch_bounds__(kPtBlock) void k_pt_nee(PtLaunch P, uint32_t b) {
  const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
  if (i >= P.counts[P.max_bounces + 2 + b] || i >= P.npaths) return;
  const uint32_t path = P.qs[i];
  if (P.shadow_hits[path].prim_id != kMiss) {
    if (P.record) P.record[(size_t)b * P.npaths + path].y = 0u;
    return;
  }
  V3 L = load3(P.L + 3 * (size_t)path) + load3(P.nee + 3 * (size_t)path);
  store3(P.L + 3 * (size_t)path, L);
}

__global__ __launch_bounds__(kPtBlock) void k_pt_finalize(PtLaunch P) {
  const uint32_t slot = blockIdx.x * kPtBlock + threadIdx.x;
  if (slot >= P.nviews * P.width * P.height) return;
  V3 s = v3(0.0f, 0.0f, 0.0f);
  for (uint32_t k = 0; k < P.spp; k++) s = s + load3(P.L + 3 * ((size_t)slot * P.spp + k));
  const float inv = 1.0f / (float)P.spp;
  store3(P.out + 3 * (size_t)image_pixel(P, slot), s * inv);
}

// standalone stable compaction (same machinery as the shade kernel)
__global__ void k_fill_u64(u64 *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

__global__ void k_fill_u32(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Sums the per-ray visit counters of the rays of one queued launch.
__global__ __launch_bounds__(kPtBlock) void k_sum_visits(const uint32_t *queue, const uint32_t *count, uint32_t n,
                                                         const uint32_t *counters, u64 *totals, uint32_t *max_visits) {
  __shared__ u64 s[3];
  __shared__ uint32_t m;
  if (threadIdx.x < 3) s[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) m = 0u;
  __syncthreads();
  const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
  if (i < n && i < *count) {
    const uint32_t r = queue[i];
    if (r < n) {
      uint32_t sum = 0;
      for (int k = 0; k < 3; k++) {
        const uint32_t c = counters[3 * (size_t)r + k];
        atomicAdd(&s[k], (u64)c);
        sum += c;
      }
      atomicMax(&m, sum);
    }
  }
  __syncthreads();
  if (threadIdx.x < 3 && s[threadIdx.x]) atomicAdd(&totals[threadIdx.x], s[threadIdx.x]);
  if (threadIdx.x == 0 && m) atomicMax(max_visits, m);
}

__global__ __launch_bounds__(kPtBlock) void k_compact(const uint32_t *mask, uint32_t n, uint32_t *tmp,
                                                      uint32_t *cnt) {
  const uint32_t bid = blockIdx.x;
  const uint32_t i = bid * kPtBlock + threadIdx.x;
  const bool keep = i < n && mask[i] != 0u;
  block_compact2<kPtBlock>(keep, false, i, bid, tmp, nullptr, cnt, nullptr);
}

// ---- adjoint ----
// Per path, from its recorded vertices k = 0..m-1 (materials a_k, e_k, light
// weights w_k) and the radiance T_m beyond the last (sky or 0), with
// beta_0 = 1, beta_{k+1} = beta_k a_k and T_k = e_k + a_k (w_k sun + T_{k+1}):
//   dC/de_k = beta_k,   dC/da_k = beta_k (w_k sun + T_{k+1}),
// scaled by the pixel's adjoint / spp.  Fully unrolled over S slots so the
// per-vertex arrays stay in registers.  The terms go into the reproducible
// sums of det_sum.h, entry 3 mi + c of acc for albedo and 3 nm + 3 mi + c
// for emission: MAXPASS only notes the largest |term|; the second pass adds
// the quantized terms, block-locally in LDS (when `lds`; 64-bit integer
// adds, so order does not matter) and then once per touched entry in
// global memory.
template <int S, bool MAXPASS>
__global__ __launch_bounds__(kPtBlock) void k_pt_backward(PtLaunch P, const float *adj, det::u64 *acc,
                                                          uint32_t *flags, det::Header *h, bool lds) {
  extern __shared__ det::u64 g_lds[];
  const uint32_t nm = P.nmaterials, slots = P.max_bounces + 2;
  if (!MAXPASS && lds) {
    for (uint32_t i = threadIdx.x; i < 6 * nm; i += kPtBlock) g_lds[i] = 0ull;
    __syncthreads();
  }
  const int32_t e = MAXPASS ? 0 : h->e;
  uint32_t bits = 0u;
  const V3 sun = load3(P.sun_radiance);
  const float inv_spp = 1.0f / (float)P.spp;
  auto add = [&](uint32_t entry, float x) {
    if (MAXPASS) {
      bits = max(bits, det::abs_bits(x));
    } else if (x != 0.0f) {
      if (!det::det_finite(x)) atomicOr(flags + entry, det::special_of(x));
      else if (lds) atomicAdd(g_lds + entry, (det::u64)det::det_quant(x, e));
      else atomicAdd(acc + entry, (det::u64)det::det_quant(x, e));
    }
  };
  for (uint32_t path = blockIdx.x * kPtBlock + threadIdx.x; path < P.npaths; path += gridDim.x * kPtBlock) {
    uint32_t code[S];
    float w[S];
    uint32_t m = 0, tail = kVertEnd;
    bool open = true;
#pragma unroll
    for (int k = 0; k < S; k++) {
      code[k] = 0u;
      w[k] = 0.0f;
      if (open && (uint32_t)k < slots) {
        const uint2 r = P.record[(size_t)k * P.npaths + path];
        if (r.x >= kVertSkyUp) {
          tail = r.x;
          open = false;
        } else {
          code[k] = r.x < nm ? r.x : nm - 1;
          w[k] = __uint_as_float(r.y);
          m = k + 1;
        }
      }
    }
    if (m == 0) continue;
    const uint32_t pixel = image_pixel(P, path / P.spp);
    const V3 a_pix = load3(adj + 3 * (size_t)pixel) * inv_spp;
    V3 beta[S];
    beta[0] = v3(1.0f, 1.0f, 1.0f);
#pragma unroll
    for (int k = 0; k + 1 < S; k++)
      if ((uint32_t)k < m) {
        const Material &mt = P.materials[code[k]];
        beta[k + 1] = beta[k] * v3(mt.albedo[0], mt.albedo[1], mt.albedo[2]);
      }
    V3 T = tail == kVertSkyUp ? load3(P.sky_up) : tail == kVertSkyDown ? load3(P.sky_down) : v3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = S - 1; k >= 0; k--)
      if ((uint32_t)k < m) {
        const uint32_t mi = code[k];
        const Material &mt = P.materials[mi];
        const V3 a = v3(mt.albedo[0], mt.albedo[1], mt.albedo[2]);
        const V3 em = v3(mt.emission[0], mt.emission[1], mt.emission[2]);
        const V3 lit = sun * w[k];
        const V3 de = a_pix * beta[k];
        const V3 da = de * (lit + T);
        add(3 * mi, da.x);
        add(3 * mi + 1, da.y);
        add(3 * mi + 2, da.z);
        add(3 * nm + 3 * mi, de.x);
        add(3 * nm + 3 * mi + 1, de.y);
        add(3 * nm + 3 * mi + 2, de.z);
        T = em + a * (lit + T);
      }
  }
  if (MAXPASS) {
    det::det_note_max(h, bits);
  } else if (lds) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 6 * nm; i += kPtBlock) {
      const det::u64 v = g_lds[i];
      if (v) atomicAdd(acc + i, v);
    }
  }
}

__global__ void k_pack_materials(const float *albedo, const float *emission, uint32_t n, Material *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Material m;
  for (int c = 0; c < 3; c++) {
    m.albedo[c] = albedo[3 * (size_t)i + c];
    m.emission[c] = emission[3 * (size_t)i + c];
  }
  m.pad0 = 0.0f;
  m.pad1 = 0.0f;
  out[i] = m;
}

// Per-block partial sums of squares (fixed lane/wave order) into part[];
// k_l2_sum adds the partials in block order.  Same grid for the same n, so
// the loss is the same bits on every call.
__global__ __launch_bounds__(kPtBlock) void k_l2_adjoint(const float *a, const float *b, uint32_t n, float *adj,
                                                         double *part) {
  __shared__ double s[kPtBlock / 64];
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * kPtBlock + threadIdx.x; i < n; i += gridDim.x * kPtBlock) {
    const float d = a[i] - b[i];
    adj[i] = 2.0f * d;
    acc += (double)d * (double)d;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63u) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kPtBlock / 64; k++) t += s[k];
    part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kPtBlock) void k_l2_sum(const double *part, uint32_t nparts, double *loss) {
  __shared__ double s[kPtBlock];
  double t = 0.0;
  for (uint32_t i = threadIdx.x; i < nparts; i += kPtBlock) t += part[i];
  s[threadIdx.x] = t;
  __syncthreads();
  for (uint32_t w = kPtBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = s[0];
}

__global__ void k_sobol(const uint32_t *index, const uint32_t *scr, uint32_t n, float *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[2 * i] = pt::van_der_corput(index[i], scr[2 * i]);
  out[2 * i + 1] = pt::sobol1(index[i], scr[2 * i + 1]);
}

// Scan of the shade kernel's block counts and scatter into the next
// extension queue and the shadow queue (sizes to counts[b + 1] and
// counts[max_bounces + 2 + b]).
void enqueue_queue_scatter(const PtLaunch &P, uint32_t b, uint32_t grid, hipStream_t st) {
  uint32_t *qout = (b & 1) ? P.qa : P.qb;
  const uint32_t s = P.blk_stride;
  hipLaunchKernelGGL(k_block_scan, dim3(1), dim3(kScanThreads), 0, st, P.counts + b, P.npaths, kPtBlock, P.blk,
                     P.blk + s, P.blk + 2 * s, P.blk + 3 * s, P.counts + b + 1, P.counts + P.max_bounces + 2 + b);
  hipLaunchKernelGGL((k_scatter2<kPtBlock>), dim3(grid), dim3(kPtBlock), 0, st, P.counts + b, P.npaths, P.tmp_a,
                     P.tmp_b, P.blk, P.blk + s, P.blk + 2 * s, P.blk + 3 * s, qout, P.qs, P.npaths);
}

}  // namespace

hipError_t launch_pt_tail(const PtLaunch &P, const TraceParams &pe, const TraceParams &ps, uint32_t push_depth,
                          const TraceTune &tune, hipStream_t st) {
  const uint32_t s = stack_entries_for(push_depth);
  if (s == 0 || !P.tflags || P.tail_b0 == 0) return hipErrorInvalidValue;
  switch (tune.wide) {
    case 1: return launch_pt_tail_1(P, pe, ps, s, tune, st);
    case 2: return launch_pt_tail_2(P, pe, ps, s, tune, st);
    default: return launch_pt_tail_0(P, pe, ps, s, tune, st);
  }
}

hipError_t enqueue_pt_frame(const PtLaunch &P, const TraceParams &scene, uint32_t push_depth,
                            const TraceTune &tune, hipStream_t st, const PtDiag *diag) {
  const uint32_t grid = (P.npaths + kPtBlock - 1) / kPtBlock;
  const bool count = diag && diag->counters;
  const uint32_t mode = count && !diag->fast_order ? RDRT_TRAVERSAL_REFERENCE : RDRT_TRAVERSAL_FAST;
  uint32_t launch = 0;
  // one traversal launch, optionally timed and visit-counted
  auto trace = [&](TraceParams t) -> hipError_t {
    if (count) t.counters = diag->counters;
    if (diag && diag->ev_begin) (void)hipEventRecord(diag->ev_begin[launch], st);
    hipError_t e = launch_trace(t, mode, push_depth, tune, st);
    if (diag && diag->ev_end) (void)hipEventRecord(diag->ev_end[launch], st);
    launch++;
    if (e == hipSuccess && count)
      hipLaunchKernelGGL(k_sum_visits, dim3((t.n + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, st, t.queue, t.count,
                         t.n, diag->counters, diag->totals, diag->max_visits + (launch - 1));
    return e;
  };
  // extension rays of bounce b and shadow rays of bounce b
  auto extension = [&](uint32_t b) {
    TraceParams tp = scene;
    tp.rays = P.rays;
    tp.hits = P.hits;
    tp.hitx = P.hitx;
    tp.meta = nullptr;
    tp.hit_meta = nullptr;
    tp.counters = nullptr;
    tp.queue = (b & 1) ? P.qb : P.qa;
    tp.count = P.counts + b;
    tp.n = P.npaths;
    tp.flags = RDRT_RAY_FLAG_CULL_BACK_FACING_TRIANGLES;  // ray_gen.rs:162-166
    tp.cull_mask = 0xffffffffu;
    tp.tlas_idx = 0;
    tp.work = P.work;
    // k_pt_begin / k_pt_shade leave the work counter at zero before each
    // traversal of an undiagnosed frame (visit counting traces twice in a row)
    tp.work_zeroed = count ? 0u : 1u;
    return tp;
  };
  auto shadow = [&](uint32_t b) {
    TraceParams sp = extension(b);
    sp.rays = P.shadow_rays;
    sp.hits = P.shadow_hits;
    sp.hitx = nullptr;
    sp.queue = P.qs;
    sp.count = P.counts + P.max_bounces + 2 + b;
    sp.flags = RDRT_RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH;  // ray_hit.rs:33-45
    return sp;
  };
  if (count) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, diag->totals, 3u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, diag->max_visits, 2u * kPtMaxBounces + 1u);
  }
  hipLaunchKernelGGL(k_pt_begin, dim3(256), dim3(256), 0, st, P);
  hipLaunchKernelGGL(k_pt_camera, dim3(grid), dim3(kPtBlock), 0, st, P);
  hipError_t e = trace(extension(0));
  if (e != hipSuccess) return e;
  for (uint32_t b = 0; b <= P.max_bounces; b++) {
    hipLaunchKernelGGL((k_pt_shade<kPtBlock, false>), dim3(grid), dim3(kPtBlock), 0, st, P, b);
    enqueue_queue_scatter(P, b, grid, st);
    if (b == P.max_bounces) break;
    if (P.tail_b0 && b == P.tail_b0 && !count) {
      // one launch finishes every path from here (its shadow ray of bounce b on)
      TraceParams pe = extension(b + 1), ps = shadow(b);
      pe.queue = ps.queue = nullptr;
      pe.count = ps.count = nullptr;
      if (diag && diag->ev_begin) (void)hipEventRecord(diag->ev_begin[launch], st);
      e = launch_pt_tail(P, pe, ps, push_depth, tune, st);  // work[1] zeroed by k_pt_shade(b)
      if (diag && diag->ev_end) (void)hipEventRecord(diag->ev_end[launch], st);
      launch++;
      if (e != hipSuccess) return e;
      break;
    }
    // The shadow rays of bounce b and the extension rays of bounce b + 1
    // both come from shade(b) and are independent: one launch traces both
    // (visit counting keeps them apart, one launch each).
    const TraceParams sp = shadow(b), ep = extension(b + 1);
    if (count) {
      e = trace(sp);
      if (e == hipSuccess) e = trace(ep);
    } else {
      TraceParams two = sp;
      two.rays2 = ep.rays;
      two.hits2 = ep.hits;
      two.hitx2 = ep.hitx;
      two.queue2 = ep.queue;
      two.count2 = ep.count;
      two.slots2 = P.npaths;
      two.flags2 = ep.flags;
      e = trace(two);
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pt_nee, dim3(grid), dim3(kPtBlock), 0, st, P, b);
  }
  hipLaunchKernelGGL(k_pt_finalize, dim3((P.nviews * P.width * P.height + kPtBlock - 1) / kPtBlock), dim3(kPtBlock),
                     0, st, P);
  return hipGetLastError();
}

hipError_t enqueue_pt_replay(const PtLaunch &P, const TraceParams &scene, uint32_t push_depth,
                             const TraceTune &tune, const DetBufs &db, hipStream_t st) {
  const uint32_t grid = (P.npaths + kPtBlock - 1) / kPtBlock;
  const size_t nv3 = 3 * (size_t)P.nvertices;
  hipError_t e = det::enqueue_clear(db.acc, db.flags, 2 * nv3, db.h, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pt_begin, dim3(256), dim3(256), 0, st, P);
  hipLaunchKernelGGL(k_pt_camera, dim3(grid), dim3(kPtBlock), 0, st, P);
  for (uint32_t b = 0; b <= P.max_bounces; b++) {
    TraceParams tp = scene;
    tp.rays = P.rays;
    tp.hits = P.hits;
    tp.hitx = P.hitx;
    tp.meta = nullptr;
    tp.hit_meta = nullptr;
    tp.counters = nullptr;
    tp.queue = (b & 1) ? P.qb : P.qa;
    tp.count = P.counts + b;
    tp.n = P.npaths;
    tp.flags = RDRT_RAY_FLAG_CULL_BACK_FACING_TRIANGLES;
    tp.cull_mask = 0xffffffffu;
    tp.tlas_idx = 0;
    tp.work = P.work;
    e = launch_trace(tp, RDRT_TRAVERSAL_FAST, push_depth, tune, st);
    if (e != hipSuccess) return e;
    if (b == P.max_bounces) break;  // the last vertex adds no geometry term
    hipLaunchKernelGGL((k_pt_shade<kPtBlock, true>), dim3(grid), dim3(kPtBlock), 0, st, P, b);
    enqueue_queue_scatter(P, b, grid, st);
  }
  // each record adds to at most 3 entries of one vertex component
  e = det::enqueue_scale(db.h, 3ull * P.max_bounces * P.npaths, st);
  if (e != hipSuccess) return e;
  const size_t nrec = (size_t)P.max_bounces * P.npaths;
  if (nrec) hipLaunchKernelGGL(k_geo_scatter, dim3((uint32_t)((nrec + kPtBlock - 1) / kPtBlock)), dim3(kPtBlock), 0,
                               st, P, db.acc, db.flags);
  e = det::enqueue_finish(db.acc, db.flags, nv3, db.h, P.g_pos, st);
  if (e == hipSuccess && P.g_nrm) e = det::enqueue_finish(db.acc + nv3, db.flags + nv3, nv3, db.h, P.g_nrm, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t enqueue_pt_backward(const PtLaunch &P, const float *adj, float *ga, float *ge, const DetBufs &db,
                               hipStream_t st) {
  const uint32_t nm = P.nmaterials;
  hipError_t e = det::enqueue_clear(db.acc, db.flags, 6 * (size_t)nm, db.h, st);
  if (e != hipSuccess) return e;
  const size_t lds_bytes = (size_t)6 * nm * sizeof(det::u64);
  const bool lds = lds_bytes <= 48 * 1024;
  const uint32_t grid = std::min<uint32_t>((P.npaths + kPtBlock - 1) / kPtBlock, 2048u);
  const uint32_t slots = P.max_bounces + 2;
  const size_t sh = lds ? lds_bytes : 0;
  auto pass = [&](auto kmax, auto kacc) {
    hipLaunchKernelGGL(kmax, dim3(grid), dim3(kPtBlock), 0, st, P, adj, db.acc, db.flags, db.h, false);
    // a path adds at most `slots` terms to one entry
    (void)det::enqueue_scale(db.h, (uint64_t)P.npaths * slots, st);
    hipLaunchKernelGGL(kacc, dim3(grid), dim3(kPtBlock), sh, st, P, adj, db.acc, db.flags, db.h, lds);
  };
  if (slots <= 4)
    pass(k_pt_backward<4, true>, k_pt_backward<4, false>);
  else if (slots <= 8)
    pass(k_pt_backward<8, true>, k_pt_backward<8, false>);
  else
    pass(k_pt_backward<kPtMaxBounces + 2, true>, k_pt_backward<kPtMaxBounces + 2, false>);
  e = det::enqueue_finish(db.acc, db.flags, 3 * (size_t)nm, db.h, ga, st);
  if (e == hipSuccess) e = det::enqueue_finish(db.acc + 3 * nm, db.flags + 3 * nm, 3 * (size_t)nm, db.h, ge, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t enqueue_pack_materials(const float *albedo, const float *emission, uint32_t n, Material *out,
                                  hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_materials, dim3((n + 255) / 256), dim3(256), 0, st, albedo, emission, n, out);
  return hipGetLastError();
}

uint32_t l2_adjoint_parts(uint32_t n) { return std::max<uint32_t>(1u, std::min<uint32_t>((n + kPtBlock - 1) / kPtBlock, 1024u)); }

hipError_t enqueue_l2_adjoint(const float *a, const float *b, uint32_t n, float *adj, double *loss, double *part,
                              hipStream_t st) {
  const uint32_t grid = l2_adjoint_parts(n);
  hipLaunchKernelGGL(k_l2_adjoint, dim3(grid), dim3(kPtBlock), 0, st, a, b, n, adj, part);
  hipLaunchKernelGGL(k_l2_sum, dim3(1), dim3(kPtBlock), 0, st, (const double *)part, grid, loss);
  return hipGetLastError();
}

size_t compact_scratch_words(uint32_t n) {
  const size_t nb = ((size_t)n + kPtBlock - 1) / kPtBlock;
  return nb * kPtBlock + 2 * nb + 1;
}

hipError_t enqueue_compact(const uint32_t *mask, uint32_t n, uint32_t *out_idx, uint32_t *out_count,
                           uint32_t *scratch, hipStream_t st) {
  const uint32_t nb = (n + kPtBlock - 1) / kPtBlock;
  if (nb == 0) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, out_count);
    return hipGetLastError();
  }
  uint32_t *tmp = scratch, *cnt = scratch + (size_t)nb * kPtBlock, *off = cnt + nb;
  hipLaunchKernelGGL(k_compact, dim3(nb), dim3(kPtBlock), 0, st, mask, n, tmp, cnt);
  hipLaunchKernelGGL(k_block_scan, dim3(1), dim3(kScanThreads), 0, st, (const uint32_t *)nullptr, n, kPtBlock, cnt,
                     (const uint32_t *)nullptr, off, (uint32_t *)nullptr, out_count, (uint32_t *)nullptr);
  hipLaunchKernelGGL((k_scatter2<kPtBlock>), dim3(nb), dim3(kPtBlock), 0, st, (const uint32_t *)nullptr, n,
                     (const uint32_t *)tmp, (const uint32_t *)nullptr, (const uint32_t *)cnt,
                     (const uint32_t *)nullptr, (const uint32_t *)off, (const uint32_t *)nullptr, out_idx,
                     (uint32_t *)nullptr, n);
  return hipGetLastError();
}

hipError_t enqueue_sobol(const uint32_t *index, const uint32_t *scr, uint32_t n, float *out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sobol, dim3((n + 255) / 256), dim3(256), 0, st, index, scr, n, out);
  return hipGetLastError();
}

}  // namespace rdrt
