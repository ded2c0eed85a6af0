// device_build.hip -- BLAS trees built on the device, level by level.
//
// Both trees of a geometry are built with the host builders' semantics
// (bvh_build.cpp), so they are node for node the same trees:
//  * the reference's tree: binned SAH on the longest axis of the node box,
//    nb <= 4 bins, first strict minimum of area(L)|L| + area(R)|R|, a stable
//    reorder by bin, the median split when every centre lands in one bin
//    (strategy.rs:99-285; the order inside each half follows a stable sort
//    by centre, where the host follows std::nth_element: only the order
//    inside leaves after a median split can differ);
//  * the fast kernels' tree: 16 bins on each axis of the centre bounds, SAH
//    in FP64 with a leaf when cheaper (at most max_leaf primitives), a stable
//    two-way partition, the object median when every centre coincides.
//
// One level is a handful of launches over all primitives (centre bounds,
// binning with block-private LDS bins when a block lies in one node, a
// node-parallel split, the partition keys, a scan and a scatter).  Node
// bounds and bin boxes are unions, exact in any order (float min / max as
// ordered-integer atomics), and the split arithmetic is the host's operation
// for operation, so the decisions are the host's.  The host converts the
// level-order nodes to the preorder Tree.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "device_build.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace {

// f(i) for i < n on the host's threads (each index once).
template <class F>
void host_for(size_t n, F f) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned threads = (unsigned)std::min<size_t>(hw, n);
  if (threads <= 1) {
    for (size_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned w = 0; w < threads; w++)
    pool.emplace_back([&] {
      for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto &th : pool) th.join();
}

constexpr uint32_t kB = 256;          // threads per block
constexpr uint32_t kFastBins = 16;    // FastTreeOptions::bins the device supports
constexpr uint32_t kMaxRefBins = 4;   // reference bins the device supports
constexpr uint32_t kMaxSlotWords = 3 * kFastBins * 7;

using u32 = uint32_t;

// ordered-integer keys of floats: u < v <=> f(u) < f(v) (finite or infinite)
__device__ __forceinline__ u32 fkey(float f) {
  const u32 u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(u32 k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
constexpr u32 kMinInit = 0xffffffffu;  // identity of a min over keys
constexpr u32 kMaxInit = 0u;           // identity of a max over keys

struct DBox {
  float mn[3], mx[3];
};
__device__ __forceinline__ DBox dempty() {
  const float inf = __builtin_inff();
  return DBox{{inf, inf, inf}, {-inf, -inf, -inf}};
}
__device__ __forceinline__ void dgrow(DBox &b, const DBox &o) {
  for (int k = 0; k < 3; k++) {
    b.mn[k] = fminf(b.mn[k], o.mn[k]);
    b.mx[k] = fmaxf(b.mx[k], o.mx[k]);
  }
}
// box3.rs:5-11
__device__ __forceinline__ float darea(const DBox &b) {
  const float w = b.mx[0] - b.mn[0], h = b.mx[1] - b.mn[1], d = b.mx[2] - b.mn[2];
  return 2.0f * (w * h + w * d + h * d);
}
// box3.rs:117-133
__device__ __forceinline__ int dlongest(const DBox &b) {
  const float x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
  if (x > y) return x > z ? 0 : 2;
  return y > z ? 1 : 2;
}

// The builder's device state for one geometry of n primitives.
struct Dev {
  u32 n;
  const float *pbox;   // 6 per primitive
  float *cen;          // 3 per primitive
  u32 *order, *order2, *node_of, *node_of2;
  uint4 *keys;         // one-hot partition key per position (n + 1)
  uint4 *scan;         // exclusive scan of keys (n + 1)
  uint4 *bsum;         // scan block sums
  // nodes
  float *nbox;         // 6 per node
  u32 *nbeg, *nend, *ndepth, *nchild, *nflags;  // flags: 1 inner, 2 median split
  u32 *nlc;            // reference: left count of a split (its split point is begin + lc)
  int *naxis;
  float *na, *ns;      // bin origin and step (reference) or centre min and scale (fast)
  u32 *nbest;          // fast: best bin
  u32 *nright, *ngeo;  // right-child steps on the path from the root; geometry
  u32 *count;          // [0] node count, [1] median splits this level, [2] error bits
  u32 cap;             // node capacity
  // per active slot
  u32 *cb;             // 6 keys per slot (centre bounds, fast)
  u32 *bins;           // kMaxSlotWords per slot: counts then box keys
  u32 lo, m;           // active nodes [lo, lo + m)
  u32 mode;            // 0 reference, 1 fast
  u32 K, axes;         // bins per axis, binned axes
  u32 max_depth, bin_size, max_leaf;
  double node_cost, prim_cost;
};

__device__ __forceinline__ DBox prim_box(const Dev &d, u32 p) {
  const float *q = d.pbox + 6 * (size_t)p;
  return DBox{{q[0], q[1], q[2]}, {q[3], q[4], q[5]}};
}
constexpr u32 kDone = 8u;    // node flag: its subtree was built whole by one thread (k_small)
constexpr u32 kSmall = 64u;  // nodes of at most this many primitives are built by one thread

// A node of this level still built by the level kernels.
__device__ __forceinline__ bool active(const Dev &d, u32 nd) {
  return nd != kInvalid && nd - d.lo < d.m && !(d.nflags[nd] & kDone);
}

// One block's primitives all lie in one node when its first and last do
// (node ranges are contiguous).  Returns that node or kInvalid.
__device__ u32 block_node(const Dev &d) {
  __shared__ u32 s_nd;
  if (threadIdx.x == 0) {
    const u32 first = blockIdx.x * kB, last = min(first + kB, d.n) - 1;
    const u32 a = d.node_of[first], b = d.node_of[last];
    s_nd = (a == b && active(d, a)) ? a : kInvalid;
  }
  __syncthreads();
  return s_nd;
}

// Centres, the identity order and every primitive's root (geometry g's
// root is node g; geo_of holds it per primitive).
__global__ __launch_bounds__(kB) void k_centres(Dev d, const u32 *geo_of) {
  const u32 i = blockIdx.x * kB + threadIdx.x;
  if (i >= d.n) return;
  const DBox b = prim_box(d, i);
  for (int k = 0; k < 3; k++) d.cen[3 * (size_t)i + k] = (b.mn[k] + b.mx[k]) * 0.5f;  // box3.rs:90-93
  d.order[i] = i;
  d.node_of[i] = geo_of[i];
}

// Union of the primitive boxes of every active node flagged `want` (1 =
// all, 2 = children of median splits, whose boxes are range unions), into
// nbox as keys first (nbox reinterpreted), decoded by k_decode_boxes.
__global__ __launch_bounds__(kB) void k_range_boxes(Dev d, u32 *keys) {
  __shared__ u32 s[6];
  const u32 nd_block = block_node(d);
  if (nd_block != kInvalid && threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? kMinInit : kMaxInit;
  __syncthreads();
  const u32 i = blockIdx.x * kB + threadIdx.x;
  if (i < d.n) {
    const u32 nd = d.node_of[i];
    if (active(d, nd)) {
      const DBox b = prim_box(d, d.order[i]);
      if (nd_block != kInvalid) {
        for (int k = 0; k < 3; k++) {
          atomicMin(&s[k], fkey(b.mn[k]));
          atomicMax(&s[3 + k], fkey(b.mx[k]));
        }
      } else {
        u32 *q = keys + 6 * (size_t)(nd - d.lo);
        for (int k = 0; k < 3; k++) {
          atomicMin(&q[k], fkey(b.mn[k]));
          atomicMax(&q[3 + k], fkey(b.mx[k]));
        }
      }
    }
  }
  __syncthreads();
  if (nd_block != kInvalid && threadIdx.x < 6) {
    u32 *q = keys + 6 * (size_t)(nd_block - d.lo);
    if (threadIdx.x < 3) atomicMin(&q[threadIdx.x], s[threadIdx.x]);
    else atomicMax(&q[threadIdx.x], s[threadIdx.x]);
  }
}

__global__ void k_decode_boxes(Dev d, const u32 *keys, u32 want) {
  const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.m) return;
  const u32 nd = d.lo + j;
  if (want == 2 && !(d.nflags[nd] & 4u)) return;  // only children of median splits
  const bool empty = d.nbeg[nd] == d.nend[nd];     // a geometry without triangles
  for (int k = 0; k < 6; k++)
    d.nbox[6 * (size_t)nd + k] = empty ? (k < 3 ? __builtin_inff() : -__builtin_inff()) : funkey(keys[6 * (size_t)j + k]);
}

// Centre bounds of every active node (fast tree).
__global__ __launch_bounds__(kB) void k_centre_bounds(Dev d) {
  __shared__ u32 s[6];
  const u32 nd_block = block_node(d);
  if (nd_block != kInvalid && threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? kMinInit : kMaxInit;
  __syncthreads();
  const u32 i = blockIdx.x * kB + threadIdx.x;
  if (i < d.n) {
    const u32 nd = d.node_of[i];
    if (active(d, nd)) {
      const u32 p = d.order[i];
      u32 *q = nd_block != kInvalid ? s : d.cb + 6 * (size_t)(nd - d.lo);
      for (int k = 0; k < 3; k++) {
        const u32 c = fkey(d.cen[3 * (size_t)p + k]);
        atomicMin(&q[k], c);
        atomicMax(&q[3 + k], c);
      }
    }
  }
  __syncthreads();
  if (nd_block != kInvalid && threadIdx.x < 6) {
    u32 *q = d.cb + 6 * (size_t)(nd_block - d.lo);
    if (threadIdx.x < 3) atomicMin(&q[threadIdx.x], s[threadIdx.x]);
    else atomicMax(&q[threadIdx.x], s[threadIdx.x]);
  }
}

// Per active node: the binning parameters.  Reference: the longest axis of
// the node box, origin and step of nb bins; fast: per axis the centre
// minimum and K / extent (0 when the axis is not live).
__global__ void k_prep(Dev d) {
  const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.m) return;
  const u32 nd = d.lo + j;
  if (d.nflags[nd] & kDone) return;
  if (d.mode == 0) {
    DBox b;
    for (int k = 0; k < 3; k++) {
      b.mn[k] = d.nbox[6 * (size_t)nd + k];
      b.mx[k] = d.nbox[6 * (size_t)nd + 3 + k];
    }
    const int axis = dlongest(b);
    d.naxis[nd] = axis;
    d.na[3 * (size_t)nd] = b.mn[axis];
    d.ns[3 * (size_t)nd] = (b.mx[axis] - b.mn[axis]) / (float)d.K;
  } else {
    for (int k = 0; k < 3; k++) {
      const float cmn = funkey(d.cb[6 * (size_t)j + k]), cmx = funkey(d.cb[6 * (size_t)j + 3 + k]);
      const float ext = cmx - cmn;
      d.na[3 * (size_t)nd + k] = cmn;
      d.ns[3 * (size_t)nd + k] = ext > 0.0f ? (float)d.K / ext : 0.0f;
    }
  }
}

// Bin of a centre: the host's formulas (bvh_build.cpp sah_split / bin_of).
__device__ __forceinline__ u32 ref_bin(float c, float a0, float step, u32 nb) {
  const float f = floorf((c - a0) / step);
  size_t w = 0;
  if (f > 0.0f) w = f >= 1.8446744e19f ? ~(size_t)0 : (size_t)f;  // Rust's saturating cast
  if (w >= nb) w = nb - 1;
  return (u32)w;
}
__device__ __forceinline__ u32 fast_bin(float c, float mn, float scale, u32 K) {
  return (u32)fminf((float)(K - 1), fmaxf(0.0f, (c - mn) * scale));
}

// Bin counts and boxes of every active node; slot layout: counts[axes * K],
// then box keys [axes * K][6].
__global__ __launch_bounds__(kB) void k_bins(Dev d) {
  __shared__ u32 s[kMaxSlotWords];
  const u32 words = d.axes * d.K * 7;
  const u32 nd_block = block_node(d);
  if (nd_block != kInvalid)
    for (u32 w = threadIdx.x; w < words; w += kB) {
      const u32 box = w - d.axes * d.K;
      s[w] = w < d.axes * d.K ? 0u : ((box % 6) < 3 ? kMinInit : kMaxInit);
    }
  __syncthreads();
  const u32 i = blockIdx.x * kB + threadIdx.x;
  if (i < d.n) {
    const u32 nd = d.node_of[i];
    if (active(d, nd)) {
      const u32 p = d.order[i];
      const DBox b = prim_box(d, p);
      u32 *q = nd_block != kInvalid ? s : d.bins + (size_t)kMaxSlotWords * (nd - d.lo);
      for (u32 a = 0; a < d.axes; a++) {
        u32 w;
        if (d.mode == 0) {
          const int axis = d.naxis[nd];
          w = ref_bin(d.cen[3 * (size_t)p + axis], d.na[3 * (size_t)nd], d.ns[3 * (size_t)nd], d.K);
        } else {
          const float sc = d.ns[3 * (size_t)nd + a];
          if (!(sc > 0.0f)) continue;  // axis not live
          w = fast_bin(d.cen[3 * (size_t)p + a], d.na[3 * (size_t)nd + a], sc, d.K);
        }
        const u32 slot = a * d.K + w;
        atomicAdd(&q[slot], 1u);
        u32 *bk = q + d.axes * d.K + 6 * slot;
        for (int k = 0; k < 3; k++) {
          atomicMin(&bk[k], fkey(b.mn[k]));
          atomicMax(&bk[3 + k], fkey(b.mx[k]));
        }
      }
    }
  }
  __syncthreads();
  if (nd_block != kInvalid) {
    u32 *q = d.bins + (size_t)kMaxSlotWords * (nd_block - d.lo);
    for (u32 w = threadIdx.x; w < words; w += kB) {
      if (w < d.axes * d.K) {
        if (s[w]) atomicAdd(&q[w], s[w]);
      } else if (((w - d.axes * d.K) % 6) < 3) {
        if (s[w] != kMinInit) atomicMin(&q[w], s[w]);
      } else if (s[w] != kMaxInit) {
        atomicMax(&q[w], s[w]);
      }
    }
  }
}

__device__ __forceinline__ DBox bin_box(const Dev &d, const u32 *q, u32 slot) {
  if (q[slot] == 0) return dempty();
  const u32 *bk = q + d.axes * d.K + 6 * slot;
  return DBox{{funkey(bk[0]), funkey(bk[1]), funkey(bk[2])}, {funkey(bk[3]), funkey(bk[4]), funkey(bk[5])}};
}

constexpr u32 kErrNodes = 1u, kErrMedian = 2u;
constexpr u32 kMaxMedianSort = 4096;  // primitives of a reference median split sorted by one thread

__device__ void new_children(Dev &d, u32 nd, const DBox &lb, const DBox &rb, u32 mid, bool median) {
  const u32 c = atomicAdd(&d.count[0], 2u);
  if (c + 2u > d.cap) {  // only with degenerate options (bin_size 0): reported, nothing written
    atomicOr(&d.count[2], kErrNodes);
    return;
  }
  if (median && d.mode == 0 && d.nend[nd] - d.nbeg[nd] > kMaxMedianSort) {
    atomicOr(&d.count[2], kErrMedian);
    return;
  }
  const u32 b = d.nbeg[nd], e = d.nend[nd], depth = d.ndepth[nd] + 1;
  const DBox kids[2] = {lb, rb};
  const u32 beg[2] = {b, mid}, end[2] = {mid, e};
  for (int s = 0; s < 2; s++) {
    for (int k = 0; k < 3; k++) {
      d.nbox[6 * (size_t)(c + s) + k] = kids[s].mn[k];
      d.nbox[6 * (size_t)(c + s) + 3 + k] = kids[s].mx[k];
    }
    d.nbeg[c + s] = beg[s];
    d.nend[c + s] = end[s];
    d.ndepth[c + s] = depth;
    d.nchild[c + s] = kInvalid;
    d.nflags[c + s] = median ? 4u : 0u;
    d.nright[c + s] = d.nright[nd] + (u32)s;
    d.ngeo[c + s] = d.ngeo[nd];
  }
  d.nchild[nd] = c;
  d.nflags[nd] |= 1u | (median ? 2u : 0u);
  d.nlc[nd] = mid - b;
  if (median) atomicAdd(&d.count[1], 1u);
}

// Split decision of every active node (one thread per node).
__global__ void k_split(Dev d) {
  const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.m) return;
  const u32 nd = d.lo + j;
  if (d.nflags[nd] & kDone) return;
  const u32 b = d.nbeg[nd], e = d.nend[nd], n = e - b, depth = d.ndepth[nd];
  const u32 *q = d.bins + (size_t)kMaxSlotWords * j;
  DBox nb;
  for (int k = 0; k < 3; k++) {
    nb.mn[k] = d.nbox[6 * (size_t)nd + k];
    nb.mx[k] = d.nbox[6 * (size_t)nd + 3 + k];
  }
  d.nchild[nd] = kInvalid;
  if (d.mode == 0) {
    if (!(depth < d.max_depth && n > d.bin_size)) return;  // utils.rs:20-37
    const u32 K = d.K;
    u32 empty = 0;
    for (u32 k = 0; k < K; k++) empty += q[k] == 0 ? 1u : 0u;
    if (empty == K - 1) {
      new_children(d, nd, dempty(), dempty(), b + n / 2, true);  // median: boxes after the sort
      return;
    }
    float best_cost = __builtin_inff();
    DBox best_l = dempty(), best_r = dempty();
    u32 best_lc = 0;
    for (u32 s = 0; s + 1 < K; s++) {
      DBox l = dempty(), r = dempty();
      u32 lc = 0, rc = 0;
      for (u32 k = 0; k <= s; k++) {
        lc += q[k];
        dgrow(l, bin_box(d, q, k));
      }
      for (u32 k = s + 1; k < K; k++) {
        rc += q[k];
        dgrow(r, bin_box(d, q, k));
      }
      const float cost = darea(l) * (float)lc + darea(r) * (float)rc;
      if (cost < best_cost) {
        best_cost = cost;
        best_l = l;
        best_r = r;
        best_lc = lc;
      }
    }
    new_children(d, nd, best_l, best_r, b + best_lc, false);
    return;
  }
  // fast tree
  if (n <= 1 || depth >= d.max_depth) return;
  bool any = false;
  for (int k = 0; k < 3; k++) any |= d.ns[3 * (size_t)nd + k] > 0.0f;
  if (!any) {
    new_children(d, nd, dempty(), dempty(), b + n / 2, true);  // object median: boxes from the ranges
    return;
  }
  const u32 K = d.K;
  const float parent_area = darea(nb);
  double best_cost = __builtin_inf();
  int best_axis = -1;
  u32 best_bin = 0;
  double right_cost[kFastBins];
  for (int axis = 0; axis < 3; axis++) {
    if (!(d.ns[3 * (size_t)nd + axis] > 0.0f)) continue;
    const u32 *c = q + axis * K;
    DBox r = dempty();
    u32 rc = 0;
    for (u32 k = K - 1; k >= 1; k--) {
      dgrow(r, bin_box(d, q, axis * K + k));
      rc += c[k];
      right_cost[k] = (double)darea(r) * rc;
    }
    DBox l = dempty();
    u32 lc = 0;
    for (u32 k = 0; k + 1 < K; k++) {
      dgrow(l, bin_box(d, q, axis * K + k));
      lc += c[k];
      if (lc == 0 || lc == n) continue;
      const double cost = (double)darea(l) * lc + right_cost[k + 1];
      if (cost < best_cost) {
        best_cost = cost;
        best_axis = axis;
        best_bin = k;
      }
    }
  }
  const double split_cost = d.node_cost + d.prim_cost * best_cost / fmax((double)parent_area, 1e-30);
  if (n <= d.max_leaf && split_cost >= d.prim_cost * n) return;  // leaf
  if (best_axis < 0) return;  // unreachable: a live axis always has a split
  DBox lb = dempty(), rb = dempty();
  u32 lc = 0;
  for (u32 k = 0; k < K; k++) {
    dgrow(k <= best_bin ? lb : rb, bin_box(d, q, best_axis * K + k));
    if (k <= best_bin) lc += q[best_axis * K + k];
  }
  d.naxis[nd] = best_axis;
  d.nbest[nd] = best_bin;
  new_children(d, nd, lb, rb, b + lc, false);
}

// Reference median split: sorts the node's range by centre on the longest
// axis (stable: ties keep their order).  One thread per node; these nodes
// are rare (every centre in one bin).
__global__ void k_median_sort(Dev d) {
  const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.m) return;
  const u32 nd = d.lo + j;
  if (d.nflags[nd] & kDone) return;
  if (!(d.nflags[nd] & 2u) || d.mode != 0) return;
  const u32 b = d.nbeg[nd], e = d.nend[nd];
  const int axis = d.naxis[nd];
  // binary insertion sort (stable)
  for (u32 i = b + 1; i < e; i++) {
    const u32 p = d.order[i];
    const float c = d.cen[3 * (size_t)p + axis];
    u32 lo = b, hi = i;
    while (lo < hi) {
      const u32 mid = (lo + hi) / 2;
      if (c < d.cen[3 * (size_t)d.order[mid] + axis]) hi = mid;
      else lo = mid + 1;
    }
    for (u32 k = i; k > lo; k--) d.order[k] = d.order[k - 1];
    d.order[lo] = p;
  }
}

// Whole subtrees of small nodes (at most kSmall primitives), one thread
// each, with the host builders' recursion (bvh_build.cpp Builder::build /
// FastBuilder::build): the same decisions as the level kernels without
// their atomics.  order2 over the node's range is the scratch of the stable
// reorders (the level's scatter rewrites it afterwards).
__device__ DBox range_union(const Dev &d, u32 b, u32 e) {
  DBox r = dempty();
  for (u32 i = b; i < e; i++) dgrow(r, prim_box(d, d.order[i]));
  return r;
}

__device__ bool small_children(Dev &d, u32 id, const DBox &lb, const DBox &rb, u32 mid, u32 *stack, int &sp) {
  const u32 c = atomicAdd(&d.count[0], 2u);
  if (c + 2u > d.cap || sp + 2 > 128) {
    atomicOr(&d.count[2], kErrNodes);
    return false;
  }
  const u32 b = d.nbeg[id], e = d.nend[id], depth = d.ndepth[id] + 1;
  const DBox kids[2] = {lb, rb};
  const u32 beg[2] = {b, mid}, end[2] = {mid, e};
  for (int q = 0; q < 2; q++) {
    for (int k = 0; k < 3; k++) {
      d.nbox[6 * (size_t)(c + q) + k] = kids[q].mn[k];
      d.nbox[6 * (size_t)(c + q) + 3 + k] = kids[q].mx[k];
    }
    d.nbeg[c + q] = beg[q];
    d.nend[c + q] = end[q];
    d.ndepth[c + q] = depth;
    d.nchild[c + q] = kInvalid;
    d.nflags[c + q] = kDone;
    d.nright[c + q] = d.nright[id] + (u32)q;
    d.ngeo[c + q] = d.ngeo[id];
  }
  d.nchild[id] = c;
  d.nflags[id] |= 1u | kDone;
  d.nlc[id] = mid - b;
  stack[sp++] = c + 1;
  stack[sp++] = c;
  return true;
}

__global__ void k_small(Dev d) {
  const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.m) return;
  const u32 root = d.lo + j;
  if ((d.nflags[root] & kDone) || d.nend[root] - d.nbeg[root] > kSmall) return;
  d.nflags[root] |= kDone;
  d.nchild[root] = kInvalid;
  u32 stack[128];
  int sp = 0;
  stack[sp++] = root;
  while (sp > 0) {
    const u32 id = stack[--sp];
    const u32 b = d.nbeg[id], e = d.nend[id], n = e - b, depth = d.ndepth[id];
    DBox box;
    for (int k = 0; k < 3; k++) {
      box.mn[k] = d.nbox[6 * (size_t)id + k];
      box.mx[k] = d.nbox[6 * (size_t)id + 3 + k];
    }
    if (d.mode == 0) {
      if (!(depth < d.max_depth && n > d.bin_size)) continue;  // utils.rs:20-37
      const u32 K = d.K;
      const int axis = dlongest(box);
      const float a0 = box.mn[axis], step = (box.mx[axis] - box.mn[axis]) / (float)K;
      u32 cnt[kMaxRefBins] = {0u, 0u, 0u, 0u};
      DBox bb[kMaxRefBins] = {dempty(), dempty(), dempty(), dempty()};
      for (u32 i = b; i < e; i++) {
        const u32 p = d.order[i];
        const u32 w = ref_bin(d.cen[3 * (size_t)p + axis], a0, step, K);
        cnt[w]++;
        dgrow(bb[w], prim_box(d, p));
      }
      u32 empty = 0;
      for (u32 k = 0; k < K; k++) empty += cnt[k] == 0 ? 1u : 0u;
      if (empty == K - 1) {  // median on the longest axis: stable sort by centre, halves
        for (u32 i = b + 1; i < e; i++) {
          const u32 p = d.order[i];
          const float c = d.cen[3 * (size_t)p + axis];
          u32 k = i;
          while (k > b && c < d.cen[3 * (size_t)d.order[k - 1] + axis]) {
            d.order[k] = d.order[k - 1];
            k--;
          }
          d.order[k] = p;
        }
        const u32 mid = b + n / 2;
        if (!small_children(d, id, range_union(d, b, mid), range_union(d, mid, e), mid, stack, sp)) return;
        continue;
      }
      float best_cost = __builtin_inff();
      DBox best_l = dempty(), best_r = dempty();
      u32 best_lc = 0;
      for (u32 q = 0; q + 1 < K; q++) {
        DBox l = dempty(), r = dempty();
        u32 lc = 0, rc = 0;
        for (u32 k = 0; k <= q; k++) {
          lc += cnt[k];
          dgrow(l, bb[k]);
        }
        for (u32 k = q + 1; k < K; k++) {
          rc += cnt[k];
          dgrow(r, bb[k]);
        }
        const float cost = darea(l) * (float)lc + darea(r) * (float)rc;
        if (cost < best_cost) {
          best_cost = cost;
          best_l = l;
          best_r = r;
          best_lc = lc;
        }
      }
      // stable reorder by bin
      u32 off[kMaxRefBins];
      off[0] = 0;
      for (u32 k = 1; k < K; k++) off[k] = off[k - 1] + cnt[k - 1];
      for (u32 i = b; i < e; i++) {
        const u32 p = d.order[i];
        d.order2[b + off[ref_bin(d.cen[3 * (size_t)p + axis], a0, step, K)]++] = p;
      }
      for (u32 i = b; i < e; i++) d.order[i] = d.order2[i];
      if (!small_children(d, id, best_l, best_r, b + best_lc, stack, sp)) return;
      continue;
    }
    // fast tree
    if (n <= 1 || depth >= d.max_depth) continue;
    const u32 K = d.K;
    float cmn[3], scale[3];
    bool live_any = false;
    for (int k = 0; k < 3; k++) {
      float lo = __builtin_inff(), hi = -__builtin_inff();
      for (u32 i = b; i < e; i++) {
        const float c = d.cen[3 * (size_t)d.order[i] + k];
        lo = fminf(lo, c);
        hi = fmaxf(hi, c);
      }
      const float ext = hi - lo;
      cmn[k] = lo;
      scale[k] = ext > 0.0f ? (float)K / ext : 0.0f;
      live_any |= ext > 0.0f;
    }
    if (!live_any) {  // every centre coincides: object median
      const u32 mid = b + n / 2;
      if (!small_children(d, id, range_union(d, b, mid), range_union(d, mid, e), mid, stack, sp)) return;
      continue;
    }
    u32 cnt[3 * kFastBins];
    DBox bb[3 * kFastBins];
    for (u32 k = 0; k < 3 * K; k++) {
      cnt[k] = 0u;
      bb[k] = dempty();
    }
    for (u32 i = b; i < e; i++) {
      const u32 p = d.order[i];
      const DBox pb = prim_box(d, p);
      for (int a = 0; a < 3; a++) {
        if (!(scale[a] > 0.0f)) continue;
        const u32 w = fast_bin(d.cen[3 * (size_t)p + a], cmn[a], scale[a], K);
        cnt[a * K + w]++;
        dgrow(bb[a * K + w], pb);
      }
    }
    const float parent_area = darea(box);
    double best_cost = __builtin_inf();
    int best_axis = -1;
    u32 best_bin = 0;
    double right_cost[kFastBins];
    for (int axis = 0; axis < 3; axis++) {
      if (!(scale[axis] > 0.0f)) continue;
      DBox r = dempty();
      u32 rc = 0;
      for (u32 k = K - 1; k >= 1; k--) {
        dgrow(r, bb[axis * K + k]);
        rc += cnt[axis * K + k];
        right_cost[k] = (double)darea(r) * rc;
      }
      DBox l = dempty();
      u32 lc = 0;
      for (u32 k = 0; k + 1 < K; k++) {
        dgrow(l, bb[axis * K + k]);
        lc += cnt[axis * K + k];
        if (lc == 0 || lc == n) continue;
        const double cost = (double)darea(l) * lc + right_cost[k + 1];
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_bin = k;
        }
      }
    }
    const double split_cost = d.node_cost + d.prim_cost * best_cost / fmax((double)parent_area, 1e-30);
    if (n <= d.max_leaf && split_cost >= d.prim_cost * n) continue;  // leaf
    if (best_axis < 0) continue;                                    // unreachable: a live axis splits
    DBox lb = dempty(), rb = dempty();
    for (u32 k = 0; k < K; k++) dgrow(k <= best_bin ? lb : rb, bb[best_axis * K + k]);
    // stable partition by the chosen bin boundary
    u32 lo = b, hi = 0;
    for (u32 i = b; i < e; i++) {
      const u32 p = d.order[i];
      if (fast_bin(d.cen[3 * (size_t)p + best_axis], cmn[best_axis], scale[best_axis], K) <= best_bin) d.order[lo++] = p;
      else d.order2[b + hi++] = p;
    }
    for (u32 i = 0; i < hi; i++) d.order[lo + i] = d.order2[b + i];
    if (!small_children(d, id, lb, rb, lo, stack, sp)) return;
  }
}

// Partition key of every position: the bin (reference, 4-way stable
// reorder), left / right (fast), or the half (median splits).
__global__ __launch_bounds__(kB) void k_keys(Dev d) {
  const u32 i = blockIdx.x * kB + threadIdx.x;
  if (i > d.n) return;
  uint4 key = make_uint4(0u, 0u, 0u, 0u);
  if (i < d.n) {
    const u32 nd = d.node_of[i];
    if (active(d, nd) && (d.nflags[nd] & 1u)) {
      u32 k;
      if (d.nflags[nd] & 2u) {
        k = i < d.nbeg[nd] + d.nlc[nd] ? 0u : 1u;
      } else {
        const u32 p = d.order[i];
        if (d.mode == 0) {
          const int axis = d.naxis[nd];
          k = ref_bin(d.cen[3 * (size_t)p + axis], d.na[3 * (size_t)nd], d.ns[3 * (size_t)nd], d.K);
        } else {
          const int axis = d.naxis[nd];
          const u32 w = fast_bin(d.cen[3 * (size_t)p + axis], d.na[3 * (size_t)nd + axis], d.ns[3 * (size_t)nd + axis],
                                 d.K);
          k = w <= d.nbest[nd] ? 0u : 1u;
        }
      }
      key.x = k == 0;
      key.y = k == 1;
      key.z = k == 2;
      key.w = k == 3;
    }
  }
  d.keys[i] = key;
}

__device__ __forceinline__ uint4 add4(uint4 a, uint4 b) { return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ u32 lane(uint4 a, u32 k) { return k == 0 ? a.x : k == 1 ? a.y : k == 2 ? a.z : a.w; }

// Exclusive scan of keys[0 .. n] in three passes: block sums, one
// workgroup over the block sums, block scans.
constexpr u32 kScanItems = 4;
constexpr u32 kScanTile = kB * kScanItems;

__global__ __launch_bounds__(kB) void k_scan_reduce(Dev d) {
  __shared__ uint4 s[kB];
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint4 t = make_uint4(0u, 0u, 0u, 0u);
  for (u32 k = 0; k < kScanItems; k++) {
    const size_t i = base + threadIdx.x * kScanItems + k;
    if (i <= d.n) t = add4(t, d.keys[i]);
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (u32 o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] = add4(s[threadIdx.x], s[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) d.bsum[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(1024) void k_scan_blocks(Dev d, u32 nblocks) {
  __shared__ uint4 s[1024];
  uint4 carry = make_uint4(0u, 0u, 0u, 0u);
  for (u32 base = 0; base < nblocks; base += 1024) {
    const u32 i = base + threadIdx.x;
    const uint4 v = i < nblocks ? d.bsum[i] : make_uint4(0u, 0u, 0u, 0u);
    s[threadIdx.x] = v;
    __syncthreads();
    for (u32 o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
      const uint4 x = threadIdx.x >= o ? s[threadIdx.x - o] : make_uint4(0u, 0u, 0u, 0u);
      __syncthreads();
      s[threadIdx.x] = add4(s[threadIdx.x], x);
      __syncthreads();
    }
    const uint4 incl = s[threadIdx.x];
    if (i < nblocks) d.bsum[i] = make_uint4(carry.x + incl.x - v.x, carry.y + incl.y - v.y, carry.z + incl.z - v.z,
                                             carry.w + incl.w - v.w);
    const uint4 total = s[1023];
    __syncthreads();
    carry = add4(carry, total);
  }
}

__global__ __launch_bounds__(kB) void k_scan_apply(Dev d) {
  __shared__ uint4 s[kB];
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint4 v[kScanItems], t = make_uint4(0u, 0u, 0u, 0u);
  for (u32 k = 0; k < kScanItems; k++) {
    const size_t i = base + threadIdx.x * kScanItems + k;
    v[k] = i <= d.n ? d.keys[i] : make_uint4(0u, 0u, 0u, 0u);
    t = add4(t, v[k]);
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (u32 o = 1; o < kB; o <<= 1) {
    const uint4 x = threadIdx.x >= o ? s[threadIdx.x - o] : make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
    s[threadIdx.x] = add4(s[threadIdx.x], x);
    __syncthreads();
  }
  uint4 run = add4(d.bsum[blockIdx.x], make_uint4(s[threadIdx.x].x - t.x, s[threadIdx.x].y - t.y,
                                                   s[threadIdx.x].z - t.z, s[threadIdx.x].w - t.w));
  for (u32 k = 0; k < kScanItems; k++) {
    const size_t i = base + threadIdx.x * kScanItems + k;
    if (i <= d.n) d.scan[i] = run;
    run = add4(run, v[k]);
  }
}

// Stable scatter of every position of a split node to its child's range;
// positions of leaves and finished nodes keep their place and leave.
__global__ __launch_bounds__(kB) void k_scatter(Dev d) {
  const u32 i = blockIdx.x * kB + threadIdx.x;
  if (i >= d.n) return;
  const u32 nd = d.node_of[i];
  const u32 p = d.order[i];
  if (!active(d, nd) || !(d.nflags[nd] & 1u)) {
    d.order2[i] = p;
    d.node_of2[i] = kInvalid;
    return;
  }
  const u32 b = d.nbeg[nd], e = d.nend[nd];
  const uint4 key = d.keys[i], sb = d.scan[b], se = d.scan[e], si = d.scan[i];
  u32 k = key.x ? 0u : key.y ? 1u : key.z ? 2u : 3u;
  u32 dest = b + (lane(si, k) - lane(sb, k));
  for (u32 q = 0; q < k; q++) dest += lane(se, q) - lane(sb, q);
  d.order2[dest] = p;
  d.node_of2[dest] = d.nchild[nd] + (dest < b + d.nlc[nd] ? 0u : 1u);
}

// Preorder from one scan.  Every inner node has two children and the
// leaves of a geometry tile its primitive range, so with L(p) the number of
// leaves starting before position p: a subtree holds 2 (L(end) - L(begin)) - 1
// nodes, and a node's preorder index is its depth plus the nodes of every
// subtree left of its path, 2 (L(begin) - L(geometry start)) - r, r being the
// right-child steps on its path.
__global__ void k_leaf_marks(Dev d, u32 total) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  if (!(d.nflags[i] & 1u) && d.nbeg[i] < d.nend[i]) d.keys[d.nbeg[i]] = make_uint4(1u, 0u, 0u, 0u);
}

__global__ void k_clear_keys(Dev d) {
  const u32 i = blockIdx.x * kB + threadIdx.x;
  if (i <= d.n) d.keys[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_emit(Dev d, u32 total, const u32 *goff, const u32 *gbase, TreeNode *out) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const u32 g = d.ngeo[i], b = d.nbeg[i], e = d.nend[i];
  const u32 lb = d.scan[b].x, le = d.scan[e].x, l0 = d.scan[goff[g]].x;
  TreeNode t;
  for (int k = 0; k < 3; k++) {
    t.box.mn[k] = d.nbox[6 * (size_t)i + k];
    t.box.mx[k] = d.nbox[6 * (size_t)i + 3 + k];
  }
  t.begin = b - goff[g];
  t.end = e - goff[g];
  t.inner = (d.nflags[i] & 1u) != 0;
  t.size = b < e ? 2u * (le - lb) - 1u : 1u;
  const u32 at = gbase[g] + d.ndepth[i] + 2u * (lb - l0) - d.nright[i];
  if (at >= total) {  // not a full binary tree over tiling leaves: reported, nothing written
    atomicOr(&d.count[2], kErrNodes);
    return;
  }
  out[at] = t;
}

u32 grid_of(size_t n, u32 b = kB) { return (u32)std::max<size_t>(1, (n + b - 1) / b); }

__global__ void k_init_keys(u32 *keys, u32 m) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)m * 6; i += (size_t)gridDim.x * blockDim.x)
    keys[i] = (i % 6) < 3 ? kMinInit : kMaxInit;
}

// counts zero, box keys empty, in every slot's first axes * K * 7 words
__global__ void k_init_bins(u32 *bins, u32 m, u32 words, u32 nbins) {
  const size_t total = (size_t)m * words;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const u32 slot = (u32)(t / words), w = (u32)(t % words);
    bins[(size_t)slot * kMaxSlotWords + w] = w < nbins ? 0u : (((w - nbins) % 6) < 3 ? kMinInit : kMaxInit);
  }
}

// One root per geometry over its primitive range [off[g], off[g + 1]).
__global__ void k_roots(Dev d, const u32 *off, u32 ngeo) {
  const u32 g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) {
    d.count[0] = ngeo;
    d.count[1] = 0u;
    d.count[2] = 0u;
  }
  if (g >= ngeo) return;
  d.nbeg[g] = off[g];
  d.nend[g] = off[g + 1];
  d.ndepth[g] = 0u;
  d.nchild[g] = kInvalid;
  d.nflags[g] = 0u;
  d.nright[g] = 0u;
  d.ngeo[g] = g;
}

__global__ void k_reset_counts(u32 *count, u32 total) {
  count[0] = total;
  count[1] = 0u;
  count[2] = 0u;
}

struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

#define DB_TRY(expr)                     \
  do {                                   \
    hipError_t e_ = (expr);              \
    if (e_ != hipSuccess) return e_;     \
  } while (0)

class Builder final : public DeviceTreeBuilder {
 public:
  explicit Builder(hipStream_t st) : st_(st) {}

  int build(const std::vector<const std::vector<Box> *> &geos, const TreeOptions &ref_opt,
            const FastTreeOptions &fast_opt, std::vector<Tree> &ref, std::vector<Tree> &fast,
            std::string &err) override {
    const u32 rb = ref_opt.bins < 2 ? 4u : ref_opt.bins;
    if (rb > kMaxRefBins) {
      err = "device build supports at most 4 reference SAH bins";
      return RDRT_ERR_CAPACITY;
    }
    if (fast_opt.bins != kFastBins) {
      err = "device build supports 16 fast-tree bins";
      return RDRT_ERR_CAPACITY;
    }
    size_t total = 0;
    for (const auto *g : geos) total += g->size();
    if (total + geos.size() >= (1u << 30)) {
      err = "device build supports fewer than 2^30 primitives per scene";
      return RDRT_ERR_CAPACITY;
    }
    ref.assign(geos.size(), Tree());
    fast.assign(geos.size(), Tree());
    if (geos.empty()) return RDRT_OK;
    if (total == 0) {  // only empty geometries: single empty leaves
      for (size_t g = 0; g < geos.size(); g++) {
        ref[g] = build_tree(*geos[g], ref_opt, 1);
        fast[g] = build_fast_tree(*geos[g], fast_opt, 1);
      }
      return RDRT_OK;
    }
    error_.clear();
    hipError_t e = upload(geos);
    if (e == hipSuccess) e = run(0, ref_opt, fast_opt, ref);
    if (e == hipSuccess) e = run(1, ref_opt, fast_opt, fast);
    if (!error_.empty()) {
      err = "device build does not support this scene: " + error_;
      return RDRT_ERR_CAPACITY;
    }
    if (e != hipSuccess) {
      err = std::string("device build: ") + hipGetErrorString(e);
      return RDRT_ERR_HIP;
    }
    return RDRT_OK;
  }

 private:
  hipStream_t st_;
  Buf pbox_, cen_, order_, order2_, node_of_, node_of2_, keys_, scan_, bsum_, nodes_, bins_, cb_, count_, geo_of_,
      off_, gbase_, emit_;
  Dev d_{};
  size_t cap_nodes_ = 0;
  std::vector<u32> off_h_;  // primitive offset of every geometry (+ total)
  std::string error_;

  hipError_t upload(const std::vector<const std::vector<Box> *> &geos) {
    const u32 ng = (u32)geos.size();
    off_h_.assign(ng + 1, 0u);
    for (u32 g = 0; g < ng; g++) off_h_[g + 1] = off_h_[g] + (u32)geos[g]->size();
    const size_t n = off_h_[ng];
    BigVec<float> h(6 * n);  // every element written below, by geometry in parallel
    BigVec<u32> gof(n);
    host_for(ng, [&](size_t g) {
      for (size_t i = 0; i < geos[g]->size(); i++) {
        const Box &b = (*geos[g])[i];
        const size_t k = off_h_[g] + i;
        std::memcpy(&h[6 * k], b.mn, 12);
        std::memcpy(&h[6 * k + 3], b.mx, 12);
        gof[k] = (u32)g;
      }
    });
    DB_TRY(pbox_.reserve(6 * n * 4));
    DB_TRY(cen_.reserve(3 * n * 4));
    DB_TRY(order_.reserve(n * 4));
    DB_TRY(order2_.reserve(n * 4));
    DB_TRY(node_of_.reserve(n * 4));
    DB_TRY(node_of2_.reserve(n * 4));
    DB_TRY(geo_of_.reserve(n * 4));
    DB_TRY(off_.reserve((ng + 1) * 4));
    DB_TRY(keys_.reserve((n + 1) * 16));
    DB_TRY(scan_.reserve((n + 1) * 16));
    DB_TRY(bsum_.reserve((size_t)(grid_of(n + 1, kScanTile) + 1) * 16));
    DB_TRY(count_.reserve(16));
    cap_nodes_ = 2 * n + ng;  // every geometry a binary tree over its primitives
    // per node: box 6 f; begin, end, depth, child, flags, lc, axis; origin 3 f, step 3 f; best
    DB_TRY(nodes_.reserve(cap_nodes_ * 22 * 4));
    DB_TRY(hipMemcpyAsync(pbox_.p, h.data(), 6 * n * 4, hipMemcpyHostToDevice, st_));
    DB_TRY(hipMemcpyAsync(geo_of_.p, gof.data(), n * 4, hipMemcpyHostToDevice, st_));
    DB_TRY(hipMemcpyAsync(off_.p, off_h_.data(), (ng + 1) * 4, hipMemcpyHostToDevice, st_));
    Dev d{};
    d.n = (u32)n;
    d.pbox = (const float *)pbox_.p;
    d.cen = (float *)cen_.p;
    d.keys = (uint4 *)keys_.p;
    d.scan = (uint4 *)scan_.p;
    d.bsum = (uint4 *)bsum_.p;
    d.count = (u32 *)count_.p;
    d.cap = (u32)cap_nodes_;
    float *f = (float *)nodes_.p;
    d.nbox = f;
    u32 *w = (u32 *)(f + 6 * cap_nodes_);
    d.nbeg = w;
    d.nend = w + cap_nodes_;
    d.ndepth = w + 2 * cap_nodes_;
    d.nchild = w + 3 * cap_nodes_;
    d.nflags = w + 4 * cap_nodes_;
    d.nlc = w + 5 * cap_nodes_;
    d.naxis = (int *)(w + 6 * cap_nodes_);
    d.na = (float *)(w + 7 * cap_nodes_);
    d.ns = (float *)(w + 10 * cap_nodes_);
    d.nbest = w + 13 * cap_nodes_;
    d.nright = w + 14 * cap_nodes_;
    d.ngeo = w + 15 * cap_nodes_;
    d_ = d;
    return hipStreamSynchronize(st_);  // before the host staging arrays go
  }

  hipError_t run(u32 mode, const TreeOptions &ro, const FastTreeOptions &fo, std::vector<Tree> &out) {
    Dev d = d_;
    const size_t n = d.n;
    const u32 ng = (u32)off_h_.size() - 1;
    d.order = (u32 *)order_.p;
    d.order2 = (u32 *)order2_.p;
    d.node_of = (u32 *)node_of_.p;
    d.node_of2 = (u32 *)node_of2_.p;
    d.mode = mode;
    d.K = mode == 0 ? (ro.bins < 2 ? 4u : ro.bins) : kFastBins;
    d.axes = mode == 0 ? 1u : 3u;
    d.max_depth = mode == 0 ? ro.max_depth : fo.max_depth;
    d.bin_size = ro.bin_size;
    d.max_leaf = std::max(1u, std::min(fo.max_leaf, kLeafMaxCount));
    d.node_cost = fo.node_cost;
    d.prim_cost = fo.prim_cost;
    // the roots: one per geometry, boxes the unions of their primitives
    hipLaunchKernelGGL(k_centres, dim3(grid_of(n)), dim3(kB), 0, st_, d, (const u32 *)geo_of_.p);
    hipLaunchKernelGGL(k_roots, dim3(grid_of(ng, 64)), dim3(64), 0, st_, d, (const u32 *)off_.p, ng);
    DB_TRY(cb_.reserve((size_t)ng * 6 * 4));
    u32 *keys = (u32 *)cb_.p;
    d.lo = 0;
    d.m = ng;
    hipLaunchKernelGGL(k_init_keys, dim3(std::min(grid_of((size_t)ng * 6), 4096u)), dim3(kB), 0, st_, keys, ng);
    hipLaunchKernelGGL(k_range_boxes, dim3(grid_of(n)), dim3(kB), 0, st_, d, keys);
    hipLaunchKernelGGL(k_decode_boxes, dim3(grid_of(ng, 64)), dim3(64), 0, st_, d, (const u32 *)keys, 1u);
    u32 lo = 0, m = ng, total = ng;
    while (m > 0) {
      d.lo = lo;
      d.m = m;
      DB_TRY(bins_.reserve((size_t)m * kMaxSlotWords * 4));
      DB_TRY(cb_.reserve((size_t)m * 6 * 4));
      d.bins = (u32 *)bins_.p;
      d.cb = (u32 *)cb_.p;
      const u32 g = grid_of(n), gn = grid_of(m, 64);
      hipLaunchKernelGGL(k_small, dim3(gn), dim3(64), 0, st_, d);
      if (mode == 1) {
        hipLaunchKernelGGL(k_init_keys, dim3(std::min(grid_of((size_t)m * 6), 4096u)), dim3(kB), 0, st_, d.cb, m);
        hipLaunchKernelGGL(k_centre_bounds, dim3(g), dim3(kB), 0, st_, d);
      }
      hipLaunchKernelGGL(k_prep, dim3(gn), dim3(64), 0, st_, d);
      const u32 nbins = d.axes * d.K;
      hipLaunchKernelGGL(k_init_bins, dim3(std::min(grid_of((size_t)m * nbins * 7), 8192u)), dim3(kB), 0, st_, d.bins,
                         m, nbins * 7, nbins);
      hipLaunchKernelGGL(k_bins, dim3(g), dim3(kB), 0, st_, d);
      hipLaunchKernelGGL(k_split, dim3(gn), dim3(64), 0, st_, d);
      if (mode == 0) hipLaunchKernelGGL(k_median_sort, dim3(gn), dim3(64), 0, st_, d);
      hipLaunchKernelGGL(k_keys, dim3(grid_of(n + 1)), dim3(kB), 0, st_, d);
      const u32 sb = grid_of(n + 1, kScanTile);
      hipLaunchKernelGGL(k_scan_reduce, dim3(sb), dim3(kB), 0, st_, d);
      hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st_, d, sb);
      hipLaunchKernelGGL(k_scan_apply, dim3(sb), dim3(kB), 0, st_, d);
      hipLaunchKernelGGL(k_scatter, dim3(g), dim3(kB), 0, st_, d);
      DB_TRY(hipGetLastError());
      u32 counts[3];
      DB_TRY(hipMemcpyAsync(counts, d.count, 12, hipMemcpyDeviceToHost, st_));
      DB_TRY(hipStreamSynchronize(st_));
      if (counts[2]) {
        error_ = counts[2] & kErrMedian ? "a median split of more than 4096 primitives (reference tree)"
                                        : "more nodes than 2n + 1 per geometry (bin_size 0)";
        return hipErrorNotSupported;
      }
      const u32 made = counts[0] - total;
      if (counts[1]) {  // children of median splits: boxes are the unions of their ranges
        Dev c = d;
        c.lo = total;
        c.m = made;
        c.node_of = d.node_of2;
        c.order = d.order2;
        DB_TRY(cb_.reserve((size_t)made * 6 * 4));
        u32 *ck = (u32 *)cb_.p;
        hipLaunchKernelGGL(k_init_keys, dim3(std::min(grid_of((size_t)made * 6), 4096u)), dim3(kB), 0, st_, ck, made);
        hipLaunchKernelGGL(k_range_boxes, dim3(g), dim3(kB), 0, st_, c, ck);
        hipLaunchKernelGGL(k_decode_boxes, dim3(grid_of(made, 64)), dim3(64), 0, st_, c, (const u32 *)ck, 2u);
      }
      std::swap(d.order, d.order2);
      std::swap(d.node_of, d.node_of2);
      lo = total;
      m = made;
      total = counts[0];
      hipLaunchKernelGGL(k_reset_counts, dim3(1), dim3(1), 0, st_, d.count, total);
    }
    return download(d, total, out);
  }

  // Preorder nodes of every geometry emitted on the device (k_emit), then
  // one copy of them and of the order to the host, split per geometry
  // (geometry-local primitive indices and ranges).
  hipError_t download(Dev &d, u32 total, std::vector<Tree> &out) {
    const size_t n = d.n;
    const u32 ng = (u32)off_h_.size() - 1;
    hipLaunchKernelGGL(k_clear_keys, dim3(grid_of(n + 1)), dim3(kB), 0, st_, d);
    hipLaunchKernelGGL(k_leaf_marks, dim3(grid_of(total)), dim3(kB), 0, st_, d, total);
    const u32 sb = grid_of(n + 1, kScanTile);
    hipLaunchKernelGGL(k_scan_reduce, dim3(sb), dim3(kB), 0, st_, d);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st_, d, sb);
    hipLaunchKernelGGL(k_scan_apply, dim3(sb), dim3(kB), 0, st_, d);
    // leaves per geometry -> node counts and bases
    std::vector<uint4> at(ng + 1);
    for (u32 g = 0; g <= ng; g++)
      DB_TRY(hipMemcpyAsync(&at[g], d.scan + off_h_[g], sizeof(uint4), hipMemcpyDeviceToHost, st_));
    DB_TRY(hipStreamSynchronize(st_));
    std::vector<u32> gbase(ng + 1, 0u);
    for (u32 g = 0; g < ng; g++) {
      const u32 leaves = at[g + 1].x - at[g].x;
      gbase[g + 1] = gbase[g] + (leaves ? 2u * leaves - 1u : 1u);
    }
    if (gbase[ng] != total) return hipErrorIllegalState;  // the node count disagrees with the leaves
    DB_TRY(gbase_.reserve((ng + 1) * 4));
    DB_TRY(emit_.reserve((size_t)total * sizeof(TreeNode)));
    DB_TRY(hipMemcpyAsync(gbase_.p, gbase.data(), (ng + 1) * 4, hipMemcpyHostToDevice, st_));
    hipLaunchKernelGGL(k_emit, dim3(grid_of(total)), dim3(kB), 0, st_, d, total, (const u32 *)off_.p,
                       (const u32 *)gbase_.p, (TreeNode *)emit_.p);
    // host copies, filled by the copies below (no zero fill first)
    BigVec<TreeNode> nodes(total);
    BigVec<u32> order(n), depth(total);
    DB_TRY(hipMemcpyAsync(nodes.data(), emit_.p, (size_t)total * sizeof(TreeNode), hipMemcpyDeviceToHost, st_));
    DB_TRY(hipMemcpyAsync(order.data(), d.order, n * 4, hipMemcpyDeviceToHost, st_));
    DB_TRY(hipMemcpyAsync(depth.data(), d.ndepth, (size_t)total * 4, hipMemcpyDeviceToHost, st_));
    BigVec<u32> ngeo(total);
    DB_TRY(hipMemcpyAsync(ngeo.data(), d.ngeo, (size_t)total * 4, hipMemcpyDeviceToHost, st_));
    u32 errs = 0;
    DB_TRY(hipMemcpyAsync(&errs, d.count + 2, 4, hipMemcpyDeviceToHost, st_));
    DB_TRY(hipStreamSynchronize(st_));
    if (errs) return hipErrorIllegalState;
    host_for(ng, [&](size_t g) {
      Tree &t = out[g];
      t.nodes.assign(nodes.begin() + gbase[g], nodes.begin() + gbase[g + 1]);
      t.order.assign(order.begin() + off_h_[g], order.begin() + off_h_[g + 1]);
      for (u32 &p : t.order) p -= off_h_[g];
      t.depth = 0;
    });
    for (u32 i = 0; i < total; i++) out[ngeo[i]].depth = std::max(out[ngeo[i]].depth, depth[i]);
    return hipSuccess;
  }
};

}  // namespace

std::unique_ptr<DeviceTreeBuilder> make_device_tree_builder(hipStream_t stream) {
  return std::unique_ptr<DeviceTreeBuilder>(new Builder(stream));
}

}  // namespace rdrt
