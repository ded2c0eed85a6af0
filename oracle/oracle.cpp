// oracle.cpp -- CPU oracle (TEST INFRASTRUCTURE ONLY, never linked by the
// product).  A plain restatement of the reference CPU ray-query path, with the
// reference's FP32 operation order.  Build with -ffp-contract=off and no
// fast-math so every float operation is a single IEEE-754 rounding, as in the
// reference's Rust code.
//
// Parity pins for each part live in tests/ (brute force, closed forms,
// invariants, golden vectors); see the file header of oracle.h for the list
// of reference files restated here.

#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t INVALID_NEXT = 0xffffffffu;  // naive/mod.rs:611

// ---------------------------------------------------------------- algebra --
// math/algebra/src/vec/vec3.rs, vec/dimension.rs, vec/mod.rs
struct V3 {
  float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator/(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
// vec3.rs:26-28
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vec3.rs:116-123
inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Rust f32::min / f32::max ignore a NaN operand, like fminf / fmaxf.
inline float rmin(float a, float b) { return std::fmin(a, b); }
inline float rmax(float a, float b) { return std::fmax(a, b); }
// dimension.rs:53-60
inline V3 vmin(V3 a, V3 b) { return v3(rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return v3(rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)); }
// vec3.rs max_channel / min_channel
inline float max_channel(V3 a) { return rmax(rmax(a.x, a.y), a.z); }
inline float min_channel(V3 a) { return rmin(rmin(a.x, a.y), a.z); }
// dimension.rs:98-105 -- multiply by the reciprocal of the length
inline V3 normalize(V3 a) {
  float mag_sq = dot(a, a);
  if (mag_sq > 0.0f) {
    float inv_sqrt = 1.0f / std::sqrt(mag_sq);
    return a * inv_sqrt;
  }
  return a;
}
inline float length(V3 a) { return std::sqrt(dot(a, a)); }

// Column-major Mat4, fields named as math/algebra/src/mat/mat4.rs:9-14.
struct M4 {
  float a1, a2, a3, a4, b1, b2, b3, b4, c1, c2, c3, c4, d1, d2, d3, d4;
};
inline M4 m4_from(const float *m) {
  M4 r;
  std::memcpy(&r, m, sizeof(M4));
  return r;
}
inline M4 m4_identity() {
  M4 r{};
  r.a1 = r.b2 = r.c3 = r.d4 = 1.0f;
  return r;
}
// mat4.rs det
float m4_det(const M4 &m) {
  return m.a1 * m.b2 * m.c3 * m.d4 - m.a1 * m.b2 * m.c4 * m.d3 + m.a1 * m.b3 * m.c4 * m.d2 -
         m.a1 * m.b3 * m.c2 * m.d4 + m.a1 * m.b4 * m.c2 * m.d3 - m.a1 * m.b4 * m.c3 * m.d2

         - m.a2 * m.b3 * m.c4 * m.d1 + m.a2 * m.b3 * m.c1 * m.d4 - m.a2 * m.b4 * m.c1 * m.d3 +
         m.a2 * m.b4 * m.c3 * m.d1 - m.a2 * m.b1 * m.c3 * m.d4 + m.a2 * m.b1 * m.c4 * m.d3

         + m.a3 * m.b4 * m.c1 * m.d2 - m.a3 * m.b4 * m.c2 * m.d1 + m.a3 * m.b1 * m.c2 * m.d4 -
         m.a3 * m.b1 * m.c4 * m.d2 + m.a3 * m.b2 * m.c4 * m.d1 - m.a3 * m.b2 * m.c1 * m.d4

         - m.a4 * m.b1 * m.c2 * m.d3 + m.a4 * m.b1 * m.c3 * m.d2 - m.a4 * m.b2 * m.c3 * m.d1 +
         m.a4 * m.b2 * m.c1 * m.d3 - m.a4 * m.b3 * m.c1 * m.d2 + m.a4 * m.b3 * m.c2 * m.d1;
}
// mat4.rs inverse + mat/dimension.rs:18-20 inverse_or_identity
M4 m4_inverse_or_identity(const M4 &m) {
  float det = m4_det(m);
  if (det == 0.0f) return m4_identity();
  float inv_det = 1.0f / det;
  float ninv = -inv_det;
  M4 r;
  r.a1 = inv_det * (m.b2 * (m.c3 * m.d4 - m.c4 * m.d3) + m.b3 * (m.c4 * m.d2 - m.c2 * m.d4) + m.b4 * (m.c2 * m.d3 - m.c3 * m.d2));
  r.a2 = ninv * (m.a2 * (m.c3 * m.d4 - m.c4 * m.d3) + m.a3 * (m.c4 * m.d2 - m.c2 * m.d4) + m.a4 * (m.c2 * m.d3 - m.c3 * m.d2));
  r.a3 = inv_det * (m.a2 * (m.b3 * m.d4 - m.b4 * m.d3) + m.a3 * (m.b4 * m.d2 - m.b2 * m.d4) + m.a4 * (m.b2 * m.d3 - m.b3 * m.d2));
  r.a4 = ninv * (m.a2 * (m.b3 * m.c4 - m.b4 * m.c3) + m.a3 * (m.b4 * m.c2 - m.b2 * m.c4) + m.a4 * (m.b2 * m.c3 - m.b3 * m.c2));

  r.b1 = ninv * (m.b1 * (m.c3 * m.d4 - m.c4 * m.d3) + m.b3 * (m.c4 * m.d1 - m.c1 * m.d4) + m.b4 * (m.c1 * m.d3 - m.c3 * m.d1));
  r.b2 = inv_det * (m.a1 * (m.c3 * m.d4 - m.c4 * m.d3) + m.a3 * (m.c4 * m.d1 - m.c1 * m.d4) + m.a4 * (m.c1 * m.d3 - m.c3 * m.d1));
  r.b3 = ninv * (m.a1 * (m.b3 * m.d4 - m.b4 * m.d3) + m.a3 * (m.b4 * m.d1 - m.b1 * m.d4) + m.a4 * (m.b1 * m.d3 - m.b3 * m.d1));
  r.b4 = inv_det * (m.a1 * (m.b3 * m.c4 - m.b4 * m.c3) + m.a3 * (m.b4 * m.c1 - m.b1 * m.c4) + m.a4 * (m.b1 * m.c3 - m.b3 * m.c1));

  r.c1 = inv_det * (m.b1 * (m.c2 * m.d4 - m.c4 * m.d2) + m.b2 * (m.c4 * m.d1 - m.c1 * m.d4) + m.b4 * (m.c1 * m.d2 - m.c2 * m.d1));
  r.c2 = ninv * (m.a1 * (m.c2 * m.d4 - m.c4 * m.d2) + m.a2 * (m.c4 * m.d1 - m.c1 * m.d4) + m.a4 * (m.c1 * m.d2 - m.c2 * m.d1));
  r.c3 = inv_det * (m.a1 * (m.b2 * m.d4 - m.b4 * m.d2) + m.a2 * (m.b4 * m.d1 - m.b1 * m.d4) + m.a4 * (m.b1 * m.d2 - m.b2 * m.d1));
  r.c4 = ninv * (m.a1 * (m.b2 * m.c4 - m.b4 * m.c2) + m.a2 * (m.b4 * m.c1 - m.b1 * m.c4) + m.a4 * (m.b1 * m.c2 - m.b2 * m.c1));

  r.d1 = ninv * (m.b1 * (m.c2 * m.d3 - m.c3 * m.d2) + m.b2 * (m.c3 * m.d1 - m.c1 * m.d3) + m.b3 * (m.c1 * m.d2 - m.c2 * m.d1));
  r.d2 = inv_det * (m.a1 * (m.c2 * m.d3 - m.c3 * m.d2) + m.a2 * (m.c3 * m.d1 - m.c1 * m.d3) + m.a3 * (m.c1 * m.d2 - m.c2 * m.d1));
  r.d3 = ninv * (m.a1 * (m.b2 * m.d3 - m.b3 * m.d2) + m.a2 * (m.b3 * m.d1 - m.b1 * m.d3) + m.a3 * (m.b1 * m.d2 - m.b2 * m.d1));
  r.d4 = inv_det * (m.a1 * (m.b2 * m.c3 - m.b3 * m.c2) + m.a2 * (m.b3 * m.c1 - m.b1 * m.c3) + m.a3 * (m.b1 * m.c2 - m.b2 * m.c1));
  return r;
}
// mat3.rs:37-42 det of to_mat3()
float m3_det(const M4 &m) {
  float t11 = m.c3 * m.b2 - m.b3 * m.c2;
  float t12 = m.b3 * m.c1 - m.c3 * m.b1;
  float t13 = m.c2 * m.b1 - m.b2 * m.c1;
  return m.a1 * t11 + m.a2 * t12 + m.a3 * t13;
}
// mat4.rs:155-169: (M * (v,1)).xyz / w
V3 m4_mul_point(const M4 &m, V3 v) {
  float w1 = 1.0f;
  float x = v.x * m.a1 + v.y * m.b1 + v.z * m.c1 + w1 * m.d1;
  float y = v.x * m.a2 + v.y * m.b2 + v.z * m.c2 + w1 * m.d2;
  float z = v.x * m.a3 + v.y * m.b3 + v.z * m.c3 + w1 * m.d3;
  float w = v.x * m.a4 + v.y * m.b4 + v.z * m.c4 + w1 * m.d4;
  return v3(x, y, z) / w;
}
// mat3.rs:103-109 (to_mat3() * v)
V3 m3_mul(const M4 &m, V3 v) {
  return v3(v.x * m.a1 + v.y * m.b1 + v.z * m.c1, v.x * m.a2 + v.y * m.b2 + v.z * m.c2,
            v.x * m.a3 + v.y * m.b3 + v.z * m.c3);
}

// ----------------------------------------------------------------- Box3 ----
// math/geometry/src/hyper_aabb.rs:20-58, dimension3/box3.rs
struct Box3 {
  V3 min, max;
};
inline Box3 box_empty() {
  const float inf = std::numeric_limits<float>::infinity();
  return Box3{v3(inf, inf, inf), v3(-inf, -inf, -inf)};
}
inline void expand_by_point(Box3 &b, V3 p) {
  b.min = vmin(b.min, p);
  b.max = vmax(b.max, p);
}
inline void expand_by_other(Box3 &b, const Box3 &o) {
  b.min = vmin(b.min, o.min);
  b.max = vmax(b.max, o.max);
}
inline bool box_is_empty(const Box3 &b) {
  return (b.max.x < b.min.x) || (b.max.y < b.min.y) || (b.max.z < b.min.z);
}
// box3.rs:90-93
inline V3 box_center(const Box3 &b) { return (b.min + b.max) * 0.5f; }
// box3.rs:5-11 (LebesgueMeasurable<T, 2>)
inline float box_surface_area(const Box3 &b) {
  float w = b.max.x - b.min.x, h = b.max.y - b.min.y, d = b.max.z - b.min.z;
  return 2.0f * (w * h + w * d + h * d);
}
// box3.rs:117-133 -- ties go to Z, then Y
inline int longest_axis(const Box3 &b) {
  float xl = b.max.x - b.min.x, yl = b.max.y - b.min.y, zl = b.max.z - b.min.z;
  if (xl > yl) return xl > zl ? 0 : 2;
  return yl > zl ? 1 : 2;
}
inline float axis_of(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
// box3.rs:22-38 apply_matrix: 8 corners in 000..111 order, then collect
Box3 box_apply_matrix(const Box3 &b, const M4 &m) {
  if (box_is_empty(b)) return b;
  V3 pts[8] = {
      m4_mul_point(m, v3(b.min.x, b.min.y, b.min.z)), m4_mul_point(m, v3(b.min.x, b.min.y, b.max.z)),
      m4_mul_point(m, v3(b.min.x, b.max.y, b.min.z)), m4_mul_point(m, v3(b.min.x, b.max.y, b.max.z)),
      m4_mul_point(m, v3(b.max.x, b.min.y, b.min.z)), m4_mul_point(m, v3(b.max.x, b.min.y, b.max.z)),
      m4_mul_point(m, v3(b.max.x, b.max.y, b.min.z)), m4_mul_point(m, v3(b.max.x, b.max.y, b.max.z)),
  };
  Box3 r = box_empty();
  for (auto &p : pts) expand_by_point(r, p);
  return r;
}

// -------------------------------------------------------- BVH build --------
// content/space/src/bvh/node.rs:124-171
struct FNode {
  Box3 bounding;
  size_t start, end;  // primitive_range
  size_t self_index;
  bool has_child;
  size_t left_count;
};

struct BuildPrim {
  Box3 bounding;
  V3 center;
};

struct TreeBuildOption {  // utils.rs:192-210
  size_t max_tree_depth, bin_size;
  bool should_continue(size_t count, size_t depth) const {
    return depth < max_tree_depth && count > bin_size;
  }
};

struct FlattenBVH {
  std::vector<FNode> nodes;
  std::vector<size_t> sorted;
  size_t depth = 0;  // deepest node depth reached (root = 0)
};

struct Split {
  Box3 lb;
  size_t ls, le;
  Box3 rb;
  size_t rs, re;
};

// utils.rs:224-238 bounding_from_build_source
Box3 bounding_of(const std::vector<size_t> &idx, const std::vector<BuildPrim> &prims, size_t s, size_t e) {
  Box3 b = box_empty();
  for (size_t i = s; i < e; i++) expand_by_other(b, prims[idx[i]].bounding);
  return b;
}

// strategy.rs:66-87 BalanceTree::split + apply.rs:304-333 median_partition_at_axis
Split balance_split(const FNode &parent, const std::vector<BuildPrim> &prims, std::vector<size_t> &idx) {
  int axis = longest_axis(parent.bounding);
  size_t s = parent.start, e = parent.end;
  size_t middle = (e + s) / 2;
  size_t range_middle = (e - s) / 2;
  if (range_middle != 0) {
    // select_nth_unstable_by with partial_cmp().unwrap_or(Less): a NaN compares Less.
    auto less = [&](size_t a, size_t b) {
      float ca = axis_of(prims[a].center, axis), cb = axis_of(prims[b].center, axis);
      if (std::isnan(ca) || std::isnan(cb)) return true;
      return ca < cb;
    };
    std::nth_element(idx.begin() + s, idx.begin() + s + range_middle, idx.begin() + e, less);
  }
  Split r;
  r.ls = s;
  r.le = middle;
  r.rs = middle;
  r.re = e;
  r.lb = bounding_of(idx, prims, s, middle);
  r.rb = bounding_of(idx, prims, middle, e);
  return r;
}

// strategy.rs:201-284 SAH::split
Split sah_split(const FNode &parent, const std::vector<BuildPrim> &prims, std::vector<size_t> &idx,
                size_t nbins) {
  std::vector<Box3> bucket_box(nbins, box_empty());
  std::vector<std::vector<size_t>> bucket(nbins);
  int axis = longest_axis(parent.bounding);
  float a0 = axis_of(parent.bounding.min, axis);
  float a1 = axis_of(parent.bounding.max, axis);
  float step = (a1 - a0) / (float)nbins;
  for (size_t i = parent.start; i < parent.end; i++) {
    size_t index = idx[i];
    float value = axis_of(prims[index].center, axis);
    float f = std::floor((value - a0) / step);
    // Rust `as usize` saturates: NaN and negatives map to 0.
    size_t which;
    if (!(f > 0.0f)) which = 0;
    else if (f >= 18446744073709551615.0f) which = SIZE_MAX;
    else which = (size_t)f;
    if (which == nbins) which -= 1;  // edge case, strategy.rs:223-226
    if (which > nbins) which = nbins - 1;  // unreachable for finite bounds (the reference panics)
    expand_by_other(bucket_box[which], prims[index].bounding);
    bucket[which].push_back(index);
  }
  size_t empty = 0;
  for (auto &b : bucket) empty += b.empty() ? 1 : 0;
  if (empty == nbins - 1) return balance_split(parent, prims, idx);  // strategy.rs:121-128,230-233

  // step 2: partition decisions (strategy.rs:236-259)
  size_t best = 0;
  float best_cost = std::numeric_limits<float>::infinity();
  std::vector<Box3> lbox(nbins - 1), rbox(nbins - 1);
  std::vector<size_t> lcnt(nbins - 1), rcnt(nbins - 1);
  for (size_t i = 0; i + 1 < nbins; i++) {
    Box3 l = box_empty(), r = box_empty();
    size_t lc = 0, rc = 0;
    for (size_t k = 0; k <= i; k++) {
      lc += bucket[k].size();
      expand_by_other(l, bucket_box[k]);
    }
    for (size_t k = i + 1; k < nbins; k++) {
      rc += bucket[k].size();
      expand_by_other(r, bucket_box[k]);
    }
    float cost = box_surface_area(l) * (float)lc + box_surface_area(r) * (float)rc;
    lbox[i] = l;
    rbox[i] = r;
    lcnt[i] = lc;
    rcnt[i] = rc;
    if (cost < best_cost) {
      best_cost = cost;
      best = i;
    }
  }
  // step 3: reorder by bucket, stable within a bucket (strategy.rs:264-271)
  size_t ptr = parent.start;
  for (size_t b = 0; b < nbins; b++)
    for (size_t i : bucket[b]) idx[ptr++] = i;
  Split r;
  r.lb = lbox[best];
  r.ls = parent.start;
  r.le = parent.start + lcnt[best];
  r.rb = rbox[best];
  r.rs = parent.start + lcnt[best];
  r.re = parent.end;
  return r;
}

// strategy.rs:11-43 BVHBuildStrategy::build (recursive, DFS preorder)
size_t build_rec(const TreeBuildOption &opt, const std::vector<BuildPrim> &prims, std::vector<size_t> &idx,
                 std::vector<FNode> &nodes, size_t depth, size_t nbins, size_t &max_depth) {
  if (depth > max_depth) max_depth = depth;
  FNode node = nodes.back();
  if (!opt.should_continue(node.end - node.start, depth)) return 1;
  Split sp = sah_split(node, prims, idx, nbins);
  size_t node_index = nodes.size() - 1;
  nodes.push_back(FNode{sp.lb, sp.ls, sp.le, nodes.size(), false, 0});
  size_t left_count = build_rec(opt, prims, idx, nodes, depth + 1, nbins, max_depth);
  nodes.push_back(FNode{sp.rb, sp.rs, sp.re, nodes.size(), false, 0});
  size_t right_count = build_rec(opt, prims, idx, nodes, depth + 1, nbins, max_depth);
  nodes[node_index].has_child = true;
  nodes[node_index].left_count = left_count;
  return left_count + right_count + 1;
}

// mod.rs:55-79 FlattenBVH::new
FlattenBVH flatten_bvh_new(const std::vector<Box3> &boxes, const TreeBuildOption &opt, size_t nbins) {
  std::vector<BuildPrim> prims(boxes.size());
  std::vector<size_t> idx(boxes.size());
  for (size_t i = 0; i < boxes.size(); i++) {
    idx[i] = i;
    prims[i].bounding = boxes[i];
    prims[i].center = box_center(boxes[i]);
  }
  Box3 root = bounding_of(idx, prims, 0, idx.size());
  FlattenBVH bvh;
  bvh.nodes.push_back(FNode{root, 0, idx.size(), 0, false, 0});
  size_t maxd = 0;
  build_rec(opt, prims, idx, bvh.nodes, 0, nbins, maxd);
  bvh.sorted = std::move(idx);
  bvh.depth = maxd;
  return bvh;
}

// naive/mod.rs:611-632 compute_bvh_next
std::vector<std::pair<uint32_t, uint32_t>> compute_bvh_next(const std::vector<FNode> &nodes) {
  std::vector<std::pair<uint32_t, uint32_t>> result;
  std::vector<uint32_t> next_stack;
  for (const FNode &node : nodes) {
    if (!next_stack.empty() && next_stack.back() == (uint32_t)node.self_index) next_stack.pop_back();
    uint32_t miss = next_stack.empty() ? INVALID_NEXT : next_stack.back();
    if (node.has_child) {
      uint32_t hit = (uint32_t)(node.self_index + 1);
      next_stack.push_back((uint32_t)(node.self_index + node.left_count + 1));
      result.push_back({hit, miss});
    } else {
      result.push_back({miss, miss});
    }
  }
  return result;
}

// naive/mod.rs:338-367 flatten_bvh_to_gpu_node
orc_node to_gpu_node(const FNode &n, uint32_t hit, uint32_t miss, uint32_t next_offset, uint32_t prim_offset) {
  orc_node g;
  std::memset(&g, 0, sizeof(g));
  g.aabb_min[0] = n.bounding.min.x;
  g.aabb_min[1] = n.bounding.min.y;
  g.aabb_min[2] = n.bounding.min.z;
  g.aabb_max[0] = n.bounding.max.x;
  g.aabb_max[1] = n.bounding.max.y;
  g.aabb_max[2] = n.bounding.max.z;
  g.hit_next = hit != INVALID_NEXT ? hit + next_offset : INVALID_NEXT;
  g.miss_next = miss != INVALID_NEXT ? miss + next_offset : INVALID_NEXT;
  g.content_range[0] = (uint32_t)n.start + prim_offset;
  g.content_range[1] = (uint32_t)n.end + prim_offset;
  return g;
}

// ------------------------------------------------------------ flags --------
// shader/ray-tracing/src/api/ty.rs:102-159, naive/flag.rs
enum : uint32_t {
  F_FORCE_OPAQUE = 0x01,
  F_FORCE_NON_OPAQUE = 0x02,
  F_ACCEPT_FIRST_HIT_AND_END_SEARCH = 0x04,
  F_CULL_BACK = 0x10,
  F_CULL_FRONT = 0x20,
  F_CULL_OPAQUE = 0x40,
  F_CULL_NON_OPAQUE = 0x80,
  F_SKIP_TRIANGLES = 0x100,
  F_TRIANGLE_FLIP_FACING = 0x400,
  GI_CULL_DISABLE = 0x1,
  GI_FLIP_FACING = 0x2,
  GI_FORCE_OPAQUE = 0x4,
  GI_FORCE_NO_OPAQUE = 0x8,
  GEOMETRY_FLAG_OPAQUE = 0x1,
};
// flag.rs:30-75
uint32_t merge_instance_flags(uint32_t f, uint32_t gi) {
  if (gi & GI_CULL_DISABLE) f &= ~(F_CULL_BACK | F_CULL_FRONT);
  if (gi & GI_FLIP_FACING) f ^= F_TRIANGLE_FLIP_FACING;
  if (gi & GI_FORCE_OPAQUE) f |= F_FORCE_OPAQUE;
  if (gi & GI_FORCE_NO_OPAQUE) f |= F_FORCE_NON_OPAQUE;
  return f;
}
// flag.rs:78-91
void cull_geometry(uint32_t f, uint32_t gflags, bool &pass, bool &is_opaque) {
  bool geometry_opaque = (gflags & GEOMETRY_FLAG_OPAQUE) != 0;
  bool force_opaque = (f & F_FORCE_OPAQUE) != 0, force_non_opaque = (f & F_FORCE_NON_OPAQUE) != 0;
  bool cull_opaque = (f & F_CULL_OPAQUE) != 0, cull_non_opaque = (f & F_CULL_NON_OPAQUE) != 0;
  is_opaque = (geometry_opaque || force_opaque) && !force_non_opaque;
  pass = (is_opaque && !cull_opaque) || (!is_opaque && !cull_non_opaque);
}
// flag.rs:94-105
void cull_triangle(uint32_t f, bool &cull_enable, bool &cull_back) {
  bool flip = (f & F_TRIANGLE_FLIP_FACING) != 0;
  bool cf = (f & F_CULL_FRONT) != 0, cb = (f & F_CULL_BACK) != 0;
  cull_enable = cf || cb;
  cull_back = (flip && cb) || (!flip && cf);
}

// ------------------------------------------------------ intersections ------
// geometry/mod.rs:44-64
bool intersect_aabb(V3 o, V3 d, float tmin, float tmax, V3 bmin, V3 bmax) {
  V3 inv_d = v3(1.0f, 1.0f, 1.0f) / d;
  V3 t0 = (bmin - o) * inv_d;
  V3 t1 = (bmax - o) * inv_d;
  V3 tn = vmin(t0, t1);
  V3 tf = vmax(t0, t1);
  float tnm = max_channel(tn);
  float tfm = min_channel(tf);
  return tnm <= tfm && tmin < tfm && tnm < tmax;
}

// geometry/mod.rs:104-155.  Returns (sign, t, u, v); miss = zeros.
// b == 0 is a miss (the reference GPU path's sign(0) = 0); the reference CPU
// path would go on with an infinite or NaN t there.
void intersect_triangle(V3 o, V3 d, float rnear, float rfar, V3 v0, V3 v1, V3 v2, bool cull_enable,
                        bool cull_back, float out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0f;
  V3 e1 = v1 - v0;
  V3 e2 = v2 - v0;
  V3 normal = normalize(cross(e1, e2));
  float b = dot(normal, d);
  if (b == 0.0f) return;
  float sign = b > 0.0f ? 1.0f : -1.0f;  // f32::signum for b != 0
  if (cull_enable) {
    bool pass = cull_back != (b < 0.0f);
    if (!pass) return;
  }
  V3 w0 = o - v0;
  float a = -dot(normal, w0);
  float t = a / b;
  if (t < rnear || t > rfar) return;
  V3 p = o + t * d;
  float uu = dot(e1, e1);
  float uv = dot(e1, e2);
  float vv = dot(e2, e2);
  V3 w = p - v0;
  float wu = dot(w, e1);
  float wv = dot(w, e2);
  float inverse_d = 1.0f / (uv * uv - uu * vv);
  float u = (uv * wv - vv * wu) * inverse_d;
  if (u < 0.0f || u > 1.0f) return;
  float v = (uv * wu - uu * wv) * inverse_d;
  if (v < 0.0f || (u + v) > 1.0f) return;
  out[0] = sign;
  out[1] = t;
  out[2] = u;
  out[3] = v;
}

}  // namespace

// any-hit behaviour bits (naive/mod.rs ANYHIT_BEHAVIOR_*): accept the hit,
// end the search
constexpr uint32_t kAnyHitAccept = 1u, kAnyHitEndSearch = 2u;

// ------------------------------------------------------------- scene -------
struct Geometry {
  std::vector<float> pos;
  std::vector<uint32_t> idx;
  bool indexed;
  uint32_t ntris, flags;
  std::vector<float> normals;         // per vertex (path tracer), optional
  std::vector<uint32_t> tri_material;  // per triangle (path tracer), optional
  std::vector<float> uvs;             // per vertex (hit attributes), optional
  std::vector<uint8_t> anyhit;        // per triangle: any-hit behaviour of non-opaque hits, optional
};

struct orc_scene {
  orc_build_opts opts;
  std::vector<bool> blas_live;
  std::vector<std::vector<Geometry>> blas_data;
  std::vector<Geometry> pending;
  std::vector<bool> tlas_live;
  std::vector<std::vector<orc_instance>> tlas_data_src;
  std::vector<uint32_t> binding;

  // built (naive/traverse_cpu.rs:9-35)
  std::vector<uint32_t> tlas_bvh_root;
  std::vector<orc_node> tlas_forest;
  std::vector<M4> inst_transform_inv;
  std::vector<uint32_t> inst_custom, inst_sbt, inst_flags, inst_blas, inst_order;
  std::vector<Box3> tlas_bounding_box;
  std::vector<uint32_t> tlas_bounding_mask, tlas_bounding_flags;
  std::vector<std::pair<uint32_t, uint32_t>> blas_meta;
  struct GeoMeta {
    uint32_t root, geometry_idx, primitive_start, flags;
  };
  std::vector<GeoMeta> geo_meta;
  std::vector<orc_node> tri_forest;
  std::vector<uint32_t> redirect, indices;
  std::vector<V3> vertices;
  uint32_t max_blas_depth = 0, max_tlas_depth = 0;
  bool built = false;
  // path-tracer shading data
  std::vector<V3> normals;            // parallel to vertices; zero = none
  std::vector<float> uvs;             // 2 per vertex, parallel to vertices (zeros = none)
  bool has_uvs = false;
  std::vector<uint32_t> tri_material;  // per global triangle
  std::vector<uint8_t> anyhit;         // per global triangle (empty: every non-opaque hit is accepted)
  std::vector<M4> inst_transform;      // object-to-world of each instance slot (as given)
  std::vector<uint32_t> tlas_inst_first;  // first instance slot of each TLAS
  std::vector<float> materials;        // 8 floats each: albedo xyz, pad, emission xyz, pad
};

// The any-hit behaviour of a non-opaque hit on global triangle `tri`: the
// caller's table, or accept (the reference tests' TEST_ANYHIT_BEHAVIOR).
inline uint32_t anyhit_of(const orc_scene *s, uint32_t tri) {
  return tri < s->anyhit.size() ? (uint32_t)s->anyhit[tri] : kAnyHitAccept;
}


extern "C" {

orc_scene *orc_scene_new(const orc_build_opts *opts) {
  orc_scene *s = new orc_scene();
  if (opts) s->opts = *opts;
  else s->opts = orc_build_opts{2, 50, 10, 50, 4};  // naive/mod.rs:173-177,280-284
  if (s->opts.sah_bins < 2) s->opts.sah_bins = 4;
  return s;
}
void orc_scene_free(orc_scene *s) { delete s; }

void orc_blas_begin(orc_scene *s) { s->pending.clear(); }
void orc_blas_add_triangles(orc_scene *s, const float *pos, uint32_t nverts, const uint32_t *indices,
                            uint32_t ntris, uint32_t gflags) {
  Geometry g;
  g.pos.assign(pos, pos + 3 * (size_t)nverts);
  g.indexed = indices != nullptr;
  if (indices) g.idx.assign(indices, indices + 3 * (size_t)ntris);
  else ntris = nverts / 3;
  g.ntris = ntris;
  g.flags = gflags;
  s->pending.push_back(std::move(g));
}
uint32_t orc_blas_end(orc_scene *s) {
  s->blas_data.push_back(std::move(s->pending));
  s->blas_live.push_back(true);
  s->pending.clear();
  s->built = false;
  return (uint32_t)(s->blas_data.size() - 1);
}
void orc_blas_delete(orc_scene *s, uint32_t h) {
  if (h < s->blas_live.size()) s->blas_live[h] = false;
  s->built = false;
}
uint32_t orc_tlas_create(orc_scene *s, const orc_instance *inst, uint32_t n) {
  s->tlas_data_src.emplace_back(inst, inst + n);
  s->tlas_live.push_back(true);
  s->built = false;
  return (uint32_t)(s->tlas_data_src.size() - 1);
}
void orc_tlas_delete(orc_scene *s, uint32_t h) {
  if (h < s->tlas_live.size()) s->tlas_live[h] = false;
  s->built = false;
}
void orc_bind_tlas(orc_scene *s, const uint32_t *h, uint32_t n) {
  s->binding.assign(h, h + n);
  s->built = false;
}

// naive/mod.rs:122-494
int orc_build(orc_scene *s) {
  s->tri_forest.clear();
  s->redirect.clear();
  s->indices.clear();
  s->vertices.clear();
  s->normals.clear();
  s->uvs.clear();
  s->has_uvs = false;
  s->inst_transform.clear();
  s->tlas_inst_first.clear();
  s->tri_material.clear();
  s->anyhit.clear();
  s->blas_meta.clear();
  s->geo_meta.clear();
  s->tlas_bvh_root.clear();
  s->tlas_forest.clear();
  s->inst_transform_inv.clear();
  s->inst_custom.clear();
  s->inst_sbt.clear();
  s->inst_flags.clear();
  s->inst_blas.clear();
  s->inst_order.clear();
  s->tlas_bounding_box.clear();
  s->tlas_bounding_mask.clear();
  s->tlas_bounding_flags.clear();
  s->max_blas_depth = s->max_tlas_depth = 0;

  TreeBuildOption blas_opt{s->opts.blas_max_depth, s->opts.blas_bin_size};
  TreeBuildOption tlas_opt{s->opts.tlas_max_depth, s->opts.tlas_bin_size};
  size_t nbins = s->opts.sah_bins;

  // build_blas (mod.rs:122-260).  The per-BLAS box is the union of its
  // geometries' root boxes; an empty or deleted BLAS has no box.
  std::vector<bool> blas_has_box;
  std::vector<Box3> blas_box;
  struct TriBvh {
    std::vector<orc_node> nodes;
    uint32_t primitive_start, geometry_idx, flags;
  };
  std::vector<TriBvh> tri_bvh;
  // Each geometry's tree depends on that geometry alone: build the trees
  // concurrently, then assemble them in the reference order below.
  std::vector<std::pair<size_t, size_t>> jobs;
  for (size_t bi = 0; bi < s->blas_data.size(); bi++)
    if (s->blas_live[bi])
      for (size_t gi = 0; gi < s->blas_data[bi].size(); gi++) jobs.push_back({bi, gi});
  std::vector<FlattenBVH> trees(jobs.size());
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int64_t j = 0; j < (int64_t)jobs.size(); j++) {
    const Geometry &g = s->blas_data[jobs[j].first][jobs[j].second];
    size_t nv = g.pos.size() / 3;
    size_t ntri = g.indexed ? g.idx.size() / 3 : nv / 3;
    std::vector<Box3> boxes(ntri);
    for (size_t t = 0; t < ntri; t++) {
      Box3 b = box_empty();
      for (int k = 0; k < 3; k++) {
        size_t vi = g.indexed ? g.idx[3 * t + k] : 3 * t + k;
        const float *p = &g.pos[3 * vi];
        expand_by_point(b, v3(p[0], p[1], p[2]));
      }
      boxes[t] = b;
    }
    trees[j] = flatten_bvh_new(boxes, blas_opt, nbins);
  }
  size_t next_tree = 0;
  for (size_t bi = 0; bi < s->blas_data.size(); bi++) {
    if (!s->blas_live[bi]) {
      s->blas_meta.push_back({0, 0});
      blas_has_box.push_back(false);
      blas_box.push_back(box_empty());
      continue;
    }
    uint32_t tri_start = (uint32_t)tri_bvh.size();
    Box3 blas_root_box = box_empty();
    for (size_t gi = 0; gi < s->blas_data[bi].size(); gi++) {
      const Geometry &g = s->blas_data[bi][gi];
      uint32_t primitive_start = (uint32_t)(s->indices.size() / 3);
      uint32_t vertex_start = (uint32_t)s->vertices.size();
      size_t nv = g.pos.size() / 3;
      for (size_t v = 0; v < nv; v++) s->vertices.push_back(v3(g.pos[3 * v], g.pos[3 * v + 1], g.pos[3 * v + 2]));
      std::vector<uint32_t> tri_idx;
      if (g.indexed) tri_idx = g.idx;
      else {
        tri_idx.resize(nv);
        for (size_t v = 0; v < nv; v++) tri_idx[v] = (uint32_t)v;
      }
      size_t ntri = tri_idx.size() / 3;
      FlattenBVH bvh = std::move(trees[next_tree++]);
      if (bvh.depth > s->max_blas_depth) s->max_blas_depth = (uint32_t)bvh.depth;
      Box3 root_box = box_empty();
      expand_by_other(root_box, bvh.nodes[0].bounding);
      expand_by_other(blas_root_box, root_box);
      auto links = compute_bvh_next(bvh.nodes);
      for (size_t k : bvh.sorted) s->redirect.push_back(primitive_start + (uint32_t)k);
      for (size_t t = 0; t < ntri; t++)
        for (int k = 0; k < 3; k++) s->indices.push_back(vertex_start + tri_idx[3 * t + k]);
      for (size_t v = 0; v < nv; v++)
        s->normals.push_back(g.normals.size() == 3 * nv ? v3(g.normals[3 * v], g.normals[3 * v + 1], g.normals[3 * v + 2])
                                                        : v3(0.0f, 0.0f, 0.0f));
      for (size_t v = 0; v < nv; v++) {
        const bool has = g.uvs.size() == 2 * nv;
        s->uvs.push_back(has ? g.uvs[2 * v] : 0.0f);
        s->uvs.push_back(has ? g.uvs[2 * v + 1] : 0.0f);
        s->has_uvs = s->has_uvs || has;
      }
      for (size_t t = 0; t < ntri; t++) s->tri_material.push_back(t < g.tri_material.size() ? g.tri_material[t] : 0u);
      for (size_t t = 0; t < ntri; t++) s->anyhit.push_back(t < g.anyhit.size() ? g.anyhit[t] : kAnyHitAccept);
      TriBvh tb;
      tb.primitive_start = primitive_start;
      tb.geometry_idx = (uint32_t)gi;
      tb.flags = g.flags;
      // keep node-local links; rebased below when placed in the forest
      for (size_t k = 0; k < bvh.nodes.size(); k++) tb.nodes.push_back(to_gpu_node(bvh.nodes[k], links[k].first, links[k].second, 0, 0));
      tri_bvh.push_back(std::move(tb));
    }
    s->blas_meta.push_back({tri_start, (uint32_t)tri_bvh.size()});
    blas_has_box.push_back(!s->blas_data[bi].empty());
    blas_box.push_back(blas_root_box);
  }
  // tri forest (mod.rs:368-385)
  for (const TriBvh &tb : tri_bvh) {
    uint32_t bvh_start = (uint32_t)s->tri_forest.size();
    s->geo_meta.push_back({bvh_start, tb.geometry_idx, tb.primitive_start, tb.flags});
    for (orc_node n : tb.nodes) {
      if (n.hit_next != INVALID_NEXT) n.hit_next += bvh_start;
      if (n.miss_next != INVALID_NEXT) n.miss_next += bvh_start;
      n.content_range[0] += tb.primitive_start;
      n.content_range[1] += tb.primitive_start;
      s->tri_forest.push_back(n);
    }
  }

  // build_tlas (mod.rs:262-320) + forest (mod.rs:435-454)
  for (size_t ti = 0; ti < s->tlas_data_src.size(); ti++) {
    if (!s->tlas_live[ti]) {
      s->tlas_bvh_root.push_back(INVALID_NEXT);
      continue;
    }
    const auto &src = s->tlas_data_src[ti];
    uint32_t bvh_start = (uint32_t)s->tlas_forest.size();
    uint32_t primitive_start = (uint32_t)s->inst_blas.size();
    while (s->tlas_inst_first.size() < ti) s->tlas_inst_first.push_back(primitive_start);
    s->tlas_inst_first.push_back(primitive_start);
    std::vector<Box3> aabbs;
    for (const orc_instance &in : src) {
      if (in.blas >= blas_box.size() || !blas_has_box[in.blas]) return -1;  // reference: unwrap() panics
      aabbs.push_back(box_apply_matrix(blas_box[in.blas], m4_from(in.transform)));
    }
    FlattenBVH bvh = flatten_bvh_new(aabbs, tlas_opt, nbins);
    if (bvh.depth > s->max_tlas_depth) s->max_tlas_depth = (uint32_t)bvh.depth;
    auto links = compute_bvh_next(bvh.nodes);
    for (size_t k : bvh.sorted) {
      const orc_instance &in = src[k];
      M4 m = m4_from(in.transform);
      uint32_t flags = in.flags;
      if (m3_det(m) < 0.0f) flags ^= GI_FLIP_FACING;
      s->inst_transform_inv.push_back(m4_inverse_or_identity(m));
      s->inst_transform.push_back(m);
      s->inst_custom.push_back(in.instance_custom_index);
      s->inst_sbt.push_back(in.sbt_offset);
      s->inst_flags.push_back(flags);
      s->inst_blas.push_back(in.blas);
      s->inst_order.push_back((uint32_t)k);
      s->tlas_bounding_box.push_back(aabbs[k]);
      s->tlas_bounding_mask.push_back(in.mask);
      s->tlas_bounding_flags.push_back(flags);
    }
    for (size_t k = 0; k < bvh.nodes.size(); k++)
      s->tlas_forest.push_back(to_gpu_node(bvh.nodes[k], links[k].first, links[k].second, bvh_start, primitive_start));
    s->tlas_bvh_root.push_back(bvh_start);
  }
  while (s->tlas_inst_first.size() < s->tlas_data_src.size()) s->tlas_inst_first.push_back((uint32_t)s->inst_blas.size());
  s->tlas_inst_first.push_back((uint32_t)s->inst_blas.size());  // end of the last TLAS
  s->built = true;
  return 0;
}

void orc_get_sizes(const orc_scene *s, orc_sizes *o) {
  o->tri_nodes = (uint32_t)s->tri_forest.size();
  o->tlas_nodes = (uint32_t)s->tlas_forest.size();
  o->tris = (uint32_t)s->redirect.size();
  o->verts = (uint32_t)s->vertices.size();
  o->instances = (uint32_t)s->inst_blas.size();
  o->geometries = (uint32_t)s->geo_meta.size();
  o->blas = (uint32_t)s->blas_meta.size();
  o->tlas = (uint32_t)s->tlas_bvh_root.size();
  o->max_blas_depth = s->max_blas_depth;
  o->max_tlas_depth = s->max_tlas_depth;
}
void orc_copy_tri_forest(const orc_scene *s, orc_node *out) {
  if (!s->tri_forest.empty()) std::memcpy(out, s->tri_forest.data(), s->tri_forest.size() * sizeof(orc_node));
}
void orc_copy_tlas_forest(const orc_scene *s, orc_node *out) {
  if (!s->tlas_forest.empty()) std::memcpy(out, s->tlas_forest.data(), s->tlas_forest.size() * sizeof(orc_node));
}
void orc_copy_indices_redirect(const orc_scene *s, uint32_t *out) {
  if (!s->redirect.empty()) std::memcpy(out, s->redirect.data(), s->redirect.size() * 4);
}
void orc_copy_instance_order(const orc_scene *s, uint32_t *out) {
  if (!s->inst_order.empty()) std::memcpy(out, s->inst_order.data(), s->inst_order.size() * 4);
}
void orc_copy_geometry_meta(const orc_scene *s, uint32_t *out) {
  for (size_t i = 0; i < s->geo_meta.size(); i++) {
    out[4 * i + 0] = s->geo_meta[i].root;
    out[4 * i + 1] = s->geo_meta[i].geometry_idx;
    out[4 * i + 2] = s->geo_meta[i].primitive_start;
    out[4 * i + 3] = s->geo_meta[i].flags;
  }
}
void orc_copy_tlas_bounding(const orc_scene *s, float *out) {
  for (size_t i = 0; i < s->tlas_bounding_box.size(); i++) {
    const Box3 &b = s->tlas_bounding_box[i];
    out[8 * i + 0] = b.min.x;
    out[8 * i + 1] = b.min.y;
    out[8 * i + 2] = b.min.z;
    std::memcpy(&out[8 * i + 3], &s->tlas_bounding_mask[i], 4);
    out[8 * i + 4] = b.max.x;
    out[8 * i + 5] = b.max.y;
    out[8 * i + 6] = b.max.z;
    std::memcpy(&out[8 * i + 7], &s->tlas_bounding_flags[i], 4);
  }
}
void orc_copy_transform_inv(const orc_scene *s, float *out) {
  for (size_t i = 0; i < s->inst_transform_inv.size(); i++)
    std::memcpy(out + 16 * i, &s->inst_transform_inv[i], 64);
}

}  // extern "C"

namespace {

struct Counters {
  uint32_t nodes = 0, tris = 0, inst = 0;
};

// traverse_cpu.rs:283-319 -- stackless iterator: yields leaf node indices.
struct BvhIter {
  const std::vector<orc_node> *bvh;
  V3 o, d;
  float near_, scaling;
  const float *far_;  // shared far (RayRange.far is an Rc<Cell<f32>>)
  uint32_t curr;
  Counters *cnt;
  bool next(uint32_t &out) {
    while (curr != INVALID_NEXT) {
      cnt->nodes++;
      const orc_node &n = (*bvh)[curr];
      // RayRange::get (traverse_cpu.rs:277-280)
      float rn = near_ * scaling, rf = *far_ * scaling;
      if (intersect_aabb(o, d, rn, rf, v3(n.aabb_min[0], n.aabb_min[1], n.aabb_min[2]),
                         v3(n.aabb_max[0], n.aabb_max[1], n.aabb_max[2]))) {
        uint32_t c = curr;
        curr = n.hit_next;
        if (n.hit_next == n.miss_next) {
          out = c;
          return true;
        }
      } else {
        curr = n.miss_next;
      }
    }
    return false;
  }
};

struct RayIn {
  V3 o, d;
  float tmin, tmax;
  uint32_t flags, mask, tlas_idx;
};

// traverse_cpu.rs:52-245
bool traverse_one(const orc_scene *s, const RayIn &ray, orc_hit &hit, orc_hit_meta &hm, Counters &cnt,
                  uint32_t *ext = nullptr) {
  bool found = false;
  uint32_t flags0 = ray.flags;
  float far_ = ray.tmax;  // RayRange::new(range.x, range.y, 1.)
  hit.prim_id = 0xffffffffu;
  hit.t = hit.u = hit.v = 0.0f;
  hm.geometry_id = hm.instance_id = hm.instance_custom_index = 0xffffffffu;
  hm.hit_kind = 0;
  if (ray.tlas_idx >= s->binding.size()) return false;
  uint32_t th = s->binding[ray.tlas_idx];
  if (th >= s->tlas_bvh_root.size()) return false;
  uint32_t root = s->tlas_bvh_root[th];
  BvhIter tlas_iter{&s->tlas_forest, ray.o, ray.d, ray.tmin, 1.0f, &far_, root, &cnt};
  uint32_t leaf;
  while (tlas_iter.next(leaf)) {
    const orc_node &node = s->tlas_forest[leaf];
    for (uint32_t ti = node.content_range[0]; ti < node.content_range[1]; ti++) {
      cnt.inst++;
      // the instance bounding uses the original ray range (traverse_cpu.rs:80-86)
      const Box3 &tb = s->tlas_bounding_box[ti];
      if (!intersect_aabb(ray.o, ray.d, ray.tmin, ray.tmax, tb.min, tb.max)) continue;
      if ((ray.mask & s->tlas_bounding_mask[ti]) == 0) continue;
      uint32_t blas_idx = s->inst_blas[ti];
      uint32_t flags = merge_instance_flags(flags0, s->inst_flags[ti]);
      const M4 &inv = s->inst_transform_inv[ti];
      V3 bo = m4_mul_point(inv, ray.o);
      V3 bd = m3_mul(inv, ray.d);
      float scaling = length(bd);
      bd = normalize(bd);
      auto bm = s->blas_meta[blas_idx];
      if ((flags & F_SKIP_TRIANGLES) != 0) continue;
      for (uint32_t gr = bm.first; gr < bm.second; gr++) {
        const auto &geo = s->geo_meta[gr];
        bool pass, is_opaque;
        cull_geometry(flags, geo.flags, pass, is_opaque);
        if (!pass) continue;
        bool cull_enable, cull_back;
        cull_triangle(flags, cull_enable, cull_back);
        BvhIter it{&s->tri_forest, bo, bd, ray.tmin, scaling, &far_, geo.root, &cnt};
        uint32_t hl;
        while (it.next(hl)) {
          const orc_node &ln = s->tri_forest[hl];
          for (uint32_t k = ln.content_range[0]; k < ln.content_range[1]; k++) {
            uint32_t tri = s->redirect[k];
            V3 v0 = s->vertices[s->indices[3 * (size_t)tri]];
            V3 v1 = s->vertices[s->indices[3 * (size_t)tri + 1]];
            V3 v2 = s->vertices[s->indices[3 * (size_t)tri + 2]];
            cnt.tris++;
            float r[4];
            intersect_triangle(bo, bd, ray.tmin * scaling, far_ * scaling, v0, v1, v2, cull_enable, cull_back, r);
            if (r[0] != 0.0f) {
              float distance = r[1] / scaling;
              // opaque hits commit; a non-opaque one takes its any-hit
              // behaviour (traverse_cpu.rs:164-185)
              uint32_t behavior = is_opaque ? kAnyHitAccept : anyhit_of(s, tri);
              if (behavior & kAnyHitAccept) {
                far_ = distance;  // update_far
                found = true;
                hit.prim_id = tri - geo.primitive_start;
                hit.t = distance;
                hit.u = r[2];
                hit.v = r[3];
                hm.geometry_id = geo.geometry_idx;
                hm.instance_id = s->inst_order[ti];
                hm.instance_custom_index = s->inst_custom[ti];
                hm.hit_kind = r[0] < 0.0f ? 0xFFu : 0xFEu;  // ty.rs:139-140
                if (ext) {
                  ext[0] = gr;  // global geometry index
                  ext[1] = ti;  // instance index in forest order
                }
                if (flags & F_ACCEPT_FIRST_HIT_AND_END_SEARCH) behavior |= kAnyHitEndSearch;
              }
              if (behavior & kAnyHitEndSearch) return found;
            }
          }
        }
      }
    }
  }
  return found;
}

RayIn load_ray(const orc_ray &r, const orc_ray_meta *m, uint32_t flags, uint32_t mask, uint32_t tlas) {
  RayIn ri;
  ri.o = v3(r.ox, r.oy, r.oz);
  ri.d = v3(r.dx, r.dy, r.dz);
  ri.tmin = r.tmin;
  ri.tmax = r.tmax;
  ri.flags = m ? m->flags : flags;
  ri.mask = m ? m->cull_mask : mask;
  ri.tlas_idx = m ? m->tlas_idx : tlas;
  return ri;
}

}  // namespace

extern "C" {

int orc_trace(const orc_scene *s, const orc_ray *rays, const orc_ray_meta *meta, uint32_t n, uint32_t flags,
              uint32_t cull_mask, uint32_t tlas_idx, orc_hit *hits, orc_hit_meta *hit_meta, uint32_t *counters,
              int nthreads) {
  if (!s->built) return -1;
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nt)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    RayIn ri = load_ray(rays[i], meta ? &meta[i] : nullptr, flags, cull_mask, tlas_idx);
    orc_hit h;
    orc_hit_meta hm;
    Counters c;
    traverse_one(s, ri, h, hm, c);
    hits[i] = h;
    if (hit_meta) hit_meta[i] = hm;
    if (counters) {
      counters[3 * i + 0] = c.nodes;
      counters[3 * i + 1] = c.tris;
      counters[3 * i + 2] = c.inst;
    }
  }
  return 0;
}

int orc_trace_brute(const orc_scene *s, const orc_ray *rays, uint32_t n, uint32_t flags, uint32_t cull_mask,
                    uint32_t tlas_idx, orc_hit *hits, int nthreads) {
  if (!s->built) return -1;
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    RayIn ray = load_ray(rays[i], nullptr, flags, cull_mask, tlas_idx);
    orc_hit h;
    h.prim_id = 0xffffffffu;
    h.t = h.u = h.v = 0.0f;
    float far_ = ray.tmax;
    uint32_t th = ray.tlas_idx < s->binding.size() ? s->binding[ray.tlas_idx] : 0xffffffffu;
    if (th < s->tlas_bvh_root.size() && s->tlas_bvh_root[th] != INVALID_NEXT) {
      const orc_node &root = s->tlas_forest[s->tlas_bvh_root[th]];
      // instances of this TLAS occupy the root's content range
      for (uint32_t ti = root.content_range[0]; ti < root.content_range[1]; ti++) {
        if ((ray.mask & s->tlas_bounding_mask[ti]) == 0) continue;
        uint32_t fl = merge_instance_flags(ray.flags, s->inst_flags[ti]);
        if (fl & F_SKIP_TRIANGLES) continue;
        const M4 &inv = s->inst_transform_inv[ti];
        V3 bo = m4_mul_point(inv, ray.o);
        V3 bd = m3_mul(inv, ray.d);
        float scaling = length(bd);
        bd = normalize(bd);
        auto bm = s->blas_meta[s->inst_blas[ti]];
        for (uint32_t gr = bm.first; gr < bm.second; gr++) {
          const auto &geo = s->geo_meta[gr];
          bool pass, opq, ce, cb;
          cull_geometry(fl, geo.flags, pass, opq);
          if (!pass) continue;
          cull_triangle(fl, ce, cb);
          const orc_node &gn = s->tri_forest[geo.root];
          for (uint32_t k = gn.content_range[0]; k < gn.content_range[1]; k++) {
            uint32_t tri = geo.primitive_start + (k - gn.content_range[0]);  // source order
            V3 v0 = s->vertices[s->indices[3 * (size_t)tri]];
            V3 v1 = s->vertices[s->indices[3 * (size_t)tri + 1]];
            V3 v2 = s->vertices[s->indices[3 * (size_t)tri + 2]];
            float r[4];
            intersect_triangle(bo, bd, ray.tmin * scaling, far_ * scaling, v0, v1, v2, ce, cb, r);
            // ignored non-opaque hits are skipped (end-search behaviours depend on visit order: not here)
            if (r[0] != 0.0f && (opq || (anyhit_of(s, tri) & kAnyHitAccept))) {
              far_ = r[1] / scaling;
              h.prim_id = tri - geo.primitive_start;
              h.t = far_;
              h.u = r[2];
              h.v = r[3];
            }
          }
        }
      }
    }
    hits[i] = h;
  }
  return 0;
}

void orc_intersect_triangle(const float o[3], const float d[3], float n, float f, const float a[3], const float b[3],
                            const float c[3], int ce, int cb, float out[4]) {
  intersect_triangle(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), n, f, v3(a[0], a[1], a[2]), v3(b[0], b[1], b[2]),
                     v3(c[0], c[1], c[2]), ce != 0, cb != 0, out);
}
int orc_intersect_aabb(const float o[3], const float d[3], float n, float f, const float bmin[3], const float bmax[3]) {
  return intersect_aabb(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), n, f, v3(bmin[0], bmin[1], bmin[2]),
                        v3(bmax[0], bmax[1], bmax[2]))
             ? 1
             : 0;
}

// sobol.rs:40-51
float orc_van_der_corput(uint32_t n, uint32_t scramble) {
  n = (n >> 16) | (n << 16);  // rotate_right(16)
  n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
  n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
  n = ((n & 0x33333333u) << 2) | ((n & 0xccccccccu) >> 2);
  n = ((n & 0x55555555u) << 1) | ((n & 0xaaaaaaaau) >> 1);
  n ^= scramble;
  return rmin((float)((n >> 8) & 0xffffffu) / 16777216.0f, 1.0f - FLT_EPSILON);
}
// sobol.rs:55-68
float orc_sobol(uint32_t n, uint32_t scramble) {
  uint32_t i = 1u << 31;
  while (n != 0) {
    if (n & 1u) scramble ^= i;
    n >>= 1;
    i ^= i >> 1;
  }
  return rmin((float)((scramble >> 8) & 0xffffffu) / 16777216.0f, 1.0f - FLT_EPSILON);
}

// stream_compaction.rs:3-45: inclusive scan of the mask; element i is kept
// where the scan steps, and lands at inclusive[i] - 1.
uint32_t orc_compact(const uint32_t *values, const uint32_t *mask, uint32_t n, uint32_t *out) {
  std::vector<uint32_t> inclusive(n);
  uint32_t acc = 0;
  for (uint32_t i = 0; i < n; i++) {
    acc += mask[i] ? 1u : 0u;
    inclusive[i] = acc;
  }
  for (uint32_t i = 0; i < n; i++) {
    uint32_t prev = i == 0 ? 0u : inclusive[i - 1];
    if (inclusive[i] != prev) out[prev] = values ? values[i] : i;
  }
  return acc;
}

}  // extern "C"

// ===================================================================
// Path tracer oracle: the estimator of the wavefront path tracer in FP64.
// Visibility (closest hit and shadow rays) uses the reference FP32
// traversal above on FP32 rays; sampling integers are bit-identical to the
// device; all shading arithmetic is double precision.
// Stage structure: scene/rendering/gpu-ray-tracing/src/feature/path_tracing/
// ray_gen.rs:89-213 (ray generation), ray_hit.rs:6-104 (closest hit: NEE
// shadow ray + BSDF sample), ray_miss.rs:136-158 (miss).
// ===================================================================
namespace {

struct D3 {
  double x, y, z;
};
inline D3 d3(double x, double y, double z) { return D3{x, y, z}; }
inline D3 d3f(V3 v) { return D3{v.x, v.y, v.z}; }
inline D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
inline D3 operator*(D3 a, D3 b) { return d3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 dcross(D3 a, D3 b) { return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline D3 dnormalize(D3 a) {
  double m = ddot(a, a);
  return m > 0.0 ? a * (1.0 / std::sqrt(m)) : a;
}

// scene/rendering/gpu-ray-tracing/src/sampler.rs:33-42
uint32_t xxhash32(uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t p1 = 2246822519u, p2 = 3266489917u, p3 = 668265263u, p4 = 374761393u;
  uint32_t h = z + p4 + x * p2;
  h = p3 * ((h << 17) | (h >> 15));
  h = h + y * p2;
  h = p3 * ((h << 17) | (h >> 15));
  h = p1 * (h ^ (h >> 15));
  h = p2 * (h ^ (h >> 13));
  return h ^ (h >> 16);
}
uint32_t scramble_of(uint32_t px, uint32_t py, uint32_t seed, uint32_t k) { return xxhash32(px, py, seed * 65536u + k); }

// ray_util.rs:6-40, on FP32 values (it moves a float by integer ulps)
float offset_axis(float p, float n) {
  const float int_scale = 256.0f, float_scale = 1.0f / 65536.0f, origin = 1.0f / 32.0f;
  int32_t of_i = (int32_t)(n * int_scale);
  int32_t pi;
  std::memcpy(&pi, &p, 4);
  pi += p < 0.0f ? -of_i : of_i;
  float pf;
  std::memcpy(&pf, &pi, 4);
  return std::fabs(p) < origin ? p + float_scale * n : pf;
}


// ---- the path's geometry in FP32 ----
// Rays, sampled directions and every decision of a path (facing, light
// gating, continuation, roulette) are computed in FP32 with the device's
// operation order, as the reference computes them in f32; only radiance,
// throughput and gradients are carried in FP64.  The device and the oracle
// then trace the same paths, and FP32-vs-FP64 differences show only in the
// shading arithmetic.

// sin / cos on [-pi/4, pi/4]: Taylor series through x^9 / x^10, Horner in FP32
void f_sincos_quarter(float x, float &s, float &c) {
  const float x2 = x * x;
  s = x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f + x2 * (1.0f / 362880.0f)))));
  c = 1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f + x2 * (-1.0f / 720.0f + x2 * (1.0f / 40320.0f +
                                                                                 x2 * (-1.0f / 3628800.0f)))));
}
// distribution_map.rs:10-28 (second wedge: angle pi/2 - phi, cos / sin swap)
void f_concentric_disk(float s0, float s1, float &x, float &y) {
  const float PI_OVER_4 = 0.785398163397448309616f;
  float ox = s0 * 2.0f - 1.0f, oy = s1 * 2.0f - 1.0f;
  if (ox == 0.0f && oy == 0.0f) {
    x = y = 0.0f;
    return;
  }
  float sn, cs;
  if (std::fabs(ox) > std::fabs(oy)) {
    f_sincos_quarter(PI_OVER_4 * (oy / ox), sn, cs);
    x = cs * ox;
    y = sn * ox;
  } else {
    f_sincos_quarter(PI_OVER_4 * (ox / oy), sn, cs);
    x = sn * oy;
    y = cs * oy;
  }
}
// sampling.rs:72-86 + diffuse.rs:24-36
V3 f_sample_cosine(V3 n, float s0, float s1) {
  float x, y;
  f_concentric_disk(s0, s1, x, y);
  const float z = std::sqrt(std::fmax(0.0f, 1.0f - x * x - y * y));
  const float sign = n.z < 0.0f ? -1.0f : 1.0f;
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  const V3 t = normalize(v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x));
  const V3 bt = normalize(v3(b, sign + n.y * n.y * a, -n.y));
  return t * x + bt * y + n * z;
}
// camera.rs:66-99: uv -> NDC at depth 1 through the inverse view-projection
V3 f_unproject(const float *m, float u, float v) {
  float x = u * 2.0f - 1.0f, y = (v * 2.0f - 1.0f) * -1.0f, z = 1.0f, w = 1.0f;
  float rx = x * m[0] + y * m[4] + z * m[8] + w * m[12];
  float ry = x * m[1] + y * m[5] + z * m[9] + w * m[13];
  float rz = x * m[2] + y * m[6] + z * m[10] + w * m[14];
  float rw = x * m[3] + y * m[7] + z * m[11] + w * m[15];
  return v3(rx / rw, ry / rw, rz / rw);
}
V3 f_normal_to_world(const M4 &m, V3 n) {
  return v3(m.a1 * n.x + m.a2 * n.y + m.a3 * n.z, m.b1 * n.x + m.b2 * n.y + m.b3 * n.z,
            m.c1 * n.x + m.c2 * n.y + m.c3 * n.z);
}

// world normal = transpose(inverse 3x3) * n, column-major inverse
D3 normal_to_world(const M4 &m, D3 n) {
  return d3(m.a1 * n.x + m.a2 * n.y + m.a3 * n.z, m.b1 * n.x + m.b2 * n.y + m.b3 * n.z,
            m.c1 * n.x + m.c2 * n.y + m.c3 * n.z);
}

struct PtCounts {
  std::vector<uint64_t> ext, shadow;
};

// The visibility side of a hit, in FP32 with the device's operation order
// (the reference computes it in f32 too): the geometric normal facing the
// ray and the continuation origin offset off the surface (ray_util.rs:6-40).
// Shading quantities stay in FP64.
struct VisFrame {
  V3 ng, origin;
  bool flipped;
};
VisFrame vis_frame(const orc_scene *s, uint32_t tri, uint32_t ti, V3 o, V3 d, float t) {
  const V3 p0 = s->vertices[s->indices[3 * (size_t)tri]], p1 = s->vertices[s->indices[3 * (size_t)tri + 1]],
           p2 = s->vertices[s->indices[3 * (size_t)tri + 2]];
  const V3 c = cross(p1 - p0, p2 - p0);
  const M4 &m = s->inst_transform_inv[ti];
  VisFrame f;
  f.ng = normalize(v3(m.a1 * c.x + m.a2 * c.y + m.a3 * c.z, m.b1 * c.x + m.b2 * c.y + m.b3 * c.z,
                      m.c1 * c.x + m.c2 * c.y + m.c3 * c.z));
  f.flipped = dot(f.ng, d) > 0.0f;
  if (f.flipped) f.ng = f.ng * -1.0f;
  const V3 pos = o + d * t;
  f.origin = v3(offset_axis(pos.x, f.ng.x), offset_axis(pos.y, f.ng.y), offset_axis(pos.z, f.ng.z));
  return f;
}

uint32_t orc_views(const orc_pt_params *P) { return P->views > 1 ? P->views : 1u; }

// One path vertex as the adjoint needs it: the material hit (or the sky the
// ray escaped to) and the unoccluded light weight cos_l / pi (0 if none).
constexpr uint32_t kVertSkyUp = 0xFFFFFFF0u, kVertSkyDown = 0xFFFFFFF1u;
struct PtVert {
  uint32_t code;
  double w;
  // geometry adjoint: the hit, the traced world ray, the sign in
  // ns = sigma * normalize(A n) (A = normal_to_world, n = the interpolated
  // object-space normal, or e1 x e2 when `interp` is false), and the
  // continuation direction with its cosine to ns
  uint32_t tri = 0, ti = 0;
  double sigma = 1.0;
  bool interp = false, cont = false;
  V3 o{}, d{};
  D3 wo{};
  double cos_wo = 0.0;
};

// The camera ray of pixel (px, py), sample n, of view `view` (ray_gen.rs:
// 89-107, camera.rs:66-99), FP32: jitter = (vdc, sobol) of the pixel's
// scrambles 0 and 1, uv = pixel / size + jitter / size, direction toward the
// unprojected uv at NDC depth 1.
void camera_ray(const orc_pt_params &P, uint32_t px, uint32_t py, uint32_t n, uint32_t view, V3 &o, V3 &dir) {
  const uint32_t seed = P.seed + view;
  const float *vp = P.view_params ? P.view_params + 20 * (size_t)view : P.view_projection_inv;
  const float *campos = P.view_params ? P.view_params + 20 * (size_t)view + 16 : P.camera_position;
  const float jx = orc_van_der_corput(n, scramble_of(px, py, seed, 0));
  const float jy = orc_sobol(n, scramble_of(px, py, seed, 1));
  const float u = (float)px / (float)P.width + jx / (float)P.width;
  const float v = (float)py / (float)P.height + jy / (float)P.height;
  o = v3(campos[0], campos[1], campos[2]);
  dir = normalize(f_unproject(vp, u, v) - o);
}

// One path; returns its radiance.  `rec`, if given, receives its vertices.
D3 pt_path(const orc_scene *s, const orc_pt_params &P, uint32_t path, uint64_t *ext_cnt, uint64_t *sh_cnt,
           std::vector<PtVert> *rec = nullptr) {
  const uint32_t gpix = path / P.spp;
  const uint32_t npix = P.width * P.height;
  const uint32_t view = gpix / npix, pixel = gpix % npix;
  const uint32_t n = P.sample_index_base + path % P.spp;
  const uint32_t px = pixel % P.width, py = pixel / P.width;
  const uint32_t seed = P.seed + view;
  V3 o32, dir;
  camera_ray(P, px, py, n, view, o32, dir);
  D3 beta = d3(1, 1, 1), L = d3(0, 0, 0);
  V3 beta_f = v3(1.0f, 1.0f, 1.0f);  // the FP32 throughput the roulette decides on
  const D3 sky_up = d3(P.sky_up[0], P.sky_up[1], P.sky_up[2]);
  const D3 sky_down = d3(P.sky_down[0], P.sky_down[1], P.sky_down[2]);
  const D3 wi = d3(-P.sun_direction[0], -P.sun_direction[1], -P.sun_direction[2]);
  const V3 wif = v3(-P.sun_direction[0], -P.sun_direction[1], -P.sun_direction[2]);
  const D3 sun = d3(P.sun_radiance[0], P.sun_radiance[1], P.sun_radiance[2]);
  const bool has_sun = P.sun_radiance[0] > 0.0f || P.sun_radiance[1] > 0.0f || P.sun_radiance[2] > 0.0f;
  for (uint32_t b = 0; b <= P.max_bounces; b++) {
    ext_cnt[b]++;
    RayIn ray;
    ray.o = o32;
    ray.d = dir;
    ray.tmin = FLT_EPSILON;
    ray.tmax = FLT_MAX;
    ray.flags = F_CULL_BACK;  // ray_gen.rs:162-166
    ray.mask = 0xffffffffu;
    ray.tlas_idx = 0;
    orc_hit h;
    orc_hit_meta hm;
    Counters cnt;
    uint32_t ext[2] = {0, 0};
    if (!traverse_one(s, ray, h, hm, cnt, ext)) {
      L = L + beta * (ray.d.y > 0.0f ? sky_up : sky_down);
      if (rec) rec->push_back(PtVert{ray.d.y > 0.0f ? kVertSkyUp : kVertSkyDown, 0.0});
      break;
    }
    const auto &geo = s->geo_meta[ext[0]];
    const uint32_t ti = ext[1];
    const uint32_t tri = geo.primitive_start + h.prim_id;
    const uint32_t i0 = s->indices[3 * (size_t)tri], i1 = s->indices[3 * (size_t)tri + 1],
                   i2 = s->indices[3 * (size_t)tri + 2];
    const M4 &inv = s->inst_transform_inv[ti];
    // FP32 frame: geometric normal facing the ray, shading normal, offset origin
    const VisFrame vf = vis_frame(s, tri, ti, ray.o, ray.d, h.t);
    const float bwf = 1.0f - h.u - h.v;
    const V3 nobj_f = s->normals[i0] * bwf + s->normals[i1] * h.u + s->normals[i2] * h.v;
    const bool interp = dot(nobj_f, nobj_f) > 0.0f;
    V3 ns_f = interp ? normalize(f_normal_to_world(inv, nobj_f)) : (vf.flipped ? vf.ng * -1.0f : vf.ng);
    double sigma = 1.0;
    if (dot(ns_f, vf.ng) < 0.0f) {
      ns_f = ns_f * -1.0f;
      sigma = -1.0;
    }
    // FP64 shading normal with the same orientation
    const D3 p0 = d3f(s->vertices[i0]), p1 = d3f(s->vertices[i1]), p2 = d3f(s->vertices[i2]);
    const double bu = h.u, bv = h.v, bw = 1.0 - bu - bv;
    const D3 nobj = interp ? d3f(s->normals[i0]) * bw + d3f(s->normals[i1]) * bu + d3f(s->normals[i2]) * bv
                           : dcross(p1 - p0, p2 - p0);
    const D3 ns = dnormalize(normal_to_world(inv, nobj)) * sigma;
    uint32_t mi = s->inst_custom[ti] + s->tri_material[tri];
    const size_t nm = s->materials.size() / 8;
    if (mi >= nm) mi = (uint32_t)nm - 1;
    const float *mat = &s->materials[8 * (size_t)mi];
    const D3 albedo = d3(mat[0], mat[1], mat[2]);
    L = L + beta * d3(mat[4], mat[5], mat[6]);
    if (rec) {
      PtVert pv{mi, 0.0};
      pv.tri = tri;
      pv.ti = ti;
      pv.sigma = sigma;
      pv.interp = interp;
      pv.o = ray.o;
      pv.d = ray.d;
      rec->push_back(pv);
    }
    if (b == P.max_bounces) break;
    const V3 origin = vf.origin;
    const float cos_lf = dot(ns_f, wif);
    if (cos_lf > 0.0f && dot(vf.ng, wif) > 0.0f && has_sun) {
      sh_cnt[b]++;
      RayIn sr;
      sr.o = origin;
      sr.d = wif;
      sr.tmin = FLT_EPSILON;
      sr.tmax = FLT_MAX;
      sr.flags = F_ACCEPT_FIRST_HIT_AND_END_SEARCH;  // ray_hit.rs:33-45
      sr.mask = 0xffffffffu;
      sr.tlas_idx = 0;
      orc_hit sh;
      orc_hit_meta shm;
      Counters sc;
      if (!traverse_one(s, sr, sh, shm, sc)) {
        const double cos_l = std::max(ddot(ns, wi), 0.0);
        L = L + beta * albedo * (cos_l / 3.14159265358979323846) * sun;
        if (rec) rec->back().w = cos_l / 3.14159265358979323846;
      }
    }
    const float s0 = orc_van_der_corput(n, scramble_of(px, py, seed, 2 + 2 * b));
    const float s1 = orc_sobol(n, scramble_of(px, py, seed, 3 + 2 * b));
    const V3 wo = f_sample_cosine(ns_f, s0, s1);
    if (!(dot(wo, vf.ng) > 0.0f)) break;
    if (rec) {
      rec->back().cont = true;
      rec->back().wo = d3f(wo);
      rec->back().cos_wo = ddot(ns, d3f(wo));
    }
    beta = beta * albedo;
    beta_f = beta_f * v3(mat[0], mat[1], mat[2]);
    if (b >= P.rr_start) {
      const float q = std::fmin(std::fmax(std::fmax(beta_f.x, beta_f.y), beta_f.z), 0.95f);
      const float ur = orc_van_der_corput(n, scramble_of(px, py, seed, 64 + b));
      if (ur >= q) break;
      beta = beta * (1.0 / (double)q);
      beta_f = v3(beta_f.x / q, beta_f.y / q, beta_f.z / q);
    }
    o32 = origin;
    dir = wo;
  }
  return L;
}

}  // namespace

extern "C" {

uint32_t orc_xxhash32(uint32_t x, uint32_t y, uint32_t z) { return xxhash32(x, y, z); }

int orc_set_vertex_uvs(orc_scene *s, uint32_t blas, uint32_t geo, const float *uv, uint32_t nverts) {
  if (blas >= s->blas_data.size() || geo >= s->blas_data[blas].size()) return -1;
  Geometry &g = s->blas_data[blas][geo];
  if ((size_t)nverts * 3 != g.pos.size()) return -1;
  g.uvs.assign(uv, uv + 2 * (size_t)nverts);
  s->built = false;
  return 0;
}

// Shading attributes of closest hits, as the reference's closest-hit stage
// fetches them through the bindless mesh bridge (scene/rendering/gpu-ray-
// tracing/src/bindless_mesh_bridge.rs:43-123), in FP32 with its operation
// order: world position o + t d (api/ctx.rs:209-211); geometric normal
// normalize(N (a - b) x (a - c)) turned to face the ray origin, N the
// transpose of the instance's world-to-object 3x3 (:95,104-114); shading
// normal normalize(N (w na + u nb + v nc)), w = 1 - u - v, on the geometric
// normal's side (:86-123; the geometric normal when the vertex normals are
// zero); uv interpolated the same way; hit group offset + stride * geometry
// + instance offset (shader/ray-tracing/src/api/ctx.rs:53-55).  And the hit
// context of the reference's closest-hit stage (HitCtxInfo, api/ctx.rs:24-
// 48, filled at traverse_gpu.rs:575-582,661-683): object-to-world (the
// instance transform), world-to-object (its inverse), the object-space ray
// (world-to-object origin with the w divide, normalized 3x3 direction;
// traverse_gpu.rs:427-431).  attr: 12 words per hit (position 3, shading
// normal 3, geometric normal 3, uv 2, hit group as u32; zeros and ~0 on a
// miss); xf (may be NULL): 40 words per hit (object-to-world 16,
// world-to-object 16, object origin 3, object direction 3, instance sbt
// offset as u32, pad; zeros on a miss).
int orc_hit_attributes(const orc_scene *s, const orc_ray *rays, const orc_ray_meta *meta, const orc_hit *hits,
                       const orc_hit_meta *hmeta, uint32_t n, uint32_t tlas_idx, uint32_t sbt_offset,
                       uint32_t sbt_stride, float *attr, float *xf) {
  if (!s->built) return -1;
  for (uint32_t i = 0; i < n; i++) {
    float *o = attr + 12 * (size_t)i;
    float *x = xf ? xf + 40 * (size_t)i : nullptr;
    for (int k = 0; k < 12; k++) o[k] = 0.0f;
    const uint32_t miss = 0xFFFFFFFFu;
    std::memcpy(o + 11, &miss, 4);
    if (x)
      for (int k = 0; k < 40; k++) x[k] = 0.0f;
    const orc_hit &h = hits[i];
    const orc_hit_meta &hm = hmeta[i];
    const uint32_t ti_bind = meta ? meta[i].tlas_idx : tlas_idx;
    if (h.prim_id == miss || ti_bind >= s->binding.size()) continue;
    const uint32_t th = s->binding[ti_bind];
    if (th + 1 >= s->tlas_inst_first.size() || th >= s->tlas_live.size() || !s->tlas_live[th]) continue;
    // the instance slot holding source instance hm.instance_id of this TLAS
    uint32_t slot = miss;
    for (uint32_t k = s->tlas_inst_first[th]; k < s->tlas_inst_first[th + 1]; k++)
      if (s->inst_order[k] == hm.instance_id) slot = k;
    if (slot == miss) continue;
    const auto &bm = s->blas_meta[s->inst_blas[slot]];
    if (hm.geometry_id >= bm.second) continue;
    const auto &geo = s->geo_meta[bm.first + hm.geometry_id];
    const uint32_t tri = geo.primitive_start + h.prim_id;
    const uint32_t ia = s->indices[3 * (size_t)tri], ib = s->indices[3 * (size_t)tri + 1],
                   ic = s->indices[3 * (size_t)tri + 2];
    const M4 &inv = s->inst_transform_inv[slot];
    const orc_ray &r = rays[i];
    const V3 org = v3(r.ox, r.oy, r.oz), dir = v3(r.dx, r.dy, r.dz);
    const V3 pos = org + dir * h.t;
    const V3 pa = s->vertices[ia], pb = s->vertices[ib], pc = s->vertices[ic];
    V3 gn = normalize(f_normal_to_world(inv, cross(pa - pb, pa - pc)));
    if (dot(org - pos, gn) < 0.0f) gn = gn * -1.0f;
    V3 sn = gn;
    const float w = 1.0f - h.u - h.v;
    if (!s->normals.empty()) {
      const V3 nn = s->normals[ia] * w + s->normals[ib] * h.u + s->normals[ic] * h.v;
      if (dot(nn, nn) > 0.0f) {
        sn = normalize(f_normal_to_world(inv, nn));
        if (dot(gn, sn) < 0.0f) sn = sn * -1.0f;
      }
    }
    if (s->has_uvs)
      for (int k = 0; k < 2; k++)
        o[9 + k] = s->uvs[2 * (size_t)ia + k] * w + s->uvs[2 * (size_t)ib + k] * h.u + s->uvs[2 * (size_t)ic + k] * h.v;
    o[0] = pos.x;
    o[1] = pos.y;
    o[2] = pos.z;
    o[3] = sn.x;
    o[4] = sn.y;
    o[5] = sn.z;
    o[6] = gn.x;
    o[7] = gn.y;
    o[8] = gn.z;
    const uint32_t group = sbt_offset + sbt_stride * hm.geometry_id + s->inst_sbt[slot];
    std::memcpy(o + 11, &group, 4);
    if (x) {
      std::memcpy(x, &s->inst_transform[slot], 64);
      std::memcpy(x + 16, &inv, 64);
      const V3 bo = m4_mul_point(inv, org);
      const V3 bd = normalize(m3_mul(inv, dir));
      x[32] = bo.x;
      x[33] = bo.y;
      x[34] = bo.z;
      x[35] = bd.x;
      x[36] = bd.y;
      x[37] = bd.z;
      std::memcpy(x + 38, &s->inst_sbt[slot], 4);
    }
  }
  return 0;
}

void orc_camera_ray(const orc_pt_params *P, uint32_t px, uint32_t py, uint32_t sample, uint32_t view, float *o3,
                    float *d3) {
  V3 o, d;
  camera_ray(*P, px, py, sample, view, o, d);
  o3[0] = o.x;
  o3[1] = o.y;
  o3[2] = o.z;
  d3[0] = d.x;
  d3[1] = d.y;
  d3[2] = d.z;
}

int orc_set_vertex_normals(orc_scene *s, uint32_t blas, uint32_t geo, const float *nrm, uint32_t nverts) {
  if (blas >= s->blas_data.size() || geo >= s->blas_data[blas].size()) return -1;
  Geometry &g = s->blas_data[blas][geo];
  if (nverts != g.pos.size() / 3) return -1;
  g.normals.assign(nrm, nrm + 3 * (size_t)nverts);
  s->built = false;
  return 0;
}

int orc_set_anyhit(orc_scene *s, uint32_t blas, uint32_t geo, const uint8_t *behavior, uint32_t ntris) {
  if (blas >= s->blas_data.size() || geo >= s->blas_data[blas].size()) return -1;
  Geometry &g = s->blas_data[blas][geo];
  g.anyhit.assign(behavior, behavior + ntris);
  s->built = false;
  return 0;
}

int orc_set_triangle_materials(orc_scene *s, uint32_t blas, uint32_t geo, const uint32_t *off, uint32_t ntris) {
  if (blas >= s->blas_data.size() || geo >= s->blas_data[blas].size()) return -1;
  Geometry &g = s->blas_data[blas][geo];
  g.tri_material.assign(off, off + ntris);
  s->built = false;
  return 0;
}

void orc_set_materials(orc_scene *s, const float *mats, uint32_t n) { s->materials.assign(mats, mats + 8 * (size_t)n); }

int orc_pt_render(const orc_scene *s, const orc_pt_params *P, double *out, uint64_t *ray_counts, int nthreads) {
  return orc_pt_render_rows(s, P, 0, orc_views(P) * P->height, out, ray_counts, nthreads);
}

int orc_pt_render_rows(const orc_scene *s, const orc_pt_params *P, uint32_t row_begin, uint32_t row_end, double *out,
                       uint64_t *ray_counts, int nthreads) {
  if (!s->built || s->materials.empty() || P->spp == 0 || P->max_bounces > 16) return -1;
  if (row_begin > row_end || row_end > orc_views(P) * P->height) return -1;
  if (P->views > 1 && !P->view_params) return -1;
  const uint32_t pix0 = row_begin * P->width;
  const uint32_t npix = (row_end - row_begin) * P->width;
  const uint32_t nb = P->max_bounces + 1;
  std::vector<uint64_t> ext_all(nb, 0), sh_all(nb, 0);
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
  {
    std::vector<uint64_t> ext(nb, 0), sh(nb, 0);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t pix = 0; pix < (int64_t)npix; pix++) {
      D3 acc = d3(0, 0, 0);
      const uint32_t gpix = pix0 + (uint32_t)pix;
      for (uint32_t k = 0; k < P->spp; k++) acc = acc + pt_path(s, *P, gpix * P->spp + k, ext.data(), sh.data());
      out[3 * pix + 0] = acc.x / P->spp;
      out[3 * pix + 1] = acc.y / P->spp;
      out[3 * pix + 2] = acc.z / P->spp;
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    for (uint32_t b = 0; b < nb; b++) {
      ext_all[b] += ext[b];
      sh_all[b] += sh[b];
    }
  }
  if (ray_counts)
    for (uint32_t b = 0; b < nb; b++) {
      ray_counts[b] = ext_all[b];
      ray_counts[nb + b] = sh_all[b];
    }
  return 0;
}

// Adjoint of the frame radiance w.r.t. the material albedo and emission, in
// FP64, by reverse accumulation over each path's recorded vertices (visibility
// and sampled directions fixed).  With T_k the radiance from vertex k on
// divided by the throughput beta_k:
//   T_k = e_k + a_k * (w_k sun + T_{k+1}),  T_{m+1} = sky or 0,
//   dC/de_k = beta_k,  dC/da_k = beta_k * (w_k sun + T_{k+1}).
// Paths that Russian roulette could end are refused (rr_start < max_bounces):
// their termination probability depends on the albedo.
int orc_pt_grad(const orc_scene *s, const orc_pt_params *P, const double *adj, double *g_albedo, double *g_emission,
                int nthreads) {
  if (!s->built || s->materials.empty() || P->spp == 0 || P->max_bounces > 16) return -1;
  if (P->rr_start < P->max_bounces) return -2;
  if (P->views > 1 && !P->view_params) return -1;
  const size_t nm = s->materials.size() / 8;
  const uint32_t npix = orc_views(P) * P->width * P->height;
  std::fill(g_albedo, g_albedo + 3 * nm, 0.0);
  std::fill(g_emission, g_emission + 3 * nm, 0.0);
  const D3 sky_up = d3(P->sky_up[0], P->sky_up[1], P->sky_up[2]);
  const D3 sky_down = d3(P->sky_down[0], P->sky_down[1], P->sky_down[2]);
  const D3 sun = d3(P->sun_radiance[0], P->sun_radiance[1], P->sun_radiance[2]);
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
  {
    std::vector<double> ga(3 * nm, 0.0), ge(3 * nm, 0.0);
    std::vector<uint64_t> ext(P->max_bounces + 1, 0), sh(P->max_bounces + 1, 0);
    std::vector<PtVert> rec;
    std::vector<D3> beta;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t pix = 0; pix < (int64_t)npix; pix++) {
      const D3 a_pix = d3(adj[3 * pix], adj[3 * pix + 1], adj[3 * pix + 2]) * (1.0 / P->spp);
      for (uint32_t k = 0; k < P->spp; k++) {
        rec.clear();
        pt_path(s, *P, (uint32_t)pix * P->spp + k, ext.data(), sh.data(), &rec);
        beta.assign(1, d3(1, 1, 1));
        size_t m = 0;  // vertices with a material
        while (m < rec.size() && rec[m].code < kVertSkyUp) {
          const float *mat = &s->materials[8 * (size_t)rec[m].code];
          beta.push_back(beta.back() * d3(mat[0], mat[1], mat[2]));
          m++;
        }
        D3 T = d3(0, 0, 0);
        if (m < rec.size()) T = rec[m].code == kVertSkyUp ? sky_up : sky_down;
        for (size_t j = m; j-- > 0;) {
          const uint32_t mi = rec[j].code;
          const float *mat = &s->materials[8 * (size_t)mi];
          const D3 a = d3(mat[0], mat[1], mat[2]), e = d3(mat[4], mat[5], mat[6]);
          const D3 lit = sun * rec[j].w;
          const D3 da = a_pix * beta[j] * (lit + T), de = a_pix * beta[j];
          ga[3 * mi] += da.x;
          ga[3 * mi + 1] += da.y;
          ga[3 * mi + 2] += da.z;
          ge[3 * mi] += de.x;
          ge[3 * mi + 1] += de.y;
          ge[3 * mi + 2] += de.z;
          T = e + a * (lit + T);
        }
      }
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    for (size_t i = 0; i < 3 * nm; i++) {
      g_albedo[i] += ga[i];
      g_emission[i] += ge[i];
    }
  }
  return 0;
}

}  // extern "C"

namespace {

// ---- geometry adjoint (FP64) ----
// The function differentiated: the frame radiance with every path's traced
// rays, hit triangles, light visibility, facing decisions and sampled
// directions held fixed, as a function of the vertex positions P and
// normals N.  At each hit the shading normal is
//   ns = sigma * normalize(A n),  n = w0 N0 + u N1 + v N2
// (or n = e1 x e2 without vertex normals), with (u, v) where the fixed
// object-space ray meets the triangle of P; the light term is
// beta a sun max(ns . wi, 0) / pi and the throughput factor of the
// continuation is a (ns . wo) / cos_wo (cos_wo: its value in the frame).
struct GeoEval {
  D3 ns, m;  // shading normal and the unnormalized A n
  double mlen, u, v;
  D3 e1, e2, c0;  // edges and -bd (object space)
  uint32_t i0, i1, i2;
};

GeoEval geo_eval(const orc_scene *s, const PtVert &pv, const double *pos, const double *nrm) {
  GeoEval g;
  g.i0 = s->indices[3 * (size_t)pv.tri];
  g.i1 = s->indices[3 * (size_t)pv.tri + 1];
  g.i2 = s->indices[3 * (size_t)pv.tri + 2];
  auto P3 = [&](uint32_t i) {
    return pos ? d3(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]) : d3f(s->vertices[i]);
  };
  auto N3 = [&](uint32_t i) {
    return nrm ? d3(nrm[3 * (size_t)i], nrm[3 * (size_t)i + 1], nrm[3 * (size_t)i + 2]) : d3f(s->normals[i]);
  };
  const M4 &m = s->inst_transform_inv[pv.ti];
  const D3 o = d3f(pv.o), d = d3f(pv.d);
  const D3 bo = d3(m.a1 * o.x + m.b1 * o.y + m.c1 * o.z + m.d1, m.a2 * o.x + m.b2 * o.y + m.c2 * o.z + m.d2,
                   m.a3 * o.x + m.b3 * o.y + m.c3 * o.z + m.d3);
  const D3 bd = d3(m.a1 * d.x + m.b1 * d.y + m.c1 * d.z, m.a2 * d.x + m.b2 * d.y + m.c2 * d.z,
                   m.a3 * d.x + m.b3 * d.y + m.c3 * d.z);
  const D3 p0 = P3(g.i0), p1 = P3(g.i1), p2 = P3(g.i2);
  g.e1 = p1 - p0;
  g.e2 = p2 - p0;
  g.c0 = bd * -1.0;
  const D3 sv = bo - p0;
  const double det = ddot(g.c0, dcross(g.e1, g.e2));
  g.u = ddot(g.c0, dcross(sv, g.e2)) / det;
  g.v = ddot(g.c0, dcross(g.e1, sv)) / det;
  const D3 n = pv.interp ? N3(g.i0) * (1.0 - g.u - g.v) + N3(g.i1) * g.u + N3(g.i2) * g.v : dcross(g.e1, g.e2);
  g.m = normal_to_world(m, n);
  g.mlen = std::sqrt(ddot(g.m, g.m));
  g.ns = g.m * (pv.sigma / g.mlen);
  return g;
}

// Replays one path; returns sum_c adj_c C_c and, if g_pos / g_nrm are given,
// adds the gradient of that value.
double replay_path(const orc_scene *s, const orc_pt_params &P, const std::vector<PtVert> &rec, D3 adj,
                   const double *pos, const double *nrm, double *g_pos, double *g_nrm) {
  const double inv_pi = 1.0 / 3.14159265358979323846;
  const D3 sun = d3(P.sun_radiance[0], P.sun_radiance[1], P.sun_radiance[2]);
  const D3 wi = d3(-P.sun_direction[0], -P.sun_direction[1], -P.sun_direction[2]);
  const size_t nv = rec.size();
  std::vector<GeoEval> ge(nv);
  std::vector<D3> beta(nv + 1), direct(nv);
  std::vector<double> ratio(nv, 1.0), cosl(nv, 0.0);
  beta[0] = d3(1, 1, 1);
  D3 tail = d3(0, 0, 0);
  size_t m = 0;
  for (; m < nv; m++) {
    const PtVert &pv = rec[m];
    if (pv.code >= kVertSkyUp) {
      tail = beta[m] * (pv.code == kVertSkyUp ? d3(P.sky_up[0], P.sky_up[1], P.sky_up[2])
                                               : d3(P.sky_down[0], P.sky_down[1], P.sky_down[2]));
      break;
    }
    const float *mat = &s->materials[8 * (size_t)pv.code];
    const D3 a = d3(mat[0], mat[1], mat[2]), e = d3(mat[4], mat[5], mat[6]);
    ge[m] = geo_eval(s, pv, pos, nrm);
    direct[m] = beta[m] * e;
    if (pv.w > 0.0) {
      cosl[m] = ddot(ge[m].ns, wi);
      if (cosl[m] > 0.0) direct[m] = direct[m] + beta[m] * a * sun * (cosl[m] * inv_pi);
    }
    if (!pv.cont) {
      m++;
      break;
    }
    ratio[m] = ddot(ge[m].ns, pv.wo) / pv.cos_wo;
    beta[m + 1] = beta[m] * a * ratio[m];
  }
  double total = ddot(adj, tail);
  for (size_t k = 0; k < m; k++) total += ddot(adj, direct[k]);
  if (!g_pos) return total;
  D3 rem = tail;  // radiance after vertex k
  for (size_t k = m; k-- > 0;) {
    const PtVert &pv = rec[k];
    const float *mat = &s->materials[8 * (size_t)pv.code];
    const D3 a = d3(mat[0], mat[1], mat[2]);
    const GeoEval &g = ge[k];
    D3 g_ns = d3(0, 0, 0);
    if (pv.w > 0.0 && cosl[k] > 0.0) g_ns = g_ns + wi * (ddot(adj, beta[k] * a * sun) * inv_pi);
    if (pv.cont && ratio[k] != 0.0) g_ns = g_ns + pv.wo * (ddot(adj, rem) / ratio[k] / pv.cos_wo);
    rem = rem + direct[k];
    // ns = sigma m / |m|
    const D3 g_m = (g_ns - g.ns * ddot(g.ns, g_ns)) * (pv.sigma / g.mlen);
    const M4 &mm = s->inst_transform_inv[pv.ti];
    const D3 g_n = d3(mm.a1 * g_m.x + mm.b1 * g_m.y + mm.c1 * g_m.z, mm.a2 * g_m.x + mm.b2 * g_m.y + mm.c2 * g_m.z,
                      mm.a3 * g_m.x + mm.b3 * g_m.y + mm.c3 * g_m.z);
    auto add = [](double *arr, uint32_t i, D3 v) {
      arr[3 * (size_t)i] += v.x;
      arr[3 * (size_t)i + 1] += v.y;
      arr[3 * (size_t)i + 2] += v.z;
    };
    if (pv.interp) {
      const double w0 = 1.0 - g.u - g.v;
      add(g_nrm, g.i0, g_n * w0);
      add(g_nrm, g.i1, g_n * g.u);
      add(g_nrm, g.i2, g_n * g.v);
      auto N3 = [&](uint32_t i) {
        return nrm ? d3(nrm[3 * (size_t)i], nrm[3 * (size_t)i + 1], nrm[3 * (size_t)i + 2]) : d3f(s->normals[i]);
      };
      const double g_u = ddot(g_n, N3(g.i1) - N3(g.i0)), g_v = ddot(g_n, N3(g.i2) - N3(g.i0));
      // (t, u, v) solves [c0 e1 e2] x = bo - p0
      const double det = ddot(g.c0, dcross(g.e1, g.e2));
      const D3 G = (dcross(g.e2, g.c0) * g_u + dcross(g.c0, g.e1) * g_v) * (1.0 / det);
      add(g_pos, g.i0, G * -w0);
      add(g_pos, g.i1, G * -g.u);
      add(g_pos, g.i2, G * -g.v);
    } else {
      // n = e1 x e2
      const D3 g_e1 = dcross(g.e2, g_n), g_e2 = dcross(g_n, g.e1);
      add(g_pos, g.i1, g_e1);
      add(g_pos, g.i2, g_e2);
      add(g_pos, g.i0, (g_e1 + g_e2) * -1.0);
    }
  }
  return total;
}

template <class F>
int for_each_path(const orc_scene *s, const orc_pt_params *P, int nthreads, F &&fn) {
  const uint32_t npix = orc_views(P) * P->width * P->height;
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    std::vector<uint64_t> ext(P->max_bounces + 1, 0), sh(P->max_bounces + 1, 0);
    std::vector<PtVert> rec;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t pix = 0; pix < (int64_t)npix; pix++)
      for (uint32_t k = 0; k < P->spp; k++) {
        rec.clear();
        pt_path(s, *P, (uint32_t)pix * P->spp + k, ext.data(), sh.data(), &rec);
        fn(tid, (uint32_t)pix, rec);
      }
  }
  return 0;
}

int nthreads_of(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

}  // namespace

extern "C" {

// One round of batched tracing: closest hit of every ray (traverse_cpu.rs:
// 52-245 with per-ray TLAS selection, traverse_cpu.rs:63), then for each hit
// a continuation ray from the offset hit point (ray_util.rs:6-40) in a
// cosine-weighted direction (sampling.rs:72-86, diffuse.rs:24-36) around the
// geometric normal facing the ray, sample (vdc, sobol)(round) of the ray's
// sequence scrambled by xxhash32(ray, seed, 2 / 3) (sampler.rs:33-42,
// sobol.rs:40-68).  Shading math in FP64, the offset on FP32 values.
int orc_batch_round(const orc_scene *s, const orc_ray *rays, const orc_ray_meta *meta, uint32_t n, uint32_t round,
                    uint32_t seed, orc_hit *hits, orc_ray *spawned, uint32_t *alive, int nthreads) {
  if (!s->built || !meta) return -1;
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const RayIn ri = load_ray(rays[i], &meta[i], 0, 0, 0);
    orc_hit h;
    orc_hit_meta hm;
    Counters c;
    uint32_t ext[2] = {0, 0};
    const bool hit = traverse_one(s, ri, h, hm, c, ext);
    hits[i] = h;
    alive[i] = hit ? 1u : 0u;
    orc_ray out{};
    if (hit) {
      const auto &geo = s->geo_meta[ext[0]];
      const uint32_t tri = geo.primitive_start + h.prim_id;
      const VisFrame vf = vis_frame(s, tri, ext[1], ri.o, ri.d, h.t);
      const float s0 = orc_van_der_corput(round, xxhash32((uint32_t)i, seed, 2u));
      const float s1 = orc_sobol(round, xxhash32((uint32_t)i, seed, 3u));
      const V3 wo = f_sample_cosine(vf.ng, s0, s1);
      out.ox = vf.origin.x;
      out.oy = vf.origin.y;
      out.oz = vf.origin.z;
      out.tmin = FLT_EPSILON;
      out.dx = (float)wo.x;
      out.dy = (float)wo.y;
      out.dz = (float)wo.z;
      out.tmax = FLT_MAX;
    }
    spawned[i] = out;
  }
  return 0;
}

uint32_t orc_vertex_count(const orc_scene *s) { return (uint32_t)s->vertices.size(); }

// Gradient of sum(adj * frame) w.r.t. vertex positions and normals (global
// vertex order, 3 doubles each), FP64, paths held fixed as above.
int orc_pt_grad_geometry(const orc_scene *s, const orc_pt_params *P, const double *adj, double *g_pos,
                         double *g_nrm, int nthreads) {
  if (!s->built || s->materials.empty() || P->spp == 0 || P->max_bounces > 16) return -1;
  if (P->rr_start < P->max_bounces) return -2;
  if (P->views > 1 && !P->view_params) return -1;
  const size_t n3 = 3 * s->vertices.size();
  const int nt = nthreads_of(nthreads);
  std::vector<std::vector<double>> gp(nt, std::vector<double>(n3, 0.0)), gn(nt, std::vector<double>(n3, 0.0));
  const double inv_spp = 1.0 / P->spp;
  for_each_path(s, P, nt, [&](int tid, uint32_t pix, const std::vector<PtVert> &rec) {
    const D3 a = d3(adj[3 * (size_t)pix], adj[3 * (size_t)pix + 1], adj[3 * (size_t)pix + 2]) * inv_spp;
    replay_path(s, *P, rec, a, nullptr, nullptr, gp[tid].data(), gn[tid].data());
  });
  for (size_t i = 0; i < n3; i++) {
    double a = 0.0, b = 0.0;
    for (int t = 0; t < nt; t++) {
      a += gp[t][i];
      b += gn[t][i];
    }
    g_pos[i] = a;
    g_nrm[i] = b;
  }
  return 0;
}

// sum(adj * frame) of the fixed paths re-shaded with positions `pos` and
// normals `nrm` (3 doubles per vertex each, NULL = the scene's own).
int orc_pt_replay_geometry(const orc_scene *s, const orc_pt_params *P, const double *adj, const double *pos,
                           const double *nrm, double *out, int nthreads) {
  if (!s->built || s->materials.empty() || P->spp == 0 || P->max_bounces > 16) return -1;
  if (P->rr_start < P->max_bounces) return -2;
  if (P->views > 1 && !P->view_params) return -1;
  const int nt = nthreads_of(nthreads);
  std::vector<double> part(nt, 0.0);
  const double inv_spp = 1.0 / P->spp;
  for_each_path(s, P, nt, [&](int tid, uint32_t pix, const std::vector<PtVert> &rec) {
    const D3 a = d3(adj[3 * (size_t)pix], adj[3 * (size_t)pix + 1], adj[3 * (size_t)pix + 2]) * inv_spp;
    part[tid] += replay_path(s, *P, rec, a, pos, nrm, nullptr, nullptr);
  });
  double t = 0.0;
  for (double v : part) t += v;
  *out = t;
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------ host picking --
// content/mesh/core/src/feature/bvh.rs:57-86 intersect_nearest_bvh over the
// math/geometry host tests.
namespace {

// math/geometry/src/dimension3/intersection.rs:3-76 (GTE IntrRay3Triangle3);
// the flipped triangle of triangle.rs:76-82 for FaceSide::Back.
float pick_tri(V3 o, V3 d, V3 a, V3 b, V3 c, uint32_t side) {
  const bool back = side == 1u, culling = side != 2u;
  if (back) {
    V3 t = a;
    a = c;
    c = t;
  }
  const V3 e1 = b - a;
  const V3 e2 = c - a;
  const V3 nrm = cross(e1, e2);
  float DdN = dot(d, nrm);
  float sign;
  if (DdN > 0.0f) {
    if (culling) return NAN;
    sign = 1.0f;
  } else if (DdN < 0.0f) {
    sign = -1.0f;
    DdN = -DdN;
  } else {
    return NAN;
  }
  const V3 q = o - a;
  const float DdQxE2 = sign * dot(d, cross(q, e2));
  if (DdQxE2 < 0.0f) return NAN;
  const float DdE1xQ = sign * dot(d, cross(e1, q));
  if (DdE1xQ < 0.0f) return NAN;
  if (DdQxE2 + DdE1xQ > DdN) return NAN;
  const float QdN = -sign * dot(q, nrm);
  if (QdN < 0.0f) return NAN;
  return QdN / DdN;
}

// intersection.rs:128-196 then is_some() (:203-208): one slab pair per axis
// in x, y, z order with the NaN-aware narrowing.
bool pick_slab(V3 o, V3 d, V3 mn, V3 mx) {
  const float oo[3] = {o.x, o.y, o.z};
  const float inv[3] = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  const float lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
  float t_min = 0.0f, t_max = 0.0f;
  for (int k = 0; k < 3; k++) {
    float a, b;
    if (inv[k] >= 0.0f) {
      a = (lo[k] - oo[k]) * inv[k];
      b = (hi[k] - oo[k]) * inv[k];
    } else {
      a = (hi[k] - oo[k]) * inv[k];
      b = (lo[k] - oo[k]) * inv[k];
    }
    if (k == 0) {
      t_min = a;
      t_max = b;
      continue;
    }
    if (t_min > b || a > t_max) return false;
    if (a > t_min || std::isnan(t_min)) t_min = a;
    if (b < t_max || std::isnan(t_max)) t_max = b;
  }
  return !(t_max < 0.0f);
}

struct PickState {
  const orc_scene *s;
  uint32_t prim_start;
  V3 o, d;
  uint32_t side;
  bool found;
  uint32_t prim;
  float dist;
  std::vector<std::pair<uint32_t, float>> *all;  // every hit in visit order, or null
};

void pick_leaf_prim(PickState &st, uint32_t tri) {
  const orc_scene *s = st.s;
  const float t = pick_tri(st.o, st.d, s->vertices[s->indices[3 * (size_t)tri]],
                           s->vertices[s->indices[3 * (size_t)tri + 1]], s->vertices[s->indices[3 * (size_t)tri + 2]],
                           st.side);
  if (std::isnan(t)) return;
  if (st.all) st.all->push_back({tri - st.prim_start, t});
  if (st.found && !(t < st.dist)) return;  // is_near_than: distance < (intersect_util.rs:15-17)
  st.found = true;
  st.dist = t;
  st.prim = tri - st.prim_start;
}

// Depth-first, right subtree before left: the order in which the reference's
// stack (push left, push right, pop) visits the nodes.
void pick_node(PickState &st, uint32_t i) {
  const orc_node &nd = st.s->tri_forest[i];
  const bool leaf = nd.hit_next == nd.miss_next;
  if (leaf) {
    for (uint32_t k = nd.content_range[0]; k < nd.content_range[1]; k++) pick_leaf_prim(st, st.s->redirect[k]);
    return;
  }
  if (!pick_slab(st.o, st.d, v3(nd.aabb_min[0], nd.aabb_min[1], nd.aabb_min[2]),
                 v3(nd.aabb_max[0], nd.aabb_max[1], nd.aabb_max[2])))
    return;
  const uint32_t left = nd.hit_next;
  const uint32_t right = st.s->tri_forest[left].miss_next;
  pick_node(st, right);
  pick_node(st, left);
}

}  // namespace

extern "C" {

float orc_pick_triangle(const float o[3], const float d[3], const float a[3], const float b[3], const float c[3],
                        uint32_t side) {
  return pick_tri(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), v3(a[0], a[1], a[2]), v3(b[0], b[1], b[2]),
                  v3(c[0], c[1], c[2]), side);
}

int orc_pick_box(const float o[3], const float d[3], const float bmin[3], const float bmax[3]) {
  return pick_slab(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), v3(bmin[0], bmin[1], bmin[2]),
                   v3(bmax[0], bmax[1], bmax[2]))
             ? 1
             : 0;
}

int orc_pick(const orc_scene *s, uint32_t blas, uint32_t geometry, const float *rays, uint32_t n, uint32_t side,
             int bvh, orc_pick_hit *out) {
  if (!s->built || blas >= s->blas_meta.size() || side > 2u) return -1;
  const auto bm = s->blas_meta[blas];
  if (geometry >= bm.second - bm.first) return -1;
  const auto &geo = s->geo_meta[bm.first + geometry];
  // the geometry's triangles in index order: its leaves cover one
  // contiguous range of the redirect array starting at its root
  const orc_node &root = s->tri_forest[geo.root];
  std::vector<uint32_t> tris(s->redirect.begin() + root.content_range[0], s->redirect.begin() + root.content_range[1]);
  std::sort(tris.begin(), tris.end());
  for (uint32_t r = 0; r < n; r++) {
    const float *q = rays + 6 * (size_t)r;
    PickState st{s, geo.primitive_start, v3(q[0], q[1], q[2]), v3(q[3], q[4], q[5]), side, false, 0u, 0.0f, nullptr};
    if (bvh) {
      pick_node(st, geo.root);
    } else {
      for (uint32_t t : tris) pick_leaf_prim(st, t);
    }
    orc_pick_hit h;
    std::memset(&h, 0, sizeof(h));
    h.prim = st.found ? st.prim : 0xffffffffu;
    if (st.found) {
      h.distance = st.dist;
      const V3 p = st.o + st.d * st.dist;  // hyper_ray.rs:11-13
      h.position[0] = p.x;
      h.position[1] = p.y;
      h.position[2] = p.z;
    }
    out[r] = h;
  }
  return 0;
}

int orc_pick_all(const orc_scene *s, uint32_t blas, uint32_t geometry, const float *ray, uint32_t side,
                 orc_pick_hit *out, uint32_t capacity, uint32_t *count) {
  if (!s->built || blas >= s->blas_meta.size() || side > 2u) return -1;
  const auto bm = s->blas_meta[blas];
  if (geometry >= bm.second - bm.first) return -1;
  const auto &geo = s->geo_meta[bm.first + geometry];
  std::vector<std::pair<uint32_t, float>> all;
  PickState st{s, geo.primitive_start, v3(ray[0], ray[1], ray[2]), v3(ray[3], ray[4], ray[5]), side, false, 0u, 0.0f,
               &all};
  pick_node(st, geo.root);
  *count = (uint32_t)all.size();
  for (uint32_t k = 0; k < all.size() && k < capacity; k++) {
    orc_pick_hit h;
    std::memset(&h, 0, sizeof(h));
    h.prim = all[k].first;
    h.distance = all[k].second;
    const V3 p = st.o + st.d * h.distance;
    h.position[0] = p.x;
    h.position[1] = p.y;
    h.position[2] = p.z;
    out[k] = h;
  }
  return 0;
}

}  // extern "C"

// ------------------------------------------------------- ambient occlusion --
// scene/rendering/gpu-ray-tracing/src/feature/ao.rs:192-304 per pixel, in
// plain control flow (ray gen -> closest hit -> occlusion ray -> running
// mean), FP32 with the reference operation order.
namespace {

// sampler.rs:11-16 pcg: one LCG step, then the RXS-M-XS hash of the new state
uint32_t ao_pcg_bits(uint32_t &state) {
  const uint32_t prev = state * 747796405u + 2891336453u;
  const uint32_t word = ((prev >> ((prev >> 28u) + 4u)) ^ prev) * 277803737u;
  state = prev;
  return (word >> 22u) ^ word;
}

// sampler.rs:64-71 next(): the top 23 hash bits as a mantissa in [1, 2), less 1
float ao_pcg_next(uint32_t &state) {
  const uint32_t r = ao_pcg_bits(state);
  const uint32_t bits = 0x3f800000u | (r >> 9u);
  float f;
  std::memcpy(&f, &bits, 4);
  return f - 1.0f;
}

// One pixel's sample value: 1 when the camera ray or the occlusion ray
// misses, 0 when both hit.
float ao_value(const orc_scene *s, const orc_ao_params &P, uint32_t x, uint32_t y, const float local[3]) {
  uint32_t state = xxhash32(x, y, P.sample_index);
  const float r0 = ao_pcg_next(state);
  const float r1 = ao_pcg_next(state);
  const float u = (float)x / (float)P.width + r0 / (float)P.width;
  const float v = (float)y / (float)P.height + r1 / (float)P.height;
  const V3 cam = v3(P.camera_position[0], P.camera_position[1], P.camera_position[2]);
  const V3 dir = normalize(f_unproject(P.view_projection_inv, u, v) - cam);
  RayIn ray;
  ray.o = cam;
  ray.d = dir;
  ray.tmin = FLT_EPSILON;  // ShaderRayRange::default (api/ty.rs:49-56)
  ray.tmax = FLT_MAX;
  ray.flags = F_CULL_BACK;  // ao.rs:209
  ray.mask = 0xffffffffu;
  ray.tlas_idx = 0;
  orc_hit h;
  orc_hit_meta hm;
  Counters cnt;
  uint32_t ext[2] = {0, 0};
  if (!traverse_one(s, ray, h, hm, cnt, ext)) return 1.0f;  // miss shader (ao.rs:304-312)
  // closest hit (ao.rs:252-291) with the bindless mesh normal
  // (bindless_mesh_bridge.rs:86-123)
  const auto &geo = s->geo_meta[ext[0]];
  const uint32_t tri = geo.primitive_start + h.prim_id;
  const uint32_t ia = s->indices[3 * (size_t)tri], ib = s->indices[3 * (size_t)tri + 1],
                 ic = s->indices[3 * (size_t)tri + 2];
  const M4 &inv = s->inst_transform_inv[ext[1]];
  const V3 pos = ray.o + ray.d * h.t;  // api/ctx.rs:209-211
  const V3 pa = s->vertices[ia], pb = s->vertices[ib], pc = s->vertices[ic];
  V3 gn = normalize(f_normal_to_world(inv, cross(pa - pb, pa - pc)));
  if (dot(ray.o - pos, gn) < 0.0f) gn = gn * -1.0f;
  V3 sn = gn;
  if (!s->normals.empty()) {
    const float w = 1.0f - h.u - h.v;
    const V3 n = s->normals[ia] * w + s->normals[ib] * h.u + s->normals[ic] * h.v;
    if (dot(n, n) > 0.0f) {
      sn = normalize(f_normal_to_world(inv, n));
      if (dot(gn, sn) < 0.0f) sn = sn * -1.0f;
    }
  }
  // tbn (sampling.rs:70-83) times the hemisphere sample
  const float sign = sn.z < 0.0f ? -1.0f : 1.0f;
  const float a = -1.0f / (sign + sn.z);
  const float b = sn.x * sn.y * a;
  const V3 t = normalize(v3(1.0f + sign * sn.x * sn.x * a, sign * b, -sign * sn.x));
  const V3 bt = normalize(v3(b, sign + sn.y * sn.y * a, -sn.y));
  const V3 aodir = t * local[0] + bt * local[1] + sn * local[2];
  RayIn occ;
  occ.o = pos;
  occ.d = aodir;
  occ.tmin = 0.01f;  // ao.rs:279-282
  occ.tmax = 100.0f;
  occ.flags = F_ACCEPT_FIRST_HIT_AND_END_SEARCH;  // ao.rs:276
  occ.mask = 0xffffffffu;
  occ.tlas_idx = 0;
  orc_hit h2;
  orc_hit_meta hm2;
  Counters cnt2;
  return traverse_one(s, occ, h2, hm2, cnt2) ? 0.0f : 1.0f;
}

}  // namespace

extern "C" {

// The per-pixel sampler of the AO ray generation: seed = xxhash32(x, y,
// sample_index) (sampler.rs:47-55), then n calls of next(); states[k],
// bits[k] and values[k] are the state after call k, its 32-bit hash and the
// float it returns.  Returns the seed.
uint32_t orc_ao_sampler(uint32_t x, uint32_t y, uint32_t sample_index, uint32_t n, uint32_t *states, uint32_t *bits,
                        float *values) {
  const uint32_t seed = xxhash32(x, y, sample_index);
  uint32_t st = seed;
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t probe = st;
    bits[k] = ao_pcg_bits(probe);
    values[k] = ao_pcg_next(st);
    states[k] = st;
  }
  return seed;
}

void orc_ao_hemisphere_sample(uint32_t s, float out[3]) {
  // hammersley_2d(s, 256) (sampling.rs:33-50): (s / 256, radical inverse)
  const float ux = (float)s / 256.0f;
  uint32_t r = 0;
  for (int k = 0; k < 32; k++) r |= ((s >> k) & 1u) << (31 - k);  // bit reversal
  const float uy = (float)r * 2.3283064e-10f;
  // sample_hemisphere_cos (sampling.rs:61-68) with 2 pi as an f32
  const float phi = (2.0f * 3.14159265358979323846f) * uy;
  const float cos_theta = std::sqrt(1.0f - ux);
  const float sin_theta = std::sqrt(1.0f - cos_theta * cos_theta);
  out[0] = (float)std::cos((double)phi) * sin_theta;
  out[1] = (float)std::sin((double)phi) * sin_theta;
  out[2] = cos_theta;
}

int orc_ao_render(const orc_scene *s, const orc_ao_params *P, float *accum, uint8_t *accum_rgba8, float *values,
                  int nthreads) {
  if (!s->built || P->width == 0 || P->height == 0) return -1;
  float local[3];
  orc_ao_hemisphere_sample(P->sample_index, local);
  const int64_t npix = (int64_t)P->width * P->height;
  const float prev_count = (float)P->sample_index;
  const float all = prev_count + 1.0f;
  const bool write = P->sample_index < 256u;  // MAX_SAMPLE (ao.rs:5, 232-243)
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
#else
  (void)nthreads;
#endif
  for (int64_t i = 0; i < npix; i++) {
    const uint32_t x = (uint32_t)(i % P->width), y = (uint32_t)(i / P->width);
    const float val = ao_value(s, *P, x, y, local);
    if (values) values[i] = val;
    if (accum && write) accum[i] = (accum[i] * prev_count + val) / all;
    if (accum_rgba8) {
      const float prev = (float)accum_rgba8[4 * i] / 255.0f;
      float nv = write ? (prev * prev_count + val) / all : prev;
      nv = std::fmin(std::fmax(nv, 0.0f), 1.0f);
      const uint8_t q = (uint8_t)std::nearbyint(nv * 255.0f);
      accum_rgba8[4 * i] = accum_rgba8[4 * i + 1] = accum_rgba8[4 * i + 2] = q;
      accum_rgba8[4 * i + 3] = 255;
    }
  }
  return 0;
}

}  // extern "C"
