// bvh_build.cpp -- host SAH BVH builder (see bvh_build.h).
//
// Compiled with -ffp-contract=off: every bound, centre, bin index and SAH cost
// is one IEEE FP32 rounding per operation, in the operation order of
// content/space/src/bvh/strategy.rs and math/geometry/src/dimension3/box3.rs,
// so the tree equals the reference's node for node.
//
// Design: centres are kept structure-of-arrays, a node's bin ids are computed
// in one pass, the stable per-bin reorder is a counting scatter through a
// scratch buffer, and subtrees of large nodes are built on separate threads
// into their own preorder vectors that are spliced afterwards (nodes store
// subtree sizes, not absolute indices, so splicing needs no fix-up).

#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <thread>

#include "rdrt_layout.h"

namespace rdrt {

Box box_empty() {
  const float inf = std::numeric_limits<float>::infinity();
  return Box{{inf, inf, inf}, {-inf, -inf, -inf}};
}

uint32_t Tree::miss(uint32_t i) const {
  uint32_t next = i + nodes[i].size;
  return next < nodes.size() ? next : kInvalid;
}
uint32_t Tree::hit(uint32_t i) const { return nodes[i].inner ? i + 1 : miss(i); }

namespace {

// Rust f32::min/max drop a NaN operand, like fminf/fmaxf (min_num, max_num).
inline void grow(Box &b, const Box &o) {
  for (int k = 0; k < 3; k++) {
    b.mn[k] = min_num(b.mn[k], o.mn[k]);
    b.mx[k] = max_num(b.mx[k], o.mx[k]);
  }
}

// box3.rs:117-133: ties resolve to Z, then Y.
inline int longest_axis(const Box &b) {
  float x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
  if (x > y) return x > z ? 0 : 2;
  return y > z ? 1 : 2;
}

// box3.rs:5-11
inline float surface_area(const Box &b) {
  float w = b.mx[0] - b.mn[0], h = b.mx[1] - b.mn[1], d = b.mx[2] - b.mn[2];
  return 2.0f * (w * h + w * d + h * d);
}

// Rust's saturating float -> usize cast (NaN and negatives -> 0).
inline size_t saturating_index(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 1.8446744e19f) return std::numeric_limits<size_t>::max();
  return (size_t)f;
}

struct Builder {
  const std::vector<Box> &boxes;
  TreeOptions opt;
  std::vector<float> centre[3];
  std::vector<uint32_t> order;
  std::vector<uint32_t> scratch;  // disjoint per node range, so safe across threads

  Builder(const std::vector<Box> &b, const TreeOptions &o) : boxes(b), opt(o) {
    size_t n = b.size();
    for (int k = 0; k < 3; k++) centre[k].resize(n);
    order.resize(n);
    scratch.resize(n);
    for (size_t i = 0; i < n; i++) {
      order[i] = (uint32_t)i;
      for (int k = 0; k < 3; k++) centre[k][i] = (b[i].mn[k] + b[i].mx[k]) * 0.5f;  // box3.rs:90-93
    }
  }

  Box range_box(uint32_t b, uint32_t e) const {
    Box r = box_empty();
    for (uint32_t i = b; i < e; i++) grow(r, boxes[order[i]]);
    return r;
  }

  struct Split {
    Box lb, rb;
    uint32_t mid;
  };

  // strategy.rs:66-87 + apply.rs:304-333.  Only the order inside each half
  // follows std::nth_element rather than Rust's select_nth_unstable_by.
  Split median_split(const Box &pb, uint32_t b, uint32_t e) {
    int axis = longest_axis(pb);
    uint32_t mid = (e + b) / 2;
    uint32_t half = (e - b) / 2;
    if (half != 0) {
      const float *c = centre[axis].data();
      std::nth_element(order.begin() + b, order.begin() + b + half, order.begin() + e,
                       [c](uint32_t x, uint32_t y) {
                         float cx = c[x], cy = c[y];
                         if (std::isnan(cx) || std::isnan(cy)) return true;  // unwrap_or(Less)
                         return cx < cy;
                       });
    }
    return Split{range_box(b, mid), range_box(mid, e), mid};
  }

  // strategy.rs:201-284
  Split sah_split(const Box &pb, uint32_t b, uint32_t e) {
    const uint32_t nb = opt.bins;
    int axis = longest_axis(pb);
    float a0 = pb.mn[axis], a1 = pb.mx[axis];
    float step = (a1 - a0) / (float)nb;
    const float *c = centre[axis].data();

    std::vector<uint32_t> bin(e - b);
    std::vector<uint32_t> count(nb, 0);
    std::vector<Box> bbox(nb, box_empty());
    for (uint32_t i = b; i < e; i++) {
      uint32_t p = order[i];
      size_t w = saturating_index(std::floor((c[p] - a0) / step));
      if (w == nb) w -= 1;
      if (w > nb) w = nb - 1;  // unreachable with finite bounds
      bin[i - b] = (uint32_t)w;
      count[w]++;
      grow(bbox[w], boxes[p]);
    }
    uint32_t empty = 0;
    for (uint32_t k = 0; k < nb; k++) empty += count[k] == 0 ? 1u : 0u;
    if (empty == nb - 1) return median_split(pb, b, e);

    // first strict minimum of area(L)*|L| + area(R)*|R|
    uint32_t best = 0;
    float best_cost = std::numeric_limits<float>::infinity();
    Box best_l = box_empty(), best_r = box_empty();
    uint32_t best_lc = 0;
    for (uint32_t s = 0; s + 1 < nb; s++) {
      Box l = box_empty(), r = box_empty();
      uint32_t lc = 0, rc = 0;
      for (uint32_t k = 0; k <= s; k++) {
        lc += count[k];
        grow(l, bbox[k]);
      }
      for (uint32_t k = s + 1; k < nb; k++) {
        rc += count[k];
        grow(r, bbox[k]);
      }
      float cost = surface_area(l) * (float)lc + surface_area(r) * (float)rc;
      if (cost < best_cost) {
        best_cost = cost;
        best = s;
        best_l = l;
        best_r = r;
        best_lc = lc;
      }
    }
    (void)best;
    // stable reorder by bin
    std::vector<uint32_t> off(nb, 0);
    for (uint32_t k = 1; k < nb; k++) off[k] = off[k - 1] + count[k - 1];
    for (uint32_t i = b; i < e; i++) scratch[b + off[bin[i - b]]++] = order[i];
    std::copy(scratch.begin() + b, scratch.begin() + e, order.begin() + b);
    return Split{best_l, best_r, b + best_lc};
  }

  // strategy.rs:11-43: preorder, left subtree before right.
  void build(const Box &box, uint32_t b, uint32_t e, uint32_t depth, std::vector<TreeNode> &out,
             uint32_t &maxd, int spawn_levels) {
    if (depth > maxd) maxd = depth;
    size_t self = out.size();
    out.push_back(TreeNode{box, b, e, 1, false});
    if (!(depth < opt.max_depth && e - b > opt.bin_size)) return;
    Split sp = sah_split(box, b, e);
    out[self].inner = true;
    if (spawn_levels > 0 && e - b >= 16384) {
      std::vector<TreeNode> left;
      uint32_t lmax = 0;
      std::thread worker([&] { build(sp.lb, b, sp.mid, depth + 1, left, lmax, spawn_levels - 1); });
      std::vector<TreeNode> right;
      uint32_t rmax = 0;
      build(sp.rb, sp.mid, e, depth + 1, right, rmax, spawn_levels - 1);
      worker.join();
      out.insert(out.end(), left.begin(), left.end());
      out.insert(out.end(), right.begin(), right.end());
      maxd = std::max(maxd, std::max(lmax, rmax));
    } else {
      build(sp.lb, b, sp.mid, depth + 1, out, maxd, 0);
      build(sp.rb, sp.mid, e, depth + 1, out, maxd, 0);
    }
    out[self].size = (uint32_t)(out.size() - self);
  }
};

}  // namespace

namespace {

// ---- fast-traversal tree: full binned SAH on the centre bounds ----
constexpr uint32_t kMaxFastBins = 32;

struct FastBuilder {
  const std::vector<Box> &boxes;
  FastTreeOptions opt;
  std::vector<float> centre[3];
  std::vector<uint32_t> order, scratch;

  FastBuilder(const std::vector<Box> &b, const FastTreeOptions &o) : boxes(b), opt(o) {
    size_t n = b.size();
    for (int k = 0; k < 3; k++) centre[k].resize(n);
    order.resize(n);
    scratch.resize(n);
    for (size_t i = 0; i < n; i++) {
      order[i] = (uint32_t)i;
      for (int k = 0; k < 3; k++) centre[k][i] = (b[i].mn[k] + b[i].mx[k]) * 0.5f;
    }
  }

  Box range_box(uint32_t b, uint32_t e) const {
    Box r = box_empty();
    for (uint32_t i = b; i < e; i++) grow(r, boxes[order[i]]);
    return r;
  }

  // Returns the split point (b < mid < e) after reordering [b, e) and the
  // child boxes, or b when a leaf is cheaper (only for count <= max_leaf).
  // One pass bins every primitive on all three axes.
  uint32_t split(const Box &nb, uint32_t b, uint32_t e, Box &lb, Box &rb) {
    const uint32_t n = e - b, K = opt.bins;
    float cmn[3], cmx[3];
    for (int k = 0; k < 3; k++) {
      cmn[k] = std::numeric_limits<float>::infinity();
      cmx[k] = -std::numeric_limits<float>::infinity();
    }
    for (uint32_t i = b; i < e; i++)
      for (int k = 0; k < 3; k++) {
        const float c = centre[k][order[i]];
        cmn[k] = min_num(cmn[k], c);
        cmx[k] = max_num(cmx[k], c);
      }
    float scale[3];
    bool live[3];
    for (int k = 0; k < 3; k++) {
      const float ext = cmx[k] - cmn[k];
      live[k] = ext > 0.0f;
      scale[k] = live[k] ? (float)K / ext : 0.0f;
    }
    if (!live[0] && !live[1] && !live[2]) {
      // every centre coincides: split the range in half (object median)
      lb = range_box(b, b + n / 2);
      rb = range_box(b + n / 2, e);
      return b + n / 2;
    }
    uint32_t cnt[3 * kMaxFastBins];
    Box bb[3 * kMaxFastBins];
    for (uint32_t k = 0; k < 3 * K; k++) {
      cnt[k] = 0u;
      bb[k] = box_empty();
    }
    for (uint32_t i = b; i < e; i++) {
      const uint32_t p = order[i];
      const Box &pb = boxes[p];
      for (int k = 0; k < 3; k++) {
        if (!live[k]) continue;
        const uint32_t w = bin_of(centre[k][p], cmn[k], scale[k], K);
        cnt[k * K + w]++;
        grow(bb[k * K + w], pb);
      }
    }
    const float parent_area = surface_area(nb);
    double best_cost = std::numeric_limits<double>::infinity();
    int best_axis = -1;
    uint32_t best_bin = 0;
    // The sweep visits the occupied bins only.  An empty bin adds nothing to
    // a side's box (min / max with an empty box are exact) or count, so the
    // cost after it equals the cost after the last occupied bin before it,
    // and the first strict minimum -- the split chosen -- is the same.
    double right_cost[kMaxFastBins];
    uint32_t occ[kMaxFastBins];
    for (int axis = 0; axis < 3; axis++) {
      if (!live[axis]) continue;
      const uint32_t *c = &cnt[axis * K];
      const Box *bx = &bb[axis * K];
      uint32_t no = 0;
      for (uint32_t k = 0; k < K; k++)
        if (c[k]) occ[no++] = k;
      Box r = box_empty();
      uint32_t rc = 0;
      for (uint32_t j = no; j-- > 1;) {  // right side from occupied bin j on
        grow(r, bx[occ[j]]);
        rc += c[occ[j]];
        right_cost[j] = (double)surface_area(r) * rc;
      }
      Box l = box_empty();
      uint32_t lc = 0;
      for (uint32_t j = 0; j + 1 < no; j++) {  // left side up to occupied bin j: 0 < lc < n
        grow(l, bx[occ[j]]);
        lc += c[occ[j]];
        const double cost = (double)surface_area(l) * lc + right_cost[j + 1];
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_bin = occ[j];
        }
      }
    }
    const double split_cost =
        opt.node_cost + opt.prim_cost * best_cost / std::max((double)parent_area, 1e-30);
    if (n <= opt.max_leaf && split_cost >= opt.prim_cost * n) return b;
    lb = box_empty();
    rb = box_empty();
    for (uint32_t k = 0; k < K; k++) grow(k <= best_bin ? lb : rb, bb[best_axis * K + k]);
    // stable partition by the chosen bin boundary
    uint32_t lo = b, hi = 0;
    for (uint32_t i = b; i < e; i++) {
      const uint32_t p = order[i];
      if (bin_of(centre[best_axis][p], cmn[best_axis], scale[best_axis], K) <= best_bin) order[lo++] = p;
      else scratch[b + hi++] = p;
    }
    std::copy(scratch.begin() + b, scratch.begin() + b + hi, order.begin() + lo);
    return lo;
  }

  static uint32_t bin_of(float c, float mn, float scale, uint32_t K) {
    return (uint32_t)std::min((float)(K - 1), std::max(0.0f, (c - mn) * scale));
  }

  void build(const Box &box, uint32_t b, uint32_t e, uint32_t depth, std::vector<TreeNode> &out, uint32_t &maxd,
             int spawn_levels) {
    if (depth > maxd) maxd = depth;
    size_t self = out.size();
    out.push_back(TreeNode{box, b, e, 1, false});
    if (e - b <= 1 || depth >= opt.max_depth) return;
    Box lb, rb;
    const uint32_t mid = split(box, b, e, lb, rb);
    if (mid == b) return;  // leaf
    out[self].inner = true;
    if (spawn_levels > 0 && e - b >= 16384) {
      std::vector<TreeNode> left;
      uint32_t lmax = 0;
      std::thread worker([&] { build(lb, b, mid, depth + 1, left, lmax, spawn_levels - 1); });
      std::vector<TreeNode> right;
      uint32_t rmax = 0;
      build(rb, mid, e, depth + 1, right, rmax, spawn_levels - 1);
      worker.join();
      out.insert(out.end(), left.begin(), left.end());
      out.insert(out.end(), right.begin(), right.end());
      maxd = std::max(maxd, std::max(lmax, rmax));
    } else {
      build(lb, b, mid, depth + 1, out, maxd, 0);
      build(rb, mid, e, depth + 1, out, maxd, 0);
    }
    out[self].size = (uint32_t)(out.size() - self);
  }
};

}  // namespace

Tree build_fast_tree(const std::vector<Box> &boxes, const FastTreeOptions &opt_in, unsigned threads) {
  FastTreeOptions opt = opt_in;
  if (opt.bins < 2) opt.bins = 16;
  opt.bins = std::min(opt.bins, kMaxFastBins);
  opt.max_leaf = std::max(1u, std::min(opt.max_leaf, kLeafMaxCount));
  FastBuilder bld(boxes, opt);
  Tree t;
  Box root = bld.range_box(0, (uint32_t)boxes.size());
  int levels = 0;
  while ((1u << levels) < threads && levels < 6) levels++;
  bld.build(root, 0, (uint32_t)boxes.size(), 0, t.nodes, t.depth, levels);
  t.order = std::move(bld.order);
  return t;
}

double sah_cost(const Tree &t, const FastTreeOptions &opt) {
  if (t.nodes.empty()) return 0.0;
  const double root = std::max((double)surface_area(t.nodes[0].box), 1e-30);
  double c = 0.0;
  for (const TreeNode &n : t.nodes) {
    const double a = surface_area(n.box) / root;
    c += n.inner ? a * opt.node_cost : a * opt.prim_cost * (n.end - n.begin);
  }
  return c;
}

Tree build_tree(const std::vector<Box> &boxes, const TreeOptions &opt_in, unsigned threads) {
  TreeOptions opt = opt_in;
  if (opt.bins < 2) opt.bins = 4;
  Builder bld(boxes, opt);
  Tree t;
  Box root = bld.range_box(0, (uint32_t)boxes.size());  // utils.rs:224-238
  int levels = 0;
  while ((1u << levels) < threads && levels < 6) levels++;
  bld.build(root, 0, (uint32_t)boxes.size(), 0, t.nodes, t.depth, levels);
  t.order = std::move(bld.order);
  return t;
}

}  // namespace rdrt
