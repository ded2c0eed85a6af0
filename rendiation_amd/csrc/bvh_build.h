// bvh_build.h -- host SAH BVH builder of the ray-query path.
//
// Builds the same tree as the reference's FlattenBVH::new with SAH(bins) and
// TreeBuildOption (content/space/src/bvh/mod.rs:55-79, strategy.rs:8-285,
// apply.rs:9-78, utils.rs:192-238): binned SAH on the longest axis of the
// node bounds, first strict cost minimum, stable bucket order, median split
// when every centre lands in one bin.  Nodes come out in DFS preorder, as in
// the reference, with the stackless hit/miss links of naive/mod.rs:611-632.
#ifndef RDRT_BVH_BUILD_H
#define RDRT_BVH_BUILD_H

#include <cstddef>
#include <cstdint>
#include <vector>

namespace rdrt {

struct Box {
  float mn[3], mx[3];
};
Box box_empty();

// fminf / fmaxf, inline: a NaN operand is dropped (both NaN give NaN) and of
// two equal operands (+0 and -0) the second is returned, as glibc's are.
// Those are out-of-line library calls, which made the box unions most of
// the host builders' time; Rust's f32::min / max follow the same NaN rule.
inline float min_num(float a, float b) { return (a < b || b != b) ? a : b; }
inline float max_num(float a, float b) { return (a > b || b != b) ? a : b; }

struct TreeOptions {
  uint32_t bin_size;   // leaf while count <= bin_size
  uint32_t max_depth;  // leaf at depth >= max_depth
  uint32_t bins;       // SAH bins
};

struct TreeNode {
  Box box;
  uint32_t begin, end;  // range into `order`
  uint32_t size;        // nodes in this subtree (1 for a leaf)
  bool inner;
};

struct Tree {
  std::vector<TreeNode> nodes;  // DFS preorder; left child = i + 1
  std::vector<uint32_t> order;  // sorted primitive index
  uint32_t depth = 0;           // deepest node (root = 0)
  // stackless links (naive/mod.rs:611-632): hit = first child or the miss
  // link of a leaf; miss = next node after the subtree, or kInvalid.
  uint32_t hit(uint32_t i) const;
  uint32_t miss(uint32_t i) const;
};

// Builds the tree over `boxes`.  `threads` > 1 builds large subtrees
// concurrently; the result does not depend on it.
Tree build_tree(const std::vector<Box> &boxes, const TreeOptions &opt, unsigned threads);

// The tree the fast traversal walks (any tree gives the same closest hit):
// binned SAH over all three axes of the centre bounds, `bins` bins, leaves of
// at most `max_leaf` primitives closed when the SAH prefers them, a median
// split when all centres coincide.  Same Tree structure as build_tree.
struct FastTreeOptions {
  uint32_t bins = 16;
  uint32_t max_leaf = 4;   // <= 31
  uint32_t max_depth = 60;
  float node_cost = 1.0f;  // SAH cost of visiting a child pair, per unit area ratio
  float prim_cost = 1.0f;  // SAH cost of one primitive test
};
Tree build_fast_tree(const std::vector<Box> &boxes, const FastTreeOptions &opt, unsigned threads);
// SAH cost of a tree (node_cost per inner node visit, prim_cost per
// primitive test, weighted by area relative to the root).
double sah_cost(const Tree &t, const FastTreeOptions &opt);

}  // namespace rdrt

#endif
