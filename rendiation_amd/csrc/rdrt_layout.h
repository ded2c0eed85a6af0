// rdrt_layout.h -- device data layouts shared by the host scene builder and
// the gfx950 kernels.  Plain structs only (no HIP or STL types) so the file is
// included from both g++-compiled host code and hipcc-compiled kernels.
#ifndef RDRT_LAYOUT_H
#define RDRT_LAYOUT_H

#include <stdint.h>

namespace rdrt {

constexpr uint32_t kInvalid = 0xffffffffu;

// Reference forest node, 48 B (naive/mod.rs:46-53 DeviceBVHNode): box,
// stackless hit/miss links and the content range (into the triangle redirect
// for BLAS forests, into the instance array for TLAS forests).
struct RefNode {
  float mn[3];
  uint32_t hit_next;
  float mx[3];
  uint32_t miss_next;
  uint32_t begin, end;
  uint32_t tail[2];
};
static_assert(sizeof(RefNode) == 48, "RefNode must be 48 B");

// Child-pair node, 64 B = one 64-byte line: the boxes of both children of one
// inner node of the same tree, so one fetch tests both children.  Child refs
// use the encoding below.
struct alignas(64) PairNode {
  float lmn[3];
  uint32_t lref;
  float lmx[3];
  uint32_t pad0;
  float rmn[3];
  uint32_t rref;
  float rmx[3];
  uint32_t pad1;
};
static_assert(sizeof(PairNode) == 64, "PairNode must be 64 B");

// Child-quad node, 128 B = one 128-byte line: the boxes of up to four
// children in SoA order (x bounds of all four, then y, then z), so eight
// 16-byte loads test four children and the tree is half as deep as the
// child-pair tree it is collapsed from.  Unused slots have ref kInvalid.
// Child refs use the encoding below, inner refs indexing QuadNodes.
struct alignas(128) QuadNode {
  float mnx[4], mxx[4], mny[4], mxy[4], mnz[4], mxz[4];
  uint32_t ref[4];
  uint32_t pad[4];
};
static_assert(sizeof(QuadNode) == 128, "QuadNode must be 128 B");

// Quantized child-quad node, 64 B: the node box's lower corner, a
// power-of-two step per axis and every child box as 8-bit multiples of the
// steps from that corner (byte c of each word = child c).  Lower bounds are
// rounded down and upper bounds up, checked with the decoding arithmetic
// (corner + q * step, one rounding), so a decoded box always contains the
// child's exact box: the traversal may enter a few more nodes, never fewer,
// and the triangle tests are unchanged.  Node i has the refs of QuadNode i.
struct alignas(64) QuadNodeQ {
  float ox, oy, oz, sx;  // lower corner, x step
  uint32_t ref[4];
  uint32_t lox, hix, loy, hiy;
  uint32_t loz, hiz;
  float sy, sz;
};
static_assert(sizeof(QuadNodeQ) == 64, "QuadNodeQ must be 64 B");

// Child ref: bit 31 clear = index of an inner PairNode; bit 31 set = leaf
// with `count` (bits 0..4, < 32) primitives from `start` (bits 5..30).
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafCountBits = 5;
constexpr uint32_t kLeafMaxCount = (1u << kLeafCountBits) - 1;
constexpr uint32_t kLeafMaxStart = (1u << (31 - kLeafCountBits)) - 1;
inline constexpr uint32_t leaf_ref(uint32_t start, uint32_t count) {
  return kLeafBit | (start << kLeafCountBits) | count;
}

// Triangle pre-pass record, 64 B, in BVH leaf order.  Every field is what the
// reference ray/triangle test computes from the three vertices before it
// touches the ray (wavefront_compute/geometry/mod.rs:115-117,138-140,144),
// computed on the host with the same single-rounding FP32 operations, so the
// per-ray arithmetic that remains gives bit-identical t, u, v.
struct alignas(16) TriPre {
  float v0[3];
  float uu;
  float e1[3];
  float uv;
  float e2[3];
  float vv;
  float n[3];
  float inverse_d;
};
static_assert(sizeof(TriPre) == 64, "TriPre must be 64 B");

// One geometry of a BLAS (naive/mod.rs:66-71 GeometryMetaInfo plus the root
// box and the child-pair root ref), 48 B.
struct alignas(16) GeoRec {
  float mn[3];
  uint32_t root_ref;  // child-pair ref of the root
  float mx[3];
  uint32_t ref_root;  // root index in the reference forest
  uint32_t primitive_start, geometry_idx, flags;
  uint32_t quad_root;  // child-quad ref of the root
};
static_assert(sizeof(GeoRec) == 48, "GeoRec must be 48 B");

// TLAS instance, 128 B: inverse transform, world bounds with mask and merged
// flags (naive/mod.rs:25-41), geometry range of its BLAS, custom index and the
// instance's position in the TLAS source array.
struct alignas(16) InstRec {
  float inv[16];
  float mn[3];
  uint32_t mask;
  float mx[3];
  uint32_t flags;
  uint32_t geo_begin, geo_end, custom_index, order;
  uint32_t sbt_offset, blas, pad[2];
};
static_assert(sizeof(InstRec) == 128, "InstRec must be 128 B");

// Root of one TLAS: its box and child-pair ref, plus the reference-forest
// root (kInvalid for a deleted TLAS).
struct alignas(16) TlasRoot {
  float mn[3];
  uint32_t root_ref;
  float mx[3];
  uint32_t ref_root;
};
static_assert(sizeof(TlasRoot) == 32, "TlasRoot must be 32 B");

// Lambertian material (content/lighting/transport/src/surface/microfacet/
// device/diffuse.rs): albedo and emitted radiance, 32 B.
struct alignas(16) Material {
  float albedo[3];
  float pad0;
  float emission[3];
  float pad1;
};
static_assert(sizeof(Material) == 32, "Material must be 32 B");

}  // namespace rdrt

#endif
