// scene.h -- host scene sources and the built device-ready arrays.
#ifndef RDRT_SCENE_H
#define RDRT_SCENE_H

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rdrt.h"
#include "bvh_build.h"
#include "rdrt_layout.h"

namespace rdrt {

// One triangle geometry of a BLAS (BottomLevelAccelerationStructureBuildSource,
// shader/ray-tracing/src/api/backend.rs:104-128), copied at creation.
struct GeometrySrc {
  std::vector<float> pos;
  std::vector<uint32_t> idx;  // empty = non-indexed
  bool indexed = false;
  uint32_t ntris = 0;
  uint32_t flags = 0;
  // shading attributes (path tracer): per-vertex normals (xyz, empty = none)
  // and per-triangle material offsets (empty = 0)
  std::vector<float> normals;
  std::vector<uint32_t> tri_material;
  std::vector<float> uvs;  // texture coordinates (uv per vertex, empty = none)
  // any-hit behaviour per triangle (RDRT_ANYHIT_*) of hits on this geometry
  // when it is not opaque for the ray (empty = accept every hit)
  std::vector<uint8_t> anyhit;
};

struct SceneSrc {
  std::vector<char> blas_live;
  std::vector<std::vector<GeometrySrc>> blas;
  std::vector<char> tlas_live;
  std::vector<std::vector<rdrt_instance>> tlas;
  std::vector<uint32_t> binding;
  TreeOptions blas_opt{2, 50, 4};
  TreeOptions tlas_opt{10, 50, 4};
};

// Allocator whose new elements are left uninitialized (resize does not zero
// them).  The scene arrays below are sized once and then every element is
// written by the parallel assembly; value-initializing them first was a
// serial zero fill of ~1.3 GB (plus its page faults) for a scene of 8.4M
// triangles, longer than the assembly itself.
template <class T>
struct UninitAlloc : std::allocator<T> {
  using value_type = T;
  template <class U>
  struct rebind {
    using other = UninitAlloc<U>;
  };
  UninitAlloc() = default;
  template <class U>
  UninitAlloc(const UninitAlloc<U> &) noexcept {}
  template <class U>
  void construct(U *) noexcept {}  // default-init: no zeroing
  template <class U, class... Args>
  void construct(U *p, Args &&...args) {
    ::new ((void *)p) U(std::forward<Args>(args)...);
  }
};
template <class T>
using BigVec = std::vector<T, UninitAlloc<T>>;

struct BuiltScene {
  // reference layout (naive/mod.rs:322-494)
  BigVec<RefNode> tri_forest;
  std::vector<RefNode> tlas_forest;
  BigVec<uint32_t> redirect, indices;
  BigVec<uint32_t> tri_ids;  // fast-traversal triangle slot -> global triangle index
  BigVec<float> vertices;    // xyz
  // shading attributes: normals parallel to vertices (empty when no geometry
  // has any), material offset per triangle (empty when all are 0), largest
  // material offset per geometry
  BigVec<float> normals;
  BigVec<uint32_t> tri_material;
  std::vector<uint32_t> geo_max_material;
  BigVec<uint8_t> anyhit;  // any-hit behaviour per triangle (empty when no geometry has a table)
  BigVec<float> uvs;       // 2 per vertex when any geometry has them (zeros elsewhere), else empty
  // instance records by source position: inst_lookup[tlas_inst_first[t] + k]
  // is the record of instance k of TLAS t
  std::vector<uint32_t> inst_lookup, tlas_inst_first;
  // shared records
  std::vector<GeoRec> geos;
  std::vector<InstRec> insts;
  std::vector<float> inst_fwd;  // 16 per instance record: its object-to-world transform as given
  std::vector<TlasRoot> tlas_roots;
  std::vector<uint32_t> binding;
  // child-pair layout
  BigVec<PairNode> tri_pairs;
  std::vector<PairNode> tlas_pairs;
  BigVec<TriPre> tris;
  uint32_t max_blas_depth = 0, max_tlas_depth = 0;
  uint32_t blas_stack = 0, tlas_stack = 0;  // deepest push count of the child-pair trees
  // child-quad layout (collapsed from the child-pair trees; same leaves)
  BigVec<QuadNode> tri_quads;
  std::vector<QuadNode> tlas_quads;
  BigVec<QuadNodeQ> tri_qquads;  // quantized twins (same indices and refs)
  std::vector<QuadNodeQ> tlas_qquads;
  std::vector<uint32_t> tlas_quad_roots;      // per TLAS (kInvalid when deleted)
  uint32_t blas_stack4 = 0, tlas_stack4 = 0;  // deepest push count of the child-quad trees
  // per BLAS: object-space box (union of its geometries), whether it has
  // one, and its geometry range (inputs of the TLAS layer)
  std::vector<Box> blas_box;
  std::vector<char> blas_has_box;
  std::vector<uint32_t> blas_geo_begin, blas_geo_end;
  double build_ms = 0.0;
  // cumulative ms at the end of: trees, forest + pairs + quads, triangle
  // records, quantized quads (then the TLAS layer and checks)
  double phase_ms[4] = {0.0, 0.0, 0.0, 0.0};
  // true: `tris` and `tri_qquads` stay empty; the commit computes them on
  // the device from the uploaded vertices, indices, tri_ids and tri_quads
  bool device_records = false;
};

// Quantized twin of a child-quad node (host; the device twin is in
// refit_kernels.hip).
QuadNodeQ quantize_quad(const QuadNode &q);

// Builds every BVH and both layouts.  Returns RDRT_OK or an RDRT_ERR_* code
// with a message in `err`.
// Builds the two trees of every geometry (the reference's tree and the
// fast kernels' tree) from their triangle boxes, with the host builders'
// semantics, all geometries together.  Implemented on the device
// (device_build.hip).
struct DeviceTreeBuilder {
  virtual ~DeviceTreeBuilder() = default;
  virtual int build(const std::vector<const std::vector<Box> *> &geos, const TreeOptions &ref_opt,
                    const FastTreeOptions &fast_opt, std::vector<Tree> &ref, std::vector<Tree> &fast,
                    std::string &err) = 0;
};

// `trees`, when given, builds every BLAS tree (one geometry after another)
// instead of the host builders.  `device_records`: leave the BLAS triangle
// records and quantized quads to the device (BuiltScene::device_records).
int build_scene(const SceneSrc &src, BuiltScene &out, std::string &err, DeviceTreeBuilder *trees = nullptr,
                bool device_records = false);
// BLAS node indices of every layout grouped by depth, deepest level first:
// the order in which a refit recomputes boxes bottom-up (a level only reads
// the level below).  *_off[l]..*_off[l + 1] is level l of that order.
struct RefitLevels {
  std::vector<uint32_t> forest, pairs, quads;
  std::vector<uint32_t> forest_off, pairs_off, quads_off;
};
RefitLevels refit_levels(const BuiltScene &b);

// Nearest hits of n host rays (origin, direction: 6 f32 each) against one
// geometry (global index) through its reference tree, with the reference's
// host picking semantics (rdrt_pick).
int pick_nearest(const BuiltScene &b, uint32_t geo, const float *rays, uint32_t n, uint32_t side, rdrt_pick_hit *out,
                 std::string &err);
// Every hit of one ray against one geometry, in the reference's visit order
// (rdrt_pick_all); *count is the number of hits, of which the first
// `capacity` are written.
int pick_all(const BuiltScene &b, uint32_t geo, const float *ray, uint32_t side, rdrt_pick_hit *out, uint32_t capacity,
             uint32_t *count, std::string &err);

// Rebuilds the TLAS layer of `out` (instances, TLAS trees in every layout)
// from out.blas_box and the BLAS geometry ranges.
int build_tlas_layer(const SceneSrc &src, BuiltScene &out, std::string &err);

}  // namespace rdrt

#endif
