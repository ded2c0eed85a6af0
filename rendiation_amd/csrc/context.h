// context.h -- the rdrt_ctx behind the C-ABI, shared by the API sources.
#ifndef RDRT_CONTEXT_H
#define RDRT_CONTEXT_H

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/rdrt.h"
#include "pt.h"
#include "scene.h"
#include "task_graph.h"
#include "trace_kernels.h"

namespace rdrt {

int32_t fail(int32_t code, const std::string &msg);

#define RDRT_HIP(expr)                                                                                        \
  do {                                                                                                        \
    hipError_t e_ = (expr);                                                                                   \
    if (e_ != hipSuccess) return ::rdrt::fail(RDRT_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
struct DevArray {
  T *ptr = nullptr;
  size_t count = 0;
};

template <class T, class A>
hipError_t upload(DevArray<T> &a, const std::vector<T, A> &h, uint64_t &bytes) {
  a.count = h.size();
  size_t n = h.empty() ? 1 : h.size();  // keep a valid pointer for empty arrays
  hipError_t e = hipMalloc(&a.ptr, n * sizeof(T));
  if (e != hipSuccess) return e;
  bytes += n * sizeof(T);
  if (!h.empty()) e = hipMemcpy(a.ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

template <class T>
void release(DevArray<T> &a) {
  if (a.ptr) (void)hipFree(a.ptr);
  a.ptr = nullptr;
  a.count = 0;
}

// Grow-only device buffer.
struct Scratch {
  void *ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&ptr, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  void free() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

struct DeviceScene {
  DevArray<PairNode> tri_pairs, tlas_pairs;
  DevArray<QuadNode> tri_quads, tlas_quads;
  DevArray<QuadNodeQ> tri_qquads, tlas_qquads;
  DevArray<TriPre> tris;
  DevArray<uint32_t> redirect, tri_ids, indices, binding, tri_material, tlas_quad_roots, inst_lookup, tlas_inst_first;
  DevArray<uint8_t> anyhit;
  DevArray<float> vertices, normals, uvs, inst_fwd;
  DevArray<GeoRec> geos;
  DevArray<InstRec> insts;
  DevArray<TlasRoot> tlas_roots;
  DevArray<RefNode> tri_forest, tlas_forest;
  uint64_t bytes = 0;
};

// Path-tracer state of one context (grown on demand).
struct PtState {
  uint32_t capacity = 0;  // paths
  uint32_t pixels = 0;
  Scratch rays, shadow_rays, hits, shadow_hits, hitx, beta, L, nee, qa, qb, qs, counts, tmp_a, tmp_b, blk;
  Scratch materials, record, views, work, tflags;
  // adjoint scratch: reproducible sums (int64 sums, flags, header), the
  // replay's per-bounce gradient terms, the loss partials
  Scratch det_acc, det_flags, det_hdr, grec, loss_part;
  std::vector<float> host_views;  // contents of `views`
  uint32_t nmaterials_uploaded = 0;
  bool materials_dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  PtLaunch cached{};
  TraceTune cached_tune{};  // launch shape the graph was captured with
  bool have_cached = false;
  uint32_t last_bounces = 0;
  float last_ms = 0.0f;
  // diagnostics (eager frames)
  std::vector<hipEvent_t> ev_begin, ev_end;
  Scratch diag_counters, diag_totals;
  uint32_t last_diag = 0, last_launches = 0;
  // last frame (for the adjoint)
  PtLaunch last_frame{};
  bool last_recorded = false;
  void free_all();
};

// Batched-tracing state of one context (grown on demand).
struct BatchState {
  Scratch wrays, hitx, qa, qb, counts, tmp, blk, work;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> cached;  // launch key of the captured graph
  std::vector<uint32_t> host_counts;
  uint32_t rounds = 0, nscenes = 1, mode = 0, launches = 0;
  void free_all();
};

void free_device_scene(DeviceScene &d);

// Ambient-occlusion pipeline state of one context (api_ao.hip).
struct AoState;
void free_ao(AoState *&s);
void drop_ao_capture(AoState *s);

}  // namespace rdrt

struct rdrt_ctx {
  rdrt_config cfg{};
  bool has_device = false;
  hipStream_t stream = nullptr;       // user stream (may be the null stream)
  hipStream_t work_stream = nullptr;  // capturable stream: the user stream, or one owned by the context
  bool owns_work_stream = false;
  rdrt::SceneSrc src;
  std::vector<rdrt::GeometrySrc> pending;
  bool in_blas = false;
  rdrt::BuiltScene built;
  bool built_valid = false;
  bool uploaded = false;
  rdrt::DeviceScene dev;
  uint32_t *status = nullptr;  // device status word of the kernels
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;       // ev0 / ev1 bracket the last rdrt_trace call
  bool call_timed = false;  // ev0 / ev1 bracket the last trace, frame or batched call
  std::vector<rdrt::Material> materials;
  rdrt::PtState pt;
  rdrt::BatchState batch;
  rdrt::Scratch compact_scratch;
  rdrt::TraceTune tune;  // launch shape of the fast traversal (config block size, rdrt_autotune)
  rdrt::Scratch trace_work;  // work counter of persistent rdrt_trace launches
  rdrt::Scratch trace_cull;  // cull pass scratch of rdrt_trace
  float tune_ms = 0.0f;      // per-launch time of the autotuned choice
  bool device_build = false;  // build the BLAS trees on the device (rdrt_set_device_build)
  // refit state (built on the first rdrt_refit after a commit)
  bool refit_ready = false;
  rdrt::DevArray<uint32_t> refit_forest, refit_pairs, refit_quads;
  std::vector<uint32_t> refit_forest_off, refit_pairs_off, refit_quads_off;
  rdrt::Scratch refit_boxes;
  rdrt::AoState *ao = nullptr;  // created by the first rdrt_ao_render
};

namespace rdrt {

int32_t need_device(rdrt_ctx *c);
void invalidate(rdrt_ctx *c);
// Device view of the committed scene for the trace kernels (no rays).
TraceParams scene_trace_params(const rdrt_ctx *c);
// Deepest push count of the traversal layout that t selects.
uint32_t push_depth(const rdrt_ctx *c, const TraceTune &t);
int32_t cull_scratch(rdrt_ctx *c, uint32_t n, CullScratch &cs);
// Looks up the device address of the fixed scratch of rdrt_compact_static
// for `device` (current) once per process; a failure is reported by that
// entry point.
void static_scratch_init(int device);
// Counters of every task group after a prepare and a compaction of each
// (read_back_execution_states, task-graph/src/runtime/mod.rs:305-370);
// flips the graph's parities.  Waits for the stream.
int32_t read_task_states(tg::Graph &gr, hipStream_t st, uint32_t *wake, uint32_t *sof, uint32_t *empty,
                         uint32_t *errors);

}  // namespace rdrt

#endif
