/*
 * rdrt.h -- C-ABI of the MI355X ray-query hot path (librdrt.so).
 *
 * The entry points replace the reference's acceleration-structure provider
 * and ray-tracing encoder for the naive SAH-BVH backend:
 *
 *   shader/ray-tracing/src/api/backend.rs:120-151
 *     GPUAccelerationStructureSystemProvider
 *       create_bottom_level_acceleration_structure -> rdrt_blas_begin /
 *                                rdrt_blas_add_triangles / rdrt_blas_end
 *       delete_bottom_level_acceleration_structure -> rdrt_blas_destroy
 *       create_top_level_acceleration_structure    -> rdrt_tlas_create
 *       delete_top_level_acceleration_structure    -> rdrt_tlas_destroy
 *       bind_tlas                                  -> rdrt_bind_tlas
 *       bind_tlas_max_len                          -> rdrt_bind_tlas_max_len
 *       create_comp_instance (lazy build)          -> rdrt_commit
 *   shader/ray-tracing/src/backend/wavefront_compute/geometry/naive/mod.rs:322-494
 *     NaiveSahBvhSource::build                     -> rdrt_build / rdrt_commit
 *   shader/ray-tracing/src/backend/wavefront_compute/geometry/naive/traverse_cpu.rs:52-245
 *   .../geometry/naive/traverse_gpu.rs:182-282
 *     traverse (closest hit)                       -> rdrt_trace, rdrt_trace_closest
 *
 * Conventions (mirroring the reference's provider contract, naive/mod.rs:98-119,497-609):
 *   - Every function returns an int32 status: RDRT_OK (0) or a negative
 *     RDRT_ERR_* code; rdrt_last_error() gives a thread-local message.
 *   - Host arrays passed in are copied; device pointers (documented as such)
 *     are borrowed for the duration of the call.
 *   - Handles are u32 indices that are never reused.
 *   - A context is not thread-safe; all device work is enqueued on the
 *     context's stream and is asynchronous (rdrt_sync waits).
 *   - Matrices are column-major Mat4 (math/algebra/src/mat/mat4.rs:9-14).
 */
#ifndef RDRT_H
#define RDRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RDRT_ABI_VERSION 1

/* status codes */
#define RDRT_OK 0
#define RDRT_ERR_INVALID_ARG (-1)
#define RDRT_ERR_HIP (-2)
#define RDRT_ERR_NOT_BUILT (-3)
#define RDRT_ERR_OUT_OF_MEMORY (-4)
#define RDRT_ERR_STATE (-5)
#define RDRT_ERR_NO_DEVICE (-6)
#define RDRT_ERR_CAPACITY (-7)
#define RDRT_ERR_UNSUPPORTED (-8)

/* ray flags (shader/ray-tracing/src/api/ty.rs:102-114) */
#define RDRT_RAY_FLAG_NONE 0x00u
#define RDRT_RAY_FLAG_FORCE_OPAQUE 0x01u
#define RDRT_RAY_FLAG_FORCE_NON_OPAQUE 0x02u
#define RDRT_RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH 0x04u
#define RDRT_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER 0x08u
#define RDRT_RAY_FLAG_CULL_BACK_FACING_TRIANGLES 0x10u
#define RDRT_RAY_FLAG_CULL_FRONT_FACING_TRIANGLES 0x20u
#define RDRT_RAY_FLAG_CULL_OPAQUE 0x40u
#define RDRT_RAY_FLAG_CULL_NON_OPAQUE 0x80u
#define RDRT_RAY_FLAG_SKIP_TRIANGLES 0x100u

/* hit kinds (ty.rs:139-140) */
#define RDRT_HIT_KIND_FRONT_FACING_TRIANGLE 0xFEu
#define RDRT_HIT_KIND_BACK_FACING_TRIANGLE 0xFFu

/* geometry-instance flags (ty.rs:142-152) and geometry flags (ty.rs:154-159) */
#define RDRT_INSTANCE_TRIANGLE_FACING_CULL_DISABLE 0x1u
#define RDRT_INSTANCE_TRIANGLE_FLIP_FACING 0x2u
#define RDRT_INSTANCE_FORCE_OPAQUE 0x4u
#define RDRT_INSTANCE_FORCE_NO_OPAQUE 0x8u
#define RDRT_GEOMETRY_FLAG_OPAQUE 0x1u
#define RDRT_GEOMETRY_FLAG_NO_DUPLICATE_ANYHIT_INVOCATION 0x2u

#define RDRT_MISS 0xffffffffu

/* Traversal kernels. */
#define RDRT_TRAVERSAL_FAST 0u      /* ordered short-stack traversal (default) */
#define RDRT_TRAVERSAL_REFERENCE 1u /* the reference's stackless hit/miss-link order */

/* Ray record (32 B): origin, t_min, direction, t_max.  The reference's
 * ShaderRayTraceCall ray + range (ty.rs:35-99). */
typedef struct rdrt_ray {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
} rdrt_ray;

/* Optional per-ray trace parameters (16 B): ray flags, cull mask and the
 * index into the bound TLAS list (traverse_cpu.rs:60-63). */
typedef struct rdrt_ray_meta {
  uint32_t flags, cull_mask, tlas_idx, pad;
} rdrt_ray_meta;

/* Closest-hit record (16 B).  prim_id = RDRT_MISS on a miss; t is the world
 * distance; (u, v) are the barycentrics weighting v1 and v2
 * (wavefront_compute/geometry/mod.rs:130-154). */
typedef struct rdrt_hit {
  uint32_t prim_id;
  float t, u, v;
} rdrt_hit;

/* Optional hit context (16 B): geometry index within the BLAS, instance
 * index within its TLAS source array, hit kind, instance custom index
 * (shader/ray-tracing/src/api/ctx.rs:24-86). */
typedef struct rdrt_hit_meta {
  uint32_t geometry_id, instance_id, hit_kind, instance_custom_index;
} rdrt_hit_meta;

/* TLAS instance (96 B): TopLevelAccelerationStructureSourceInstance
 * (shader/ray-tracing/src/api/backend.rs:160-177). */
typedef struct rdrt_instance {
  float transform[16];
  uint32_t instance_custom_index, mask, sbt_offset, flags, blas;
  uint32_t pad[3];
} rdrt_instance;

/* BVH build options (content/space/src/utils.rs:192-210 TreeBuildOption,
 * SAH::new(bins)).  The reference uses BLAS (bin 2, depth 50), TLAS (bin 10,
 * depth 50), 4 SAH bins (naive/mod.rs:173-177,280-284). */
typedef struct rdrt_build_opts {
  uint32_t blas_bin_size, blas_max_depth;
  uint32_t tlas_bin_size, tlas_max_depth;
  uint32_t sah_bins;
} rdrt_build_opts;

typedef struct rdrt_config {
  int32_t device;        /* HIP device ordinal; -1 = host-only context (build, no trace) */
  void *stream;          /* hipStream_t to enqueue on; NULL = the null stream */
  rdrt_build_opts build; /* zero fields take the reference defaults */
  uint32_t traversal;    /* RDRT_TRAVERSAL_* default for rdrt_trace_closest */
  uint32_t block_size;   /* 0 = tuned default */
} rdrt_config;

typedef struct rdrt_ctx rdrt_ctx;

/* One closest-hit launch.  All pointers are device pointers. */
typedef struct rdrt_trace_args {
  const rdrt_ray *rays;          /* n rays */
  const rdrt_ray_meta *ray_meta; /* n records or NULL (then flags/cull_mask/tlas_idx below) */
  rdrt_hit *hits;                /* n hits */
  rdrt_hit_meta *hit_meta;       /* n records or NULL */
  uint32_t *counters;            /* 3 u32 per ray or NULL: node visits, triangle tests, instance tests */
  uint32_t n;
  uint32_t flags, cull_mask, tlas_idx;
  uint32_t traversal; /* RDRT_TRAVERSAL_* */
  uint32_t pad;
} rdrt_trace_args;

/* Built-scene statistics. */
typedef struct rdrt_scene_info {
  uint32_t tri_nodes, tlas_nodes, tris, verts, instances, geometries, blas, tlas;
  uint32_t max_blas_depth, max_tlas_depth;
  uint32_t stack_entries;       /* traversal stack per lane of the child-pair layout (0 = too deep) */
  uint32_t quad_stack_entries;  /* the same for the child-quad layout */
  uint64_t device_bytes;
  double build_ms;
  uint32_t tri_quads, tlas_quads;  /* child-quad nodes (and their quantized twins) */
  double build_phase_ms[4];        /* cumulative at the end of: BLAS trees, forest + pairs + quads,
                                      triangle records, quantized quads */
} rdrt_scene_info;

/* ---- errors / version ---- */
const char *rdrt_last_error(void);
int32_t rdrt_abi_version(void);

/* ---- context ---- */
int32_t rdrt_config_default(rdrt_config *out);
int32_t rdrt_ctx_create(const rdrt_config *cfg, rdrt_ctx **out);
int32_t rdrt_ctx_destroy(rdrt_ctx *ctx);
int32_t rdrt_sync(rdrt_ctx *ctx);

/* ---- acceleration structures (backend.rs:120-151) ---- */
int32_t rdrt_blas_begin(rdrt_ctx *ctx);
/* A one-geometry BLAS in one call: rdrt_blas_begin, rdrt_blas_add_triangles,
 * rdrt_blas_end.  opts_or_null, if given, must equal the context's BLAS
 * build options (zero fields = the reference defaults); the trees of one
 * context share them (RDRT_ERR_UNSUPPORTED otherwise). */
int32_t rdrt_blas_create(rdrt_ctx *ctx, const float *pos_xyz, uint32_t nverts, const uint32_t *indices_or_null,
                         uint32_t ntris, uint32_t geometry_flags, const rdrt_build_opts *opts_or_null,
                         uint32_t *out_handle);
/* One geometry of the pending BLAS: positions are xyz f32 triples (copied);
 * indices_or_null = NULL means a non-indexed triangle list.  A rejected
 * geometry aborts the pending BLAS (rdrt_blas_end then fails). */
/* As rdrt_blas_add_triangles with 16-bit indices (the bindless mesh's u16
 * index format, bindless_mesh_bridge.rs:11-41). */
int32_t rdrt_blas_add_triangles_u16(rdrt_ctx *ctx, const float *pos_xyz, uint32_t nverts, const uint16_t *idx,
                                    uint32_t ntris, uint32_t flags);
int32_t rdrt_blas_add_triangles(rdrt_ctx *ctx, const float *pos_xyz, uint32_t nverts,
                                const uint32_t *indices_or_null, uint32_t ntris,
                                uint32_t geometry_flags);
int32_t rdrt_blas_end(rdrt_ctx *ctx, uint32_t *out_handle);
int32_t rdrt_blas_destroy(rdrt_ctx *ctx, uint32_t handle);
int32_t rdrt_tlas_create(rdrt_ctx *ctx, const rdrt_instance *inst, uint32_t n, uint32_t *out_handle);
int32_t rdrt_tlas_destroy(rdrt_ctx *ctx, uint32_t handle);
int32_t rdrt_bind_tlas(rdrt_ctx *ctx, const uint32_t *handles, uint32_t n);
uint32_t rdrt_bind_tlas_max_len(const rdrt_ctx *ctx);
/* Host-side build of every BVH (no device work), unless device builds are
 * on (rdrt_set_device_build). */
int32_t rdrt_build(rdrt_ctx *ctx);
/* enable = 1: the next rdrt_build / rdrt_commit builds every BLAS tree (the
 * reference's 4-bin SAH tree and the fast kernels' tree) on the device,
 * level by level, with the host builders' split decisions (same nodes; only
 * the order inside leaves under a reference median split may differ), then
 * the layouts on the host as before.  Not supported (RDRT_ERR_CAPACITY):
 * more than 4 reference SAH bins, a reference median split of more than 4096
 * primitives.  enable = 0 (default): host builds.  (No counterpart in the
 * reference, whose build is single-threaded on the host.) */
int32_t rdrt_set_device_build(rdrt_ctx *ctx, uint32_t enable);
/* Build if needed and upload to the device. */
int32_t rdrt_commit(rdrt_ctx *ctx);
/* Refit after a vertex update (no counterpart in the reference, which
 * rebuilds): positions_dev holds every vertex of the committed scene in the
 * scene vertex order (rdrt_geometry_vertex_range), 3 f32 each; normals_dev,
 * if not NULL, the vertex normals likewise.  Every BLAS keeps its trees and
 * gets its boxes and triangle records recomputed on the device (the
 * reference forest, the child pairs, the child quads and their quantized
 * twins); the TLAS layer (instances and TLAS trees) is rebuilt on the host
 * from the new BLAS boxes.  Synchronizes.  Captured frames and batched calls
 * are re-captured and a recorded frame can no longer be differentiated.  The
 * host-side copies (rdrt_copy_built) keep the committed geometry, and a later
 * rebuild starts from it. */
int32_t rdrt_refit(rdrt_ctx *ctx, const float *positions_dev, const float *normals_dev);
int32_t rdrt_get_scene_info(const rdrt_ctx *ctx, rdrt_scene_info *out);
/* Copies of the built reference-layout arrays (host memory, sized from
 * rdrt_scene_info), for inspection and parity checks.  which:
 * 0 = BLAS forest (48 B nodes), 1 = TLAS forest (48 B nodes),
 * 2 = triangle redirect (u32), 3 = instance source order (u32),
 * 4 = BLAS child quads (128 B nodes), 5 = their quantized twins (64 B). */
int32_t rdrt_copy_built(const rdrt_ctx *ctx, uint32_t which, void *host_out, size_t bytes);

/* Shading attributes of one closest hit (48 B): the bindless mesh fetch
 * of the reference's closest-hit stage (bindless_mesh_bridge.rs:43-123) and
 * its shader-binding-table hit group (shader/ray-tracing/src/api/ctx.rs:53-55). */
typedef struct rdrt_hit_attributes {
  float position[3];         /* world hit point: origin + t * direction */
  float shading_normal[3];   /* barycentric vertex normal through the normal matrix, normalized, on the
                                geometric normal's side (the geometric normal without vertex normals) */
  float geometric_normal[3]; /* (a - b) x (a - c) through the normal matrix, normalized, facing the ray origin */
  float uv[2];               /* barycentric texture coordinates (0, 0 without them) */
  uint32_t sbt_record;       /* sbt_offset + sbt_stride * geometry + instance sbt offset; RDRT_MISS on a miss */
} rdrt_hit_attributes;

/* The rest of the reference's closest-hit context (HitCtxInfo,
 * shader/ray-tracing/src/api/ctx.rs:24-48; HitCtxStorage,
 * wavefront_compute/ctx.rs:63-74), 160 B. */
typedef struct rdrt_hit_transform {
  float object_to_world[16];     /* gl_ObjectToWorldEXT: the instance transform as given (column-major) */
  float world_to_object[16];     /* gl_WorldToObjectEXT: its inverse as the traversal computed it */
  float object_ray_origin[3];    /* gl_ObjectRayOriginEXT: world_to_object * (origin, 1), w-divided */
  float object_ray_direction[3]; /* gl_ObjectRayDirectionEXT: 3x3 of world_to_object * direction, normalized */
  uint32_t instance_sbt_offset;  /* the instance's hit-group offset */
  uint32_t pad;
} rdrt_hit_transform;

/* ---- device task graph (shader/task-graph/src/runtime) ---- */
/* Groups of tasks that spawn, sleep on, wake and finish one another, polled
 * round by round on the context's stream.  A task is in one of the states of
 * task_pool.rs:85-107 (free, awake, going to sleep, asleep, finished).  The
 * generic graph offers two task kinds, enough for the reference's own task
 * graph tests (shader/task-graph/src/test.rs): */
#define RDRT_TASK_READY 0u      /* resolves on its first poll (BaseShaderFuture) */
#define RDRT_TASK_SPAWN_THEN 1u /* first poll spawns a task of `child_group` with payload + 1 and itself as parent,
                                   sleeps until it finishes, takes its payload and resolves */
typedef struct rdrt_task_group_desc {
  uint32_t kind;          /* RDRT_TASK_* */
  uint32_t child_group;   /* RDRT_TASK_SPAWN_THEN: the group it spawns into (may be itself) */
  uint32_t max_in_flight; /* slots = max_in_flight * capacity (mod.rs:121) */
  uint32_t pad;
} rdrt_task_group_desc;
typedef struct rdrt_task_graph rdrt_task_graph;
/* DeviceTaskGraphBuildSource::build (mod.rs:97-160); at most 8 groups. */
int32_t rdrt_task_graph_create(rdrt_ctx *ctx, const rdrt_task_group_desc *groups, uint32_t ngroups, uint32_t capacity,
                               rdrt_task_graph **out);
int32_t rdrt_task_graph_destroy(rdrt_task_graph *g); /* before its context is destroyed */
/* dispatch_allocate_init_task (mod.rs:223-303): `count` root tasks of
 * `group`, payload = spawn index, no parent. */
int32_t rdrt_task_graph_spawn(rdrt_task_graph *g, uint32_t group, uint32_t count);
/* execute (mod.rs:396-460): `rounds` rounds; each polls every group in
 * order (prepare, poll the active tasks, compact the active list). */
int32_t rdrt_task_graph_execute(rdrt_task_graph *g, uint32_t rounds);
/* read_back_execution_states (mod.rs:305-370), per group: awake tasks,
 * tasks asleep or finished (slots - empty - awake) and empty slots; errors
 * (optional): bit 0 pool exhausted, bit 1 active list full, bit 2 free list
 * full.  Waits for the stream. */
int32_t rdrt_task_graph_states(rdrt_task_graph *g, uint32_t *wake, uint32_t *sleep_or_finished, uint32_t *empty,
                               uint32_t *errors);

/* ---- ambient occlusion on the task graph (scene/rendering/gpu-ray-tracing/src/feature/ao.rs:176-315) ---- */
#define RDRT_AO_FORMAT_F32 0u   /* one float per pixel */
#define RDRT_AO_FORMAT_RGBA8 1u /* RGBA8 unorm per pixel, as the reference's AO texture */
typedef struct rdrt_ao_params {
  uint32_t width, height;
  uint32_t sample_index; /* samples accumulated so far (the reference's AO sample count) */
  uint32_t format;       /* RDRT_AO_FORMAT_* */
  float view_projection_inv[16];
  float camera_position[3];
  uint32_t rounds;      /* task-graph rounds per tile; 0 = 8 (ao.rs:300).  Rounds past the fifth are
                           skipped: every AO task has finished by then, whatever the scene */
  uint32_t tile_pixels; /* pixels per task-graph execution, whole rows; 0 = 512 * 512 (wavefront_compute/mod.rs:134) */
  uint32_t use_graph;   /* replay a captured graph of the call */
  uint32_t pad[2];
} rdrt_ao_params;
/* One AO sample for every pixel of the TLAS bound at index 0: camera ray
 * (primary hits cull back faces), then an occlusion ray from the hit,
 * cosine-distributed around the shading normal by the Hammersley point of
 * the sample index, first hit in [0.01, 100]; the pixel's running mean in
 * `accum` (device memory) becomes (mean * s + v) / (s + 1) while s < 256
 * (v = 1 when the occlusion ray or the camera ray misses, 0 otherwise).
 * Runs as the reference's ray-tracing pipeline does: trace, ray-gen,
 * closest-hit and miss task groups on the device task graph. */
int32_t rdrt_ao_render(rdrt_ctx *ctx, const rdrt_ao_params *params, void *accum);
/* States of the AO task graph after the last rdrt_ao_render, as
 * rdrt_task_graph_states (5 groups: trace, ray gen, closest hit, occlusion
 * closest hit, miss); errors bit 3 of the trace group: a task open after
 * round 3 that is not a pixel's occlusion trace awaiting only the
 * resolution of its chain (rounds 4 and 5 run as one pass over those
 * chains; not reached). */
int32_t rdrt_ao_states(rdrt_ctx *ctx, uint32_t *wake, uint32_t *sleep_or_finished, uint32_t *empty, uint32_t *errors);
/* Times every traversal launch shape on eager AO calls with these
 * parameters (`accum` is overwritten) and keeps the fastest for the
 * context, as rdrt_pt_autotune does for path-tracing frames. */
int32_t rdrt_ao_autotune(rdrt_ctx *ctx, const rdrt_ao_params *params, void *accum, uint32_t reps,
                         struct rdrt_tune *best);
/* Camera rays and occlusion rays traced by the last rdrt_ao_render (waits
 * for the stream). */
int32_t rdrt_ao_ray_counts(rdrt_ctx *ctx, uint32_t out[2]);
/* The unit-hemisphere direction (tangent space, z = normal) of sample
 * index s: hammersley_2d(s, 256) through sample_hemisphere_cos
 * (shader/library/src/sampling.rs:33-68), sine and cosine rounded from
 * double precision. */
int32_t rdrt_ao_hemisphere_sample(uint32_t s, float out[3]);

/* ---- programmable ray-tracing pipelines on the device task graph ----
 * The reference's stage pipelines (shader/ray-tracing/src/api/pipeline.rs:
 * 75-130: a ray-gen stage, closest-hit stages per hit group, miss stages;
 * operator/mod.rs:34-69: then_trace chains) run by its wavefront executor
 * (wavefront_compute/{mod.rs:113-190, pipeline.rs:95-207, trace_task.rs}).
 * A stage is a HIP kernel of the caller, `__global__ void f(rdrt_stage_args
 * a)` (compiled for gfx950 into any code object of the process; pass its
 * host-side handle, `(const void *)f`).  Each round the library launches it
 * over the stage's polled tasks -- slots[i] for i < *count, a grid of
 * 256-thread blocks covering `capacity` threads -- and every polled task
 * either resolves (its payload goes back to the task that traced it) or
 * traces one more ray
 * (then_trace): it writes the trace call and the payload the traced ray's
 * stages receive, and is polled again once that ray's closest-hit or miss
 * stage has resolved, with polls[slot] one higher and the returned payload
 * in trace_result.  The trace task between them is the library's: it
 * traverses the ray, runs the closest-hit stage of the hit group
 * sbt_offset + sbt_stride * geometry + instance sbt offset (ctx.rs:53-55)
 * or the miss stage miss_index, and hands the payload back. */
#define RDRT_STAGE_RESOLVE 0u
#define RDRT_STAGE_TRACE 1u
#define RDRT_PIPELINE_MAX_PAYLOAD 16u
#define RDRT_PIPELINE_HIT_ATTRIBUTES 1u /* fill rdrt_stage_args.attributes for every traced ray */
#define RDRT_PIPELINE_HIT_TRANSFORMS 2u /* fill rdrt_stage_args.transforms likewise */
typedef struct rdrt_trace_call { /* ShaderRayTraceCall (api/ty.rs:35-99) */
  float origin[3], tmin;
  float direction[3], tmax;
  uint32_t flags, cull_mask, tlas_idx, miss_index;
  uint32_t sbt_offset, sbt_stride, pad[2];
} rdrt_trace_call; /* 64 B */
typedef struct rdrt_stage_args {
  const uint32_t *count;         /* device word: tasks polled this round */
  const uint32_t *slots;         /* their slots: slots[i] for i < *count */
  uint32_t *payload;             /* payload_words per slot: ray gen: the task's own words (zero at launch);
                                    closest hit / miss: in = the payload of the traced ray, out = returned */
  const uint32_t *trace_result;  /* payload_words per slot: the payload the task's last trace returned */
  const uint32_t *polls;         /* per slot: traces of the task that have returned (0 at its first poll) */
  const uint32_t *launch;        /* per slot: launch index x + y * launch_size[0] of the chain's ray gen */
  const uint32_t *trace_of;      /* per slot (closest hit / miss): trace slot of the ray and hit below */
  const rdrt_ray *rays;          /* per trace slot: the traced ray */
  const rdrt_ray_meta *ray_meta; /* per trace slot: its flags, cull mask and TLAS index */
  const rdrt_hit *hits;          /* per trace slot: the closest hit (prim_id ~0 on a miss) */
  const rdrt_hit_meta *hit_meta; /* per trace slot */
  const rdrt_hit_attributes *attributes; /* per trace slot with RDRT_PIPELINE_HIT_ATTRIBUTES, else NULL */
  const rdrt_hit_transform *transforms;  /* per trace slot with RDRT_PIPELINE_HIT_TRANSFORMS, else NULL */
  uint32_t *action;              /* out per slot: RDRT_STAGE_RESOLVE (preset) or RDRT_STAGE_TRACE */
  rdrt_trace_call *call;         /* out per slot, with RDRT_STAGE_TRACE */
  uint32_t *call_payload;        /* out per slot, with RDRT_STAGE_TRACE: payload_words for the ray's stages */
  void *user;                    /* the pointer given to rdrt_pipeline_trace_rays */
  uint32_t launch_size[2];
  uint32_t payload_words;
  uint32_t capacity;             /* slots of the stage (the grid covers this many threads) */
} rdrt_stage_args;
typedef struct rdrt_pipeline_desc {
  const void *ray_gen;        /* stage kernel handles */
  const void *closest_hit[4]; /* per hit group; NULL: a hit resolves the trace with the payload as it was */
  const void *miss[4];        /* per miss index; NULL: likewise */
  uint32_t nclosest, nmiss;   /* nclosest + nmiss <= 6 */
  uint32_t payload_words;     /* 1 .. RDRT_PIPELINE_MAX_PAYLOAD */
  uint32_t max_in_flight;     /* trace tasks per launch entry in flight (pipeline.rs:211-257); 0 = 2 */
  uint32_t round_hint;        /* task-graph rounds per tile (execution_round_hint); 0 = 8 */
  uint32_t tile_pixels;       /* launch entries per tile, whole rows (the reference: 512 * 512,
                                 wavefront_compute/mod.rs:134); 0 = 2^22 */
  uint32_t hit_context;       /* RDRT_PIPELINE_HIT_*: per-hit data the closest-hit stages read (ctx.rs:24-48) */
  uint32_t use_graph;         /* 1: capture a call's launches into a hipGraph and replay it while the launch
                                 size, user pointer, scene and launch shape stay the same */
} rdrt_pipeline_desc;
typedef struct rdrt_pipeline rdrt_pipeline;
/* GPURaytracingPipelineAndBindingSource + executor: groups trace, ray gen,
 * the closest-hit stages, the miss stages. */
int32_t rdrt_pipeline_create(rdrt_ctx *ctx, const rdrt_pipeline_desc *desc, rdrt_pipeline **out);
int32_t rdrt_pipeline_destroy(rdrt_pipeline *p); /* before its context is destroyed */
/* RayTracingEncoderProvider::trace_ray (backend.rs:48-56): one ray-gen task
 * per launch entry (width x height, tiles of whole rows), round_hint rounds
 * per tile, on the context's stream and its TLAS bindings. */
int32_t rdrt_pipeline_trace_rays(rdrt_pipeline *p, uint32_t width, uint32_t height, void *user);
/* Task states per group after the last call, as rdrt_task_graph_states
 * (groups: trace, ray gen, closest hit..., miss...). */
int32_t rdrt_pipeline_states(rdrt_pipeline *p, uint32_t *wake, uint32_t *sleep_or_finished, uint32_t *empty,
                             uint32_t *errors);
/* Rays the trace group traversed in the last call (waits for the stream). */
int32_t rdrt_pipeline_ray_count(rdrt_pipeline *p, uint64_t *out);

/* ---- buffers shared with another graphics API ----
 * The reference's rays, hits and render targets live in wgpu buffers
 * (shader/ray-tracing/src/api/backend.rs:48-56).  A Vulkan device exports
 * a buffer's memory as an opaque POSIX file descriptor
 * (VK_EXTERNAL_MEMORY_HANDLE_TYPE_OPAQUE_FD_BIT, vkGetMemoryFdKHR); this
 * maps bytes [offset, offset + length) of that allocation (`size` bytes in
 * all) into the context's device and returns their device address, which
 * every rdrt_* call that takes device memory accepts.  On success the
 * library owns `fd` (closed by the release); on failure the caller still
 * does.  Work on the buffer is ordered by the caller (the context's stream
 * against the other API's queue, e.g. with an exported semaphore). */
typedef struct rdrt_external_buffer rdrt_external_buffer;
int32_t rdrt_external_buffer_import(rdrt_ctx *ctx, int fd, uint64_t size, uint64_t offset, uint64_t length,
                                    rdrt_external_buffer **out, void **dev_ptr);
/* Waits for the context's stream, unmaps the buffer and releases the
 * imported memory.  NULL is a no-op. */
int32_t rdrt_external_buffer_release(rdrt_external_buffer *buffer);

/* ---- host picking (content/mesh/core/src/feature/bvh.rs:57-86) ---- */
#define RDRT_FACE_FRONT 0u
#define RDRT_FACE_BACK 1u
#define RDRT_FACE_DOUBLE 2u
typedef struct rdrt_pick_hit {
  uint32_t primitive; /* triangle of the geometry, RDRT_MISS when nothing is hit */
  float distance;     /* along the direction (normalize it for a metric distance) */
  float position[3];  /* origin + direction * distance */
  uint32_t pad;
} rdrt_pick_hit;
/* Nearest hits of n rays (host array: origin xyz, direction xyz per ray)
 * against one geometry of a BLAS, on the host through the geometry's
 * reference tree, with the reference's host tests: the GTE ray/triangle test
 * with a face side (math/geometry/src/dimension3/intersection.rs:3-76), the
 * ray/box test of the same file (:128-196), no range limit beyond t >= 0,
 * the first of equally near hits kept.  Needs a built scene (rdrt_build; a
 * host-only context is enough); uses the committed positions. */
int32_t rdrt_pick(rdrt_ctx *ctx, uint32_t blas, uint32_t geometry, const float *rays, uint32_t n, uint32_t face_side,
                  rdrt_pick_hit *out);
/* Every hit of one ray (6 floats) against one geometry, in the order the
 * reference's walk finds them (intersect_list_bvh,
 * content/mesh/core/src/feature/bvh.rs:23-55).  *count receives the number
 * of hits; the first min(count, capacity) are written to out. */
int32_t rdrt_pick_all(rdrt_ctx *ctx, uint32_t blas, uint32_t geometry, const float *ray, uint32_t face_side,
                      rdrt_pick_hit *out, uint32_t capacity, uint32_t *count);

/* ---- tracing ---- */
int32_t rdrt_trace(rdrt_ctx *ctx, const rdrt_trace_args *args);
/* Attributes of n closest hits traced before (device arrays: the rays, their
 * metadata or NULL for tlas_idx, the hits and hit metadata rdrt_trace
 * wrote); misses give zeros and sbt_record RDRT_MISS.  Enqueued on the
 * context stream. */
int32_t rdrt_fetch_hit_attributes(rdrt_ctx *ctx, const rdrt_ray *rays, const rdrt_ray_meta *ray_meta,
                                  const rdrt_hit *hits, const rdrt_hit_meta *hit_meta, uint32_t n, uint32_t tlas_idx,
                                  uint32_t sbt_offset, uint32_t sbt_stride, rdrt_hit_attributes *out);
/* The hit context of closest hits (transforms and object-space ray, as
 * the reference's closest-hit stage sees them; traverse_gpu.rs:575-582,
 * 661-683), for the same inputs as rdrt_fetch_hit_attributes; misses give
 * zeros.  Enqueued on the context stream. */
int32_t rdrt_fetch_hit_transforms(rdrt_ctx *ctx, const rdrt_ray *rays, const rdrt_ray_meta *ray_meta,
                                  const rdrt_hit *hits, const rdrt_hit_meta *hit_meta, uint32_t n, uint32_t tlas_idx,
                                  rdrt_hit_transform *out);
/* Closest hit with flags 0, cull mask ~0, TLAS binding 0, the context's
 * default traversal. */
int32_t rdrt_trace_closest(rdrt_ctx *ctx, const rdrt_ray *rays_dev, uint32_t n, rdrt_hit *hits_dev);
/* Average duration (ms) of the traversal kernel launches of the last
 * rdrt_trace call, measured with HIP events on the context stream; requires
 * rdrt_sync first. */
int32_t rdrt_last_trace_ms(rdrt_ctx *ctx, float *ms);
/* Context metrics in one record. */
typedef struct rdrt_metrics {
  float last_call_ms;    /* device time of the last trace, frame or batched call (waits for it; -1 if none) */
  float tune_ms;         /* the chosen shape's time in the last autotune (0 if none) */
  uint64_t device_bytes; /* device memory of the committed scene */
  double build_ms;       /* host time of the last build */
  uint32_t status;       /* device status word: bit 0 = a traversal stack would have overflowed */
  uint32_t pad;
} rdrt_metrics;
int32_t rdrt_get_metrics(rdrt_ctx *ctx, rdrt_metrics *out);

/* ---- launch tuning of the fast traversal (no counterpart in the reference:
 *      its workgroup size is fixed at 128, traverse_gpu.rs) ---- */
typedef struct rdrt_tune {
  uint32_t block_size;  /* threads per block: 64, 128, 256 */
  uint32_t lds_entries; /* traversal stack entries per lane in LDS (8, 16); deeper ones in private memory */
  uint32_t waves;       /* 8 / 6 = compiled for 8 / 6 waves per SIMD (at most 64 / 80 VGPRs); 0 = the
                           compiler's choice */
  uint32_t persistent;  /* 1 = resident-size grid whose waves pull 64-ray chunks from a counter */
  uint32_t cull;        /* 1 = a first pass finishes rays that miss their scene's root box and queues the rest */
  uint32_t wide;        /* node layout: 0 = child pairs (64 B), 1 = child quads (four boxes, 128 B),
                           2 = quantized child quads (four 8-bit boxes, 64 B) */
  float ms;             /* rdrt_autotune: per-launch time of the chosen shape on the sample */
  uint32_t coop;        /* launches with at most this many live rays (device-side count) are traced by the
                           cooperative kernel, 8 lanes per ray (16 for launches of at most 32,768 live rays):
                           the TLAS walk in lockstep, the instances it accepts dealt to the lanes; 0 = never.
                           Fast layouts only, not with visit counting */
  uint32_t tail;        /* path-tracer frames: from bounce `tail` (1 .. max_bounces - 1) on, one launch finishes
                           every path, 8 lanes per path running its remaining bounces back to back (the frame
                           waits for its slowest path instead of the slowest ray of every bounce); 0 = one set
                           of launches per bounce.  Not with visit counting */
} rdrt_tune;
/* Times every shape of the search space (3 block sizes x 2 LDS stack depths
 * x 2 occupancy targets x grid / persistent x with / without the cull pass
 * x the three node layouts) on a sample launch (the args of
 * rdrt_trace; its hits are overwritten), `reps` timed launches each, and
 * keeps the fastest for every later traversal of this context (rdrt_trace,
 * the path tracer, batched tracing).  Synchronizes.  best may be NULL. */
int32_t rdrt_autotune(rdrt_ctx *ctx, const rdrt_trace_args *sample, uint32_t reps, rdrt_tune *best);
int32_t rdrt_set_tune(rdrt_ctx *ctx, const rdrt_tune *tune);
int32_t rdrt_get_tune(rdrt_ctx *ctx, rdrt_tune *tune);

/* ---- shading attributes (bindless mesh data of
 *      scene/rendering/gpu-ray-tracing/src/bindless_mesh_bridge.rs:11-135) ---- */
/* Lambertian material (content/lighting/transport/src/surface/microfacet/
 * device/diffuse.rs): albedo and emitted radiance. */
typedef struct rdrt_material {
  float albedo[3];
  float pad0;
  float emission[3];
  float pad1;
} rdrt_material;

/* Per-vertex normals of one geometry of a BLAS (xyz per vertex, copied);
 * interpolated with the hit barycentrics for shading.  Without them the
 * geometric normal is used. */
/* Where a geometry's vertices sit in the scene vertex order (all live BLAS
 * in creation order, their geometries in order): the layout of the geometry
 * gradients of rdrt_pt_backward_geometry. */
int32_t rdrt_geometry_vertex_range(rdrt_ctx *ctx, uint32_t blas, uint32_t geometry, uint32_t *first,
                                   uint32_t *count);
int32_t rdrt_blas_set_vertex_normals(rdrt_ctx *ctx, uint32_t blas, uint32_t geometry, const float *normals_xyz,
                                     uint32_t nverts);
/* Per-triangle material offsets of one geometry (copied).  A hit's material
 * is materials[instance_custom_index + offset], as the reference looks the
 * surface up by the instance's custom id (feature/path_tracing/ray_hit.rs:65-75). */
/* Any-hit behaviour of each triangle of a geometry, for hits while the
 * geometry is not opaque for the ray (no GEOMETRY_FLAG_OPAQUE, or forced
 * non-opaque by the ray or instance): RDRT_ANYHIT_ACCEPT commits the hit,
 * RDRT_ANYHIT_END_SEARCH ends the traversal (with or without it), 0 ignores
 * it -- the result of the reference's any-hit stage (ANYHIT_BEHAVIOR_*,
 * traverse_cpu.rs:158-185) expressed per primitive, as an alpha-tested
 * cutout needs.  Without a table every such hit is accepted.  Which hit an
 * end-search behaviour stops at depends on the visit order (as in the
 * reference); accept / ignore do not.  Takes effect at the next commit. */
#define RDRT_ANYHIT_ACCEPT 1u
#define RDRT_ANYHIT_END_SEARCH 2u
int32_t rdrt_blas_set_anyhit(rdrt_ctx *ctx, uint32_t blas, uint32_t geometry, const uint8_t *behavior,
                             uint32_t ntris);
/* Per-vertex texture coordinates of a geometry (uv pairs), the uv channel
 * of the bindless mesh (scene/rendering/gpu-ray-tracing/src/
 * bindless_mesh_bridge.rs:65-77). */
int32_t rdrt_blas_set_vertex_uvs(rdrt_ctx *ctx, uint32_t blas, uint32_t geometry, const float *uvs, uint32_t nverts);
int32_t rdrt_blas_set_triangle_materials(rdrt_ctx *ctx, uint32_t blas, uint32_t geometry,
                                         const uint32_t *offsets, uint32_t ntris);
int32_t rdrt_set_materials(rdrt_ctx *ctx, const rdrt_material *materials, uint32_t n);

/* ---- wavefront path tracing (scene/rendering/gpu-ray-tracing/src/feature/
 *      path_tracing/mod.rs:127-235, ray_gen.rs, ray_hit.rs, ray_miss.rs) ---- */
typedef struct rdrt_pt_params {
  float view_projection_inv[16]; /* column-major NDC -> world */
  float camera_position[3];
  uint32_t width, height, spp;
  uint32_t max_bounces; /* scattering events per path (<= 16) */
  uint32_t rr_start;    /* first bounce with Russian roulette (a frame with rr_start < max_bounces is not differentiable) */
  uint32_t seed;        /* per-pixel sample scrambles: xxhash32(px, py, seed * 2^16 + dimension pair) */
  uint32_t sample_index_base;
  float sky_up[3], sky_down[3];            /* miss radiance for direction.y > 0 / <= 0 */
  float sun_direction[3], sun_radiance[3]; /* directional light (direction of travel) */
  uint32_t use_graph;                      /* 1 = replay a captured hipGraph of the frame */
  uint32_t diagnostics;                    /* RDRT_PT_DIAG_* bits; any bit forces an eager frame */
  uint32_t record;                         /* 1 = keep the path vertices for rdrt_pt_backward */
  uint32_t pad;
  /* Several cameras in one frame: views > 1 renders views * height * width
   * pixels (view-major); view_params (host) holds 20 f32 per view:
   * view_projection_inv[16], camera_position[3], pad.  View v scrambles its
   * samples with seed + v.  views <= 1 with view_params NULL: the camera above. */
  uint32_t views;
  uint32_t pad2;
  const float *view_params;
} rdrt_pt_params;

/* time every traversal launch of the frame with HIP events (a frame traces
 * the shadow rays of bounce b and the extension rays of bounce b + 1 in one
 * launch: max_bounces + 1 launches) */
#define RDRT_PT_DIAG_TIME_TRACES 0x1u
/* trace in the reference order, one launch per ray set (2 * max_bounces + 1),
 * and count node / triangle / instance visits */
#define RDRT_PT_DIAG_COUNT_VISITS 0x2u
/* as RDRT_PT_DIAG_COUNT_VISITS with the fast traversal: its own tree and
 * order (node-pair visits, triangle tests, instance tests) */
#define RDRT_PT_DIAG_COUNT_FAST 0x4u

typedef struct rdrt_pt_stats {
  uint32_t bounces;             /* entries below that are valid: max_bounces + 1 */
  uint32_t extension_rays[17];  /* closest-hit rays traced at each bounce */
  uint32_t shadow_rays[17];     /* shadow rays traced after each bounce */
  float frame_ms;               /* HIP-event time of the last frame */
  float trace_ms;               /* sum over its traversal launches (RDRT_PT_DIAG_TIME_TRACES) */
  uint32_t trace_launches;
  uint64_t visits[3];           /* node visits, triangle tests, instance tests (RDRT_PT_DIAG_COUNT_*) */
  uint32_t max_ray_visits[33];  /* per launch in frame order: the largest visit sum of one ray (RDRT_PT_DIAG_COUNT_*) */
  float launch_ms[33];          /* per launch in frame order: its HIP-event time (RDRT_PT_DIAG_TIME_TRACES) */
} rdrt_pt_stats;

/* Renders one frame: radiance_dev receives views*width*height RGB f32 (the
 * mean of spp paths per pixel).  Scene committed and materials set first. */
int32_t rdrt_pt_render(rdrt_ctx *ctx, const rdrt_pt_params *params, float *radiance_dev);
/* The same call under the name of the forward half of the differentiable
 * frame (rdrt_pt_backward is the other). */
int32_t rdrt_pt_forward(rdrt_ctx *ctx, const rdrt_pt_params *params, float *radiance_dev);
/* Ray counts of the last frame (synchronizes). */
int32_t rdrt_pt_get_stats(rdrt_ctx *ctx, rdrt_pt_stats *out);
/* Times every launch shape of the fast traversal (3 block sizes x 2 LDS
 * stack depths x 3 occupancy targets x grid / persistent x the three node
 * layouts x 3 cooperative thresholds) on eager frames of these parameters
 * (record and diagnostics off), `reps` timed frames each, then the tail
 * kernel from bounce 1, 2 or 3 with the fastest shape, and keeps the
 * fastest for the context, as rdrt_autotune does for single launches;
 * best->ms is the frame time.  radiance_dev is overwritten.  Synchronizes.
 * best may be NULL.  (No counterpart in the reference.) */
int32_t rdrt_pt_autotune(rdrt_ctx *ctx, const rdrt_pt_params *params, float *radiance_dev, uint32_t reps,
                         rdrt_tune *best);

/* ---- differentiable path tracing (no counterpart in the reference: the
 *      adjoint of the path tracer's frame radiance) ---- */
/* Replaces the materials with device arrays (albedo and emission, n * 3 f32
 * each), packed on the context's stream: no host round trip, so a caller
 * can update them from an optimizer each step. */
int32_t rdrt_pt_set_material_params(rdrt_ctx *ctx, const float *albedo_dev, const float *emission_dev, uint32_t n);
/* Gradient of sum(adj * frame) over the last frame rendered with record = 1,
 * w.r.t. the material albedo and emission (n * 3 f32 each, overwritten),
 * with visibility and sampled directions held fixed.  adj_dev holds
 * views * height * width RGB f32 (dLoss/dframe) in the frame's image order:
 * view after view, each row-major, pixel (x, y) of view k at
 * ((k * height + y) * width + x) * 3 -- not the order paths or tiles are
 * traced in.  Frames with Russian roulette (rr_start < max_bounces, such as
 * the default rr_start = 2 of the 4-bounce benchmark frame) are not
 * differentiable and return RDRT_ERR_UNSUPPORTED: the termination
 * probability depends on the albedo.
 * The result is bit-reproducible: every term is rounded once to a 64-bit
 * fixed-point integer with a per-call scale (from the largest |term| and the
 * term count) and the integers are summed, so the order the device adds
 * them in does not change a bit; each term is exact to about 1e-10 of the
 * largest term.  A NaN term (or both infinities) makes its entry NaN, one
 * infinity that infinity.  The gradient arrays are written with plain
 * stores (no atomics on caller memory). */
int32_t rdrt_pt_backward(rdrt_ctx *ctx, const float *adj_dev, float *grad_albedo_dev, float *grad_emission_dev);
/* Gradient of sum(adj * frame) over the last frame rendered with record = 1
 * w.r.t. vertex positions and normals (scene vertex order, see
 * rdrt_geometry_vertex_range; vertex_count * 3 f32 each, overwritten), with
 * every path held fixed: traced rays, hit triangles, light visibility, facing
 * and sampled directions.  The shading normal -- the barycentric
 * interpolation of the vertex normals, or the face normal without them --
 * enters the light term and the continuation factor a (ns . wo) / cos_wo;
 * positions act through the barycentrics of the fixed ray's hit (or the face
 * normal).  The paths are recomputed by replaying the frame, not stored.
 * grad_nrm_dev may be NULL when no geometry has vertex normals.  Same frame
 * conditions, adjoint layout and bit-reproducible sums as rdrt_pt_backward
 * (the replay keeps 48 bytes of gradient terms per path and bounce, and 12
 * bytes of sums per vertex component, in context scratch). */
int32_t rdrt_pt_backward_geometry(rdrt_ctx *ctx, const float *adj_dev, float *grad_positions_dev,
                                  float *grad_normals_dev);
/* L2 loss helper: adj_dev[i] = 2 (a[i] - b[i]); *loss_dev (one f64) =
 * sum (a[i] - b[i])^2, added in a fixed order (per-block partials, then the
 * partials in block order): the same n and inputs give the same bits.
 * Device pointers; enqueued on the context's stream. */
int32_t rdrt_l2_adjoint(rdrt_ctx *ctx, const float *a_dev, const float *b_dev, uint32_t n, float *adj_dev,
                        double *loss_dev);

/* ---- batched tracing: many scenes and uneven ray sets in one queue ---- */
/* Each ray's rdrt_ray_meta.tlas_idx selects its scene among the bound TLAS
 * (the reference's per-ray TLAS selection, traverse_cpu.rs:63).  Round 0
 * traces the given rays; each later round traces, for every ray that hit, a
 * new ray from the offset hit point (ray_util.rs:6-40) in a cosine-weighted
 * direction around the geometric normal facing the ray, sampled from the
 * ray's Sobol (0,2)-sequence at the round index, XOR-scrambled by
 * xxhash32(ray index, seed, 2 / 3).  Survivors are compacted (stable).
 * A ray that misses in round r gets miss records in every later round. */
#define RDRT_BATCH_PERSISTENT 0u /* one persistent launch per round over all scenes, graph-captured */
#define RDRT_BATCH_SEQUENTIAL 1u /* per scene, per round: plain launches (baseline) */

typedef struct rdrt_batch_params {
  const rdrt_ray *rays;          /* n primary rays (device), grouped by scene */
  const rdrt_ray_meta *ray_meta; /* n records (device): flags, cull mask, tlas_idx = scene */
  rdrt_hit *hits;                /* (bounces + 1) * n (device): round r, ray i at r * n + i */
  rdrt_ray *spawned;             /* bounces * n (device) or NULL: rays of round r >= 1 at (r - 1) * n + i */
  const uint32_t *scene_offsets; /* host, nscenes + 1 (sequential mode): scene k = rays [off[k], off[k+1]) */
  uint32_t n, nscenes, bounces, seed, mode, pad;
} rdrt_batch_params;

typedef struct rdrt_batch_stats {
  uint32_t rounds;
  uint32_t rays[17]; /* rays traced in each round */
  uint32_t launches; /* kernel launches of the last call */
  float ms;          /* HIP-event time of the last call */
} rdrt_batch_stats;

int32_t rdrt_trace_batched(rdrt_ctx *ctx, const rdrt_batch_params *params);
int32_t rdrt_batch_get_stats(rdrt_ctx *ctx, rdrt_batch_stats *out);
/* Times every launch shape (3 block sizes x 2 LDS stack depths x 2 occupancy
 * targets x with / without the cull pass x the three node layouts)
 * on whole persistent batched calls of these parameters, `reps` timed calls
 * each, and keeps the fastest for the context; best->ms is the call time.
 * The outputs of params are overwritten.  Synchronizes.  best may be NULL.
 * (No counterpart in the reference.) */
int32_t rdrt_batch_autotune(rdrt_ctx *ctx, const rdrt_batch_params *params, uint32_t reps, rdrt_tune *best);

/* ---- device primitives of the path (parity helpers) ---- */
/* Stable stream compaction (shader/parallel-compute/src/stream_compaction.rs:
 * 3-45): out_idx_dev receives the indices i with mask_dev[i] != 0 in
 * increasing order, count_dev (device u32) their number. */
int32_t rdrt_compact(rdrt_ctx *ctx, const uint32_t *mask_dev, uint32_t n, uint32_t *out_idx_dev,
                     uint32_t *count_dev);
/* rdrt_compact with no device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops).  Same outputs as rdrt_compact, for n up to
 * RDRT_COMPACT_STATIC_MAX (larger n returns RDRT_ERR_CAPACITY).  Its scratch
 * is a fixed device array of the library (about 8.5 MB, held by each device
 * the library runs on until the process ends); every call on one device
 * shares it, so calls that overlap on one device (other streams, other
 * contexts, other threads) race and must be ordered by the caller.  Enqueued
 * on the context's stream. */
#define RDRT_COMPACT_STATIC_MAX 2097152u
int32_t rdrt_compact_static(rdrt_ctx *ctx, const uint32_t *mask_dev, uint32_t n, uint32_t *out_idx_dev,
                            uint32_t *count_dev);
/* One scrambled 2-D Sobol point on the host, bit-identical to the device
 * samplers: out2[0] = van_der_corput(n, scramble0), out2[1] = sobol(n,
 * scramble1) (math/statistics/src/sampling/sobol.rs:40-68). */
int32_t rdrt_sobol(uint32_t n, uint32_t scramble0, uint32_t scramble1, float out2[2]);
/* Scrambled Sobol samples (math/statistics/src/sampling/sobol.rs:40-68):
 * out_dev[2i] = van_der_corput(index[i], scramble[2i]),
 * out_dev[2i+1] = sobol(index[i], scramble[2i+1]).  Device pointers. */
int32_t rdrt_sample_sobol(rdrt_ctx *ctx, const uint32_t *index_dev, const uint32_t *scramble_dev, uint32_t n,
                          float *out_dev);

#ifdef __cplusplus
}
#endif
#endif
