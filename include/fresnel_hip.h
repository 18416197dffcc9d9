/* fresnel_hip.h -- C ABI of the MI355X render backend (libfresnel_hip.so).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never
 * throws, and returns FH_OK (0) or a negative status; fh_last_error() then
 * holds a message for the calling thread.  Handles are opaque.
 *
 * The entry points mirror the class set of the reference's `_gpu` binding
 * module (fresnel/gpu/module-gpu.cc:22-43), which a pybind11 module
 * (fresnel_amd/csrc/bindings/gpu_module.cpp) binds one-to-one.  Each
 * declaration cites the reference interface it replaces.
 *
 * All calls are synchronous for the caller: a render returns once the output
 * is complete on the device, as the reference's render() does
 * (fresnel/cpu/TracerPath.cc:49-54, fresnel/gpu/TracerPath.cc:101).
 */
#ifndef FRESNEL_HIP_H
#define FRESNEL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define FH_OK 0
#define FH_ERR_INVALID (-1)   /* bad argument (maps to RuntimeError) */
#define FH_ERR_HIP (-2)       /* HIP runtime error */
#define FH_ERR_STATE (-3)     /* object in the wrong state (e.g. removed geometry) */
#define FH_ERR_NOMEM (-4)     /* device or host allocation failed / over budget */

/* Message for the last failing call on this thread ("" if none). */
const char* fh_last_error(void);

/* --------------------------------------------------------- plain data types
 * Field order and meaning follow the reference's POD types exported by
 * `_common` (fresnel/common/module-common.cc:37-115). */
typedef struct { float r, g, b; } fh_rgb;                 /* RGB<float>, ColorMath.h:37-63 */
typedef struct { float x, y, z; } fh_vec3;                /* vec3<float>, VectorMath.h:330 */
typedef struct {                                           /* Material, Material.h:47-56 */
    float solid;
    fh_rgb color;
    float primitive_color_mix;
    float roughness;
    float specular;
    float metal;
    float spec_trans;
} fh_material;
typedef struct {                                           /* UserCamera, Camera.h:55-77 */
    fh_vec3 position;
    fh_vec3 look_at;
    fh_vec3 up;
    float h;
    float f;
    float f_stop;
    float focus_distance;
    int32_t model;                                         /* 0 orthographic, 1 perspective */
} fh_user_camera;
typedef struct {                                           /* Lights, Light.h:36-45 */
    fh_vec3 direction[4];
    fh_rgb color[4];
    float theta[4];
    uint32_t N;
} fh_lights;

typedef struct fh_device_s* fh_device;
typedef struct fh_scene_s* fh_scene;
typedef struct fh_geometry_s* fh_geometry;
typedef struct fh_tracer_s* fh_tracer;

/* ------------------------------------------------------------------ device */
/* get_num_available_devices(), fresnel/gpu/module-gpu.cc:22-25 */
int fh_get_num_available_devices(int* n);
/* Device(ptx_root, n), fresnel/gpu/Device.cc:29-54.  limit = -1: all GPUs.
 * Only the first GPU is used: a Device drives one GPU. */
int fh_device_create(int limit, fh_device* out);
int fh_device_destroy(fh_device dev);
/* Device.describe(), fresnel/gpu/Device.cc:114-118 */
int fh_device_describe(fh_device dev, char* buf, size_t cap);
/* Device.getAllGPUs(), fresnel/gpu/Device.cc:120-128 (newline-terminated lines) */
int fh_device_get_all_gpus(char* buf, size_t cap);

/* ------------------------------------------------------------------- scene */
/* Scene(device), fresnel/gpu/Scene.cc (cpu/Scene.cc:15-27 defaults) */
int fh_scene_create(fh_device dev, fh_scene* out);
int fh_scene_destroy(fh_scene scene);
/* set/getCamera, set/getBackgroundColor, set/getBackgroundAlpha,
 * set/getLights: fresnel/gpu/Scene.cc:40-52 */
int fh_scene_set_camera(fh_scene scene, const fh_user_camera* cam);
int fh_scene_get_camera(fh_scene scene, fh_user_camera* cam);
int fh_scene_set_background(fh_scene scene, fh_rgb color, float alpha);
int fh_scene_get_background(fh_scene scene, fh_rgb* color, float* alpha);
int fh_scene_set_lights(fh_scene scene, const fh_lights* lights);
int fh_scene_get_lights(fh_scene scene, fh_lights* lights);
/* BVH maintenance (no reference counterpart: the reference rebuilds on every
 * commit, fresnel/gpu/Scene.cc).  With refit enabled, a change that keeps the
 * set of primitives (positions, radii, orientations moved) only recomputes the
 * boxes of the existing hierarchy.  Closest hits are defined independently of
 * the hierarchy, so images are bit-identical either way; only traversal speed
 * differs. */
int fh_scene_set_refit(fh_scene scene, int enable);
/* A refit is kept while the hierarchy's surface-area cost (sum of child-box
 * half areas) stays within `limit` times its value right after the last
 * build; beyond that the scene is rebuilt (default 1.5). */
int fh_scene_set_refit_limit(fh_scene scene, float limit);
int fh_scene_bvh_cost(fh_scene scene, float* built, float* current);
int fh_scene_bvh_info(fh_scene scene, uint64_t* builds, uint64_t* refits, uint32_t* n_prims);

/* ---------------------------------------------------------------- geometry */
#define FH_GEOM_SPHERE 0
#define FH_GEOM_CYLINDER 1
#define FH_GEOM_POLYGON 2
#define FH_GEOM_CONVEX_POLYHEDRON 3
#define FH_GEOM_MESH 4

/* GeometrySphere(scene, N), fresnel/gpu/GeometrySphere.cc */
int fh_geometry_sphere_create(fh_scene scene, uint32_t N, fh_geometry* out);
/* GeometryCylinder(scene, N), fresnel/gpu/GeometryCylinder.cc */
int fh_geometry_cylinder_create(fh_scene scene, uint32_t N, fh_geometry* out);
/* GeometryPolygon(scene, vertices[V,2], rounding_radius, N),
 * fresnel/cpu/GeometryPolygon.cc:20-79 */
int fh_geometry_polygon_create(fh_scene scene, const float* vertices, uint32_t n_verts,
                               float rounding_radius, uint32_t N, fh_geometry* out);
/* GeometryConvexPolyhedron(scene, origins[P,3], normals[P,3], colors[P,3], N, r),
 * fresnel/cpu/GeometryConvexPolyhedron.cc:19-109 (normals are normalized here) */
int fh_geometry_convex_polyhedron_create(fh_scene scene, const float* origins,
                                         const float* normals, const float* colors,
                                         uint32_t n_planes, uint32_t N, float radius,
                                         fh_geometry* out);
/* GeometryMesh(scene, vertices[3F,3], N), fresnel/cpu/GeometryMesh.cc:18-73 */
int fh_geometry_mesh_create(fh_scene scene, const float* vertices, uint32_t n_verts,
                            uint32_t N, fh_geometry* out);
/* Releases the handle; the geometry leaves the scene (Geometry::~Geometry,
 * fresnel/cpu/Geometry.cc:19-22). */
int fh_geometry_destroy(fh_geometry geom);

/* Buffer ids for fh_geometry_buffer */
#define FH_BUF_POSITION 0     /* sphere/polyhedron/mesh (N,3); polygon (N,2) */
#define FH_BUF_RADIUS 1       /* sphere/cylinder (N) */
#define FH_BUF_COLOR 2        /* sphere/polyhedron/polygon (N,3); cylinder (N,2,3); mesh (3F,3) */
#define FH_BUF_POINTS 3       /* cylinder (N,2,3) */
#define FH_BUF_ORIENTATION 4  /* polyhedron/mesh (N,4) quaternion (s,x,y,z) */
#define FH_BUF_ANGLE 5        /* polygon (N) */
/* Host view of a geometry attribute buffer (getPositionBuffer() etc.,
 * fresnel/cpu/Array.h:98-229).  The host copy is authoritative; it stays valid
 * for the life of the geometry.  shape[0..ndim) receives the numpy shape. */
int fh_geometry_buffer(fh_geometry geom, int which, float** data, uint32_t* ndim,
                       uint64_t shape[3]);
/* Geometry.update(): buffers changed, rebuild on the next render
 * (fresnel/cpu/Geometry.h:91-94, fresnel/gpu/Geometry.h:85-88). */
int fh_geometry_update(fh_geometry geom);
/* Partial form of update(): rows [first, first + count) along the first axis
 * of buffer `which` changed (the reference's Array __setitem__ followed by
 * update(), fresnel/util.py:41-48, which re-commits the whole geometry).  The
 * next render uploads only the marked rows; colour rows leave the hierarchy
 * untouched, other rows refit or rebuild it as fh_scene_set_refit selects.
 * FH_ERR_INVALID for an unknown buffer or rows beyond the buffer. */
int fh_geometry_update_rows(fh_geometry geom, int which, uint64_t first, uint64_t count);
/* get/setMaterial, get/setOutlineMaterial, get/setOutlineWidth, enable,
 * disable, remove: fresnel/gpu/Geometry.cc:125-139 */
int fh_geometry_set_material(fh_geometry geom, const fh_material* m);
int fh_geometry_get_material(fh_geometry geom, fh_material* m);
int fh_geometry_set_outline_material(fh_geometry geom, const fh_material* m);
int fh_geometry_get_outline_material(fh_geometry geom, fh_material* m);
int fh_geometry_set_outline_width(fh_geometry geom, float w);
int fh_geometry_get_outline_width(fh_geometry geom, float* w);
int fh_geometry_enable(fh_geometry geom);
int fh_geometry_disable(fh_geometry geom);
int fh_geometry_remove(fh_geometry geom);
/* ConvexPolyhedron set/getColorByFace, fresnel/cpu/GeometryConvexPolyhedron.h:63-72 */
int fh_geometry_set_color_by_face(fh_geometry geom, float f);
int fh_geometry_get_color_by_face(fh_geometry geom, float* f);
/* Polygon getRadius, fresnel/cpu/GeometryPolygon.h:56-59 */
int fh_geometry_get_radius(fh_geometry geom, float* r);

/* ----------------------------------------------------------------- tracers */
/* TracerPath(device, w, h, light_samples), fresnel/cpu/TracerPath.cc:20-27 */
int fh_tracer_path_create(fh_device dev, uint32_t w, uint32_t h, uint32_t light_samples,
                          fh_tracer* out);
/* TracerDirect(device, w, h), fresnel/cpu/TracerDirect.cc */
int fh_tracer_direct_create(fh_device dev, uint32_t w, uint32_t h, fh_tracer* out);
int fh_tracer_destroy(fh_tracer tr);
/* Tracer.resize(w, h), fresnel/cpu/Tracer.cc:23-30 and cpu/TracerPath.h:48-53 */
int fh_tracer_resize(fh_tracer tr, uint32_t w, uint32_t h);
int fh_tracer_get_size(fh_tracer tr, uint32_t* w, uint32_t* h);
/* Tracer.render(scene): one camera sample per pixel for a Path tracer
 * (cpu/TracerPath.cc:56-256), the full image for a Direct tracer
 * (cpu/TracerDirect.cc:33-225). */
int fh_tracer_render(fh_tracer tr, fh_scene scene);
/* `samples` consecutive Path render() calls in one submission; the result is
 * bit-identical to calling fh_tracer_render `samples` times (tracer.py:249-255). */
int fh_tracer_render_samples(fh_tracer tr, fh_scene scene, uint32_t samples);
/* getSRGBOutputBuffer / getLinearOutputBuffer + map(): copies the device
 * output into the tracer's host staging buffer and returns it
 * (fresnel/gpu/Array.h:203-215).  which: 0 = sRGB uint8 (h,w,4),
 * 1 = linear float32 (h,w,4). */
#define FH_OUT_SRGB 0
#define FH_OUT_LINEAR 1
int fh_tracer_map_output(fh_tracer tr, int which, void** data);
/* Asynchronous readback (no reference counterpart: the reference copies on
 * map(), fresnel/gpu/Array.h:203-215).  fh_tracer_output_async enqueues a copy
 * of the output (as rendered by the tracer's last completed render) into a
 * pinned host buffer of the tracer on the device's copy stream and returns at
 * once; fh_tracer_output_wait waits for that copy and returns the buffer and
 * the image size it was taken at (h, w, 4 values of uint8 sRGB or float32
 * linear), valid until the next fh_tracer_output_async of the same output or the
 * tracer's destruction; fh_tracer_output_ready tells whether it has completed.
 * A render of the tracer started meanwhile runs its traversal concurrently with
 * the copy and writes the output only after it, so a progressive client can
 * read frame k back while frame k + 1 is traced.  Not while a served request of
 * the tracer is active. */
int fh_tracer_output_async(fh_tracer tr, int which);
int fh_tracer_output_ready(fh_tracer tr, int which, int* ready);
int fh_tracer_output_wait(fh_tracer tr, int which, void** data, uint32_t* w, uint32_t* h);
/* The same host staging buffer without copying (no reference counterpart):
 * what the last fh_tracer_map_output left there, for the buffer protocol of an
 * array that was mapped before it is read. */
int fh_tracer_output_host(fh_tracer tr, int which, void** data);
/* Tracer.histogram() (fresnel/tracer.py:79-108) on the device-resident
 * linear output: counts[bin * 4 + column], 512 bins over [0, 1] of
 * lightness**(1/2.2), R**(1/2.2), G**(1/2.2), B**(1/2.2) of the pixels with
 * alpha > 0, lightness = 0.21 R + 0.72 G + 0.07 B.  Only the 8 KiB of counts
 * cross PCIe; the bin positions are k / 512, k = 1..512. */
int fh_tracer_histogram(fh_tracer tr, uint32_t counts[2048]);
/* fh_tracer_histogram without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The bin
 * counts accumulate in a file-scope device array of 8 KiB (the library's
 * fixed maximum, exactly the 4 x 512 counts; every GPU the library runs on
 * holds one until the process ends) instead of a buffer the tracer
 * allocates.  Images of 1 to 2^32 - 1 pixels; FH_ERR_INVALID otherwise.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several threads or several Device handles) race.  Same result as
 * fh_tracer_histogram. */
int fh_tracer_histogram_noalloc(fh_tracer tr, uint32_t counts[2048]);
/* enable/disableHighlightWarning, get/setSeed: cpu/Tracer.h:57-87 */
int fh_tracer_enable_highlight_warning(fh_tracer tr, fh_rgb color);
int fh_tracer_disable_highlight_warning(fh_tracer tr);
int fh_tracer_set_seed(fh_tracer tr, uint32_t seed);
int fh_tracer_get_seed(fh_tracer tr, uint32_t* seed);
/* TracerPath reset / setLightSamples / getNumSamples, cpu/TracerPath.cc:31-47,
 * cpu/TracerPath.h:55-65 */
int fh_tracer_path_reset(fh_tracer tr);
int fh_tracer_path_set_light_samples(fh_tracer tr, uint32_t light_samples);
int fh_tracer_path_get_num_samples(fh_tracer tr, uint32_t* n);
/* TracerDirect set/getAntialiasingN, cpu/TracerDirect.h:33-42 */
int fh_tracer_direct_set_aa_n(fh_tracer tr, uint32_t n);
int fh_tracer_direct_get_aa_n(fh_tracer tr, uint32_t* n);

/* ------------------------------------------------------ batching (no reference
 * counterpart: the reference renders one (scene, tracer) pair per call).
 * Renders `samples` passes for every (scenes[k], tracers[k]) pair in one
 * submission: all pairs' rays share scene-tagged queues in one set of
 * launches.  Each pair's output is bit-identical to fh_tracer_render_samples
 * on that pair alone. */
int fh_batch_render(fh_scene* scenes, fh_tracer* tracers, uint32_t n, uint32_t samples);

/* ------------------------------------------------------ serving (no reference
 * counterpart).  A server keeps admitted requests' scenes (BVH, primitives)
 * and accumulation buffers resident and advances all active requests together:
 * each fh_server_step admits queued requests first-come first-served while
 * their HBM footprints fit the budget (less the render scratches, see
 * fh_policy_scratch_bytes), retires started batches that have completed,
 * plans a step (every active request takes some camera samples) when none is
 * waiting, starts planned batches on the server's two render scratches and
 * the device's two render streams (a started batch is enqueued whole; the
 * device finds each bounce's rays itself; a Path render on the same device
 * between steps waits for the batches in flight), and retires requests whose
 * samples have all been
 * accumulated.  When nothing can start it waits for the oldest started batch.
 * Work stays in flight between calls; fh_server_destroy finishes the batches
 * already started.
 * A request's image is bit-identical to Path.sample(scene, samples,
 * light_samples) on its own: submission resets the tracer like sample()'s
 * reset, and scheduling never changes per-path arithmetic.
 * The scene and tracer of a request must stay alive until it is done; the
 * tracer's light-sample count is set back to 1 at completion, as
 * Path.sample leaves it.
 * Times are seconds; pass now < 0 to use the server's monotonic clock.
 * target_paths_per_step = 0 selects 256 Mi; the server's two render scratches
 * (about 2 x 160 B x target_paths_per_step) are charged against the budget. */
typedef struct fh_server_s* fh_server;
int fh_server_create(fh_device dev, uint64_t hbm_budget_bytes, uint64_t target_paths_per_step,
                     fh_server* out);
int fh_server_destroy(fh_server srv);
int fh_server_submit(fh_server srv, fh_scene scene, fh_tracer tracer, uint32_t samples,
                     uint32_t light_samples, double now, uint64_t* id);
int fh_server_step(fh_server srv, double now, uint32_t* n_active, uint32_t* n_queued);
/* state: 0 queued, 1 active, 2 done, 3 rejected (footprint above the budget),
 * 4 failed (a batch holding some of its samples could not start; that
 * fh_server_step returned the error); times[0..2] = arrival, admission,
 * completion (-1 if not yet). */
int fh_server_poll(fh_server srv, uint64_t id, int* state, uint32_t* samples_done, double times[3]);
/* Forget a done or rejected request (its id becomes unknown). */
int fh_server_release(fh_server srv, uint64_t id);
int fh_server_memory(fh_server srv, uint64_t* in_use, uint64_t* budget);
/* The server's policy without a device: footprint formula and a simulation of
 * admission and completion iterations for synthetic requests. */
int fh_policy_footprint(uint64_t n_prims, uint64_t prim_bytes, uint64_t W, uint64_t H, uint64_t L,
                        uint64_t* out);
/* Bytes of one render scratch for batches of `paths` light-sample paths plus
 * its size-independent part `fixed`.  A server's admission budget is its HBM
 * budget minus two such scratches (sized for target_paths_per_step) and any
 * render scratch the device holds for Path.sample (released when it would
 * keep the queue's head out). */
int fh_policy_scratch_bytes(uint64_t paths, uint64_t fixed, uint64_t* out);
int fh_policy_simulate(uint32_t n, const uint64_t* footprint, const uint64_t* paths_per_sample,
                       const uint32_t* samples, const uint32_t* arrive, uint64_t budget,
                       uint64_t target_paths, uint32_t max_iterations, int32_t* admit_iter,
                       int32_t* done_iter, uint64_t* peak_in_use);

/* ------------------------------------------------------ tuning (no reference
 * counterpart).  key "batch_paths": light-sample paths per wavefront batch
 * (default 256 Mi; sizes the ray queues, ≈ 160 B per path per render scratch); "tail_rays": below this many live
 * rays a batch finishes its paths in one kernel (default 524288); "stack_lds":
 * traversal stack entries kept in LDS per thread, 8 or 16 (default 8);
 * "trav_chunk": rays a secondary-ray traversal wavefront claims per
 * work-counter atomic, a multiple of 64 up to 65536 (default 256);
 * "primary_chunk": the most primary rays a primary-kernel workgroup claims at
 * once, a multiple of 64 up to 65536 (default 1024; a batch with fewer than 16
 * such claims per CU claims less, down to 256); "trav_refill": a traversal lane
 * whose ray is done takes a new one once this many lanes of its wavefront are
 * idle, 1..64, or 65 for the batch-synchronous kernel that traces 64 rays at
 * a time (default 65); "trav_packet": primary rays traced as wavefront
 * packets (2, default) or one by one (1); "primary_tiles": primary rays of
 * pinhole-perspective and orthographic cameras tested against per-8x8-tile
 * candidate lists (2, default) or searched in the BVH (1); "tile_min_rays":
 * with lists on, a tracer builds them once this many primary rays have been
 * rendered with one scene, camera and image size (this render's included;
 * default 4 Mi: building costs about what the BVH search loses on a few
 * million rays) and searches the BVH before; 1 builds them at the first
 * render; "tile_group": in
 * scenes other than spheres-only, those lists hold hierarchy subtrees whose
 * tile rectangle covers at most this many tiles (a mesh instance's triangles
 * become one entry, searched only by the rays that pass its box), 2..1024, or
 * 1 for one entry per primitive (default 4); "fuse_primary":
 * primary traversal and depth-0 shading in one kernel (2, default) or two (1);
 * "fuse_bounce": the same for each bounce of the secondary rays (default 1:
 * the fused kernel holds 88 registers, and the latency-bound secondary
 * traversal runs faster at the 8 wavefronts per SIMD of the separate kernel);
 * "inflight": batches of one render kept in flight on the device's two
 * render streams (2, default: one batch's kernels run while the host waits
 * for the other's queue counts) or 1; "xcd_bands": the primary kernel fills
 * the 16 queue segments with 16 image bands and the secondary traversal's
 * workgroups claim the rays of their own XCD's two bands first (2, default)
 * or the segments hold interleaved chunks and claims are global (1);
 * "graph": a render's batch kernels (primary rays and the bounce steps) are
 * captured into a HIP graph on first use and replayed for batches of the
 * same shape (2), or launched one by one (1, default); renders with
 * per-kernel timing or work counting always launch.
 * value <= 0 only reads the current value into *previous. */
int fh_device_tune(fh_device dev, const char* key, int64_t value, int64_t* previous);

/* ------------------------------------------------------ diagnostics (no
 * reference counterpart; used by the parity tests and the benchmark) */
/* Copy `bytes` (a multiple of 16) between two fresh device buffers with a
 * plain streaming kernel (k_debug_copy); *ms = its duration.  It reads and
 * writes exactly `bytes`, so HBM byte counters of other kernels can be scaled
 * by what the same counters read for it. */
int fh_debug_copy(fh_device dev, uint64_t bytes, float* ms);
/* Philox4x32-10 on the device: in[6k..6k+5] = ctr0..3, key0, key1 -> out[4k..] */
int fh_debug_philox(fh_device dev, const uint32_t* in, uint32_t n, uint32_t* out);
/* The shading code's sinf/cosf of lobe and aperture angles (correctly
 * rounded for 0 <= x <= 8) for n arguments: s[k] = sinf(x[k]), c[k] = cosf(x[k]). */
int fh_debug_sincos(fh_device dev, const float* x, uint32_t n, float* s, float* c);
/* Primary rays and closest hits of pass `sample` for every pixel of a Path
 * tracer's image: rays[6p] = org.xyz, dir.xyz; geom_id/prim = 0xffffffff on a
 * miss; t = +inf on a miss.  Any pointer may be NULL. */
int fh_debug_primary(fh_tracer tr, fh_scene scene, uint32_t sample, float* rays,
                     uint32_t* geom_id, uint32_t* prim, float* t);
/* The same hits through the renderer's own primary-ray kernel (tile lists or
 * BVH packets, per the tuning knobs), for camera sample number `sample` >= 1. */
int fh_debug_primary_render(fh_tracer tr, fh_scene scene, uint32_t sample, uint32_t* geom_id,
                            uint32_t* prim, float* t);
/* Closest hits for arbitrary rays (6 floats each) with tnear. */
int fh_debug_trace(fh_scene scene, const float* rays, uint32_t n, float tnear,
                   uint32_t* geom_id, uint32_t* prim, float* t);
/* The scene's BVH after bringing it up to date: *n_nodes 64-byte nodes
 * (child 0 box lo.xyz hi.xyz, child 1 box lo.xyz hi.xyz, then links c0, c1,
 * g0, g1; a link with the top bit set is a leaf holding primitive c & 0x7fffffff
 * of geometry slot g; 0xffffffff is an empty child).  Copies at most `cap`
 * nodes into `nodes` (NULL to query the count). */
int fh_debug_bvh(fh_scene scene, void* nodes, uint64_t cap, uint64_t* n_nodes);
/* Per-kernel device time of the most recent fh_tracer_render_samples /
 * fh_batch_render call, from HIP events recorded on the render stream.
 * names: '\n'-separated kernel names; ms[k]: summed time; launches[k]: count. */
int fh_stats_last(fh_device dev, char* names, size_t names_cap, double* ms,
                  uint32_t* launches, uint32_t cap, uint32_t* n_kernels);
/* Batches whose kernels were captured into a HIP graph, and captured graphs
 * replayed, since the device was created (tuning key "graph" = 2). */
int fh_stats_graphs(fh_device dev, uint64_t* captures, uint64_t* replays);
/* Enable per-kernel event timing (adds one event pair per launch). */
int fh_stats_enable(fh_device dev, int enable);
/* Deterministic work counters of the most recent render: rays traced, BVH
 * nodes visited and primitive tests of the secondary-ray traversal kernel
 * ([0], [1], [2]), the primary-ray traversal kernel ([3], [4], [5]) and the
 * tail kernel that finishes the last paths of a batch ([6], [7], [8]); [9]
 * paths stopped at the depth cap; [10] tile-list entries examined by primary
 * rays; [15] non-zero if a traversal stack overflowed.  Counting is enabled with fh_counters_enable and uses
 * instrumented kernels (slower; the counts do not depend on timing). */
int fh_counters_enable(fh_device dev, int enable);
int fh_counters_last(fh_device dev, uint64_t out[16]);

#ifdef __cplusplus
}
#endif

#endif /* FRESNEL_HIP_H */
