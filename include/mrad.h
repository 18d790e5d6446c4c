/*
 * mrad: MI355X (gfx950) path-tracing hot path behind a plain C ABI.
 *
 * This is the drop-in boundary for the reference's CUDA render path.  Every
 * entry point names the reference interface it replaces (paths relative to
 * the reference repository, src/xyz/marsavic/gfxlab/).  Arguments are plain
 * pointers and sizes; all arrays are borrowed and copied during the call.
 *
 * Status codes: MRAD_OK (0) or a negative MRAD_ERR_*; the message of the
 * last failure on the calling thread is returned by mrad_last_error().  A
 * Java binding maps MRAD_ERR_INVALID_ARGUMENT to IllegalArgumentException and
 * MRAD_ERR_ILLEGAL_STATE / MRAD_ERR_UNAVAILABLE / MRAD_ERR_HIP to
 * IllegalStateException, as GpuRayTracer does (gpu/GpuRayTracer.java:324-337,
 * :470-476, :572-580).
 *
 * Threading: every call on one tracer handle is serialized by an internal
 * mutex (the reference's synchronized(this), GpuRayTracer.java:339); each
 * handle owns one HIP stream on its device.
 */
#ifndef MRAD_H
#define MRAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	MRAD_OK = 0,
	MRAD_ERR_INVALID_ARGUMENT = -1,
	MRAD_ERR_ILLEGAL_STATE = -2,
	MRAD_ERR_UNAVAILABLE = -3,
	MRAD_ERR_HIP = -4,
	MRAD_ERR_OUT_OF_MEMORY = -5,
	MRAD_ERR_UNSUPPORTED = -6
};

/* BVH build modes built here (gpu/BvhBuildMode.java:4-12). */
enum { MRAD_BVH_UNIFORM_SAH = 0, MRAD_BVH_WEIGHTED_SAH = 1 };

/* Primitive kinds (refs' 2-bit type, gpu/GpuKernelSources.java:19-24). */
enum { MRAD_PRIM_SPHERE = 0, MRAD_PRIM_BOX = 1, MRAD_PRIM_AFFINE_SPHERE = 2, MRAD_PRIM_AFFINE_BOX = 3 };

/* Counter slots of mrad_frame_stats.counters (GpuKernelSources.java:30-43). */
enum {
	MRAD_STAT_RAYS = 0, MRAD_STAT_AABB_TESTS, MRAD_STAT_PRIMITIVE_TESTS, MRAD_STAT_SPHERE_TESTS,
	MRAD_STAT_BOX_TESTS, MRAD_STAT_PLANE_TESTS, MRAD_STAT_AFFINE_SPHERE_TESTS,
	MRAD_STAT_AFFINE_BOX_TESTS, MRAD_STAT_STACK_OVERFLOWS, MRAD_STAT_MAX_STACK_SIZE,
	MRAD_STAT_INTERNAL_NODES, MRAD_STAT_LEAF_NODES, MRAD_STAT_HOMOGENEOUS_LEAF_NODES,
	MRAD_STAT_MIXED_LEAF_NODES, MRAD_STAT_COUNT
};

/* ------------------------------------------------------------------ scene */

/* BvhBuildConfig record (gpu/BvhBuildConfig.java:6-15). */
typedef struct {
	int mode;                 /* MRAD_BVH_* */
	int leafSize;
	double traversalWeight;   /* PrimitiveCostModel.INTERIOR_TRAVERSAL = 1.275576713 */
	double sphereWeight, boxWeight, affineSphereWeight, affineBoxWeight;
	double lambda;            /* 0 for UNIFORM_SAH */
} mrad_bvh_config;

/* GpuScene.BvhStats (gpu/GpuScene.java:113-133), the fields the flat build has. */
typedef struct {
	int node_count, root_count, leaf_count, max_depth, primitive_ref_count, leaf_size;
	double generalized_sah_cost, uniform_sah_cost, weighted_sah_cost;
	int64_t bytes;
	int min_leaf_occupancy, max_leaf_occupancy;
	double mean_leaf_occupancy;
	int64_t build_nanos;
} mrad_bvh_stats;

/* The reference's packed wire format (GpuScene.java:24-31, :507-607). */
typedef struct {
	const float *spheres; int sphere_count;                /* f32 x 5  */
	const float *boxes; int box_count;                     /* f32 x 7  */
	const float *planes; int plane_count;                  /* f32 x 26 */
	const float *affine_spheres; int affine_sphere_count;  /* f32 x 22 */
	const float *affine_boxes; int affine_box_count;       /* f32 x 22 */
	const float *node_bounds; int node_count;              /* f32 x 6  */
	const int32_t *node_data;                              /* i32 x 4  */
	const int32_t *primitive_refs; int primitive_ref_count;
	const int32_t *root_indices; int root_count;           /* <= 4, roots[0] == 0 */
	const float *materials; int material_count;            /* f32 x 13 */
	float background[3];
	uint64_t revision;
} mrad_scene_arrays;

typedef struct mrad_scene mrad_scene;

/* new GpuScene(config, options) (GpuScene.java:165-170). */
int mrad_scene_create(const mrad_bvh_config *config, mrad_scene **out);
void mrad_scene_destroy(mrad_scene *scene);
/* GpuScene.setBackground (GpuScene.java:200-203). */
int mrad_scene_set_background(mrad_scene *scene, double r, double g, double b);
/* Materials are 13 doubles per primitive in packed order: diffuse.rgb,
 * reflective.rgb, refractive.rgb, emittance.rgb, refractive index; equal
 * materials are deduplicated (GpuScene.java:850-859). */
/* GpuScene.addSphere (GpuScene.java:205-221): center.xyz, radius per sphere. */
int mrad_scene_add_spheres(mrad_scene *scene, int n, const double *center_radius, const double *materials);
/* GpuScene.addBox (GpuScene.java:223-233): min.xyz, max.xyz per box. */
int mrad_scene_add_boxes(mrad_scene *scene, int n, const double *min_max, const double *materials);
/* GpuScene.addPlane (GpuScene.java:235-253): point.xyz, normal.xyz per plane. */
int mrad_scene_add_planes(mrad_scene *scene, int n, const double *point_normal, const double *materials);
/* GpuScene.addAffineSphere / addAffineBox (GpuScene.java:457-479): 3x4
 * inverse and inverse-transpose, world center, bounds min.xyz max.xyz. */
int mrad_scene_add_affine(mrad_scene *scene, int kind, int n, const double *inverse12,
                          const double *inverse_transpose12, const double *center3,
                          const double *bounds6, const double *materials);
/* GpuSceneBuilder.convertSolid for affine SceneSphere/SceneBox
 * (gpu/GpuSceneBuilder.java:37-78, :130-159): forward 3x4 transform. */
int mrad_scene_add_affine_transforms(mrad_scene *scene, int kind, int n, const double *transform12,
                                     const double *materials);
/* GpuScene.bvhData() (GpuScene.java:609-700): packs and builds lazily. */
int mrad_scene_build(mrad_scene *scene);
/* GpuScene.*Data(), bvh*() (GpuScene.java:507-607): views valid until the
 * scene is modified or destroyed. */
int mrad_scene_view(mrad_scene *scene, mrad_scene_arrays *out);
/* GpuScene.bvhStats() (GpuScene.java:607). */
int mrad_scene_bvh_stats(mrad_scene *scene, mrad_bvh_stats *out);

/* ----------------------------------------------------------------- tracer */

/* GpuRayTracer constructor arguments (gpu/GpuRayTracer.java:256-282) plus
 * the -Dgfxlab.gpu.* properties it reads (:1100-1103, GpuLaunchProvenance). */
typedef struct {
	int width, height;
	int samples_per_frame;
	int bvh_traversal;        /* boolean: BVH (1) or linear scan (0) */
	int collect_metrics;      /* boolean: counter variant of the kernels */
	int bvh_stack_size;       /* gfxlab.gpu.bvhStackSize, 8..256 (0 = 32) */
	int samples_per_launch;   /* physical launches per frame = ceil(spp / this); 0 = all */
	int device;               /* HIP device ordinal */
	int pixels_per_launch;    /* gfxlab.gpu.renderPixelsPerLaunch (GpuLaunchProvenance.java:44-60): 0 = the
	                             whole image per launch; else tiles of min(W, n) x max(1, n / that) pixels,
	                             1..262144 */
	int replay_rays_per_launch; /* gfxlab.gpu.replayRaysPerLaunch (GpuLaunchProvenance.java:30-33): rays per
	                               physical replay launch, 1..262144 (0 = 262144); a linear replay also keeps
	                               each launch within 2^20 primitive-work units per ray and 8192 launches
	                               (GpuRayTracer.java:743-778) */
	int64_t max_physical_kernel_nanos;  /* 0 = unchecked; else a frame with a physical launch at or above
	                                       this fails with MRAD_ERR_ILLEGAL_STATE (the reference's 250 ms
	                                       watchdog guard, GpuRayTracer.java:572-580) */
} mrad_tracer_config;

/* GpuRayTracer.FrameStats (GpuRayTracer.java:79-99). */
typedef struct {
	int64_t total_nanos, kernel_nanos, maximum_physical_kernel_nanos, copy_nanos, upload_nanos;
	uint64_t counters[MRAD_STAT_COUNT];
	int physical_launches;
} mrad_frame_stats;

typedef struct mrad_tracer mrad_tracer;

/* new GpuRayTracer(w, h, spp, bvh, metrics) (GpuRayTracer.java:256-282). */
int mrad_create(const mrad_tracer_config *config, mrad_tracer **out);
/* GpuRayTracer.close() (GpuRayTracer.java:1223-1310); NULL is a no-op. */
void mrad_destroy(mrad_tracer *tracer);
/* GpuRayTracer.isAvailable() (GpuRayTracer.java:284-287): 1 or 0. */
int mrad_is_available(mrad_tracer *tracer);
/* ensureSceneResident/uploadScene (GpuRayTracer.java:873-948): uploads and
 * re-lays-out the wire arrays unless (scene_key, arrays->revision) is the
 * resident pair. */
int mrad_upload_scene(mrad_tracer *tracer, const mrad_scene_arrays *scene, uint64_t scene_key);
/* Russian roulette after bounce start_depth with survival clamp(max T,
 * 0.05, 0.95); -1 (the default) renders exactly the reference integrator.
 * No counterpart in the reference: its integrator has no roulette. */
int mrad_set_russian_roulette(mrad_tracer *tracer, int start_depth);
/* Leaf batching of the megakernel's traversal: 0 (the default) tests a leaf
 * as soon as a lane pops it; 16 holds popped leaves until 16 lanes of the
 * wave (or every lane still traversing) hold one and tests them together.
 * Hits, images and counters are the same either way; leaf-heavy scenes (the
 * Layered Hero frames) render faster with 16, the seeded fields without.
 * No counterpart in the reference.  Other values: invalid argument. */
int mrad_set_leaf_batch(mrad_tracer *tracer, int lanes);
/* The form mrad_render_sample renders in.  MRAD_RENDER_FORM_MEGAKERNEL (the
 * default) is the reference's per-pixel kernel under its launch contract
 * (pixel tiles, sample ranges, the watchdog limit).  MRAD_RENDER_FORM_WAVEFRONT
 * renders the same frame with the wavefront form's bounded path pool and
 * persistent trace kernel over the tracer's resident scene: the image and the
 * 14 counters are bit-identical, large frames render faster, the physical
 * launches are its bounce steps (bounded by the pool, not by pixel tiles) and
 * the per-launch watchdog limit is not checked.  No counterpart in the
 * reference.  Other values: invalid argument. */
#define MRAD_RENDER_FORM_MEGAKERNEL 0
#define MRAD_RENDER_FORM_WAVEFRONT 1
int mrad_set_render_form(mrad_tracer *tracer, int form);
/* GpuRayTracer.renderSample(rgb, scene, camera, maxDepth, frameIndex,
 * frameSeed) (GpuRayTracer.java:317-432).  camera = position, forward,
 * right, up (GpuCamera.java:8-22).  rgb_out (W*H*3 floats) may be NULL to
 * leave the frame on the device (read it later with mrad_copy_frame).  The
 * frame runs as (pixel tiles) x (sample ranges) physical launches; a plan of
 * more than 8192 launches is an invalid argument, checked before the device
 * (renderLaunchCount, GpuRayTracer.java:550-564). */
int mrad_render_sample(mrad_tracer *tracer, float *rgb_out, const float camera[12], int max_depth,
                       int frame_index, uint64_t frame_seed);
/* GpuRayTracer.renderLaunchCount (GpuRayTracer.java:556-564): the number of
 * physical launches mrad_render_sample would make for this tracer's
 * configuration (pixel tiles x sample ranges), or MRAD_ERR_INVALID_ARGUMENT
 * when it exceeds 8192.  Needs no device. */
int mrad_render_launch_count(const mrad_tracer *tracer, int *launches);
/* The device-to-host copy of renderSample (GpuRayTracer.java:397-399). */
int mrad_copy_frame(mrad_tracer *tracer, float *rgb_out);
/* GpuRayTracer.traceReplay(scene, rays, n, metrics) (GpuRayTracer.java:593-716)
 * for the resident scene; outputs as TraceReplayResult (:119-126).
 * per_ray_counters (n x 14 u32, slot order as MRAD_STAT_*) may be NULL;
 * it requires a tracer created with collect_metrics.  The rays run in
 * physical launches of at most replay_rays_per_launch; a linear tracer's plan
 * (linearReplayRaysPerLaunch, :743-769) that exceeds the per-ray work budget
 * or 8192 launches is MRAD_ERR_INVALID_ARGUMENT. */
int mrad_trace_replay(mrad_tracer *tracer, const float *rays, int n, int collect_metrics,
                      int32_t *hit_flags, float *hit_distances, int32_t *hit_primitive_orders,
                      int32_t *hit_material_indices, float *hit_normals, uint32_t *per_ray_counters);
/* No counterpart in the reference.  mrad_trace_replay for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments, results and errors, but the
 * device staging of the rays, results and counters lives in one static device
 * array per GPU (6.8 MiB, held until the process ends), so the call makes no
 * hipMalloc or hipFree.  n above MRAD_NOALLOC_MAX_RAYS is an invalid argument.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (two tracer handles used from two threads) race. */
#define MRAD_NOALLOC_MAX_RAYS 65536
int mrad_trace_replay_noalloc(mrad_tracer *tracer, const float *rays, int n, int collect_metrics,
                              int32_t *hit_flags, float *hit_distances, int32_t *hit_primitive_orders,
                              int32_t *hit_material_indices, float *hit_normals, uint32_t *per_ray_counters);
/* GpuRayTracer.lastFrameStats() (GpuRayTracer.java:298-307). */
int mrad_get_frame_stats(mrad_tracer *tracer, mrad_frame_stats *out);
/* GpuRayTracer.deviceInfo() (GpuRayTracer.java:309-311): name and gfx arch. */
int mrad_device_info(mrad_tracer *tracer, char *name, int name_len, char *arch, int arch_len);
/* GpuRayTracer.compiledPtxSha256() (GpuRayTracer.java:314-316): SHA-256 hex
 * (64 chars + NUL) of the gfx950 code object embedded in this library. */
int mrad_code_object_sha256(char *out65);
/* GpuRayTracer.validateFullPathDepth (GpuRayTracer.java:470-476). */
int mrad_validate_full_path_depth(int max_depth);
/* Message of the last failed call on this thread ("" after success). */
const char *mrad_last_error(void);

/* ------------------------------------------------------ path-replay adjoint */
/* No counterpart in the reference.  Backward pass of mrad_render_sample for
 * the resident scene and this tracer's width, height, samples and roulette:
 * given grad_image = dLoss/dImage (W*H*3 floats, host), returns dLoss/d(diffuse
 * albedo) and dLoss/d(emittance) per primitive (primitive_count x 3 floats each,
 * primitives in the reference's global order: spheres, boxes, affine spheres,
 * affine boxes, planes).  Paths are re-traced from their seeds, not stored. */
int mrad_backward(mrad_tracer *tracer, const float *grad_image, const float camera[12], int max_depth,
                  int frame_index, uint64_t frame_seed, float *grad_albedo, float *grad_emission,
                  int primitive_count);
/* Replaces the resident material table (13 floats each, same count) until
 * another scene or revision is uploaded: the parameter update of an
 * optimisation step.  No counterpart in the reference. */
int mrad_update_materials(mrad_tracer *tracer, const float *materials13, int count);

/* ------------------------------------------------------ wavefront batch */
/* No counterpart in the reference (its only form is the per-pixel
 * renderKernel): many scenes rendered by one bounded pool of in-flight paths,
 * one persistent kernel per bounce step, captured as a hipGraph.  Every image
 * is bit-identical to mrad_render_sample's for the same view. */

/* One scene's render request: the renderSample arguments plus the
 * GpuRayTracer constructor's width, height and samples per frame. */
typedef struct {
	int width, height, samples;
	int max_depth;           /* 0..32 */
	int rr_start;            /* Russian roulette start depth, -1 = none */
	int frame_index;
	uint64_t frame_seed;
	float camera[12];        /* position, forward, right, up */
} mrad_view;

typedef struct {
	int device;
	int queue_capacity;      /* paths in flight per pool (0 = 4M) */
	int waves_per_cu;        /* trace kernel persistent-grid occupancy (0 = 28) */
	int bvh_stack_size;      /* 8..256 (0 = 32) */
	int linear_scan;         /* 1 = the reference's GPU_LINEAR scan instead of the BVH */
	int collect_metrics;     /* counter variant of the trace kernel */
	int use_graph;           /* capture the bounce steps as a hipGraph */
	int time_kernels;        /* time each kernel with HIP events (no graph; for measurement) */
	int lds_stack_entries;   /* traversal-stack entries per lane held in LDS: 4 or 10 (0 = 10) */
	int leaf_regroup;        /* 1 = regroup each wave's leaf tests by primitive class */
	int result_chunk_paths;  /* paths of whole pixels whose radiance is held for the ordered per-pixel sum, per
	                            result buffer (two are kept): 0 = 8 pools' worth in [2^24, 2^30]; at most 2^30 */
	int stack_storage;       /* traversal-stack entries per lane compiled ahead of time: 32, 64 or 256, at least
	                            bvh_stack_size (0 = 64, or 256 above a stack size of 64); the autotuned depth */
	int trace_passes;        /* closest-hit passes per bounce step: 1, 2 or 3 (0 = 2 for a queue_capacity of 2^25
	                            paths or more, else 1).  Every pass but the last suspends the queries still
	                            traversing after pass_pops pops (waves of camera paths only are not capped), and the
	                            next pass resumes them packed into full waves; images and counters are the same */
	int pass_pops;           /* traversal-stack pops of a capped pass, in [1, 2^20] (0 = 80) */
	int suspend_capacity;    /* suspended queries one pass can hold, in [64, queue_capacity] (0 = half the
	                            queue); a query that finds the list full finishes in its pass */
	int pass_dense_lanes;    /* a wave with at least this many lanes still traversing when its pop budget runs
	                            out is not suspended: it keeps going and is checked again every 32 pops; in
	                            [1, 64], -1 = every wave is suspended at the budget (0 = 16) */
} mrad_wavefront_config;

typedef struct {
	int64_t total_nanos, kernel_nanos, upload_nanos;
	int64_t steps;           /* bounce steps run (generate + trace + shade) */
	uint64_t paths;
	uint64_t counters[MRAD_STAT_COUNT];
	int64_t trace_nanos, shade_nanos, generate_nanos;   /* with time_kernels */
} mrad_batch_stats;

typedef struct mrad_batch mrad_batch;

int mrad_batch_create(const mrad_wavefront_config *config, mrad_batch **out);
void mrad_batch_destroy(mrad_batch *batch);
/* Appends a scene (wire arrays copied) with its view; index_out may be NULL.
 * A view has at most 2^32 pixels and 2^30 samples; the batch's paths are
 * numbered with 64 bits, so its size is bounded by device memory for the
 * scenes and images only. */
int mrad_batch_add_scene(mrad_batch *batch, const mrad_scene_arrays *scene, const mrad_view *view, int *index_out);
/* Renders every scene of the batch; images stay on the device. */
int mrad_batch_render(mrad_batch *batch);
/* New launch parameters for a loaded batch (the autotuner's knob); device,
 * stack size and traversal mode stay as created. */
int mrad_batch_configure(mrad_batch *batch, const mrad_wavefront_config *config);
/* Normalized RGB image (width*height*3 floats) of one scene. */
int mrad_batch_copy_image(mrad_batch *batch, int scene, float *rgb_out);
int mrad_batch_get_stats(mrad_batch *batch, mrad_batch_stats *out);
/* The batch's launch parameters with every default resolved (the zero fields
 * of the create or configure call replaced by the values in use, the queue
 * capacity rounded up to whole waves).  No counterpart in the reference. */
int mrad_batch_get_config(mrad_batch *batch, mrad_wavefront_config *out);

/* ------------------------------------------- primitive-cost calibration */
/* The measurement kernel of the reference's calibration campaign
 * (gpu/PrimitiveCostBenchmark.java:717-873) on this library's
 * traversal-time operations, for the weighted SAH's per-primitive weights
 * (gpu/PrimitiveCostModel.java:11-19).  The workload uses the reference's
 * calibration layout (PrimitiveCostBenchmark.java:488-580): profiles x records
 * rays (6 floats), then records spheres (5), boxes (7), affine records (22),
 * planes (26), node AABBs (6), node pairs (12); records a power of two. */
typedef struct mrad_cost mrad_cost;
/* compute_units receives the device's CU count (for sizing the grid). */
int mrad_cost_create(int device, const float *workload, int records, int profiles, int *compute_units,
                     mrad_cost **out);
/* One timed launch (PrimitiveCostBenchmark.runChunk, :158-199): operation 0
 * setup, 1 sphere, 2 box, 3 affine sphere, 4 affine box, 5 node AABB,
 * 6 interior traversal, 7 plane; elapsed_ms from HIP events around it. */
int mrad_cost_time(mrad_cost *cost, int operation, int profile, int iteration_offset, int iterations, int copies,
                   int blocks, int threads, float *elapsed_ms);
/* The accumulators the last mrad_cost_time launch wrote, one float per
 * thread in launch order (n <= blocks * threads of that launch): every
 * operation's results folded as the reference's primitiveCostLoop folds its
 * own (PrimitiveCostBenchmark.java:820-841), so each operation can be checked
 * against its intersector. */
int mrad_cost_output(mrad_cost *cost, float *out, int n);
void mrad_cost_destroy(mrad_cost *cost);

#ifdef __cplusplus
}
#endif

#endif /* MRAD_H */
