/*
 * mcpt.h — C ABI of the MI355X (gfx950) path-tracing core.
 *
 * The reference renderer (MonoCUDA-PT, src/path_tracer.cu) is one CUDA
 * translation unit with no library boundary; its in-TU entry points are
 *   BVH::generateHierarchy          (src/path_tracer.cu:1063)
 *   trace_sample_kernel             (src/path_tracer.cu:2900, looped at :4338)
 *   renderWithWavefront             (src/path_tracer.cu:3757, called at :4336)
 *   resolve_kernel                  (src/path_tracer.cu:3092)
 * This header is the boundary that replaces them. Every function returns an
 * int status (MCPT_OK = 0, negative on error); mcpt_last_error() returns the
 * calling thread's last message. The reference instead prints and exits
 * (CUDA_CHECK, src/path_tracer.cu:55-69).
 *
 * Threading: one mcpt_ctx per host thread. A context owns one HIP stream on
 * one device and an HBM budget; scenes, films and servers belong to one
 * context. There is no global device state (the reference keeps the
 * environment map in a __constant__ global, src/path_tracer.cu:333; here it
 * is part of the scene).
 *
 * Memory: inputs passed as host pointers are copied into device memory the
 * library owns; the caller keeps its arrays. Films (accumulation buffers)
 * stay resident in HBM between render calls.
 */
#ifndef MCPT_H
#define MCPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define MCPT_OK 0
#define MCPT_DEFERRED 1          /* server: job accepted but waiting for HBM */
#define MCPT_PENDING 2           /* server: job not finished yet */
#define MCPT_ERR_INVALID (-1)    /* bad argument or state */
#define MCPT_ERR_HIP (-2)        /* HIP runtime error */
#define MCPT_ERR_BUDGET (-3)     /* request exceeds the context's HBM budget */
#define MCPT_ERR_NODEVICE (-4)   /* no usable gfx950 device */

/* ----------------------------------------------------------- scene records */
/* Material kinds, in the order of MaterialType (src/path_tracer.cu:484). */
#define MCPT_MAT_LAMBERTIAN 0
#define MCPT_MAT_DIELECTRIC 1
#define MCPT_MAT_CONDUCTOR 2
#define MCPT_MAT_EMISSIVE 3
#define MCPT_MAT_ISOTROPIC 4

/* triangle_hot (src/path_tracer.cu:660-670): the 36-byte record read by
 * every leaf test. */
typedef struct mcpt_tri_hot {
    float v0[3], v1[3], v2[3];
} mcpt_tri_hot;

/* triangle_cold (src/path_tracer.cu:672-680) without the inline Material and
 * AABB: shading data read once per confirmed hit. `material` indexes the
 * scene's material table. 64 bytes. */
typedef struct mcpt_tri_cold {
    float uv0[2], uv1[2], uv2[2];
    float n0[3], n1[3], n2[3];
    int32_t material;
} mcpt_tri_cold;

/* Material (src/path_tracer.cu:486-496) with a solid albedo that is already
 * linear (make_solid_texture converts sRGB at construction, :207-209). */
typedef struct mcpt_material {
    int32_t type;
    float albedo[3];
    float emission[3];
    float fuzz;
    float ior;
    float density;
    float phase_g;
    int32_t texture; /* 0: the solid albedo above; k > 0: the scene's texture
                        k - 1 (mcpt_scene_add_texture) gives the albedo     */
} mcpt_material;

/* Sphere (src/path_tracer.cu:845-875), 48 bytes. A moving sphere's centre
 * goes from `center` at time 0 to `center1` at time 1 (motion blur: every
 * camera ray draws a time in [0, 1) and its path keeps it). Spheres are
 * never on the light list; an emissive sphere is seen by rays that hit it. */
typedef struct mcpt_sphere {
    float center[3];
    float radius;
    float center1[3];
    int32_t moving;   /* 0: static (center1 unused) */
    int32_t material;
    int32_t pad[3];
} mcpt_sphere;

/* Light (src/path_tracer.cu:942-947): an emissive triangle. 64 bytes. */
typedef struct mcpt_light {
    mcpt_tri_hot tri;
    float emission[3];
    float area;
    float normal[3];
} mcpt_light;

/* Equirectangular environment map, RGBA FP32 rows top (v = 0, +y) first
 * (the layout loadPfmAsRgba32f produces, src/path_tracer.cu:335-368). */
typedef struct mcpt_env_desc {
    int32_t width, height;
    float intensity;
    float rotation_radians;
    const float* rgba; /* width*height*4 floats, host memory */
} mcpt_env_desc;

/* Ray for the trace hooks: origin, time, direction, tmax (32 bytes). */
typedef struct mcpt_ray {
    float origin[3];
    float time;
    float dir[3];
    float tmax;
} mcpt_ray;

/* PackedHit (src/path_tracer.cu:991-996), the deferred 16-byte hit record. */
typedef struct mcpt_packed_hit {
    float t;
    int32_t prim_type;  /* 0 = sphere, 1 = triangle, -1 = miss */
    int32_t prim_index; /* original sphere or triangle index, -1 = miss */
    int32_t did_hit;
} mcpt_packed_hit;

/* Camera as set up by Camera::initialize (src/path_tracer.cu:1610-1631). */
typedef struct mcpt_camera_desc {
    float look_from[3], look_at[3], up[3];
    float vfov_degrees;
    float aperture;
    float focus_distance;
    int32_t width, height;
    float background[3];
    float rr_cap; /* Russian-roulette cap, 0.95 in the reference (:1606) */
} mcpt_camera_desc;

typedef struct mcpt_camera {
    float center[3], pixel00[3], delta_u[3], delta_v[3], u[3], v[3], w[3];
    float lens_radius;
    int32_t width, height;
    float background[3];
    float rr_cap;
} mcpt_camera;

/* Reference AABB (src/path_tracer.cu:612-625): min, pad, max, pad. */
typedef struct mcpt_aabb {
    float min[3];
    float pad0;
    float max[3];
    float pad1;
} mcpt_aabb;

/* ------------------------------------------------------------- opaque ids */
typedef struct mcpt_ctx_s* mcpt_ctx;
typedef struct mcpt_scene_s* mcpt_scene;
typedef struct mcpt_film_s* mcpt_film;

const char* mcpt_last_error(void);
const char* mcpt_version(void);

/* ----------------------------------------------------------------- context */
/* hbm_budget_bytes = 0 means "all free memory at creation". */
int mcpt_ctx_create(int device, uint64_t hbm_budget_bytes, mcpt_ctx* out);
/* Returns MCPT_ERR_INVALID, and destroys nothing, while scenes, films,
 * servers or networks made from the context are still alive: they use its
 * stream and allocator. */
int mcpt_ctx_destroy(mcpt_ctx ctx);
/* bytes the context has allocated / may allocate */
int mcpt_ctx_memory(mcpt_ctx ctx, uint64_t* used, uint64_t* budget);
/* device name, CU count, clock, HBM size into a caller buffer (JSON) */
int mcpt_ctx_describe(mcpt_ctx ctx, char* buf, size_t buflen);

/* ------------------------------------------------------------------- scene */
/* Copies triangles, materials and the optional environment map into HBM and
 * builds the light list from emissive triangles, in triangle order
 * (finalizeScene, src/path_tracer.cu:1973-1985). Does not build the BVH. */
int mcpt_scene_create(mcpt_ctx ctx, const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, int32_t n_tris,
                      const mcpt_material* materials, int32_t n_materials, const mcpt_env_desc* env,
                      mcpt_scene* out);
/* The same with spheres. The BVH is built over the primitive list the
 * reference's finalizeScene forms (:1955-1972): the spheres first, then the
 * triangles, so sorted ids and leaf order count spheres from 0 and
 * triangles from n_spheres. n_tris may be 0 when there are spheres. The
 * batched render and the server take at most 256 scenes when one of them
 * has moving spheres (their rays then carry scene and time in one word). */
int mcpt_scene_create_ex(mcpt_ctx ctx, const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, int32_t n_tris,
                         const mcpt_sphere* spheres, int32_t n_spheres, const mcpt_material* materials,
                         int32_t n_materials, const mcpt_env_desc* env, mcpt_scene* out);
/* Returns MCPT_ERR_INVALID, and destroys nothing, while a server renders
 * the scene. */
int mcpt_scene_destroy(mcpt_scene scene);
int mcpt_scene_light_count(mcpt_scene scene, int32_t* n_lights);
int mcpt_scene_lights(mcpt_scene scene, mcpt_light* out /* n_lights */);
/* environment sampling tables (buildEnvironmentSamplingTables,
 * src/path_tracer.cu:370-385): row_cdf[h+1], col_cdf[h*(w+1)], total */
int mcpt_scene_env_tables(mcpt_scene scene, float* row_cdf, float* col_cdf, float* total_weight);

/* Optional copy-back of every LBVH stage for parity tests. Each pointer may
 * be NULL. Arrays are host memory sized as noted (n = triangle count). */
typedef struct mcpt_lbvh_dump {
    uint32_t* morton;       /* [n] per triangle, original order            */
    uint32_t* sorted_keys;  /* [n]                                          */
    int32_t* sorted_ids;    /* [n] stable (key, index) order                */
    int32_t* left;          /* [n-1] internal node children                 */
    int32_t* right;         /* [n-1]                                        */
    uint8_t* left_is_leaf;  /* [n-1]                                        */
    uint8_t* right_is_leaf; /* [n-1]                                        */
    int32_t* parent;        /* [n-1] internal parents, root = -1            */
    int32_t* leaf_parent;   /* [n]                                          */
    mcpt_aabb* internal_bounds; /* [n-1] refit bounds                       */
    mcpt_aabb* leaf_bounds; /* [n] triangle AABB padded by 1e-3, leaf order */
    mcpt_aabb* scene_bounds; /* [1]                                         */
    int32_t max_depth;      /* out: deepest leaf, in internal levels        */
    float build_ms;         /* out: device time of the whole build          */
} mcpt_lbvh_dump;

/* Morton codes, stable radix sort, Karras radix tree and refit, all on the
 * GPU (BVH::generateHierarchy + the host part of finalizeScene and the host
 * std::sort, src/path_tracer.cu:1930-1972, :4224-4245, :1063-1090); the
 * refit runs level by level, and the build also collapses the binary tree
 * into the four-wide quantized tree the fast traversal walks. The dump holds
 * the binary tree in the reference's node convention. */
int mcpt_lbvh_build(mcpt_scene scene, mcpt_lbvh_dump* dump /* may be NULL */);

/* Parity hook for the four-wide tree the fast traversal walks (no
 * counterpart in the reference, which walks its binary tree). nodes
 * receives up to cap 64-byte records: float origin[3]; uint32 meta (bits
 * 0-7, 8-15, 16-23: x, y, z scale exponents biased by 127; bits 24-31: child
 * count); uint8 q[6][4] (lo.x, lo.y, lo.z, hi.x, hi.y, hi.z of children 0-3:
 * child k's box is origin + q * 2^(e - 127) per axis); 8 unused bytes;
 * int32 child[4] (>= 0 wide node, < 0 leaf ~r). Wide node i is the i-th
 * binary internal node at an even level (root 0). n_wide receives the
 * count; tie (scene size ints, may be NULL) each leaf's rank in the
 * reference's leaf-test order. The scene must be built. */
int mcpt_lbvh_dump_wide(mcpt_scene scene, void* nodes, int32_t cap, int32_t* n_wide, int32_t* tie);

/* ------------------------------------------------------------ trace hooks */
/* Closest hit for n rays (BVH::traverseRay / stepTraversal,
 * src/path_tracer.cu:1092-1150, :1262-1321). `steps` and `leaves`, when not
 * NULL, receive per-ray internal-node visits and leaf tests. Host arrays. */
int mcpt_trace_closest(mcpt_scene scene, const mcpt_ray* rays, int32_t n, mcpt_packed_hit* hits, uint32_t* steps,
                       uint32_t* leaves);
/* Any hit in (1e-4, tmax) (BVH::traverseAnyHit, src/path_tracer.cu:1152-1207). */
int mcpt_trace_any(mcpt_scene scene, const mcpt_ray* rays, int32_t n, uint8_t* occluded);

/* ------------------------------------------------------------------ render */
#define MCPT_MODE_MEGAKERNEL 0 /* trace_sample_kernel (:2900)                 */
#define MCPT_MODE_WAVEFRONT 1  /* renderWithWavefront (:3757)                 */
#define MCPT_MODE_PERSISTENT 2 /* renderWithWavefront + extendKernelPersistent */

#define MCPT_TONEMAP_NONE 0
#define MCPT_TONEMAP_REINHARD 1
#define MCPT_TONEMAP_ACES 2

/* flags */
#define MCPT_RENDER_COUNT_TRAVERSAL 1u /* count extend steps/leaf tests      */
#define MCPT_RENDER_TIME_KERNELS 2u    /* per-kernel HIP event timing        */
#define MCPT_RENDER_REFERENCE_ORDER 4u /* extend in the reference's visit
                                          order (left child first) instead of
                                          near child first; same hits, the
                                          reference's step/leaf counts       */
#define MCPT_RENDER_PERSISTENT_SHADOW 8u /* accepted for compatibility: the
                                            persistent mode now traces shadow
                                            rays on the persistent scheme by
                                            default (scenes without media)   */
#define MCPT_RENDER_SHADOW_PER_RAY 16u   /* persistent mode: shadow rays one
                                            lane per ray instead            */

typedef struct mcpt_render_opts {
    int32_t mode;
    int32_t spp;          /* samples rendered by this call                  */
    int32_t sample_begin; /* index of the first sample (resume / sharding)  */
    int32_t max_depth;
    uint64_t seed;
    int32_t refill_lanes; /* persistent extend and shadow rays: idle lanes
                             of 64 before a refill; 0 picks the default (24
                             when the render runs two lanes, else 32)      */
    uint32_t flags;
    int32_t batch;        /* wavefront modes: samples in flight per pass
                             (paths = W*H*batch); 0 picks the default     */
    int32_t group_below;  /* persistent mode: a closest-hit or shadow queue
                             shorter than this is walked by 16 lanes per
                             ray (same hits); 0 picks the default (32768),
                             < 0 never                                     */
} mcpt_render_opts;

typedef struct mcpt_render_stats {
    double total_ms;      /* device time of the whole call                 */
    double camera_ms, extend_ms, shade_ms, shadow_ms, finalize_ms;
    uint64_t extend_launches;
    uint64_t extend_rays; /* closest-hit rays traced                       */
    uint64_t shadow_rays; /* any-hit rays traced                           */
    uint64_t paths;       /* camera paths = W*H*spp                        */
    uint64_t extend_steps;  /* with MCPT_RENDER_COUNT_TRAVERSAL            */
    uint64_t extend_leaves; /* with MCPT_RENDER_COUNT_TRAVERSAL            */
} mcpt_render_stats;

/* Camera::initialize (src/path_tracer.cu:1610-1631) on the host. */
int mcpt_camera_init(const mcpt_camera_desc* desc, mcpt_camera* out);

/* Films: the linear accumulation contract. accum[p] holds W*H*3 FP32 sums of
 * firefly-clamped samples (src/path_tracer.cu:3050-3057, :3740-3755). */
int mcpt_film_create(mcpt_ctx ctx, int32_t width, int32_t height, mcpt_film* out);
int mcpt_film_destroy(mcpt_film film);
int mcpt_film_clear(mcpt_film film);
int mcpt_film_read(mcpt_film film, float* rgb /* W*H*3 */);
int mcpt_film_write(mcpt_film film, const float* rgb /* W*H*3 */);

/* Renders opts->spp samples starting at opts->sample_begin and adds them to
 * the film. stats may be NULL. The call returns when the device is done. */
int mcpt_render(mcpt_scene scene, const mcpt_camera* camera, const mcpt_render_opts* opts, mcpt_film film,
                mcpt_render_stats* stats);

/* No counterpart in the reference. mcpt_render for callers that cannot
 * allocate device memory (serving loops that must not call hipMalloc,
 * harnesses that forbid it): the same film, bit for bit, from a fixed
 * 64 MiB workspace that the library holds on each device from the first
 * context on that device until the process ends. Nothing is allocated or
 * freed during the call. The wavefront and persistent modes carry at most
 * MCPT_NOALLOC_MAX_PATHS paths per pass: images with more pixels are
 * refused with MCPT_ERR_INVALID (the megakernel has no such limit);
 * opts->flags may not ask for kernel timing. Calls on one device are
 * serialised by the library (they share the workspace); the call returns
 * when the device is done. */
#define MCPT_NOALLOC_MAX_PATHS 299584
int mcpt_render_noalloc(mcpt_scene scene, const mcpt_camera* camera, const mcpt_render_opts* opts, mcpt_film film);

/* resolve_kernel (src/path_tracer.cu:3092-3100): accum / spp * exposure,
 * tone map, sRGB OETF. out_rgb is host memory, W*H*3. */
int mcpt_resolve(mcpt_film film, int32_t samples_so_far, int32_t tonemap, float exposure, float* out_rgb);

/* resolve_kernel followed by the output conversion of main
 * (src/path_tracer.cu:4366-4380) on the device: NaN shield, clamp to
 * [0, 1], (unsigned char)(255.99 x). out_rgb8 is host memory, W*H*3 bytes,
 * top row first. */
int mcpt_resolve_rgb8(mcpt_film film, int32_t samples_so_far, int32_t tonemap, float exposure, uint8_t* out_rgb8);

/* computeAutoExposure (src/path_tracer.cu:3101-3124): target_gray (0.18 in
 * the reference) / exp(mean over pixels of ln(1e-4 + luminance)), where a
 * non-finite luminance counts as 0. The sum is FP64 in a fixed order (the
 * reference adds FP32 warp sums atomically, in no fixed order). */
int mcpt_auto_exposure(mcpt_film film, int32_t samples_so_far, float target_gray, float* exposure);

/* PNG file of W*H RGB8 pixels, top row first (saveImageRgb8Png,
 * src/image_io.cpp:62-65); rows stored uncompressed in the zlib stream. */
int mcpt_write_png(const char* path, int32_t width, int32_t height, const uint8_t* rgb8);

/* ------------------------------------------------------------ render server
 * Continuous batching. The reference renders one image per process run;
 * these entry points have no counterpart in it. A server owns a pool of
 * paths in flight shared by every admitted job; each bounce step of the
 * pool refills freed slots with camera paths of admitted jobs, so work of
 * jobs that arrive while others run overlaps on the device. A job is
 * admitted when a job slot is free and its film (plus its frame buffers)
 * fits the server's HBM budget; otherwise it waits (MCPT_DEFERRED) and is
 * admitted as memory is released. Each job's film is bit-identical to
 * mcpt_render of the same job alone (samples of a pixel are summed in
 * sample order). Scenes stay resident; a server runs its own host thread
 * and HIP stream. */
typedef struct mcpt_server_s* mcpt_server;

typedef struct mcpt_server_opts {
    int32_t path_capacity;  /* paths in flight (0 = 8 Mi)                   */
    int32_t max_jobs;       /* admitted jobs at once (0 = 64)               */
    int32_t frames_per_job; /* samples of one job in flight (0 = auto)      */
    int32_t refill_lanes;   /* persistent extend refill threshold (0 = 32)  */
    uint64_t film_budget_bytes; /* HBM for films of admitted jobs (0 = half
                                   of the context's free budget)            */
    int32_t policy;         /* which admitted job's paths enter the pool
                               first: MCPT_SCHED_* (0 = shortest first)     */
    int32_t group_below;    /* a closest-hit or shadow queue shorter than
                               this is walked by 16 lanes per ray (0 =
                               32768, < 0 never)                           */
} mcpt_server_opts;

#define MCPT_SCHED_SHORTEST_FIRST 0 /* least remaining paths to start first */
#define MCPT_SCHED_FIFO 1           /* admission order                      */
#define MCPT_SCHED_SHORTEST_FIRST_AGED 2 /* shortest first, but the oldest
                                        running job may take up to half of
                                        each step's free slots before the
                                        others (bounds a large job's wait) */

typedef struct mcpt_job {
    int32_t scene;     /* index into the server's scene list */
    int32_t spp;
    int32_t max_depth;
    int32_t reserved;
    uint64_t seed;
    mcpt_camera camera;
} mcpt_job;

typedef struct mcpt_job_result {
    int32_t status; /* MCPT_OK done, MCPT_PENDING running, MCPT_DEFERRED waiting */
    int32_t width, height, spp;
    double submit_ms, start_ms, finish_ms; /* server clock                */
    uint64_t paths;
} mcpt_job_result;

typedef struct mcpt_server_stats {
    uint64_t iterations;
    uint64_t jobs_submitted, jobs_done, jobs_waiting, jobs_running;
    uint64_t paths_started, extend_rays, shadow_rays;
    uint64_t film_bytes_in_use, film_budget;
    double busy_ms; /* device time of the bounce steps */
} mcpt_server_stats;

int mcpt_server_create(mcpt_ctx ctx, const mcpt_scene* scenes, int32_t n_scenes, const mcpt_server_opts* opts,
                       mcpt_server* out);
/* returns MCPT_OK when admitted at once, MCPT_DEFERRED when queued */
int mcpt_server_submit(mcpt_server srv, const mcpt_job* job, uint64_t* job_id);
/* MCPT_OK when done (result filled), MCPT_PENDING / MCPT_DEFERRED otherwise */
int mcpt_server_poll(mcpt_server srv, uint64_t job_id, mcpt_job_result* result);
/* blocks until the job is done; copies its film (W*H*3 sums) when rgb != NULL */
int mcpt_server_wait(mcpt_server srv, uint64_t job_id, mcpt_job_result* result, float* rgb);
/* frees a finished job's film so waiting jobs can be admitted */
int mcpt_server_release(mcpt_server srv, uint64_t job_id);
int mcpt_server_stats_get(mcpt_server srv, mcpt_server_stats* stats);
int mcpt_server_destroy(mcpt_server srv);

/* --------------------------------------------------------- batched scenes
 * No counterpart in the reference, which builds and renders one scene per
 * run. Every scene's BVH is bit-identical to its own mcpt_lbvh_build and
 * every film to mcpt_render of its scene with the same options. */

/* Builds the BVHs of n scenes together: each LBVH stage is one launch over
 * the triangles of all of them (scenes of one triangle take the single
 * path). build_ms (optional) receives the device time. */
int mcpt_lbvh_build_batch(const mcpt_scene* scenes, int32_t n, float* build_ms);

typedef struct mcpt_batch_tune {
    int32_t samples_per_pass; /* samples of every scene in flight per pass
                                 (0 = as many as fit a quarter of the budget) */
    int32_t refill_lanes;     /* persistent extend refill threshold (0 = 32)   */
    int32_t blocks_per_cu;    /* persistent extend blocks per CU (0 = the
                                 occupancy limit)                              */
    int32_t reserved;
} mcpt_batch_tune;

/* Renders scene i through cameras[i] into films[i]; the rays of all scenes
 * share one wavefront (persistent extend). opts applies to every scene and
 * its mode must be MCPT_MODE_PERSISTENT. tune may be NULL. */
int mcpt_render_batch(const mcpt_scene* scenes, const mcpt_camera* cameras, int32_t n, const mcpt_render_opts* opts,
                      mcpt_film* films, const mcpt_batch_tune* tune, mcpt_render_stats* stats);

/* Chooses a tune for mcpt_render_batch. Queue capacity: every sample of
 * every scene in one pass when a quarter of the HBM budget holds it (a pass
 * ends with its slowest rays, so fewer, fuller passes are faster), else as
 * many samples as fit. Launch shape: the refill threshold, then the
 * persistent blocks per CU, each kept when it shortens the closest-hit
 * launch of a calibration bounce: the camera rays of the sample after the
 * requested range are traced and shaded once, and the next bounce's extend
 * launch is timed on up to 2 Mi of its rays for each candidate. No film is
 * written. A tune depends on the workload's shape and is meant to be reused
 * across batches of that shape. search_ms (optional) receives the device
 * time the search spent. */
int mcpt_batch_autotune(const mcpt_scene* scenes, const mcpt_camera* cameras, int32_t n, const mcpt_render_opts* opts,
                        mcpt_batch_tune* best, double* search_ms);

/* --------------------------------------------------------- radiance cache
 * No counterpart in the reference. An MLP trained online on the tracer's
 * own paths predicts the radiance scattered at a path vertex (next-event
 * light plus everything the continuation would gather). Features, 11
 * floats per vertex: position normalised to the scene box (3), outgoing
 * direction and normal as (theta/pi, phi/(2pi) + 1/2) (2 + 2), roughness
 * (1), albedo (3). They are encoded into 64 inputs (sin/cos at 6 octaves
 * for the position, 2 for the angles, the raw values and a constant 1),
 * then 4 hidden layers of 64 ReLU units and a linear RGB output; no
 * biases. Weights and activations are BF16 on the matrix cores with FP32
 * accumulation; master weights and Adam moments are FP32. */
typedef struct mcpt_mlp_s* mcpt_mlp;

#define MCPT_MLP_FEATURES 11
#define MCPT_MLP_INPUTS 64
#define MCPT_MLP_WIDTH 64
#define MCPT_MLP_HIDDEN 4
#define MCPT_MLP_OUTPUTS 3
/* W0..W3 as [64 out][64 in], then W4 as [3 out][64 in], row-major */
#define MCPT_MLP_PARAMS (4 * 64 * 64 + 3 * 64)

/* He-uniform initial weights from a PCG32 stream keyed by seed */
int mcpt_mlp_create(mcpt_ctx ctx, uint64_t seed, mcpt_mlp* out);
int mcpt_mlp_destroy(mcpt_mlp mlp);
int mcpt_mlp_get_params(mcpt_mlp mlp, float* params /* MCPT_MLP_PARAMS */);
/* replaces the weights and resets the optimiser */
int mcpt_mlp_set_params(mcpt_mlp mlp, const float* params);
/* host arrays: features n*11, rgb n*3 */
int mcpt_mlp_forward(mcpt_mlp mlp, const float* features, int32_t n, float* rgb);
/* Relative L2 loss of Mueller et al. 2021, mean over samples and channels
 * of (y - t)^2 / (y^2 + 0.01) with the denominator held constant, and its
 * gradient with respect to the parameters (MCPT_MLP_PARAMS floats). */
int mcpt_mlp_gradient(mcpt_mlp mlp, const float* features, const float* targets, int32_t n, float* loss,
                      float* grad);
/* Adam (beta1 0.9, beta2 0.99, eps 1e-8, bias-corrected) over the batch in
 * consecutive mini-batches of `batch` samples; loss = the last step's */
int mcpt_mlp_train(mcpt_mlp mlp, const float* features, const float* targets, int32_t n, int32_t batch, float lr,
                   float* loss);

typedef struct mcpt_cache_opts {
    int32_t train_paths;     /* image paths per frame traced without the
                                cache and recorded for training (0 = 65536) */
    int32_t train_batch;     /* samples per Adam step (0 = 16384)             */
    int32_t train_steps;     /* at most this many steps per frame (0 = 4)     */
    int32_t terminate_depth; /* diffuse hits at this bounce or later end in
                                the cache (0 = 1: the first secondary hit)   */
    float lr;                /* Adam step size (0 = 0.005)                    */
    float ema_decay;         /* queries use an exponential average of the
                                weights, this decay per Adam step (0 = the
                                latest weights)                              */
    int32_t frame_spp;       /* samples per pixel traced together in one
                                frame, between two training rounds (0 = 1)  */
} mcpt_cache_opts;

typedef struct mcpt_cache_stats {
    double total_ms;       /* device time of the call                        */
    double train_trace_ms; /* film update and the training batch          */
    double train_ms;       /* MLP forward + backward + Adam                  */
    double render_ms;      /* tracing the frame's paths                      */
    double infer_ms;       /* MLP queries of the image paths                 */
    uint64_t train_samples, queries, paths;
    float loss;            /* last training step                             */
    int32_t reserved;
} mcpt_cache_stats;

/* Renders opts->spp frames (samples sample_begin ...), one sample per pixel
 * each, as Mueller et al. 2021 train a neural radiance cache: the paths of
 * cache_opts->train_paths pixels drawn at random are traced without the
 * cache (unbiased samples of the image like any other) and every diffuse
 * vertex on them becomes a training sample (target: the radiance the path
 * gathered after the vertex divided by its throughput there); every other
 * path ends in a cache query at its first diffuse hit from terminate_depth
 * on. After the frame the network takes up to train_steps Adam steps on the
 * frame's samples. Until the network has taken a step, frames are traced
 * without the cache. The film receives the same clamped linear sums as
 * mcpt_render. Requires MCPT_MODE_PERSISTENT or MCPT_MODE_WAVEFRONT. */
int mcpt_render_cached(mcpt_scene scene, const mcpt_camera* camera, const mcpt_render_opts* opts, mcpt_film film,
                       mcpt_mlp mlp, const mcpt_cache_opts* cache_opts, mcpt_cache_stats* stats);

/* ------------------------------------------------------- scene generators */
#define MCPT_GEN_CORNELL 1 /* 36-triangle Cornell box (10 walls, 24 box, 2 light) */
#define MCPT_GEN_SOUP 2    /* seeded incoherent triangle soup + 16 emitters      */
#define MCPT_GEN_ICOSPHERE 3 /* displaced icosphere + random-triangle fill       */
#define MCPT_GEN_MEDIUM 5  /* soup + homogeneous isotropic medium box            */

typedef struct mcpt_gen_sizes {
    int32_t n_tris;
    int32_t n_materials;
} mcpt_gen_sizes;

/* Two-call pattern: pass NULL arrays to get sizes, then arrays of that size.
 * param = soup/sphere triangle count (0 = the workload default). */
int mcpt_gen_scene(int32_t kind, uint64_t seed, int64_t param, mcpt_tri_hot* hot, mcpt_tri_cold* cold,
                   mcpt_material* materials, mcpt_gen_sizes* sizes);
/* ------------------------------------------------------------- textures
 * sampleTexture (src/path_tracer.cu:216-234): a checker of two sRGB colours
 * (converted to linear like make_checker_texture, :211-213) or an RGBA8
 * image (sRGB encoded, top row first) read bilinearly with wrap in u and v
 * at (u, 1 - v), each texel decoded from sRGB before filtering. The
 * reference filters in the texture unit; here the filter is FP32 in
 * software, so image-texture results can differ from it in the last bits. */
#define MCPT_TEX_CHECKER 1
#define MCPT_TEX_IMAGE 2

typedef struct mcpt_texture_desc {
    int32_t type;         /* MCPT_TEX_CHECKER or MCPT_TEX_IMAGE            */
    float color_a[3];     /* checker: sRGB colour of even cells            */
    float color_b[3];     /* checker: sRGB colour of odd cells             */
    float scale;          /* checker: cells per unit of u and of v         */
    int32_t width, height;
    const uint8_t* rgba8; /* image: width*height*4 bytes, top row first    */
} mcpt_texture_desc;

/* Adds a texture to a scene; *index receives k - 1 for the material field
 * `texture = k`. A render refuses materials that name a texture not added. */
int mcpt_scene_add_texture(mcpt_scene scene, const mcpt_texture_desc* desc, int32_t* index);

/* Heterogeneous medium: a dense density grid, 48 bytes. Voxel (i, j, k)
 * holds density[(k * dims[1] + j) * dims[0] + i] at world position
 * origin + (i, j, k) * voxel_size; between voxels the density is trilinear
 * and outside the grid it is 0. It stands in for the reference's NanoVDB
 * float grid (VolumeGrid, src/volume_io.h:7-12; sampleVolumeDensity,
 * src/path_tracer.cu:784-789), stored dense rather than as a sparse tree. */
typedef struct mcpt_density_grid {
    int32_t dims[3];      /* voxels along x, y, z (>= 1)                   */
    float origin[3];      /* world position of voxel (0, 0, 0)             */
    float voxel_size[3];  /* world size of a voxel along x, y, z (> 0)     */
    int32_t reserved;
    const float* density; /* host values, finite and >= 0                  */
} mcpt_density_grid;

/* Gives every isotropic material of the scene the density grid, as the
 * reference does for a loaded volume (src/path_tracer.cu:2509-2519): such a
 * medium's extinction at p is material.density * grid(p); free flight is
 * delta tracking (:796-812) and shadow rays ratio tracking (:817-835)
 * against the majorant material.density * max(grid). *max_density (may be
 * NULL) receives max(grid). Replaces an earlier grid; not allowed while a
 * server holds the scene. */
int mcpt_scene_set_density_grid(mcpt_scene scene, const mcpt_density_grid* grid, float* max_density);

/* loadPpmAsRgba8 (src/path_tracer.cu:236-294): a binary PPM (P6, maxval
 * 255, '#' comments in the header) as RGBA8 with alpha 255. Two-call
 * pattern: rgba8 NULL to get the size. */
int mcpt_load_ppm(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8);

/* A PNG texture as RGBA8, top row first: the files the reference reads with
 * stb_image (loadImageRgba8, src/image_io.cpp:30-39; the texture fallback
 * at src/path_tracer.cu:2602). Every colour type and bit depth, Adam7,
 * tRNS; 16-bit samples keep their high byte, 1/2/4-bit grey is scaled to
 * 0..255. Two-call pattern: rgba8 NULL to get the size. mcpt_load_obj
 * (load_textures) and mcpt_load_pbrt read PNG textures the same way. */
int mcpt_load_png(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8);

/* A Truevision TGA texture as RGBA8, top row first (another stb_image
 * format the reference's texture fallback reads): colour-mapped (8-bit
 * index), true-colour (15/16/24/32-bit) and grey (8-bit) images, raw or
 * run-length encoded, either row origin. Two-call pattern: rgba8 NULL to
 * get the size. The texture loaders read .tga files the same way. */
int mcpt_load_tga(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8);

/* A baseline JPEG texture as RGBA8, top row first (the reference reads
 * JPEG textures through stb_image, src/image_io.cpp:1-40): sequential
 * Huffman-coded 8-bit files, grey or YCbCr with any sampling factors,
 * restart intervals; progressive and arithmetic-coded files are refused.
 * Chroma at half resolution is upsampled with the triangle filter.
 * Two-call pattern: rgba8 NULL to get the size. The texture loaders read
 * JPEG files (by their signature) the same way. */
int mcpt_load_jpeg(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8);

/* Environment map files for mcpt_env_desc (RGBA FP32, top row first, alpha
 * 1). Two-call pattern: rgba NULL to get the size.
 * mcpt_load_pfm: loadPfmAsRgba32f (src/path_tracer.cu:335-368), "PF"
 * little-endian files only (a positive scale is refused), bottom row first
 * in the file.
 * mcpt_load_hdr: Radiance RGBE, the file type `--env` reads with
 * loadImageRgbaFloat (src/image_io.cpp:41-60) and scripts/hdr2pfm.py:14-61
 * converts: "-Y h +X w" only, flat or new-style run-length rows, each texel
 * byte * 2^(e - 136), 0 when e = 0.
 * mcpt_load_exr: OpenEXR, the other file type loadImageRgbaFloat reads
 * (src/image_io.cpp:41-52, tinyexr's LoadEXR): single-part scanline files,
 * UINT / HALF / FLOAT channels, NONE / RLE / ZIPS / ZIP compression; R, G,
 * B, A by name (A = 1 when absent, one channel repeated in R, G, B). */
int mcpt_load_pfm(const char* path, int32_t* width, int32_t* height, float* rgba);
int mcpt_load_hdr(const char* path, int32_t* width, int32_t* height, float* rgba);
int mcpt_load_exr(const char* path, int32_t* width, int32_t* height, float* rgba);

/* A NanoVDB float grid (.nvdb) as the dense grid of
 * mcpt_scene_set_density_grid: the file's first float grid over its root's
 * index bounding box, with the grid's axis-aligned scale and translation as
 * voxel size and origin (the world position of the box's min voxel);
 * *max_density (may be NULL) receives the root's maximum statistic, the
 * majorant the reference takes. The reference reads the file with
 * nanovdb::io::readGrid (loadNanoVdbGrid, src/volume_io.cpp:45-53; the
 * statistics, takeFloatGrid, :11-43) and samples the sparse grid
 * trilinearly in index space (sampleVolumeDensity,
 * src/path_tracer.cu:784-789), which the dense grid reproduces. Raw and
 * ZIP-compressed grids; refuses BLOSC, non-float grids, a non-zero
 * background and maps with rotation. Two-call pattern: density NULL fills
 * grid->dims, origin and voxel_size; then density (dims[0]*dims[1]*dims[2]
 * floats, x fastest) is filled and grid->density points at it. */
int mcpt_load_nvdb(const char* path, mcpt_density_grid* grid, float* density, float* max_density);

/* ------------------------------------------------------------ OBJ import
 * Wavefront OBJ + MTL into triangle records (loadObject / loadMtlFile,
 * src/path_tracer.cu:2527-2859): fan triangulation, 1-based and negative
 * indices, the model normalised (bottom centre at the origin, largest side
 * 1) then scaled, rotated about +y and translated; area-weighted vertex
 * normals when the file has none; degenerate and out-of-range faces
 * skipped. Materials: index 0 is `fallback` (faces without a known
 * material), 1 + i the i-th MTL material (Ke > 0: emissive; otherwise
 * Lambertian with Kd converted from sRGB). With load_textures, a map_Kd
 * image is read as the PPM beside it (the path's extension replaced by
 * .ppm, convertTexturePathToPpm :2580-2591; one read per path,
 * getOrLoadObjTexture :2594-2611) and its material gets texture = 1 + the
 * image's position in the list mcpt_obj_texture hands out; a material
 * whose image cannot be read keeps its Kd colour, as in the reference.
 * Two-call pattern: NULL arrays to get the counts in `stats`, then arrays
 * of that size. */
typedef struct mcpt_obj_opts {
    float translation[3];    /* where the model's bottom centre ends up  */
    float uniform_scale;     /* its largest side after normalisation (> 0) */
    float rotation_y;        /* radians about +y                          */
    mcpt_material fallback;
    int32_t load_textures;   /* 1: read map_Kd images (PPM)               */
} mcpt_obj_opts;

typedef struct mcpt_obj_stats {
    int32_t n_tris;
    int32_t n_materials;       /* fallback + MTL materials               */
    int32_t skipped_faces;     /* out-of-range indices or zero area      */
    int32_t texture_requests;  /* map_Kd materials that are not emissive */
    int32_t missing_libraries; /* mtllib files that could not be opened  */
    int32_t textures_loaded;   /* distinct images read (load_textures)   */
    int32_t missing_textures;  /* map_Kd materials left with Kd (load_textures) */
    int32_t reserved;
} mcpt_obj_stats;

int mcpt_load_obj(const char* path, const mcpt_obj_opts* opts, mcpt_tri_hot* hot, mcpt_tri_cold* cold,
                  mcpt_material* materials, mcpt_obj_stats* stats);

/* The index-th image of this thread's last successful mcpt_load_obj call
 * (RGBA8, top row first), kept until that thread's next load. Add the
 * images to the scene in order with mcpt_scene_add_texture so that the
 * materials' texture fields name them. Two-call pattern: rgba8 NULL to get
 * the size. No reference counterpart: the reference uploads each image to
 * a texture object inside its loader. */
int mcpt_obj_texture(int32_t index, int32_t* width, int32_t* height, uint8_t* rgba8);

/* ------------------------------------------------------------ pbrt import
 * A pbrt-v3 scene file into triangle, sphere and material records
 * (loadPbrtScene, src/path_tracer.cu:1990-2483): trianglemesh, plymesh and
 * sphere shapes with their transforms and instances; materials, area
 * lights, media, colour textures (image maps read as PPM), lights and the
 * camera mapped as the reference maps them. Two-call pattern: NULL arrays
 * to get the counts in `info`, then arrays of that size. max_prims bounds
 * triangles + spheres (0: 2^30). */
typedef struct mcpt_pbrt_info {
    int32_t n_tris, n_spheres, n_materials;
    int32_t textures_loaded;    /* images for mcpt_pbrt_texture              */
    int32_t has_camera;         /* look_from .. vfov_degrees are valid       */
    int32_t width, height;      /* Film resolution (0 without a Film)        */
    float look_from[3], look_at[3], up[3];
    float vfov_degrees;
    float background[3];        /* an infinite light without a map           */
    int32_t has_env_map;        /* env_path names the infinite light's map   */
    float env_intensity, env_rotation_degrees;
    char env_path[1024];
    int32_t shapes, instances, skipped_shapes, approximated_lights;
    int32_t unsupported_textures, approximated_materials;
} mcpt_pbrt_info;

int mcpt_load_pbrt(const char* path, int64_t max_prims, mcpt_tri_hot* hot, mcpt_tri_cold* cold,
                   mcpt_sphere* spheres, mcpt_material* materials, mcpt_pbrt_info* info);
/* The index-th image texture of this thread's last successful
 * mcpt_load_pbrt call (material.texture = index + 1), RGBA8, top row first;
 * two-call pattern. No reference counterpart: the reference uploads each
 * image to a texture object inside its loader. */
int mcpt_pbrt_texture(int32_t index, int32_t* width, int32_t* height, uint8_t* rgba8);

/* Procedural equirect sky: gradient + sun disk + glow (closed form of
 * scripts/make_sky.py:5-27), width*height*4 floats, top row first. */
int mcpt_gen_sky(int32_t width, int32_t height, float* rgba);
/* Camera preset for a generated scene. */
int mcpt_gen_camera(int32_t kind, int32_t width, int32_t height, mcpt_camera_desc* out);

#ifdef __cplusplus
}
#endif

#endif /* MCPT_H */
