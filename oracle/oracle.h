/*
 * oracle.h — CPU oracle for the path-tracing core. TEST INFRASTRUCTURE ONLY:
 * the test suite, the smoke check and the CPU-baseline leg of the benchmark
 * call it as the thing results are compared against. The product library
 * never links it.
 *
 * It restates the reference's device functions (MonoCUDA-PT
 * src/path_tracer.cu, cited per function in oracle.cpp) in host C++17,
 * built with g++ -O2 -ffp-contract=off and no fast-math. Elementary
 * functions come from include/mcpt_fmath.h and random numbers from
 * include/mcpt_pcg32.h; both are pinned independently (libm error bounds,
 * an independent PCG32 restatement) by the tests.
 */
#ifndef MCPT_ORACLE_H
#define MCPT_ORACLE_H

#include "mcpt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene_s* oracle_scene;

/* Camera::initialize (src/path_tracer.cu:1610-1631) */
int oracle_camera_init(const mcpt_camera_desc* desc, mcpt_camera* out);

/* buildEnvironmentSamplingTables (src/path_tracer.cu:370-385) */
void oracle_env_tables(const float* rgba, int w, int h, float* row_cdf, float* col_cdf, float* total_weight);

/* finalizeScene (src/path_tracer.cu:1930-1987), triangles only: padded
 * boxes, scene bounds, centroids and Morton codes. centroids may be NULL. */
void oracle_finalize(const mcpt_tri_hot* hot, int n, mcpt_aabb* tri_bounds, mcpt_aabb* scene_bounds, float* centroids,
                     uint32_t* morton);
/* expandBits / quantize / morton3D (src/path_tracer.cu:913-935) */
uint32_t oracle_expand_bits(uint32_t v);
uint32_t oracle_quantize(float v, float lo, float hi);

/* stable (key, index) sort: the tie rule of this build */
void oracle_sort_stable(const uint32_t* morton, int n, uint32_t* keys, int32_t* ids);
/* std::sort with the reference comparator (src/path_tracer.cu:4224-4226) */
void oracle_sort_reference(const uint32_t* morton, int n, uint32_t* keys, int32_t* ids);

/* buildInternalNodesKernel (src/path_tracer.cu:1337-1386, :1417-1447),
 * sequential. Arrays sized n-1 (internal) and n (leaf_parent). */
void oracle_karras(const uint32_t* keys, int n, int32_t* left, int32_t* right, uint8_t* left_leaf, uint8_t* right_leaf,
                   int32_t* parent, int32_t* leaf_parent);

/* Whole scene: BVH (finalize + stable sort + Karras + refit), light list,
 * environment (env may be NULL). */
int oracle_scene_create(const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, int n, const mcpt_material* mats,
                        int n_mats, const mcpt_env_desc* env, oracle_scene* out);
/* computeSphereUvFromOutwardNormal (src/path_tracer.cu:877-884) for n unit normals */
void oracle_sphere_uv(const float* normals, int n, float* uv);
/* the same with spheres: the BVH's primitives are the spheres, then the
 * triangles (finalizeScene, src/path_tracer.cu:1955-1972) */
int oracle_scene_create_ex(const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, int n_tris, const mcpt_sphere* spheres,
                           int n_spheres, const mcpt_material* mats, int n_mats, const mcpt_env_desc* env,
                           oracle_scene* out);
void oracle_scene_destroy(oracle_scene s);
/* sampleTexture's checker and image textures (src/path_tracer.cu:216-234);
 * returns the texture's index (materials name it as texture = index + 1). */
int oracle_scene_add_texture(oracle_scene s, const mcpt_texture_desc* desc);
/* heterogeneous media: the grid, its trilinear density at n points, and
 * delta tracking (fn 0) / ratio tracking (fn 1) along n rays of 7 floats */
int oracle_scene_set_density_grid(oracle_scene s, const mcpt_density_grid* g, float* max_density);
void oracle_grid_density(oracle_scene s, const float* pts, int n, float* out);
void oracle_grid_track(oracle_scene s, int fn, float sigma, const float* rays, int n, uint64_t seed, float* out);
/* the albedo material m gives at n (u, v) points (closed-form tests) */
void oracle_sample_albedo(oracle_scene s, const mcpt_material* m, const float* uv, int n, float* out);
int oracle_scene_dump(oracle_scene s, mcpt_lbvh_dump* dump);
int oracle_scene_lights(oracle_scene s, mcpt_light* out, int cap);

/* traverseRay (closest) / traverseAnyHit, with per-ray visit counts. */
void oracle_trace_closest(oracle_scene s, const mcpt_ray* rays, int n, mcpt_packed_hit* hits, uint32_t* steps,
                          uint32_t* leaves, int threads);
void oracle_trace_any(oracle_scene s, const mcpt_ray* rays, int n, uint8_t* occluded, int threads);
/* every triangle in index order, strict '<' (brute force) */
void oracle_brute_closest(const mcpt_tri_hot* hot, int n_tris, const mcpt_ray* rays, int n, mcpt_packed_hit* hits);

typedef struct oracle_render_stats {
    uint64_t extend_rays, shadow_rays, paths;
    uint64_t extend_steps, extend_leaves;
    double seconds;
    int threads;
} oracle_render_stats;

/* trace_sample_kernel (src/path_tracer.cu:2900-3059) for samples
 * [sample_begin, sample_begin + spp) of rows [row_begin, row_end).
 * accum (W*H*3 FP32) is added to in sample order; accum64 (may be NULL)
 * receives the same sums in FP64. */
int oracle_render(oracle_scene s, const mcpt_camera* cam, int spp, int sample_begin, int max_depth, uint64_t seed,
                  int row_begin, int row_end, float* accum, double* accum64, int threads, oracle_render_stats* stats);

/* Elementary functions of mcpt_fmath.h over arrays (for accuracy tests):
 * fn 0 sin, 1 cos, 2 asin, 3 acos, 4 atan2(x, y), 5 log, 6 exp, 7 pow(x, y) */
void oracle_fmath(int fn, const float* x, const float* y, float* out, int n);
/* first n outputs of the PCG32 stream of (pixel, sample, seed) */
void oracle_pcg32(uint32_t pixel, uint32_t sample, uint64_t seed, int n, uint32_t* out);

/* Sampling functions on given uniforms, one record per call (closed-form
 * tests). fn: 0 refract(d[3], n[3], eta) -> dir[3]; 1 schlick(cos, eta) -> r;
 * 2 dielectric_scatter_direction(d[3], n[3], eta, u) -> dir[3];
 * 3 henyeyGreenstein_scatter_direction(in[3], g, u1, u2) -> dir[3];
 * 4 HG pdf(cos, g) -> pdf. Returns -1 for an unknown fn. */
int oracle_sampling(int fn, const float* in, int n, float* out);
/* The Conductor scatter of shadeKernel (src/path_tracer.cu:3657-3661):
 * normalize(reflect(normalize(d), n) + fuzz * random_unit_vector) for n
 * draws, draw i from the path stream (pixel i, sample 0, seed); out n*3. */
void oracle_conductor(const float* in_dir, const float* normal, float fuzz, int n, uint64_t seed, float* out);
/* Camera::getRay (src/path_tracer.cu:1633-1658) for n (x, y) pixels, each
 * from the path stream (y*width + x, sample, seed) the renders use; out n*7:
 * origin, direction, time. */
void oracle_camera_rays(const mcpt_camera* cam, const int32_t* pixels, int n, uint32_t sample, uint64_t seed,
                        float* out);
/* computeTriangleUvFromHitPoint and computeTriangleShadingNormal
 * (src/path_tracer.cu:687-713) of one triangle at n points; the face
 * normal is the unit normal of hit_triangle (:716-718). out uv n*2, normal n*3. */
void oracle_tri_interp(const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, const float* points, int n, float* uv,
                       float* normal);
/* Environment sampling on an RGBA map (w x h, rotation in radians):
 * fn 0 sample(u[4]) -> dir[3], pdf; 1 pdf(dir[3]) -> pdf;
 * 2 uv(dir[3]) -> u, v; 3 direction(u, v) -> dir[3]. */
int oracle_env_probe(const float* rgba, int w, int h, float rot, int fn, const float* in, int n, float* out);

/* resolve_kernel (src/path_tracer.cu:3062-3100) */
void oracle_resolve(const float* accum, int w, int h, int spp, int tonemap, float exposure, float* out);
/* resolve, then main's NaN shield and 8-bit conversion (:4366-4380) */
void oracle_resolve_rgb8(const float* accum, int w, int h, int spp, int tonemap, float exposure, uint8_t* out);
/* computeAutoExposure (:3101-3124), FP64 sum in pixel order */
float oracle_auto_exposure(const float* accum, int w, int h, int spp, float target);

#ifdef __cplusplus
}
#endif

#endif
