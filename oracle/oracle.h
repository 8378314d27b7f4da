/*
 * oracle.h -- C API of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * The oracle is a plain C++ restatement of MiTrace's CPU Tracer path
 * (src/Tracer/*.cpp, src/Scene/*.cpp of the reference).  It is used by the test-suite as the
 * checker of the GPU library and by bench.py as the CPU baseline.  Nothing in the product
 * (mitrace_amd/, include/) links or calls it.
 */
#ifndef MITRACE_ORACLE_H
#define MITRACE_ORACLE_H

#include <stdint.h>

#include "../include/mitrace_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_RNG_COUNTER 0 /* the keyed counter RNG the GPU uses (per (pixel, sample, dim)) */
#define ORC_RNG_MT19937 1 /* std::mt19937 per tile, the reference's consumption order (Tracer.cpp:21, 563-586) */

typedef struct orc_render_opts {
    int32_t threads;          /* worker threads (0 = hardware_concurrency)          */
    int32_t rng;              /* ORC_RNG_*                                          */
    int32_t faithful_inverse; /* 1 = glm::inverse per ray and instance, as Intersect.cpp:152 */
    uint32_t row_begin;       /* render only rows [row_begin, row_end) (0,0 = all)  */
    uint32_t row_end;
    int32_t tlas;             /* 1 = skip instances through a tree over their world boxes
                                 (same entered instances, order and results)          */
} orc_render_opts;

typedef struct orc_work {
    uint64_t rays, samples;
    uint64_t inst_aabb_tests, inst_entered, node_pops, tri_tests;
    uint64_t hits_shaded, misses;
} orc_work;

void* orc_scene_create(const mt_scene_desc* desc);
void orc_scene_destroy(void* scene);
const char* orc_last_error(void);

/* flattened BVH of one mesh, same export format as mt_scene_mesh_bvh */
int orc_scene_mesh_bvh(const void* scene, uint32_t mesh, uint32_t* nodes_out, uint32_t* node_count,
                       uint32_t* indices_out, uint32_t* index_count);
/* instance world AABBs: 6 floats per instance */
int orc_scene_world_aabbs(const void* scene, float* out);

/* IntersectScene for n rays {o.xyz, d.xyz}; work (optional) receives summed counters */
int orc_trace_closest(const void* scene, uint32_t n, const float* rays, mt_hit* hits, orc_work* work);
/* IntersectRayAABB (Intersect.cpp:9-18): ray {o, d}, box {min, max} -> exit t or -1 */
float orc_slab(const float* ray, const float* box);
/* glm::intersectRayTriangle contract: ray {o, d}, tri {v0, v1, v2} -> hit, uvt {u, v, t} */
int orc_triangle_test(const float* ray, const float* tri, float* uvt);
/* the 4x4 inverse used for the per-ray local transform (column-major) */
void orc_invert4(const float* m, float* out);
/* exact closest hit in double precision over every triangle (no BVH, no culling slack) */
int orc_trace_bruteforce(const void* scene, uint32_t n, const float* rays, mt_hit* hits);

/* BRDF: in 14 floats per case, out 5 floats (same layout as mt_brdf_eval) */
void orc_brdf_eval(uint32_t n, const float* in, float* out);
/* sampling helpers for distribution tests: in {V.xyz, N.xyz, base.rgb, rough, metal, u0..u4}
 * (16 floats) -> out {dir.xyz, pdf, energy.rgb, specular(0/1)} (8 floats) */
void orc_sample_bounce(uint32_t n, const float* in, float* out);
/* equirect sample of the scene environment, n directions -> rgb (intensity multiplier 1) */
int orc_env_sample(const void* scene, uint32_t n, const float* dirs, int32_t apply_rotation,
                   float rotation_deg, float* rgb);
/* the path's elementary functions: in = n pairs {x, y}; MT_MATH_ATAN2 computes atan2(x, y),
 * the others use x only */
void orc_math(int32_t op, uint32_t n, const float* in, float* out);
/* One bounce of the differentiable estimator with given uniforms: in = 22 floats per case
 * {V.xyz, N.xyz, base'.rgb, rough, metal, tangent {dbase'.rgb, drough, dmetal}, u0..u4 (lobe,
 * direction, fallback), pad}; out {dir.xyz, w.rgb = f cos / pdf, its tangent with
 * the sampling detached, score (tangent of log p(decisions)), tangent of log pdf(dir), p_spec,
 * fallback (1 when the specular reflection fell below the surface)}: 13 floats */
void orc_bounce_terms(uint32_t n, const float* in, float* out);
/* Image::Sample (Image.cpp:11-61) of a w x h RGB8 image at n uv pairs -> n RGBA floats */
int orc_sample_image(uint32_t w, uint32_t h, const uint8_t* rgb8, uint32_t n, const float* uv,
                     int32_t filter, float* out);
/* ComputeNormal (Tracer.cpp:95-105): surface normals and tangent-space normals, 3 floats each */
void orc_compute_normal(uint32_t n, const float* surf, const float* tex, float* out);
/* One light sample of ProcessRayForward (Tracer.cpp:257-315) at n hit positions, each with
 * the counter RNG of (seed, pixel i, sample 0) from dimension base_dim: the chosen light (-1 =
 * none), its selection probability, the sampled light point and the light colour there. */
int orc_light_sample(const void* scene, uint32_t n, const float* pos, uint64_t seed, uint32_t base_dim,
                     int32_t* chosen, float* pdf, float* lpos, float* lcol);
/* camera rays for (u,v) pairs */
void orc_camera_rays(const mt_camera_desc* cam, uint32_t n, const float* uv, float aspect,
                     float* rays);
/* ACES of L*2^-ev (Tracer.cpp:212-222) */
void orc_aces(uint32_t n, const float* rgb, float ev, float* out);
/* CleanupFireflies (Tracer.cpp:490-548) on one pixel's samples; returns the divisor and
 * zeroes rejected samples in place */
uint32_t orc_cleanup_fireflies(float* samples_rgb, uint32_t count, float threshold);
/* the counter RNG: n draws for (seed, pixel, sample, dim) */
void orc_rng_draws(uint64_t seed, uint32_t n, const uint32_t* pixel, const uint32_t* sample,
                   const uint32_t* dim, float* out);

/* full render (RenderBlock/ProcessRayForward); rgb_out W*H*3; sample_rgb_out optional
 * [pixel][sample][3] */
int orc_render(const void* scene, const mt_camera_desc* cam, const mt_settings* settings,
               const orc_render_opts* opts, float* rgb_out, float* sample_rgb_out,
               orc_work* work, double* seconds);

/* differentiable parameters for orc_render_diff (HOST pointers, NULL = the scene's value);
 * same meaning as mt_diff_params */
typedef struct orc_diff_params {
    const float* tex; /* tex_h*tex_w*3 RGB float, replaces the base colour of materials
                         [tex_first, tex_first + tex_count) */
    uint32_t tex_w, tex_h, tex_first, tex_count;
    const float* rough; /* per material */
    const float* metal; /* per material */
    const float* emis;  /* per material, RGB */
} orc_diff_params;

/* EvaluateBRDF over tangent numbers: in as orc_brdf_eval (14 floats), tin 5 floats
 * {d base.rgb, d rough, d metal} -> out {f.rgb, df.rgb} */
void orc_brdf_tangent(uint32_t n, const float* in, const float* tin, float* out);

/* The differentiable estimator (linear radiance, no firefly rejection, counter RNG):
 * image = mean radiance per pixel; dimage = its directional derivative along tangent_dir
 * (same layout as params, NULL members = 0), forward mode.  sampling_terms = 0: the sampling
 * detached (decisions and pdfs held fixed); 1: plus the dependence of the sampling decisions
 * on the parameters (the lobe choice's probability, the GGX half-vector density and the
 * mixture pdf), times the radiance collected after each decision -- the score-function form
 * of the derivative of the expected image. */
int orc_render_diff(const void* scene, const mt_camera_desc* cam, const mt_settings* settings,
                    const orc_diff_params* params, const orc_diff_params* tangent_dir,
                    int32_t sampling_terms, int32_t threads, float* image, float* dimage);

#ifdef __cplusplus
}
#endif
#endif
