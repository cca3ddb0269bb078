/*
 * oracle.h - CPU restatement of swpathtracer's hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP core: only
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product never links it.
 *
 * Every function follows the reference's semantics (operation order, constants,
 * branches); the file:line each one restates is cited beside it in oracle.cpp.
 * Inputs use the POD formats of include/spt.h so the same buffers feed both.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include "../include/spt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* accel: 0 = NoAccelStructure (brute force), 1 = SBVH object-split SAH. */
orc_scene* orc_scene_create(int accel);
void orc_scene_destroy(orc_scene* s);
/* Adds triangles in order (prim id = insertion index). normals may be NULL
 * (flat Triangle) or n x 9 floats (SmoothTriangle; an all-zero row is a flat
 * Triangle). mat_ids may be NULL (0). */
int orc_scene_add_triangles(orc_scene* s, const float* verts, const float* normals, const uint32_t* mat_ids,
                            uint32_t n);
/* Spheres: n x 4 floats (center xyz, radius). */
int orc_scene_add_spheres(orc_scene* s, const float* spheres, const uint32_t* mat_ids, uint32_t n);
int orc_scene_build(orc_scene* s);
int orc_set_materials(orc_scene* s, const spt_material* mats, uint32_t n);

/* out[0..7] = nodes, inner, leaves, height, references, min_leaf, max_leaf, unreferenced */
int orc_bvh_stats(const orc_scene* s, uint64_t* out);
int orc_bvh_dump(const orc_scene* s, const char* path);
/* prim ids in depth-first left-first leaf order (n_prims entries). */
int orc_bvh_leaf_order(const orc_scene* s, uint32_t* prim_ids);

/* Closest hit per ray with the reference's counters. Any output may be NULL.
 * normals receives the hit's shading normal (0 on miss). */
void orc_trace(const orc_scene* s, const spt_ray32* rays, uint64_t n, spt_hit* hits, uint32_t* tests,
               uint32_t* isects, float* normals, int n_threads);

/* Primary ray through image point (x, y) (getRay of the reference cameras/ directory). */
int orc_camera_ray(const spt_camera* cam, float x, float y, spt_ray32* out);

/* One pixel-sampler sample: draws from (seed, path) and returns the screen
 * offset (sx, sy) added to the pixel's integer coordinates. */
int orc_pixel_sample(uint32_t sampler, uint32_t spp, uint32_t sample, uint64_t seed, uint64_t path, float* sx,
                     float* sy);

/* The counter-based uniform draw u(seed, path, k) (include/spt.h). */
float orc_rng_u(uint64_t seed, uint64_t path, uint32_t k);

/* Samples material m's layered BSDF for local wo with draws (seed, path):
 * out = {wi.x, wi.y, wi.z, pdf, f.r, f.g, f.b, refl.r, refl.g, refl.b, draws} */
int orc_material_sample(const spt_material* m, const float* wo, uint64_t seed, uint64_t path, float* out);

/* Single-layer BSDF value and pdf for a given (wo, wi) (no sampling):
 * out = {f.r, f.g, f.b, pdf}. Only for layers with a density
 * (Lambertian, Cook-Torrance, GGX). */
int orc_layer_eval(const spt_layer* l, const float* wo, const float* wi, float* out);

/* Exact conductor Fresnel reflectance for complex IOR eta + i k (GGX layer). */
float orc_fresnel_conductor(float cos_i, float eta, float k);
/* The float transcendentals both sides use: fn 0 sin, 1 cos, 2 exp, 3 log, 4 atan
   of x[0..n) into y. */
int orc_fmath(int fn, const float* x, float* y, uint64_t n);
/* FresnelSchlick with two IORs (materials/fresnel_schlick.cpp:20-38), red channel. */
float orc_schlick_ior(float ior_i, float ior_t, float cos_theta);

/* Path tracing (integrators/pathtracer.cpp). rgb32: w*h*3 pixel values as the
 * reference leaves them (sum / spp, FP32); rgb64: same with FP64 sums; path_rgb:
 * per path radiance. params->d_stream, when used, is a HOST pointer here.
 * overflows receives the number of paths that ran out of replay-stream draws. */
int orc_render(const orc_scene* s, const spt_camera* cam, const spt_render_params* p, float* rgb32, double* rgb64,
               float* path_rgb, uint64_t* overflows, int n_threads);

/* omp_get_max_threads() of the library (the thread count the reference's
 * render loop reports, integrators/pathtracer.cpp:74). */
int orc_omp_max_threads(void);

/* Ray casters (integrators/raycaster.h and the three ray caster sources). */
int orc_raycast(const orc_scene* s, const spt_camera* cam, uint32_t kind, void* out, int n_threads);

#ifdef __cplusplus
}
#endif

#endif
