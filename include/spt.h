/*
 * spt.h - C ABI of the MI355X (gfx950) path-tracing core.
 *
 * The core sits behind swpathtracer's AccelStructure / Integrator interfaces:
 *   AccelStructure  acceleration/accel_structure.h:15-35   (build + intersect)
 *   Integrator      integrators/integrator.h:8-55            (render + stats)
 * Host C++ keeps scene assembly and the SAH BVH build (acceleration/sbvh.cpp);
 * everything per ray / per path runs as hand-written HIP kernels.
 *
 * Conventions
 *   - Every entry point returns spt_status (0 = SPT_OK). No C++ exception crosses
 *     this boundary; spt_last_error() gives the message of the last failure.
 *   - Buffers named d_* are DEVICE pointers (hipMalloc'd, or spt_malloc'd).
 *     Buffers without the prefix are host pointers and are copied; the caller
 *     keeps ownership.
 *   - A context owns one HIP stream. Launches are asynchronous on that stream;
 *     spt_sync() waits. A context is used by one host thread at a time.
 *   - Floating-point semantics follow the reference: IEEE binary32, no FMA
 *     contraction, correctly rounded divide and sqrt, denormals kept.
 */
#ifndef SPT_H
#define SPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t spt_status;
enum {
    SPT_OK = 0,
    SPT_ERR_INVALID = 1,     /* bad argument / shape */
    SPT_ERR_DEVICE = 2,      /* HIP runtime error or missing device */
    SPT_ERR_NOMEM = 3,       /* allocation failed */
    SPT_ERR_STATE = 4,       /* call made in the wrong state (no scene, ...) */
    SPT_ERR_UNSUPPORTED = 5, /* feature not available for this input */
    SPT_ERR_OVERFLOW = 6     /* a bounded resource (replay stream, stack) ran out */
};

#define SPT_MISS 0xFFFFFFFFu

/* ------------------------------------------------------------------------- */
/* Data formats                                                              */
/* ------------------------------------------------------------------------- */

/* A ray: Ray (core/ray.h:6-19) padded to 32 B so a lane loads it as 2 x 16 B. */
typedef struct spt_ray32 {
    float ox, oy, oz, pad0;
    float dx, dy, dz, pad1;
} spt_ray32;

/* Closest hit: t and the hit primitive's id in scene insertion order
 * (core/scene.cpp:26-30). prim_id == SPT_MISS when nothing is hit, t = +inf. */
typedef struct spt_hit {
    float t;
    uint32_t prim_id;
} spt_hit;

/* One BVH node of the flattened tree: the boxes of BOTH children (an AABB pair)
 * and their references, 64 B, loaded by one lane as 4 x 16 B.
 * A child with cnt == 0 is the inner node nodes[ref]; with cnt > 0 it is a leaf
 * holding triangles tris[ref .. ref+cnt). Boxes are the reference's padded child
 * boxes (acceleration/sbvh.cpp:409-412). Node 0 holds the root's children; the
 * root's own (unpadded) box is kept separately. */
typedef struct spt_node64 {
    float lmin[3];
    float lmax[3];
    float rmin[3];
    float rmax[3];
    uint32_t lref, lcnt, rref, rcnt;
} spt_node64;

/* Primitive kinds (primitives/ directory of the reference). A primitive is
 * 9 floats: a triangle's v0 v1 v2, or a sphere's centre xyz and radius
 * (primitives/sphere.cpp) followed by 5 unused floats. */
enum { SPT_PRIM_TRIANGLE = 0, SPT_PRIM_SPHERE = 1 };

/* One primitive in leaf (depth-first, left-first) order, 48 B = 3 x 16 B.
 * Triangle: v0 and the edges e1 = v1 - v0, e2 = v2 - v0 exactly as
 * primitives/triangle.cpp:22-23 computes them. Sphere: v0 = centre,
 * e1[0] = radius, other fields 0. prim_id: the primitive's scene index;
 * pad[0]: material id (filled by spt_set_scene); pad[1]: SPT_PRIM_* kind. */
typedef struct spt_tri48 {
    float v0[3];
    float e1[3];
    float e2[3];
    uint32_t prim_id;
    uint32_t pad[2];
} spt_tri48;

/* BSDF layer kinds (the reference materials/ directory). GGX_CONDUCTOR has no reference
 * counterpart: Walter et al. 2007 GGX with an exact complex-IOR conductor
 * Fresnel term. */
enum {
    SPT_BSDF_NONE = 0,
    SPT_BSDF_LAMBERTIAN = 1,        /* materials/lambertian.cpp            */
    SPT_BSDF_COOK_TORRANCE = 2,     /* materials/cook_torrance.cpp (Beckmann) */
    SPT_BSDF_SMOOTH_CONDUCTOR = 3,  /* materials/smooth_conductor.cpp      */
    SPT_BSDF_SMOOTH_DIELECTRIC = 4, /* materials/smooth_dielectric.cpp     */
    SPT_BSDF_GGX_CONDUCTOR = 5
};
/* bsdf_sampler = "importance" | "uniform" (lua_core/bsdfs.lua:9). */
enum { SPT_SAMPLER_IMPORTANCE = 0, SPT_SAMPLER_UNIFORM = 1 };

typedef struct spt_layer {
    uint32_t type;     /* SPT_BSDF_* */
    uint32_t sampler;  /* SPT_SAMPLER_* */
    float kd[3];       /* Lambertian reflectance */
    float r0[3];       /* Schlick reflectance at normal incidence */
    float roughness;   /* Beckmann m (Cook-Torrance) or GGX alpha */
    float ior;         /* smooth dielectric IOR (outside medium: 1.0) */
    float eta[3];      /* GGX conductor: complex IOR, real part per RGB */
    float k[3];        /* GGX conductor: complex IOR, imaginary part per RGB */
} spt_layer;           /* 64 B */

#define SPT_MAX_LAYERS 4

/* Material = layered BSDF + emission (materials/material.h, layered_bsdf.h).
 * layers[0] is the top layer. */
typedef struct spt_material {
    uint32_t n_layers;
    float emission[3];
    spt_layer layers[SPT_MAX_LAYERS];
} spt_material;

enum { SPT_CAMERA_PINHOLE = 0, SPT_CAMERA_ORTHOGRAPHIC = 1 };

/* cameras/pinhole_camera.cpp, cameras/orthographic_camera.cpp, with the
 * image size of core/image_buffer.h. */
typedef struct spt_camera {
    uint32_t type;
    uint32_t width, height;
    uint32_t pad;
    float position[3];
    float look_at[3];
    float up[3];
    float fov_deg;                  /* pinhole: vertical field of view */
    float min_x, max_x, min_y, max_y; /* orthographic window */
} spt_camera;

/* pixel_samplers/ directory of the reference */
enum { SPT_PIXEL_REGULAR = 0, SPT_PIXEL_JITTERED = 1, SPT_PIXEL_UNIFORM = 2 };

/* Random numbers. The reference draws from per-thread mt19937 streams with
 * non-deterministic seeds (core/prng.h:12-51); here every draw is a pure
 * function of (seed, path, k), k = index of the draw within the path:
 *   mix(z)  = SplitMix64 finaliser (Steele, Lea, Flood 2014)
 *   key     = mix(seed ^ (path * 0x9E3779B97F4A7C15))
 *   u(k)    = (mix(key + (k + 1) * 0x9E3779B97F4A7C15) >> 40) * 2^-24   in [0, 1)
 * path = (pixel_y * width + pixel_x) * spp + sample. In replay mode the draws
 * instead come from a host stream u_stream[path_local * stride + k]. */
enum { SPT_RNG_COUNTER = 0, SPT_RNG_STREAM = 1 };

/* Flags of spt_render_params.flags */
enum {
    SPT_RENDER_FILM32 = 1u << 0,    /* accumulate FP32 film into d_rgb32 */
    SPT_RENDER_FILM64 = 1u << 1,    /* accumulate FP64 film into d_rgb64 (oracle film) */
    SPT_RENDER_PER_PATH = 1u << 2,  /* write each path's radiance into d_path_rgb */
    SPT_RENDER_NO_SORT = 1u << 3,   /* skip the material sort between stages */
    SPT_RENDER_GUIDED = 1u << 4,    /* sample with the learned guide (spt_guide_*) */
    SPT_RENDER_FETCH_STATS = 1u << 5, /* count tree data in 32-B units (a 64-B node is 2,
                                         a 32-B leaf record 1) and 48-B triangle records
                                         fetched by the extension rays (slower variant) */
    SPT_RENDER_TIME_KERNELS = 1u << 6 /* time every extend launch with HIP events */
};

/* Integrator settings. The reference has two modes (integrators/pathtracer.cpp:18,43):
 *   MAX_DEPTH:        max_depth = path_length, rr_depth = -1
 *   RUSSIAN_ROULETTE: max_depth = -1, rr_depth = path_length (1000-bounce cap kept)
 * Setting both gives RR from rr_depth with a hard cap at max_depth. */
typedef struct spt_render_params {
    uint32_t pixel_sampler; /* SPT_PIXEL_* */
    uint32_t spp;           /* regular/jittered round up to ceil(sqrt(spp))^2 */
    int32_t max_depth;      /* depth >= max_depth returns 0 before tracing; < 0: none */
    int32_t rr_depth;       /* Russian roulette when depth >= rr_depth; < 0: none */
    uint64_t seed;
    uint32_t spp_begin;     /* render samples [spp_begin, spp_begin + spp_count) */
    uint32_t spp_count;     /* 0 = all */
    uint32_t viewport[4];   /* top, left, width, height; width == 0: whole image */
    float background[3];    /* returned by a primary miss (core/scene.h:149) */
    uint32_t flags;         /* SPT_RENDER_* */
    uint32_t rng;           /* SPT_RNG_* */
    uint32_t stream_stride; /* SPT_RNG_STREAM: draws per path */
    const float* d_stream;  /* SPT_RNG_STREAM: device buffer, paths x stride */
} spt_render_params;

/* Statistics of the built tree: StatsSBVH (acceleration/sbvh.h:19-40) plus the
 * flattened-layout figures. */
typedef struct spt_bvh_info {
    uint64_t n_prims;
    uint64_t n_nodes;       /* reference nodes (inner + leaves) */
    uint64_t n_inner;
    uint64_t n_leaves;
    uint64_t tree_height;   /* levels - 1, as collectTreeStatistics counts */
    uint64_t n_references;
    uint64_t min_leaf, max_leaf;
    uint64_t n_pair_nodes;  /* spt_node64 records */
    uint64_t max_depth;     /* longest root-to-leaf path in inner nodes */
    float root_min[3];
    float root_max[3];
    double build_ms;
} spt_bvh_info;

typedef struct spt_bvh spt_bvh;
typedef struct spt_ctx spt_ctx;

/* ------------------------------------------------------------------------- */
/* Host: BVH build (replaces SBVH::build, acceleration/sbvh.cpp:12-37)        */
/* ------------------------------------------------------------------------- */

/* Builds the reference's object-split full-sweep SAH BVH over n_tris triangles
 * (verts: n_tris x 9 floats, v0 v1 v2) with the reference's sort sequence, SAH
 * constants (1.2 / 1.0), leaf rule and +-1e-3 child padding, so the tree is the
 * one the reference would build for the same primitive order. n_threads > 1
 * builds independent subtrees concurrently (same tree). */
spt_status spt_bvh_build(const float* verts, uint32_t n_tris, uint32_t n_threads, spt_bvh** out);
/* The same build over a scene of triangles and spheres in insertion order
 * (core/scene.cpp:26-30): prims is n x 9 floats (see SPT_PRIM_*), kinds[i]
 * the kind of primitive i (NULL: all triangles). A sphere's box is centre -+
 * radius and its centroid the centre (primitives/sphere.cpp:47-60). */
spt_status spt_bvh_build_prims(const float* prims, const uint8_t* kinds, uint32_t n, uint32_t n_threads,
                               spt_bvh** out);
spt_status spt_bvh_get_info(const spt_bvh* bvh, spt_bvh_info* info);
/* Copies out the flattened tree: info.n_pair_nodes nodes, n_prims triangles,
 * and leaf_prim[i] = prim id of the i-th triangle in leaf order. Any pointer may
 * be NULL. root_box = {min xyz, max xyz}. */
spt_status spt_bvh_export(const spt_bvh* bvh, spt_node64* nodes, spt_tri48* tris, uint32_t* leaf_prim,
                          float* root_box);
/* Writes the reference's tree dump text format (acceleration/sbvh.cpp:577-620). */
spt_status spt_bvh_dump(const spt_bvh* bvh, const char* path);
/* The four-wide compressed tree the throughput traversal reads, for
 * inspection (no reference counterpart). Node (64 B, 16 u32): grid origin
 * x y z (f32); biased exponent bytes of the per-axis grid step 2^e (bytes
 * 0-2); six words of child bytes q (lo.x hi.x lo.y hi.y lo.z hi.z, byte i =
 * child i; bound = origin + q * 2^e, lo rounded down, hi up); two words of
 * 16-bit child triangle counts (0: inner or unused, 0xFFFF: see the box
 * record); four children (node index; first triangle in leaf order for a
 * leaf; 0xFFFFFFFF unused). Leaf boxes (32 B, four per node, slot i of node
 * j at 4 j + i): exact box min xyz (f32), full triangle count, box max xyz,
 * unused. NULL arrays: the counts are returned; otherwise *n_nodes and
 * *n_leaves give the arrays' capacities on entry. */
spt_status spt_bvh_wide(const spt_bvh* bvh, void* nodes, uint64_t* n_nodes, void* leaves, uint64_t* n_leaves);
void spt_bvh_destroy(spt_bvh* bvh);

/* ------------------------------------------------------------------------- */
/* Device context                                                            */
/* ------------------------------------------------------------------------- */

spt_status spt_create(int device, spt_ctx** out);
void spt_destroy(spt_ctx* ctx);
const char* spt_last_error(const spt_ctx* ctx);
spt_status spt_sync(spt_ctx* ctx);
/* The context's hipStream_t, for callers that time or order work around it. */
void* spt_stream(spt_ctx* ctx);

spt_status spt_malloc(spt_ctx* ctx, size_t bytes, void** d_ptr);
spt_status spt_free(spt_ctx* ctx, void* d_ptr);
spt_status spt_memcpy_h2d(spt_ctx* ctx, void* d_dst, const void* src, size_t bytes);
spt_status spt_memcpy_d2h(spt_ctx* ctx, void* dst, const void* d_src, size_t bytes);
spt_status spt_memcpy_d2d(spt_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* Uploads a built tree (copied). verts: the primitives the tree was built
 * over (n_prims x 9 floats, as spt_bvh_build / spt_bvh_build_prims took them).
 * normals: n_prims x 9 floats (n0 n1 n2, smooth triangle,
 * primitives/smooth_triangle.h:11-13) or NULL for flat triangles (normal =
 * normalize(cross(e1, e2)), primitives/triangle.cpp:5-6); a row of nine zeros
 * marks one flat triangle among smooth ones; a sphere's normal is always
 * normalize(hit - centre) (primitives/sphere.cpp:35). mat_ids: n_prims
 * material indices or NULL (all 0). */
spt_status spt_set_scene(spt_ctx* ctx, const spt_bvh* bvh, const float* verts, const float* normals,
                         const uint32_t* mat_ids);
/* Same, from the flattened arrays (as produced by spt_bvh_export). shading is
 * n_tris x 12 floats in leaf order: n0 n1 n2 (9), material id (as u32), normal
 * kind (as u32: 0 flat n0, 1 smooth from n0 n1 n2, 2 sphere with centre n0),
 * pad. */
spt_status spt_upload_scene(spt_ctx* ctx, const spt_node64* nodes, uint32_t n_nodes, const spt_tri48* tris,
                            uint32_t n_tris, const float* root_box, const float* shading);
spt_status spt_set_materials(spt_ctx* ctx, const spt_material* mats, uint32_t n_mats);

/* ------------------------------------------------------------------------- */
/* Closest-hit ray casting (replaces AccelStructure::intersect, batched)      */
/* ------------------------------------------------------------------------- */

enum {
    SPT_TRACE_CLOSEST = 0,   /* pruned near-first traversal; same closest hit */
    SPT_TRACE_COUNT = 1,     /* visits the reference's exact node set and also
                                returns its counters (sbvh.cpp:478-510) */
    SPT_TRACE_UNPRUNED = 2,  /* like COUNT without writing counters */
    SPT_TRACE_FETCH = 3      /* CLOSEST, and d_tests receives the tree data each ray
                                fetched in 32-B units (a 64-B node is 2, a 32-B
                                leaf record 1), d_isects its 48-B triangle records */
};

/* d_rays: n rays; d_hits: n hits. In SPT_TRACE_COUNT mode d_tests/d_isects
 * (n each, may be NULL) receive num_intersection_tests / num_intersections. */
spt_status spt_trace(spt_ctx* ctx, const spt_ray32* d_rays, uint64_t n, spt_hit* d_hits, uint32_t mode,
                     uint32_t* d_tests, uint32_t* d_isects);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): spt_trace in SPT_TRACE_CLOSEST mode with no allocation
 * at all. Its scratch (ray counter, overflow flag, traversal-stack spill) is
 * a static device array of SPT_NOALLOC_SCRATCH_WORDS 32-bit words (32 MiB)
 * per GPU, held until the process ends; the persistent grid is capped to what
 * fits in it. Rejects n > SPT_NOALLOC_MAX_RAYS (SPT_ERR_INVALID) and trees
 * whose stack does not fit one block (SPT_ERR_UNSUPPORTED). Calls that overlap
 * on one GPU (several contexts or threads) race on the shared scratch and must
 * be ordered by the caller. A stack overflow is reported by spt_sync. */
#define SPT_NOALLOC_SCRATCH_WORDS (8u * 1024u * 1024u)
#define SPT_NOALLOC_MAX_RAYS 0xFFFFFF00ull
spt_status spt_trace_noalloc(spt_ctx* ctx, const spt_ray32* d_rays, uint64_t n, spt_hit* d_hits);

/* Primary rays: d_rays[y * w + x] = camera ray through image point
 * (x + sx, y + sy) with (sx, sy) = (0.5, 0.5), the single sample of
 * RegularPixelSampler(1) that the ray casters use (integrators/raycaster.h:45). */
spt_status spt_camera_rays(spt_ctx* ctx, const spt_camera* cam, spt_ray32* d_rays);

/* Ray casters (integrators/raycaster.h:61-99): one ray per pixel centre.
 * kind 0 normal (d_out: w*h*3 floats, 0 on miss), 1 depth (w*h floats, t or -1),
 * 2 intersection-test count (w*h uint32). Pixel (x, y) is at index y*w + x. */
enum { SPT_RAYCAST_NORMAL = 0, SPT_RAYCAST_DEPTH = 1, SPT_RAYCAST_COUNT = 2 };
spt_status spt_raycast(spt_ctx* ctx, const spt_camera* cam, uint32_t kind, void* d_out);

/* ------------------------------------------------------------------------- */
/* Path tracing (replaces PathTracer::render, integrators/pathtracer.cpp)     */
/* ------------------------------------------------------------------------- */

/* Renders samples [spp_begin, spp_begin+spp_count) of every viewport pixel.
 * d_rgb32 (w*h*3 floats) / d_rgb64 (w*h*3 doubles) accumulate radiance SUMS
 * (divide by the sample count to get the reference's pixel value). d_path_rgb
 * receives 3 floats per path, path index = (pixel_in_viewport * spp_count +
 * sample - spp_begin). Unused outputs may be NULL. */
spt_status spt_render(spt_ctx* ctx, const spt_camera* cam, const spt_render_params* params, float* d_rgb32,
                      double* d_rgb64, float* d_path_rgb);

/* Counters of the last spt_render call: paths, extension rays traced, bounce
 * iterations, host wall ms; with SPT_RENDER_TIME_KERNELS the summed device time
 * and count of the extend (closest-hit) launches; with SPT_RENDER_FETCH_STATS the
 * tree data (node_fetches, in 32-B units) and triangle records those rays fetched. */
typedef struct spt_render_stats {
    uint64_t paths;
    uint64_t rays;
    uint64_t bounces;
    uint64_t stream_overflows;
    double ms;
    double extend_ms;
    uint64_t extend_launches;
    uint64_t node_fetches;
    uint64_t tri_fetches;
    double guide_ms;       /* with SPT_RENDER_TIME_KERNELS and GUIDED: guide inference time */
    uint64_t guide_queries; /* GUIDED with the material sort: hits the guide network evaluated */
} spt_render_stats;
spt_status spt_render_get_stats(spt_ctx* ctx, spt_render_stats* out);

/* HIP-event timing on the context stream: spt_event_record(ctx, slot) with
 * slot in [0, 8); spt_event_elapsed() waits for slot b and returns ms a -> b. */
spt_status spt_event_record(spt_ctx* ctx, int slot);
spt_status spt_event_elapsed(spt_ctx* ctx, int a, int b, float* ms);

/* ------------------------------------------------------------------------- */
/* Launch tuning and the autotuner (no reference counterpart)                */
/* ------------------------------------------------------------------------- */

/* Queue and tile sizes of the wavefront pipeline. A zero field means the
 * default. Results never depend on these, only speed. */
typedef struct spt_tuning {
    uint32_t path_capacity;        /* paths in flight (queue size, at most 2^27; default 134,217,728:
                                      about 250 B of device memory per slot, 410 B when guided;
                                      a render allocates min(paths of the call, capacity), halved
                                      while the device has no room, down to 2^20) */
    uint32_t extend_blocks_per_cu; /* persistent traversal grid (default: the occupancy limit on
                                      one wavefront, at most 4 each for two, 2 for more) */
    uint32_t shade_blocks_per_cu;  /* shading grid, at most 1024 (default 128 on one wavefront,
                                      64 with several) */
    uint32_t use_graph;            /* 1: bounce loop captured once in a HIP graph and replayed;
                                      2: plain launches (default) */
    uint32_t tile_pixels;          /* multi-scene calls: pixels per generation tile (default 4096) */
    uint32_t tile_samples;         /* multi-scene calls: samples per tile (default 16) */
    uint32_t check_every;          /* bounces the host may run ahead of the device (default 4) */
    uint32_t refill_lanes;         /* a traversal wave fetches new rays once this many of
                                      its 64 lanes are idle (default in path tracing 48 on one
                                      wavefront, 20 with several; 16 in spt_trace) */
    uint32_t wavefronts;           /* unguided renders: independent wavefronts run at once on
                                      their own streams, each with a share of the paths and of
                                      the queue capacity, so one's traversal overlaps another's
                                      shading (1..4, default 3) */
} spt_tuning;
spt_status spt_set_tuning(spt_ctx* ctx, const spt_tuning* t);
spt_status spt_get_tuning(spt_ctx* ctx, spt_tuning* t);
/* Times candidate tunings on a sample of the given render (about sample_paths
 * paths per trial; outputs discarded), keeps the fastest in the context and
 * returns it. Coordinate descent over capacity, grids, graph and check
 * interval. */
spt_status spt_autotune(spt_ctx* ctx, const spt_camera* cam, const spt_render_params* params, uint64_t sample_paths,
                        spt_tuning* out);

/* ------------------------------------------------------------------------- */
/* Batched rendering of many scenes (no reference counterpart)               */
/* ------------------------------------------------------------------------- */

/* Scenes of any size are concatenated in one set of device arrays (tree
 * indices and material ids rebased per scene) and rendered by one wavefront:
 * every bounce is one launch of each stage over the live paths of all scenes,
 * the traversal grid is persistent, and paths are generated in tiles taken
 * round-robin from the scenes. Per path the result equals a single-scene
 * spt_render of the same scene with the same parameters. GUIDED and replay
 * streams are per-scene options the batch does not take; max_depth must not be 0. */
typedef struct spt_batch_scene {
    const spt_bvh* bvh;
    const float* verts;      /* n_prims x 9 (as spt_set_scene) */
    const float* normals;    /* or NULL */
    const uint32_t* mat_ids; /* or NULL */
    const spt_material* mats;
    uint32_t n_mats;
    uint32_t pad;
    spt_camera camera;
    spt_render_params params; /* FILM32 / FILM64 / PER_PATH flags select the outputs below */
    float* d_rgb32;
    double* d_rgb64;
    float* d_path_rgb;
} spt_batch_scene;

typedef struct spt_batch spt_batch;

typedef struct spt_batch_stats {
    uint64_t scenes;
    uint64_t paths;
    uint64_t rays;
    uint64_t bounces;
    double ms;                /* host wall time of the render call */
    uint64_t tiles;
    uint64_t triangles;
    uint64_t device_bytes;    /* scene arrays of the batch */
} spt_batch_stats;

spt_status spt_batch_create(spt_ctx* ctx, const spt_batch_scene* scenes, uint32_t n, spt_batch** out);
/* Renders every scene of the batch (tuning: the context's). */
spt_status spt_batch_render(spt_ctx* ctx, spt_batch* batch);
/* scene_ms (n scenes, may be NULL): time from the start of the last
 * spt_batch_render until each scene's last path finished (device clock). */
spt_status spt_batch_get_stats(spt_ctx* ctx, const spt_batch* batch, spt_batch_stats* out, double* scene_ms);
/* As spt_autotune, on a sample of every scene's samples (1/sample_div of each). */
spt_status spt_batch_autotune(spt_ctx* ctx, spt_batch* batch, uint32_t sample_div, spt_tuning* out);
void spt_batch_destroy(spt_batch* batch);

/* ------------------------------------------------------------------------- */
/* Learned path guide (no reference counterpart)                             */
/* ------------------------------------------------------------------------- */

/* A small MLP maps the encoded hit (position in the scene box, shading normal,
 * outgoing direction; 32 features) through three 64-unit ReLU layers to an
 * 8-lobe von Mises-Fisher mixture (40 outputs: per lobe mu_xyz, log kappa,
 * logit). BF16 weights with FP32 accumulation (MFMA), FP32 Adam, loss scaling.
 * Rendering with SPT_RENDER_GUIDED samples single-layer Lambertian and GGX
 * vertices from the one-sample mixture bsdf_fraction * BSDF + (1 -
 * bsdf_fraction) * guide, so the estimator stays unbiased; other vertices use
 * the reference's BSDF sampling. SPT_RENDER_GUIDE_TRAIN additionally records
 * training samples (position, normal, wo, wi, weight = incident radiance x
 * f cos / pdf, luminance) for spt_guide_train_collected(). */
#define SPT_GUIDE_PARAMS 13552
#define SPT_GUIDE_RECORD_FLOATS 24
enum { SPT_RENDER_GUIDE_TRAIN = 1u << 7 };

typedef struct spt_guide_config {
    uint64_t seed;         /* parameter initialisation */
    float learning_rate;   /* Adam step (default 5e-3) */
    float loss_scale;      /* initial loss scale (default 65536; halved on overflow) */
    float bsdf_fraction;   /* one-sample mixture weight of BSDF sampling (default 0.5) */
    uint32_t batch;        /* training records per step (default 262144) */
    uint32_t steps_per_pass; /* a training render records up to batch x steps_per_pass
                                records; spt_guide_train_collected takes one Adam step
                                per batch of them (default 1) */
    uint32_t flags;        /* SPT_GUIDE_* options (default 0) */
    uint32_t train_every;  /* > 0: a GUIDE_TRAIN render also trains inside the call, one
                              round every train_every bounces (Adam steps on the records of
                              paths recorded at least one round earlier; the record buffer
                              is a ring of batch x steps_per_pass records); 0: records are
                              kept for spt_guide_train_collected (default 0) */
} spt_guide_config;

/* Guide options. COATED: also guide a Lambertian / GGX layer under one smooth
 * dielectric coat (the first visit of the base layer in the layered walk
 * draws from the mixture, the guide's exit direction mapped under the coat).
 * PRODUCT: learn incident radiance (training weight L_i / p) and propose
 * directions from its product with a vMF stand-in of the BSDF lobe, instead
 * of learning the product f L_i cos directly. */
#define SPT_GUIDE_COATED (1u << 0)
#define SPT_GUIDE_PRODUCT (1u << 1)

typedef struct spt_guide_stats {
    uint64_t steps;
    uint64_t skipped_steps;
    uint64_t records;      /* records collected since the last training step */
    float last_loss;
    float loss_scale;
    double train_ms;       /* host wall time spent in training steps */
    double train_kernel_ms;/* device time of the fused training kernel (HIP events) */
    uint64_t trained_records; /* records consumed by successful steps */
    uint64_t candidates;   /* guided vertices of the last training render */
    uint32_t record_stride;/* 1 in record_stride candidates is recorded (set from the
                              previous render so a pass fills about one batch) */
    uint32_t pad;
} spt_guide_stats;

/* (Re)initialises the guide (cfg NULL: defaults). */
spt_status spt_guide_init(spt_ctx* ctx, const spt_guide_config* cfg);
/* Host copies of the FP32 master parameters (SPT_GUIDE_PARAMS floats):
 * W1 64x32 | b1 64 | W2 64x64 | b2 64 | W3 64x64 | b3 64 | W4 48x64 | b4 48
 * (row-major out x in; rows 40..47 of W4 and b4 are padding, kept zero).
 * Setting parameters resets the Adam moments. */
spt_status spt_guide_set_params(spt_ctx* ctx, const float* theta);
spt_status spt_guide_get_params(spt_ctx* ctx, float* theta);
/* Raw network outputs (n x 40) for queries d_queries (n x 9: position,
 * normal, outgoing direction; world space). Needs an uploaded scene (its box
 * normalises positions). */
spt_status spt_guide_eval(spt_ctx* ctx, const float* d_queries, uint64_t n, float* d_out);
/* One Adam step on n records (loss may be NULL: no wait) (n x SPT_GUIDE_RECORD_FLOATS: position 0-2,
 * normal 3-5, wo 6-8, wi 9-11, weight 12). loss may be NULL. */
spt_status spt_guide_train(spt_ctx* ctx, const float* d_records, uint32_t n, float* loss);
/* The (unscaled) gradient the step would apply, without applying it (host
 * grad: SPT_GUIDE_PARAMS floats). */
spt_status spt_guide_gradient(spt_ctx* ctx, const float* d_records, uint32_t n, float* grad, float* loss);
/* Adam steps on the records collected by the last SPT_RENDER_GUIDE_TRAIN
 * render: one per `batch` records (up to steps_per_pass). Steps are queued on
 * the context stream; with loss != NULL the call waits and returns the loss of
 * the last step, with NULL it returns at once. Skips and loss-scale changes
 * are decided on the device (spt_guide_get_stats reports them). */
spt_status spt_guide_train_collected(spt_ctx* ctx, float* loss);
spt_status spt_guide_get_stats(spt_ctx* ctx, spt_guide_stats* out);
/* Device pointer and count of the collected records. */
spt_status spt_guide_records(spt_ctx* ctx, float** d_records, uint32_t* n);

/* ------------------------------------------------------------------------- */
/* Host: Lua scene scripts (replaces lb::LuaBind, lua_bind/lua_bind.cpp, and  */
/* the scene assembly of SwpathtracerApp, core/swpathtracer_app.cpp:40-262)   */
/* ------------------------------------------------------------------------- */

/* What the script's Integrator{} selects (lua_core/integrators.lua). */
enum {
    SPT_LUA_NONE = 0,
    SPT_LUA_PATHTRACER = 1,     /* type = "path-tracer"                            */
    SPT_LUA_RAYCAST_NORMAL = 2, /* type = "ray-caster", pixel_value = "normal"     */
    SPT_LUA_RAYCAST_DEPTH = 3,  /* ... "depth" (min_value/max_value: minimum/maximum_depth) */
    SPT_LUA_RAYCAST_COUNT = 4   /* ... "intersection-test-count" (minimum/maximum_count)   */
};
/* GlobalSettings{acceleration_structure} (lua_bind/lua_bind.cpp:122-155). */
enum { SPT_ACCEL_NONE = 0, SPT_ACCEL_BVH_SAH = 1, SPT_ACCEL_SBVH_SAH = 2 };
/* spt_lua_load flags: the scene functions of the lua_core scripts come from the
 * library's built-in restatement even when package.path would find the files. */
#define SPT_LUA_BUILTIN_CORE 1u

typedef struct spt_lua_info {
    spt_camera camera;         /* Camera{} with ImageBuffer{}'s width and height */
    spt_render_params params;  /* PixelSampler{}, path termination, viewport, background;
                                  FILM32, counter RNG, seed 0 */
    uint32_t integrator;       /* SPT_LUA_* */
    uint32_t accel;            /* SPT_ACCEL_* */
    float overlap_threshold;   /* inf for bvh-sah (lua_core/global_settings.lua:14-16) */
    float min_value, max_value;/* depth / count ray caster ranges (-1: from the image) */
    int32_t path_length;       /* Integrator{path_length} */
    uint64_t n_prims;          /* primitives after mesh expansion, scene order */
    uint64_t n_materials;
    uint64_t n_objects;        /* Triangle/SmoothTriangle/Sphere/Mesh statements */
    uint64_t n_bsdfs;
    char file_name[256];       /* ImageBuffer{file_name} */
} spt_lua_info;

typedef struct spt_lua_scene spt_lua_scene;

/* Runs a scene script as lb::LuaBind::runScript does (lua_bind.cpp:18-31) on
 * the library's own interpreter of the Lua 5.3 subset scene scripts use, then
 * extracts the scene as lua_bind's getters and SwpathtracerApp's loaders do.
 * Relative paths (the script, package.path entries, Mesh{filename}) resolve
 * against workdir (NULL: the process's working directory), the directory the
 * reference's app runs from. `require "core"` loads the lua_core scripts through
 * package.path; when no file is found (or with SPT_LUA_BUILTIN_CORE) the
 * library's restatement of those functions is used: same defaults,
 * validation messages and registry layout. math.random() is Lua 5.3's on
 * glibc (random() / 2^31, unseeded = srandom(1)). A validation failure (the
 * reference prints "Lua ERROR: ..." and exits) or a lua_bind error (the
 * reference throws) returns SPT_ERR_INVALID with the message in err. */
spt_status spt_lua_load(const char* script, const char* workdir, uint32_t flags, spt_lua_scene** out, char* err,
                        size_t err_len);
spt_status spt_lua_get_info(const spt_lua_scene* sc, spt_lua_info* info);
/* prims/normals: n_prims x 9 floats; kinds: n_prims SPT_PRIM_* (spt_bvh_build_prims,
 * spt_set_scene); mat_ids: n_prims. Any pointer may be NULL. */
spt_status spt_lua_get_prims(const spt_lua_scene* sc, float* prims, uint8_t* kinds, float* normals,
                             uint32_t* mat_ids);
/* n_materials records for spt_set_materials: Material{bsdf = LayeredBSDF{...}}
 * layers top first, with each BSDF's sampler as the app picks it. */
spt_status spt_lua_get_materials(const spt_lua_scene* sc, spt_material* mats);
/* The registry (_G.__RENDERER_ELEMENTS__) as JSON text (keys sorted, array
 * entries as "#000000000001"), and what the script printed. *len: capacity
 * on entry (buf may be NULL to query), length + 1 on return. */
spt_status spt_lua_registry_json(const spt_lua_scene* sc, char* buf, uint64_t* len);
spt_status spt_lua_printed(const spt_lua_scene* sc, char* buf, uint64_t* len);
void spt_lua_destroy(spt_lua_scene* sc);

#ifdef __cplusplus
}
#endif

#endif /* SPT_H */
