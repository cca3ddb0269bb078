// spt_batch.hip - many scenes in one wavefront (spt_batch_*), launch tuning
// and the autotuner (spt_set_tuning, spt_autotune, spt_batch_autotune).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>

#include "spt_device.h"

struct spt_batch {
    spt::BatchData d;
};

namespace spt {
namespace {

void free_batch(BatchData& b) {
    void* ptrs[] = {b.d_nodes, b.d_nodesq, b.d_leaves,     b.d_tris, b.d_shading, b.d_mats,
                    b.d_jobs,  b.d_tiles,  b.d_tile_start, b.d_done, b.d_clock};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    b = BatchData{};
}

spt_status check_materials(spt_ctx* ctx, const spt_material* mats, uint32_t n) {
    if (!mats || n == 0) return set_error(ctx, SPT_ERR_INVALID, "batch scene without materials");
    for (uint32_t i = 0; i < n; ++i) {
        if (mats[i].n_layers == 0 || mats[i].n_layers > SPT_MAX_LAYERS)
            return set_error(ctx, SPT_ERR_INVALID, "material layer count must be 1..4");
        for (uint32_t l = 0; l < mats[i].n_layers; ++l)
            if (mats[i].layers[l].type < SPT_BSDF_LAMBERTIAN || mats[i].layers[l].type > SPT_BSDF_GGX_CONDUCTOR)
                return set_error(ctx, SPT_ERR_INVALID, "unknown BSDF layer type");
    }
    return SPT_OK;
}

template <typename T>
spt_status upload_vec(spt_ctx* ctx, T** dst, const std::vector<T>& v, const char* what, uint64_t& bytes) {
    const size_t n = std::max<size_t>(v.size() * sizeof(T), 16);
    spt_status s = check_hip(ctx, hipMalloc(reinterpret_cast<void**>(dst), n), what);
    if (s == SPT_OK && !v.empty())
        s = check_hip(ctx, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), what);
    bytes += n;
    return s;
}

// Wall time of one trial (device synchronised on both sides).
template <typename F>
double time_ms(spt_ctx* ctx, F run, spt_status& s) {
    (void)hipStreamSynchronize(ctx->stream);
    auto t0 = std::chrono::steady_clock::now();
    s = run();
    (void)hipStreamSynchronize(ctx->stream);
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Coordinate descent over the tuning fields: each field in turn takes the
// candidate value with the lowest time (best of two runs) with the others
// held at their current best.
template <typename F>
spt_status tune(spt_ctx* ctx, bool batch, F run, spt_tuning* out) {
    const spt_tuning saved = ctx->tuning;
    spt_tuning best = resolve_tuning(ctx->tuning);
    spt_status s = SPT_OK;
    auto trial = [&](const spt_tuning& t) {
        ctx->tuning = t;
        double ms = 1e300;
        for (int r = 0; r < 2 && s == SPT_OK; ++r) ms = std::min(ms, time_ms(ctx, run, s));
        return ms;
    };
    double best_ms = trial(best);  // also warms up allocations
    best_ms = std::min(best_ms, trial(best));
    struct Dim {
        uint32_t spt_tuning::*field;
        std::vector<uint32_t> values;
    };
    std::vector<Dim> dims = {
        {&spt_tuning::path_capacity, {1u << 23, 1u << 24, 1u << 25, 1u << 26, 1u << 27}},
        {&spt_tuning::extend_blocks_per_cu, {0u, 6u, 4u, 3u, 2u}},
        {&spt_tuning::shade_blocks_per_cu, {16u, 32u, 64u, 128u}},
        {&spt_tuning::use_graph, {2u, 1u}},
        {&spt_tuning::check_every, {2u, 4u, 8u}},
        {&spt_tuning::refill_lanes, {16u, 32u, 48u, 56u}},
        {&spt_tuning::wavefronts, {3u, 2u, 1u}},
    };
    if (batch) {
        dims.push_back({&spt_tuning::tile_pixels, {1024u, 4096u, 16384u, 65536u}});
        dims.push_back({&spt_tuning::tile_samples, {4u, 16u, 64u}});
    }
    for (const Dim& d : dims) {
        for (uint32_t v : d.values) {
            if (best.*(d.field) == v) continue;
            spt_tuning t = best;
            t.*(d.field) = v;
            const double ms = trial(t);
            if (s != SPT_OK) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
        if (s != SPT_OK) break;
    }
    if (s != SPT_OK) {
        ctx->tuning = saved;
        return s;
    }
    ctx->tuning = best;
    if (out) *out = best;
    return SPT_OK;
}

}  // namespace
}  // namespace spt

using namespace spt;

extern "C" {

spt_status spt_set_tuning(spt_ctx* ctx, const spt_tuning* t) {
    if (!ctx || !t) return SPT_ERR_INVALID;
    if (t->path_capacity > (1u << 27) || t->use_graph > 2 || t->shade_blocks_per_cu > 1024 ||
        t->extend_blocks_per_cu > 64 || t->check_every > 64 || t->refill_lanes > 64 || t->wavefronts > 4)
        return set_error(ctx, SPT_ERR_INVALID, "tuning value out of range");
    ctx->tuning = *t;
    return SPT_OK;
}

spt_status spt_get_tuning(spt_ctx* ctx, spt_tuning* t) {
    if (!ctx || !t) return SPT_ERR_INVALID;
    *t = resolve_tuning(ctx->tuning);
    return SPT_OK;
}

spt_status spt_autotune(spt_ctx* ctx, const spt_camera* cam, const spt_render_params* params, uint64_t sample_paths,
                        spt_tuning* out) {
    if (!ctx || !cam || !params) return SPT_ERR_INVALID;
    if (!ctx->has_scene || ctx->n_mats == 0) return set_error(ctx, SPT_ERR_STATE, "upload a scene and materials first");
    // trials render into a scratch film, never into the caller's outputs
    spt_render_params p = *params;
    p.flags = (p.flags & SPT_RENDER_NO_SORT) | SPT_RENDER_FILM32;
    if (cam->width == 0 || cam->height == 0) return SPT_ERR_INVALID;
    float* d_film = nullptr;
    spt_status s = check_hip(ctx, hipMalloc(&d_film, (size_t)cam->width * cam->height * 3 * sizeof(float)),
                             "tuning film");
    if (s != SPT_OK) return s;
    Job jb;
    uint64_t total = 0;
    s = make_job(ctx, cam, &p, d_film, nullptr, nullptr, jb, total);
    if (s != SPT_OK) {
        (void)hipFree(d_film);
        return s;
    }
    const uint64_t npix = (uint64_t)jb.vw * jb.vh;
    const uint64_t want = sample_paths ? sample_paths : (8ull << 20);
    p.spp_begin = jb.spp_begin;
    p.spp_count = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(jb.spp_count, want / std::max<uint64_t>(npix, 1)));
    s = tune(ctx, false, [&]() { return launch_render(ctx, cam, &p, d_film, nullptr, nullptr); }, out);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_film);
    return s;
}

spt_status spt_batch_create(spt_ctx* ctx, const spt_batch_scene* scenes, uint32_t n, spt_batch** out) {
    if (!ctx || !scenes || n == 0 || !out) return SPT_ERR_INVALID;
    *out = nullptr;
    std::vector<spt_node64> nodes;
    std::vector<NodeQ> nodesq;
    std::vector<LeafRec> leaves;
    std::vector<spt_tri48> tris;
    std::vector<float> sh;
    std::vector<spt_material> mats;
    spt_batch* b = new (std::nothrow) spt_batch();
    if (!b) return SPT_ERR_NOMEM;
    BatchData& d = b->d;
    auto fail = [&](spt_status s) {
        free_batch(d);
        delete b;
        return s;
    };
    for (uint32_t i = 0; i < n; ++i) {
        const spt_batch_scene& sc = scenes[i];
        if (!sc.bvh || !sc.verts) return fail(set_error(ctx, SPT_ERR_INVALID, "batch scene without geometry"));
        spt_status s = check_materials(ctx, sc.mats, sc.n_mats);
        if (s != SPT_OK) return fail(s);
        if (sc.params.flags & (SPT_RENDER_GUIDED | SPT_RENDER_GUIDE_TRAIN | SPT_RENDER_FETCH_STATS |
                               SPT_RENDER_TIME_KERNELS))
            return fail(set_error(ctx, SPT_ERR_UNSUPPORTED, "batch scenes take only film / per-path flags"));
        if (sc.params.max_depth == 0) return fail(set_error(ctx, SPT_ERR_INVALID, "batch scene with max_depth 0"));
        if (sc.camera.type > SPT_CAMERA_ORTHOGRAPHIC || sc.camera.width == 0 || sc.camera.height == 0)
            return fail(set_error(ctx, SPT_ERR_INVALID, "batch scene camera"));
        Job jb;
        uint64_t total = 0;
        s = make_job(ctx, &sc.camera, &sc.params, sc.d_rgb32, sc.d_rgb64, sc.d_path_rgb, jb, total);
        if (s != SPT_OK) return fail(s);
        const std::vector<spt_node64>& bn = bvh_nodes(sc.bvh);
        std::vector<spt_tri48> bt;
        std::vector<float> bs;
        scene_arrays(sc.bvh, sc.verts, sc.normals, sc.mat_ids, bt, bs);
        const uint64_t node_base = nodes.size(), tri_base = tris.size(), mat_base = mats.size();
        if (node_base + bn.size() >= (1u << 30) || tri_base + bt.size() >= (1ull << 32) || mat_base + sc.n_mats >= 65536)
            return fail(set_error(ctx, SPT_ERR_UNSUPPORTED, "batch exceeds 2^30 nodes / 2^32 triangles"));
        for (const spt_tri48& t : bt)
            if (t.pad[0] >= sc.n_mats) return fail(set_error(ctx, SPT_ERR_INVALID, "a triangle references a missing material"));
        // rebase child references, triangle materials and shading materials
        for (spt_node64 nd : bn) {
            nd.lref += (nd.lcnt == 0) ? (uint32_t)node_base : (uint32_t)tri_base;
            nd.rref += (nd.rcnt == 0) ? (uint32_t)node_base : (uint32_t)tri_base;
            nodes.push_back(nd);
        }
        std::vector<Node4> b4;
        d.stack_cap4 = std::max(d.stack_cap4, collapse4(bn.data(), (uint32_t)bn.size(), b4));
        const uint64_t node4_base = nodesq.size();
        if (!compress4(b4, (uint32_t)tri_base, nodesq, leaves))
            return fail(set_error(ctx, SPT_ERR_UNSUPPORTED, "batch exceeds 2^31 four-wide nodes or leaves"));
        for (spt_tri48 t : bt) {
            t.pad[0] += (uint32_t)mat_base;
            d.has_spheres = d.has_spheres || t.pad[1] == SPT_PRIM_SPHERE;
            tris.push_back(t);
        }
        for (size_t k = 0; k < bt.size(); ++k) {
            float* o = &bs[k * 12];
            uint32_t m;
            std::memcpy(&m, &o[9], 4);
            m += (uint32_t)mat_base;
            std::memcpy(&o[9], &m, 4);
        }
        sh.insert(sh.end(), bs.begin(), bs.end());
        mats.insert(mats.end(), sc.mats, sc.mats + sc.n_mats);
        const float* rb = bvh_root_box(sc.bvh);
        for (int k = 0; k < 3; ++k) {
            jb.rmin[k] = rb[k];
            jb.rmax[k] = rb[3 + k];
        }
        jb.root_leaf = bn.empty() ? (uint32_t)bt.size() : 0u;
        jb.node0 = (uint32_t)node4_base;  // persistent traversal starts at the scene's NodeQ root
        jb.tri0 = (uint32_t)tri_base;
        d.stack_cap = std::max(d.stack_cap, tree_stack_cap(bn.data(), (uint32_t)bn.size()));
        d.jobs.push_back(jb);
    }
    (void)hipSetDevice(ctx->device);
    uint64_t bytes = 0;
    spt_status s = upload_vec(ctx, reinterpret_cast<spt_node64**>(&d.d_nodes), nodes, "batch nodes", bytes);
    if (s == SPT_OK) s = upload_vec(ctx, &d.d_nodesq, nodesq, "batch compressed nodes", bytes);
    if (s == SPT_OK) s = upload_vec(ctx, &d.d_leaves, leaves, "batch leaf records", bytes);
    if (s == SPT_OK) s = upload_vec(ctx, reinterpret_cast<spt_tri48**>(&d.d_tris), tris, "batch triangles", bytes);
    if (s == SPT_OK) s = upload_vec(ctx, reinterpret_cast<float**>(&d.d_shading), sh, "batch shading", bytes);
    if (s == SPT_OK) s = upload_vec(ctx, &d.d_mats, mats, "batch materials", bytes);
    if (s == SPT_OK) s = upload_vec(ctx, &d.d_jobs, d.jobs, "batch jobs", bytes);
    if (s == SPT_OK) s = check_hip(ctx, hipMalloc(&d.d_done, n * sizeof(uint32_t)), "batch counters");
    if (s == SPT_OK) s = check_hip(ctx, hipMalloc(&d.d_clock, (n + 1) * sizeof(uint64_t)), "batch clocks");
    if (s != SPT_OK) return fail(s);
    d.n_nodes = (uint32_t)nodes.size();
    d.n_tris = (uint32_t)tris.size();
    d.n_mats = (uint32_t)mats.size();
    d.device_bytes = bytes;
    *out = b;
    return SPT_OK;
}

spt_status spt_batch_render(spt_ctx* ctx, spt_batch* batch) {
    if (!ctx || !batch) return SPT_ERR_INVALID;
    return launch_batch(ctx, batch->d, 1);
}

spt_status spt_batch_get_stats(spt_ctx* ctx, const spt_batch* batch, spt_batch_stats* out, double* scene_ms) {
    if (!ctx || !batch || !out) return SPT_ERR_INVALID;
    *out = batch->d.stats;
    if (scene_ms)
        for (size_t i = 0; i < batch->d.jobs.size(); ++i)
            scene_ms[i] = i < batch->d.scene_ms.size() ? batch->d.scene_ms[i] : 0.0;
    return SPT_OK;
}

spt_status spt_batch_autotune(spt_ctx* ctx, spt_batch* batch, uint32_t sample_div, spt_tuning* out) {
    if (!ctx || !batch) return SPT_ERR_INVALID;
    const uint32_t div = sample_div ? sample_div : 16;
    return tune(ctx, true, [&]() { return launch_batch(ctx, batch->d, div); }, out);
}

void spt_batch_destroy(spt_batch* batch) {
    if (!batch) return;
    free_batch(batch->d);
    delete batch;
}

}  // extern "C"
