// spt_render.hip - wavefront path tracer (replaces PathTracer::render and the
// recursive estimator of integrators/pathtracer.cpp:12-147).
//
// Pipeline per bounce, each a separate kernel over a queue of live paths:
//   generate (camera sample + primary ray) -> extend (closest hit) ->
//   [material histogram -> scatter: paths sorted by hit material] ->
//   shade (emission, layered BSDF walk, Russian roulette, next ray; finished
//   paths write the film) -> the survivors form the next queue.
// The recursion L = Le + f * L' * cos / pdf is evaluated forwards with a
// throughput: L = sum_d (prod_{j<d} f_j cos_j / pdf_j) * Le_d.
#include <algorithm>
#include <chrono>

#include "spt_bsdf.h"

namespace spt {

namespace {

// Path state, structure of arrays (one slot per in-flight path).
struct Paths {
    float4* o;      // ray origin
    float4* d;      // ray direction
    float4* thr;    // throughput
    float4* rad;    // radiance so far
    float4* hit;    // t, leaf index (bits), u, v
    uint64_t* key;  // RNG key or replay row
    uint64_t* local;  // path index within this render call
    uint32_t* pix;  // pixel index (y * width + x)
    uint32_t* k;    // draws consumed
    uint32_t* depth;
    uint32_t* mat;  // sort key of the current hit (n_mats = miss)
    uint32_t* q0;
    uint32_t* q1;
    uint32_t* qs;
    uint32_t* counters;  // [0] queue length out, [1] stream overflow count, [2..] material histogram
};

constexpr size_t kSlotBytes = 16 * 5 + 8 * 2 + 4 * 8;
constexpr uint32_t kMaxSortKeys = 1024;

struct RenderArgs {
    DevScene sc;
    DevCamera cam;
    uint32_t sampler, n_sub, spp, spp_begin, spp_count;
    float w_sub;
    int32_t max_depth, rr_depth;
    uint64_t seed;
    uint32_t vt, vl, vw, vh;
    float bg[3];
    uint32_t flags;
    const float* stream;
    uint32_t stride;
    float* film32;
    double* film64;
    float* path_rgb;
};

__device__ inline void finish_path(const RenderArgs& a, const Paths& p, uint32_t s, float4 L) {
    if (a.film32) {
        float* f = a.film32 + 3 * (size_t)p.pix[s];
        unsafeAtomicAdd(f + 0, L.x);
        unsafeAtomicAdd(f + 1, L.y);
        unsafeAtomicAdd(f + 2, L.z);
    }
    if (a.film64) {
        double* f = a.film64 + 3 * (size_t)p.pix[s];
        unsafeAtomicAdd(f + 0, (double)L.x);
        unsafeAtomicAdd(f + 1, (double)L.y);
        unsafeAtomicAdd(f + 2, (double)L.z);
    }
    if (a.path_rgb) {
        float* o = a.path_rgb + 3 * (size_t)p.local[s];
        o[0] = L.x;
        o[1] = L.y;
        o[2] = L.z;
    }
}

__device__ inline Draws make_draws(const RenderArgs& a, const Paths& p, uint32_t s) {
    Draws r;
    r.k = p.k[s];
    r.overflow = false;
    r.stride = a.stride;
    if (a.stream) {
        r.key = 0;
        r.stream = a.stream + p.key[s] * (uint64_t)a.stride;
    } else {
        r.key = p.key[s];
        r.stream = nullptr;
    }
    return r;
}

// Camera samples for paths [base, base + n): pathtracer.cpp:104-113.
__global__ __launch_bounds__(kBlock) void k_generate(RenderArgs a, Paths p, uint64_t base, uint32_t n) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint64_t g = base + s;
    const uint64_t pl = g / a.spp_count;
    const uint32_t smp = a.spp_begin + (uint32_t)(g % a.spp_count);
    const uint32_t x = a.vl + (uint32_t)(pl % a.vw), y = a.vt + (uint32_t)(pl / a.vw);
    const uint64_t path = ((uint64_t)y * a.cam.width + x) * a.spp + smp;
    p.pix[s] = y * a.cam.width + x;
    p.local[s] = g;
    p.key[s] = a.stream ? g : rng_key(a.seed, path);
    p.k[s] = 0;
    Draws r = make_draws(a, p, s);
    float sx, sy;
    pixel_offset(a.sampler, a.n_sub, a.w_sub, smp, r, sx, sy);
    const float xs = x + sx, ys = y + sy;
    V3 o, d;
    camera_ray(a.cam, xs, ys, o, d);
    p.k[s] = r.k;
    if (r.overflow) atomicAdd(&p.counters[1], 1u);
    p.o[s] = make_float4(o.x, o.y, o.z, 0.0f);
    p.d[s] = make_float4(d.x, d.y, d.z, 0.0f);
    p.thr[s] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    p.rad[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    p.depth[s] = 0;
    if (a.max_depth == 0) {  // depth >= path_length before the first trace
        finish_path(a, p, s, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        return;
    }
    const uint32_t slot = atomicAdd(&p.counters[0], 1u);
    p.q1[slot] = s;
}

// Closest hit for every queued path. FETCH adds the node / triangle record
// fetch counts of all rays to the two 64-bit counters at flags + 4.
template <bool FETCH>
__global__ __launch_bounds__(kBlock) void k_extend(RenderArgs a, Paths p, const uint32_t* __restrict__ q, uint32_t n,
                                                   uint32_t* __restrict__ spill, uint32_t* __restrict__ flags,
                                                   uint32_t sort_keys) {
    __shared__ uint32_t lds[kLdsStack * kBlock];
    __shared__ unsigned long long fetched[2];
    const uint32_t gthreads = gridDim.x * kBlock;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st{lds + threadIdx.x, spill + gtid, gthreads, a.sc.stack_cap};
    bool overflow = false;
    unsigned long long my_nodes = 0, my_tris = 0;
    if (FETCH) {
        if (threadIdx.x < 2) fetched[threadIdx.x] = 0;
        __syncthreads();
    }
    for (uint32_t i = gtid; i < n; i += gthreads) {
        const uint32_t s = q[i];
        const float4 o = p.o[s], d = p.d[s];
        Hit h;
        uint32_t tests, isects;
        trace_ray<true, false, FETCH>(a.sc, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z}, st, h, tests, isects, overflow);
        if (FETCH) {
            my_nodes += tests;
            my_tris += isects;
        }
        p.hit[s] = make_float4(h.t, __uint_as_float(h.idx), h.u, h.v);
        const uint32_t key = (h.idx == kNoHit) ? a.sc.n_mats : h.mat;
        p.mat[s] = key;
        if (sort_keys) (void)wave_add(p.counters + 2, key);
    }
    if (overflow) atomicOr(flags, 1u);
    if (FETCH) {
        atomicAdd(&fetched[0], my_nodes);
        atomicAdd(&fetched[1], my_tris);
        __syncthreads();
        if (threadIdx.x < 2)
            atomicAdd(reinterpret_cast<unsigned long long*>(flags + 4) + threadIdx.x, fetched[threadIdx.x]);
    }
}

// Persistent variant of k_extend (the throughput path).
template <bool FETCH>
__global__ __launch_bounds__(kBlock) void k_extend_p(RenderArgs a, Paths p, const uint32_t* __restrict__ q, uint32_t n,
                                                     uint32_t* __restrict__ spill, uint32_t* __restrict__ flags,
                                                     uint32_t sort_keys) {
    __shared__ uint32_t lds[kLdsStack * kBlock];
    __shared__ unsigned long long fetched[2];
    const uint32_t gthreads = gridDim.x * kBlock;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st{lds + threadIdx.x, spill + gtid, gthreads, a.sc.stack_cap};
    bool overflow = false;
    unsigned long long my_nodes = 0, my_tris = 0;
    if (FETCH) {
        if (threadIdx.x < 2) fetched[threadIdx.x] = 0;
        __syncthreads();
    }
    trace_persistent<FETCH>(
        a.sc, n, flags + 9, st, overflow,
        [&](uint32_t i, V3& o, V3& d) {
            const uint32_t s = q[i];
            const float4 o4 = p.o[s], d4 = p.d[s];
            o = V3{o4.x, o4.y, o4.z};
            d = V3{d4.x, d4.y, d4.z};
        },
        [&](uint32_t i, const Hit& h, uint32_t nn, uint32_t nt) {
            const uint32_t s = q[i];
            p.hit[s] = make_float4(h.t, __uint_as_float(h.idx), h.u, h.v);
            const uint32_t key = (h.idx == kNoHit) ? a.sc.n_mats : h.mat;
            p.mat[s] = key;
            if (sort_keys) (void)wave_add(p.counters + 2, key);
            if (FETCH) {
                my_nodes += nn;
                my_tris += nt;
            }
        });
    if (overflow) atomicOr(flags, 1u);
    if (FETCH) {
        atomicAdd(&fetched[0], my_nodes);
        atomicAdd(&fetched[1], my_tris);
        __syncthreads();
        if (threadIdx.x < 2)
            atomicAdd(reinterpret_cast<unsigned long long*>(flags + 4) + threadIdx.x, fetched[threadIdx.x]);
    }
}

// Exclusive scan of the material histogram (one block).
__global__ void k_scan(Paths p, uint32_t keys) {
    if (threadIdx.x != 0) return;
    uint32_t acc = 0;
    for (uint32_t k = 0; k < keys; ++k) {
        const uint32_t c = p.counters[2 + k];
        p.counters[2 + k] = acc;
        acc += c;
    }
}

__global__ __launch_bounds__(kBlock) void k_scatter(Paths p, const uint32_t* __restrict__ q, uint32_t n,
                                                    uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = q[i];
    const uint32_t pos = wave_add(p.counters + 2, p.mat[s]);
    out[pos] = s;
}

// Shading of one bounce: pathtracer.cpp:21-54 for every queued path.
__global__ __launch_bounds__(kBlock) void k_shade(RenderArgs a, Paths p, const uint32_t* __restrict__ q, uint32_t n,
                                                  uint32_t* __restrict__ qout) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = q[i];
    const float4 hv = p.hit[s];
    float4 rad = p.rad[s];
    const float4 thr = p.thr[s];
    const uint32_t depth = p.depth[s];
    const uint32_t idx = __float_as_uint(hv.y);
    if (idx == kNoHit) {  // miss: background at depth 0, else 0 (pathtracer.cpp:21-26)
        if (depth == 0) {
            rad.x += thr.x * a.bg[0];
            rad.y += thr.y * a.bg[1];
            rad.z += thr.z * a.bg[2];
        }
        finish_path(a, p, s, rad);
        return;
    }
    const float4 o4 = p.o[s], d4 = p.d[s];
    const V3 o{o4.x, o4.y, o4.z}, d{d4.x, d4.y, d4.z};
    Hit h;
    h.t = hv.x;
    h.idx = idx;
    h.u = hv.z;
    h.v = hv.w;
    const V3 pos = o + d * h.t;  // triangle.cpp:69
    const V3 nrm = shading_normal(a.sc, h);
    const uint32_t mid = p.mat[s];
    const spt_material& m = a.sc.mats[mid];
    const Frame fr = frame_from_up(nrm);              // pathtracer.cpp:28
    const V3 wo = normalize(to_local(fr, -d));        // pathtracer.cpp:30
    Draws r = make_draws(a, p, s);
    V3 bsdf, refl, wi;
    float pdf;
    layered_eval(m, wo, bsdf, refl, pdf, wi, r);      // pathtracer.cpp:36
    const V3 wwi = to_world(fr, wi);                  // pathtracer.cpp:37
    const V3 no = pos + wwi * 1e-3f;                  // pathtracer.cpp:38
    rad.x += thr.x * m.emission[0];
    rad.y += thr.y * m.emission[1];
    rad.z += thr.z * m.emission[2];
    const float kk = gmax(gmax(refl.x, refl.y), refl.z);  // pathtracer.cpp:41
    if ((a.rr_depth >= 0) && ((int32_t)depth >= a.rr_depth)) {  // pathtracer.cpp:43-49
        if ((r.next() < kk) && (depth < 1000u)) {
            bsdf = bsdf / kk;
        } else {
            if (r.overflow) atomicAdd(&p.counters[1], 1u);
            finish_path(a, p, s, rad);
            return;
        }
    }
    if (r.overflow) atomicAdd(&p.counters[1], 1u);
    // pathtracer.cpp:52-54, forwards: throughput *= f * cos / pdf
    const V3 wgt = (bsdf * wi.y) / pdf;
    const float4 nthr = make_float4(thr.x * wgt.x, thr.y * wgt.y, thr.z * wgt.z, 0.0f);
    const uint32_t nd = depth + 1;
    if ((a.max_depth >= 0) && ((int32_t)nd >= a.max_depth)) {  // pathtracer.cpp:18
        finish_path(a, p, s, rad);
        return;
    }
    p.rad[s] = rad;
    p.thr[s] = nthr;
    p.depth[s] = nd;
    p.k[s] = r.k;
    p.o[s] = make_float4(no.x, no.y, no.z, 0.0f);
    p.d[s] = make_float4(wwi.x, wwi.y, wwi.z, 0.0f);
    const uint32_t slot = atomicAdd(&p.counters[0], 1u);
    qout[slot] = s;
}

uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace

spt_status launch_render(spt_ctx* ctx, const spt_camera* cam, const spt_render_params* prm, float* d_rgb32,
                         double* d_rgb64, float* d_path_rgb) {
    auto t0 = std::chrono::steady_clock::now();
    RenderArgs a{};
    a.sc = ctx->scene();
    a.cam = make_camera(*cam);
    a.sampler = prm->pixel_sampler;
    if (a.sampler > SPT_PIXEL_UNIFORM) return set_error(ctx, SPT_ERR_INVALID, "unknown pixel sampler");
    if (prm->spp == 0) return set_error(ctx, SPT_ERR_INVALID, "spp must be > 0");
    if (a.sampler == SPT_PIXEL_UNIFORM) {
        a.n_sub = 1;
        a.spp = prm->spp;
        a.w_sub = 1.0f;
    } else {  // n = ceil(sqrt(spp)), spp := n^2 (jittered_pixel_sampler.cpp:6-8)
        a.n_sub = (uint32_t)std::ceil(std::sqrt((double)prm->spp));
        a.spp = a.n_sub * a.n_sub;
        a.w_sub = 1.0f / a.n_sub;
    }
    a.spp_begin = prm->spp_begin;
    a.spp_count = prm->spp_count ? prm->spp_count : a.spp - prm->spp_begin;
    if ((uint64_t)a.spp_begin + a.spp_count > a.spp || a.spp_count == 0)
        return set_error(ctx, SPT_ERR_INVALID, "sample range outside [0, spp)");
    a.vt = prm->viewport[0];
    a.vl = prm->viewport[1];
    a.vw = prm->viewport[2];
    a.vh = prm->viewport[3];
    if (a.vw == 0) {
        a.vt = a.vl = 0;
        a.vw = cam->width;
        a.vh = cam->height;
    }
    if ((uint64_t)a.vl + a.vw > cam->width || (uint64_t)a.vt + a.vh > cam->height || a.vh == 0)
        return set_error(ctx, SPT_ERR_INVALID, "viewport outside the image");
    a.max_depth = prm->max_depth;
    a.rr_depth = prm->rr_depth;
    if (a.max_depth < 0 && a.rr_depth < 0)
        return set_error(ctx, SPT_ERR_INVALID, "need max_depth >= 0 or rr_depth >= 0 (paths must terminate)");
    a.seed = prm->seed;
    for (int i = 0; i < 3; ++i) a.bg[i] = prm->background[i];
    a.flags = prm->flags;
    a.film32 = (prm->flags & SPT_RENDER_FILM32) ? d_rgb32 : nullptr;
    a.film64 = (prm->flags & SPT_RENDER_FILM64) ? d_rgb64 : nullptr;
    a.path_rgb = (prm->flags & SPT_RENDER_PER_PATH) ? d_path_rgb : nullptr;
    if ((prm->flags & SPT_RENDER_FILM32) && !d_rgb32) return set_error(ctx, SPT_ERR_INVALID, "FILM32 without buffer");
    if ((prm->flags & SPT_RENDER_FILM64) && !d_rgb64) return set_error(ctx, SPT_ERR_INVALID, "FILM64 without buffer");
    if ((prm->flags & SPT_RENDER_PER_PATH) && !d_path_rgb)
        return set_error(ctx, SPT_ERR_INVALID, "PER_PATH without buffer");
    if (prm->flags & SPT_RENDER_GUIDED) return set_error(ctx, SPT_ERR_UNSUPPORTED, "guided rendering not enabled");
    if (prm->rng == SPT_RNG_STREAM) {
        if (!prm->d_stream || prm->stream_stride == 0)
            return set_error(ctx, SPT_ERR_INVALID, "replay stream missing");
        a.stream = prm->d_stream;
        a.stride = prm->stream_stride;
    } else if (prm->rng != SPT_RNG_COUNTER) {
        return set_error(ctx, SPT_ERR_INVALID, "unknown rng mode");
    }
    const uint64_t total = (uint64_t)a.vw * a.vh * a.spp_count;
    const uint32_t keys = ctx->n_mats + 1;
    const bool sort = !(prm->flags & SPT_RENDER_NO_SORT) && keys <= kMaxSortKeys;

    // path slots: enough for the whole call, at most 4M in flight
    const uint64_t cap = std::min<uint64_t>(total, 1u << 22);
    const size_t bytes = (size_t)cap * kSlotBytes + (2 + kMaxSortKeys) * sizeof(uint32_t) + 1024;
    if (bytes > ctx->paths_bytes) {
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->d_paths) (void)hipFree(ctx->d_paths);
        ctx->d_paths = nullptr;
        ctx->paths_bytes = 0;
        spt_status s = check_hip(ctx, hipMalloc(&ctx->d_paths, bytes), "path state");
        if (s != SPT_OK) return s;
        ctx->paths_bytes = bytes;
    }
    Paths p;
    char* base = reinterpret_cast<char*>(ctx->d_paths);
    auto take = [&](size_t sz) {
        char* r = base;
        base += (sz + 255) & ~size_t(255);
        return r;
    };
    p.o = (float4*)take(cap * 16);
    p.d = (float4*)take(cap * 16);
    p.thr = (float4*)take(cap * 16);
    p.rad = (float4*)take(cap * 16);
    p.hit = (float4*)take(cap * 16);
    p.key = (uint64_t*)take(cap * 8);
    p.local = (uint64_t*)take(cap * 8);
    p.pix = (uint32_t*)take(cap * 4);
    p.k = (uint32_t*)take(cap * 4);
    p.depth = (uint32_t*)take(cap * 4);
    p.mat = (uint32_t*)take(cap * 4);
    p.q0 = (uint32_t*)take(cap * 4);
    p.q1 = (uint32_t*)take(cap * 4);
    p.qs = (uint32_t*)take(cap * 4);
    p.counters = (uint32_t*)take((2 + kMaxSortKeys) * 4);

    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_extend_p<true>),
                                                     kBlock, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    const uint32_t ext_grid = std::min<uint32_t>(blocks_for(cap), (uint32_t)(per_cu * ctx->num_cus));
    spt_status s = ensure_spill(ctx, (size_t)ext_grid * kBlock *
                                         (ctx->stack_cap > kLdsStack ? ctx->stack_cap - kLdsStack : 1));
    if (s != SPT_OK) return s;
    hipStream_t st = ctx->stream;
    const bool timed = (prm->flags & SPT_RENDER_TIME_KERNELS) != 0;
    const bool fetch_stats = (prm->flags & SPT_RENDER_FETCH_STATS) != 0;
    if (fetch_stats) {
        s = check_hip(ctx, hipMemsetAsync(ctx->d_flags + 4, 0, 16, st), "fetch counters");
        if (s != SPT_OK) return s;
    }
    spt_render_stats stats{};
    stats.paths = total;
    uint32_t overflow_total = 0;
    for (uint64_t done = 0; done < total; done += cap) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(cap, total - done);
        s = check_hip(ctx, hipMemsetAsync(p.counters, 0, 2 * sizeof(uint32_t), st), "counters");
        if (s != SPT_OK) return s;
        hipLaunchKernelGGL(k_generate, dim3(blocks_for(n)), dim3(kBlock), 0, st, a, p, done, n);
        uint32_t live = 0;
        s = check_hip(ctx, hipMemcpyAsync(&live, p.counters, 4, hipMemcpyDeviceToHost, st), "count");
        if (s == SPT_OK) s = check_hip(ctx, hipStreamSynchronize(st), "generate");
        if (s != SPT_OK) return s;
        uint32_t* qin = p.q1;
        uint32_t* qout = p.q0;
        while (live > 0) {
            ++stats.bounces;
            stats.rays += live;
            if (sort) (void)hipMemsetAsync(p.counters + 2, 0, keys * sizeof(uint32_t), st);
            const uint32_t g = std::min<uint32_t>(blocks_for(live), ext_grid);
            (void)hipMemsetAsync(ctx->d_flags + 9, 0, 4, st);
            if (timed) (void)hipEventRecord(ctx->ev[0], st);
            if (fetch_stats)
                hipLaunchKernelGGL(k_extend_p<true>, dim3(g), dim3(kBlock), 0, st, a, p, qin, live, ctx->d_spill,
                                   ctx->d_flags, sort ? 1u : 0u);
            else
                hipLaunchKernelGGL(k_extend_p<false>, dim3(g), dim3(kBlock), 0, st, a, p, qin, live, ctx->d_spill,
                                   ctx->d_flags, sort ? 1u : 0u);
            if (timed) (void)hipEventRecord(ctx->ev[1], st);
            const uint32_t* qshade = qin;
            if (sort) {
                hipLaunchKernelGGL(k_scan, dim3(1), dim3(64), 0, st, p, keys);
                hipLaunchKernelGGL(k_scatter, dim3(blocks_for(live)), dim3(kBlock), 0, st, p, qin, live, p.qs);
                qshade = p.qs;
            }
            (void)hipMemsetAsync(p.counters, 0, sizeof(uint32_t), st);
            hipLaunchKernelGGL(k_shade, dim3(blocks_for(live)), dim3(kBlock), 0, st, a, p, qshade, live, qout);
            s = check_hip(ctx, hipMemcpyAsync(&live, p.counters, 4, hipMemcpyDeviceToHost, st), "count");
            if (s == SPT_OK) s = check_hip(ctx, hipStreamSynchronize(st), "bounce");
            if (s != SPT_OK) return s;
            if (timed) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) stats.extend_ms += ms;
                ++stats.extend_launches;
            }
            std::swap(qin, qout);
        }
        uint32_t ov = 0;
        s = check_hip(ctx, hipMemcpy(&ov, p.counters + 1, 4, hipMemcpyDeviceToHost), "overflow count");
        if (s != SPT_OK) return s;
        overflow_total += ov;
    }
    stats.stream_overflows = overflow_total;
    stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->stats = stats;
    uint32_t flags = 0;
    s = check_hip(ctx, hipMemcpy(&flags, ctx->d_flags, 4, hipMemcpyDeviceToHost), "flags");
    if (s != SPT_OK) return s;
    if (fetch_stats) {
        unsigned long long f[2] = {0, 0};
        s = check_hip(ctx, hipMemcpy(f, ctx->d_flags + 4, 16, hipMemcpyDeviceToHost), "fetch counters");
        if (s != SPT_OK) return s;
        ctx->stats.node_fetches = f[0];
        ctx->stats.tri_fetches = f[1];
    }
    if (flags & 1u) {
        (void)hipMemset(ctx->d_flags, 0, 4);
        return set_error(ctx, SPT_ERR_OVERFLOW, "traversal stack overflow");
    }
    if (overflow_total) return set_error(ctx, SPT_ERR_OVERFLOW, "replay stream exhausted");
    return SPT_OK;
}

}  // namespace spt
