// spt_trace.hip - closest-hit ray casting kernels (batched
// AccelStructure::intersect, acceleration/accel_structure.h:29-30) and the
// three ray casters (integrators/raycaster.h:61-99).
#include <mutex>

#include "spt_device.h"

namespace spt {

namespace {

template <bool PRUNE, bool COUNT, bool FETCH>
__global__ __launch_bounds__(kBlock) void k_trace(DevScene sc, const float4* __restrict__ rays, uint64_t n,
                                                  uint2* __restrict__ hits, uint32_t* __restrict__ tests_out,
                                                  uint32_t* __restrict__ isects_out, uint32_t* __restrict__ spill,
                                                  uint32_t* __restrict__ flags) {
    __shared__ uint32_t lds[kLdsStack * kBlock];
    const uint32_t gthreads = gridDim.x * kBlock;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st(lds + threadIdx.x, spill + gtid, gthreads, sc.stack_cap);
    bool overflow = false;
    for (uint64_t i = gtid; i < n; i += gthreads) {
        const float4 ro = rays[2 * i], rd = rays[2 * i + 1];
        Hit h;
        uint32_t tests, isects;
        trace_ray<PRUNE, COUNT, FETCH>(sc, V3{ro.x, ro.y, ro.z}, V3{rd.x, rd.y, rd.z}, st, h, tests, isects,
                                       overflow);
        hits[i] = make_uint2(__float_as_uint(h.t), h.prim);
        if (COUNT || FETCH) {
            if (tests_out) tests_out[i] = tests;
            if (isects_out) isects_out[i] = isects;
        }
    }
    if (overflow) atomicOr(flags, 1u);
}

// Persistent closest-hit kernel: rays[i] -> hits[i] (FETCH: per-ray node and
// triangle record counts into nodes_out / tris_out).
template <bool FETCH>
__global__ __launch_bounds__(kBlock) void k_trace_p(DevScene sc, const float4* __restrict__ rays, uint32_t n,
                                                    uint2* __restrict__ hits, uint32_t* __restrict__ nodes_out,
                                                    uint32_t* __restrict__ tris_out, uint32_t* __restrict__ spill,
                                                    uint32_t* __restrict__ flags, uint32_t refill) {
    __shared__ uint32_t lds[kLdsStack * kBlock];
    const uint32_t gthreads = gridDim.x * kBlock;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st(lds + threadIdx.x, spill + gtid, gthreads, sc.stack_cap4);
    bool overflow = false;
    trace_persistent<FETCH>(
        sc, n, flags + 8, st, overflow, refill,
        [&](uint32_t i, V3& o, V3& d, Root&) {
            const float4 ro = rays[2 * (size_t)i], rd = rays[2 * (size_t)i + 1];
            o = V3{ro.x, ro.y, ro.z};
            d = V3{rd.x, rd.y, rd.z};
        },
        [&](uint32_t i, const Hit& h, uint32_t nn, uint32_t nt) {
            hits[i] = make_uint2(__float_as_uint(h.t), h.prim);
            if (FETCH) {
                nodes_out[i] = nn;
                tris_out[i] = nt;
            }
        });
    if (overflow) atomicOr(flags, 1u);
}

// One ray per pixel centre (RegularPixelSampler(1) gives (0.5, 0.5),
// regular_pixel_sampler.cpp:3-14); value per normal_raycaster.cpp:19-31,
// depth_raycaster.cpp:36-48, int_test_count_raycaster.cpp:75-85.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_raycast(DevScene sc, DevCamera cam, void* __restrict__ out,
                                                    uint32_t* __restrict__ spill, uint32_t* __restrict__ flags) {
    __shared__ uint32_t lds[kLdsStack * kBlock];
    const uint32_t gthreads = gridDim.x * kBlock;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st(lds + threadIdx.x, spill + gtid, gthreads, sc.stack_cap);
    bool overflow = false;
    const uint64_t n = (uint64_t)cam.width * cam.height;
    for (uint64_t i = gtid; i < n; i += gthreads) {
        const uint32_t x = (uint32_t)(i % cam.width), y = (uint32_t)(i / cam.width);
        V3 o, d;
        camera_ray(cam, x + 0.5f, y + 0.5f, o, d);
        Hit h;
        uint32_t tests, isects;
        if (KIND == SPT_RAYCAST_COUNT) {
            trace_ray<false, true>(sc, o, d, st, h, tests, isects, overflow);
            reinterpret_cast<uint32_t*>(out)[i] = tests;
        } else {
            trace_ray<true, false>(sc, o, d, st, h, tests, isects, overflow);
            if (KIND == SPT_RAYCAST_DEPTH) {
                reinterpret_cast<float*>(out)[i] = (h.idx != kNoHit) ? h.t : -1.0f;
            } else {
                V3 nrm = (h.idx != kNoHit) ? shading_normal(sc, h, o + d * h.t) : V3{0.0f, 0.0f, 0.0f};
                float* o3 = reinterpret_cast<float*>(out) + 3 * i;
                o3[0] = nrm.x;
                o3[1] = nrm.y;
                o3[2] = nrm.z;
            }
        }
    }
    if (overflow) atomicOr(flags, 1u);
}

__global__ __launch_bounds__(kBlock) void k_camera_rays(DevCamera cam, float4* __restrict__ rays) {
    const uint64_t n = (uint64_t)cam.width * cam.height;
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = (uint32_t)(i % cam.width), y = (uint32_t)(i / cam.width);
    V3 o, d;
    camera_ray(cam, x + 0.5f, y + 0.5f, o, d);
    rays[2 * i] = make_float4(o.x, o.y, o.z, 0.0f);
    rays[2 * i + 1] = make_float4(d.x, d.y, d.z, 0.0f);
}

// Scratch of the allocation-free closest-hit entry point
// (spt_trace_noalloc): words [0, 16) hold the overflow flag (word 0) and the
// persistent kernel's ray counter (word 8), words [64, ...) the traversal
// stacks' spill area. Each device holds its own copy; host code reaches it
// only through the address hipGetSymbolAddress returns for that device.
static __device__ __align__(256) uint32_t scratch[SPT_NOALLOC_SCRATCH_WORDS];
constexpr uint32_t kScratchSpill = 64;

__global__ void k_scratch_reset(uint32_t* __restrict__ ctl) {
    if (threadIdx.x < kScratchSpill) ctl[threadIdx.x] = 0u;
}

__global__ void k_scratch_fold_overflow(const uint32_t* __restrict__ ctl, uint32_t* __restrict__ ctx_flags) {
    if (threadIdx.x == 0 && ctl[0] != 0u) atomicOr(ctx_flags, 1u);
}

uint32_t grid_for(const spt_ctx* ctx, uint64_t n) {
    uint64_t want = (n + kBlock - 1) / kBlock;
    uint64_t cap = (uint64_t)ctx->num_cus * 16;  // resident-size grid, stride loop beyond
    if (want > cap) want = cap;
    if (want == 0) want = 1;
    return (uint32_t)want;
}

}  // namespace

uint32_t persistent_grid(spt_ctx* ctx, const void* kernel) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu <= 0)
        per_cu = 4;
    return (uint32_t)(per_cu * ctx->num_cus);
}

spt_status launch_trace(spt_ctx* ctx, const spt_ray32* d_rays, uint64_t n, spt_hit* d_hits, uint32_t mode,
                        uint32_t* d_tests, uint32_t* d_isects) {
    if (n == 0) return SPT_OK;
    if ((mode == SPT_TRACE_CLOSEST || mode == SPT_TRACE_FETCH) && n < 0xFFFFFF00ull) {
        const void* kern = (mode == SPT_TRACE_FETCH) ? reinterpret_cast<const void*>(&k_trace_p<true>)
                                                    : reinterpret_cast<const void*>(&k_trace_p<false>);
        const uint32_t grid = persistent_grid(ctx, kern);
        spt_status s =
            ensure_spill(ctx, (size_t)grid * kBlock * (ctx->stack_cap4 > kLdsStack ? ctx->stack_cap4 - kLdsStack : 1));
        if (s != SPT_OK) return s;
        s = check_hip(ctx, hipMemsetAsync(ctx->d_flags + 8, 0, 4, ctx->stream), "ray counter");
        if (s != SPT_OK) return s;
        DevScene sc = ctx->scene();
        const float4* rays = reinterpret_cast<const float4*>(d_rays);
        uint2* hits = reinterpret_cast<uint2*>(d_hits);
        // incoherent ray batches prefer eager refills (16 idle lanes); the path
        // tracer's default (48) is for its extension rays
        const uint32_t want = ctx->tuning.refill_lanes ? ctx->tuning.refill_lanes : 16u;
        const uint32_t refill = std::min<uint32_t>(64, std::max<uint32_t>(1, want));
        if (mode == SPT_TRACE_FETCH)
            hipLaunchKernelGGL(k_trace_p<true>, dim3(grid), dim3(kBlock), 0, ctx->stream, sc, rays, (uint32_t)n, hits,
                               d_tests, d_isects, ctx->d_spill, ctx->d_flags, refill);
        else
            hipLaunchKernelGGL(k_trace_p<false>, dim3(grid), dim3(kBlock), 0, ctx->stream, sc, rays, (uint32_t)n,
                               hits, nullptr, nullptr, ctx->d_spill, ctx->d_flags, refill);
        return check_hip(ctx, hipGetLastError(), "trace launch");
    }
    const uint32_t grid = grid_for(ctx, n);
    spt_status s = ensure_spill(ctx, (size_t)grid * kBlock * (ctx->stack_cap > kLdsStack ? ctx->stack_cap - kLdsStack : 1));
    if (s != SPT_OK) return s;
    DevScene sc = ctx->scene();
    const float4* rays = reinterpret_cast<const float4*>(d_rays);
    uint2* hits = reinterpret_cast<uint2*>(d_hits);
    if (mode == SPT_TRACE_CLOSEST)
        hipLaunchKernelGGL((k_trace<true, false, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, sc, rays, n,
                           hits, nullptr, nullptr, ctx->d_spill, ctx->d_flags);
    else if (mode == SPT_TRACE_COUNT)
        hipLaunchKernelGGL((k_trace<false, true, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, sc, rays, n,
                           hits, d_tests, d_isects, ctx->d_spill, ctx->d_flags);
    else if (mode == SPT_TRACE_FETCH)
        hipLaunchKernelGGL((k_trace<true, false, true>), dim3(grid), dim3(kBlock), 0, ctx->stream, sc, rays, n,
                           hits, d_tests, d_isects, ctx->d_spill, ctx->d_flags);
    else
        hipLaunchKernelGGL((k_trace<false, false, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, sc, rays, n,
                           hits, nullptr, nullptr, ctx->d_spill, ctx->d_flags);
    return check_hip(ctx, hipGetLastError(), "trace launch");
}

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
bool g_scratch_done[kMaxScratchDevices] = {};
}  // namespace

spt_status init_static_scratch(spt_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxScratchDevices) {
        ctx->scratch = nullptr;
        return SPT_OK;  // the allocation-free entry point reports it when called
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[ctx->device]) {
        // the calling thread has ctx->device current (spt_create set it)
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ctx->scratch = nullptr;
            return check_hip(ctx, e, "static scratch address");
        }
        g_scratch_ptr[ctx->device] = ptr;
        g_scratch_done[ctx->device] = true;
    }
    ctx->scratch = static_cast<uint32_t*>(g_scratch_ptr[ctx->device]);
    return SPT_OK;
}

spt_status launch_trace_noalloc(spt_ctx* ctx, const spt_ray32* d_rays, uint64_t n, spt_hit* d_hits) {
    if (n > SPT_NOALLOC_MAX_RAYS)
        return set_error(ctx, SPT_ERR_INVALID, "spt_trace_noalloc: more than SPT_NOALLOC_MAX_RAYS rays");
    if (!ctx->scratch) return set_error(ctx, SPT_ERR_STATE, "spt_trace_noalloc: no static scratch on this device");
    if (n == 0) return SPT_OK;
    const void* kern = reinterpret_cast<const void*>(&k_trace_p<false>);
    // spill words per lane: the stack entries beyond the LDS part
    const uint64_t per_lane = ctx->stack_cap4 > kLdsStack ? ctx->stack_cap4 - kLdsStack : 1;
    const uint64_t room = (uint64_t)SPT_NOALLOC_SCRATCH_WORDS - kScratchSpill;
    const uint64_t max_blocks = room / (per_lane * kBlock);
    if (max_blocks == 0)
        return set_error(ctx, SPT_ERR_UNSUPPORTED, "spt_trace_noalloc: the tree is too deep for the static scratch");
    const uint32_t grid = (uint32_t)std::min<uint64_t>(persistent_grid(ctx, kern), max_blocks);
    uint32_t* ctl = ctx->scratch;
    uint32_t* spill = ctx->scratch + kScratchSpill;
    // a fill kernel, not a memset node: callers may capture this in a graph
    hipLaunchKernelGGL(k_scratch_reset, dim3(1), dim3(kScratchSpill), 0, ctx->stream, ctl);
    const uint32_t want = ctx->tuning.refill_lanes ? ctx->tuning.refill_lanes : 16u;
    const uint32_t refill = std::min<uint32_t>(64, std::max<uint32_t>(1, want));
    hipLaunchKernelGGL(k_trace_p<false>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->scene(),
                       reinterpret_cast<const float4*>(d_rays), (uint32_t)n, reinterpret_cast<uint2*>(d_hits), nullptr,
                       nullptr, spill, ctl, refill);
    // a stack overflow is reported by spt_sync, as for spt_trace
    hipLaunchKernelGGL(k_scratch_fold_overflow, dim3(1), dim3(64), 0, ctx->stream, ctl, ctx->d_flags);
    return check_hip(ctx, hipGetLastError(), "trace (no allocation) launch");
}

spt_status launch_camera_rays(spt_ctx* ctx, const spt_camera* cam, spt_ray32* d_rays) {
    const uint64_t n = (uint64_t)cam->width * cam->height;
    const uint32_t grid = (uint32_t)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_camera_rays, dim3(grid), dim3(kBlock), 0, ctx->stream, make_camera(*cam),
                       reinterpret_cast<float4*>(d_rays));
    return check_hip(ctx, hipGetLastError(), "camera rays launch");
}

spt_status launch_raycast(spt_ctx* ctx, const spt_camera* cam, uint32_t kind, void* d_out) {
    const uint64_t n = (uint64_t)cam->width * cam->height;
    if (n == 0) return SPT_OK;
    const uint32_t grid = grid_for(ctx, n);
    spt_status s = ensure_spill(ctx, (size_t)grid * kBlock * (ctx->stack_cap > kLdsStack ? ctx->stack_cap - kLdsStack : 1));
    if (s != SPT_OK) return s;
    DevScene sc = ctx->scene();
    DevCamera dc = make_camera(*cam);
    if (kind == SPT_RAYCAST_NORMAL)
        hipLaunchKernelGGL((k_raycast<SPT_RAYCAST_NORMAL>), dim3(grid), dim3(kBlock), 0, ctx->stream, sc, dc, d_out,
                           ctx->d_spill, ctx->d_flags);
    else if (kind == SPT_RAYCAST_DEPTH)
        hipLaunchKernelGGL((k_raycast<SPT_RAYCAST_DEPTH>), dim3(grid), dim3(kBlock), 0, ctx->stream, sc, dc, d_out,
                           ctx->d_spill, ctx->d_flags);
    else
        hipLaunchKernelGGL((k_raycast<SPT_RAYCAST_COUNT>), dim3(grid), dim3(kBlock), 0, ctx->stream, sc, dc, d_out,
                           ctx->d_spill, ctx->d_flags);
    return check_hip(ctx, hipGetLastError(), "raycast launch");
}

}  // namespace spt
