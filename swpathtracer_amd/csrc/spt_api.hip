// spt_api.hip - the C ABI (include/spt.h): context, memory, scene upload and
// the entry points that stand in for AccelStructure::intersect and
// Integrator::render. No exception crosses this boundary.
#include <algorithm>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <new>

#include "spt_device.h"

namespace spt {

spt_status set_error(spt_ctx* ctx, spt_status code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

spt_status check_hip(spt_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return SPT_OK;
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(ctx, e == hipErrorOutOfMemory ? SPT_ERR_NOMEM : SPT_ERR_DEVICE, m);
}

spt_status stream_wait(spt_ctx* ctx, const char* what) {
    while (true) {
        const hipError_t e = hipStreamQuery(ctx->stream);
        if (e == hipSuccess) return SPT_OK;
        if (e != hipErrorNotReady) return check_hip(ctx, e, what);
        __builtin_ia32_pause();
    }
}

spt_status read_small(spt_ctx* ctx, void* dst, const void* d_src, size_t bytes, const char* what) {
    if (bytes > 4096 || !ctx->h_pin) return set_error(ctx, SPT_ERR_INVALID, "read_small: more than 4 KiB");
    spt_status s = check_hip(ctx, hipMemcpyAsync(ctx->h_pin, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream), what);
    if (s == SPT_OK) s = stream_wait(ctx, what);
    if (s == SPT_OK) std::memcpy(dst, ctx->h_pin, bytes);
    return s;
}

spt_status ensure_spill(spt_ctx* ctx, size_t words) {
    if (words <= ctx->spill_words) return SPT_OK;
    if (ctx->d_spill) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return check_hip(ctx, e, "sync before spill resize");
        (void)hipFree(ctx->d_spill);
        ctx->d_spill = nullptr;
        ctx->spill_words = 0;
    }
    hipError_t e = hipMalloc(&ctx->d_spill, words * sizeof(uint32_t));
    if (e != hipSuccess) return check_hip(ctx, e, "spill buffer");
    ctx->spill_words = words;
    return SPT_OK;
}

// cameras/camera.cpp:3-11, pinhole_camera.cpp:3-12, orthographic_camera.cpp:5-19
DevCamera make_camera(const spt_camera& c) {
    DevCamera d{};
    d.type = c.type;
    d.width = c.width;
    d.height = c.height;
    d.position = V3{c.position[0], c.position[1], c.position[2]};
    const V3 look_at{c.look_at[0], c.look_at[1], c.look_at[2]};
    const V3 up{c.up[0], c.up[1], c.up[2]};
    const V3 direction = normalize(look_at - d.position);
    d.frame = frame_from_uw(normalize(cross(up, -direction)), -direction);
    if (c.type == SPT_CAMERA_PINHOLE) {
        const float a = c.width * 0.5f;
        const float b = (float)(c.height * 0.5);
        // tan() of a float argument resolves to ::tan(double) in the reference
        // translation unit; the quotient is narrowed to float on assignment.
        const float vpd = (float)((double)b / std::tan((double)(c.fov_deg * 0.5f * kPi / 180.0f)));
        d.ulc = V3{-a, b, -vpd};
    } else {
        const float ww = c.max_x - c.min_x, wh = c.max_y - c.min_y;
        d.ulc = V3{c.min_x, c.max_y, 0.0f};
        d.pdx = ww / c.width;
        d.pdy = wh / c.height;
        d.ortho_dir = to_world(d.frame, V3{0.0f, 0.0f, -1.0f});
    }
    return d;
}

namespace {

void free_scene(spt_ctx* ctx) {
    if (ctx->d_nodes) (void)hipFree(ctx->d_nodes);
    if (ctx->d_nodesq) (void)hipFree(ctx->d_nodesq);
    if (ctx->d_leaves) (void)hipFree(ctx->d_leaves);
    if (ctx->d_tris) (void)hipFree(ctx->d_tris);
    if (ctx->d_shading) (void)hipFree(ctx->d_shading);
    ctx->d_nodes = nullptr;
    ctx->d_nodesq = nullptr;
    ctx->d_leaves = nullptr;
    ctx->n_nodesq = ctx->n_leaves = 0;
    ctx->d_tris = nullptr;
    ctx->d_shading = nullptr;
    ctx->n_nodes = ctx->n_tris = 0;
    ctx->has_scene = false;
}

template <typename T>
spt_status upload(spt_ctx* ctx, T** dst, const void* src, size_t bytes, const char* what) {
    // an empty array still gets a (16-B) allocation so kernels have a valid
    // address; only the caller's real bytes are copied
    hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), bytes ? bytes : 16);
    if (e != hipSuccess) return check_hip(ctx, e, what);
    if (src && bytes) {
        e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return check_hip(ctx, e, what);
    }
    return SPT_OK;
}

// inner-node depth of the flattened tree (longest chain of pair nodes)
uint32_t tree_depth(const spt_node64* nodes, uint32_t n) {
    if (n == 0) return 0;
    std::vector<uint32_t> depth(n, 0);
    uint32_t best = 1;
    depth[0] = 1;
    // pre-order: parents come before children
    for (uint32_t i = 0; i < n; ++i) {
        const spt_node64& nd = nodes[i];
        if (nd.lcnt == 0 && nd.lref < n) depth[nd.lref] = depth[i] + 1;
        if (nd.rcnt == 0 && nd.rref < n) depth[nd.rref] = depth[i] + 1;
        if (depth[i] > best) best = depth[i];
    }
    return best;
}

}  // namespace

// Triangles in leaf order with their material ids, and the 12-float shading
// records (normals, material, smooth flag) the kernels read per hit.
void scene_arrays(const spt_bvh* bvh, const float* verts, const float* normals, const uint32_t* mat_ids,
                  std::vector<spt_tri48>& tris, std::vector<float>& sh) {
    tris = bvh_tris(bvh);
    const std::vector<uint32_t>& order = bvh_leaf_prim(bvh);
    sh.assign((size_t)tris.size() * 12, 0.0f);
    for (size_t i = 0; i < tris.size(); ++i) {
        const uint32_t p = order[i];
        const uint32_t m = mat_ids ? mat_ids[p] : 0u;
        tris[i].pad[0] = m;
        float* o = &sh[i * 12];
        if (tris[i].pad[1] == SPT_PRIM_SPHERE) {  // normal from the hit point and the centre
            const uint32_t kind = kShadeSphere;
            std::memcpy(o, verts + 9ull * p, 3 * sizeof(float));
            std::memcpy(&o[9], &m, 4);
            std::memcpy(&o[10], &kind, 4);
            continue;
        }
        uint32_t smooth = 0u;
        if (normals)
            for (int j = 0; j < 9; ++j) smooth |= (normals[9ull * p + j] != 0.0f) ? 1u : 0u;
        if (smooth) {
            std::memcpy(o, normals + 9ull * p, 9 * sizeof(float));
        } else {
            const float* v = verts + 9ull * p;
            V3 a{v[0], v[1], v[2]}, b{v[3], v[4], v[5]}, c{v[6], v[7], v[8]};
            V3 n = normalize(cross(b - a, c - a));  // triangle.cpp:5-6
            o[0] = n.x;
            o[1] = n.y;
            o[2] = n.z;
        }
        std::memcpy(&o[9], &m, 4);
        std::memcpy(&o[10], &smooth, 4);
    }
}

uint32_t tree_stack_cap(const spt_node64* nodes, uint32_t n) { return tree_depth(nodes, n) + 1; }

namespace {

struct Entry4 {
    float lo[3], hi[3];
    uint32_t ref, cnt;  // cnt > 0: leaf (first triangle, count); 0: binary node index
};

float area4(const Entry4& e) {
    const float dx = e.hi[0] - e.lo[0], dy = e.hi[1] - e.lo[1], dz = e.hi[2] - e.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

void side(const spt_node64& nd, int s, Entry4& e) {
    const float* lo = s == 0 ? nd.lmin : nd.rmin;
    const float* hi = s == 0 ? nd.lmax : nd.rmax;
    for (int k = 0; k < 3; ++k) {
        e.lo[k] = lo[k];
        e.hi[k] = hi[k];
    }
    e.ref = s == 0 ? nd.lref : nd.rref;
    e.cnt = s == 0 ? nd.lcnt : nd.rcnt;
}

}  // namespace

uint32_t collapse4(const spt_node64* nodes, uint32_t n, std::vector<Node4>& out) {
    out.clear();
    if (n == 0) return 1;
    // work list: binary node, the Node4 slot it fills, its depth
    struct Item {
        uint32_t bin, slot, depth;
    };
    std::vector<Item> work{{0u, 0u, 1u}};
    out.emplace_back();
    uint32_t max_depth = 1;
    while (!work.empty()) {
        const Item j = work.back();
        work.pop_back();
        max_depth = std::max(max_depth, j.depth);
        // the binary node's two children, then the inner child with the
        // largest surface area is opened until there are four
        Entry4 e[4];
        int m = 0;
        side(nodes[j.bin], 0, e[m++]);
        side(nodes[j.bin], 1, e[m++]);
        while (m < 4) {
            int best = -1;
            float ba = -1.0f;
            for (int i = 0; i < m; ++i)
                if (e[i].cnt == 0 && area4(e[i]) > ba) {
                    ba = area4(e[i]);
                    best = i;
                }
            if (best < 0) break;
            const spt_node64& nd = nodes[e[best].ref];
            side(nd, 0, e[best]);
            side(nd, 1, e[m++]);
        }
        Node4 q;
        for (int i = 0; i < 4; ++i) {
            if (i < m) {
                q.lox[i] = e[i].lo[0];
                q.hix[i] = e[i].hi[0];
                q.loy[i] = e[i].lo[1];
                q.hiy[i] = e[i].hi[1];
                q.loz[i] = e[i].lo[2];
                q.hiz[i] = e[i].hi[2];
                q.count[i] = e[i].cnt;
                if (e[i].cnt > 0) {
                    q.child[i] = e[i].ref;
                } else {
                    q.child[i] = (uint32_t)out.size();
                    out.emplace_back();
                    work.push_back({e[i].ref, q.child[i], j.depth + 1});
                }
            } else {
                q.lox[i] = q.loy[i] = q.loz[i] = 0.0f;
                q.hix[i] = q.hiy[i] = q.hiz[i] = 0.0f;
                q.child[i] = kNoChild;
                q.count[i] = 0;
            }
        }
        out[j.slot] = q;
    }
    return 3 * max_depth + 1;
}

namespace {

// The value the traversal decodes for byte q (qdec): origin + q * scale in
// float, q * scale exact.
float grid_value(float origin, uint32_t q, float scale) {
    const volatile float p = (float)q * scale;
    return origin + p;
}

// One axis of a node: the grid (origin = smallest lo, scale = a power of two
// with origin + 255 scale >= largest hi) and each child's bytes, lo rounded
// down and hi rounded up on the grid. Returns the biased exponent byte.
uint32_t quantize_axis(const float* lo, const float* hi, const bool* used, float& origin, uint32_t& qlo,
                       uint32_t& qhi) {
    float mn = std::numeric_limits<float>::infinity(), mx = -std::numeric_limits<float>::infinity();
    for (int i = 0; i < 4; ++i)
        if (used[i]) {
            mn = std::min(mn, lo[i]);
            mx = std::max(mx, hi[i]);
        }
    if (!(mn <= mx)) mn = mx = 0.0f;
    origin = mn;
    int e = -126;
    const float ext = (mx - mn) / 255.0f;
    if (ext > 0.0f) {
        (void)std::frexp(ext, &e);  // ext < 2^e
        e = std::max(e, -126);
    }
    while (grid_value(origin, 255u, std::ldexp(1.0f, e)) < mx) ++e;
    const float s = std::ldexp(1.0f, e);
    qlo = qhi = 0;
    for (int i = 0; i < 4; ++i) {
        if (!used[i]) continue;
        double fl = std::floor(((double)lo[i] - origin) / s), fh = std::ceil(((double)hi[i] - origin) / s);
        uint32_t a = (uint32_t)std::min(std::max(fl, 0.0), 255.0), b = (uint32_t)std::min(std::max(fh, 0.0), 255.0);
        while (a > 0 && grid_value(origin, a, s) > lo[i]) --a;
        while (b < 255 && grid_value(origin, b, s) < hi[i]) ++b;
        qlo |= a << (8 * i);
        qhi |= b << (8 * i);
    }
    return (uint32_t)(e + 127);
}

}  // namespace

bool compress4(const std::vector<Node4>& in, uint32_t tri_base, std::vector<NodeQ>& nodes, std::vector<LeafRec>& leaves) {
    const uint64_t node_base = nodes.size();
    // leaf stack entries hold node << 2 | slot below kLeafBit
    if (node_base + in.size() > (1ull << 29)) return false;
    leaves.resize(4 * (node_base + in.size()), LeafRec{});
    for (size_t j = 0; j < in.size(); ++j) {
        const Node4& n = in[j];
        NodeQ q{};
        bool used[4];
        for (int i = 0; i < 4; ++i) used[i] = n.child[i] != kNoChild;
        const uint32_t ex = quantize_axis(n.lox, n.hix, used, q.ox, q.q[0], q.q[1]);
        const uint32_t ey = quantize_axis(n.loy, n.hiy, used, q.oy, q.q[2], q.q[3]);
        const uint32_t ez = quantize_axis(n.loz, n.hiz, used, q.oz, q.q[4], q.q[5]);
        q.exps = ex | (ey << 8) | (ez << 16);
        for (int i = 0; i < 4; ++i) {
            if (!used[i]) {
                q.child[i] = kNoChild;
            } else if (n.count[i] > 0) {
                if ((uint64_t)n.child[i] + tri_base + n.count[i] > 0xFFFFFFFFull) return false;
                LeafRec& r = leaves[4 * (node_base + j) + i];
                r.mn[0] = n.lox[i];
                r.mn[1] = n.loy[i];
                r.mn[2] = n.loz[i];
                r.mx[0] = n.hix[i];
                r.mx[1] = n.hiy[i];
                r.mx[2] = n.hiz[i];
                r.cnt = n.count[i];
                q.child[i] = n.child[i] + tri_base;
                q.cnt[i >> 1] |= std::min<uint32_t>(n.count[i], 0xFFFFu) << (16 * (i & 1));
            } else {
                q.child[i] = n.child[i] + (uint32_t)node_base;
            }
        }
        nodes.push_back(q);
    }
    return true;
}

}  // namespace spt

using namespace spt;

extern "C" {

spt_status spt_bvh_wide(const spt_bvh* bvh, void* nodes, uint64_t* n_nodes, void* leaves, uint64_t* n_leaves) {
    if (!bvh || !n_nodes || !n_leaves) return SPT_ERR_INVALID;
    const std::vector<spt_node64>& bn = bvh_nodes(bvh);
    std::vector<Node4> n4;
    (void)collapse4(bn.data(), (uint32_t)bn.size(), n4);
    std::vector<NodeQ> nq;
    std::vector<LeafRec> lr;
    if (!compress4(n4, 0u, nq, lr)) return SPT_ERR_UNSUPPORTED;
    if (nodes) {
        if (*n_nodes < nq.size()) return SPT_ERR_INVALID;
        std::memcpy(nodes, nq.data(), nq.size() * sizeof(NodeQ));
    }
    if (leaves) {
        if (*n_leaves < lr.size()) return SPT_ERR_INVALID;
        std::memcpy(leaves, lr.data(), lr.size() * sizeof(LeafRec));
    }
    *n_nodes = nq.size();
    *n_leaves = lr.size();
    return SPT_OK;
}


spt_status spt_create(int device, spt_ctx** out) {
    if (!out) return SPT_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return SPT_ERR_DEVICE;
    if (device < 0 || device >= count) return SPT_ERR_INVALID;
    spt_ctx* ctx = new (std::nothrow) spt_ctx();
    if (!ctx) return SPT_ERR_NOMEM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return SPT_ERR_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        ctx->wall_clock_khz = khz;
    if (hipHostMalloc(&ctx->h_pin, 4096, hipHostMallocDefault) != hipSuccess) ctx->h_pin = nullptr;
    if (!ctx->h_pin || hipMalloc(&ctx->d_flags, 64) != hipSuccess || hipMemset(ctx->d_flags, 0, 64) != hipSuccess) {
        if (ctx->h_pin) (void)hipHostFree(ctx->h_pin);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return SPT_ERR_DEVICE;
    }
    // status words the bounce loop writes for the host (mapped host memory)
    void* hs = nullptr;
    if (hipHostMalloc(&hs, 256, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) hs = nullptr;
    ctx->h_status = static_cast<volatile uint32_t*>(hs);
    if (!hs || hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->d_status), hs, 0) != hipSuccess) {
        spt_destroy(ctx);
        return SPT_ERR_DEVICE;
    }
    std::memset(hs, 0, 256);
    if (init_static_scratch(ctx) != SPT_OK) ctx->scratch = nullptr;  // only spt_trace_noalloc needs it
    for (hipEvent_t& e : ctx->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            spt_destroy(ctx);
            return SPT_ERR_DEVICE;
        }
    }
    *out = ctx;
    return SPT_OK;
}

void spt_destroy(spt_ctx* ctx) {
    if (!ctx) return;
    for (spt_ctx* f : ctx->fronts) spt_destroy(f);
    ctx->fronts.clear();
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_scene(ctx);
    if (ctx->d_mats) (void)hipFree(ctx->d_mats);
    if (ctx->d_spill) (void)hipFree(ctx->d_spill);
    if (ctx->d_flags) (void)hipFree(ctx->d_flags);
    if (ctx->d_paths) (void)hipFree(ctx->d_paths);
    if (ctx->guide.d_theta) (void)hipFree(ctx->guide.d_theta);
    if (ctx->guide.d_records) (void)hipFree(ctx->guide.d_records);
    if (ctx->guide.d_gout) (void)hipFree(ctx->guide.d_gout);
    if (ctx->guide.d_scratch) (void)hipFree(ctx->guide.d_scratch);
    ctx->guide.destroy_events();
    if (ctx->h_pin) (void)hipHostFree(ctx->h_pin);
    if (ctx->h_status) (void)hipHostFree(const_cast<uint32_t*>(ctx->h_status));
    if (ctx->d_guidable) (void)hipFree(ctx->d_guidable);
    for (hipEvent_t e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->tev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

spt_status spt_event_record(spt_ctx* ctx, int slot) {
    if (!ctx || slot < 0 || slot >= 8) return SPT_ERR_INVALID;
    return check_hip(ctx, hipEventRecord(ctx->ev[2 + slot], ctx->stream), "event record");
}

spt_status spt_event_elapsed(spt_ctx* ctx, int a, int b, float* ms) {
    if (!ctx || !ms || a < 0 || a >= 8 || b < 0 || b >= 8) return SPT_ERR_INVALID;
    spt_status s = check_hip(ctx, hipEventSynchronize(ctx->ev[2 + b]), "event sync");
    if (s != SPT_OK) return s;
    return check_hip(ctx, hipEventElapsedTime(ms, ctx->ev[2 + a], ctx->ev[2 + b]), "event elapsed");
}

const char* spt_last_error(const spt_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

spt_status spt_sync(spt_ctx* ctx) {
    if (!ctx) return SPT_ERR_INVALID;
    uint32_t flags = 0;
    spt_status s = read_small(ctx, &flags, ctx->d_flags, sizeof(flags), "stream sync");
    if (s != SPT_OK) return s;
    if (flags & 1u) {
        (void)hipMemset(ctx->d_flags, 0, 64);
        return set_error(ctx, SPT_ERR_OVERFLOW, "traversal stack overflow");
    }
    return SPT_OK;
}

void* spt_stream(spt_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

spt_status spt_malloc(spt_ctx* ctx, size_t bytes, void** d_ptr) {
    if (!ctx || !d_ptr) return SPT_ERR_INVALID;
    return check_hip(ctx, hipMalloc(d_ptr, bytes ? bytes : 16), "spt_malloc");
}

spt_status spt_free(spt_ctx* ctx, void* d_ptr) {
    if (!ctx) return SPT_ERR_INVALID;
    if (!d_ptr) return SPT_OK;
    return check_hip(ctx, hipFree(d_ptr), "spt_free");
}

spt_status spt_memcpy_h2d(spt_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
    if (!ctx || (!d_dst && bytes) || (!src && bytes)) return SPT_ERR_INVALID;
    return check_hip(ctx, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "h2d");
}

spt_status spt_memcpy_d2h(spt_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!d_src && bytes)) return SPT_ERR_INVALID;
    spt_status s = check_hip(ctx, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    if (s != SPT_OK) return s;
    return stream_wait(ctx, "d2h sync");
}

spt_status spt_memcpy_d2d(spt_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    if (!ctx || (!d_dst && bytes) || (!d_src && bytes)) return SPT_ERR_INVALID;
    return check_hip(ctx, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream), "d2d");
}

spt_status spt_upload_scene(spt_ctx* ctx, const spt_node64* nodes, uint32_t n_nodes, const spt_tri48* tris,
                            uint32_t n_tris, const float* root_box, const float* shading) {
    if (!ctx || !tris || n_tris == 0 || !root_box || (n_nodes && !nodes)) return SPT_ERR_INVALID;
    if (n_nodes >= (1u << 30)) return set_error(ctx, SPT_ERR_UNSUPPORTED, "more than 2^30 nodes");
    for (uint32_t i = 0; i < n_tris; ++i)
        if (tris[i].pad[1] > SPT_PRIM_SPHERE) return set_error(ctx, SPT_ERR_INVALID, "unknown primitive kind");
    // every child reference must stay inside the arrays the kernels read
    for (uint32_t i = 0; i < n_nodes; ++i) {
        const spt_node64& nd = nodes[i];
        const uint32_t refs[2] = {nd.lref, nd.rref}, cnts[2] = {nd.lcnt, nd.rcnt};
        for (int c = 0; c < 2; ++c) {
            if (cnts[c] == 0) {
                if (refs[c] >= n_nodes || refs[c] <= i)
                    return set_error(ctx, SPT_ERR_INVALID, "node child index out of range (pre-order expected)");
            } else if ((uint64_t)refs[c] + cnts[c] > n_tris) {
                return set_error(ctx, SPT_ERR_INVALID, "leaf triangle range out of bounds");
            }
        }
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_scene(ctx);
    std::vector<Node4> n4;
    const uint32_t cap4 = collapse4(nodes, n_nodes, n4);
    std::vector<NodeQ> nq;
    std::vector<LeafRec> leaves;
    if (!compress4(n4, 0u, nq, leaves)) return set_error(ctx, SPT_ERR_INVALID, "tree over 2^29 four-wide nodes");
    spt_status s = upload(ctx, &ctx->d_nodes, nodes, (size_t)n_nodes * sizeof(spt_node64), "nodes");
    if (s == SPT_OK) s = upload(ctx, &ctx->d_nodesq, nq.data(), nq.size() * sizeof(NodeQ), "compressed nodes");
    if (s == SPT_OK) s = upload(ctx, &ctx->d_leaves, leaves.data(), leaves.size() * sizeof(LeafRec), "leaf records");
    if (s == SPT_OK) s = upload(ctx, &ctx->d_tris, tris, (size_t)n_tris * sizeof(spt_tri48), "triangles");
    std::vector<float> flat;
    if (s == SPT_OK) {
        if (shading) {
            s = upload(ctx, &ctx->d_shading, shading, (size_t)n_tris * 12 * sizeof(float), "shading");
        } else {
            // flat normals from the triangle edges (triangle.cpp:5-6)
            flat.assign((size_t)n_tris * 12, 0.0f);
            for (uint32_t i = 0; i < n_tris; ++i) {
                const spt_tri48& t = tris[i];
                float* o = &flat[(size_t)i * 12];
                if (t.pad[1] == SPT_PRIM_SPHERE) {
                    const uint32_t kind = kShadeSphere;
                    std::memcpy(o, t.v0, 3 * sizeof(float));
                    std::memcpy(&o[10], &kind, 4);
                } else {
                    V3 n = normalize(cross(V3{t.e1[0], t.e1[1], t.e1[2]}, V3{t.e2[0], t.e2[1], t.e2[2]}));
                    o[0] = n.x;
                    o[1] = n.y;
                    o[2] = n.z;
                }
                uint32_t m = t.pad[0];
                std::memcpy(&o[9], &m, 4);
            }
            s = upload(ctx, &ctx->d_shading, flat.data(), flat.size() * sizeof(float), "shading");
        }
    }
    if (s != SPT_OK) {
        free_scene(ctx);
        return s;
    }
    ctx->n_nodes = n_nodes;
    ctx->n_tris = n_tris;
    std::memcpy(ctx->root_box, root_box, sizeof(ctx->root_box));
    ctx->root_leaf = n_nodes == 0 ? n_tris : 0;
    ctx->stack_cap = tree_depth(nodes, n_nodes) + 1;
    ctx->n_nodesq = (uint32_t)nq.size();
    ctx->n_leaves = (uint32_t)leaves.size();
    ctx->stack_cap4 = cap4;
    ctx->max_mat = 0;
    ctx->has_spheres = false;
    for (uint32_t i = 0; i < n_tris; ++i) {
        ctx->max_mat = std::max(ctx->max_mat, tris[i].pad[0]);
        ctx->has_spheres = ctx->has_spheres || tris[i].pad[1] == SPT_PRIM_SPHERE;
    }
    ctx->has_scene = true;
    return SPT_OK;
}

spt_status spt_set_scene(spt_ctx* ctx, const spt_bvh* bvh, const float* verts, const float* normals,
                         const uint32_t* mat_ids) {
    if (!ctx || !bvh || !verts) return SPT_ERR_INVALID;
    const std::vector<spt_node64>& nodes = bvh_nodes(bvh);
    std::vector<spt_tri48> tris;
    std::vector<float> sh;
    scene_arrays(bvh, verts, normals, mat_ids, tris, sh);
    return spt_upload_scene(ctx, nodes.data(), (uint32_t)nodes.size(), tris.data(), (uint32_t)tris.size(),
                            bvh_root_box(bvh), sh.data());
}

spt_status spt_set_materials(spt_ctx* ctx, const spt_material* mats, uint32_t n_mats) {
    if (!ctx || !mats || n_mats == 0) return SPT_ERR_INVALID;
    for (uint32_t i = 0; i < n_mats; ++i) {
        if (mats[i].n_layers == 0 || mats[i].n_layers > SPT_MAX_LAYERS)
            return set_error(ctx, SPT_ERR_INVALID, "material layer count must be 1..4");
        for (uint32_t l = 0; l < mats[i].n_layers; ++l)
            if (mats[i].layers[l].type < SPT_BSDF_LAMBERTIAN || mats[i].layers[l].type > SPT_BSDF_GGX_CONDUCTOR)
                return set_error(ctx, SPT_ERR_INVALID, "unknown BSDF layer type");
    }
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_mats) (void)hipFree(ctx->d_mats);
    ctx->d_mats = nullptr;
    spt_status s = upload(ctx, &ctx->d_mats, mats, (size_t)n_mats * sizeof(spt_material), "materials");
    if (s != SPT_OK) return s;
    ctx->n_mats = n_mats;
    ctx->h_mats.assign(mats, mats + n_mats);
    // vertices the learned guide may sample: a layer with an evaluable
    // density, alone or under one smooth dielectric coat
    auto evaluable = [](const spt_layer& l) {
        return l.type == SPT_BSDF_LAMBERTIAN || (l.type == SPT_BSDF_GGX_CONDUCTOR && l.sampler == SPT_SAMPLER_IMPORTANCE);
    };
    std::vector<uint8_t> g(n_mats, 0);
    for (uint32_t i = 0; i < n_mats; ++i) {
        const spt_material& m = mats[i];
        if (m.n_layers == 1 && evaluable(m.layers[0]))
            g[i] = kGuideSingle;
        else if (m.n_layers == 2 && m.layers[0].type == SPT_BSDF_SMOOTH_DIELECTRIC && m.layers[0].ior >= 1.0f &&
                 evaluable(m.layers[1]))
            g[i] = kGuideCoated;
    }
    if (ctx->d_guidable) (void)hipFree(ctx->d_guidable);
    ctx->d_guidable = nullptr;
    s = upload(ctx, &ctx->d_guidable, g.data(), g.size(), "guidable flags");
    if (s != SPT_OK) return s;
    return SPT_OK;
}

spt_status spt_trace(spt_ctx* ctx, const spt_ray32* d_rays, uint64_t n, spt_hit* d_hits, uint32_t mode,
                     uint32_t* d_tests, uint32_t* d_isects) {
    if (!ctx || (n && (!d_rays || !d_hits)) || mode > SPT_TRACE_FETCH) return SPT_ERR_INVALID;
    if (mode == SPT_TRACE_FETCH && (!d_tests || !d_isects)) return SPT_ERR_INVALID;
    if (!ctx->has_scene) return set_error(ctx, SPT_ERR_STATE, "no scene uploaded");
    return launch_trace(ctx, d_rays, n, d_hits, mode, d_tests, d_isects);
}

spt_status spt_trace_noalloc(spt_ctx* ctx, const spt_ray32* d_rays, uint64_t n, spt_hit* d_hits) {
    if (!ctx || (n && (!d_rays || !d_hits))) return SPT_ERR_INVALID;
    if (!ctx->has_scene) return set_error(ctx, SPT_ERR_STATE, "no scene uploaded");
    return launch_trace_noalloc(ctx, d_rays, n, d_hits);
}

spt_status spt_camera_rays(spt_ctx* ctx, const spt_camera* cam, spt_ray32* d_rays) {
    if (!ctx || !cam || !d_rays || cam->type > SPT_CAMERA_ORTHOGRAPHIC || cam->width == 0 || cam->height == 0)
        return SPT_ERR_INVALID;
    return launch_camera_rays(ctx, cam, d_rays);
}

spt_status spt_raycast(spt_ctx* ctx, const spt_camera* cam, uint32_t kind, void* d_out) {
    if (!ctx || !cam || !d_out || kind > SPT_RAYCAST_COUNT || cam->type > SPT_CAMERA_ORTHOGRAPHIC)
        return SPT_ERR_INVALID;
    if (cam->width == 0 || cam->height == 0) return SPT_ERR_INVALID;
    if (!ctx->has_scene) return set_error(ctx, SPT_ERR_STATE, "no scene uploaded");
    return launch_raycast(ctx, cam, kind, d_out);
}

spt_status spt_render(spt_ctx* ctx, const spt_camera* cam, const spt_render_params* p, float* d_rgb32,
                      double* d_rgb64, float* d_path_rgb) {
    if (!ctx || !cam || !p || cam->type > SPT_CAMERA_ORTHOGRAPHIC || cam->width == 0 || cam->height == 0)
        return SPT_ERR_INVALID;
    if (!ctx->has_scene) return set_error(ctx, SPT_ERR_STATE, "no scene uploaded");
    if (ctx->n_mats == 0) return set_error(ctx, SPT_ERR_STATE, "no materials uploaded");
    if (ctx->max_mat >= ctx->n_mats) return set_error(ctx, SPT_ERR_INVALID, "a triangle references a missing material");
    return launch_render(ctx, cam, p, d_rgb32, d_rgb64, d_path_rgb);
}

spt_status spt_render_get_stats(spt_ctx* ctx, spt_render_stats* out) {
    if (!ctx || !out) return SPT_ERR_INVALID;
    *out = ctx->stats;
    return SPT_OK;
}

}  // extern "C"
