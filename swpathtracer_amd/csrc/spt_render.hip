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
#include <mutex>
#include <thread>

#include "spt_bsdf.h"
#include "spt_guide.h"

namespace spt {

namespace {

constexpr uint32_t kLdsMats = 48;         // shading stages up to this many materials in LDS
constexpr uint32_t kMaxSortKeys = 1024;  // materials + 1 (miss) sorted; more keys: unsorted
constexpr uint32_t kMaxSortBlocks = 2048;  // most chunks (blocks) of the histogram / scatter passes
constexpr uint32_t kScanThreads = 1024;

// Path state of one buffer (structure of arrays, one slot per live path).
// Two buffers alternate: shade reads buffer A and appends survivors to B.
struct PathBuf {
    float4* o;       // ray origin; .w: pixel index (bits)
    float4* d;       // ray direction; .w: scene of the path in a multi-scene call (bits)
    float4* thr;     // throughput; .w: draws consumed (bits)
    float4* rad;     // radiance so far; .w: depth (bits)
    uint64_t* key;   // RNG key (counter mode) or replay-stream row
    uint64_t* local; // path index within this render call (per-path output)
    uint4* grec;     // guide training records of this path: 3 slots + count
};
constexpr size_t kBufBytes = 16 * 5 + 8 * 2;
constexpr uint32_t kNoRec = 0xFFFFFFFFu;

// Per-bounce scratch shared by both buffers.
struct Work {
    float4* hit;       // t, leaf-order triangle index (bits), u, v   (indexed like buffer A)
    uint32_t* skey;    // sort key: material of the hit, n_mats for a miss
    uint32_t* perm;    // shading order (stable counting sort by skey)
    uint32_t* bhist;   // [key][block] histogram, then exclusive offsets
    uint32_t* cnt;     // [0], [1]: live counts of the two buffers; [2]: replay overflows;
                       // [8]: ray-fetch counter of the extend kernel; u64 at [10]: rays, u64 at [12]: bounces
};
constexpr size_t kWorkBytes = 16 + 4 + 4;

struct RenderArgs {
    DevScene sc;
    Job job;                    // single-scene call
    const Job* jobs;            // multi-scene call: per-job parameters (path d.w = job)
    const BatchTile* tiles;     // multi-scene call: generation order
    const uint64_t* tile_start; // first path of each tile (n_tiles + 1 entries)
    uint32_t n_tiles;
    uint32_t* job_done;         // paths finished per job
    uint64_t* job_clock;        // [0] start, [1 + j] finish of job j (device wall clock)
    uint32_t flags;
    uint32_t per_path;  // some job writes per-path radiance: paths carry their output index
    uint32_t refill;  // idle lanes that trigger a ray refill in the traversal
    uint32_t keys;  // sort keys (n_mats + 1), 0 = no sort
    uint64_t gen_base;  // number of this wavefront's first path in the call's numbering
    // learned guide
    uint32_t guided, train;
    uint32_t gflags;          // SPT_GUIDE_*
    const uint8_t* guidable;  // per material: kGuideSingle / kGuideCoated / 0
    float bsdf_frac;
    unsigned short* gout;     // per path slot: 40 raw network outputs (IEEE half)
    float* records;           // training records (guide::kRecord floats each)
    uint32_t* rec_count;  // [0]: records written, [1]: candidate vertices (one 8-B aligned u64)
    uint32_t* rec_shared; // concurrent wavefronts: one rec_count for all (else nullptr: the work words)
    const float* theta;          // guide parameters (FP32 master copy)
    const unsigned short* wbf;   // guide weights in bf16
    GuideBox gbox;               // encoding box of the guide
    uint32_t rec_cap;
    uint32_t rec_stride;  // record 1 in rec_stride candidates (hash of key and depth)
    uint32_t rec_ring;    // in-call training: record slots wrap around rec_cap
};

__device__ inline float lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// Tag of record k of the path with RNG key `key`, kept in the record: with a
// ring of record slots a slot can be reused while its first path is alive,
// and that path must then leave the new record alone.
__device__ inline uint32_t rec_tag(uint64_t key, uint32_t k) {
    return (uint32_t)mix64(key * 3u + k + 1u);
}

// Training targets of a finished path: the radiance that arrived along each
// recorded wi is (L - radiance before it) / throughput after it.
__device__ inline void finish_records(const RenderArgs& a, uint4 gr, float4 L, uint64_t key) {
    const uint32_t sl[3] = {gr.x, gr.y, gr.z};
    for (uint32_t k = 0; k < gr.w && k < 3; ++k) {
        float* r = a.records + (size_t)sl[k] * guide::kRecord;
        if (__float_as_uint(r[20]) != rec_tag(key, k)) continue;  // slot reused
        const float lt = lum(V3{r[17], r[18], r[19]});
        const float li = lum(V3{L.x - r[14], L.y - r[15], L.z - r[16]});
        r[12] = (lt > 0.0f && li > 0.0f) ? r[13] * li / lt : 0.0f;
    }
}

// Guide mode of a material in this call: coated layers only with SPT_GUIDE_COATED.
__device__ inline uint32_t guide_mode(const RenderArgs& a, uint32_t mid) {
    const uint32_t g = a.guidable[mid];
    return (g == kGuideCoated && !(a.gflags & SPT_GUIDE_COATED)) ? 0u : g;
}

template <bool MULTI>
__device__ inline const Job& job_of(const RenderArgs& a, uint32_t j) {
    if (MULTI) return a.jobs[j];
    return a.job;
}

// Film and per-path output of a finished path; in a multi-scene call also
// counts the job's finished paths and stamps the job's finish time when its
// last path ends. Called by every active lane of the branch together.
template <bool MULTI>
__device__ inline void finish_path(const RenderArgs& a, const Job& jb, uint32_t jid, uint32_t pix, uint64_t local,
                                   float4 L) {
    if (MULTI) {
        const uint32_t k = wave_add(a.job_done, jid);
        if ((uint64_t)k + 1 == jb.n_paths) a.job_clock[1 + jid] = wall_clock64();
    }
    if (jb.film32) {
        float* f = jb.film32 + 3 * (size_t)pix;
        unsafeAtomicAdd(f + 0, L.x);
        unsafeAtomicAdd(f + 1, L.y);
        unsafeAtomicAdd(f + 2, L.z);
    }
    if (jb.film64) {
        double* f = jb.film64 + 3 * (size_t)pix;
        unsafeAtomicAdd(f + 0, (double)L.x);
        unsafeAtomicAdd(f + 1, (double)L.y);
        unsafeAtomicAdd(f + 2, (double)L.z);
    }
    if (jb.path_rgb) {
        float* o = jb.path_rgb + 3 * (size_t)local;
        o[0] = L.x;
        o[1] = L.y;
        o[2] = L.z;
    }
}

__device__ inline Draws make_draws(const Job& jb, uint64_t key, uint32_t k) {
    Draws r;
    r.k = k;
    r.overflow = false;
    r.stride = jb.stride;
    if (jb.stream) {
        r.key = 0;
        r.stream = jb.stream + key * (uint64_t)jb.stride;
    } else {
        r.key = key;
        r.stream = nullptr;
    }
    return r;
}

// Camera sample of path g into slot s of buffer A:
// integrators/pathtracer.cpp:104-113. A single-scene call numbers its paths
// sample by sample over the viewport's pixels; a multi-scene call walks its
// tile list. The per-path output index is pixel * spp_count + sample either way.
// With max depth 0 (single-scene calls only) the path finishes here.
template <bool MULTI>
__device__ inline void gen_path(const RenderArgs& a, PathBuf& A, const Work& w, uint32_t s, uint64_t g) {
    g += a.gen_base;  // a wavefront numbers its own paths from 0
    uint32_t jid = 0;
    uint64_t pl, local;
    uint32_t smp;
    if (MULTI) {
        uint32_t lo = 0, hi = a.n_tiles - 1;  // last tile starting at or before g
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (a.tile_start[mid] <= g)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchTile tl = a.tiles[lo];
        const uint64_t q = g - a.tile_start[lo];
        jid = tl.job;
        pl = tl.p0 + q % tl.np;
        smp = tl.s0 + (uint32_t)(q / tl.np);
        local = pl * a.jobs[jid].spp_count + (smp - a.jobs[jid].spp_begin);
    } else {
        // pixel-minor: consecutive paths are neighbouring pixels (coherent
        // primary rays, and no pile-up of one pixel's film atomics in a wave)
        const uint64_t npl = (uint64_t)a.job.vw * a.job.vh;
        pl = g % npl;
        smp = a.job.spp_begin + (uint32_t)(g / npl);
        local = pl * a.job.spp_count + (smp - a.job.spp_begin);
    }
    const Job& jb = job_of<MULTI>(a, jid);
    const uint32_t x = jb.vl + (uint32_t)(pl % jb.vw), y = jb.vt + (uint32_t)(pl / jb.vw);
    const uint64_t path = ((uint64_t)y * jb.cam.width + x) * jb.spp + smp;
    const uint32_t pix = y * jb.cam.width + x;
    const uint64_t key = jb.stream ? local : rng_key(jb.seed, path);
    Draws r = make_draws(jb, key, 0);
    float sx, sy;
    pixel_offset(jb.sampler, jb.n_sub, jb.w_sub, smp, r, sx, sy);
    const float xs = x + sx, ys = y + sy;
    V3 o, d;
    camera_ray(jb.cam, xs, ys, o, d);
    if (r.overflow) atomicAdd(&w.cnt[2], 1u);
    if (jb.max_depth == 0) {  // depth >= path_length before the first trace
        finish_path<MULTI>(a, jb, jid, pix, local, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        return;
    }
    A.o[s] = make_float4(o.x, o.y, o.z, __uint_as_float(pix));
    A.d[s] = make_float4(d.x, d.y, d.z, __uint_as_float(jid));
    A.thr[s] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(r.k));
    A.rad[s] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
    A.key[s] = key;
    if (a.per_path) A.local[s] = local;
    if (a.train) A.grec[s] = make_uint4(kNoRec, kNoRec, kNoRec, 0u);
}

// Paths [base, base + n) (of this wavefront) into slots [0, n) of buffer A;
// a call with max depth 0 finishes them all here.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_generate(RenderArgs a, PathBuf A, Work w, uint64_t base, uint32_t n) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s < n) gen_path<MULTI>(a, A, w, s, base + s);
}

// Work-word layout (w.cnt, 64 words): [0], [1] live counts of the two
// buffers; [2] replay overflows; [8] ray-fetch counter of the extend kernel;
// u64 [10] rays, u64 [12] bounces; [16] guide records, [17] candidates;
// u64 [20] next path to generate; [24] first slot and [25] number of the
// planned new paths, u64 [26] their first path index; [28] plans made;
// u64 [30] hits evaluated by the guide network.
constexpr int kGen = 20, kPlanSlot = 24, kPlanN = 25, kPlanPath = 26, kPlanSeq = 28;

// Between bounces (one thread): the consumed buffer's count and the
// ray-fetch counter go back to zero (a kernel, not a memset, so the bounce
// can be graph-captured), and new camera paths are planned into the free
// tail of the survivors' buffer B: n = min(capacity - survivors, paths not
// yet generated). Progress goes to mapped host memory: status[1] = live
// paths after the refill, status[2] = all paths generated, then status[0] =
// the plan's sequence number, so the host can stop issuing bounces without
// waiting on the stream.
__global__ void k_plan(uint32_t* cntA, uint32_t* cntB, uint32_t* cnt, uint32_t cap, uint64_t total,
                       uint32_t* status, uint64_t* job_clock) {
    if (threadIdx.x != 0) return;
    if (cntA) *cntA = 0;
    cnt[8] = 0;
    const uint32_t have = *cntB;
    uint64_t* gen = reinterpret_cast<uint64_t*>(cnt + kGen);
    const uint64_t g0 = *gen;
    if (g0 == 0 && job_clock) job_clock[0] = wall_clock64();
    const uint32_t n = (uint32_t)min((uint64_t)(cap - have), total - g0);
    cnt[kPlanSlot] = have;
    cnt[kPlanN] = n;
    *reinterpret_cast<uint64_t*>(cnt + kPlanPath) = g0;
    *cntB = have + n;
    *gen = g0 + n;
    const uint32_t seq = cnt[kPlanSeq] + 1;
    cnt[kPlanSeq] = seq;
    if (status) {
        volatile uint32_t* hs = status;
        hs[1] = have + n;
        hs[2] = (g0 + n == total) ? 1u : 0u;
        hs[3] = cnt[16];  // guide records written so far
        __threadfence_system();
        hs[0] = seq;
    }
}

// The paths k_plan planned, generated into buffer B (grid-stride).
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_refill(RenderArgs a, PathBuf B, Work w) {
    const uint32_t slot0 = w.cnt[kPlanSlot], n = w.cnt[kPlanN];
    const uint64_t g0 = *reinterpret_cast<const uint64_t*>(w.cnt + kPlanPath);
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock)
        gen_path<MULTI>(a, B, w, slot0 + s, g0 + s);
}

// Closest hit of every live path of buffer A (count *cntA). Persistent
// traversal; hit record and sort key written at the path's slot. FETCH adds
// node / triangle record fetch counts to the two 64-bit counters at flags + 4.
template <bool FETCH, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_extend(RenderArgs a, PathBuf A, Work w, const uint32_t* __restrict__ cntA,
                                                   uint32_t* __restrict__ spill, uint32_t* __restrict__ flags) {
    __shared__ uint32_t lds[kLdsStack * kBlock];
    __shared__ unsigned long long fetched[2];
    const uint32_t n = *cntA;
    if (n == 0) return;
    const uint32_t gthreads = gridDim.x * kBlock;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    if (gtid == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(w.cnt + 10), (unsigned long long)n);
        atomicAdd(reinterpret_cast<unsigned long long*>(w.cnt + 12), 1ull);
    }
    Stack st(lds + threadIdx.x, spill + gtid, gthreads, a.sc.stack_cap4);
    bool overflow = false;
    unsigned long long my_nodes = 0, my_tris = 0;
    if (FETCH) {
        if (threadIdx.x < 2) fetched[threadIdx.x] = 0;
        __syncthreads();
    }
    trace_persistent<FETCH>(
        a.sc, n, w.cnt + 8, st, overflow, a.refill,
        [&](uint32_t i, V3& o, V3& d, Root& root) {
            const float4 o4 = A.o[i], d4 = A.d[i];
            o = V3{o4.x, o4.y, o4.z};
            d = V3{d4.x, d4.y, d4.z};
            if (MULTI) {
                const Job& jb = a.jobs[__float_as_uint(d4.w)];
                root = Root{V3{jb.rmin[0], jb.rmin[1], jb.rmin[2]}, V3{jb.rmax[0], jb.rmax[1], jb.rmax[2]},
                            jb.root_leaf, jb.node0, jb.tri0};
            }
        },
        [&](uint32_t i, const Hit& h, uint32_t nn, uint32_t nt) {
            w.hit[i] = make_float4(h.t, __uint_as_float(h.idx), h.u, h.v);
            w.skey[i] = (h.idx == kNoHit) ? a.sc.n_mats : h.mat;
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

// Stable counting sort of the live slots by material, in three passes over
// contiguous chunks (one per block): per-chunk histogram (LDS), one exclusive scan
// over [key][chunk], then a scatter that ranks equal keys in slot order. No
// global atomics: many lanes with one material would serialise on them.
__device__ inline void chunk_range(uint32_t n, uint32_t b, uint32_t nb, uint32_t& c0, uint32_t& c1) {
    const uint32_t per = (n + nb - 1) / nb;
    c0 = min(n, b * per);
    c1 = min(n, c0 + per);
}

__global__ __launch_bounds__(kBlock) void k_hist(RenderArgs a, Work w, const uint32_t* __restrict__ cntA) {
    __shared__ uint32_t hist[kMaxSortKeys];
    const uint32_t n = *cntA;
    for (uint32_t k = threadIdx.x; k < a.keys; k += kBlock) hist[k] = 0;
    __syncthreads();
    uint32_t c0, c1;
    chunk_range(n, blockIdx.x, gridDim.x, c0, c1);
    for (uint32_t e = c0 + threadIdx.x; e < c1; e += kBlock) atomicAdd(&hist[w.skey[e]], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < a.keys; k += kBlock) w.bhist[k * gridDim.x + blockIdx.x] = hist[k];
}

// Exclusive prefix sum of the [key][chunk] counts in place, one block:
// tiles of kScanThreads counts read and written coalesced, each scanned by
// wave shuffles and one pass over the wave totals, with a running carry.
__global__ __launch_bounds__(kScanThreads) void k_scan(RenderArgs a, Work w, uint32_t nb) {
    constexpr uint32_t kWaves = kScanThreads / 64;
    __shared__ uint32_t wbase[kWaves];
    __shared__ uint32_t tile_sum;
    const uint32_t total = a.keys * nb;
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint32_t t0 = 0; t0 < total; t0 += kScanThreads) {
        const uint32_t i = t0 + threadIdx.x;
        const uint32_t v = i < total ? w.bhist[i] : 0u;
        uint32_t x = v;  // inclusive scan within the wave
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wbase[wave] = x;
        __syncthreads();
        if (wave == 0) {  // exclusive scan of the wave totals
            const uint32_t s = lane < kWaves ? wbase[lane] : 0u;
            uint32_t t = s;
            for (uint32_t off = 1; off < kWaves; off <<= 1) {
                const uint32_t y = __shfl_up(t, off);
                if (lane >= off) t += y;
            }
            if (lane < kWaves) wbase[lane] = t - s;
            if (lane == kWaves - 1) tile_sum = t;
        }
        __syncthreads();
        if (i < total) w.bhist[i] = carry + wbase[wave] + x - v;
        carry += tile_sum;
        __syncthreads();  // wbase and tile_sum are rewritten by the next tile
    }
}

__global__ __launch_bounds__(kBlock) void k_scatter(RenderArgs a, Work w, const uint32_t* __restrict__ cntA) {
    __shared__ uint32_t base[kMaxSortKeys];
    __shared__ uint32_t wcnt[kBlock / 64][kMaxSortKeys];
    const uint32_t n = *cntA;
    const uint32_t wave = threadIdx.x / 64;
    for (uint32_t k = threadIdx.x; k < a.keys; k += kBlock) {
        base[k] = w.bhist[k * gridDim.x + blockIdx.x];
        for (uint32_t v = 0; v < kBlock / 64; ++v) wcnt[v][k] = 0;
    }
    __syncthreads();
    uint32_t c0, c1;
    chunk_range(n, blockIdx.x, gridDim.x, c0, c1);
    const uint64_t lt = __lanemask_lt();
    for (uint32_t r0 = c0; r0 < c1; r0 += kBlock) {
        const uint32_t e = r0 + threadIdx.x;
        const bool valid = e < c1;
        const uint32_t key = valid ? w.skey[e] : 0u;
        uint32_t rank = 0;
        uint64_t todo = __ballot(valid);
        while (todo) {  // rank among equal keys of this wave, count per key
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const uint32_t lkey = __shfl(key, leader);
            const uint64_t group = __ballot(valid && key == lkey) & todo;
            if (__lane_id() == (unsigned)leader) wcnt[wave][lkey] = (uint32_t)__popcll(group);
            if (valid && key == lkey) rank = (uint32_t)__popcll(group & lt);
            todo &= ~group;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = base[key] + rank;
            for (uint32_t v = 0; v < wave; ++v) pos += wcnt[v][key];
            w.perm[pos] = e;
        }
        __syncthreads();
        todo = __ballot(valid);
        while (todo) {  // advance the running bases, clear this wave's counts
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const uint32_t lkey = __shfl(key, leader);
            const uint64_t group = __ballot(valid && key == lkey) & todo;
            if (__lane_id() == (unsigned)leader) {
                atomicAdd(&base[lkey], wcnt[wave][lkey]);
                wcnt[wave][lkey] = 0;
            }
            todo &= ~group;
        }
        __syncthreads();
    }
}

// Guided calls: the hits the guide network evaluates this bounce (those on
// guidable materials), read off the sort's exclusive offsets ([key][chunk]:
// the count of key m is the first offset of m + 1 minus that of m), added to
// the u64 at work word 30. One thread.
__global__ void k_count_guided(RenderArgs a, Work w, uint32_t nb) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long q = 0;
    for (uint32_t m = 0; m + 1 < a.keys; ++m)  // the last key is the misses
        if (guide_mode(a, m) != 0u) q += w.bhist[(m + 1) * nb] - w.bhist[m * nb];
    if (q) atomicAdd(reinterpret_cast<unsigned long long*>(w.cnt + 30), q);
}

// Guide inference for the live hits of buffer A whose material is guidable,
// in shading order: each wave encodes 64 hits (one per lane) and runs the
// register-resident MLP; raw outputs go to gout[slot] as half floats.
__global__ __launch_bounds__(kBlock) void k_guide_paths(RenderArgs a, PathBuf A, Work w,
                                                        const uint32_t* __restrict__ cntA,
                                                        const float* __restrict__ theta,
                                                        const unsigned short* __restrict__ wbf, GuideBox box) {
    using namespace guide;
    extern __shared__ float dyn[];
    FwdWeights& sw = *reinterpret_cast<FwdWeights*>(dyn);
    WaveStage* stages = reinterpret_cast<WaveStage*>(reinterpret_cast<char*>(dyn) + sizeof(FwdWeights));
    const uint32_t n = *cntA;
    if (n == 0) return;
    load_weights(sw, wbf, theta);
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = __lane_id(), r = l & 15;
    WaveStage& st = stages[wave];
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t j = base + threadIdx.x;
        bool mine = false;
        uint32_t slot = 0;
        float f[kEnc];
        if (j < n) {
            slot = a.keys ? w.perm[j] : j;
            const float4 hv = w.hit[slot];
            const uint32_t idx = __float_as_uint(hv.y);
            mine = (idx != kNoHit) && guide_mode(a, w.skey[slot]) != 0u;
            if (mine) {
                const float4 o4 = A.o[slot], d4 = A.d[slot];
                const V3 d{d4.x, d4.y, d4.z};
                Hit h;
                h.t = hv.x;
                h.idx = idx;
                h.u = hv.z;
                h.v = hv.w;
                const V3 pos = V3{o4.x, o4.y, o4.z} + d * h.t;
                encode(pos, shading_normal(a.sc, h, pos), -d, box.bmin, box.bext, f);
            }
        }
        if (!mine)
            for (int i = 0; i < kEnc; ++i) f[i] = 0.0f;
        const uint64_t any = __ballot(mine);
        stage_row(st, f);
        __syncthreads();
        if (any)
            mlp_stage_fwd(sw, st, [&](int g16, const f32x4 y[3]) {
                const int src = 16 * g16 + r;
                const uint32_t s = __shfl(slot, src);
                if ((any >> src) & 1ull) store_outputs_h(a.gout + (size_t)s * kOut, y);
            });
        __syncthreads();
    }
}

// Density of a single BSDF layer's own sampler (guidable layers only).
__device__ inline float layer_density(const spt_layer& l, V3 wo, V3 wi) {
    if (l.type == SPT_BSDF_LAMBERTIAN) {
        if (wi.y <= 0.0f) return 0.0f;
        return (l.sampler == SPT_SAMPLER_UNIFORM) ? 1.0f / (2.0f * kPi) : wi.y / kPi;
    }
    return ggx_pdf(l, wo, wi);
}

// Guide density of a direction w2 inside a smooth dielectric coat of IOR eta
// (local frame, pointing up): the guide proposes the exit direction above the
// coat, and w2 is the direction that refracts to it (tangential part scaled
// by 1/eta). p(w2) = p(exit) |d exit / d w2| = p(exit) eta^2 cos(w2) /
// cos(exit). Directions past the critical angle are never proposed.
__device__ inline float coat_guide_density(const guide::Lobe* lobes, const Frame& fr, V3 w2, float eta) {
    const float t2 = eta * eta * (w2.x * w2.x + w2.z * w2.z);
    if (w2.y <= 0.0f || t2 >= 1.0f) return 0.0f;
    const float cy = sqrtf(1.0f - t2);
    const V3 out{eta * w2.x, cy, eta * w2.z};
    return guide::mixture_pdf(lobes, to_world(fr, out)) * (eta * eta * w2.y / cy);
}

// vMF stand-in (local frame) for the lobe of a guided layer, multiplied with
// the learned incident radiance to propose directions: Lambertian, the
// normal with kappa 3 (mean cosine 2/3, that of the cosine lobe); GGX, the
// mirror direction with kappa 1 / (4 alpha^2) (reflected spread about twice
// the microfacet spread alpha), widened by 'widen' (the IOR of a smooth coat
// above the layer, which scales exit angles).
__device__ inline void bsdf_lobe(const spt_layer& l, V3 wo, float widen, V3& mb, float& kb) {
    if (l.type == SPT_BSDF_LAMBERTIAN) {
        mb = V3{0.0f, 1.0f, 0.0f};
        kb = 3.0f;
        return;
    }
    mb = V3{-wo.x, wo.y, -wo.z};
    const float s = l.roughness * widen;
    kb = fminf(fmaxf(0.25f / fmaxf(s * s, 1e-8f), guide::kKappaMin), 1e3f);
}

// The layered random walk of layered_eval (materials/layered_bsdf.cpp:10-74)
// for a smooth dielectric coat over one opaque layer with an evaluable
// density, where the first visit of the opaque layer draws its direction from
// a one-sample mixture: the layer's own sampler with probability c, else the
// guide's exit direction mapped under the coat. The step density is the
// mixture's, so the walk weight estimates the same BSDF without bias. Later
// visits (after total internal reflection under the coat, where the guide
// has no density) use the layer's own sampler: mixing there would double the
// weight of every such revisit. 'via' is set when the walk reached the
// opaque layer.
__device__ inline void layered_eval_guided(const spt_material& m, V3 wo, const Frame& frm, const guide::Lobe* lobes,
                                           float c, V3& bsdf, V3& refl, float& pdf, V3& wi, Draws& r, bool& via) {
    const spt_layer& coat = m.layers[0];
    const spt_layer& base = m.layers[1];
    const float eta = coat.ior, ieta = 1.0f / coat.ior;
    V3 fb{1.0f, 1.0f, 1.0f}, fr{1.0f, 1.0f, 1.0f};
    float fp = 1.0f;
    V3 nwo = wo;
    bool reached_inf = false;
    int bounce = 0;
    int id = (nwo.y > 0.0f) ? 0 : 1;
    const float inf = __builtin_inff();
    via = false;
    while ((id > -1) && (id < 2) && (fp > 0.0f) && (!reached_inf) && (bounce < 500)) {
        if (id == 0) {  // coat: the layer's own sampler, as in layered_eval
            const BSample s = layer_sample(coat, nwo, r);
            wi = s.wi;
            fp *= s.pdf;
            if (!(fp < inf)) reached_inf = true;
            if (wi.y > 0.0f) {
                fb = fb * layer_f(coat, nwo, wi);
                fr = fr * layer_reflectance(coat, nwo, wi);
                --id;
            } else if (wi.y < 0.0f) {
                fb = fb * (layer_f(coat, nwo, wi) * wi.y);
                fr = fr * layer_reflectance(coat, nwo, wi);
                ++id;
            } else {
                fp = 0.0f;
            }
        } else if (via) {  // opaque base, revisited: its own sampler, as in layered_eval
            const BSample s = layer_sample(base, nwo, r);
            wi = s.wi;
            fp *= s.pdf;
            if (!(fp < inf)) reached_inf = true;
            if (wi.y > 0.0f) {
                fb = fb * (layer_f(base, nwo, wi) * wi.y);
                fr = fr * layer_reflectance(base, nwo, wi);
                --id;
            } else {
                fp = 0.0f;
            }
        } else {  // opaque base, first visit: one-sample mixture with the guide
            via = true;
            if (r.next() < c) {
                wi = layer_sample(base, nwo, r).wi;
            } else {
                const float u0 = r.next(), u1 = r.next(), u2 = r.next();
                const V3 g = to_local(frm, guide::mixture_sample(lobes, u0, u1, u2));
                if (g.y > 0.0f) {
                    const float tx = g.x * ieta, tz = g.z * ieta;
                    wi = V3{tx, sqrtf(gmax(0.0f, 1.0f - tx * tx - tz * tz)), tz};
                } else {
                    wi = V3{0.0f, -1.0f, 0.0f};  // exit below the surface: no contribution
                }
            }
            const float p = c * layer_density(base, nwo, wi) + (1.0f - c) * coat_guide_density(lobes, frm, wi, eta);
            if ((wi.y > 0.0f) && (p > 0.0f)) {
                fp *= p;
                if (!(fp < inf)) reached_inf = true;
                fb = fb * (layer_f(base, nwo, wi) * wi.y);
                fr = fr * layer_reflectance(base, nwo, wi);
                --id;
            } else {
                fp = 0.0f;
            }
        }
        if (!((fb.x < inf) && (fb.y < inf) && (fb.z < inf))) reached_inf = true;
        nwo = -wi;
        ++bounce;
    }
    if ((bounce < 500) && !reached_inf && (fp > 0.0f)) {
        pdf = fp;
        bsdf = fb;
        refl = fr;
    } else {
        pdf = 1.0f;
        bsdf = V3{0.0f, 0.0f, 0.0f};
        refl = V3{0.0f, 0.0f, 0.0f};
    }
}

// Shading of one bounce (integrators/pathtracer.cpp:21-54) for every live
// path of A, in material order when sorted; survivors are appended to B.
// GUIDED instantiates the guide's sampling and record capture; the plain
// instantiation carries none of their registers. Four waves per SIMD (128
// VGPRs, a few spills) measured 2% faster than the two the register count
// allows: the kernel waits on gathers from a path queue no cache holds.
template <bool MULTI, bool GUIDED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_shade(RenderArgs a, PathBuf A, PathBuf B, Work w,
                                                  const uint32_t* __restrict__ cntA, uint32_t* __restrict__ cntB) {
    // the material table in LDS when it is small (each lane reads its
    // layered material field by field)
    __shared__ spt_material smats[kLdsMats];
    const bool lds_mats = a.sc.n_mats <= kLdsMats;
    if (lds_mats) {
        const uint32_t words = a.sc.n_mats * (uint32_t)(sizeof(spt_material) / 4);
        const uint32_t* src = reinterpret_cast<const uint32_t*>(a.sc.mats);
        uint32_t* dst = reinterpret_cast<uint32_t*>(smats);
        for (uint32_t k = threadIdx.x; k < words; k += kBlock) dst[k] = src[k];
    }
    __syncthreads();
    const uint32_t n = *cntA;
    const uint32_t gthreads = gridDim.x * kBlock;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gthreads) {
        const uint32_t i = a.keys ? w.perm[j] : j;
        const float4 hv = w.hit[i];
        float4 rad = A.rad[i];
        const float4 thr = A.thr[i];
        const float4 o4 = A.o[i], d4 = A.d[i];
        const uint32_t pix = __float_as_uint(o4.w), jid = __float_as_uint(d4.w);
        const Job& jb = job_of<MULTI>(a, jid);
        const uint32_t depth = __float_as_uint(rad.w);
        const uint32_t idx = __float_as_uint(hv.y);
        if (idx == kNoHit) {  // miss: background at depth 0, else 0 (pathtracer.cpp:21-26)
            if (depth == 0) {
                rad.x += thr.x * jb.bg[0];
                rad.y += thr.y * jb.bg[1];
                rad.z += thr.z * jb.bg[2];
            }
            if (GUIDED && a.train) finish_records(a, A.grec[i], rad, A.key[i]);
            finish_path<MULTI>(a, jb, jid, pix, a.per_path ? A.local[i] : 0ull, rad);
            continue;
        }
        const V3 o{o4.x, o4.y, o4.z}, d{d4.x, d4.y, d4.z};
        Hit h;
        h.t = hv.x;
        h.idx = idx;
        h.u = hv.z;
        h.v = hv.w;
        const V3 pos = o + d * h.t;  // triangle.cpp:69
        const V3 nrm = shading_normal(a.sc, h, pos);
        const uint32_t mid = w.skey[i];
        const spt_material& m = lds_mats ? smats[mid] : a.sc.mats[mid];
        const Frame fr = frame_from_up(nrm);          // pathtracer.cpp:28
        const V3 wo = normalize(to_local(fr, -d));    // pathtracer.cpp:30
        const uint64_t key = A.key[i];
        Draws r = make_draws(jb, key, __float_as_uint(thr.w));
        V3 bsdf, refl, wi;
        float pdf;
        const uint32_t gmode = GUIDED ? guide_mode(a, mid) : 0u;
        const bool prod = GUIDED && (a.gflags & SPT_GUIDE_PRODUCT);
        const bool guided = gmode != 0u;
        bool via = true;  // the sampled direction came from a guided step
        float pmarg = 0.0f;  // PRODUCT: density of wi, the training weight's 1/p
        if (gmode == kGuideCoated) {
            // with PRODUCT: the learned incident radiance times the coated
            // lobe in exit space
            using namespace guide;
            Lobe lobes[kLobes];
            mixture_params(HalfRow{a.gout + (size_t)i * kOut}, lobes);
            V3 mb;
            float kb;
            bsdf_lobe(m.layers[1], wo, m.layers[0].ior, mb, kb);
            const V3 mbw = to_world(fr, mb);
            if (prod) mixture_product(lobes, mbw, kb);
            layered_eval_guided(m, wo, fr, lobes, a.bsdf_frac, bsdf, refl, pdf, wi, r, via);
            if (prod && a.train && via && wi.y > 0.0f) {  // exit density, the lobe standing in for the walk
                const V3 e = to_world(fr, wi);
                Lobe bl;
                bl.mu = mbw;
                bl.kappa = kb;
                pmarg = a.bsdf_frac * expf(log_vmf(bl, e)) + (1.0f - a.bsdf_frac) * mixture_pdf(lobes, e);
            }
        } else if (guided) {
            // one-sample mixture of the layer's sampler and the learned
            // incident radiance times a vMF stand-in of the layer's lobe
            using namespace guide;
            const spt_layer& l = m.layers[0];
            Lobe lobes[kLobes];
            mixture_params(HalfRow{a.gout + (size_t)i * kOut}, lobes);
            if (prod) {
                V3 mb;
                float kb;
                bsdf_lobe(l, wo, 1.0f, mb, kb);
                mixture_product(lobes, to_world(fr, mb), kb);
            }
            if (r.next() < a.bsdf_frac) {
                wi = layer_sample(l, wo, r).wi;
            } else {
                const float u0 = r.next(), u1 = r.next(), u2 = r.next();
                wi = to_local(fr, mixture_sample(lobes, u0, u1, u2));
            }
            pdf = a.bsdf_frac * layer_density(l, wo, wi) +
                  (1.0f - a.bsdf_frac) * mixture_pdf(lobes, to_world(fr, wi));
            pmarg = pdf;
            if (wi.y > 0.0f && pdf > 0.0f) {
                bsdf = layer_f(l, wo, wi);
                refl = layer_reflectance(l, wo, wi);
            } else {
                bsdf = V3{0.0f, 0.0f, 0.0f};
                refl = V3{0.0f, 0.0f, 0.0f};
                pdf = 1.0f;
            }
        } else {
            layered_eval(m, wo, bsdf, refl, pdf, wi, r);  // pathtracer.cpp:36
        }
        const V3 wwi = to_world(fr, wi);              // pathtracer.cpp:37
        const V3 no = pos + wwi * 1e-3f;              // pathtracer.cpp:38
        rad.x += thr.x * m.emission[0];
        rad.y += thr.y * m.emission[1];
        rad.z += thr.z * m.emission[2];
        const float kk = gmax(gmax(refl.x, refl.y), refl.z);  // pathtracer.cpp:41
        // training weight before roulette: L_i / p (PRODUCT, radiance target)
        // or f cos / p L_i (the product target)
        const float rec_w = prod ? ((pmarg > 0.0f) ? 1.0f / pmarg : 0.0f) : lum(bsdf) * wi.y / pdf;
        bool alive = true;
        if ((jb.rr_depth >= 0) && ((int32_t)depth >= jb.rr_depth)) {  // pathtracer.cpp:43-49
            if ((r.next() < kk) && (depth < 1000u))
                bsdf = bsdf / kk;
            else
                alive = false;
        }
        if (r.overflow) atomicAdd(&w.cnt[2], 1u);
        const uint32_t nd = depth + 1;
        if (alive && (jb.max_depth >= 0) && ((int32_t)nd >= jb.max_depth)) alive = false;  // pathtracer.cpp:18
        if (!alive) {
            if (GUIDED && a.train) finish_records(a, A.grec[i], rad, A.key[i]);
            finish_path<MULTI>(a, jb, jid, pix, a.per_path ? A.local[i] : 0ull, rad);
            continue;
        }
        // pathtracer.cpp:52-54, forwards: throughput *= f * cos / pdf
        const V3 wgt = (bsdf * wi.y) / pdf;
        uint4 gr = (GUIDED && a.train) ? A.grec[i] : make_uint4(0u, 0u, 0u, 0u);
        const bool cand = GUIDED && a.train && guided && via && gr.w < 3u && wi.y > 0.0f && bsdf.x + bsdf.y + bsdf.z > 0.0f;
        const bool take = cand && (a.rec_stride <= 1u ||
                                   mix64(key + (uint64_t)(depth + 1) * 0x9E3779B97F4A7C15ull) % a.rec_stride == 0u);
        uint32_t rs = 0;
        if (GUIDED && a.train) {
            // one 64-bit atomic per wave adds the wave's candidates (high
            // word) and records (low word); the record slots of the wave are
            // consecutive from the old low word
            const uint64_t cb = __ballot(cand), tb = __ballot(take);
            const int leader = __ffsll((unsigned long long)cb) - 1;
            const unsigned long long inc = ((unsigned long long)__popcll(cb) << 32) | (unsigned long long)__popcll(tb);
            unsigned long long* rc = reinterpret_cast<unsigned long long*>(a.rec_count);
            if (tb) {
                unsigned long long old = 0;
                if (__lane_id() == (unsigned)leader) old = atomicAdd(rc, inc);
                rs = __shfl((uint32_t)old, leader) + (uint32_t)__popcll(tb & __lanemask_lt());
            } else if (cb && __lane_id() == (unsigned)leader) {
                atomicAdd(rc, inc);
            }
        }
        if (take) {
            if (a.rec_ring) rs %= a.rec_cap;
            if (rs < a.rec_cap) {
                float* rec = a.records + (size_t)rs * guide::kRecord;
                rec[20] = __uint_as_float(rec_tag(key, gr.w));
                const V3 wdir = to_world(fr, wi);
                const float vals[20] = {pos.x, pos.y, pos.z, nrm.x, nrm.y, nrm.z, -d.x, -d.y, -d.z, wdir.x, wdir.y,
                                        wdir.z, 0.0f, rec_w, rad.x, rad.y, rad.z,
                                        thr.x * wgt.x, thr.y * wgt.y, thr.z * wgt.z};
                for (int k = 0; k < 20; ++k) rec[k] = vals[k];
                if (gr.w == 0) gr.x = rs;
                else if (gr.w == 1) gr.y = rs;
                else gr.z = rs;
                ++gr.w;
            }
        }
        const uint32_t slot = atomicAdd(cntB, 1u);
        B.o[slot] = make_float4(no.x, no.y, no.z, __uint_as_float(pix));
        B.d[slot] = make_float4(wwi.x, wwi.y, wwi.z, __uint_as_float(jid));
        B.thr[slot] = make_float4(thr.x * wgt.x, thr.y * wgt.y, thr.z * wgt.z, __uint_as_float(r.k));
        B.rad[slot] = make_float4(rad.x, rad.y, rad.z, __uint_as_float(nd));
        B.key[slot] = key;
        if (a.per_path) B.local[slot] = A.local[i];
        if (GUIDED && a.train) B.grec[slot] = gr;
    }
}


uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

size_t guide_paths_lds_bytes() {
    return sizeof(guide::FwdWeights) + sizeof(guide::WaveStage) * (kBlock / 64);
}

}  // namespace

// Defaults for the zero fields of a tuning.
// Shading grid and traversal refill threshold by wavefront count: one
// wavefront waits for 48 idle lanes before refilling and shades with 128
// blocks per CU; with other wavefronts to cover a refill's stall, eager
// refills and a smaller shading grid measured faster (902-908 against 843
// Mpaths/s at three wavefronts).
void resolve_front_defaults(const spt_tuning& raw, spt_tuning& r, uint32_t fronts) {
    r.shade_blocks_per_cu = raw.shade_blocks_per_cu ? raw.shade_blocks_per_cu : (fronts > 1 ? 64u : 128u);
    r.refill_lanes = raw.refill_lanes ? raw.refill_lanes : (fronts > 1 ? 20u : 48u);
}

spt_tuning resolve_tuning(const spt_tuning& t) {
    spt_tuning r = t;
    if (r.path_capacity == 0) r.path_capacity = 1u << 27;
    if (r.use_graph == 0) r.use_graph = 2;
    if (r.tile_pixels == 0) r.tile_pixels = 4096;
    if (r.tile_samples == 0) r.tile_samples = 16;
    if (r.check_every == 0) r.check_every = 4;
    if (r.wavefronts == 0) r.wavefronts = 3;
    resolve_front_defaults(t, r, r.wavefronts);
    return r;
}

// Validates one render's parameters and fills its job (everything but the
// tree root). total = paths of the call.
spt_status make_job(spt_ctx* ctx, const spt_camera* cam, const spt_render_params* prm, float* d_rgb32, double* d_rgb64,
                    float* d_path_rgb, Job& jb, uint64_t& total) {
    jb = Job{};
    jb.cam = make_camera(*cam);
    jb.sampler = prm->pixel_sampler;
    if (jb.sampler > SPT_PIXEL_UNIFORM) return set_error(ctx, SPT_ERR_INVALID, "unknown pixel sampler");
    if (prm->spp == 0) return set_error(ctx, SPT_ERR_INVALID, "spp must be > 0");
    if (jb.sampler == SPT_PIXEL_UNIFORM) {
        jb.n_sub = 1;
        jb.spp = prm->spp;
        jb.w_sub = 1.0f;
    } else {  // n = ceil(sqrt(spp)), spp := n^2 (jittered_pixel_sampler.cpp:6-8)
        jb.n_sub = (uint32_t)std::ceil(std::sqrt((double)prm->spp));
        jb.spp = jb.n_sub * jb.n_sub;
        jb.w_sub = 1.0f / jb.n_sub;
    }
    jb.spp_begin = prm->spp_begin;
    jb.spp_count = prm->spp_count ? prm->spp_count : jb.spp - std::min(prm->spp_begin, jb.spp);
    if ((uint64_t)jb.spp_begin + jb.spp_count > jb.spp || jb.spp_count == 0)
        return set_error(ctx, SPT_ERR_INVALID, "sample range outside [0, spp)");
    jb.vt = prm->viewport[0];
    jb.vl = prm->viewport[1];
    jb.vw = prm->viewport[2];
    jb.vh = prm->viewport[3];
    if (jb.vw == 0) {
        jb.vt = jb.vl = 0;
        jb.vw = cam->width;
        jb.vh = cam->height;
    }
    if ((uint64_t)jb.vl + jb.vw > cam->width || (uint64_t)jb.vt + jb.vh > cam->height || jb.vh == 0)
        return set_error(ctx, SPT_ERR_INVALID, "viewport outside the image");
    jb.max_depth = prm->max_depth;
    jb.rr_depth = prm->rr_depth;
    if (jb.max_depth < 0 && jb.rr_depth < 0)
        return set_error(ctx, SPT_ERR_INVALID, "need max_depth >= 0 or rr_depth >= 0 (paths must terminate)");
    jb.seed = prm->seed;
    for (int i = 0; i < 3; ++i) jb.bg[i] = prm->background[i];
    jb.film32 = (prm->flags & SPT_RENDER_FILM32) ? d_rgb32 : nullptr;
    jb.film64 = (prm->flags & SPT_RENDER_FILM64) ? d_rgb64 : nullptr;
    jb.path_rgb = (prm->flags & SPT_RENDER_PER_PATH) ? d_path_rgb : nullptr;
    if ((prm->flags & SPT_RENDER_FILM32) && !d_rgb32) return set_error(ctx, SPT_ERR_INVALID, "FILM32 without buffer");
    if ((prm->flags & SPT_RENDER_FILM64) && !d_rgb64) return set_error(ctx, SPT_ERR_INVALID, "FILM64 without buffer");
    if ((prm->flags & SPT_RENDER_PER_PATH) && !d_path_rgb)
        return set_error(ctx, SPT_ERR_INVALID, "PER_PATH without buffer");
    if (prm->rng == SPT_RNG_STREAM) {
        if (!prm->d_stream || prm->stream_stride == 0)
            return set_error(ctx, SPT_ERR_INVALID, "replay stream missing");
        jb.stream = prm->d_stream;
        jb.stride = prm->stream_stride;
    } else if (prm->rng != SPT_RNG_COUNTER) {
        return set_error(ctx, SPT_ERR_INVALID, "unknown rng mode");
    }
    total = (uint64_t)jb.vw * jb.vh * jb.spp_count;
    jb.n_paths = total;
    return SPT_OK;
}

namespace {

// The dynamic LDS size of the guide inference kernel, set once per process
// (thread-safe: concurrent wavefronts start their calls together).
spt_status allow_guide_paths_lds(spt_ctx* ctx) {
    static std::once_flag once;
    static hipError_t e = hipSuccess;
    std::call_once(once, []() {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_guide_paths),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)guide_paths_lds_bytes());
    });
    return check_hip(ctx, e, "guide LDS size");
}

// Runs `total` paths through the wavefront: the queue is filled to its
// capacity, then every bounce refills it from the paths not yet generated,
// until every path is generated and none is live.
template <bool MULTI>
spt_status run_paths(spt_ctx* ctx, RenderArgs& a, uint64_t total, const spt_tuning& tn, spt_render_stats& stats) {
    uint64_t cap = std::min<uint64_t>(total, tn.path_capacity);
    auto state_bytes = [](uint64_t c) {
        return (size_t)c * (2 * kBufBytes + kWorkBytes + 4 * 8) + (size_t)kMaxSortKeys * kMaxSortBlocks * 4 + 4096;
    };
    if (state_bytes(cap) > ctx->paths_bytes) {
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->d_paths) (void)hipFree(ctx->d_paths);
        ctx->d_paths = nullptr;
        ctx->paths_bytes = 0;
        // a device without room for the queue gets a smaller one (slower,
        // same results), down to 2^20 paths
        while (true) {
            const hipError_t e = hipMalloc(&ctx->d_paths, state_bytes(cap));
            if (e == hipSuccess) break;
            (void)hipGetLastError();  // the failed allocation is not the next launch's error
            ctx->d_paths = nullptr;
            if (e != hipErrorOutOfMemory || cap <= (1u << 20)) return check_hip(ctx, e, "path state");
            cap = std::max<uint64_t>(cap / 2, 1u << 20);
        }
        ctx->paths_bytes = state_bytes(cap);
    }
    char* base = reinterpret_cast<char*>(ctx->d_paths);
    auto take = [&](size_t sz) {
        char* r = base;
        base += (sz + 255) & ~size_t(255);
        return r;
    };
    PathBuf buf[2];
    for (PathBuf& b : buf) {
        b.o = (float4*)take(cap * 16);
        b.d = (float4*)take(cap * 16);
        b.thr = (float4*)take(cap * 16);
        b.rad = (float4*)take(cap * 16);
        b.key = (uint64_t*)take(cap * 8);
        b.local = (uint64_t*)take(cap * 8);
        b.grec = (uint4*)take(cap * 16);
    }
    Work w;
    w.hit = (float4*)take(cap * 16);
    w.skey = (uint32_t*)take(cap * 4);
    w.perm = (uint32_t*)take(cap * 4);
    w.bhist = (uint32_t*)take((size_t)kMaxSortKeys * kMaxSortBlocks * 4);
    w.cnt = (uint32_t*)take(256);
    if ((size_t)(base - reinterpret_cast<char*>(ctx->d_paths)) > ctx->paths_bytes)
        return set_error(ctx, SPT_ERR_NOMEM, "path state layout exceeds its allocation");
    a.rec_count = a.rec_shared ? a.rec_shared : w.cnt + 16;
    a.refill = std::min<uint32_t>(64, std::max<uint32_t>(1, tn.refill_lanes));
    if (a.guided) {
        const size_t need = (size_t)cap * guide::kOut * sizeof(unsigned short);
        if (need > ctx->guide.gout_cap) {
            (void)hipStreamSynchronize(ctx->stream);
            if (ctx->guide.d_gout) (void)hipFree(ctx->guide.d_gout);
            ctx->guide.d_gout = nullptr;
            ctx->guide.gout_cap = 0;
            spt_status s = check_hip(ctx, hipMalloc(&ctx->guide.d_gout, need), "guide outputs");
            if (s != SPT_OK) return s;
            ctx->guide.gout_cap = need;
        }
        a.gout = ctx->guide.d_gout;
    }

    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_extend<true, MULTI>),
                                                     kBlock, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    if (tn.extend_blocks_per_cu)
        per_cu = std::min<int>(per_cu, (int)tn.extend_blocks_per_cu);
    else if (tn.wavefronts > 1)  // concurrent wavefronts share the CUs (measured: 4 each for two, 2 for three)
        per_cu = std::min<int>(per_cu, tn.wavefronts == 2 ? 4 : 2);
    const uint32_t ext_grid = std::min<uint32_t>(blocks_for(cap), (uint32_t)(per_cu * ctx->num_cus));
    const uint32_t shade_grid = std::min<uint32_t>(blocks_for(cap), (uint32_t)ctx->num_cus * tn.shade_blocks_per_cu);
    const uint32_t refill_grid = std::min<uint32_t>(blocks_for(cap), (uint32_t)ctx->num_cus * 8);
    // sort chunks: many (latency hiding) unless many keys would make the scan long
    const uint32_t sort_blocks =
        a.keys ? std::max<uint32_t>(1024, std::min<uint32_t>(kMaxSortBlocks, 16384 / a.keys)) : 256;
    spt_status s = ensure_spill(ctx, (size_t)ext_grid * kBlock *
                                         (a.sc.stack_cap4 > kLdsStack ? a.sc.stack_cap4 - kLdsStack : 1));
    if (s != SPT_OK) return s;
    hipStream_t st = ctx->stream;
    const bool timed = (a.flags & SPT_RENDER_TIME_KERNELS) != 0;
    const bool fetch_stats = (a.flags & SPT_RENDER_FETCH_STATS) != 0;
    s = check_hip(ctx, hipMemsetAsync(w.cnt, 0, 256, st), "counters");
    if (s != SPT_OK) return s;
    if (a.train && !a.rec_shared) {  // records continue after those collected since the last training step
        s = check_hip(ctx,
                      hipMemsetD32Async((hipDeviceptr_t)(w.cnt + 16), a.rec_ring ? 0 : (int)ctx->guide.collected, 1, st),
                      "record counter");
        if (s != SPT_OK) return s;
    }
    if (a.guided) {
        s = allow_guide_paths_lds(ctx);
        if (s != SPT_OK) return s;
    }
    if (fetch_stats) {
        s = check_hip(ctx, hipMemsetAsync(ctx->d_flags + 4, 0, 16, st), "fetch counters");
        if (s != SPT_OK) return s;
    }
    std::vector<hipEvent_t> gev;  // guide inference launches (TIME_KERNELS)
    // One bounce over buffer (b & 1) into the other buffer.
    auto bounce = [&](uint32_t b) -> spt_status {
        PathBuf& A = buf[b & 1];
        PathBuf& B = buf[(b + 1) & 1];
        uint32_t* cntA = w.cnt + (b & 1);
        uint32_t* cntB = w.cnt + ((b + 1) & 1);
        if (timed) {
            while (ctx->tev.size() < 2 * (stats.extend_launches + 1)) {
                hipEvent_t e;
                spt_status es = check_hip(ctx, hipEventCreate(&e), "event");
                if (es != SPT_OK) return es;
                ctx->tev.push_back(e);
            }
            (void)hipEventRecord(ctx->tev[2 * stats.extend_launches], st);
        }
        if (fetch_stats)
            hipLaunchKernelGGL((k_extend<true, MULTI>), dim3(ext_grid), dim3(kBlock), 0, st, a, A, w, cntA,
                               ctx->d_spill, ctx->d_flags);
        else
            hipLaunchKernelGGL((k_extend<false, MULTI>), dim3(ext_grid), dim3(kBlock), 0, st, a, A, w, cntA,
                               ctx->d_spill, ctx->d_flags);
        if (timed) {
            (void)hipEventRecord(ctx->tev[2 * stats.extend_launches + 1], st);
            ++stats.extend_launches;
        }
        if (a.keys) {
            hipLaunchKernelGGL(k_hist, dim3(sort_blocks), dim3(kBlock), 0, st, a, w, cntA);
            hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, a, w, sort_blocks);
            hipLaunchKernelGGL(k_scatter, dim3(sort_blocks), dim3(kBlock), 0, st, a, w, cntA);
            if (a.guided) hipLaunchKernelGGL(k_count_guided, dim3(1), dim3(64), 0, st, a, w, sort_blocks);
        }
        if (a.guided) {
            if (timed) {
                hipEvent_t e0, e1;
                if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
                    gev.push_back(e0);
                    gev.push_back(e1);
                    (void)hipEventRecord(e0, st);
                }
            }
            hipLaunchKernelGGL(k_guide_paths, dim3((uint32_t)ctx->num_cus * 3), dim3(kBlock), guide_paths_lds_bytes(),
                               st, a, A, w, cntA, a.theta, a.wbf, a.gbox);
            if (timed && !gev.empty()) (void)hipEventRecord(gev.back(), st);
        }
        if (a.guided)
            hipLaunchKernelGGL((k_shade<MULTI, true>), dim3(shade_grid), dim3(kBlock), 0, st, a, A, B, w, cntA, cntB);
        else
            hipLaunchKernelGGL((k_shade<MULTI, false>), dim3(shade_grid), dim3(kBlock), 0, st, a, A, B, w, cntA, cntB);
        // regeneration: the survivors' buffer is topped up with new camera paths
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, st, cntA, cntB, w.cnt, (uint32_t)cap, total, ctx->d_status,
                           a.job_clock);
        hipLaunchKernelGGL((k_refill<MULTI>), dim3(refill_grid), dim3(kBlock), 0, st, a, B, w);
        return check_hip(ctx, hipGetLastError(), "bounce launch");
    };
    // Graph mode: two bounces (one per buffer parity) captured once and
    // replayed; every kernel reads its live count from device memory, so the
    // same graph serves every bounce. No memset nodes are recorded.
    const bool graph_on = tn.use_graph == 1 && !timed && !fetch_stats;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto drop_graph = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    };
    if (graph_on) {
        s = check_hip(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
        if (s != SPT_OK) return s;
        spt_status s0 = bounce(0);
        spt_status s1 = bounce(1);
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (s0 != SPT_OK || s1 != SPT_OK) {
            drop_graph();
            return s0 != SPT_OK ? s0 : s1;
        }
        s = check_hip(ctx, e, "graph capture end");
        if (s == SPT_OK) s = check_hip(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
        if (s != SPT_OK) {
            drop_graph();
            return s;
        }
    }
    // The host stays at most `lag` bounces ahead of the device and stops once
    // a plan reports every path generated and none live; the bounces issued
    // after that one find empty queues (a few empty launches each).
    const uint32_t lag = std::max<uint32_t>(2, graph_on ? (tn.check_every + 1) & ~1u : tn.check_every);
    if (!MULTI && a.job.max_depth == 0) {  // every path finishes at generation
        for (uint64_t done = 0; done < total && s == SPT_OK; done += cap) {
            const uint32_t n = (uint32_t)std::min<uint64_t>(cap, total - done);
            hipLaunchKernelGGL((k_generate<MULTI>), dim3(blocks_for(n)), dim3(kBlock), 0, st, a, buf[0], w, done, n);
            s = check_hip(ctx, hipGetLastError(), "generate launch");
        }
    } else {
        ctx->h_status[0] = 0;
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, st, nullptr, w.cnt, w.cnt, (uint32_t)cap, total, ctx->d_status,
                           a.job_clock);
        hipLaunchKernelGGL((k_refill<MULTI>), dim3(refill_grid), dim3(kBlock), 0, st, a, buf[0], w);
        s = check_hip(ctx, hipGetLastError(), "generate launch");
        uint32_t b = 0;
        // in-call training (ring of records): train aligned batches of the
        // records counted one round earlier; skip what is about to be reused
        uint32_t next_train = ctx->guide.train_every, prev_count = 0, trained = 0;
        while (s == SPT_OK) {
            if (graph_on) {
                s = check_hip(ctx, hipGraphLaunch(exec, st), "graph launch");
                b += 2;
            } else {
                s = bounce(b);
                ++b;
            }
            if (s != SPT_OK || b <= lag) continue;
            // plan number 1 is the initial one; bounce j's plan is number j + 2
            const uint32_t want = b - lag + 1;
            bool idle = false;
            while (ctx->h_status[0] < want && !idle) {
                const hipError_t q = hipStreamQuery(st);
                if (q == hipSuccess) idle = true;  // drained: the status below is final
                else if (q != hipErrorNotReady) s = check_hip(ctx, q, "bounce");
                if (s != SPT_OK) break;
                __builtin_ia32_pause();
            }
            if (s == SPT_OK && ctx->h_status[2] && ctx->h_status[1] == 0) break;
            if (idle && ctx->h_status[0] < want) {
                s = set_error(ctx, SPT_ERR_DEVICE, "bounce status did not advance");
                break;
            }
            if (a.rec_ring && s == SPT_OK && b >= next_train) {
                const uint32_t now = ctx->h_status[3], upto = prev_count, bt = ctx->guide.batch;
                prev_count = now;
                next_train = b + ctx->guide.train_every;
                if (now > a.rec_cap / 2) trained = std::max(trained, (now - a.rec_cap / 2 + bt - 1) / bt * bt);
                while (s == SPT_OK && trained + bt <= upto) {
                    s = guide_train_launch(ctx, ctx->guide.d_records + (size_t)(trained % a.rec_cap) * guide::kRecord, bt,
                                           nullptr, 1u);
                    trained += bt;
                }
            }
        }
    }
    if (s != SPT_OK) {
        (void)stream_wait(ctx, "render abort");
        drop_graph();
        return s;
    }
    uint32_t cnt[32];  // the work words above
    s = read_small(ctx, cnt, w.cnt, sizeof(cnt), "render");
    drop_graph();
    if (s != SPT_OK) return s;
    stats.stream_overflows = cnt[2];
    if (a.train && !a.rec_shared) {
        GuideState& gs = ctx->guide;
        // a ring of records was trained inside the call; otherwise the
        // records wait for spt_guide_train_collected
        gs.collected = a.rec_ring ? 0 : std::min<uint64_t>(cnt[16], a.rec_cap);
        gs.candidates = cnt[17];
        // next pass: record about one batch worth of candidates
        gs.cand_per_path = total ? (double)cnt[17] / (double)total : 0.0;
    }
    for (size_t i = 0; i + 1 < gev.size(); i += 2) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, gev[i], gev[i + 1]) == hipSuccess) stats.guide_ms += ms;
    }
    for (hipEvent_t e : gev) (void)hipEventDestroy(e);
    stats.rays = (uint64_t)cnt[10] | ((uint64_t)cnt[11] << 32);
    if (a.train && !a.rec_shared && total) ctx->guide.rays_per_path = std::max(1.0, (double)stats.rays / (double)total);
    stats.bounces = (uint64_t)cnt[12] | ((uint64_t)cnt[13] << 32);
    stats.guide_queries = (uint64_t)cnt[30] | ((uint64_t)cnt[31] << 32);
    if (timed) {
        for (uint64_t i = 0; i < stats.extend_launches; ++i) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ctx->tev[2 * i], ctx->tev[2 * i + 1]) == hipSuccess) stats.extend_ms += ms;
        }
    }
    uint32_t flags = 0;
    s = read_small(ctx, &flags, ctx->d_flags, 4, "flags");
    if (s != SPT_OK) return s;
    if (fetch_stats) {
        unsigned long long f[2] = {0, 0};
        s = read_small(ctx, f, ctx->d_flags + 4, 16, "fetch counters");
        if (s != SPT_OK) return s;
        stats.node_fetches = f[0];
        stats.tri_fetches = f[1];
    }
    if (flags & 1u) {
        (void)hipMemset(ctx->d_flags, 0, 4);
        return set_error(ctx, SPT_ERR_OVERFLOW, "traversal stack overflow");
    }
    if (stats.stream_overflows) return set_error(ctx, SPT_ERR_OVERFLOW, "replay stream exhausted");
    return SPT_OK;
}

// Runs `total` paths as tn.wavefronts independent wavefronts at once: each
// on its own stream, with its own queue (the capacity split between them),
// status word, flags and stack spill buffer, taking a contiguous share of
// the call's path numbering. One wavefront's traversal (bound by VALU issue)
// then overlaps another's shading (bound by memory latency). Per-path results
// do not depend on the wavefront a path runs in. Guided calls (one record
// ring, one output buffer) and calls of fewer than 2^21 paths per wavefront
// run as one.
template <bool MULTI>
spt_status run_fronts(spt_ctx* ctx, RenderArgs& a, uint64_t total, const spt_tuning& tn, spt_render_stats& stats) {
    spt_tuning t1 = tn;
    uint32_t k = std::max<uint32_t>(1, std::min<uint32_t>(4, tn.wavefronts));
    if (a.rec_ring) k = 1;  // in-call training trains from one wavefront's ring
    while (k > 1 && total / k < (1ull << 21)) --k;
    if (k == 1) {
        t1.wavefronts = 1;
        resolve_front_defaults(ctx->tuning, t1, 1);
        return run_paths<MULTI>(ctx, a, total, t1, stats);
    }
    resolve_front_defaults(ctx->tuning, t1, k);
    while (ctx->fronts.size() < k - 1) {
        spt_ctx* f = nullptr;
        const spt_status cs = spt_create(ctx->device, &f);
        if (cs != SPT_OK) return set_error(ctx, cs, "context of a concurrent wavefront");
        ctx->fronts.push_back(f);
    }
    t1.wavefronts = k;
    t1.path_capacity = std::max<uint32_t>(1u << 20, tn.path_capacity / k);
    std::vector<RenderArgs> args(k, a);
    std::vector<spt_render_stats> st(k, stats);
    std::vector<spt_status> res(k, SPT_OK);
    std::vector<uint64_t> n(k);
    for (uint32_t i = 0; i < k; ++i) {
        const uint64_t lo = total * i / k, hi = total * (i + 1) / k;
        args[i].gen_base = lo;
        n[i] = hi - lo;
    }
    // training records of all wavefronts go to one buffer through one
    // counter pair (records written, candidate vertices) in the calling
    // context's flag words 12-13, starting after the records collected so far
    if (a.train) {
        const uint32_t init[2] = {(uint32_t)ctx->guide.collected, 0u};
        spt_status cs = check_hip(ctx, hipMemcpyAsync(ctx->d_flags + 12, init, sizeof(init), hipMemcpyHostToDevice,
                                                      ctx->stream), "record counter");
        if (cs != SPT_OK) return cs;
        for (RenderArgs& x : args) x.rec_shared = ctx->d_flags + 12;
    }
    if (a.guided) {
        spt_status ls = allow_guide_paths_lds(ctx);
        if (ls != SPT_OK) return ls;
    }
    // the stream of the calling context orders this call after earlier work
    // on it; the other wavefronts start once that work is done
    spt_status s = stream_wait(ctx, "render start");
    if (s != SPT_OK) return s;
    auto run = [&](uint32_t i) {
        if (hipSetDevice(ctx->device) != hipSuccess) {
            res[i] = SPT_ERR_DEVICE;
            return;
        }
        res[i] = run_paths<MULTI>(i ? ctx->fronts[i - 1] : ctx, args[i], n[i], t1, st[i]);
    };
    std::vector<std::thread> th;
    std::vector<uint32_t> inline_fronts;
    for (uint32_t i = 1; i < k; ++i) {
        try {
            th.emplace_back(run, i);
        } catch (...) {  // no thread: that wavefront runs after the others
            inline_fronts.push_back(i);
        }
    }
    run(0);
    for (std::thread& x : th) x.join();
    for (uint32_t i : inline_fronts) run(i);
    for (uint32_t i = 0; i < k; ++i) {
        if (res[i] != SPT_OK) {
            if (i > 0) ctx->err = ctx->fronts[i - 1]->err;
            return res[i];
        }
    }
    stats = st[0];
    for (uint32_t i = 1; i < k; ++i) {
        stats.rays += st[i].rays;
        stats.bounces = std::max(stats.bounces, st[i].bounces);
        stats.stream_overflows += st[i].stream_overflows;
        stats.extend_ms += st[i].extend_ms;
        stats.extend_launches += st[i].extend_launches;
        stats.node_fetches += st[i].node_fetches;
        stats.tri_fetches += st[i].tri_fetches;
        stats.guide_ms += st[i].guide_ms;
        stats.guide_queries += st[i].guide_queries;
    }
    if (a.train) {
        uint32_t rc[2];
        s = read_small(ctx, rc, ctx->d_flags + 12, sizeof(rc), "record counter");
        if (s != SPT_OK) return s;
        GuideState& gs = ctx->guide;
        gs.collected = std::min<uint64_t>(rc[0], a.rec_cap);
        gs.candidates = rc[1];
        gs.cand_per_path = total ? (double)rc[1] / (double)total : 0.0;
        if (total) gs.rays_per_path = std::max(1.0, (double)stats.rays / (double)total);
    }
    return SPT_OK;
}

}  // namespace

spt_status launch_render(spt_ctx* ctx, const spt_camera* cam, const spt_render_params* prm, float* d_rgb32,
                         double* d_rgb64, float* d_path_rgb) {
    auto t0 = std::chrono::steady_clock::now();
    RenderArgs a{};
    a.sc = ctx->scene();
    uint64_t total = 0;
    spt_status s = make_job(ctx, cam, prm, d_rgb32, d_rgb64, d_path_rgb, a.job, total);
    if (s != SPT_OK) return s;
    for (int i = 0; i < 3; ++i) {
        a.job.rmin[i] = ctx->root_box[i];
        a.job.rmax[i] = ctx->root_box[3 + i];
    }
    a.job.root_leaf = ctx->root_leaf;
    a.flags = prm->flags;
    a.per_path = a.job.path_rgb ? 1u : 0u;
    a.guided = (prm->flags & SPT_RENDER_GUIDED) ? 1u : 0u;
    a.train = (a.guided && (prm->flags & SPT_RENDER_GUIDE_TRAIN)) ? 1u : 0u;
    if ((prm->flags & SPT_RENDER_GUIDE_TRAIN) && !a.guided)
        return set_error(ctx, SPT_ERR_INVALID, "GUIDE_TRAIN needs GUIDED");
    if (a.guided && !ctx->guide.ready) return set_error(ctx, SPT_ERR_STATE, "guided rendering needs spt_guide_init");
    a.guidable = ctx->d_guidable;
    a.theta = ctx->guide.d_theta;
    a.wbf = ctx->guide.d_wbf;
    a.gbox = guide_box(ctx);
    a.bsdf_frac = ctx->guide.bsdf_fraction;
    a.gflags = ctx->guide.flags;
    a.records = ctx->guide.d_records;
    a.rec_cap = ctx->guide.records_cap;
    // record about records_cap of this render's candidate vertices, estimated
    // from the candidates per path of the previous training render. With
    // in-call training the ring must not wrap between the count a training
    // round reads and the moment its steps run: the host runs up to
    // check_every bounces ahead of the device, so one round spans
    // train_every + check_every + 1 bounces, and that span writes at most a
    // quarter of the ring (a bounce of a full queue meets about capacity /
    // rays-per-path paths); the host skips batches older than half a ring
    a.rec_ring = (a.train && ctx->guide.train_every) ? 1u : 0u;
    if (ctx->guide.cand_per_path > 0.0 && ctx->guide.records_cap > 0) {
        double want = ctx->guide.cand_per_path * (double)total;
        if (a.rec_ring) {
            const spt_tuning tn = resolve_tuning(ctx->tuning);
            const double per_bounce = ctx->guide.cand_per_path / ctx->guide.rays_per_path *
                                      (double)std::min<uint64_t>(total, tn.path_capacity);
            want = per_bounce * (double)(ctx->guide.train_every + tn.check_every + 1) * 4.0;
        }
        ctx->guide.rec_stride =
            (uint32_t)std::min<double>(1 << 30, std::max(1.0, std::ceil(want / ctx->guide.records_cap)));
    }
    a.rec_stride = ctx->guide.rec_stride;
    const uint32_t keys = ctx->n_mats + 1;
    a.keys = (!(prm->flags & SPT_RENDER_NO_SORT) && keys <= kMaxSortKeys) ? keys : 0u;
    spt_render_stats stats{};
    stats.paths = total;
    s = run_fronts<false>(ctx, a, total, resolve_tuning(ctx->tuning), stats);
    stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->stats = stats;
    return s;
}

spt_status launch_batch(spt_ctx* ctx, BatchData& b, uint32_t sample_div) {
    auto t0 = std::chrono::steady_clock::now();
    const spt_tuning tn = resolve_tuning(ctx->tuning);
    const uint32_t nj = (uint32_t)b.jobs.size();
    std::vector<Job> jobs = b.jobs;
    for (Job& jb : jobs) {
        if (sample_div > 1) {  // tuning trial: a share of the samples, outputs untouched
            jb.spp_count = std::max<uint32_t>(1, jb.spp_count / sample_div);
            jb.film32 = nullptr;
            jb.film64 = nullptr;
            jb.path_rgb = nullptr;
        }
        jb.n_paths = (uint64_t)jb.vw * jb.vh * jb.spp_count;
    }
    // generation tiles, taken round-robin from the jobs; within a job samples
    // advance slowest, so every pixel gets its first samples early
    std::vector<std::vector<BatchTile>> per(nj);
    for (uint32_t j = 0; j < nj; ++j) {
        const Job& jb = jobs[j];
        const uint32_t npix = jb.vw * jb.vh;
        for (uint32_t s0 = 0; s0 < jb.spp_count; s0 += tn.tile_samples)
            for (uint32_t p0 = 0; p0 < npix; p0 += tn.tile_pixels) {
                BatchTile t{};
                t.job = j;
                t.p0 = p0;
                t.np = std::min(tn.tile_pixels, npix - p0);
                t.s0 = jb.spp_begin + s0;
                t.ns = std::min(tn.tile_samples, jb.spp_count - s0);
                per[j].push_back(t);
            }
    }
    std::vector<BatchTile> tiles;
    std::vector<uint64_t> start;
    uint64_t total = 0;
    for (size_t r = 0;; ++r) {
        bool any = false;
        for (uint32_t j = 0; j < nj; ++j) {
            if (r >= per[j].size()) continue;
            any = true;
            tiles.push_back(per[j][r]);
            start.push_back(total);
            total += (uint64_t)per[j][r].np * per[j][r].ns;
        }
        if (!any) break;
    }
    start.push_back(total);
    if (tiles.empty() || tiles.size() >= (1u << 31)) return set_error(ctx, SPT_ERR_INVALID, "empty or oversized batch");
    if (tiles.size() > b.tiles_cap) {
        if (b.d_tiles) (void)hipFree(b.d_tiles);
        if (b.d_tile_start) (void)hipFree(b.d_tile_start);
        b.d_tiles = nullptr;
        b.d_tile_start = nullptr;
        b.tiles_cap = 0;
        spt_status s = check_hip(ctx, hipMalloc(&b.d_tiles, tiles.size() * sizeof(BatchTile)), "batch tiles");
        if (s == SPT_OK)
            s = check_hip(ctx, hipMalloc(&b.d_tile_start, (tiles.size() + 1) * sizeof(uint64_t)), "batch tiles");
        if (s != SPT_OK) return s;
        b.tiles_cap = tiles.size();
    }
    hipStream_t st = ctx->stream;
    spt_status s = check_hip(ctx, hipMemcpyAsync(b.d_tiles, tiles.data(), tiles.size() * sizeof(BatchTile),
                                                 hipMemcpyHostToDevice, st), "tiles upload");
    if (s == SPT_OK)
        s = check_hip(ctx, hipMemcpyAsync(b.d_tile_start, start.data(), start.size() * sizeof(uint64_t),
                                          hipMemcpyHostToDevice, st), "tiles upload");
    if (s == SPT_OK)
        s = check_hip(ctx, hipMemcpyAsync(b.d_jobs, jobs.data(), nj * sizeof(Job), hipMemcpyHostToDevice, st),
                      "jobs upload");
    if (s == SPT_OK) s = check_hip(ctx, hipMemsetAsync(b.d_done, 0, nj * sizeof(uint32_t), st), "job counters");
    if (s == SPT_OK) s = check_hip(ctx, hipMemsetAsync(b.d_clock, 0, (nj + 1) * sizeof(uint64_t), st), "job clocks");
    if (s == SPT_OK) s = check_hip(ctx, hipStreamSynchronize(st), "batch upload");  // host vectors go out of scope
    if (s != SPT_OK) return s;

    RenderArgs a{};
    a.sc.nodes = b.d_nodes;
    a.sc.nodesq = reinterpret_cast<const uint4*>(b.d_nodesq);
    a.sc.leaves = reinterpret_cast<const float4*>(b.d_leaves);
    a.sc.stack_cap4 = b.stack_cap4;
    a.sc.tris = b.d_tris;
    a.sc.shading = b.d_shading;
    a.sc.mats = b.d_mats;
    a.sc.n_tris = b.n_tris;
    a.sc.n_nodes = b.n_nodes;
    a.sc.n_mats = b.n_mats;
    a.sc.stack_cap = b.stack_cap;
    a.sc.has_spheres = b.has_spheres ? 1u : 0u;
    a.jobs = b.d_jobs;
    a.tiles = b.d_tiles;
    a.tile_start = b.d_tile_start;
    a.n_tiles = (uint32_t)tiles.size();
    a.job_done = b.d_done;
    a.job_clock = b.d_clock;
    a.flags = 0;
    for (const Job& jb : jobs) a.per_path |= jb.path_rgb ? 1u : 0u;
    const uint32_t keys = b.n_mats + 1;
    a.keys = keys <= kMaxSortKeys ? keys : 0u;
    spt_render_stats stats{};
    stats.paths = total;
    s = run_fronts<true>(ctx, a, total, tn, stats);
    if (s != SPT_OK) return s;
    std::vector<uint64_t> clk(nj + 1, 0);
    std::vector<uint32_t> done(nj, 0);
    s = check_hip(ctx, hipMemcpy(clk.data(), b.d_clock, clk.size() * sizeof(uint64_t), hipMemcpyDeviceToHost),
                  "job clocks");
    if (s == SPT_OK)
        s = check_hip(ctx, hipMemcpy(done.data(), b.d_done, done.size() * sizeof(uint32_t), hipMemcpyDeviceToHost),
                      "job counters");
    if (s != SPT_OK) return s;
    b.scene_ms.assign(nj, 0.0);
    for (uint32_t j = 0; j < nj; ++j) {
        if (done[j] != jobs[j].n_paths) return set_error(ctx, SPT_ERR_DEVICE, "a scene did not finish all its paths");
        b.scene_ms[j] = (double)(clk[1 + j] - clk[0]) / (double)ctx->wall_clock_khz;
    }
    b.stats = spt_batch_stats{};
    b.stats.scenes = nj;
    b.stats.paths = total;
    b.stats.rays = stats.rays;
    b.stats.bounces = stats.bounces;
    b.stats.tiles = tiles.size();
    b.stats.triangles = b.n_tris;
    b.stats.device_bytes = b.device_bytes;
    b.stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->stats = stats;
    ctx->stats.ms = b.stats.ms;
    return SPT_OK;
}

}  // namespace spt
