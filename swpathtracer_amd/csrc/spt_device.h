// spt_device.h - device-side scene layout, context and closest-hit traversal.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "spt.h"
#include "spt_internal.h"
#include "spt_math.h"

namespace spt {

constexpr int kBlock = 256;       // threads per block of the ray kernels (4 wave64)
// normal kind of a shading record (word 10): flat n0, smooth n0 n1 n2, sphere centre n0
constexpr uint32_t kShadeFlat = 0, kShadeSmooth = 1, kShadeSphere = 2;
constexpr int kLdsStack = 16;     // traversal stack entries per lane kept in LDS
constexpr uint32_t kNoHit = SPT_MISS;

// Four-wide node of the throughput traversal, collapsed from the binary
// tree (children are binary-tree nodes or leaves two levels down, with their
// own padded boxes): child boxes as six float4 (lo.x of the 4 children, hi.x,
// lo.y, hi.y, lo.z, hi.z), then the child words and leaf counts.
struct Node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t child[4];  // inner: Node4 index; leaf: first triangle; unused: kNoChild
    uint32_t count[4];  // leaf: triangle count (> 0); inner or unused: 0
};
static_assert(sizeof(Node4) == 128, "Node4 layout");
constexpr uint32_t kNoChild = 0xFFFFFFFFu;

// Compressed four-wide node (64 B, what the throughput traversal reads):
// each axis has a per-node grid origin + q * 2^e, and every child bound is
// rounded outward to a byte q (lo down, hi up), so the decoded boxes contain
// the exact padded child boxes (the traversal visits a superset of the exact
// four-wide tree's boxes; results do not depend on visit order).
//   word 0-2: grid origin x y z (float); word 3: biased exponent bytes of
//   2^e_x, 2^e_y, 2^e_z (float exponent field, bytes 0-2)
//   words 4-9: q of lo.x, hi.x, lo.y, hi.y, lo.z, hi.z (byte i = child i)
//   words 10-11: 16-bit triangle counts of the children (0: inner or unused;
//   0xFFFF: 65535 or more, the full count is in the leaf's box record)
//   words 12-15: children: NodeQ index (inner), first triangle (leaf), kNoChild
// A leaf's range is read from the node that holds it, which the traversal
// has just visited; its exact box (the reference's hit-point-inside-leaf
// test) sits in a side array at 4 * node + slot and is read only when a
// triangle of the leaf passes the t tests.
struct NodeQ {
    float ox, oy, oz;
    uint32_t exps;
    uint32_t q[6];
    uint32_t cnt[2];
    uint32_t child[4];
};
static_assert(sizeof(NodeQ) == 64, "NodeQ layout");

// Exact padded box of the leaf in slot s of node i, at index 4 i + s (32 B).
struct LeafRec {
    float mn[3];
    uint32_t cnt;  // full triangle count
    float mx[3];
    uint32_t pad;
};
static_assert(sizeof(LeafRec) == 32, "LeafRec layout");

// Materials the learned guide samples: one Lambertian / importance-sampled
// GGX layer (kGuideSingle), or such a layer under one smooth dielectric coat
// of IOR >= 1 (kGuideCoated: the guide drives the base layer's directions).
constexpr uint8_t kGuideSingle = 1, kGuideCoated = 2;

// Scene as the kernels see it.
struct DevScene {
    const uint4* nodesq;    // NodeQ x 4 uint4, persistent traversal
    const float4* leaves;   // LeafRec x 2 float4, 4 per node (leaf slots used)
    const float4* nodes;    // n_nodes x 4 float4 (spt_node64)
    const float4* tris;     // n_tris x 3 float4 (spt_tri48; .pad[0] = material id)
    const float4* shading;  // n_tris x 3 float4: n0 n1 n2, material, smooth flag, pad
    const spt_material* mats;
    float rmin[3], rmax[3];
    uint32_t root_leaf;  // > 0: the root is a leaf holding this many triangles
    uint32_t n_tris;
    uint32_t n_nodes;
    uint32_t n_mats;
    uint32_t stack_cap;  // traversal stack entries a lane may need (inner-node depth + 1)
    uint32_t stack_cap4; // same for the four-wide tree (3 per level + 1)
    uint32_t has_spheres;  // some primitive record is a sphere (uniform branch in the leaf loops)
};

// Closest-hit test of one primitive record (a, b, c = its three float4):
// Moller-Trumbore for a triangle, Hultquist for a sphere (centre a.xyz,
// radius a.w, kind in c.w). u, v are the triangle's barycentrics (0 for a
// sphere). `spheres` is the scene's uniform flag: without spheres the kind is
// never read.
__device__ inline bool prim_test(bool spheres, V3 o, V3 d, float4 a, float4 b, float4 c, float& t, float& u,
                                 float& v) {
    if (spheres && __float_as_uint(c.w) == SPT_PRIM_SPHERE) {
        u = v = 0.0f;
        return sphere_test(o, d, V3{a.x, a.y, a.z}, a.w, t);
    }
    return tri_test(o, d, V3{a.x, a.y, a.z}, V3{a.w, b.x, b.y}, V3{b.z, b.w, c.x}, t, u, v);
}

// Closest-hit result of one ray.
struct Hit {
    float t;
    uint32_t idx;   // triangle index in leaf order (kNoHit: miss)
    uint32_t prim;  // primitive id in scene order
    uint32_t mat;
    float u, v;
};

// Traversal stack: the first kLdsStack entries of a lane live in LDS
// (layout [entry][lane], conflict-free), deeper entries spill to a global
// buffer laid out [entry - kLdsStack][thread] so a wave's accesses coalesce.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(1))) uint32_t glb_u32;

struct Stack {
    lds_u32* lds;       // &lds[threadIdx.x], stride kBlock
    glb_u32* spill;     // &spill[global thread], stride spill_stride
    uint32_t spill_stride;
    uint32_t cap;       // total entries available
    // Address-space-qualified pointers keep LDS and global accesses as ds_* and
    // global_* instructions instead of one merged flat access.
    __device__ Stack(uint32_t* lds_lane, uint32_t* spill_lane, uint32_t stride, uint32_t capacity)
        : lds((lds_u32*)lds_lane), spill((glb_u32*)spill_lane), spill_stride(stride), cap(capacity) {}
    __device__ void put(int i, uint32_t v) {
        if (i < kLdsStack)
            lds[i * kBlock] = v;
        else
            spill[(size_t)(i - kLdsStack) * spill_stride] = v;
    }
    __device__ uint32_t get(int i) const {
        if (i < kLdsStack) return lds[i * kBlock];
        return spill[(size_t)(i - kLdsStack) * spill_stride];
    }
};

// Leaf: tests triangles [first, first+cnt) against the ray. A candidate is
// accepted exactly when the reference accepts it (sbvh.cpp:494-496: t > 0,
// t < best, hit point inside the leaf box +-1e-3); among equal t the smaller
// leaf-order index wins, which is the reference's depth-first left-first
// first-come rule (strict '<').
template <bool COUNT, bool FETCH>
__device__ inline void leaf_test(const DevScene& sc, V3 o, V3 d, uint32_t first, uint32_t cnt, V3 bmn, V3 bmx,
                                 Hit& h, uint32_t& tests, uint32_t& isects) {
    if (FETCH) isects += cnt;  // triangle records fetched
    for (uint32_t k = first; k < first + cnt; ++k) {
        const float4* tp = sc.tris + 3 * (size_t)k;
        const float4 a = tp[0], b = tp[1], c = tp[2];
        if (COUNT) ++tests;
        float t, u, v;
        if (prim_test(sc.has_spheres != 0u, o, d, a, b, c, t, u, v)) {
            if (COUNT) ++isects;
            if ((t > 0.0f) && ((t < h.t) || ((t == h.t) && (k < h.idx)))) {
                V3 p = o + d * t;
                if (point_inside(p, bmn, bmx)) {
                    h.t = t;
                    h.idx = k;
                    h.prim = __float_as_uint(c.y);
                    h.mat = __float_as_uint(c.z);
                    h.u = u;
                    h.v = v;
                }
            }
        }
    }
}

// Closest hit of one ray (acceleration/sbvh.cpp:423-515 semantics).
//   PRUNE = true : near-first order, children skipped when their slab interval
//                  starts after the best hit or ends behind the origin. Child
//                  boxes are padded by 1e-3 around their triangles, so a
//                  skipped box cannot hold an accepted candidate; the result
//                  equals the reference's.
//   PRUNE = false: visits exactly the nodes the reference visits (every box
//                  the ray's line meets) and, with COUNT, reproduces its
//                  num_intersection_tests / num_intersections.
//   FETCH = true : (without COUNT) tests/isects receive the number of 64-B
//                  node records and 48-B triangle records this ray fetched.
template <bool PRUNE, bool COUNT, bool FETCH = false>
__device__ inline void trace_ray(const DevScene& sc, V3 o, V3 d, Stack& st, Hit& h, uint32_t& tests,
                                 uint32_t& isects, bool& overflow) {
    h.t = __builtin_inff();
    h.idx = kNoHit;
    h.prim = kNoHit;
    h.mat = 0;
    h.u = h.v = 0.0f;
    tests = 0;
    isects = 0;
    const V3 inv{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float tn, tf;
    const V3 rmn{sc.rmin[0], sc.rmin[1], sc.rmin[2]}, rmx{sc.rmax[0], sc.rmax[1], sc.rmax[2]};
    if (COUNT) tests = 1;
    if (!slab(o, inv, rmn, rmx, tn, tf)) return;
    if (COUNT) isects = 1;
    if (sc.root_leaf) {
        leaf_test<COUNT, FETCH>(sc, o, d, 0, sc.root_leaf, rmn, rmx, h, tests, isects);
        return;
    }
    uint32_t node = 0;
    int sp = 0;
    while (true) {
        const float4* np = sc.nodes + 4 * (size_t)node;
        const float4 a = np[0], b = np[1], c = np[2];
        const uint4 r = reinterpret_cast<const uint4*>(np)[3];
        if (FETCH) ++tests;
        const V3 lmn{a.x, a.y, a.z}, lmx{a.w, b.x, b.y}, rmn2{b.z, b.w, c.x}, rmx2{c.y, c.z, c.w};
        float tnl = 0.0f, tfl = 0.0f, tnr = 0.0f, tfr = 0.0f;
        bool hl = slab(o, inv, lmn, lmx, tnl, tfl);
        bool hr = slab(o, inv, rmn2, rmx2, tnr, tfr);
        if (COUNT) {
            tests += 2;
            isects += (uint32_t)hl + (uint32_t)hr;
        }
        if (PRUNE) {
            hl = hl && !(tnl > h.t) && !(tfl < 0.0f);
            hr = hr && !(tnr > h.t) && !(tfr < 0.0f);
        }
        if (hl && r.y) {
            leaf_test<COUNT, FETCH>(sc, o, d, r.x, r.y, lmn, lmx, h, tests, isects);
            hl = false;
        }
        if (hr && r.w) {
            leaf_test<COUNT, FETCH>(sc, o, d, r.z, r.w, rmn2, rmx2, h, tests, isects);
            hr = false;
        }
        if (hl && hr) {
            const bool rfirst = PRUNE && (tnr < tnl);
            if (sp >= (int)st.cap) {
                overflow = true;
                return;
            }
            st.put(sp++, rfirst ? r.x : r.z);
            node = rfirst ? r.z : r.x;
        } else if (hl) {
            node = r.x;
        } else if (hr) {
            node = r.z;
        } else {
            if (sp == 0) break;
            node = st.get(--sp);
        }
    }
}

// ---------------------------------------------------------------------------
// Persistent closest-hit traversal (the throughput path).
//
// Each wave loops until the ray pool is empty: idle lanes take the next rays
// with one wave-aggregated atomic, then every lane runs an inner-node phase
// (descend near-first, push the far child) followed by a leaf phase (test the
// triangles of postponed and consecutive leaves). Leaves are stack entries
// too, encoded as kLeafBit | node << 2 | slot. The inner phase is
// speculative (Aila and Laine, "Understanding the efficiency of ray traversal
// on GPUs", 2009): a lane that meets a leaf postpones it and keeps
// descending until every lane holds a leaf, so lanes wait less for each
// other; a lane holds up to two postponed leaves before it waits (faster than
// one on the 1M-triangle path-tracing workload). Nodes are the compressed NodeQ; a leaf's range is reread from its
// node and its exact box (the reference's hit-point-inside-leaf test) is
// read only when one of its triangles passes the t tests.
// Same candidate acceptance and tie rule as trace_ray<true, false>: the
// closest hit is the minimum of (t, leaf index), independent of visit order,
// and boxes are only skipped when they start beyond the best t found so far.
// ---------------------------------------------------------------------------
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kDone = 0xFFFFFFFFu;

// Where traversal starts: the root box, and either a root leaf of `leaf`
// triangles at tri0 or the inner node node0. One per scene; rays of several
// scenes share the node and triangle arrays (indices rebased per scene).
struct Root {
    V3 mn, mx;
    uint32_t leaf, node0, tri0;
};

__device__ inline Root scene_root(const DevScene& sc) {
    return Root{V3{sc.rmin[0], sc.rmin[1], sc.rmin[2]}, V3{sc.rmax[0], sc.rmax[1], sc.rmax[2]}, sc.root_leaf, 0u, 0u};
}

// Reciprocal for the conservative box tests: a zero component gives a large
// finite value, so (lo - o) / d products never form inf - inf.
__device__ inline float safe_rcp(float d) {
    return (fabsf(d) > 1e-30f) ? 1.0f / d : copysignf(1e30f, d);
}

// One child box against the ray: key = entry distance (+inf when the box is
// missed, lies behind the origin or beyond the best hit t_best, or the slot
// is unused). The test is one FMA per slab plane and IEEE min/max (a NaN
// product keeps the box): conservative, it may visit a box the reference's
// exact slab test skips but never skips one holding a closer candidate,
// since every box is padded 1e-3 around its triangles
// (acceleration/sbvh.cpp:409-412) and the compressed boxes contain the
// padded ones.
__device__ inline float child_test(V3 inv, V3 oi, float t_best, float lox, float hix, float loy, float hiy, float loz,
                                   float hiz, uint32_t child) {
    const float t0x = fmaf(lox, inv.x, oi.x), t1x = fmaf(hix, inv.x, oi.x);
    const float t0y = fmaf(loy, inv.y, oi.y), t1y = fmaf(hiy, inv.y, oi.y);
    const float t0z = fmaf(loz, inv.z, oi.z), t1z = fmaf(hiz, inv.z, oi.z);
    const float tmin = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), 0.0f));
    const float tmax = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), t_best));
    return (child != kNoChild && tmin <= tmax) ? tmin : __builtin_inff();
}

// Byte b of word w as a float.
__device__ inline float qbyte(uint32_t w, int b) { return (float)((w >> (8 * b)) & 0xFFu); }

// Slab test of one compressed child in ray space: for a plane at origin +
// q * scale, t = (origin + q * scale - o) / d = q * (scale / d) + (origin / d
// - o / d), one FMA per plane with a = scale * inv and b = origin * inv + oi
// per node and axis. Rounding differs from the world-space form by a few
// ulps of the same terms, far inside the 1e-3 box padding. The planes come
// in near/far order for the ray: t = q * a + b is monotonic in q, so for
// a >= 0 (a has the sign of the ray's reciprocal) the low plane is the near
// one and for a < 0 the high one, which gives the same interval as taking
// the per-axis min and max of the two plane values.
__device__ inline float child_test_q(float ax, float bx, float ay, float by, float az, float bz, float t_best,
                                     float qnx, float qfx, float qny, float qfy, float qnz, float qfz,
                                     uint32_t child) {
    const float tnx = fmaf(qnx, ax, bx), tfx = fmaf(qfx, ax, bx);
    const float tny = fmaf(qny, ay, by), tfy = fmaf(qfy, ay, by);
    const float tnz = fmaf(qnz, az, bz), tfz = fmaf(qfz, az, bz);
    const float tmin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
    const float tmax = fminf(fminf(tfx, tfy), fminf(tfz, t_best));
    return (child != kNoChild && tmin <= tmax) ? tmin : __builtin_inff();
}

__device__ inline void cswap(float k[4], uint32_t e[4], int a, int b) {
    const bool sw = k[b] < k[a];
    const float ka = k[a], kb = k[b];
    const uint32_t ea = e[a], eb = e[b];
    k[a] = sw ? kb : ka;
    k[b] = sw ? ka : kb;
    e[a] = sw ? eb : ea;
    e[b] = sw ? ea : eb;
}

template <bool FETCH>
struct PLane {
    V3 o, d;
    V3 inv4, oi;    // box tests: finite reciprocal, -o * inv4
    Hit h;
    uint32_t cur;   // node index, leaf entry, or kDone
    uint32_t leaf;  // postponed leaf entry, or kDone
    uint32_t leaf2;  // a second postponed leaf entry, or kDone
    int sp;
    uint32_t ray;   // index in the batch (kDone: idle)
    uint32_t nodes, tris;  // FETCH counters
};

// GetRay(i, o, d, root) loads ray i (and may set the root of its scene; it
// starts as the single scene's root); PutHit(i, h, nodes, tris) stores its result.
template <bool FETCH, typename GetRay, typename PutHit>
__device__ inline void trace_persistent(const DevScene& sc, uint32_t n, uint32_t* ray_counter, Stack& st,
                                        bool& overflow, uint32_t refill, GetRay get_ray, PutHit put_hit) {
    Root R = scene_root(sc);
    const uint64_t lt = __lanemask_lt();
    PLane<FETCH> L;
    L.ray = kDone;
    L.cur = kDone;
    L.leaf = kDone;
    L.leaf2 = kDone;
    L.sp = 0;
    bool pool_empty = false;
    while (true) {
        // ---- refill idle lanes
        const uint64_t idle = __ballot(L.ray == kDone);
        if (!pool_empty && (uint32_t)__popcll(idle) >= refill) {
            uint32_t base = 0;
            const int leader = __ffsll((unsigned long long)idle) - 1;
            if (__lane_id() == (unsigned)leader) base = atomicAdd(ray_counter, (uint32_t)__popcll(idle));
            base = __shfl(base, leader);
            if (base + __popcll(idle) >= n) pool_empty = true;
            if (L.ray == kDone) {
                const uint32_t i = base + (uint32_t)__popcll(idle & lt);
                if (i < n) {
                    L.ray = i;
                    get_ray(i, L.o, L.d, R);
                    L.inv4 = V3{safe_rcp(L.d.x), safe_rcp(L.d.y), safe_rcp(L.d.z)};
                    L.oi = V3{-L.o.x * L.inv4.x, -L.o.y * L.inv4.y, -L.o.z * L.inv4.z};
                    L.h.t = __builtin_inff();
                    L.h.idx = kNoHit;
                    L.h.prim = kNoHit;
                    L.h.mat = 0;
                    L.h.u = L.h.v = 0.0f;
                    L.sp = 0;
                    L.leaf = kDone;
                    L.leaf2 = kDone;
                    if (FETCH) L.nodes = L.tris = 0;
                    // a ray with a non-finite component cannot hit anything (every
                    // t comparison fails) but would pass every box test: a miss
                    const float sum = L.o.x + L.o.y + L.o.z + L.d.x + L.d.y + L.d.z;
                    const float troot = child_test(L.inv4, L.oi, __builtin_inff(), R.mn.x, R.mx.x, R.mn.y, R.mx.y,
                                                   R.mn.z, R.mx.z, 0u);
                    if (troot == __builtin_inff() || !(sum - sum == 0.0f))
                        L.cur = kDone;
                    else
                        L.cur = R.leaf ? kLeafBit : R.node0;
                }
            }
        }
        if (__ballot(L.ray != kDone) == 0) {
            if (pool_empty) break;
            continue;
        }
        // ---- inner-node phase, speculative: a lane that meets a leaf
        // postpones it (up to two) and keeps descending; the phase ends when no lane is
        // still searching for a leaf (or none has inner work left)
        while (true) {
            if (L.ray != kDone && L.cur != kDone && !(L.cur & kLeafBit)) {
                const uint4* np = sc.nodesq + 4 * (size_t)L.cur;
                const uint4 g = np[0], qa = np[1], qb = np[2], ch = np[3];
                if (FETCH) L.nodes += 2;  // 32-B units
                const float ox = __uint_as_float(g.x), oy = __uint_as_float(g.y), oz = __uint_as_float(g.z);
                const float sx = __uint_as_float((g.w & 0xFFu) << 23), sy = __uint_as_float(((g.w >> 8) & 0xFFu) << 23),
                            sz = __uint_as_float(((g.w >> 16) & 0xFFu) << 23);
                const float ax = sx * L.inv4.x, ay = sy * L.inv4.y, az = sz * L.inv4.z;
                const float bx = fmaf(ox, L.inv4.x, L.oi.x), by = fmaf(oy, L.inv4.y, L.oi.y),
                            bz = fmaf(oz, L.inv4.z, L.oi.z);
                // plane words in near/far order for the ray's direction signs
                const bool px = L.inv4.x >= 0.0f, py = L.inv4.y >= 0.0f, pz = L.inv4.z >= 0.0f;
                const uint32_t nx = px ? qa.x : qa.y, fx = px ? qa.y : qa.x;
                const uint32_t ny = py ? qa.z : qa.w, fy = py ? qa.w : qa.z;
                const uint32_t nz = pz ? qb.x : qb.y, fz = pz ? qb.y : qb.x;
                float k[4];
                uint32_t e[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    k[c] = child_test_q(ax, bx, ay, by, az, bz, L.h.t, qbyte(nx, c), qbyte(fx, c), qbyte(ny, c),
                                        qbyte(fy, c), qbyte(nz, c), qbyte(fz, c), e[c]);
                    const uint32_t cw = (c < 2) ? qb.z : qb.w;
                    if ((cw >> (16 * (c & 1))) & 0xFFFFu) e[c] = kLeafBit | (L.cur << 2) | (uint32_t)c;
                }
                // nearest first: sort the four (entry, entry distance) pairs, misses (+inf) last
                cswap(k, e, 0, 1);
                cswap(k, e, 2, 3);
                cswap(k, e, 0, 2);
                cswap(k, e, 1, 3);
                cswap(k, e, 1, 2);
                const float inf = __builtin_inff();
                if (k[0] == inf) {
                    L.cur = (L.sp > 0) ? st.get(--L.sp) : kDone;
                } else {
                    const int need = (int)(k[1] < inf) + (int)(k[2] < inf) + (int)(k[3] < inf);
                    if (L.sp + need > (int)st.cap) {
                        overflow = true;
                        L.cur = kDone;
                        L.leaf = kDone;
                        L.leaf2 = kDone;
                        L.sp = 0;
                    } else {
                        if (k[3] < inf) st.put(L.sp++, e[3]);
                        if (k[2] < inf) st.put(L.sp++, e[2]);
                        if (k[1] < inf) st.put(L.sp++, e[1]);
                        L.cur = e[0];
                    }
                }
            }
            // up to two leaves are postponed; a lane waits at its third
            if (L.ray != kDone && L.cur != kDone && (L.cur & kLeafBit)) {
                if (L.leaf == kDone) {
                    L.leaf = L.cur;
                    L.cur = (L.sp > 0) ? st.get(--L.sp) : kDone;
                } else if (L.leaf2 == kDone) {
                    L.leaf2 = L.cur;
                    L.cur = (L.sp > 0) ? st.get(--L.sp) : kDone;
                }
            }
            const bool inner = L.ray != kDone && L.cur != kDone && !(L.cur & kLeafBit);
            if (!__any(inner && L.leaf == kDone)) break;
        }
        // ---- leaf phase: the postponed leaves, then consecutive leaf entries
        while (true) {
            if (L.ray != kDone && L.leaf == kDone) {
                if (L.leaf2 != kDone) {
                    L.leaf = L.leaf2;
                    L.leaf2 = kDone;
                } else if (L.cur != kDone && (L.cur & kLeafBit)) {
                    L.leaf = L.cur;
                    L.cur = (L.sp > 0) ? st.get(--L.sp) : kDone;
                }
            }
            const bool has = L.ray != kDone && L.leaf != kDone;
            if (!__any(has)) break;
            if (has) {
                uint32_t first, cnt;
                const float4* box = nullptr;  // leaf box record, read when a candidate needs it
                V3 bmn = R.mn, bmx = R.mx;
                if (R.leaf) {
                    first = R.tri0;
                    cnt = R.leaf;
                } else {
                    // the range from the node that holds the leaf (visited just before)
                    const uint32_t slot = L.leaf & 3u, li = L.leaf & ~kLeafBit;
                    const uint32_t* np = reinterpret_cast<const uint32_t*>(sc.nodesq + 4 * (size_t)(li >> 2));
                    first = np[12 + slot];
                    cnt = (np[10 + (slot >> 1)] >> (16 * (slot & 1u))) & 0xFFFFu;
                    box = sc.leaves + 2 * (size_t)li;
                    if (cnt == 0xFFFFu) cnt = __float_as_uint(box[0].w);
                }
                bool have_box = R.leaf != 0u;
                // software-pipelined: triangle k + 1 is requested before k is tested
                // (every leaf holds at least one triangle)
                const uint32_t end = first + cnt;
                const float4* tp = sc.tris + 3 * (size_t)first;
                float4 na = tp[0], nb = tp[1], nc = tp[2];
                for (uint32_t k = first; k < end; ++k) {
                    const float4 a = na, b = nb, c = nc;
                    if (k + 1 < end) {
                        tp += 3;
                        na = tp[0];
                        nb = tp[1];
                        nc = tp[2];
                    }
                    float t, u, v;
                    if (prim_test(sc.has_spheres != 0u, L.o, L.d, a, b, c, t, u, v) && (t > 0.0f) &&
                        ((t < L.h.t) || ((t == L.h.t) && (k < L.h.idx)))) {
                        if (!have_box) {
                            const float4 b0 = box[0], b1 = box[1];
                            bmn = V3{b0.x, b0.y, b0.z};
                            bmx = V3{b1.x, b1.y, b1.z};
                            have_box = true;
                            if (FETCH) ++L.nodes;
                        }
                        // leaf_test's acceptance (sbvh.cpp:494-496)
                        if (point_inside(L.o + L.d * t, bmn, bmx)) {
                            L.h.t = t;
                            L.h.idx = k;
                            L.h.prim = __float_as_uint(c.y);
                            L.h.mat = __float_as_uint(c.z);
                            L.h.u = u;
                            L.h.v = v;
                        }
                    }
                }
                if (FETCH) L.tris += cnt;
                L.leaf = kDone;
            }
        }
        // ---- finished rays
        if (L.ray != kDone && L.cur == kDone && L.leaf == kDone && L.leaf2 == kDone) {
            put_hit(L.ray, L.h, FETCH ? L.nodes : 0u, FETCH ? L.tris : 0u);
            L.ray = kDone;
        }
    }
}

// Camera ray through image point (x, y): PinholeCamera::getRay
// (cameras/pinhole_camera.cpp:14-18), OrthographicCamera::getRay
// (cameras/orthographic_camera.cpp:21-25).
__device__ inline void camera_ray(const DevCamera& c, float x, float y, V3& o, V3& d) {
    if (c.type == SPT_CAMERA_PINHOLE) {
        V3 ld = normalize(c.ulc + V3{x, -y, 0.0f});
        o = c.position;
        d = to_world(c.frame, ld);
    } else {
        V3 lo = c.ulc + V3{x * c.pdx, y * -c.pdy, 0.0f};
        o = to_world(c.frame, lo) + c.position;
        d = c.ortho_dir;
    }
}

// Shading normal of a hit at pos: SmoothTriangle::getNormal
// (primitives/smooth_triangle.h:11-13), the stored flat normal
// (primitives/triangle.cpp:5-6) or a sphere's normalize(pos - centre)
// (primitives/sphere.cpp:35).
__device__ inline V3 shading_normal(const DevScene& sc, const Hit& h, V3 pos) {
    const float4* sp = sc.shading + 3 * (size_t)h.idx;
    const float4 a = sp[0], b = sp[1], c = sp[2];
    const V3 n0{a.x, a.y, a.z};
    const uint32_t kind = __float_as_uint(c.z);
    if (kind == kShadeFlat) return n0;
    if (kind == kShadeSphere) return normalize(pos - n0);
    const V3 n1{a.w, b.x, b.y}, n2{b.z, b.w, c.x};
    return normalize(n0 * (1.0f - h.u - h.v) + n1 * h.u + n2 * h.v);
}

// Wave-aggregated counter increment: lanes of the wave that carry the same
// key share one global atomicAdd (one per distinct key per wave instead of one
// per lane); returns this lane's slot (old value + rank among equal keys).
// Must be called by every active lane of the wave together.
__device__ inline uint32_t wave_add(uint32_t* counters, uint32_t key) {
    uint64_t todo = __ballot(1);
    const uint64_t lt = __lanemask_lt();
    uint32_t slot = 0;
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lkey = __shfl(key, leader);
        const uint64_t group = __ballot(key == lkey) & todo;
        uint32_t base = 0;
        if (__lane_id() == (unsigned)leader) base = atomicAdd(&counters[lkey], (uint32_t)__popcll(group));
        base = __shfl(base, leader);
        if (key == lkey) slot = base + (uint32_t)__popcll(group & lt);
        todo &= ~group;
    }
    return slot;
}

// One scene's share of a render call: camera, sampling, termination,
// outputs and where its tree starts in the (shared) node/triangle arrays.
struct Job {
    DevCamera cam;
    uint32_t sampler, n_sub, spp, spp_begin, spp_count;
    float w_sub;
    int32_t max_depth, rr_depth;
    uint64_t seed;
    uint32_t vt, vl, vw, vh;
    float bg[3];
    const float* stream;
    uint32_t stride;
    float* film32;
    double* film64;
    float* path_rgb;
    float rmin[3], rmax[3];
    uint32_t root_leaf, node0, tri0;
    uint64_t n_paths;  // paths of this job in the call
};

// A run of consecutive paths of one job: pixels [p0, p0 + np) of its
// viewport (row-major), samples [s0, s0 + ns) of each; the paths run over the
// pixels for one sample, then the next sample.
struct BatchTile {
    uint32_t job, p0, np, s0, ns, pad[3];
};

// Device arrays and per-scene jobs of a batch (spt_batch_*).
struct BatchData {
    float4* d_nodes = nullptr;
    NodeQ* d_nodesq = nullptr;
    LeafRec* d_leaves = nullptr;
    float4* d_tris = nullptr;
    float4* d_shading = nullptr;
    spt_material* d_mats = nullptr;
    uint32_t n_nodes = 0, n_tris = 0, n_mats = 0, stack_cap = 0, stack_cap4 = 0;
    bool has_spheres = false;
    std::vector<Job> jobs;  // host copy: sampling, outputs, roots
    Job* d_jobs = nullptr;
    BatchTile* d_tiles = nullptr;
    uint64_t* d_tile_start = nullptr;
    size_t tiles_cap = 0;
    uint32_t* d_done = nullptr;   // finished paths per job
    uint64_t* d_clock = nullptr;  // start, then finish time per job
    uint64_t device_bytes = 0;
    spt_batch_stats stats{};
    std::vector<double> scene_ms;
};

// Scene box used by the guide's input encoding.
struct GuideBox {
    float bmin[3], bext[3];
};

// Device-resident training state of the guide (64-float area after the
// gradient): the host never waits for a step; it reads this block when it
// asks for statistics or a loss.
struct GuideCtl {
    float wsum;       // sum of the current step's record weights
    float loss;       // loss of the current step
    uint32_t bad;     // a gradient entry was not finite
    float scale;      // loss scale (halved on overflow)
    uint32_t steps;   // successful steps (Adam step number - 1)
    uint32_t skipped; // skipped steps
    float last_loss;  // loss of the last successful step
    uint32_t pad;
    unsigned long long trained;  // records consumed by successful steps
};

// Learned guide state of a context (spt_guide.hip).
struct GuideState {
    bool ready = false;
    float* d_theta = nullptr;  // fp32 master parameters, then m, v, grad, control block (64 floats)
    float* d_m = nullptr;
    float* d_v = nullptr;
    float* d_grad = nullptr;
    float* d_acc = nullptr;
    unsigned short* d_wbf = nullptr;  // bf16 copy of the parameters
    float* d_records = nullptr;
    uint32_t records_cap = 0;    // batch x steps_per_pass
    uint32_t records_alloc = 0;  // allocated record slots
    uint64_t collected = 0;
    float lr = 5e-3f, loss_scale = 65536.0f, bsdf_fraction = 0.5f;  // loss_scale: the configured start
    uint32_t flags = 0;  // SPT_GUIDE_*
    uint32_t train_every = 0;      // in-call training rounds every this many bounces (0: none)
    double rays_per_path = 4.0;    // extension rays per path of the last training render
    uint32_t batch = 262144;
    double train_ms = 0.0;
    unsigned short* d_gout = nullptr;  // per-path raw outputs (half floats) during a guided render
    size_t gout_cap = 0;
    uint32_t rec_stride = 1;      // 1 in rec_stride candidate vertices is recorded
    double cand_per_path = 0.0;   // candidate vertices per path of the last training render
    uint64_t candidates = 0;      // candidate vertices of the last training render
    double kernel_ms = 0.0;       // device time of the training kernels (settled steps)
    // training scratch, feature-major (feature f of sample s at f * rows + s):
    // x^T bf16 [32], h1..h3^T bf16 [64] each, dy^T f32 [48], dz3..dz1^T bf16
    // [64] each, then per-split partial gradients f32 [splits][params]
    char* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    uint32_t scratch_rows = 0;
    std::vector<hipEvent_t> ev_free, ev_pending;  // event pairs around training kernels

    GuideCtl* ctl() const { return reinterpret_cast<GuideCtl*>(d_acc); }
    // an event pair for one training kernel (false: none available)
    bool event_pair(hipEvent_t& a, hipEvent_t& b) {
        for (int k = 0; k < 2; ++k) {
            if (ev_free.empty()) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return false;
                ev_free.push_back(e);
            }
            (k == 0 ? a : b) = ev_free.back();
            ev_free.pop_back();
            ev_pending.push_back(k == 0 ? a : b);
        }
        return true;
    }
    // adds the device time of completed pairs (call after the stream is idle)
    void settle_events() {
        for (size_t i = 0; i + 1 < ev_pending.size(); i += 2) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev_pending[i], ev_pending[i + 1]) == hipSuccess) kernel_ms += ms;
        }
        ev_free.insert(ev_free.end(), ev_pending.begin(), ev_pending.end());
        ev_pending.clear();
    }
    void destroy_events() {
        for (hipEvent_t e : ev_free) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_pending) (void)hipEventDestroy(e);
        ev_free.clear();
        ev_pending.clear();
    }
};

}  // namespace spt

// The context: one device, one stream, the uploaded scene and scratch buffers.
struct spt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cus = 256;
    // scene
    float4* d_nodes = nullptr;
    spt::NodeQ* d_nodesq = nullptr;
    spt::LeafRec* d_leaves = nullptr;
    uint32_t n_nodesq = 0, n_leaves = 0, stack_cap4 = 0;
    float4* d_tris = nullptr;
    float4* d_shading = nullptr;
    spt_material* d_mats = nullptr;
    uint32_t n_nodes = 0, n_tris = 0, n_mats = 0;
    float root_box[6] = {0, 0, 0, 0, 0, 0};
    uint32_t root_leaf = 0;
    uint32_t stack_cap = 0;
    uint32_t max_mat = 0;  // largest material id referenced by a triangle
    bool has_spheres = false;
    bool has_scene = false;
    std::vector<spt_material> h_mats;
    uint8_t* d_guidable = nullptr;  // per material: kGuideSingle / kGuideCoated, 0 = not guide-sampled
    // traversal stack spill buffer
    uint32_t* d_spill = nullptr;
    size_t spill_words = 0;
    // status word written by kernels (1 = stack overflow), then two 64-bit
    // fetch counters (nodes, triangles) at byte offset 16
    uint32_t* d_flags = nullptr;
    uint32_t* scratch = nullptr;  // this device's static scratch (spt_trace_noalloc), never freed
    // HIP events: [0..1] bracket extend launches, [2..9] user slots (spt_event_*)
    hipEvent_t ev[10] = {};
    // event pairs bracketing the extend launches of one render call
    std::vector<hipEvent_t> tev;
    // path-tracing state (grown on demand)
    void* d_paths = nullptr;
    size_t paths_bytes = 0;
    spt_render_stats stats{};
    spt::GuideState guide;
    spt_tuning tuning{};  // zero fields: defaults
    int wall_clock_khz = 100000;
    void* h_pin = nullptr;  // 4 KiB of pinned host memory for small read-backs
    // mapped host memory the bounce loop reports its progress to (see k_plan)
    volatile uint32_t* h_status = nullptr;
    uint32_t* d_status = nullptr;
    // contexts of the extra concurrent wavefronts of a render (their own
    // stream, queue, status, flags and spill buffer), created on first use
    std::vector<spt_ctx*> fronts;

    spt::DevScene scene() const {
        spt::DevScene s;
        s.nodes = d_nodes;
        s.nodesq = reinterpret_cast<const uint4*>(d_nodesq);
        s.leaves = reinterpret_cast<const float4*>(d_leaves);
        s.stack_cap4 = stack_cap4;
        s.tris = d_tris;
        s.shading = d_shading;
        s.mats = d_mats;
        for (int i = 0; i < 3; ++i) {
            s.rmin[i] = root_box[i];
            s.rmax[i] = root_box[3 + i];
        }
        s.root_leaf = root_leaf;
        s.n_tris = n_tris;
        s.n_nodes = n_nodes;
        s.n_mats = n_mats;
        s.stack_cap = stack_cap;
        s.has_spheres = has_spheres ? 1u : 0u;
        return s;
    }
};

namespace spt {
spt_status set_error(spt_ctx* ctx, spt_status code, const std::string& msg);
spt_status check_hip(spt_ctx* ctx, hipError_t e, const char* what);
// Waits for the context stream by polling it (no blocking wait on an
// interrupt: those have been seen to wake up hundreds of ms late).
spt_status stream_wait(spt_ctx* ctx, const char* what);
// Copies up to 4 KiB from the device through pinned memory and waits.
spt_status read_small(spt_ctx* ctx, void* dst, const void* d_src, size_t bytes, const char* what);
// grows ctx->d_spill to hold `words` entries; returns false on failure
spt_status ensure_spill(spt_ctx* ctx, size_t words);
// launch helpers implemented by the kernel translation units
spt_status launch_trace(spt_ctx* ctx, const spt_ray32* d_rays, uint64_t n, spt_hit* d_hits, uint32_t mode,
                        uint32_t* d_tests, uint32_t* d_isects);
spt_status launch_raycast(spt_ctx* ctx, const spt_camera* cam, uint32_t kind, void* d_out);
// The allocation-free closest-hit path (spt_trace_noalloc) and the lookup of
// its static scratch on the context's device (from spt_create).
constexpr int kMaxScratchDevices = 64;
spt_status init_static_scratch(spt_ctx* ctx);
spt_status launch_trace_noalloc(spt_ctx* ctx, const spt_ray32* d_rays, uint64_t n, spt_hit* d_hits);
spt_status launch_camera_rays(spt_ctx* ctx, const spt_camera* cam, spt_ray32* d_rays);
GuideBox guide_box(const spt_ctx* ctx);
spt_status guide_eval_launch(spt_ctx* ctx, const float* d_q, uint32_t stride, uint32_t n, float* d_out);
// One Adam step on n records, record i at d_rec + i * rs records (queued on
// the context stream; loss_out, when given, waits for the step).
spt_status guide_train_launch(spt_ctx* ctx, const float* d_rec, uint32_t n, float* loss_out, uint32_t rs);
spt_status launch_render(spt_ctx* ctx, const spt_camera* cam, const spt_render_params* p, float* d_rgb32,
                         double* d_rgb64, float* d_path_rgb);
spt_tuning resolve_tuning(const spt_tuning& t);
void scene_arrays(const spt_bvh* bvh, const float* verts, const float* normals, const uint32_t* mat_ids,
                  std::vector<spt_tri48>& tris, std::vector<float>& sh);
uint32_t tree_stack_cap(const spt_node64* nodes, uint32_t n);
// Collapses a pre-order pair-node tree into four-wide nodes (pre-order, root
// first); returns the stack entries the persistent traversal needs.
uint32_t collapse4(const spt_node64* nodes, uint32_t n, std::vector<Node4>& out);
// Appends the compressed form of a four-wide tree (and four leaf box slots
// per node): inner children rebased by the node array's size, leaf ranges by
// tri_base. Returns false when an index would not fit its field.
bool compress4(const std::vector<Node4>& in, uint32_t tri_base, std::vector<NodeQ>& nodes, std::vector<LeafRec>& leaves);
spt_status make_job(spt_ctx* ctx, const spt_camera* cam, const spt_render_params* prm, float* d_rgb32, double* d_rgb64,
                    float* d_path_rgb, Job& jb, uint64_t& total);
// A multi-scene render (spt_batch_*): every job of the batch, or with
// sample_div > 1 the first 1/sample_div of each job's samples.
spt_status launch_batch(spt_ctx* ctx, BatchData& b, uint32_t sample_div);
}  // namespace spt
