// bvh_build.cpp - host SAH build and flattening (replaces SBVH::build,
// acceleration/sbvh.cpp:12-37, 58-421).
//
// The tree is the reference's: a full sweep over the three centroid axes per
// node, each preceded by std::sort of the node's references by centroid
// (sbvh_node.h:26-38), SAH cost 2*1.2 + (A_L/A)*N_L + (A_R/A)*N_R with A the
// node's own box area (sbvh.cpp:47-56, 103-104), a leaf when no split beats N
// (sbvh.cpp:70, 117), a final re-sort on the best axis (sbvh.cpp:310) and child
// boxes padded by 1e-3 (sbvh.cpp:409-412). Spatial splits are dead code in the
// reference and are not built.
//
// Design differs from the reference where the result cannot change:
//   - references live in one array and every node sorts its own sub-range in
//     place. std::sort's permutation depends only on the input order and the
//     comparisons, so sorting a sub-range equals sorting the reference's copied
//     child vector;
//   - independent subtrees are built on several host threads.
// The flattened form is what the kernels read: 64-B AABB-pair nodes in
// depth-first pre-order and triangles in depth-first left-first leaf order.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <fstream>
#include <memory>
#include <queue>
#include <stdexcept>
#include <thread>
#include <vector>

#include "spt.h"
#include "spt_internal.h"
#include "spt_math.h"

using namespace spt;

namespace {

struct Ref {  // PrimitiveRef (primitive_ref.h:11-17) without the unused centroid of the box
    V3 mn, mx;  // primitive AABB
    V3 c;       // primitive centroid (triangle.cpp:85-87)
    uint32_t id;
};

struct Node {
    V3 mn, mx;  // this node's box (padded unless root)
    uint32_t begin = 0, count = 0;
    std::unique_ptr<Node> left, right;
    bool leaf() const { return !left; }
};

struct Builder {
    std::vector<Ref> refs;
    std::atomic<int> threads_left{0};
    const float cost_tri = 1.0f, cost_aabb = 1.2f;  // sbvh.cpp:6-7

    // sbvh.cpp:47-56 (infinite-area branches cannot trigger for finite boxes but
    // are kept for identical semantics)
    float sah(std::size_t n1, float a1, std::size_t n2, float a2, float area) const {
        const float inf = std::numeric_limits<float>::infinity();
        if ((a1 == inf) && (n1 == 0)) return inf;
        if ((a2 == inf) && (n2 == 0)) return inf;
        return 2.0f * cost_aabb + ((a1 / area) * n1 * cost_tri) + ((a2 / area) * n2 * cost_tri);
    }

    void sort_axis(uint32_t b, uint32_t e, int axis) {
        Ref* first = refs.data() + b;
        Ref* last = refs.data() + e;
        if (axis == 0)
            std::sort(first, last, [](const Ref& a, const Ref& c) { return a.c.x < c.c.x; });
        else if (axis == 1)
            std::sort(first, last, [](const Ref& a, const Ref& c) { return a.c.y < c.c.y; });
        else
            std::sort(first, last, [](const Ref& a, const Ref& c) { return a.c.z < c.c.z; });
    }

    // union in reference order: box = box + ref  (aabb.h:30-32)
    void range_box(uint32_t b, uint32_t e, V3& mn, V3& mx) const {
        mn = refs[b].mn;
        mx = refs[b].mx;
        for (uint32_t i = b + 1; i < e; ++i) {
            mn = vmin(mn, refs[i].mn);
            mx = vmax(mx, refs[i].mx);
        }
    }

    void split(Node* node, std::vector<float>& left_area) {
        const uint32_t b = node->begin;
        const std::size_t n = node->count;
        float best_cost = cost_tri * n;
        int best_axis = -1;
        long best_event = -1;
        const float node_area = box_area(node->mn, node->mx);
        if (n > 1) {
            if (left_area.size() < n - 1) left_area.resize(n - 1);
            for (int axis = 0; axis < 3; ++axis) {
                sort_axis(b, b + (uint32_t)n, axis);
                V3 lmn = refs[b].mn, lmx = refs[b].mx;
                left_area[0] = box_area(lmn, lmx);
                for (std::size_t i = 1; i < n - 1; ++i) {
                    lmn = vmin(lmn, refs[b + i].mn);
                    lmx = vmax(lmx, refs[b + i].mx);
                    left_area[i] = box_area(lmn, lmx);
                }
                V3 rmn = refs[b + n - 1].mn, rmx = refs[b + n - 1].mx;
                for (long i = (long)n - 2; i >= 0; --i) {
                    if (i != (long)n - 2) {
                        rmn = vmin(rmn, refs[b + i + 1].mn);
                        rmx = vmax(rmx, refs[b + i + 1].mx);
                    }
                    float cost = sah((std::size_t)i + 1, left_area[i], n - 1 - (std::size_t)i, box_area(rmn, rmx),
                                     node_area);
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_event = i;
                        best_axis = axis;
                    }
                }
            }
        } else {
            // a single reference: the reference still sorts (a no-op) and finds no split
        }
        if (best_axis == -1) return;  // leaf
        sort_axis(b, b + (uint32_t)n, best_axis);
        const uint32_t mid = b + (uint32_t)best_event + 1;
        node->left.reset(new Node());
        node->right.reset(new Node());
        Node* l = node->left.get();
        Node* r = node->right.get();
        l->begin = b;
        l->count = mid - b;
        r->begin = mid;
        r->count = b + (uint32_t)n - mid;
        range_box(l->begin, l->begin + l->count, l->mn, l->mx);
        range_box(r->begin, r->begin + r->count, r->mn, r->mx);
        l->mn = l->mn - 1e-3f;
        l->mx = l->mx + 1e-3f;
        r->mn = r->mn - 1e-3f;
        r->mx = r->mx + 1e-3f;
        // Children are independent: large left subtrees go to another thread.
        bool spawn = false;
        if (l->count > 32768) {
            if (threads_left.fetch_sub(1) > 0)
                spawn = true;
            else
                threads_left.fetch_add(1);
        }
        if (spawn) {
            std::thread t([this, l]() {
                std::vector<float> scratch;
                split(l, scratch);
            });
            split(r, left_area);
            t.join();
            threads_left.fetch_add(1);
        } else {
            split(l, left_area);
            split(r, left_area);
        }
    }
};

}  // namespace

struct spt_bvh {
    spt_bvh_info info{};
    std::vector<spt_node64> nodes;
    std::vector<spt_tri48> tris;
    std::vector<uint32_t> leaf_prim;
    std::vector<uint8_t> kinds;  // SPT_PRIM_* per scene primitive (empty: all triangles)
    float root_box[6] = {0, 0, 0, 0, 0, 0};
    std::unique_ptr<Node> root;  // kept for dump / statistics
};

namespace {

void collect_stats(spt_bvh* out) {  // sbvh.cpp:517-575
    std::queue<Node*> q;
    spt_bvh_info& s = out->info;
    s.min_leaf = std::numeric_limits<uint64_t>::max();
    q.push(out->root.get());
    q.push(nullptr);
    while (!q.empty()) {
        Node* n = q.front();
        q.pop();
        if (n) {
            if (n->leaf()) {
                ++s.n_leaves;
                s.n_references += n->count;
                s.min_leaf = std::min<uint64_t>(s.min_leaf, n->count);
                s.max_leaf = std::max<uint64_t>(s.max_leaf, n->count);
            } else {
                ++s.n_inner;
            }
            ++s.n_nodes;
            if (n->left) q.push(n->left.get());
            if (n->right) q.push(n->right.get());
        } else if (!q.empty()) {
            q.push(nullptr);
            ++s.tree_height;
        }
    }
}

void flatten(spt_bvh* out, const std::vector<Ref>& refs, const float* verts, const uint8_t* kinds) {
    // Pre-order indices for inner nodes, leaf-order offsets for triangles.
    struct Item {
        Node* n;
        uint32_t depth;
    };
    std::vector<Node*> inner;  // pre-order
    uint64_t max_depth = 0;
    // first pass: assign pre-order inner indices and leaf tri offsets
    std::vector<Item> stack;
    std::vector<std::pair<Node*, uint32_t>> leaf_off;  // leaf -> offset
    uint32_t cursor = 0;
    stack.push_back({out->root.get(), 0});
    struct Mapped {
        Node* n;
        uint32_t idx;
    };
    std::vector<Mapped> inner_idx;
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        if (it.n->leaf()) {
            leaf_off.push_back({it.n, cursor});
            for (uint32_t i = 0; i < it.n->count; ++i) out->leaf_prim.push_back(refs[it.n->begin + i].id);
            cursor += it.n->count;
            continue;
        }
        max_depth = std::max<uint64_t>(max_depth, it.depth + 1);
        inner_idx.push_back({it.n, (uint32_t)inner.size()});
        inner.push_back(it.n);
        stack.push_back({it.n->right.get(), it.depth + 1});
        stack.push_back({it.n->left.get(), it.depth + 1});
    }
    // node -> index lookup: store in begin/count-free side tables
    std::vector<std::pair<const Node*, uint32_t>> idx_sorted;
    idx_sorted.reserve(inner_idx.size() + leaf_off.size());
    for (auto& m : inner_idx) idx_sorted.push_back({m.n, m.idx});
    std::sort(idx_sorted.begin(), idx_sorted.end());
    std::vector<std::pair<const Node*, uint32_t>> leaf_sorted(leaf_off.begin(), leaf_off.end());
    std::sort(leaf_sorted.begin(), leaf_sorted.end());
    auto find_idx = [](const std::vector<std::pair<const Node*, uint32_t>>& v, const Node* n) {
        auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(n, 0u));
        return it->second;
    };
    out->nodes.resize(inner.size());
    for (std::size_t k = 0; k < inner.size(); ++k) {
        Node* n = inner[k];
        spt_node64& o = out->nodes[k];
        const Node* ch[2] = {n->left.get(), n->right.get()};
        float* mins[2] = {o.lmin, o.rmin};
        float* maxs[2] = {o.lmax, o.rmax};
        uint32_t* refp[2] = {&o.lref, &o.rref};
        uint32_t* cntp[2] = {&o.lcnt, &o.rcnt};
        for (int c = 0; c < 2; ++c) {
            mins[c][0] = ch[c]->mn.x;
            mins[c][1] = ch[c]->mn.y;
            mins[c][2] = ch[c]->mn.z;
            maxs[c][0] = ch[c]->mx.x;
            maxs[c][1] = ch[c]->mx.y;
            maxs[c][2] = ch[c]->mx.z;
            if (ch[c]->leaf()) {
                *refp[c] = find_idx(leaf_sorted, ch[c]);
                *cntp[c] = ch[c]->count;
            } else {
                *refp[c] = find_idx(idx_sorted, ch[c]);
                *cntp[c] = 0;
            }
        }
    }
    // primitives in leaf order: triangle edges as triangle.cpp:22-23 computes
    // them; a sphere's centre and radius
    const std::size_t nt = out->leaf_prim.size();
    out->tris.resize(nt);
    for (std::size_t i = 0; i < nt; ++i) {
        const float* p = verts + 9ull * out->leaf_prim[i];
        if (kinds && kinds[out->leaf_prim[i]] == SPT_PRIM_SPHERE) {
            spt_tri48& t = out->tris[i];
            std::memset(&t, 0, sizeof(t));
            t.v0[0] = p[0];
            t.v0[1] = p[1];
            t.v0[2] = p[2];
            t.e1[0] = p[3];
            t.prim_id = out->leaf_prim[i];
            t.pad[1] = SPT_PRIM_SPHERE;
            continue;
        }
        V3 a{p[0], p[1], p[2]}, b{p[3], p[4], p[5]}, c{p[6], p[7], p[8]};
        V3 e1 = b - a, e2 = c - a;
        spt_tri48& t = out->tris[i];
        t.v0[0] = a.x;
        t.v0[1] = a.y;
        t.v0[2] = a.z;
        t.e1[0] = e1.x;
        t.e1[1] = e1.y;
        t.e1[2] = e1.z;
        t.e2[0] = e2.x;
        t.e2[1] = e2.y;
        t.e2[2] = e2.z;
        t.prim_id = out->leaf_prim[i];
        t.pad[0] = 0;  // material id, filled when shading data is attached
        t.pad[1] = SPT_PRIM_TRIANGLE;
    }
    out->info.n_pair_nodes = inner.size();
    out->info.max_depth = max_depth;
}

}  // namespace

extern "C" spt_status spt_bvh_build(const float* verts, uint32_t n_tris, uint32_t n_threads, spt_bvh** out) {
    return spt_bvh_build_prims(verts, nullptr, n_tris, n_threads, out);
}

extern "C" spt_status spt_bvh_build_prims(const float* verts, const uint8_t* kinds, uint32_t n_tris,
                                          uint32_t n_threads, spt_bvh** out) {
    if (!out || (!verts && n_tris) || n_tris == 0) return SPT_ERR_INVALID;
    bool any_sphere = false;
    if (kinds) {
        for (uint32_t i = 0; i < n_tris; ++i) {
            if (kinds[i] > SPT_PRIM_SPHERE) return SPT_ERR_INVALID;
            any_sphere = any_sphere || kinds[i] == SPT_PRIM_SPHERE;
        }
    }
    auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<spt_bvh> bvh(new spt_bvh());
    try {
        if (any_sphere) bvh->kinds.assign(kinds, kinds + n_tris);
        Builder bld;
        bld.threads_left = (int)(n_threads > 0 ? n_threads - 1 : 0);
        bld.refs.resize(n_tris);
        V3 rmn{0, 0, 0}, rmx{0, 0, 0};
        for (uint32_t i = 0; i < n_tris; ++i) {
            const float* p = verts + 9ull * i;
            Ref& r = bld.refs[i];
            if (any_sphere && kinds[i] == SPT_PRIM_SPHERE) {
                const V3 c{p[0], p[1], p[2]};
                r.mn = c - p[3];  // sphere.cpp:47-55
                r.mx = c + p[3];
                r.c = c;  // sphere.cpp:58-60
            } else {
                V3 a{p[0], p[1], p[2]}, b{p[3], p[4], p[5]}, c{p[6], p[7], p[8]};
                r.mn = vmin(vmin(a, b), c);  // triangle.cpp:76-83
                r.mx = vmax(vmax(a, b), c);
                r.c = (a + b + c) * (1.0f / 3.0f);  // triangle.cpp:85-87
            }
            r.id = i;
            if (!i) {  // sbvh.cpp:20-23
                rmn = r.mn;
                rmx = r.mx;
            } else {
                rmn = vmin(rmn, r.mn);
                rmx = vmax(rmx, r.mx);
            }
        }
        bvh->root.reset(new Node());
        bvh->root->begin = 0;
        bvh->root->count = n_tris;
        bvh->root->mn = rmn;
        bvh->root->mx = rmx;
        std::vector<float> scratch;
        bld.split(bvh->root.get(), scratch);
        bvh->root_box[0] = rmn.x;
        bvh->root_box[1] = rmn.y;
        bvh->root_box[2] = rmn.z;
        bvh->root_box[3] = rmx.x;
        bvh->root_box[4] = rmx.y;
        bvh->root_box[5] = rmx.z;
        collect_stats(bvh.get());
        flatten(bvh.get(), bld.refs, verts, any_sphere ? kinds : nullptr);
    } catch (const std::bad_alloc&) {
        return SPT_ERR_NOMEM;
    } catch (const std::exception&) {
        return SPT_ERR_INVALID;
    }
    bvh->info.n_prims = n_tris;
    for (int i = 0; i < 3; ++i) {
        bvh->info.root_min[i] = bvh->root_box[i];
        bvh->info.root_max[i] = bvh->root_box[3 + i];
    }
    bvh->info.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = bvh.release();
    return SPT_OK;
}

extern "C" spt_status spt_bvh_get_info(const spt_bvh* bvh, spt_bvh_info* info) {
    if (!bvh || !info) return SPT_ERR_INVALID;
    *info = bvh->info;
    return SPT_OK;
}

extern "C" spt_status spt_bvh_export(const spt_bvh* bvh, spt_node64* nodes, spt_tri48* tris, uint32_t* leaf_prim,
                                     float* root_box) {
    if (!bvh) return SPT_ERR_INVALID;
    if (nodes && !bvh->nodes.empty()) std::memcpy(nodes, bvh->nodes.data(), bvh->nodes.size() * sizeof(spt_node64));
    if (tris) std::memcpy(tris, bvh->tris.data(), bvh->tris.size() * sizeof(spt_tri48));
    if (leaf_prim) std::memcpy(leaf_prim, bvh->leaf_prim.data(), bvh->leaf_prim.size() * sizeof(uint32_t));
    if (root_box) std::memcpy(root_box, bvh->root_box, sizeof(bvh->root_box));
    return SPT_OK;
}

// Text dump in the reference's format (sbvh.cpp:577-620): a blank line, then per
// BFS level one line per node "l|i minx miny minz maxx maxy maxz" (default
// ostream float formatting) and a blank line between levels.
extern "C" spt_status spt_bvh_dump(const spt_bvh* bvh, const char* path) {
    if (!bvh || !path) return SPT_ERR_INVALID;
    std::ofstream f(path);
    if (!f) return SPT_ERR_INVALID;
    std::queue<const Node*> q;
    q.push(bvh->root.get());
    q.push(nullptr);
    f << "\n";
    while (!q.empty()) {
        const Node* n = q.front();
        q.pop();
        if (n) {
            f << (n->leaf() ? "l" : "i");
            f << " " << n->mn.x << " " << n->mn.y << " " << n->mn.z;
            f << " " << n->mx.x << " " << n->mx.y << " " << n->mx.z;
            f << "\n";
            if (n->left) q.push(n->left.get());
            if (n->right) q.push(n->right.get());
        } else if (!q.empty()) {
            q.push(nullptr);
            f << "\n";
        }
    }
    return f ? SPT_OK : SPT_ERR_INVALID;
}

extern "C" void spt_bvh_destroy(spt_bvh* bvh) { delete bvh; }

namespace spt {
const std::vector<spt_node64>& bvh_nodes(const spt_bvh* b) { return b->nodes; }
const std::vector<spt_tri48>& bvh_tris(const spt_bvh* b) { return b->tris; }
const std::vector<uint32_t>& bvh_leaf_prim(const spt_bvh* b) { return b->leaf_prim; }
const float* bvh_root_box(const spt_bvh* b) { return b->root_box; }
const std::vector<uint8_t>& bvh_kinds(const spt_bvh* b) { return b->kinds; }
}  // namespace spt
