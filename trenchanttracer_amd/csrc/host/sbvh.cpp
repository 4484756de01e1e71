// SBVH builder (Stich, Friedrich, Dietrich: "Spatial Splits in Bounding Volume
// Hierarchies", HPG 2009), making the same decisions as the reference builder
// src/bvh/BVH.cpp:3-341 so that the flattened tree is identical:
//   - root references are triangle boxes, minOverlap = area(root) * alpha (BVH.cpp:24-35)
//   - leaf when n <= 1 or level >= 64 (BVH.cpp:71)
//   - object split: per axis sort by (min+max, tri id), SAH sweep, strict < (BVH.cpp:122-154,343-351)
//   - spatial split: 32 bins, only for level < 48 and overlap >= minOverlap (BVH.cpp:82-87,166-231)
//   - leaf if minSAH == leafSAH; spatial if minSAH == spatial SAH, with
//     unsplit-left / unsplit-right / duplicate choices (BVH.cpp:89-104,233-302)
//   - leaf triangle order = references popped from the stack top (BVH.cpp:113-120)
//
// Data structure differs from the reference (one reference vector per node
// instead of one shared reference stack, index sorts instead of in-place
// quicksort, subtrees built on several threads). The sort key is a strict total
// order, so any correct sort yields the same permutation, and every node sees its
// references in the same order as the reference's stack segment; the output is
// therefore independent of the thread count.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <future>
#include <memory>
#include <thread>

#include "tt_host.h"

namespace tt {
namespace {

constexpr int kMaxDepth = 64;         // BVH.h:37
constexpr int kMaxSpatialDepth = 48;  // BVH.h:38
constexpr int kBins = 32;             // BVH.h:39
constexpr int kMaxLeafSize = 0x7FFFFFF;  // SAHHelper.h:19
constexpr int kParallelMinRefs = 4096;

// SAHHelper with default costs (SAHHelper.h:11-19, SAHHelper.cpp:19-25).
inline float tri_cost(int n) { return (float)n * 1.0f; }
inline float node_cost(int n) { return (float)n * 1.0f; }

// min1f / min1f3 (LinearMath.h:66-72).
inline float min1f(float a, float b) { return (a > b) ? b : a; }
inline float min1f3(float a, float b, float c) { return min1f(min1f(a, b), c); }

// (int) conversion of a float as x86 cvttss2si does it: NaN or out of range
// gives INT_MIN. Reached by flat nodes (0 * inf) in the spatial binning
// (Vec3i(Vec3f) at BVH.cpp:183-186).
inline int32_t f2i(float f) {
    if (f >= -2147483648.0f && f < 2147483648.0f) return (int32_t)f;
    return INT32_MIN;
}
inline int clamp1i(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline float clamp1f(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct Ref {
    int32_t tri;
    Box b;
};

struct BNode {
    Box bounds;
    std::unique_ptr<BNode> c[2];
    std::vector<int32_t> tris;  // leaf triangles in leaf order
    bool leaf = false;
    int level = 0;
};

struct ObjectSplit {
    float sah = FLT_MAX;
    int dim = 0;
    int num_left = 0;
    Box left, right;
};

struct SpatialSplit {
    float sah = FLT_MAX;
    int dim = 0;
    float pos = 0.0f;
};

struct SpatialBin {
    Box b;
    int enter = 0, exit = 0;
};

class Builder {
  public:
    Builder(const Mesh &m, float min_overlap, int nthreads)
        : m_(m), min_overlap_(min_overlap), budget_(nthreads - 1) {}

    std::unique_ptr<BNode> build(std::vector<Ref> refs, const Box &bounds, int level);
    int64_t duplicates() const { return dups_.load(); }

  private:
    std::unique_ptr<BNode> make_leaf(const std::vector<Ref> &refs, const std::vector<int32_t> *order,
                                     const Box &bounds, int level);
    ObjectSplit find_object_split(const std::vector<Ref> &refs, float node_sah,
                                  std::vector<int32_t> order[3]);
    SpatialSplit find_spatial_split(const std::vector<Ref> &refs, const std::vector<int32_t> &zorder,
                                    const Box &bounds, float node_sah);
    void perform_spatial_split(std::vector<Ref> &refs, const SpatialSplit &s, std::vector<Ref> &left,
                               Box &lb, std::vector<Ref> &right, Box &rb);
    void split_reference(Ref &left, Ref &right, const Ref &ref, int dim, float pos) const;

    const Mesh &m_;
    float min_overlap_;
    std::atomic<int> budget_;
    std::atomic<int64_t> dups_{0};
};

std::unique_ptr<BNode> Builder::make_leaf(const std::vector<Ref> &refs, const std::vector<int32_t> *order,
                                          const Box &bounds, int level) {
    // createLeaf (BVH.cpp:113-120): refs are popped from the top of the stack
    // segment, i.e. the leaf lists them in reverse of the current order.
    auto n = std::make_unique<BNode>();
    n->leaf = true;
    n->bounds = bounds;
    n->level = level;
    size_t cnt = refs.size();
    n->tris.resize(cnt);
    for (size_t i = 0; i < cnt; ++i) {
        size_t src = cnt - 1 - i;
        n->tris[i] = order ? refs[(*order)[src]].tri : refs[src].tri;
    }
    return n;
}

ObjectSplit Builder::find_object_split(const std::vector<Ref> &refs, float node_sah,
                                       std::vector<int32_t> order[3]) {
    // BVH.cpp:122-154. The sweep order of each axis is the sort of the node's
    // references by (min+max along the axis, triangle id).
    const int n = (int)refs.size();
    ObjectSplit split;
    struct Key {
        float k;
        int32_t tri;
        int32_t idx;
    };
    std::vector<Key> keys(n);
    std::vector<Box> right_bounds(n > 1 ? n - 1 : 0);
    for (int dim = 0; dim < 3; ++dim) {
        for (int i = 0; i < n; ++i) keys[i] = {refs[i].b.lo[dim] + refs[i].b.hi[dim], refs[i].tri, i};
        std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
            if (a.k < b.k) return true;
            if (a.k > b.k) return false;
            return a.tri < b.tri;
        });
        order[dim].resize(n);
        for (int i = 0; i < n; ++i) order[dim][i] = keys[i].idx;
        const std::vector<int32_t> &o = order[dim];

        Box rb;
        for (int i = n - 1; i > 0; --i) {
            rb.grow(refs[o[i]].b);
            right_bounds[i - 1] = rb;
        }
        Box lb;
        for (int i = 1; i < n; ++i) {
            lb.grow(refs[o[i - 1]].b);
            float sah = node_sah + lb.area() * tri_cost(i) + right_bounds[i - 1].area() * tri_cost(n - i);
            if (sah < split.sah) {
                split.sah = sah;
                split.dim = dim;
                split.num_left = i;
                split.left = lb;
                split.right = right_bounds[i - 1];
            }
        }
    }
    return split;
}

void Builder::split_reference(Ref &left, Ref &right, const Ref &ref, int dim, float pos) const {
    // BVH.cpp:304-341: clip the triangle against the plane, vertices first, then
    // the edge crossings, and intersect with the reference's own box.
    left.tri = right.tri = ref.tri;
    left.b = Box();
    right.b = Box();
    const int32_t *ind = &m_.tri_v[3 * (size_t)ref.tri];
    V3 v1 = m_.verts[ind[2]];
    for (int k = 0; k < 3; ++k) {
        V3 v0 = v1;
        v1 = m_.verts[ind[k]];
        float v0p = v0[dim];
        float v1p = v1[dim];
        if (v0p <= pos) left.b.grow(v0);
        if (v0p >= pos) right.b.grow(v0);
        if ((v0p < pos && v1p > pos) || (v0p > pos && v1p < pos)) {
            float t = clamp1f((pos - v0p) / (v1p - v0p), 0.0f, 1.0f);
            float s = 1.0f - t;  // lerp (Sort.h:26-29): a * (1 - t) + b * t
            V3 p = {v0.x * s + v1.x * t, v0.y * s + v1.y * t, v0.z * s + v1.z * t};
            left.b.grow(p);
            right.b.grow(p);
        }
    }
    left.b.hi[dim] = pos;
    right.b.lo[dim] = pos;
    left.b.intersect(ref.b);
    right.b.intersect(ref.b);
}

SpatialSplit Builder::find_spatial_split(const std::vector<Ref> &refs, const std::vector<int32_t> &zorder,
                                         const Box &bounds, float node_sah) {
    // BVH.cpp:166-231. References are binned in the current stack order (sorted
    // along z by the preceding object-split search).
    float origin[3] = {bounds.lo[0], bounds.lo[1], bounds.lo[2]};
    float bin_size[3], inv_bin[3];
    for (int k = 0; k < 3; ++k) bin_size[k] = (bounds.hi[k] - origin[k]) * (1.0f / (float)kBins);
    for (int k = 0; k < 3; ++k) inv_bin[k] = 1.0f / bin_size[k];

    SpatialBin bins[3][kBins];
    const int n = (int)refs.size();
    for (int r = 0; r < n; ++r) {
        const Ref &ref = refs[zorder[r]];
        int first[3], last[3];
        for (int k = 0; k < 3; ++k) first[k] = clamp1i(f2i((ref.b.lo[k] - origin[k]) * inv_bin[k]), 0, kBins - 1);
        for (int k = 0; k < 3; ++k)
            last[k] = clamp1i(f2i((ref.b.hi[k] - origin[k]) * inv_bin[k]), first[k], kBins - 1);
        for (int dim = 0; dim < 3; ++dim) {
            Ref cur = ref;
            for (int i = first[dim]; i < last[dim]; ++i) {
                Ref l, rr;
                split_reference(l, rr, cur, dim, origin[dim] + bin_size[dim] * (float)(i + 1));
                bins[dim][i].b.grow(l.b);
                cur = rr;
            }
            bins[dim][last[dim]].b.grow(cur.b);
            bins[dim][first[dim]].enter++;
            bins[dim][last[dim]].exit++;
        }
    }

    SpatialSplit split;
    Box right_bounds[kBins - 1];
    for (int dim = 0; dim < 3; ++dim) {
        Box rb;
        for (int i = kBins - 1; i > 0; --i) {
            rb.grow(bins[dim][i].b);
            right_bounds[i - 1] = rb;
        }
        Box lb;
        int left_num = 0, right_num = n;
        for (int i = 1; i < kBins; ++i) {
            lb.grow(bins[dim][i - 1].b);
            left_num += bins[dim][i - 1].enter;
            right_num -= bins[dim][i - 1].exit;
            float sah = node_sah + lb.area() * tri_cost(left_num) + right_bounds[i - 1].area() * tri_cost(right_num);
            if (sah < split.sah) {
                split.sah = sah;
                split.dim = dim;
                split.pos = origin[dim] + bin_size[dim] * (float)i;
            }
        }
    }
    return split;
}

void Builder::perform_spatial_split(std::vector<Ref> &refs, const SpatialSplit &s, std::vector<Ref> &left,
                                    Box &lb, std::vector<Ref> &right, Box &rb) {
    // BVH.cpp:233-302, on the node's own segment [0, refs.size()): left part
    // [0, left_end), undecided [left_end, right_start), right part
    // [right_start, size) with duplicates appended at the end.
    int left_end = 0;
    int right_start = (int)refs.size();
    lb = Box();
    rb = Box();
    for (int i = left_end; i < right_start; ++i) {
        if (refs[i].b.hi[s.dim] <= s.pos) {
            lb.grow(refs[i].b);
            std::swap(refs[i], refs[left_end++]);
        } else if (refs[i].b.lo[s.dim] >= s.pos) {
            rb.grow(refs[i].b);
            std::swap(refs[i--], refs[--right_start]);
        }
    }
    while (left_end < right_start) {
        Ref lref, rref;
        split_reference(lref, rref, refs[left_end], s.dim, s.pos);
        Box lub = lb, rub = rb, ldb = lb, rdb = rb;
        lub.grow(refs[left_end].b);
        rub.grow(refs[left_end].b);
        ldb.grow(lref.b);
        rdb.grow(rref.b);
        const int size = (int)refs.size();
        float lac = tri_cost(left_end);
        float rac = tri_cost(size - right_start);
        float lbc = tri_cost(left_end + 1);
        float rbc = tri_cost(size - right_start + 1);
        float unsplit_left = lub.area() * lbc + rb.area() * rac;
        float unsplit_right = lb.area() * lac + rub.area() * rbc;
        float duplicate = ldb.area() * lbc + rdb.area() * rbc;
        float mn = min1f3(unsplit_left, unsplit_right, duplicate);
        if (mn == unsplit_left) {
            lb = lub;
            left_end++;
        } else if (mn == unsplit_right) {
            rb = rub;
            std::swap(refs[left_end], refs[--right_start]);
        } else {
            lb = ldb;
            rb = rdb;
            refs[left_end++] = lref;
            refs.push_back(rref);
        }
    }
    left.assign(refs.begin(), refs.begin() + left_end);
    right.assign(refs.begin() + right_start, refs.end());
}

std::unique_ptr<BNode> Builder::build(std::vector<Ref> refs, const Box &bounds, int level) {
    // buildNode, BVH.cpp:70-111.
    const int n = (int)refs.size();
    if (n <= 1 || level >= kMaxDepth) return make_leaf(refs, nullptr, bounds, level);

    float area = bounds.area();
    float leaf_sah = area * tri_cost(n);
    float node_sah = area * node_cost(2);
    std::vector<int32_t> order[3];
    ObjectSplit object = find_object_split(refs, node_sah, order);

    SpatialSplit spatial;
    if (level < kMaxSpatialDepth) {
        Box overlap = object.left;
        overlap.intersect(object.right);
        if (overlap.area() >= min_overlap_) spatial = find_spatial_split(refs, order[2], bounds, node_sah);
    }

    float min_sah = min1f3(leaf_sah, object.sah, spatial.sah);
    if (min_sah == leaf_sah && n <= kMaxLeafSize) return make_leaf(refs, &order[2], bounds, level);

    std::vector<Ref> left, right;
    Box lb, rb;
    if (min_sah == spatial.sah) {
        // The node's segment is in z order when the spatial partition runs.
        std::vector<Ref> zref(n);
        for (int i = 0; i < n; ++i) zref[i] = refs[order[2][i]];
        perform_spatial_split(zref, spatial, left, lb, right, rb);
    }
    if (left.empty() || right.empty()) {
        // performObjectSplit (BVH.cpp:156-164). An empty side means the spatial
        // partition only permuted the references, so the sort below sees the same set.
        const std::vector<int32_t> &o = order[object.dim];
        left.resize(object.num_left);
        right.resize(n - object.num_left);
        for (int i = 0; i < object.num_left; ++i) left[i] = refs[o[i]];
        for (int i = object.num_left; i < n; ++i) right[i - object.num_left] = refs[o[i]];
        lb = object.left;
        rb = object.right;
    }
    dups_ += (int64_t)left.size() + (int64_t)right.size() - n;
    refs.clear();
    refs.shrink_to_fit();
    for (auto &o : order) {
        o.clear();
        o.shrink_to_fit();
    }

    auto node = std::make_unique<BNode>();
    node->bounds = bounds;
    node->level = level;
    // The reference recurses into the right child first (BVH.cpp:108-109); with
    // per-node reference vectors the two subtrees are independent.
    bool spawn = false;
    if ((int)right.size() >= kParallelMinRefs && (int)left.size() >= kParallelMinRefs) {
        int b = budget_.load();
        while (b > 0 && !budget_.compare_exchange_weak(b, b - 1)) {
        }
        spawn = b > 0;
    }
    if (spawn) {
        auto fut = std::async(std::launch::async, [this, r = std::move(right), rb, level]() mutable {
            auto res = build(std::move(r), rb, level + 1);
            budget_++;
            return res;
        });
        node->c[0] = build(std::move(left), lb, level + 1);
        node->c[1] = fut.get();
    } else {
        node->c[1] = build(std::move(right), rb, level + 1);
        node->c[0] = build(std::move(left), lb, level + 1);
    }
    return node;
}

// Converts the pointer tree to the index tree (pre-order, child 0 first).
int32_t emit(const BNode *bn, Tree &t, int &max_depth) {
    int32_t id = (int32_t)t.nodes.size();
    t.nodes.emplace_back();
    t.nodes[id].bounds = bn->bounds;
    if (bn->leaf) {
        t.nodes[id].leaf = true;
        t.nodes[id].lo = (int32_t)t.leaf_tris.size();
        t.leaf_tris.insert(t.leaf_tris.end(), bn->tris.begin(), bn->tris.end());
        t.nodes[id].hi = (int32_t)t.leaf_tris.size();
        max_depth = std::max(max_depth, bn->level);
        return id;
    }
    int32_t c0 = emit(bn->c[0].get(), t, max_depth);
    int32_t c1 = emit(bn->c[1].get(), t, max_depth);
    t.nodes[id].child[0] = c0;
    t.nodes[id].child[1] = c1;
    return id;
}

// BVHNode::computeSubtreeProbabilities (BVHNode.cpp:56-66): pre-order float sum.
void subtree_sah(const Tree &t, int32_t id, float prob, float &sah) {
    const TreeNode &n = t.nodes[id];
    float cost = n.leaf ? node_cost(0) + tri_cost(n.hi - n.lo) : node_cost(2) + tri_cost(0);
    sah += prob * cost;
    if (n.leaf) return;
    float a = n.bounds.area();
    for (int i = 0; i < 2; ++i) {
        const TreeNode &c = t.nodes[n.child[i]];
        subtree_sah(t, n.child[i], prob * c.bounds.area() / a, sah);
    }
}

}  // namespace

void build_sbvh(const Mesh &m, float split_alpha, int nthreads, BVH &out) {
    auto t0 = std::chrono::steady_clock::now();
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = m.ntris();
    std::vector<Ref> refs(nt);
    Box root;
    for (int i = 0; i < nt; ++i) {
        refs[i].tri = i;
        for (int k = 0; k < 3; ++k) refs[i].b.grow(m.verts[m.tri_v[3 * (size_t)i + k]]);
        root.grow(refs[i].b);
    }
    float min_overlap = root.area() * split_alpha;
    Builder b(m, min_overlap, nthreads);
    std::unique_ptr<BNode> top = b.build(std::move(refs), root, 0);

    out.tree = Tree();
    int max_depth = 0;
    out.tree.root = emit(top.get(), out.tree, max_depth);
    top.reset();

    tt_bvh_stats &s = out.stats;
    s = tt_bvh_stats{};
    s.num_tris = nt;
    s.num_refs = (int32_t)out.tree.leaf_tris.size();
    s.num_duplicates = (int32_t)b.duplicates();
    for (const TreeNode &n : out.tree.nodes) (n.leaf ? s.num_leaves : s.num_inner)++;
    s.max_depth = max_depth;
    float sah = 0.0f;
    if (nt > 0) subtree_sah(out.tree, out.tree.root, 1.0f, sah);
    s.sah = sah;
    for (int k = 0; k < 3; ++k) {
        s.root_lo[k] = root.lo[k];
        s.root_hi[k] = root.hi[k];
    }
    s.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace tt
