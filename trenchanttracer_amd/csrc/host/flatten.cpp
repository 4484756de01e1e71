// Flattening of the SBVH tree into
//  (1) the reference's compact arrays (BVHCompact::createCompact,
//      src/bvh/BVHCompact.cpp:88-250): DFS from a LIFO stack, inner child index =
//      float4 offset of the child's 64-byte node, leaf = ~vertex offset, 3 float4
//      raw vertices per triangle (v0.x of -0.0 rewritten to +0.0, :143) and a
//      {0x80000000,0,0,0} terminator per leaf, tri/mat index arrays padded to the
//      vertex layout;
//  (2) the device layout streamed by the HIP kernels: the same 64-byte nodes in
//      the same order, but child = node index (inner) or ~triangle index (leaf)
//      with the leaf's triangle counts in the two spare words, and one 48-byte
//      record per triangle {v0, id}, {v1 - v0, material}, {v2 - v0, last} where
//      last = 1 on a leaf's final triangle, so a leaf is read without a
//      terminator fetch and the edges are not recomputed. A leaf without
//      triangles holds one degenerate record (zero edges: det = 0, never
//      accepted).
// The device nodes are then renumbered so that inner siblings share one 128-byte
// line (pair_align); node order does not change any traversal decision.
#include <climits>
#include <cstring>

#include "tt_host.h"

namespace tt {
namespace {

inline int32_t fbits(float f) {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i;
}
inline float bitsf(int32_t i) {
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}

}  // namespace

// Renumbers the device nodes so that two inner siblings share one 128-byte
// L2 line (slots 2k, 2k+1), in depth-first order from the root. The root takes
// slot 0 and slot 1 stays unused; a lone inner child takes half of a fresh
// pair and the next lone child fills the other half. Only memory positions
// change: children keep their order (c0, c1), so traversal decisions, and
// therefore results, are the same.
void pair_align(std::vector<int32_t> &nodes) {
    const int32_t n = (int32_t)(nodes.size() / 16);
    if (n <= 1) return;
    std::vector<int32_t> slot(n, -1);
    int32_t next = 2, pending = -1;
    slot[0] = 0;
    std::vector<int32_t> stack = {0};
    while (!stack.empty()) {
        int32_t k = stack.back();
        stack.pop_back();
        int32_t c0 = nodes[(size_t)k * 16 + 12], c1 = nodes[(size_t)k * 16 + 13];
        if (c0 >= 0 && c1 >= 0) {
            slot[c0] = next;
            slot[c1] = next + 1;
            next += 2;
        } else {
            for (int32_t c : {c0, c1}) {
                if (c < 0) continue;
                if (pending >= 0) {
                    slot[c] = pending;
                    pending = -1;
                } else {
                    slot[c] = next;
                    pending = next + 1;
                    next += 2;
                }
            }
        }
        if (c1 >= 0) stack.push_back(c1);
        if (c0 >= 0) stack.push_back(c0);
    }
    std::vector<int32_t> out((size_t)next * 16, 0);
    for (int32_t k = 0; k < n; ++k) {
        int32_t *d = &out[(size_t)slot[k] * 16];
        std::memcpy(d, &nodes[(size_t)k * 16], 64);
        for (int c = 0; c < 2; ++c)
            if (d[12 + c] >= 0) d[12 + c] = slot[d[12 + c]];
    }
    // unused slots stay zero; no child pointer refers to them
    nodes.swap(out);
}

void flatten(const Mesh &m, BVH &b) {
    Flat &f = b.flat;
    f = Flat();
    const Tree &t = b.tree;
    if (t.root < 0 || t.nodes.empty()) return;

    // A root that is itself a leaf (0 or 1 triangle) has no 64-byte node in the
    // reference layout (createCompact asserts two children). It is wrapped in an
    // inner node whose second child is a leaf with no triangles (visiting it
    // tests nothing).
    Tree wrapped;
    const Tree *tp = &t;
    int32_t root = t.root;
    if (t.nodes[t.root].leaf) {
        wrapped = t;
        TreeNode empty;
        empty.leaf = true;
        empty.lo = empty.hi = (int32_t)wrapped.leaf_tris.size();
        wrapped.nodes.push_back(empty);
        TreeNode top;
        top.bounds = t.nodes[t.root].bounds;
        top.child[0] = t.root;
        top.child[1] = (int32_t)wrapped.nodes.size() - 1;
        wrapped.nodes.push_back(top);
        root = (int32_t)wrapped.nodes.size() - 1;
        tp = &wrapped;
    }
    const Tree &tr = *tp;

    struct Entry {
        int32_t node;
        int32_t idx;  // node index (64-byte units)
    };
    std::vector<Entry> stack;
    stack.push_back({root, 0});
    int32_t num_nodes = 1;
    f.ref_nodes.assign(16, 0);
    f.dev_nodes.assign(16, 0);
    int32_t dev_tri = 0;

    while (!stack.empty()) {
        Entry e = stack.back();
        stack.pop_back();
        const TreeNode &n = tr.nodes[e.node];
        int32_t ref_child[2], dev_child[2], dev_cnt[2] = {0, 0};
        const Box *cb[2];
        for (int i = 0; i < 2; ++i) {
            const TreeNode &c = tr.nodes[n.child[i]];
            cb[i] = &c.bounds;
            if (!c.leaf) {
                ref_child[i] = num_nodes * 4;  // float4 offset (getNumBytes()/16)
                dev_child[i] = num_nodes;
                stack.push_back({n.child[i], num_nodes});
                num_nodes++;
                f.ref_nodes.resize((size_t)num_nodes * 16, 0);
                f.dev_nodes.resize((size_t)num_nodes * 16, 0);
                continue;
            }
            ref_child[i] = ~(int32_t)(f.ref_vertices.size() / 4);
            dev_child[i] = ~dev_tri;
            dev_cnt[i] = c.hi - c.lo;
            for (int32_t j = c.lo; j < c.hi; ++j) {
                int32_t tri = tr.leaf_tris[j];
                const int32_t *vi = &m.tri_v[3 * (size_t)tri];
                const int32_t *ni = &m.tri_n[3 * (size_t)tri];
                V3 v[3] = {m.verts[vi[0]], m.verts[vi[1]], m.verts[vi[2]]};
                if (v[0].x == 0.0f) v[0].x = 0.0f;  // BVHCompact.cpp:143
                for (int k = 0; k < 3; ++k) {
                    V3 nn = m.normals[ni[k]];
                    float vv[4] = {v[k].x, v[k].y, v[k].z, 0.0f};
                    float nv[4] = {nn.x, nn.y, nn.z, 0.0f};
                    f.ref_vertices.insert(f.ref_vertices.end(), vv, vv + 4);
                    f.ref_normals.insert(f.ref_normals.end(), nv, nv + 4);
                }
                f.ref_tri_idx.push_back(tri);
                f.ref_tri_idx.push_back(0);
                f.ref_tri_idx.push_back(0);
                f.ref_mat_idx.push_back(m.tri_mat[tri]);
                f.ref_mat_idx.push_back(0);
                f.ref_mat_idx.push_back(0);
                f.tri_count++;

                V3 e1 = v[1] - v[0];
                V3 e2 = v[2] - v[0];
                float rec[12] = {v[0].x, v[0].y, v[0].z, bitsf(tri),
                                 e1.x,   e1.y,   e1.z,   bitsf(m.tri_mat[tri]),
                                 e2.x,   e2.y,   e2.z,   bitsf(j + 1 == c.hi ? 1 : 0)};
                f.dev_tris.insert(f.dev_tris.end(), rec, rec + 12);
                dev_tri++;
            }
            if (c.hi == c.lo) {  // empty leaf: one degenerate record (det = 0, never accepted)
                float rec[12] = {0.0f, 0.0f, 0.0f, bitsf(-1), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, bitsf(1)};
                f.dev_tris.insert(f.dev_tris.end(), rec, rec + 12);
                dev_tri++;
            }
            float term[4] = {bitsf((int32_t)0x80000000), 0.0f, 0.0f, 0.0f};
            f.ref_vertices.insert(f.ref_vertices.end(), term, term + 4);
            f.ref_normals.insert(f.ref_normals.end(), term, term + 4);
            f.ref_tri_idx.push_back(0);
            f.ref_mat_idx.push_back(0);
            f.leaf_count++;
        }
        int32_t words[12] = {fbits(cb[0]->lo[0]), fbits(cb[0]->hi[0]), fbits(cb[0]->lo[1]), fbits(cb[0]->hi[1]),
                             fbits(cb[1]->lo[0]), fbits(cb[1]->hi[0]), fbits(cb[1]->lo[1]), fbits(cb[1]->hi[1]),
                             fbits(cb[0]->lo[2]), fbits(cb[0]->hi[2]), fbits(cb[1]->lo[2]), fbits(cb[1]->hi[2])};
        int32_t *rn = &f.ref_nodes[(size_t)e.idx * 16];
        int32_t *dn = &f.dev_nodes[(size_t)e.idx * 16];
        std::memcpy(rn, words, sizeof(words));
        std::memcpy(dn, words, sizeof(words));
        rn[12] = ref_child[0];
        rn[13] = ref_child[1];
        rn[14] = 0;
        rn[15] = 0;
        dn[12] = dev_child[0];
        dn[13] = dev_child[1];
        dn[14] = dev_cnt[0];
        dn[15] = dev_cnt[1];
    }
    pair_align(f.dev_nodes);
}

}  // namespace tt
