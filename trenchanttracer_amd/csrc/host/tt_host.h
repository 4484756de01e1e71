// Host-side types of the rendering hot path: mesh, SBVH tree, flattened layouts.
// Compiled with -ffp-contract=off and no -march flags, so every float operation
// is a single IEEE binary32 operation in source order (as the reference's host
// build of src/bvh and include/math/LinearMath.h).
#pragma once

#include <cfloat>
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/tt.h"

namespace tt {

// Thread-local error slot behind tt_last_error().
void set_error(const std::string &msg);
tt_status fail(tt_status code, const std::string &msg);

struct V3 {
    float x, y, z;
    float operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
    float &operator[](int k) { return k == 0 ? x : (k == 1 ? y : z); }
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
// cross / dot with the operand order of include/math/LinearMath.h:260-266.
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Axis-aligned box with the reference's select-based min/max
// (AABB.cpp:11-43 over min3f/max3f, LinearMath.h:252-258).
struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float p[3]) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = lo[k] < p[k] ? lo[k] : p[k];
            hi[k] = hi[k] > p[k] ? hi[k] : p[k];
        }
    }
    void grow(V3 v) {
        float p[3] = {v.x, v.y, v.z};
        grow(p);
    }
    void grow(const Box &b) {
        grow(b.lo);
        grow(b.hi);
    }
    void intersect(const Box &b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = lo[k] > b.lo[k] ? lo[k] : b.lo[k];
            hi[k] = hi[k] < b.hi[k] ? hi[k] : b.hi[k];
        }
    }
    bool valid() const { return lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]; }
    float area() const {
        if (!valid()) return 0.0f;
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return (dx * dy + dy * dz + dz * dx) * 2.0f;
    }
};

struct Mesh {
    std::vector<V3> verts;
    std::vector<V3> normals;
    std::vector<int32_t> tri_v;     // 3 per triangle
    std::vector<int32_t> tri_n;     // 3 per triangle (normal indices)
    std::vector<int32_t> tri_mat;   // 1 per triangle
    int32_t ntris() const { return (int32_t)tri_mat.size(); }
};

// Binary SBVH tree as produced by the builder.
struct TreeNode {
    Box bounds;
    int32_t child[2] = {-1, -1};  // indices into Tree::nodes (inner nodes)
    int32_t lo = 0, hi = 0;       // leaf: range in Tree::leaf_tris
    bool leaf = false;
};

struct Tree {
    std::vector<TreeNode> nodes;
    std::vector<int32_t> leaf_tris;
    int32_t root = -1;
};

// Flattened arrays: reference compact layout + device layout.
struct Flat {
    // reference compact layout (BVHCompact.cpp:88-250)
    std::vector<int32_t> ref_nodes;     // 16 per inner node
    std::vector<float> ref_vertices;    // 4 per entry
    std::vector<float> ref_normals;     // 4 per entry
    std::vector<int32_t> ref_tri_idx;
    std::vector<int32_t> ref_mat_idx;
    uint32_t leaf_count = 0, tri_count = 0;
    // device layout
    std::vector<int32_t> dev_nodes;     // 16 per node
    std::vector<float> dev_tris;        // 12 per triangle reference
};

struct BVH {
    Tree tree;
    Flat flat;
    tt_bvh_stats stats{};
};

// sbvh.cpp
void build_sbvh(const Mesh &m, float split_alpha, int nthreads, BVH &out);
// flatten.cpp
void flatten(const Mesh &m, BVH &b);
void pair_align(std::vector<int32_t> &dev_nodes);
// api_host.cpp: structural check of a device layout (a tree from node 0 with every
// index in range, depth <= 64, leaves inside the triangle array, the last record
// ending a leaf) and of its triangle ids against a mesh of ntris triangles
tt_status check_device_layout(const Flat &f);
tt_status check_tri_ids(const Flat &f, int64_t ntris);
// gen.cpp helpers
void face_normal_fill(Mesh &m);

}  // namespace tt

struct tt_mesh {
    tt::Mesh m;
};
struct tt_bvh {
    tt::BVH b;
};
