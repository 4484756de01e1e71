// Device-side shared definitions (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/tt.h"

namespace tt {

constexpr int kSentinel = 0x76543210;  // IntersectKernel.cuh:9
constexpr int kStackCap = 72;          // >= 64 outstanding pushes (tree depth <= 64)

struct TraceArgs {
    const float4 *nodes;   // 4 float4 per node
    const float4 *tris;    // 3 float4 per triangle record
    const tt_ray *rays;
    tt_hit *hits;          // closest-hit output (or nullptr)
    int32_t *flags;        // any-hit output (or nullptr)
    int64_t n;
    int32_t num_nodes;
    const int32_t *ray_root;  // if set: root node of each ray (several scenes in one node array)
    const int32_t *n_dev;  // if set: ray count read on the device (capped at n)
    int32_t *reset0;       // if set: zeroed by the first lane of block 0
    int32_t *reset1;
    int32_t xcd_order;     // 1: each XCD (blocks b = x mod 8) holds one contiguous range of waves
    int32_t refill;        // a wave takes new rays once this many of its lanes are idle (1..64)
    const uint32_t *order; // if set: the r-th ray traced is ray order[r] (a permutation of 0..n-1)
};

// Triangle acceptance: the reference's test (front faces only, det > 1e-5) or
// a two-sided test with no epsilon (accept any det != 0) for the path tracer.
enum TriMode { kTriReference = 0, kTriTwoSided = 1 };

// Ray reordering (reorder.hip): workspace size for n rays (sort_bytes: the part the
// radix sort uses), and the permutation that sorts the rays by a Morton code of
// their origins within the box lo..hi.
hipError_t reorder_workspace_bytes(int64_t n, size_t *bytes, size_t *sort_bytes);
hipError_t reorder_rays(const tt_ray *rays, int64_t n, const float lo[3], const float hi[3], void *ws,
                        size_t sort_bytes, hipStream_t s, const uint32_t **order);

// Launchers (trace.hip).
hipError_t launch_trace(const TraceArgs &a, int variant, bool any, int lds_stack, int block, int xcd_order,
                        int refill, int num_cus, hipStream_t s, int tri_mode = kTriReference);

}  // namespace tt
