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
    // node-fetch probe (closest-hit loop): per ray the counts of node visits,
    // triangle tests and leaf visits, and from rec_off[ray] (32-bit words) the 4-bit
    // slab-test outcome of every node visit, 8 per word
    int32_t *rec_nodes;
    int32_t *rec_tris;
    int32_t *rec_leaves;
    int32_t *rec_log;
    const int64_t *rec_off;
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
// The same for a queue whose length is read on the device (n_dev, capped at cap); the
// workspace is reorder_workspace_bytes(cap); launches sized for cap, blocks past the
// count leave at once.
hipError_t reorder_rays_dev(const tt_ray *rays, int64_t cap, const int32_t *n_dev, const float lo[3],
                            const float hi[3], void *ws, hipStream_t s, const uint32_t **order);

// Inversion of a float CDF (cdf[0] = 0, nondecreasing): the largest k in [lo, hi]
// with cdf[k] < u, or lo when there is none. Over [0, n - 1] this is the sampled
// entry: cdf[k] < u <= cdf[k + 1].
__device__ __forceinline__ int cdf_search_range(const float *cdf, int lo, int hi, float u) {
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cdf[mid] < u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Guide-table cell of u in [0, 1) for a table of m cells; the build and the
// lookup use these same float operations.
__device__ __forceinline__ int guide_cell(float u, int m) { return min(m - 1, max(0, (int)(u * (float)m))); }

// CDF inversion through a guide table of n cells (Chen & Asau's index table):
// guide[c] holds the smallest and largest index the full search returns for any
// float u of cell c, so the search runs inside that range (usually 0-2 steps
// instead of log2 n) and returns the full search's index for every u.
__device__ __forceinline__ int cdf_search_guided(const float *cdf, const int2 *guide, int n, float u) {
    const int2 r = guide[guide_cell(u, n)];
    return cdf_search_range(cdf, r.x, r.y, u);
}

// One cell of a row's conditional guide table together with everything a sample drawn
// in it reads: the index range [lo, hi] the CDF search can return for a u of the cell,
// the CDF entries lo + 1 and lo + 2, and the texels (rgb, sampling weight) of columns
// lo, lo + 1 and lo + 2 (zero past the row's end). 64 bytes, so a next-event sample of
// the environment reads one record; ranges of more than three entries fall back to the
// CDF and texel arrays.
struct alignas(64) EnvCell {
    int32_t lo, hi;
    float c1, c2;
    float4 t[3];
};
static_assert(sizeof(EnvCell) == 64, "EnvCell is one 64-byte record");

// One cell of the marginal CDF's guide table: its search range and the CDF entries
// lo + 1, lo + 2 (16 bytes; ranges of more than three entries read the CDF).
struct alignas(16) MargCell {
    int32_t lo, hi;
    float c1, c2;
};

// Launchers (trace.hip).
hipError_t launch_trace(const TraceArgs &a, int variant, bool any, int lds_stack, int block, int xcd_order,
                        int refill, int num_cus, hipStream_t s, int tri_mode = kTriReference);
// The closest-hit loop (variant 0 or 1, LDS stack 16, reference triangle test) for
// the node-fetch probe: pass 1 counts (rec_nodes / rec_tris / rec_leaves), pass 2
// records the node visits' outcomes (rec_off, rec_log), pass 3 replays them,
// writing one 16-byte record per ray to hits.
hipError_t launch_trace_record(const TraceArgs &a, int variant, int pass, int block, int refill, int num_cus,
                               hipStream_t s);

}  // namespace tt
