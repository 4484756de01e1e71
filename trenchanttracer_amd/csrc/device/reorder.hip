// Ray reordering for tt_trace_closest / tt_trace_any (tt_trace_cfg.reorder = 1):
// a stable sort of the ray indices by a 24-bit Morton code of the ray origins.
//
// The traversal kernel then takes ray order[r] as its r-th ray and writes the
// result at that index, so the caller's arrays keep their order and every hit is
// the one the unsorted launch computes (each ray's traversal depends on that ray
// alone). Waves of rays that start near each other walk the same top of the tree
// and the same leaves near their origins, and fetch fewer distinct nodes. Camera
// rays share one origin and so one key: the sort is stable, and they stay in the
// caller's order.
//
// The sort: LSD radix, 8 bits per pass, 3 passes, written for wave64. A pass over
// tiles of 4096 keys (256 threads, 16 keys per thread) is three launches:
//   hist    each tile's 256-bin histogram of the pass's digit, stored digit-major
//           (hist[d * T + t]); the first pass's histogram comes with the key pass;
//   scan    one block per digit: exclusive prefix of its row over the tiles, and the
//           digit's total;
//   scatter each tile ranks its keys stably (per wave: digit-match masks from eight
//           ballots, a popcount below the lane, running per-wave digit counts in LDS),
//           stages them in LDS in digit order, and writes each digit's run to its
//           global position (digit base + the tile's prefix), so a run of equal digits
//           is written by consecutive lanes.
// Every histogram entry and total is written by its kernel, so nothing is cleared
// before a call: the sort issues no memset and allocates nothing, and the pass reads
// only what the previous launch on the stream wrote. The key pass reads each ray's
// origin; the first scatter takes the ray index as the value (no index array is read),
// and the last writes the permutation only.
#include <cstdint>

#include "dev_common.h"

namespace tt {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;                         // keys per thread
constexpr int kSortTile = kSortThreads * kSortItems;   // 4096 keys per tile
constexpr int kWaveChunk = 64 * kSortItems;            // 1024 keys per wave, in 16 rounds of 64
static_assert(kSortThreads == 256, "one thread per digit in the scan and scatter");

// bit i of an 8-bit value moves to bit 3i
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0xffu;
    v = (v | (v << 8)) & 0x0000f00fu;
    v = (v | (v << 4)) & 0x000c30c3u;
    v = (v | (v << 2)) & 0x00249249u;
    return v;
}

// cell 0..255 of a coordinate; NaN and values outside the box clamp to the nearest end
// (fmaxf returns the non-NaN operand)
__device__ __forceinline__ uint32_t cell8(float v, float lo, float scale) {
    return (uint32_t)fminf(fmaxf((v - lo) * scale, 0.0f), 255.0f);
}

// Lanes of this wave whose (valid) digit equals this lane's.
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return valid ? peers : 0ull;
}

__device__ __forceinline__ uint64_t lanes_below() { return (1ull << __lane_id()) - 1ull; }

// Adds a key to the block histogram with one LDS atomic per lane. (Gathering each
// digit's lanes first, with eight ballots per round as the scatter's ranking does, took
// 27 us per histogram pass of 8.3M keys against 22 us.)
__device__ __forceinline__ void hist_add(uint32_t *h, uint32_t d, bool valid) {
    if (valid) atomicAdd(&h[d], 1u);
}

__device__ __forceinline__ void write_hist(const uint32_t *h, uint32_t *hist, int64_t tiles, int64_t tile) {
    for (int d = threadIdx.x; d < 256; d += kSortThreads) hist[(int64_t)d * tiles + tile] = h[d];
}

// Keys of the rays (24-bit Morton code of the origin) and the first pass's tile
// histograms.
// The number of keys: n, or the device count n_dev (capped at n) when given; tiles at or
// past it are neither histogrammed nor scanned nor scattered.
__device__ __forceinline__ int64_t sort_count(int64_t n, const int32_t *n_dev) {
    return n_dev ? min((int64_t)max(*n_dev, 0), n) : n;
}

__global__ __launch_bounds__(kSortThreads) void k_sort_keys(const tt_ray *__restrict__ rays, int64_t n_cap,
                                                            const int32_t *n_dev, float lox, float loy, float loz,
                                                            float sx, float sy, float sz, uint32_t *__restrict__ keys,
                                                            uint32_t *__restrict__ hist, int64_t tiles) {
    const int64_t n = sort_count(n_cap, n_dev);
    if ((int64_t)blockIdx.x * kSortTile >= n) return;   // the whole block: past the keys
    __shared__ uint32_t h[256];
    for (int d = threadIdx.x; d < 256; d += kSortThreads) h[d] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    // all of the thread's origins in flight at once, then the keys
    float4 o[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = base + (int64_t)r * kSortThreads + threadIdx.x;
        o[r] = i < n ? reinterpret_cast<const float4 *>(rays + i)[0] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = base + (int64_t)r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const uint32_t k = (spread3(cell8(o[r].x, lox, sx)) << 2) | (spread3(cell8(o[r].y, loy, sy)) << 1) |
                           spread3(cell8(o[r].z, loz, sz));
        if (valid) keys[i] = k;
        hist_add(h, k & 0xffu, valid);
    }
    __syncthreads();
    write_hist(h, hist, tiles, blockIdx.x);
}

// Tile histograms of the digit at `shift`.
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t *__restrict__ keys, int64_t n_cap,
                                                            const int32_t *n_dev, int shift, uint32_t *__restrict__ hist,
                                                            int64_t tiles) {
    const int64_t n = sort_count(n_cap, n_dev);
    if ((int64_t)blockIdx.x * kSortTile >= n) return;
    __shared__ uint32_t h[256];
    for (int d = threadIdx.x; d < 256; d += kSortThreads) h[d] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    uint32_t k[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = base + (int64_t)r * kSortThreads + threadIdx.x;
        k[r] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = base + (int64_t)r * kSortThreads + threadIdx.x;
        hist_add(h, (k[r] >> shift) & 0xffu, i < n);
    }
    __syncthreads();
    write_hist(h, hist, tiles, blockIdx.x);
}

// Exclusive block scan of one value per thread; the block's total in *total.
__device__ __forceinline__ uint32_t block_scan_u32(uint32_t x, uint32_t *sh, uint32_t *total) {
    const int t = threadIdx.x;
    sh[t] = x;
    __syncthreads();
    for (int off = 1; off < kSortThreads; off <<= 1) {
        const uint32_t y = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += y;
        __syncthreads();
    }
    *total = sh[kSortThreads - 1];
    const uint32_t incl = sh[t];
    __syncthreads();
    return incl - x;
}

// One block per digit: hist[d][*] (rows of `tiles` entries) becomes its exclusive
// prefix over the tiles that hold keys, and total[d] the digit's count.
__global__ __launch_bounds__(kSortThreads) void k_sort_scan(uint32_t *__restrict__ hist, int64_t tiles, int64_t n_cap,
                                                            const int32_t *n_dev, uint32_t *__restrict__ total) {
    __shared__ uint32_t sh[kSortThreads];
    uint32_t *row = hist + (int64_t)blockIdx.x * tiles;
    const int64_t used = (sort_count(n_cap, n_dev) + kSortTile - 1) / kSortTile;
    const int64_t per = (used + kSortThreads - 1) / kSortThreads;
    const int64_t lo = min(used, (int64_t)threadIdx.x * per), hi = min(used, lo + per);
    uint32_t s = 0u;
    for (int64_t t = lo; t < hi; ++t) s += row[t];
    uint32_t all;
    uint32_t run = block_scan_u32(s, sh, &all);
    for (int64_t t = lo; t < hi; ++t) {
        const uint32_t c = row[t];
        row[t] = run;
        run += c;
    }
    if (threadIdx.x == 0) total[blockIdx.x] = all;
}

// One pass's stable scatter. vals_in null: the value is the key's own index (first
// pass); keys_out null: only the values are written (last pass).
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint32_t *__restrict__ keys_in,
                                                               const uint32_t *__restrict__ vals_in, int64_t n_cap,
                                                               const int32_t *n_dev, int shift,
                                                               const uint32_t *__restrict__ prefix,
                                                               const uint32_t *__restrict__ total, int64_t tiles,
                                                               uint32_t *__restrict__ keys_out,
                                                               uint32_t *__restrict__ vals_out) {
    const int64_t n = sort_count(n_cap, n_dev);
    if ((int64_t)blockIdx.x * kSortTile >= n) return;
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t wcnt[kSortWaves][256];   // per wave: running, then total count of each digit
    __shared__ uint32_t tstart[256];             // first LDS slot of each digit in the tile
    __shared__ uint32_t gstart[256];             // global position of the tile's first key of each digit
    __shared__ uint32_t sh[kSortThreads];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int d = threadIdx.x; d < kSortWaves * 256; d += kSortThreads) (&wcnt[0][0])[d] = 0u;
    __syncthreads();

    // each wave ranks its 1024 consecutive keys: round r holds keys w*1024 + r*64 + lane,
    // so (wave, round, lane) is the input order
    const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)wave * kWaveChunk;
    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
    // all of the lane's keys and values in flight at once, then the ranking rounds
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = base + r * 64 + lane;
        const bool valid = i < n;
        key[r] = valid ? keys_in[i] : 0u;
        val[r] = valid ? (vals_in ? vals_in[i] : (uint32_t)i) : 0u;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = base + r * 64 + lane;
        const bool valid = i < n;
        const uint32_t d = (key[r] >> shift) & 0xffu;
        const uint64_t peers = digit_peers(d, valid);
        const uint32_t before = valid ? wcnt[wave][d] : 0u;
        rank[r] = before + (uint32_t)__popcll(peers & lanes_below());
        if (valid && (peers & lanes_below()) == 0ull) wcnt[wave][d] = before + (uint32_t)__popcll(peers);
        __asm__ volatile("" ::: "memory");   // the next round reads the counts this one wrote
    }
    __syncthreads();

    // per digit: the tile's count, each wave's offset within the digit, the digit's
    // first LDS slot (exclusive scan over digits) and its global position
    const int d = threadIdx.x;
    uint32_t cnt = 0u;
    for (int w = 0; w < kSortWaves; ++w) {
        const uint32_t c = wcnt[w][d];
        wcnt[w][d] = cnt;
        cnt += c;
    }
    uint32_t tile_n;
    const uint32_t start = block_scan_u32(cnt, sh, &tile_n);
    uint32_t all;
    const uint32_t dbase = block_scan_u32(total[d], sh, &all);
    tstart[d] = start;
    gstart[d] = dbase + prefix[(int64_t)d * tiles + blockIdx.x];
    __syncthreads();

    // stage in digit order
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = base + r * 64 + lane;
        if (i < n) {
            const uint32_t dd = (key[r] >> shift) & 0xffu;
            const uint32_t p = tstart[dd] + wcnt[wave][dd] + rank[r];
            if (p < (uint32_t)kSortTile) {
                s_keys[p] = key[r];
                s_vals[p] = val[r];
            }
        }
    }
    __syncthreads();

    // write each digit's run: slot p of digit d goes to gstart[d] + (p - tstart[d])
    for (int p = threadIdx.x; p < (int)tile_n; p += kSortThreads) {
        const uint32_t k = s_keys[p];
        const uint32_t dd = (k >> shift) & 0xffu;
        const uint64_t g = (uint64_t)gstart[dd] + (uint64_t)(p - tstart[dd]);
        if (g < (uint64_t)n) {
            if (keys_out) keys_out[g] = k;
            vals_out[g] = s_vals[p];
        }
    }
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

hipError_t reorder_workspace_bytes(int64_t n, size_t *bytes, size_t *sort_bytes) {
    const int64_t tiles = (n + kSortTile - 1) / kSortTile;
    *sort_bytes = align256((size_t)256 * tiles * 4) + align256(256 * 4);
    *bytes = 4 * align256((size_t)n * 4) + *sort_bytes;
    return hipSuccess;
}

static hipError_t sort_rays(const tt_ray *rays, int64_t n, const int32_t *n_dev, const float lo[3], const float hi[3],
                            void *ws, hipStream_t s, const uint32_t **order) {
    const size_t arr = align256((size_t)n * 4);
    const int64_t tiles = (n + kSortTile - 1) / kSortTile;
    char *p = static_cast<char *>(ws);
    uint32_t *k0 = reinterpret_cast<uint32_t *>(p), *k1 = reinterpret_cast<uint32_t *>(p + arr);
    uint32_t *v0 = reinterpret_cast<uint32_t *>(p + 2 * arr), *v1 = reinterpret_cast<uint32_t *>(p + 3 * arr);
    uint32_t *hist = reinterpret_cast<uint32_t *>(p + 4 * arr);
    uint32_t *total = reinterpret_cast<uint32_t *>(p + 4 * arr + align256((size_t)256 * tiles * 4));
    float scale[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = hi[a] - lo[a];
        scale[a] = ext > 0.0f ? 256.0f / ext : 0.0f;
    }
    const dim3 grid((unsigned)tiles), block(kSortThreads);
    hipLaunchKernelGGL(k_sort_keys, grid, block, 0, s, rays, n, n_dev, lo[0], lo[1], lo[2], scale[0], scale[1],
                       scale[2], k0, hist, tiles);
    // pass 0: k0 -> (k1, v1); pass 1: (k1, v1) -> (k0, v0); pass 2: (k0, v0) -> v1
    const uint32_t *kin[3] = {k0, k1, k0};
    const uint32_t *vin[3] = {nullptr, v1, v0};
    uint32_t *kout[3] = {k1, k0, nullptr};
    uint32_t *vout[3] = {v1, v0, v1};
    for (int pass = 0; pass < 3; ++pass) {
        if (pass > 0) hipLaunchKernelGGL(k_sort_hist, grid, block, 0, s, kin[pass], n, n_dev, 8 * pass, hist, tiles);
        hipLaunchKernelGGL(k_sort_scan, dim3(256), block, 0, s, hist, tiles, n, n_dev, total);
        hipLaunchKernelGGL(k_sort_scatter, grid, block, 0, s, kin[pass], vin[pass], n, n_dev, 8 * pass,
                           (const uint32_t *)hist, (const uint32_t *)total, tiles, kout[pass], vout[pass]);
    }
    *order = v1;
    return hipGetLastError();
}

hipError_t reorder_rays(const tt_ray *rays, int64_t n, const float lo[3], const float hi[3], void *ws,
                        size_t sort_bytes, hipStream_t s, const uint32_t **order) {
    (void)sort_bytes;
    if (n <= 0) {
        *order = nullptr;
        return hipSuccess;
    }
    return sort_rays(rays, n, nullptr, lo, hi, ws, s, order);
}

hipError_t reorder_rays_dev(const tt_ray *rays, int64_t cap, const int32_t *n_dev, const float lo[3],
                            const float hi[3], void *ws, hipStream_t s, const uint32_t **order) {
    if (cap <= 0) {
        *order = nullptr;
        return hipSuccess;
    }
    return sort_rays(rays, cap, n_dev, lo, hi, ws, s, order);
}

}  // namespace tt
