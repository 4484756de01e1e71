// Ray reordering for tt_trace_closest / tt_trace_any (tt_trace_cfg.reorder = 1):
// a stable sort of the ray indices by a 24-bit Morton code of the ray origins.
//
// The traversal kernel then takes ray order[r] as its r-th ray and writes the
// result at that index, so the caller's arrays keep their order and every hit is
// the one the unsorted launch computes (each ray's traversal depends on that ray
// alone). Waves of rays that start near each other walk the same top of the tree
// and the same leaves near their origins, and fetch fewer distinct nodes.
//
// Cost per ray: the key pass reads the 16-byte origin half of the ray record and
// writes a 4-byte key and a 4-byte index; the sort (rocPRIM's LSD radix sort, 8
// bits per pass, 3 passes for 24 bits) reads and writes 8 bytes per pass. Camera
// rays share one origin and so one key: the sort is stable, and they stay in the
// caller's order.
#include <rocprim/device/device_radix_sort.hpp>

#include "dev_common.h"

namespace tt {

namespace {

constexpr unsigned kKeyBits = 24;  // 8 bits per axis

// bit i of an 8-bit value moves to bit 3i
__device__ inline uint32_t spread3(uint32_t v) {
    v &= 0xffu;
    v = (v | (v << 8)) & 0x0000f00fu;  // bits 4..7 -> 12..15
    v = (v | (v << 4)) & 0x000c30c3u;  // pairs at 0, 6, 12, 18
    v = (v | (v << 2)) & 0x00249249u;  // bits at 0, 3, ..., 21
    return v;
}

// cell 0..255 of a coordinate; NaN and values outside the box clamp to the
// nearest end (fmaxf returns the non-NaN operand)
__device__ inline uint32_t cell8(float v, float lo, float scale) {
    return (uint32_t)fminf(fmaxf((v - lo) * scale, 0.0f), 255.0f);
}

__global__ __launch_bounds__(256) void k_ray_keys(const tt_ray *__restrict__ rays, int64_t n, float lox, float loy,
                                                  float loz, float sx, float sy, float sz, uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 o = reinterpret_cast<const float4 *>(rays + i)[0];
    keys[i] = (spread3(cell8(o.x, lox, sx)) << 2) | (spread3(cell8(o.y, loy, sy)) << 1) | spread3(cell8(o.z, loz, sz));
    idx[i] = (uint32_t)i;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

hipError_t reorder_workspace_bytes(int64_t n, size_t *bytes, size_t *sort_bytes) {
    rocprim::double_buffer<uint32_t> kb(nullptr, nullptr), vb(nullptr, nullptr);
    size_t tmp = 0;
    const hipError_t e = rocprim::radix_sort_pairs(nullptr, tmp, kb, vb, (unsigned)n, 0u, kKeyBits, hipStream_t(0));
    if (e != hipSuccess) return e;
    *sort_bytes = tmp;
    *bytes = 4 * align256((size_t)n * 4) + align256(tmp);
    return hipSuccess;
}

hipError_t reorder_rays(const tt_ray *rays, int64_t n, const float lo[3], const float hi[3], void *ws,
                        size_t sort_bytes, hipStream_t s, const uint32_t **order) {
    const size_t arr = align256((size_t)n * 4);
    char *p = static_cast<char *>(ws);
    uint32_t *k0 = reinterpret_cast<uint32_t *>(p), *k1 = reinterpret_cast<uint32_t *>(p + arr);
    uint32_t *v0 = reinterpret_cast<uint32_t *>(p + 2 * arr), *v1 = reinterpret_cast<uint32_t *>(p + 3 * arr);
    void *tmp = p + 4 * arr;
    float scale[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = hi[a] - lo[a];
        scale[a] = ext > 0.0f ? 256.0f / ext : 0.0f;
    }
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_ray_keys, dim3((unsigned)blocks), dim3(256), 0, s, rays, n, lo[0], lo[1], lo[2], scale[0],
                       scale[1], scale[2], k0, v0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    rocprim::double_buffer<uint32_t> kb(k0, k1), vb(v0, v1);
    size_t tb = sort_bytes;
    e = rocprim::radix_sort_pairs(tmp, tb, kb, vb, (unsigned)n, 0u, kKeyBits, s);
    if (e != hipSuccess) return e;
    *order = vb.current();
    return hipSuccess;
}

}  // namespace tt
