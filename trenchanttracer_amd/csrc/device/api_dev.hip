// C ABI, device half: context, scene upload, traversal entry points.
#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../host/tt_host.h"
#include "dev_common.h"
#include "scene_dev.h"

using namespace tt;

#define TT_HIP(call, what)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return fail(TT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace tt {

// Octahedral encoding of a unit vector in two 16-bit signed-normalised halves
// (x in the low half): n / (|x| + |y| + |z|), the lower hemisphere folded over the
// diagonals, each coordinate rounded to the nearest of 32767 steps. The shade stage
// decodes it (oct_decode in render.hip).
uint32_t oct_encode(float x, float y, float z) {
    const float l1 = std::fabs(x) + std::fabs(y) + std::fabs(z);
    if (!(l1 > 0.0f)) return 0x7FFF0000u;   // degenerate: +y
    float u = x / l1, v = y / l1;
    if (z < 0.0f) {
        const float fu = (1.0f - std::fabs(v)) * (u >= 0.0f ? 1.0f : -1.0f);
        const float fv = (1.0f - std::fabs(u)) * (v >= 0.0f ? 1.0f : -1.0f);
        u = fu;
        v = fv;
    }
    auto q = [](float t) -> uint32_t {
        const float c = std::min(1.0f, std::max(-1.0f, t));
        return (uint32_t)(uint16_t)(int16_t)std::lround(c * 32767.0f);
    };
    return q(u) | (q(v) << 16);
}

tt_status dev_alloc(tt_ctx *ctx, void **p, size_t bytes, const char *what) {
    (void)ctx;
    *p = nullptr;
    if (bytes == 0) return TT_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? TT_ERR_NOMEM : TT_ERR_DEVICE,
                    std::string(what) + ": hipMalloc failed: " + hipGetErrorString(e));
    }
    return TT_OK;
}

tt_status resolve_cfg(const tt_trace_cfg *cfg, tt_trace_cfg &out) {
    out.variant = 0;
    out.lds_stack = 16;
    out.block = 128;
    out.xcd_order = 0;
    out.refill = 0;
    out.reorder = 0;
    if (cfg) out = *cfg;
    if (out.reorder < 0 || out.reorder > 1) return fail(TT_ERR_INVALID, "trace cfg: reorder must be 0 or 1");
    if (out.refill < 0 || out.refill > 64) return fail(TT_ERR_INVALID, "trace cfg: refill must be 0..64");
    if (out.variant < 0 || out.variant > 1) return fail(TT_ERR_INVALID, "trace cfg: variant must be 0 or 1");
    if (out.lds_stack != 8 && out.lds_stack != 16 && out.lds_stack != 24 && out.lds_stack != 32)
        return fail(TT_ERR_INVALID, "trace cfg: lds_stack must be 8, 16, 24 or 32");
    if (out.block < 64 || out.block > 256 || out.block % 64 != 0)
        return fail(TT_ERR_INVALID, "trace cfg: block must be 64, 128, 192 or 256");
    return TT_OK;
}

// Workspace of tt_trace_closest_noalloc: a file-scope device array, so that the entry
// point allocates nothing. Each GPU the library runs on holds its own copy at its own
// address until the process ends. In host code the array's name is not a device
// address: the address of the current device's copy comes from hipGetSymbolAddress
// (once per device, when a context is created on it) and only that address is used.
constexpr int64_t kNoallocMaxRays = (int64_t)1 << 22;
constexpr size_t kNoallocScratchBytes = (size_t)72 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
static void *g_scratch_ptr[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];
static std::once_flag g_scratch_once[kMaxDevices];

// Called with `device` current.
static void scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::call_once(g_scratch_once[device], [device] {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch_err[device] = e;
        g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
    });
}

}  // namespace tt

extern "C" {

tt_status tt_ctx_create(int32_t device, tt_ctx **out) {
    if (!out) return fail(TT_ERR_INVALID, "tt_ctx_create: null out");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(TT_ERR_UNSUPPORTED, "tt_ctx_create: no HIP device available");
    if (device < 0 || device >= count) return fail(TT_ERR_INVALID, "tt_ctx_create: device index out of range");
    TT_HIP(hipSetDevice(device), "hipSetDevice");
    auto *c = new (std::nothrow) tt_ctx();
    if (!c) return fail(TT_ERR_NOMEM, "tt_ctx_create: host allocation failed");
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->num_cus = prop.multiProcessorCount;
        std::strncpy(c->arch, prop.gcnArchName, sizeof(c->arch) - 1);
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(TT_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    e = hipEventCreateWithFlags(&c->reorder_done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->shade_done, hipEventDisableTiming);
    for (int k = 0; k < 2; ++k)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->shadow_done[k], hipEventDisableTiming);
    if (e != hipSuccess) {
        if (c->shadow_done[0]) hipEventDestroy(c->shadow_done[0]);
        if (c->shade_done) hipEventDestroy(c->shade_done);
        if (c->aux) hipStreamDestroy(c->aux);
        if (c->reorder_done) hipEventDestroy(c->reorder_done);
        hipStreamDestroy(c->stream);
        delete c;
        return fail(TT_ERR_DEVICE, std::string("tt_ctx_create: ") + hipGetErrorString(e));
    }
    scratch_setup(device);
    *out = c;
    return TT_OK;
}

void tt_ctx_destroy(tt_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    hipStreamSynchronize(ctx->aux);
    if (ctx->scratch) hipFree(ctx->scratch);
    if (ctx->reorder_ws) hipFree(ctx->reorder_ws);
    if (ctx->reorder_done) hipEventDestroy(ctx->reorder_done);
    if (ctx->shade_done) hipEventDestroy(ctx->shade_done);
    for (hipEvent_t ev : ctx->shadow_done)
        if (ev) hipEventDestroy(ev);
    hipStreamDestroy(ctx->aux);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

tt_status tt_ctx_sync(tt_ctx *ctx) {
    if (!ctx) return fail(TT_ERR_INVALID, "tt_ctx_sync: null ctx");
    TT_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TT_OK;
}

tt_status tt_dev_alloc(tt_ctx *ctx, size_t bytes, void **out) {
    if (!ctx || !out) return fail(TT_ERR_INVALID, "tt_dev_alloc: bad arguments");
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    return dev_alloc(ctx, out, bytes, "tt_dev_alloc");
}

tt_status tt_dev_free(tt_ctx *ctx, void *p) {
    if (!ctx) return fail(TT_ERR_INVALID, "tt_dev_free: null ctx");
    if (!p) return TT_OK;
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TT_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    TT_HIP(hipFree(p), "hipFree");
    return TT_OK;
}

tt_status tt_memcpy_h2d(tt_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(TT_ERR_INVALID, "tt_memcpy_h2d: bad arguments");
    if (!bytes) return TT_OK;
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync h2d");
    TT_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TT_OK;
}

tt_status tt_memcpy_d2h(tt_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(TT_ERR_INVALID, "tt_memcpy_d2h: bad arguments");
    if (!bytes) return TT_OK;
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync d2h");
    TT_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TT_OK;
}

tt_status tt_memset_d(tt_ctx *ctx, void *dst, int value, size_t bytes) {
    if (!ctx || (bytes && !dst)) return fail(TT_ERR_INVALID, "tt_memset_d: bad arguments");
    if (!bytes) return TT_OK;
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TT_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream), "hipMemsetAsync");
    TT_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TT_OK;
}

namespace {
__global__ __launch_bounds__(256) void k_copy16(float4 *__restrict__ dst, const float4 *__restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}
constexpr int kGatherPerLane = 16;

__global__ __launch_bounds__(256) void k_gather64(const float4 *__restrict__ buf, uint64_t nblocks, uint32_t seed,
                                                  int64_t lanes, uint32_t *__restrict__ sink) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lanes) return;
    uint32_t h = (uint32_t)t * 0x9E3779B9u ^ seed;
    uint32_t acc = 0;
    float4 v[kGatherPerLane][4];
#pragma unroll
    for (int i = 0; i < kGatherPerLane; ++i) {
        h ^= h << 13;
        h ^= h >> 17;
        h ^= h << 5;
        uint64_t blk = ((uint64_t)h * (uint64_t)(h ^ 0x5bd1e995u)) % nblocks;
        const float4 *p = buf + blk * 4;
        v[i][0] = p[0];
        v[i][1] = p[1];
        v[i][2] = p[2];
        v[i][3] = p[3];
    }
#pragma unroll
    for (int i = 0; i < kGatherPerLane; ++i)
        for (int k = 0; k < 4; ++k)
            acc ^= __float_as_uint(v[i][k].x) ^ __float_as_uint(v[i][k].y) ^ __float_as_uint(v[i][k].z) ^
                   __float_as_uint(v[i][k].w);
    sink[t] = acc;
}

}  // namespace

tt_status tt_bench_gather64(tt_ctx *ctx, const void *buf, size_t bytes, int64_t fetches, uint32_t seed, float *ms) {
    if (!ctx || !buf || !ms || bytes < 64 || fetches <= 0 || ((uintptr_t)buf % 64))
        return fail(TT_ERR_INVALID, "tt_bench_gather64: bad arguments");
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    int64_t lanes = (fetches + kGatherPerLane - 1) / kGatherPerLane;
    if ((lanes + 255) / 256 > 0x7FFFFFFF) return fail(TT_ERR_INVALID, "tt_bench_gather64: too many fetches");
    void *sink = nullptr;
    tt_status st = dev_alloc(ctx, &sink, (size_t)lanes * 4, "gather sink");
    if (st != TT_OK) return st;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_gather64, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, ctx->stream,
                           (const float4 *)buf, (uint64_t)(bytes / 64), seed, lanes, (uint32_t *)sink);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    hipFree(sink);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("gather probe: ") + hipGetErrorString(e));
    return TT_OK;
}

tt_status tt_dev_copy(tt_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src)) || bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return fail(TT_ERR_INVALID, "tt_dev_copy: bad arguments");
    if (!bytes) return TT_OK;
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    size_t n = bytes / 16;
    size_t blocks = (n + 255) / 256;
    if (blocks > (size_t)ctx->num_cus * 64) blocks = (size_t)ctx->num_cus * 64;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (float4 *)dst, (const float4 *)src, n);
    TT_HIP(hipGetLastError(), "copy launch");
    return TT_OK;
}

void *tt_ctx_stream(tt_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

tt_status tt_timer_create(tt_ctx *ctx, int32_t n, tt_timer **out) {
    if (!ctx || !out || n <= 0 || n > (1 << 20)) return fail(TT_ERR_INVALID, "tt_timer_create: bad arguments");
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    auto *t = new (std::nothrow) tt_timer();
    if (!t) return fail(TT_ERR_NOMEM, "tt_timer_create: host allocation failed");
    t->ctx = ctx;
    t->ev.resize(n, nullptr);
    for (int32_t i = 0; i < n; ++i) {
        hipError_t e = hipEventCreate(&t->ev[i]);
        if (e != hipSuccess) {
            tt_timer_destroy(t);
            return fail(TT_ERR_DEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    *out = t;
    return TT_OK;
}

tt_status tt_timer_record(tt_timer *t, int32_t slot) {
    if (!t || slot < 0 || slot >= (int32_t)t->ev.size()) return fail(TT_ERR_INVALID, "tt_timer_record: bad slot");
    TT_HIP(hipEventRecord(t->ev[slot], t->ctx->stream), "hipEventRecord");
    return TT_OK;
}

tt_status tt_timer_elapsed(tt_timer *t, int32_t a, int32_t b, float *ms) {
    if (!t || !ms || a < 0 || b < 0 || a >= (int32_t)t->ev.size() || b >= (int32_t)t->ev.size())
        return fail(TT_ERR_INVALID, "tt_timer_elapsed: bad slots");
    TT_HIP(hipEventSynchronize(t->ev[b]), "hipEventSynchronize");
    TT_HIP(hipEventElapsedTime(ms, t->ev[a], t->ev[b]), "hipEventElapsedTime");
    return TT_OK;
}

void tt_timer_destroy(tt_timer *t) {
    if (!t) return;
    for (hipEvent_t e : t->ev)
        if (e) hipEventDestroy(e);
    delete t;
}

tt_status tt_scene_create(tt_ctx *ctx, const tt_mesh *mesh, const tt_bvh *bvh, const tt_material *mats, int32_t nmats,
                          const float *env_rgb, int32_t env_w, int32_t env_h, float env_scale, tt_scene **out) {
    if (!ctx || !bvh || !out) return fail(TT_ERR_INVALID, "tt_scene_create: null argument");
    if (nmats < 0 || (nmats > 0 && !mats)) return fail(TT_ERR_INVALID, "tt_scene_create: bad materials");
    if (env_rgb && (env_w <= 0 || env_h <= 0 || (int64_t)env_w * env_h > (1LL << 28)))
        return fail(TT_ERR_INVALID, "tt_scene_create: bad env map size");
    const Flat &f = bvh->b.flat;
    if (f.dev_nodes.size() % 16 != 0 || f.dev_tris.size() % 12 != 0)
        return fail(TT_ERR_INVALID, "tt_scene_create: corrupt device layout");
    if (mesh) {   // shading reads the mesh's per-triangle arrays by the records' ids
        const tt_status cs = check_tri_ids(f, mesh->m.ntris());
        if (cs != TT_OK) return cs;
    }
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    auto s = new (std::nothrow) tt_scene();
    if (!s) return fail(TT_ERR_NOMEM, "tt_scene_create: host allocation failed");
    s->ctx = ctx;
    s->num_nodes = (int64_t)(f.dev_nodes.size() / 16);
    s->num_tris = (int64_t)(f.dev_tris.size() / 12);
    if (s->num_nodes > 0x7FFFFFFF / 4 || s->num_tris > 0x7FFFFFFF / 3) {
        delete s;
        return fail(TT_ERR_INVALID, "tt_scene_create: tree too large for 32-bit node / triangle addressing");
    }
    tt_status st = TT_OK;
    auto up = [&](void **dst, const void *src, size_t bytes, const char *what) {
        if (st != TT_OK) return;
        st = dev_alloc(ctx, dst, bytes, what);
        if (st != TT_OK || bytes == 0) return;
        s->bytes += bytes;
        hipError_t e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) st = fail(TT_ERR_DEVICE, std::string(what) + ": hipMemcpy: " + hipGetErrorString(e));
    };
    up((void **)&s->nodes, f.dev_nodes.data(), f.dev_nodes.size() * 4, "scene nodes");
    up((void **)&s->tris, f.dev_tris.data(), f.dev_tris.size() * 4, "scene triangles");
    if (!mesh && s->num_nodes > 0) {
        // no mesh: the scene box (ray-reordering keys) is the union of the root's child
        // boxes, words {c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y}, {c1 ...}, {c0.lo.z, c0.hi.z,
        // c1.lo.z, c1.hi.z}; an empty child box (lo > hi) adds nothing
        float w[12];
        std::memcpy(w, f.dev_nodes.data(), sizeof(w));
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int ch = 0; ch < 2; ++ch) {
            const float cl[3] = {w[4 * ch + 0], w[4 * ch + 2], w[8 + 2 * ch]};
            const float chh[3] = {w[4 * ch + 1], w[4 * ch + 3], w[9 + 2 * ch]};
            if (!(cl[0] <= chh[0] && cl[1] <= chh[1] && cl[2] <= chh[2])) continue;
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], cl[k]);
                hi[k] = std::max(hi[k], chh[k]);
            }
        }
        if (lo[0] <= hi[0])
            for (int k = 0; k < 3; ++k) {
                s->bmin[k] = lo[k];
                s->bmax[k] = hi[k];
            }
    }
    if (mesh && st == TT_OK) {
        // per-triangle shading normals (3 float4) and material ids, by original triangle id
        const Mesh &m = mesh->m;
        std::vector<float> tn((size_t)m.ntris() * 12);
        for (int32_t i = 0; i < m.ntris(); ++i)
            for (int k = 0; k < 3; ++k) {
                V3 n = m.normals[m.tri_n[3 * (size_t)i + k]];
                float *d = &tn[(size_t)i * 12 + 4 * k];
                d[0] = n.x;
                d[1] = n.y;
                d[2] = n.z;
                d[3] = 0.0f;
            }
        s->num_orig_tris = m.ntris();
        up((void **)&s->tri_normals, tn.data(), tn.size() * 4, "scene normals");
        up((void **)&s->tri_mat, m.tri_mat.data(), m.tri_mat.size() * 4, "scene material ids");
        // geometric normals and areas; emissive triangles for light sampling
        std::vector<float> geo((size_t)m.ntris() * 4);
        std::vector<int32_t> etri;
        std::vector<float> ecdf(1, 0.0f), everts;
        double earea = 0.0;
        float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
        for (int32_t i = 0; i < m.ntris(); ++i) {
            V3 v0 = m.verts[m.tri_v[3 * (size_t)i]], v1 = m.verts[m.tri_v[3 * (size_t)i + 1]],
               v2 = m.verts[m.tri_v[3 * (size_t)i + 2]];
            for (const V3 &v : {v0, v1, v2}) {
                lo[0] = std::min(lo[0], v.x), lo[1] = std::min(lo[1], v.y), lo[2] = std::min(lo[2], v.z);
                hi[0] = std::max(hi[0], v.x), hi[1] = std::max(hi[1], v.y), hi[2] = std::max(hi[2], v.z);
            }
            double ax = (double)v1.x - v0.x, ay = (double)v1.y - v0.y, az = (double)v1.z - v0.z;
            double bx = (double)v2.x - v0.x, by = (double)v2.y - v0.y, bz = (double)v2.z - v0.z;
            double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
            double len = std::sqrt(cx * cx + cy * cy + cz * cz);
            double area = 0.5 * len;
            float *g = &geo[(size_t)i * 4];
            g[0] = len > 0 ? (float)(cx / len) : 0.0f;
            g[1] = len > 0 ? (float)(cy / len) : 0.0f;
            g[2] = len > 0 ? (float)(cz / len) : 1.0f;
            g[3] = (float)area;
            int32_t mid = m.tri_mat[i];
            if (nmats > 0 && mid >= 0 && mid < nmats && area > 0) {
                const tt_material &mm = mats[mid];
                if (mm.emission[0] > 0 || mm.emission[1] > 0 || mm.emission[2] > 0) {
                    etri.push_back(i);
                    earea += area;
                    ecdf.push_back((float)earea);
                    const V3 vs[3] = {v0, v1, v2};
                    for (const V3 &v : vs) {
                        everts.push_back(v.x);
                        everts.push_back(v.y);
                        everts.push_back(v.z);
                        everts.push_back(0.0f);
                    }
                }
            }
        }
        if (m.ntris() > 0)
            for (int k = 0; k < 3; ++k) {
                s->bmin[k] = lo[k];
                s->bmax[k] = hi[k];
            }
        up((void **)&s->tri_geo, geo.data(), geo.size() * 4, "scene geometric normals");
        {
            // the path-mode shading record: one 32-byte gather per hit instead of three
            std::vector<uint32_t> rec((size_t)m.ntris() * 8, 0u);
            for (int32_t i = 0; i < m.ntris(); ++i) {
                uint32_t *r = &rec[(size_t)i * 8];
                std::memcpy(r, &geo[(size_t)i * 4], 12);
                r[3] = (uint32_t)m.tri_mat[i];
                for (int k = 0; k < 3; ++k) {
                    const V3 n = m.normals[m.tri_n[3 * (size_t)i + k]];
                    r[4 + k] = oct_encode(n.x, n.y, n.z);
                }
            }
            up((void **)&s->shade_tri, rec.data(), rec.size() * 4, "scene shading records");
        }
        if (!etri.empty()) {
            for (float &c : ecdf) c = (float)(c / earea);
            ecdf.back() = 1.0f;
            s->num_emit = (int32_t)etri.size();
            s->emit_area = (float)earea;
            up((void **)&s->emit_tri, etri.data(), etri.size() * 4, "emitter ids");
            up((void **)&s->emit_cdf, ecdf.data(), ecdf.size() * 4, "emitter cdf");
            up((void **)&s->emit_verts, everts.data(), everts.size() * 4, "emitter vertices");
        }
        for (int32_t x : m.tri_mat)
            if (x < 0 || (nmats > 0 && x >= nmats)) {
                if (st == TT_OK) st = fail(TT_ERR_INVALID, "tt_scene_create: triangle material id out of range");
                break;
            }
    }
    if (nmats > 0) up((void **)&s->mats, mats, sizeof(tt_material) * (size_t)nmats, "scene materials");
    s->num_mats = nmats;
    if (st == TT_OK && env_rgb) st = build_env(s, env_rgb, env_w, env_h, env_scale);
    if (st != TT_OK) {
        tt_scene_destroy(s);
        return st;
    }
    *out = s;
    return TT_OK;
}

void tt_scene_destroy(tt_scene *s) {
    if (!s) return;
    hipSetDevice(s->ctx->device);
    hipStreamSynchronize(s->ctx->stream);
    free_render_state(s);
    void *ptrs[] = {s->nodes, s->tris, s->tri_normals, s->tri_mat, s->mats, s->tri_geo, s->shade_tri, s->emit_tri, s->emit_cdf,
                    s->emit_verts, s->env_rgb, s->env_marg_cdf, s->env_cond_cdf, s->env_pdf_scale, s->geoms,
                    s->env_marg_cells, s->env_cells};
    for (void *p : ptrs)
        if (p) hipFree(p);
    delete s;
}

tt_status tt_scene_set_geometries(tt_scene *s, const tt_geometry *g, int32_t n) {
    if (!s || n < 0 || n > 4096 || (n > 0 && !g)) return fail(TT_ERR_INVALID, "tt_scene_set_geometries: bad arguments");
    for (int32_t k = 0; k < n; ++k) {
        if (g[k].type != TT_GEOM_SPHERE && g[k].type != TT_GEOM_PLANE)
            return fail(TT_ERR_INVALID, "tt_scene_set_geometries: unknown geometry type");
        if (g[k].material < 0 || g[k].material >= s->num_mats)
            return fail(TT_ERR_INVALID, "tt_scene_set_geometries: material index out of range");
        if (g[k].type == TT_GEOM_SPHERE && !(g[k].b[0] > 0.0f))
            return fail(TT_ERR_INVALID, "tt_scene_set_geometries: sphere radius must be > 0");
    }
    TT_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    TT_HIP(hipStreamSynchronize(s->ctx->stream), "sync");
    if (s->geoms) hipFree(s->geoms);
    s->geoms = nullptr;
    s->num_geoms = 0;
    if (n == 0) return TT_OK;
    tt_status st = dev_alloc(s->ctx, (void **)&s->geoms, sizeof(tt_geometry) * (size_t)n, "scene geometries");
    if (st != TT_OK) return st;
    TT_HIP(hipMemcpy(s->geoms, g, sizeof(tt_geometry) * (size_t)n, hipMemcpyHostToDevice), "copy geometries");
    s->num_geoms = n;
    return TT_OK;
}

tt_status tt_scene_info(const tt_scene *s, int64_t *num_nodes, int64_t *num_tris, int64_t *device_bytes) {
    if (!s) return fail(TT_ERR_INVALID, "tt_scene_info: null scene");
    if (num_nodes) *num_nodes = s->num_nodes;
    if (num_tris) *num_tris = s->num_tris;
    if (device_bytes) *device_bytes = (int64_t)s->bytes;
    return TT_OK;
}

static tt_status trace_impl(tt_scene *s, const tt_ray *rays, int64_t n, tt_hit *hits, int32_t *flags,
                            const tt_trace_cfg *cfg, void *stream, bool any, bool noalloc = false) {
    if (!s) return fail(TT_ERR_INVALID, "trace: null scene");
    if (n < 0 || (n > 0 && (!rays || (any ? !flags : !hits)))) return fail(TT_ERR_INVALID, "trace: bad buffers");
    if (n >= (int64_t)1 << 40) return fail(TT_ERR_INVALID, "trace: too many rays");
    tt_trace_cfg c;
    tt_status st = resolve_cfg(cfg, c);
    if (st != TT_OK) return st;
    if (n == 0) return TT_OK;
    TraceArgs a{};
    a.nodes = reinterpret_cast<const float4 *>(s->nodes);
    a.tris = reinterpret_cast<const float4 *>(s->tris);
    a.rays = rays;
    a.hits = hits;
    a.flags = flags;
    a.n = n;
    a.num_nodes = (int32_t)s->num_nodes;
    hipStream_t hs = stream ? (hipStream_t)stream : s->ctx->stream;
    TT_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    if (c.reorder && noalloc) {
        // the device array's copy on this device; calls that overlap on one GPU share it
        const int dev = s->ctx->device;
        if (dev >= kMaxDevices || !g_scratch_ptr[dev])
            return fail(TT_ERR_DEVICE, std::string("trace (no-alloc): workspace address unavailable: ") +
                                           hipGetErrorString(dev < kMaxDevices ? g_scratch_err[dev] : hipErrorInvalidDevice));
        size_t need = 0, sort_bytes = 0;
        TT_HIP(reorder_workspace_bytes(n, &need, &sort_bytes), "trace: reorder workspace size");
        if (need > kNoallocScratchBytes)
            return fail(TT_ERR_INVALID, "trace (no-alloc): workspace of " + std::to_string(need) +
                                            " bytes exceeds the fixed " + std::to_string(kNoallocScratchBytes));
        TT_HIP(reorder_rays(rays, n, s->bmin, s->bmax, g_scratch_ptr[dev], sort_bytes, hs, &a.order),
               "trace: reorder");
    } else if (c.reorder) {
        if (n >= ((int64_t)1 << 31)) return fail(TT_ERR_INVALID, "trace: reorder supports at most 2^31 - 1 rays per call");
        tt_ctx *ctx = s->ctx;
        size_t need = 0, sort_bytes = 0;
        TT_HIP(reorder_workspace_bytes(n, &need, &sort_bytes), "trace: reorder workspace size");
        if (ctx->reorder_ws_bytes < need) {
            // launches on any stream may still read the old workspace
            TT_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
            if (ctx->reorder_ws) hipFree(ctx->reorder_ws);
            ctx->reorder_ws = nullptr;
            ctx->reorder_ws_bytes = 0;
            st = dev_alloc(ctx, &ctx->reorder_ws, need, "trace: reorder workspace");
            if (st != TT_OK) return st;
            ctx->reorder_ws_bytes = need;
        }
        // the workspace is the context's: a call on another stream waits for the last
        // launch that read it
        if (ctx->reorder_pending) TT_HIP(hipStreamWaitEvent(hs, ctx->reorder_done, 0), "hipStreamWaitEvent");
        TT_HIP(reorder_rays(rays, n, s->bmin, s->bmax, ctx->reorder_ws, sort_bytes, hs, &a.order), "trace: reorder");
    }
    TT_HIP(launch_trace(a, c.variant, any, c.lds_stack, c.block, c.xcd_order, c.refill, s->ctx->num_cus, hs), "trace launch");
    if (c.reorder && !noalloc) {
        TT_HIP(hipEventRecord(s->ctx->reorder_done, hs), "hipEventRecord");
        s->ctx->reorder_pending = true;
    }
    return TT_OK;
}

tt_status tt_debug_ray_order(tt_scene *s, const tt_ray *rays, int64_t n, int64_t cap, int32_t *order) {
    if (!s || !rays || !order || n <= 0 || n >= ((int64_t)1 << 31) || cap >= ((int64_t)1 << 31))
        return fail(TT_ERR_INVALID, "tt_debug_ray_order: bad arguments");
    tt_ctx *ctx = s->ctx;
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    const bool dev_count = cap > n;
    size_t need = 0, sort_bytes = 0;
    TT_HIP(reorder_workspace_bytes(dev_count ? cap : n, &need, &sort_bytes), "tt_debug_ray_order: workspace size");
    void *ws = nullptr;
    tt_status st = dev_alloc(ctx, &ws, need + 256, "tt_debug_ray_order workspace");
    if (st != TT_OK) return st;
    int32_t *d_n = reinterpret_cast<int32_t *>(static_cast<char *>(ws) + need);
    const int32_t n32 = (int32_t)n;
    const uint32_t *d_order = nullptr;
    hipError_t e = hipSuccess;
    if (dev_count) {
        e = hipMemcpyAsync(d_n, &n32, 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = reorder_rays_dev(rays, cap, d_n, s->bmin, s->bmax, ws, ctx->stream, &d_order);
    } else {
        e = reorder_rays(rays, n, s->bmin, s->bmax, ws, sort_bytes, ctx->stream, &d_order);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(order, d_order, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(ws);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("tt_debug_ray_order: ") + hipGetErrorString(e));
    return TT_OK;
}

tt_status tt_trace_closest(tt_scene *s, const tt_ray *rays, int64_t n, tt_hit *hits, const tt_trace_cfg *cfg,
                           void *stream) {
    return trace_impl(s, rays, n, hits, nullptr, cfg, stream, false);
}

tt_status tt_trace_any(tt_scene *s, const tt_ray *rays, int64_t n, int32_t *hit_flags, const tt_trace_cfg *cfg,
                       void *stream) {
    return trace_impl(s, rays, n, nullptr, hit_flags, cfg, stream, true);
}

tt_status tt_trace_closest_noalloc(tt_scene *s, const tt_ray *rays, int64_t n, tt_hit *hits, const tt_trace_cfg *cfg,
                                   void *stream) {
    if (n > kNoallocMaxRays)
        return fail(TT_ERR_INVALID, "tt_trace_closest_noalloc: at most " + std::to_string(kNoallocMaxRays) +
                                        " rays per call (got " + std::to_string(n) + ")");
    return trace_impl(s, rays, n, hits, nullptr, cfg, stream, false, true);
}

tt_status tt_bench_fetch_replay(tt_scene *s, const tt_ray *rays, int64_t n, const tt_trace_cfg *cfg, int32_t reps,
                                float *ms, int64_t *node_fetches, int64_t *tri_fetches, int32_t *counts_out) {
    if (!s || !rays || n <= 0 || n >= ((int64_t)1 << 31) || reps < 1 || !ms)
        return fail(TT_ERR_INVALID, "tt_bench_fetch_replay: bad arguments");
    tt_trace_cfg c;
    tt_status st = resolve_cfg(cfg, c);
    if (st != TT_OK) return st;
    tt_ctx *ctx = s->ctx;
    hipStream_t hs = ctx->stream;
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TraceArgs a{};
    a.nodes = reinterpret_cast<const float4 *>(s->nodes);
    a.tris = reinterpret_cast<const float4 *>(s->tris);
    a.rays = rays;
    a.n = n;
    a.num_nodes = (int32_t)s->num_nodes;
    if (c.reorder) {
        size_t need = 0, sort_bytes = 0;
        TT_HIP(reorder_workspace_bytes(n, &need, &sort_bytes), "fetch replay: reorder workspace size");
        if (ctx->reorder_ws_bytes < need) {
            TT_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
            if (ctx->reorder_ws) hipFree(ctx->reorder_ws);
            ctx->reorder_ws = nullptr;
            ctx->reorder_ws_bytes = 0;
            st = dev_alloc(ctx, &ctx->reorder_ws, need, "fetch replay: reorder workspace");
            if (st != TT_OK) return st;
            ctx->reorder_ws_bytes = need;
        }
        if (ctx->reorder_pending) TT_HIP(hipStreamWaitEvent(hs, ctx->reorder_done, 0), "hipStreamWaitEvent");
        TT_HIP(reorder_rays(rays, n, s->bmin, s->bmax, ctx->reorder_ws, sort_bytes, hs, &a.order),
               "fetch replay: reorder");
    }
    // per-ray counts (node visits, triangle tests, leaf visits), then the outcome of
    // every node visit, 4 bits each, at the counts' prefix sums (in 32-bit words)
    void *d_counts = nullptr, *d_off = nullptr, *d_log = nullptr, *d_out = nullptr;
    auto cleanup = [&] {
        hipStreamSynchronize(hs);
        for (void *p : {d_counts, d_off, d_log, d_out})
            if (p) hipFree(p);
    };
    st = dev_alloc(ctx, &d_counts, (size_t)n * 12, "fetch replay: counts");
    if (st != TT_OK) return st;
    a.rec_nodes = static_cast<int32_t *>(d_counts);
    a.rec_tris = a.rec_nodes + n;
    a.rec_leaves = a.rec_nodes + 2 * n;
    hipError_t e = launch_trace_record(a, c.variant, 1, c.block, c.refill, ctx->num_cus, hs);
    std::vector<int32_t> cnt((size_t)n * 3);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), d_counts, (size_t)n * 12, hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    if (e != hipSuccess) {
        cleanup();
        return fail(TT_ERR_DEVICE, std::string("fetch replay: counting pass: ") + hipGetErrorString(e));
    }
    std::vector<int64_t> off((size_t)n + 1);
    int64_t nodes_total = 0, tris_total = 0;
    off[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        nodes_total += cnt[(size_t)i];
        tris_total += cnt[(size_t)(n + i)];
        off[(size_t)i + 1] = off[(size_t)i] + (cnt[(size_t)i] + 7) / 8;
    }
    if (counts_out)
        for (int64_t i = 0; i < n; ++i) {
            counts_out[2 * i] = cnt[(size_t)i];
            counts_out[2 * i + 1] = cnt[(size_t)(n + i)];
        }
    const size_t log_words = (size_t)std::max<int64_t>(1, off[(size_t)n]);
    st = dev_alloc(ctx, &d_off, off.size() * 8, "fetch replay: offsets");
    if (st == TT_OK) st = dev_alloc(ctx, &d_log, log_words * 4, "fetch replay: log");
    if (st == TT_OK) st = dev_alloc(ctx, &d_out, (size_t)n * 16, "fetch replay: output");
    if (st != TT_OK) {
        cleanup();
        return st;
    }
    e = hipMemcpyAsync(d_off, off.data(), off.size() * 8, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemsetAsync(d_log, 0, log_words * 4, hs);
    a.rec_off = static_cast<const int64_t *>(d_off);
    a.rec_log = static_cast<int32_t *>(d_log);
    if (e == hipSuccess) e = launch_trace_record(a, c.variant, 2, c.block, c.refill, ctx->num_cus, hs);
    TraceArgs r = a;
    r.hits = static_cast<tt_hit *>(d_out);
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (e == hipSuccess) e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    // with reorder each timed replay is preceded by the same key pass and sort as a
    // traversal call, so the time bounds the whole call
    size_t sort_bytes = 0, ws_need = 0;
    if (e == hipSuccess && c.reorder) e = reorder_workspace_bytes(n, &ws_need, &sort_bytes);
    auto one = [&]() -> hipError_t {
        if (c.reorder) {
            const hipError_t er = reorder_rays(rays, n, s->bmin, s->bmax, ctx->reorder_ws, sort_bytes, hs, &r.order);
            if (er != hipSuccess) return er;
        }
        return launch_trace_record(r, c.variant, 3, c.block, c.refill, ctx->num_cus, hs);
    };
    if (e == hipSuccess) e = one();   // warm-up
    if (e == hipSuccess) e = hipEventRecord(ev[0], hs);
    for (int k = 0; k < reps && e == hipSuccess; ++k) e = one();
    if (e == hipSuccess) e = hipEventRecord(ev[1], hs);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    float total_ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&total_ms, ev[0], ev[1]);
    for (hipEvent_t x : ev)
        if (x) hipEventDestroy(x);
    cleanup();
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("fetch replay: ") + hipGetErrorString(e));
    if (c.reorder) {
        TT_HIP(hipEventRecord(ctx->reorder_done, hs), "hipEventRecord");
        ctx->reorder_pending = true;
    }
    *ms = total_ms / (float)reps;
    if (node_fetches) *node_fetches = nodes_total;
    if (tri_fetches) *tri_fetches = tris_total;
    return TT_OK;
}

static tt_status host_trace(tt_scene *s, const tt_ray *rays, int64_t n, void *out, size_t out_elem,
                            const tt_trace_cfg *cfg, bool any) {
    if (!s) return fail(TT_ERR_INVALID, "trace_host: null scene");
    if (n < 0 || (n > 0 && (!rays || !out))) return fail(TT_ERR_INVALID, "trace_host: bad buffers");
    if (n == 0) return TT_OK;
    tt_ctx *ctx = s->ctx;
    TT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    size_t need = (size_t)n * (sizeof(tt_ray) + out_elem);
    if (ctx->scratch_bytes < need) {
        if (ctx->scratch) hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        tt_status st = dev_alloc(ctx, &ctx->scratch, need, "trace_host scratch");
        if (st != TT_OK) return st;
        ctx->scratch_bytes = need;
    }
    tt_ray *dr = (tt_ray *)ctx->scratch;
    void *dout = (char *)ctx->scratch + (size_t)n * sizeof(tt_ray);
    TT_HIP(hipMemcpyAsync(dr, rays, (size_t)n * sizeof(tt_ray), hipMemcpyHostToDevice, ctx->stream), "copy rays");
    tt_status st = any ? tt_trace_any(s, dr, n, (int32_t *)dout, cfg, ctx->stream)
                       : tt_trace_closest(s, dr, n, (tt_hit *)dout, cfg, ctx->stream);
    if (st != TT_OK) return st;
    TT_HIP(hipMemcpyAsync(out, dout, (size_t)n * out_elem, hipMemcpyDeviceToHost, ctx->stream), "copy hits");
    TT_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TT_OK;
}

tt_status tt_trace_closest_host(tt_scene *s, const tt_ray *rays, int64_t n, tt_hit *hits, const tt_trace_cfg *cfg) {
    return host_trace(s, rays, n, hits, sizeof(tt_hit), cfg, false);
}

tt_status tt_trace_any_host(tt_scene *s, const tt_ray *rays, int64_t n, int32_t *hit_flags, const tt_trace_cfg *cfg) {
    return host_trace(s, rays, n, hit_flags, sizeof(int32_t), cfg, true);
}

}  // extern "C"
