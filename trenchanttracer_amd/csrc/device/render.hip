// Wavefront path tracing on gfx950: a pass keeps k samples of every pixel in
// flight (up to 2^27 paths); each bounce = extend (closest-hit traversal of the ray
// queue) -> shade (materials, emission, next-event rays, continuation rays appended
// to the next queue with a wave64 ballot + prefix popcount and one atomic per
// 256-thread block) -> shadow (any-hit traversal of the next-event queue) -> apply
// (unoccluded contributions); the pass ends by adding each pixel's paths in sample
// order. Queue lengths stay on the device; kernels read them, so the host never
// waits between stages.
//
// Reference mode restates the reference megakernel (src/core/RenderKernel.cu:27-218)
// and its samplers (include/material/DiffImpl.cuh:12-36, MetalImpl.cuh:13-41,
// RefrImpl.cuh:12-61); for the same Philox stream it computes the same bits as
// the host restatement. Path mode is the physically based tracer (no reference
// counterpart): Lambert, GGX conductor with visible-normal sampling (Heitz 2018),
// smooth dielectric, next-event estimation of emitters and of the environment
// map (marginal / conditional CDF, FP64 prefix sums), power-heuristic MIS,
// Russian roulette.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../common/philox.h"
#include "../host/tt_host.h"
#include "dev_common.h"
#include "guide_dev.h"
#include "scene_dev.h"
#include "ttmath.h"

#pragma clang fp contract(off)

struct RenderState {
    int32_t cap = 0;
    tt_ray *rays[2] = {nullptr, nullptr};
    int32_t *path[2] = {nullptr, nullptr};
    int32_t *root[2] = {nullptr, nullptr};   // root node of each queued ray's scene
    // next-event queues of even and odd bounces (a bounce's shadow work may still run
    // while the next bounce fills the other set)
    int32_t *sroot[2] = {nullptr, nullptr};
    int32_t *sample_dev = nullptr;            // first sample index of the current pass
    void *table = nullptr;                    // device SceneTable
    tt_hit *hits = nullptr;
    // reference: mask | path: throughput, w = density of the last BSDF sample, or -1 when
    // that sample was a specular (delta) one
    float4 *thr = nullptr;
    float4 *L = nullptr;         // reference: path radiance | path: x roulette weight, yzw radiance
    float4 *Lsh = nullptr;       // path: next-event radiance (xyz), added to L's at the end of the pass
    tt_ray *srays[2] = {nullptr, nullptr};     // next-event rays
    float4 *scontrib[2] = {nullptr, nullptr};  // contribution (rgb) and path id (w, as int bits)
    int32_t *socc[2] = {nullptr, nullptr};     // any-hit flags
    int32_t *counts = nullptr;   // [0], [1] queues, [2 + (b & 3)] next-event queue of bounce b
    unsigned long long *stats = nullptr;  // [0] extension rays, [1] shadow rays
    float *acc = nullptr;        // per pixel rgb
    double *acc64 = nullptr;     // per pixel rgb (FP64 accumulation)
    int32_t acc64_cap = 0;
    int32_t pix_cap = 0;
};

namespace tt {

tt_status dev_alloc(tt_ctx *ctx, void **p, size_t bytes, const char *what);

namespace {

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 F3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Vec3f::normalize (LinearMath.h:101-107)
__device__ __forceinline__ f3 vnorm(f3 a) {
    float n = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / n, a.y / n, a.z / n};
}
__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }

__device__ __forceinline__ float u01c(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

// Four uniforms in (0, 1] for (pixel, sample, slot, lane); lane 0 is the stream of
// the reference mode (the host restatement draws the same).
__device__ __forceinline__ void draws(uint32_t pix, uint32_t sample, uint32_t slot, uint32_t lane, uint32_t seed,
                                      float u[4]) {
    u32x4 o = philox4x32_10({pix, sample, slot, lane}, seed, 0x5EEDu);
    u[0] = u01c(o.x);
    u[1] = u01c(o.y);
    u[2] = u01c(o.z);
    u[3] = u01c(o.w);
}

// Slots in up to two queues for the lanes of a block of kAppendBlock threads that
// want them, with one counter update per block and queue instead of one per wave:
// a wavefront bounce appends to the same two counters from millions of waves, and
// updates of one address are serialised in L2. Every thread of the block must call
// it (it synchronises the block). Slot order follows the lanes within a block;
// blocks take their ranges in the order their updates reach L2. Back-to-back
// calls are safe: cnt is read only between the two barriers, and base is written
// only after the next call's first barrier, which every wave reaches after its
// last read of base.
constexpr int kAppendBlock = 256;
__device__ __forceinline__ void block_append2(bool w0, int32_t *c0, bool w1, int32_t *c1, int *s0, int *s1) {
    constexpr int kWaves = kAppendBlock / 64;
    __shared__ int cnt[2][kWaves];
    __shared__ int base[2];
    const unsigned long long m0 = __ballot(w0), m1 = __ballot(w1);
    const int lane = __lane_id(), wv = (int)threadIdx.x >> 6;
    if (lane == 0) {
        cnt[0][wv] = __popcll(m0);
        cnt[1][wv] = __popcll(m1);
    }
    __syncthreads();
    int p0 = 0, p1 = 0;
    for (int k = 0; k < wv; ++k) {
        p0 += cnt[0][k];
        p1 += cnt[1][k];
    }
    if (threadIdx.x < 2) {
        const int q = (int)threadIdx.x;
        int t = 0;
        for (int k = 0; k < kWaves; ++k) t += cnt[q][k];
        int32_t *c = q ? c1 : c0;
        base[q] = (t > 0 && c) ? atomicAdd(c, t) : 0;
    }
    __syncthreads();
    p0 += base[0];
    p1 += base[1];
    const unsigned long long below = (1ull << lane) - 1ull;
    *s0 = w0 ? p0 + __popcll(m0 & below) : -1;
    *s1 = w1 ? p1 + __popcll(m1 & below) : -1;
}

struct CamArgs {
    f3 pos, hax, vax, middle, horizontal, vertical;
    float focal, aperture, two_pi;
};

// Up to kMaxScenes scenes rendered in one wavefront: per scene its pixel range
// in the concatenated image, resolution, root node and camera.
constexpr int kMaxScenes = 64;
// Path slots walk 8x8 pixel tiles: on the 1080p path-trace case (64 spp, depth 16,
// sample-major path ids) the render took 0.1605 s against 0.1650 s in pixel-row
// order (4x4: 0.1606, 16x16: 0.1607), with bit-identical images.
constexpr int kSlotTile = 8;
struct SceneTable {
    int32_t n;
    int32_t tile;   // slot order inside a scene: 0 = pixel rows, T > 1 = T x T tiles
    int32_t pix_off[kMaxScenes + 1];
    int32_t w[kMaxScenes], h[kMaxScenes], root[kMaxScenes];
    CamArgs cam[kMaxScenes];
};

__device__ __forceinline__ int scene_of(const SceneTable *t, int32_t pix) {
    int lo = 0, hi = t->n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (t->pix_off[mid] <= pix) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Pixel rendered by path slot q (0..npix-1; the slot's k samples are consecutive
// path ids): scene sid, local pixel index Y * W + X within that scene. Slots walk
// bands of `tile` rows, and each band tile by tile (tile x band-height, the last
// tile of a band narrower), so consecutive slots are neighbouring pixels of one
// 8x8 block (a wave spans several pixels when k < 64) and their later rays start
// close together. The pixel's random numbers and its accumulation slot follow the
// pixel, not the slot, so images do not depend on the order.
__device__ __forceinline__ int32_t pixel_of_slot(const SceneTable *t, int32_t q, int *sid_out, int32_t *x_out,
                                                 int32_t *y_out) {
    const int sid = scene_of(t, q);
    const int32_t ql = q - t->pix_off[sid];
    const int32_t W = t->w[sid], H = t->h[sid], T = t->tile;
    int32_t X, Y;
    if (T <= 1) {
        X = ql % W;
        Y = ql / W;
    } else {
        const int32_t y0 = (ql / (W * T)) * T;
        const int32_t hb = min(T, H - y0);
        const int32_t r = ql - y0 * W;
        const int32_t x0 = (r / (T * hb)) * T;
        const int32_t wc = min(T, W - x0);
        const int32_t j = r - x0 * hb;
        X = x0 + j % wc;
        Y = y0 + j / wc;
    }
    *sid_out = sid;
    *x_out = X;
    *y_out = Y;
    return Y * W + X;
}

// Path ids of a pass of k samples per pixel are pixel-major: path p = q * k + s is
// sample s of slot q. The k samples of one pixel are neighbours in the queues, so a
// wave's camera rays pass through one pixel (or a few, when k < 64) and its later
// rays start from nearly one point: on the 1080p path-trace case (64 spp in one pass)
// the render takes 0.1450 s against 0.1566 s with sample-major ids (s * npix + q),
// with bit-identical images.
__device__ __forceinline__ int32_t path_slot(int32_t p, int32_t k) { return p / k; }
__device__ __forceinline__ int32_t path_sample(int32_t p, int32_t k) { return p % k; }
__device__ __forceinline__ int64_t path_id(int32_t q, int32_t s, int32_t k) { return (int64_t)q * k + s; }

struct GenArgs {
    const SceneTable *tab;
    int32_t npix, mode, k;   // k samples per pixel in flight: path id = slot * k + j (pixel_of_slot)
    const int32_t *sample_dev;
    uint32_t seed;
    float tmin;
    tt_ray *rays;
    int32_t *path, *root;
    float4 *thr, *L;
    int32_t *counts;
};

__global__ __launch_bounds__(256) void k_generate(GenArgs a) {
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t npix = a.npix;
    if (p == 0) {
        a.counts[0] = npix * a.k;
        a.counts[1] = 0;
        a.counts[2] = 0;
    }
    if (p >= npix * a.k) return;
    int sid;
    int32_t X, Y;
    const int32_t local = pixel_of_slot(a.tab, path_slot(p, a.k), &sid, &X, &Y);
    const int32_t W = a.tab->w[sid], H = a.tab->h[sid];
    float u[4];
    draws((uint32_t)local, (uint32_t)*a.sample_dev + (uint32_t)path_sample(p, a.k), 0u, 0u, a.seed, u);
    // camera ray, RenderKernel.cu:152-197
    const float jx = u[0] - 0.5f, jy = u[1] - 0.5f;
    const float sx = (jx + (float)X) / (float)(W - 1);
    const float sy = (jy + (float)Y) / (float)(H - 1);
    const CamArgs &c = a.tab->cam[sid];
    f3 on_plane = (c.middle + c.horizontal * ((2 * sx) - 1)) + c.vertical * ((2 * sy) - 1);
    f3 on_image = c.pos + (on_plane - c.pos) * c.focal;
    f3 ap = c.pos;
    if (c.aperture > 1e-5f) {
        float angle = c.two_pi * u[2];
        float dist = c.aperture * sqrtf(u[3]);
        float ca, sa = tm_sincos(angle, &ca);
        ap = (c.pos + c.hax * (ca * dist)) + c.vax * (sa * dist);
    }
    f3 d = vnorm(vnorm(on_image - ap));
    a.rays[p] = {ap.x, ap.y, ap.z, a.tmin, d.x, d.y, d.z, 1e20f};
    a.path[p] = p;
    a.root[p] = a.tab->root[sid];
    // reference mode: mask and path radiance. Path mode starts every path from
    // throughput 1 and roulette weight 1, which its camera-bounce shading takes as
    // constants (every generated path is in that bounce's queue), so nothing is
    // written here.
    if (a.mode != TT_RENDER_PATH) {
        a.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        if (a.L) a.L[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

__global__ void k_advance_sample(int32_t *sample_dev, int32_t k) { *sample_dev += k; }
__global__ void k_set_sample(int32_t *sample_dev, int32_t v) { *sample_dev = v; }

struct SceneView {
    const float4 *tri_normals;
    const float4 *tri_geo;
    const int32_t *tri_mat;
    const uint4 *shade_tri;   // 2 x uint4 per triangle: {ng.xyz, material}, {oct n0, n1, n2, 0}
    const tt_material *mats;
    int32_t num_mats;
    // environment
    const float4 *env;
    const float *env_marg, *env_cond;
    const MargCell *env_gmarg;           // guide cells of the marginal CDF
    const EnvCell *env_cells;            // the conditional CDFs' guide cells with their texels
    int32_t env_w, env_h;
    float env_scale, env_inv_total;
    // emitters
    const int32_t *emit_tri;
    const float *emit_cdf;
    const float4 *emit_verts;
    int32_t num_emit;
    float emit_inv_area;
    // analytic primitives
    const tt_geometry *geoms;
    int32_t num_geoms;
};

// ------------------------------------------------------------------ analytic primitives
// Parallelogram test, planeIntersect (include/geometry/PlaneImpl.cuh:10-52):
// two-sided Moller-Trumbore with |det| >= 1e-5, u and v each in [0, 1], t > 1e-5
// (double comparisons, as the literals are double). +inf when not hit.
__device__ __forceinline__ float geom_plane_t(const tt_geometry &g, f3 o, f3 d) {
    const f3 p0 = F3(g.a[0], g.a[1], g.a[2]);
    const f3 e1 = F3(g.b[0], g.b[1], g.b[2]) - p0;
    const f3 e2 = F3(g.c[0], g.c[1], g.c[2]) - p0;
    const f3 P = cross(d, e2);
    const float det = dot(e1, P);
    if ((double)det > -0.00001 && (double)det < 0.00001) return INFINITY;
    const float inv_det = 1.f / det;
    const f3 T = o - p0;
    const f3 Q = cross(T, e1);
    const float u = dot(T, P) * inv_det;
    const float v = dot(d, Q) * inv_det;
    if (u < 0.f || u > 1.f || v < 0.f || v > 1.f) return INFINITY;
    const float t = dot(e2, Q) * inv_det;
    return (double)t > 0.00001 ? t : INFINITY;
}

// Sphere test, sphereIntersect (include/geometry/SphereImpl.cuh:10-32): the nearer
// root above 0.01 (else the farther), accepted when 0.01 < t < tmax.
__device__ __forceinline__ float geom_sphere_t(const tt_geometry &g, f3 o, f3 d, float tmax) {
    const f3 op = F3(g.a[0], g.a[1], g.a[2]) - o;
    const float eps = 0.01f;
    const float b = dot(op, d);
    float disc = b * b - dot(op, op) + g.b[0] * g.b[0];
    if (disc >= 0.0f) {
        disc = sqrtf(disc);
        float t;
        const float dist = (t = b - disc) > eps ? t : ((t = b + disc) > eps ? t : 0.0f);
        if (0.01f < dist && dist < tmax) return dist;
    }
    return INFINITY;
}

__device__ __forceinline__ float geom_t(const tt_geometry &g, f3 o, f3 d, float tmax) {
    return g.type == TT_GEOM_SPHERE ? geom_sphere_t(g, o, d, tmax) : geom_plane_t(g, o, d);
}

struct GeomArgs {
    const tt_ray *rays;
    tt_hit *hits;        // closest: replaced when a primitive is strictly nearer
    int32_t *occ;        // any: set when a primitive lies inside the ray interval
    const int32_t *n_dev;
    const tt_geometry *geoms;
    int32_t num_geoms;
};

// After the triangle traversal: the per-ray loop over the geometry list
// (src/core/RenderKernel.cu:38-61); the nearest primitive below tmax wins when
// strictly nearer than the triangle hit (a miss carries t = tmax).
__global__ __launch_bounds__(256) void k_geom_hits(GeomArgs a) {
    const int32_t n = *a.n_dev;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const tt_ray r = a.rays[i];
        const tt_hit h = a.hits[i];
        const f3 o = F3(r.ox, r.oy, r.oz), d = F3(r.dx, r.dy, r.dz);
        float best = h.t;
        int32_t id = h.tri_id;
        for (int32_t k = 0; k < a.num_geoms; ++k) {
            const float t = geom_t(a.geoms[k], o, d, r.tmax);
            if (t < r.tmax && t < best) {
                best = t;
                id = -2 - k;
            }
        }
        if (id != h.tri_id) a.hits[i] = tt_hit{id, best, 0.0f, 0.0f};
    }
}

__global__ __launch_bounds__(256) void k_geom_occluded(GeomArgs a) {
    const int32_t n = *a.n_dev;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (a.occ[i]) continue;
        const tt_ray r = a.rays[i];
        const f3 o = F3(r.ox, r.oy, r.oz), d = F3(r.dx, r.dy, r.dz);
        for (int32_t k = 0; k < a.num_geoms; ++k)
            if (geom_t(a.geoms[k], o, d, r.tmax) < r.tmax) {
                a.occ[i] = 1;
                break;
            }
    }
}

// Surface at a hit: unnormalised normal (interpolated vertex normals of a
// triangle; the primitive's own normal, as the reference computes it, for analytic
// geometry), unit geometric normal and material id.
struct Surf {
    f3 nrm, ng;
    int32_t mat;
    bool analytic;
};
__device__ __forceinline__ Surf surface(const SceneView &sv, const tt_hit &h, f3 point) {
    Surf s;
    if (h.tri_id >= 0) {
        const float4 *tn = sv.tri_normals + 3 * (int64_t)h.tri_id;
        const float w = 1.0f - h.u - h.v;
        s.nrm = (xyz(tn[0]) * w + xyz(tn[1]) * h.u) + xyz(tn[2]) * h.v;
        s.ng = xyz(sv.tri_geo[h.tri_id]);
        s.mat = sv.tri_mat[h.tri_id];
        s.analytic = false;
    } else {
        const tt_geometry &g = sv.geoms[-2 - h.tri_id];
        if (g.type == TT_GEOM_SPHERE) {
            s.nrm = F3(point.x - g.a[0], point.y - g.a[1], point.z - g.a[2]);
        } else {
            const f3 p0 = F3(g.a[0], g.a[1], g.a[2]);
            s.nrm = cross(F3(g.b[0], g.b[1], g.b[2]) - p0, F3(g.c[0], g.c[1], g.c[2]) - p0);
        }
        s.ng = vnorm(s.nrm);
        s.mat = g.material;
        s.analytic = true;
    }
    return s;
}

// Octahedral decoding of oct_encode (api_dev.hip): two 16-bit signed-normalised
// coordinates, the lower hemisphere unfolded, normalised (angular error below 1e-4).
__device__ __forceinline__ f3 oct_decode(uint32_t w) {
    float u = (float)(int16_t)(w & 0xFFFFu) * (1.0f / 32767.0f);
    float v = (float)(int16_t)(w >> 16) * (1.0f / 32767.0f);
    const float z = 1.0f - fabsf(u) - fabsf(v);
    if (z < 0.0f) {
        const float fu = (1.0f - fabsf(v)) * (u >= 0.0f ? 1.0f : -1.0f);
        const float fv = (1.0f - fabsf(u)) * (v >= 0.0f ? 1.0f : -1.0f);
        u = fu;
        v = fv;
    }
    return vnorm(F3(u, v, z));
}

// The path tracer's surface at a hit: one 32-byte record per triangle (geometric
// normal, material, octahedral vertex normals) instead of the reference mode's float
// normals, geometric normals and material ids in three arrays.
__device__ __forceinline__ Surf surface_path(const SceneView &sv, const tt_hit &h, f3 point) {
    if (h.tri_id < 0 || !sv.shade_tri) return surface(sv, h, point);
    const uint4 *r = sv.shade_tri + 2 * (int64_t)h.tri_id;
    const uint4 a = r[0], b = r[1];
    Surf s;
    const float w = 1.0f - h.u - h.v;
    s.nrm = (oct_decode(b.x) * w + oct_decode(b.y) * h.u) + oct_decode(b.z) * h.v;
    s.ng = F3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
    s.mat = (int32_t)a.w;
    s.analytic = false;
    return s;
}

// Training record of a path vertex: inputs (9), wi (3), pdf, 1 / throughput after the
// vertex (3), radiance credited by the shade kernels (3), BSDF density of wi, radiance
// credited by the next-event apply kernels (3), unused. The two kinds of credit land in
// separate fields because a bounce's apply kernel may run beside the next bounce's shade
// kernel.
constexpr int kRecFloats = 24;
constexpr int kRecNext = 20;   // first field of the next-event credit
static_assert(kRecFloats == 24, "records are six 16-byte pieces");

// Guiding inputs of the shade stage (all null / zero when unguided).
struct GuideShade {
    const float *raw;        // [slot][48] mixture of this bounce's queue slots (null: unguided)
    float alpha;             // probability of sampling the guide (one-sample MIS with the BSDF)
    int32_t learn_alpha;     // 1: alpha from the network's output kGuideAlphaOut per vertex
    float *rec;              // training records [ntrain][max_depth][kRecFloats] (null: no training)
    int32_t stride, ntrain;  // training paths: (path id + toff) % stride == 0, (id + toff) / stride < ntrain
    int32_t toff;            // per-pass offset, so that the recorded pixels change from pass to pass
    float min_rough;         // conductors with at least this roughness are guided (diffuse always)
    float bmin[3], binv[3];  // scene box, for the position input
};

struct ShadeArgs {
    GuideShade gd;
    SceneView sv;
    const SceneTable *tab;
    const int32_t *sample_dev;
    tt_ray *rays_in, *rays_out;
    int32_t *path_in, *path_out;
    const int32_t *root_in;
    int32_t *root_out, *sroot;
    const tt_hit *hits;
    float4 *thr, *L;
    float4 *Lsh;       // zeroed for every path on the camera bounce
    int32_t *counts;
    int32_t cur, nxt;
    int32_t scount;    // counts[scount]: this bounce's next-event queue
    tt_ray *srays;
    float4 *scontrib;
    float *acc;
    double *acc64;
    unsigned long long *stats;
    int32_t bounce, max_depth, rr_depth;
    int32_t npix;     // pixel slots
    int32_t k;        // samples per pixel in this pass: path id = slot * k + j, sample = *sample_dev + j
    uint32_t seed;
    float inv_spp, pi, two_pi;
};

// hdrSample (HDRKernel.cuh:6-42): nearest texel, x2; the row index is clamped
// (the reference reads one row past the end for d.y near -1 with its PI).
// hdrSample (HDRKernel.cuh:6-42): longitude atan2(d.x, d.z) wrapped into [0, TWO_PI),
// latitude acos(d.y); u = lon / TWO_PI, v = lat / PI (the reference's PI, slightly
// below pi); nearest texel (int)(u * width) + (int)(v * height) * width read as a
// linear index, so u * width reaching width reads the next row's first texel as the
// reference does; times 2 and the scene's env scale. Where the reference reads past
// the end of its texture (v * height reaching height for d.y near -1, or the wrap in
// the last row) the index is clamped to the last texel.
__device__ f3 env_reference(const SceneView &sv, f3 d, float pi, float two_pi) {
    if (!sv.env) return F3(0, 0, 0);
    float lx = tm_atan2(d.x, d.z);
    lx = lx < 0.f ? lx + two_pi : lx;
    const float ly = tm_acos(d.y);
    const float u = lx / two_pi, v = ly / pi;
    const float fu = u * (float)sv.env_w, fv = v * (float)sv.env_h;
    const int64_t u2 = fu >= 0.0f ? (int64_t)fu : 0, v2 = fv >= 0.0f ? (int64_t)fv : 0;
    const int64_t last = (int64_t)sv.env_w * sv.env_h - 1;
    const int64_t idx = min(u2 + v2 * (int64_t)sv.env_w, last);
    const float4 t = sv.env[idx];
    return F3(t.x, t.y, t.z) * (2.0f * sv.env_scale);
}

// ------------------------------------------------------------------ reference shade
// One slot of the hit queue (i >= n: no path; the thread still joins the block's
// queue append).
__device__ __forceinline__ void shade_reference_one(const ShadeArgs &a, const int32_t i, const int32_t n) {
    bool cont = false;
    int32_t p = 0, root = 0;
    tt_ray nr{};
    if (i < n) {
        p = a.path_in[i];
        root = a.root_in[i];
        int sid_;
        int32_t x_, y_;
        const int32_t local = pixel_of_slot(a.tab, path_slot(p, a.k), &sid_, &x_, &y_);
        const uint32_t smp = (uint32_t)*a.sample_dev + (uint32_t)path_sample(p, a.k);
        const tt_ray r = a.rays_in[i];
        const tt_hit h = a.hits[i];
        f3 mask = xyz(a.thr[p]);
        f3 Lp = xyz(a.L[p]);
        const f3 ro = F3(r.ox, r.oy, r.oz), rd = F3(r.dx, r.dy, r.dz);
        bool finish = false;
        if (h.tri_id == -1 || h.t > 1e19f) {
            Lp = Lp + mask * env_reference(a.sv, rd, a.pi, a.two_pi);
            finish = true;
        } else {
            f3 point = ro + rd * h.t;
            const Surf sf = surface(a.sv, h, point);
            f3 nn = vnorm(sf.nrm);
            f3 nl = dot(nn, rd) < 0 ? nn : nn * -1.0f;
            const tt_material m = a.sv.mats[sf.mat];
            Lp = Lp + mask * F3(m.emission[0], m.emission[1], m.emission[2]);
            if (a.bounce == a.max_depth - 1 || m.type == TT_MAT_EMITTER) {
                finish = true;
            } else {
                float u[4];
                draws((uint32_t)local, smp, (uint32_t)(a.bounce + 1), 0u, a.seed, u);
                f3 no = ro, nd = rd;
                const f3 alb = F3(m.albedo[0], m.albedo[1], m.albedo[2]);
                if (m.type == TT_MAT_DIFFUSE) {
                    // diffSample
                    float phi = (float)(2.0 * 3.14159265358979323846 * (double)u[0]);
                    float r2 = u[1];
                    float r2s = sqrtf(r2);
                    f3 ww = vnorm(nl);
                    f3 ax = fabsf(ww.x) >= 0.1f ? F3(0, 1, 0) : F3(1, 0, 0);
                    f3 uu = vnorm(cross(ax, ww));
                    f3 vv = cross(ww, uu);
                    float c, s = tm_sincos(phi, &c);
                    nd = vnorm(((uu * c) * r2s + (vv * s) * r2s) + ww * sqrtf(1 - r2));
                    no = point + nl * 0.001f;
                    mask = mask * alb;
                } else if (m.type == TT_MAT_CONDUCTOR) {
                    // metalSample (Phong lobe, exponent 30)
                    float phi = (float)(2.0 * 3.14159265358979323846 * (double)u[0]);
                    float r2 = u[1];
                    float cos_t = tm_pow(1 - r2, 1.0f / (30.0f + 1));
                    float sin_t = sqrtf(1 - cos_t * cos_t);
                    f3 ww = vnorm(rd - (nn * 2.0f) * dot(nn, rd));
                    f3 ax = fabsf(ww.x) >= 0.1f ? F3(0, 1, 0) : F3(1, 0, 0);
                    f3 uu = vnorm(cross(ax, ww));
                    f3 vv = cross(ww, uu);
                    float c, s = tm_sincos(phi, &c);
                    nd = vnorm(((uu * c) * sin_t + (vv * s) * sin_t) + ww * cos_t);
                    no = point + nl * 0.0001f;
                    mask = mask * alb;
                } else if (m.type == TT_MAT_SPECULAR) {
                    // specSample: mirror about the unoriented normal, origin off the oriented one
                    nd = vnorm(rd - (nn * dot(nn, rd)) * 2.0f);
                    no = point + nl * 0.001f;
                    mask = mask * alb;
                } else if (m.type == TT_MAT_COAT) {
                    // coatSample: reflect with the Fresnel probability of air (1.000293) over
                    // the coat (1.291), else Lambert; reflect / refract / fresnel follow
                    // CoatImpl.cuh:12-69 (powf(x, 2) taken as the rounded square x * x)
                    const float n1 = 1.000293f, n2 = 1.291f;
                    const f3 refl = rd - (nn * 2.0f) * dot(nn, rd);
                    const float cos1 = dot(nn, rd * -1.0f);
                    const float n12 = n1 / n2;
                    const float radicand = 1 - (n12 * n12) * (1 - cos1 * cos1);
                    f3 tdir = F3(0, 0, 0);
                    if (!(radicand < 0)) {
                        const float cos2 = sqrtf(radicand);
                        tdir = cos1 > 0 ? rd * n12 + nn * (cos1 * n12 - cos2) : rd * n12 + nn * (cos1 * n12 + cos2);
                    }
                    float thr_f;
                    if ((double)sqrtf(dot(tdir, tdir)) <= 0.12345 || dot(nn, tdir) > 0) {
                        thr_f = 1.0f;
                    } else {
                        const float ci = dot(nn, rd * -1.0f), ct = dot(nn * -1.0f, tdir);
                        const float rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct);
                        const float rp = (n1 * ct - n2 * ci) / (n1 * ct + n2 * ci);
                        thr_f = (rs * rs + rp * rp) / 2.0f;
                    }
                    if (u[0] < thr_f) {
                        mask = mask * F3(m.specular[0], m.specular[1], m.specular[2]);
                        nd = vnorm(refl);
                    } else {
                        float phi = (float)(2.0 * 3.14159265358979323846 * (double)u[1]);
                        float r2 = u[2];
                        float r2s = sqrtf(r2);
                        f3 ww = vnorm(nl);
                        f3 ax = fabsf(ww.x) >= 0.1f ? F3(0, 1, 0) : F3(1, 0, 0);
                        f3 uu = vnorm(cross(ax, ww));
                        f3 vv = cross(ww, uu);
                        float c, s = tm_sincos(phi, &c);
                        nd = vnorm(((uu * c) * r2s + (vv * s) * r2s) + ww * sqrtf(1 - r2));
                        mask = mask * alb;
                    }
                    no = point + nl * 0.001f;
                } else {
                    // refrSample
                    bool into = dot(nn, nl) > 0;
                    float nc = 1.0f, nt = m.ior;
                    float nnt = into ? nc / nt : nt / nc;
                    float ddn = dot(rd, nl);
                    float cos2t = 1.0f - nnt * nnt * (1.f - ddn * ddn);
                    if (cos2t < 0.0f) {
                        nd = vnorm(rd - (nn * 2.0f) * dot(nn, rd));
                        no = point + nl * 0.001f;
                    } else {
                        f3 tdir = rd * nnt;
                        tdir = tdir - nn * ((float)(into ? 1 : -1) * (ddn * nnt + sqrtf(cos2t)));
                        tdir = vnorm(tdir);
                        float R0 = (nt - nc) * (nt - nc) / (nt + nc) * (nt + nc);
                        float c = 1.f - (into ? -ddn : dot(tdir, nn));
                        float Re = R0 + (1.f - R0) * c * c * c * c * c;
                        float Tr = 1 - Re;
                        float P = .25f + .5f * Re;
                        float RP = Re / P;
                        float TP = Tr / (1.f - P);
                        if (u[0] < 0.2f) {
                            mask = mask * RP;
                            nd = vnorm(rd - (nn * 2.0f) * dot(nn, rd));
                        } else {
                            mask = mask * TP;
                            nd = vnorm(tdir);
                        }
                        no = point + nl * 0.001f;
                    }
                }
                nr = {no.x, no.y, no.z, r.tmin, nd.x, nd.y, nd.z, r.tmax};
                cont = true;
                a.thr[p] = make_float4(mask.x, mask.y, mask.z, 0.0f);
                a.L[p] = make_float4(Lp.x, Lp.y, Lp.z, 0.0f);
            }
        }
        // the finished path's radiance stays in its slot; k_resolve_reference adds
        // it to the pixel in sample order
        if (finish) a.L[p] = make_float4(Lp.x, Lp.y, Lp.z, 0.0f);
    }
    int slot, unused;
    block_append2(cont, &a.counts[a.nxt], false, nullptr, &slot, &unused);
    if (cont) {
        a.rays_out[slot] = nr;
        a.path_out[slot] = p;
        a.root_out[slot] = root;
    }
}

// One thread per queue slot over the whole capacity; blocks wholly past the queue's
// end leave at once (block-uniform, before the append's barriers). A loop over the
// queue would save those blocks' launch but takes the shade kernels from 73 to 97
// VGPRs (6 to 4 waves per SIMD), which costs more at the full early bounces.
__global__ __launch_bounds__(kAppendBlock) void k_shade_reference(ShadeArgs a) {
    const int32_t n = a.counts[a.cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a.stats[0], (unsigned long long)n);
    if ((int64_t)blockIdx.x * kAppendBlock >= n) return;
    shade_reference_one(a, (int32_t)(blockIdx.x * kAppendBlock + threadIdx.x), n);
}

// ------------------------------------------------------------------ path mode
struct Frame {
    f3 t, b, n;
};
// Duff et al. 2017, "Building an orthonormal basis, revisited"
__device__ __forceinline__ Frame make_frame(f3 n) {
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float b = n.x * n.y * a;
    return {F3(1.0f + s * n.x * n.x * a, s * b, -s * n.x), F3(b, s + n.y * n.y * a, -n.y), n};
}
__device__ __forceinline__ f3 to_local(const Frame &f, f3 v) { return F3(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
__device__ __forceinline__ f3 to_world(const Frame &f, f3 v) { return (f.t * v.x + f.b * v.y) + f.n * v.z; }

__device__ __forceinline__ float power_heuristic(float a, float b) {
    float a2 = a * a, b2 = b * b;
    return a2 + b2 > 0.0f ? a2 / (a2 + b2) : 0.0f;
}

// GGX with alpha = roughness^2 (clamped), height-correlated Smith masking.
__device__ __forceinline__ float ggx_alpha(float rough) { return fmaxf(rough * rough, 1e-4f); }
__device__ __forceinline__ float ggx_D(f3 h, float al) {
    float x = h.x / al, y = h.y / al;
    float q = x * x + y * y + h.z * h.z;
    return 1.0f / (3.14159265f * al * al * q * q);
}
__device__ __forceinline__ float ggx_lambda(f3 w, float al) {
    float z2 = w.z * w.z;
    if (z2 <= 0.0f) return 0.0f;
    float t2 = fmaxf(0.0f, 1.0f - z2) / z2;
    return 0.5f * (-1.0f + sqrtf(1.0f + al * al * t2));
}
__device__ __forceinline__ f3 schlick(f3 f0, float c) {
    float m = fminf(fmaxf(1.0f - c, 0.0f), 1.0f);
    float m5 = m * m * m * m * m;
    return f0 + (F3(1, 1, 1) - f0) * m5;
}
// Fresnel reflectance of a smooth dielectric, eta = eta_i / eta_t.
__device__ __forceinline__ float fresnel_dielectric(float cosi, float eta, float *cost_out) {
    float sin2t = eta * eta * fmaxf(0.0f, 1.0f - cosi * cosi);
    if (sin2t >= 1.0f) {
        *cost_out = 0.0f;
        return 1.0f;
    }
    float cost = sqrtf(1.0f - sin2t);
    float rs = (eta * cosi - cost) / (eta * cosi + cost);
    float rp = (cosi - eta * cost) / (cosi + eta * cost);
    *cost_out = cost;
    return 0.5f * (rs * rs + rp * rp);
}

// Fresnel reflectance of the coat layer (air 1.000293 over 1.291) at cos(wo).
__device__ __forceinline__ float coat_fresnel(float cos_o) {
    float cost;
    return fresnel_dielectric(fminf(cos_o, 1.0f), 1.000293f / 1.291f, &cost);
}

// BSDF value times |cos(wi)| and the sampling pdf for a non-delta material;
// wo, wi in the local frame of the shading normal (z up, wo.z > 0).
__device__ void bsdf_eval(const tt_material &m, f3 wo, f3 wi, f3 *f_cos, float *pdf) {
    const f3 alb = F3(m.albedo[0], m.albedo[1], m.albedo[2]);
    *f_cos = F3(0, 0, 0);
    *pdf = 0.0f;
    if (wi.z <= 0.0f || wo.z <= 0.0f) return;
    if (m.type == TT_MAT_DIFFUSE) {
        *f_cos = alb * (wi.z * 0.318309886f);
        *pdf = wi.z * 0.318309886f;
    } else if (m.type == TT_MAT_COAT) {
        // Lambert layer under a smooth coat: the share 1 - F(wo) that enters it
        const float keep = 1.0f - coat_fresnel(wo.z);
        *f_cos = alb * (keep * wi.z * 0.318309886f);
        *pdf = keep * wi.z * 0.318309886f;
    } else if (m.type == TT_MAT_CONDUCTOR) {
        float al = ggx_alpha(m.roughness);
        f3 h = vnorm(wo + wi);
        float D = ggx_D(h, al);
        float G2 = 1.0f / (1.0f + ggx_lambda(wo, al) + ggx_lambda(wi, al));
        float G1 = 1.0f / (1.0f + ggx_lambda(wo, al));
        f3 F = schlick(alb, dot(wo, h));
        *f_cos = F * (D * G2 / (4.0f * wo.z));
        *pdf = G1 * D / (4.0f * wo.z);
    }
}

// Samples wi; returns weight = f |cos| / pdf. *delta for specular events.
__device__ bool bsdf_sample(const tt_material &m, f3 wo, bool front, const float u[4], f3 *wi, f3 *weight, float *pdf,
                            bool *delta) {
    const f3 alb = F3(m.albedo[0], m.albedo[1], m.albedo[2]);
    *delta = false;
    if (m.type == TT_MAT_DIFFUSE) {
        float r = sqrtf(u[0]), phi = 6.28318531f * u[1];
        float c, s = tm_sincos(phi, &c);
        f3 d = F3(r * c, r * s, sqrtf(fmaxf(0.0f, 1.0f - u[0])));
        *wi = d;
        *weight = alb;
        *pdf = d.z * 0.318309886f;
        return d.z > 0.0f;
    }
    if (m.type == TT_MAT_CONDUCTOR) {
        // visible-normal sampling (Heitz 2018, "Sampling the GGX Distribution of Visible Normals")
        float al = ggx_alpha(m.roughness);
        f3 vh = vnorm(F3(al * wo.x, al * wo.y, wo.z));
        float lensq = vh.x * vh.x + vh.y * vh.y;
        f3 t1 = lensq > 0.0f ? F3(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(lensq)) : F3(1, 0, 0);
        f3 t2 = cross(vh, t1);
        float r = sqrtf(u[0]), phi = 6.28318531f * u[1];
        float c, s = tm_sincos(phi, &c);
        float p1 = r * c, p2 = r * s;
        float sv = 0.5f * (1.0f + vh.z);
        p2 = (1.0f - sv) * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1)) + sv * p2;
        f3 nh = (t1 * p1 + t2 * p2) + vh * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
        f3 h = vnorm(F3(al * nh.x, al * nh.y, fmaxf(1e-6f, nh.z)));
        float woh = dot(wo, h);
        f3 d = h * (2.0f * woh) - wo;
        if (d.z <= 0.0f || woh <= 0.0f) return false;
        float D = ggx_D(h, al);
        float l_o = ggx_lambda(wo, al), l_i = ggx_lambda(d, al);
        float G1 = 1.0f / (1.0f + l_o);
        float G2 = 1.0f / (1.0f + l_o + l_i);
        *wi = d;
        *weight = schlick(alb, woh) * (G2 / G1);
        *pdf = G1 * D / (4.0f * wo.z);
        return true;
    }
    if (m.type == TT_MAT_SPECULAR) {
        *delta = true;
        *pdf = 0.0f;
        *wi = F3(-wo.x, -wo.y, wo.z);
        *weight = alb;
        return true;
    }
    if (m.type == TT_MAT_COAT) {
        // mirror with the coat's Fresnel probability, else the Lambert layer
        const float F = coat_fresnel(wo.z);
        if (u[2] <= F) {
            *delta = true;
            *pdf = 0.0f;
            *wi = F3(-wo.x, -wo.y, wo.z);
            *weight = F3(m.specular[0], m.specular[1], m.specular[2]);
            return true;
        }
        float r = sqrtf(u[0]), phi = 6.28318531f * u[1];
        float c, s = tm_sincos(phi, &c);
        f3 d = F3(r * c, r * s, sqrtf(fmaxf(0.0f, 1.0f - u[0])));
        *wi = d;
        *weight = alb;
        *pdf = (1.0f - F) * d.z * 0.318309886f;
        return d.z > 0.0f;
    }
    if (m.type == TT_MAT_DIELECTRIC) {
        float eta = front ? 1.0f / m.ior : m.ior;
        float cost;
        float F = fresnel_dielectric(wo.z, eta, &cost);
        *delta = true;
        *pdf = 0.0f;
        if (u[2] <= F) {
            *wi = F3(-wo.x, -wo.y, wo.z);
        } else {
            *wi = F3(-eta * wo.x, -eta * wo.y, -cost);
        }
        *weight = alb;
        return true;
    }
    return false;
}

// Texel (row j, column i) direction d looks up, nearest texel.
__device__ __forceinline__ void env_dir_texel(const SceneView &sv, f3 d, int *j, int *i) {
    float phi = atan2f(d.x, d.z);
    if (phi < 0.0f) phi += 6.28318531f;
    float u = phi * 0.159154943f;
    float v = acosf(fminf(1.0f, fmaxf(-1.0f, d.y))) * 0.318309886f;
    *i = min(max((int)(u * sv.env_w), 0), sv.env_w - 1);
    *j = min(max((int)(v * sv.env_h), 0), sv.env_h - 1);
}

// Radiance of texel t and the solid-angle density of sampling direction d through it.
__device__ __forceinline__ f3 env_texel_value(const SceneView &sv, f3 d, float4 t, float *pdf) {
    float sin_t = sqrtf(d.x * d.x + d.z * d.z);  // no cancellation near the poles
    *pdf = 0.0f;
    if (sin_t > 0.0f && sv.env_inv_total > 0.0f)
        *pdf = t.w * sv.env_inv_total * (float)sv.env_w * (float)sv.env_h / (19.7392088f * sin_t);
    return F3(t.x, t.y, t.z) * sv.env_scale;
}

// Environment lookup (nearest texel) and the density of sampling direction d.
__device__ f3 env_eval(const SceneView &sv, f3 d, float *pdf) {
    *pdf = 0.0f;
    if (!sv.env) return F3(0, 0, 0);
    int j, i;
    env_dir_texel(sv, d, &j, &i);
    return env_texel_value(sv, d, sv.env[(int64_t)j * sv.env_w + i], pdf);
}

// Largest index k in [0, n) with cdf[k] < u (cdf[0] = 0 < u).
__device__ __forceinline__ int cdf_search(const float *cdf, int n, float u) { return cdf_search_range(cdf, 0, n - 1, u); }

// Texel (row j, column i) the env distribution gives for (u0, u1): inversion of the
// marginal CDF through its guide table (small: it stays in cache), then of row j's
// conditional CDF through the cell record of u1, which holds the search range and,
// when the range spans at most three entries, the CDF entries and texels the search
// can need: one 64-byte read where the binary searches take 21 dependent loads at
// 2048 x 1024. The same texel as the full search for every u. Returns true with the
// texel's value in *t when it came with the record.
__device__ __forceinline__ bool env_texel(const SceneView &sv, float u0, float u1, int *j, int *i, float4 *t) {
    // cdf_search_range over [lo, hi] of a nondecreasing CDF, when the range spans at
    // most three entries: one step past lo per entry below u
    const float4 m = *(const float4 *)(sv.env_gmarg + guide_cell(u0, sv.env_h));
    const int mlo = __float_as_int(m.x), mhi = __float_as_int(m.y);
    *j = mhi - mlo <= 2 ? mlo + (mhi > mlo && m.z < u0 ? 1 : 0) + (mhi > mlo + 1 && m.w < u0 ? 1 : 0)
                        : cdf_search_range(sv.env_marg, mlo, mhi, u0);
    // the whole cell record at once (four independent 16-byte loads)
    const float4 *rec = (const float4 *)(sv.env_cells + ((int64_t)*j * sv.env_w + guide_cell(u1, sv.env_w)));
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
    const int lo = __float_as_int(r0.x), hi = __float_as_int(r0.y);
    if (hi - lo <= 2) {
        const int k = (hi > lo && r0.z < u1 ? 1 : 0) + (hi > lo + 1 && r0.w < u1 ? 1 : 0);
        *i = lo + k;
        *t = k == 0 ? r1 : k == 1 ? r2 : r3;
        return true;
    }
    *i = cdf_search_range(sv.env_cond + (int64_t)*j * (sv.env_w + 1), lo, hi, u1);
    return false;
}

__device__ f3 env_sample(const SceneView &sv, const float u[4], f3 *dir, float *pdf) {
    // texel from the marginal / conditional CDFs, uniform (u, v) inside it; the
    // density and radiance are those of the texel the direction looks up, so
    // sampling and env_eval agree even where rounding moves a direction across a
    // texel edge
    int j, i;
    float4 t;
    if (!env_texel(sv, u[0], u[1], &j, &i, &t)) t = sv.env[(int64_t)j * sv.env_w + i];
    float theta = 3.14159265f * ((float)j + (1.0f - u[2])) / (float)sv.env_h;
    float phi = 6.28318531f * ((float)i + (1.0f - u[3])) / (float)sv.env_w;
    float ct, st = tm_sincos(theta, &ct);
    float cp, sp = tm_sincos(phi, &cp);
    *dir = F3(st * sp, ct, st * cp);
    // the texel the direction looks up: the sampled one, unless rounding carried the
    // direction across a texel edge
    int jd, id;
    env_dir_texel(sv, *dir, &jd, &id);
    if (jd != j || id != i) t = sv.env[(int64_t)jd * sv.env_w + id];
    return env_texel_value(sv, *dir, t, pdf);
}

__device__ __forceinline__ int train_slot(const GuideShade &gd, int32_t p) {
    if (!gd.rec || (p + gd.toff) % gd.stride) return -1;
    const int t = (p + gd.toff) / gd.stride;
    return t < gd.ntrain ? t : -1;
}

// Adds a radiance contribution found at segment `bounce` to the incident-radiance
// estimates of the path's earlier recorded vertices: L_i += c / T_{i+1}.
// Adds contribution c, arriving at vertex `bounce`, to the radiance of the path's earlier
// recorded vertices (field f: 16 from the shade kernels, kRecNext from the apply kernels).
__device__ __forceinline__ void credit(float *rec, int t, int bounce, int max_depth, f3 c, int f) {
    for (int i = 0; i < bounce; ++i) {
        float *r = rec + ((int64_t)t * max_depth + i) * kRecFloats;
        if (r[12] > 0.0f) {
            r[f] += c.x * r[13];
            r[f + 1] += c.y * r[14];
            r[f + 2] += c.z * r[15];
        }
    }
}

__device__ __forceinline__ void guide_inputs(const GuideShade &gd, f3 point, f3 ns, f3 wo, float *in) {
    in[0] = fminf(1.0f, fmaxf(0.0f, (point.x - gd.bmin[0]) * gd.binv[0]));
    in[1] = fminf(1.0f, fmaxf(0.0f, (point.y - gd.bmin[1]) * gd.binv[1]));
    in[2] = fminf(1.0f, fmaxf(0.0f, (point.z - gd.bmin[2]) * gd.binv[2]));
    in[3] = 0.5f * (ns.x + 1.0f);
    in[4] = 0.5f * (ns.y + 1.0f);
    in[5] = 0.5f * (ns.z + 1.0f);
    in[6] = 0.5f * (wo.x + 1.0f);
    in[7] = 0.5f * (wo.y + 1.0f);
    in[8] = 0.5f * (wo.z + 1.0f);
}

// Shading normal oriented like the geometric normal towards wo (as the shade stage).
struct GuideInArgs {
    GuideShade gd;
    SceneView sv;
    const tt_ray *rays;
    const tt_hit *hits;
    const int32_t *counts;
    int32_t cur;
    float *in;          // [slot][9], written for the guided slots
    uint64_t *mask;     // [slot / 64]: bit slot % 64 set where the vertex is guided
    int32_t *block_n;   // [block]: guided slots of each block's span
    const int32_t *given;   // check hook only: the guided flag of each slot (inputs already in in)
};

// The vertices the guide samples at: diffuse surfaces and conductors at least
// min_rough rough (smooth conductors and dielectrics are delta-like: BSDF sampling).
__device__ __forceinline__ bool guide_eligible(const tt_material &m, float min_rough) {
    return m.type == TT_MAT_DIFFUSE || (m.type == TT_MAT_CONDUCTOR && m.roughness >= min_rough);
}

// Whether slot i's vertex is guided; if so, its network inputs into a.in[i].
__device__ __forceinline__ bool guide_vertex_inputs(const GuideInArgs &a, int64_t i) {
    const tt_hit h = a.hits[i];
    if (h.tri_id == -1) return false;
    const tt_ray r = a.rays[i];
    const f3 rd = F3(r.dx, r.dy, r.dz), wo = rd * -1.0f;
    const f3 point = F3(r.ox, r.oy, r.oz) + rd * h.t;
    const Surf sf = surface_path(a.sv, h, point);
    if (!guide_eligible(a.sv.mats[sf.mat], a.gd.min_rough)) return false;
    const f3 ng = sf.ng;
    const f3 ngf = dot(wo, ng) > 0.0f ? ng : ng * -1.0f;
    f3 ns = sf.analytic ? ngf : vnorm(sf.nrm);
    if (dot(ns, ngf) < 0.0f) ns = ns * -1.0f;
    guide_inputs(a.gd, point, ns, wo, a.in + i * kGuideIn);
    return true;
}

// The queue split between the blocks of the guide-row kernels (k_guide_inputs and
// k_guide_slots run the same grid): block b takes slots [b span, (b + 1) span), a span
// of whole 256-slot steps, so a short queue still gets a thread per slot.
__device__ __forceinline__ int64_t guide_span(int32_t n) {
    const int64_t steps = ((int64_t)n + 255) / 256;
    return (steps + gridDim.x - 1) / gridDim.x * 256;
}

// Network inputs of the hit queue for the vertices the guide serves, a flag bit per
// slot (one ballot word per 64 slots) and each block's number of guided slots.
// k_guide_slots turns the flags into the list of guided slots in slot order, and the
// inference runs over that list alone: on the 1080p path-trace case only about a third
// of the queue is guided (sky, glass and smooth conductors are not), and skipping only
// the 16-row tiles without a guided row still left 41% of the inferred rows unneeded.
// GIVEN (tt_debug_guide_infer_slots): the flags are given per slot instead.
template <bool GIVEN>
__global__ __launch_bounds__(256) void k_guide_inputs(GuideInArgs a) {
    __shared__ int32_t wc[4];
    const int32_t n = a.counts[a.cur];
    const int64_t span = guide_span(n);
    const int64_t lo = (int64_t)blockIdx.x * span, hi = min((int64_t)n, lo + span);
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    int32_t cnt = 0;
    for (int64_t base = lo; base < hi; base += 256) {
        const int64_t i = base + threadIdx.x;
        const bool take = i < hi && (GIVEN ? a.given[i] != 0 : guide_vertex_inputs(a, i));
        const uint64_t m = __ballot(take);
        if (lane == 0) {
            a.mask[i >> 6] = m;
            cnt += __popcll(m);
        }
    }
    if (lane == 0) wc[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) a.block_n[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// One block of 1024 threads: the counts off[0, nb) -> exclusive offsets in place (fixed
// order); *total = their sum.
__device__ __forceinline__ void scan_counts(int32_t *off, int64_t nb, int32_t *total) {
    __shared__ int32_t sh[1024];
    const int t = threadIdx.x;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t lo = min(nb, (int64_t)t * per), hi = min(nb, lo + per);
    int32_t sum = 0;
#pragma unroll 8
    for (int64_t i = lo; i < hi; ++i) sum += off[i];
    sh[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int32_t y = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += y;
        __syncthreads();
    }
    int32_t run = sh[t] - sum;
    for (int64_t i = lo; i < hi; ++i) {
        const int32_t c = off[i];
        off[i] = run;
        run += c;
    }
    if (t == 1023) *total = sh[1023];
}

// The guided slots in slot order: slot_of[r] = the queue slot of inference row r. Each
// block adds the guided counts of the blocks before it (at most 16 per CU, in a fixed
// order) for its first row, then walks its span 256 slots at a time; the last block
// writes the row total.
__global__ __launch_bounds__(256) void k_guide_slots(const uint64_t *mask, const int32_t *block_n,
                                                     const int32_t *n_dev, int32_t *slot_of, int32_t *total) {
    __shared__ int32_t part[4];
    __shared__ int32_t wc[2][4];
    const int32_t n = *n_dev;
    const int64_t span = guide_span(n);
    const int64_t lo = (int64_t)blockIdx.x * span, hi = min((int64_t)n, lo + span);
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    int32_t pre = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += 256) pre += block_n[b];
    for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o);
    if (lane == 0) part[wave] = pre;
    __syncthreads();
    int32_t run = part[0] + part[1] + part[2] + part[3];
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = run + block_n[blockIdx.x];
    int par = 0;
    // wc alternates between steps: a block's next writes go to the other half, and the
    // one after that follows the barrier every reader of this half has passed
    for (int64_t base = lo; base < hi; base += 256, par ^= 1) {
        const int64_t i = base + threadIdx.x;
        const uint64_t m = mask[i >> 6];
        if (lane == 0) wc[par][wave] = __popcll(m);
        __syncthreads();
        int32_t r = run;
        for (int w = 0; w < wave; ++w) r += wc[par][w];
        if ((m >> lane) & 1ull) slot_of[r + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)i;
        run += wc[par][0] + wc[par][1] + wc[par][2] + wc[par][3];
    }
}

// GUIDED: the guide's mixture, training records and roulette rule are compiled in
// only for guided renders, so the plain kernel keeps its registers for occupancy.
template <bool GUIDED>
__device__ __forceinline__ void shade_path_one(const ShadeArgs &a, const int32_t i, const int32_t n) {
    bool cont = false, shadow = false;
    int32_t p = 0, root = 0;
    tt_ray nr{}, sr{};
    f3 scontrib = F3(0, 0, 0);
    const SceneView &sv = a.sv;
    const float p_env = sv.env ? (sv.num_emit > 0 ? 0.5f : 1.0f) : 0.0f;
    const float p_emit = sv.num_emit > 0 ? 1.0f - p_env : 0.0f;
    if (i < n) {
        p = a.path_in[i];
        root = a.root_in[i];
        int sid_;
        int32_t x_, y_;
        const int32_t local = pixel_of_slot(a.tab, path_slot(p, a.k), &sid_, &x_, &y_);
        const uint32_t smp = (uint32_t)*a.sample_dev + (uint32_t)path_sample(p, a.k);
        const tt_ray r = a.rays_in[i];
        const tt_hit h = a.hits[i];
        // the camera bounce's state is the generated one: throughput 1, roulette weight 1
        const bool first = a.bounce == 0;
        const float4 tp = first ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : a.thr[p];
        f3 thr = xyz(tp);
        const float last_pdf = tp.w;
        const bool last_delta = tp.w < 0.0f;
        float4 Ls = first ? make_float4(1.0f, 0.0f, 0.0f, 0.0f) : a.L[p];   // x: roulette weight, yzw: radiance
        if (first) a.Lsh[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const f3 ro = F3(r.ox, r.oy, r.oz), rd = F3(r.dx, r.dy, r.dz);
        f3 add = F3(0, 0, 0);
        if (h.tri_id == -1) {
            float pdf_env;
            f3 le = env_eval(sv, rd, &pdf_env);
            float w = (a.bounce == 0 || last_delta) ? 1.0f : power_heuristic(last_pdf, pdf_env * p_env);
            add = thr * le * w;
        } else {
            const f3 point = ro + rd * h.t;
            const Surf sf = surface_path(sv, h, point);
            const f3 ng = sf.ng;
            const f3 wo = rd * -1.0f;
            const bool front = dot(wo, ng) > 0.0f;
            const f3 ngf = front ? ng : ng * -1.0f;
            f3 ns = sf.analytic ? ngf : vnorm(sf.nrm);
            if (dot(ns, ngf) < 0.0f) ns = ns * -1.0f;
            const tt_material m = sv.mats[sf.mat];
            const f3 em = F3(m.emission[0], m.emission[1], m.emission[2]);
            if (front && maxc(em) > 0.0f) {
                // analytic emitters are not in the next-event list: full weight
                float w = 1.0f;
                if (a.bounce > 0 && !last_delta && !sf.analytic) {
                    float cos_l = dot(wo, ng);
                    float pdf_l = p_emit * sv.emit_inv_area * h.t * h.t / fmaxf(cos_l, 1e-12f);
                    w = power_heuristic(last_pdf, pdf_l);
                }
                add = thr * em * w;
            }
            const int ts = GUIDED ? train_slot(a.gd, p) : -1;
            float *rec_b = ts >= 0 ? a.gd.rec + ((int64_t)ts * a.max_depth + a.bounce) * kRecFloats : nullptr;
            if (rec_b) rec_b[12] = 0.0f;
            if (a.bounce < a.max_depth - 1 && m.type != TT_MAT_EMITTER) {
                const float eps = 1e-4f * fmaxf(1.0f, fmaxf(fabsf(point.x), fmaxf(fabsf(point.y), fabsf(point.z))));
                const Frame fr = make_frame(ns);
                const f3 wol = to_local(fr, wo);
                float ul[4], ub[4];
                draws((uint32_t)local, smp, (uint32_t)(a.bounce + 1), 1u, a.seed, ul);
                draws((uint32_t)local, smp, (uint32_t)(a.bounce + 1), 2u, a.seed, ub);
                const bool smooth = m.type == TT_MAT_DIELECTRIC || m.type == TT_MAT_SPECULAR;
                // guiding: mixture of this slot (the vertices k_guide_inputs flagged, by
                // the same rule, so their rows were inferred)
                const bool guided = GUIDED && a.gd.raw != nullptr && guide_eligible(m, a.gd.min_rough);
                Mixture lobes;
                if (guided) decode_mixture(a.gd.raw + (int64_t)i * kGuideNOutPad16, lobes);
                const float alpha = !guided ? 0.0f
                                            : a.gd.learn_alpha
                                                  ? learned_alpha(a.gd.raw[(int64_t)i * kGuideNOutPad16 + kGuideAlphaOut])
                                                  : a.gd.alpha;
                // next-event estimation
                if (!smooth && (p_env > 0.0f || p_emit > 0.0f) && wol.z > 0.0f) {
                    f3 wi, le, ylight = F3(0, 0, 0);
                    float pdf_l = 0.0f;
                    bool pick_env = ul[0] <= p_env;
                    float uu[4] = {pick_env ? ul[0] / p_env : (ul[0] - p_env) / p_emit, ul[1], ul[2], ul[3]};
                    uu[0] = fminf(fmaxf(uu[0], 1e-7f), 1.0f);
                    if (pick_env) {
                        le = env_sample(sv, uu, &wi, &pdf_l);
                        pdf_l *= p_env;
                    } else {
                        int k = cdf_search(sv.emit_cdf, sv.num_emit, uu[0]);
                        const float4 *v = sv.emit_verts + 3 * (int64_t)k;
                        float su = sqrtf(uu[1]);
                        f3 y = (xyz(v[0]) * (1.0f - su) + xyz(v[1]) * (su * (1.0f - uu[2]))) + xyz(v[2]) * (su * uu[2]);
                        int t = sv.emit_tri[k];
                        f3 nl = xyz(sv.tri_geo[t]);
                        f3 dv = y - point;
                        float d2 = dot(dv, dv);
                        wi = dv * (1.0f / sqrtf(d2));
                        ylight = y;
                        float cos_l = -dot(wi, nl);
                        const tt_material lm = sv.mats[sv.tri_mat[t]];
                        le = F3(lm.emission[0], lm.emission[1], lm.emission[2]);
                        pdf_l = cos_l > 0.0f ? p_emit * sv.emit_inv_area * d2 / cos_l : 0.0f;
                    }
                    if (pdf_l > 0.0f && dot(wi, ngf) > 0.0f) {
                        f3 f_cos;
                        float pdf_b;
                        bsdf_eval(m, wol, to_local(fr, wi), &f_cos, &pdf_b);
                        if (guided) pdf_b = alpha * mixture_density(lobes, wi.x, wi.y, wi.z) + (1.0f - alpha) * pdf_b;
                        if (maxc(f_cos) > 0.0f) {
                            float w = power_heuristic(pdf_l, pdf_b);
                            scontrib = thr * f_cos * le * (w / pdf_l);
                            // shadow ray from the offset origin; towards an emitter it aims at
                            // the sampled point and stops short of it by a relative 1e-3 of
                            // its own length (measured from the offset origin, so that the
                            // offset cannot carry the emitter inside the tested interval)
                            f3 so = point + ngf * eps;
                            if (pick_env) {
                                sr = {so.x, so.y, so.z, 0.0f, wi.x, wi.y, wi.z, 1e20f};
                            } else {
                                f3 ds = ylight - so;
                                float ls = sqrtf(dot(ds, ds));
                                ds = ds * (1.0f / ls);
                                sr = {so.x, so.y, so.z, 0.0f, ds.x, ds.y, ds.z, ls * (1.0f - 1e-3f)};
                            }
                            shadow = maxc(scontrib) > 0.0f;
                        }
                    }
                }
                // BSDF sampling, or one-sample MIS of guide and BSDF on guided vertices
                // (defined even where no direction is sampled: such paths end)
                f3 wil = F3(0.0f, 0.0f, 1.0f), weight = F3(0.0f, 0.0f, 0.0f);
                float pdf_b = 0.0f;
                float pdf_bsdf = 0.0f;   // the BSDF's own density of the direction (guided vertices)
                bool delta = false;
                bool sampled = false;
                float w_rr;  // BSDF-sampling weight of the direction (max component), for roulette
                if (guided) {
                    float ug[4];
                    draws((uint32_t)local, smp, (uint32_t)(a.bounce + 1), 3u, a.seed, ug);
                    delta = false;
                    f3 f_cos;
                    if (ug[0] <= alpha) {
                        float d[3];
                        mixture_sample(lobes, ug[1], ug[2], ug[3], d);
                        wil = to_local(fr, F3(d[0], d[1], d[2]));
                        bsdf_eval(m, wol, wil, &f_cos, &pdf_b);
                    } else {
                        sampled = wol.z > 0.0f && bsdf_sample(m, wol, front, ub, &wil, &weight, &pdf_b, &delta);
                        f_cos = sampled ? weight * pdf_b : F3(0, 0, 0);
                    }
                    const f3 wiw = to_world(fr, wil);
                    pdf_bsdf = pdf_b;
                    const float pdf_c = alpha * mixture_density(lobes, wiw.x, wiw.y, wiw.z) + (1.0f - alpha) * pdf_b;
                    sampled = pdf_c > 0.0f && maxc(f_cos) > 0.0f && wol.z > 0.0f;
                    weight = sampled ? f_cos * (1.0f / pdf_c) : F3(0, 0, 0);
                    w_rr = sampled && pdf_b > 0.0f ? maxc(f_cos) / pdf_b : 0.0f;
                    pdf_b = pdf_c;
                } else {
                    sampled = wol.z > 0.0f && bsdf_sample(m, wol, front, ub, &wil, &weight, &pdf_b, &delta);
                    w_rr = maxc(weight);
                }
                if (sampled) {
                    f3 wi = to_world(fr, wil);
                    float side = dot(wi, ngf);
                    bool transmit = delta && wil.z < 0.0f;
                    bool ok = transmit ? side < 0.0f : side > 0.0f;
                    if (ok) {
                        thr = thr * weight;
                        bool alive = maxc(thr) > 0.0f;
                        // roulette: survival min(0.95, max throughput); guided renders use
                        // the product of BSDF-sampling weights instead, so that a direction
                        // the guide favours (small f cos / pdf, large expected radiance) is
                        // not culled for its small throughput
                        float rr = Ls.x * w_rr;
                        if (alive && a.bounce + 1 >= a.rr_depth) {
                            float q = fminf(0.95f, GUIDED ? rr : maxc(thr));
                            if (ub[3] > q) {
                                alive = false;
                            } else {
                                thr = thr * (1.0f / q);
                                rr = rr / q;
                            }
                        }
                        if (alive) {
                            f3 no = point + (side > 0.0f ? ngf : ngf * -1.0f) * eps;
                            nr = {no.x, no.y, no.z, 0.0f, wi.x, wi.y, wi.z, 1e20f};
                            cont = true;
                            a.thr[p] = make_float4(thr.x, thr.y, thr.z, delta ? -1.0f : pdf_b);
                            Ls.x = rr;
                            if (rec_b && !delta && pdf_b > 0.0f) {
                                guide_inputs(a.gd, point, ns, wo, rec_b);
                                rec_b[9] = wi.x;
                                rec_b[10] = wi.y;
                                rec_b[11] = wi.z;
                                rec_b[13] = thr.x > 0.0f ? 1.0f / thr.x : 0.0f;
                                rec_b[14] = thr.y > 0.0f ? 1.0f / thr.y : 0.0f;
                                rec_b[15] = thr.z > 0.0f ? 1.0f / thr.z : 0.0f;
                                rec_b[16] = rec_b[17] = rec_b[18] = 0.0f;
                                rec_b[kRecNext] = rec_b[kRecNext + 1] = rec_b[kRecNext + 2] = 0.0f;
                                rec_b[19] = pdf_bsdf;
                                rec_b[12] = pdf_b;
                            }
                        }
                    }
                }
            }
        }
        if (maxc(add) > 0.0f) {
            const int tc = GUIDED ? train_slot(a.gd, p) : -1;
            if (tc >= 0) credit(a.gd.rec, tc, a.bounce, a.max_depth, add, 16);
            Ls.y += add.x;
            Ls.z += add.y;
            Ls.w += add.z;
        }
        a.L[p] = Ls;   // the path's own slot: no atomics; k_resolve sums a pixel's paths
    }
    int slot, sslot;
    block_append2(cont, &a.counts[a.nxt], shadow, &a.counts[a.scount], &slot, &sslot);
    if (cont) {
        a.rays_out[slot] = nr;
        a.path_out[slot] = p;
        a.root_out[slot] = root;
    }
    if (shadow) {
        a.srays[sslot] = sr;
        a.sroot[sslot] = root;
        a.scontrib[sslot] = make_float4(scontrib.x, scontrib.y, scontrib.z, __int_as_float(p));
    }
}

// as k_shade_reference: one thread per slot, empty blocks leave at once
template <bool GUIDED>
__global__ __launch_bounds__(kAppendBlock) void k_shade_path(ShadeArgs a) {
    const int32_t n = a.counts[a.cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a.stats[0], (unsigned long long)n);
    if ((int64_t)blockIdx.x * kAppendBlock >= n) return;
    shade_path_one<GUIDED>(a, (int32_t)(blockIdx.x * kAppendBlock + threadIdx.x), n);
}

struct ApplyArgs {
    const int32_t *n_dev;   // the bounce's next-event queue length
    const int32_t *occ;
    const float4 *contrib;
    float4 *Lsh;         // per-path next-event radiance in xyz
    unsigned long long *stats;
    int32_t bounce, max_depth;
    GuideShade gd;
};

__global__ __launch_bounds__(256) void k_apply_shadow(ApplyArgs a) {
    const int32_t n = *a.n_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a.stats[1], (unsigned long long)n);
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (a.occ[i]) continue;
        const float4 c = a.contrib[i];
        const int32_t path = __float_as_int(c.w);
        const int tc = train_slot(a.gd, path);
        if (tc >= 0) credit(a.gd.rec, tc, a.bounce, a.max_depth, F3(c.x, c.y, c.z), kRecNext);
        // a path has at most one next-event ray per bounce, and the bounces' apply
        // launches run in order: plain read-modify-write
        float4 l = a.Lsh[path];
        l.x += c.x;
        l.y += c.y;
        l.z += c.z;
        a.Lsh[path] = l;
    }
}

// End of a path-mode pass: each pixel's radiance is the sum of its kk paths'
// radiance (paths q * kk + s for the pixel's slot q), added in sample order to the
// float or FP64 accumulation, so images do not depend on the order threads finish.
// Reference mode: acc += L * (1 / spp) for the pass's kk samples of each pixel in
// sample order, one multiply and one add per sample and channel as the reference
// accumulates (src/core/RenderKernel.cu:198-199), so a pass of many samples gives
// the float sums of one sample per pass bit for bit.
__device__ __forceinline__ int32_t slot_pixel(const SceneTable *t, int32_t q) {
    int sid;
    int32_t x, y;
    const int32_t local = pixel_of_slot(t, q, &sid, &x, &y);
    return t->pix_off[sid] + local;
}

// A block's 256 slots have their samples side by side (path ids q * kk + s); the
// block stages them in LDS kResolveChunk samples at a time, each pixel's chunk read
// as one 128-byte segment (a thread per slot reading its own samples would put the
// wave's 64 loads 16 * kk bytes apart). Rows are padded to kResolveChunk + 1 float4
// so that both the staging writes and each thread's row reads hit distinct banks.
constexpr int kResolveChunk = 8;
struct ResolveTile {
    float4 v[256][kResolveChunk + 1];
};

// Lsh (path mode): each path's next-event radiance, added to its other radiance (yzw).
__device__ __forceinline__ void stage_chunk(ResolveTile &t, const float4 *L, const float4 *Lsh, int32_t q0,
                                            int32_t npix, int32_t kk, int32_t s0) {
    for (int i = threadIdx.x; i < 256 * kResolveChunk; i += 256) {
        const int pl = i / kResolveChunk, j = i - pl * kResolveChunk;
        const int32_t q = q0 + pl, s = s0 + j;
        if (q < npix && s < kk) {
            const int32_t id = path_id(q, s, kk);
            float4 v = L[id];
            if (Lsh) {
                const float4 e = Lsh[id];
                v.y += e.x;
                v.z += e.y;
                v.w += e.z;
            }
            t.v[pl][j] = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_resolve_reference(const SceneTable *tab, const float4 *L, int32_t npix,
                                                           int32_t kk, float inv_spp, float *acc) {
    __shared__ ResolveTile t;
    const int32_t q0 = blockIdx.x * 256, q = q0 + threadIdx.x;
    const bool live = q < npix;
    float *o = acc + 3 * (int64_t)(live ? slot_pixel(tab, q) : 0);
    float r = 0.0f, g = 0.0f, b = 0.0f;
    if (live) {
        r = o[0];
        g = o[1];
        b = o[2];
    }
    for (int32_t s0 = 0; s0 < kk; s0 += kResolveChunk) {
        __syncthreads();
        stage_chunk(t, L, nullptr, q0, npix, kk, s0);
        __syncthreads();
        const int32_t m = min(kResolveChunk, kk - s0);
        for (int j = 0; j < m; ++j) {
            const float4 v = t.v[threadIdx.x][j];
            r = r + v.x * inv_spp;
            g = g + v.y * inv_spp;
            b = b + v.z * inv_spp;
        }
    }
    if (live) {
        o[0] = r;
        o[1] = g;
        o[2] = b;
    }
}

// Path mode: each pixel's running sum (float, or FP64 when acc64 is set) takes the
// pass's samples one at a time in sample order, continuing from the previous pass, so
// the sum is the same left fold over the pixel's samples for any pass structure
// (a batch of scenes, whose passes split the samples differently, gives each scene
// the bits of its own render).
__global__ __launch_bounds__(256) void k_resolve(const SceneTable *tab, const float4 *L, const float4 *Lsh,
                                                 int32_t npix, int32_t kk, float *acc, double *acc64) {
    __shared__ ResolveTile t;
    const int32_t q0 = blockIdx.x * 256, q = q0 + threadIdx.x;
    const bool live = q < npix;
    const int32_t pix = live ? slot_pixel(tab, q) : 0;
    double rd = 0.0, gd = 0.0, bd = 0.0;
    float rf = 0.0f, gf = 0.0f, bf = 0.0f;
    if (live) {
        if (acc64) {
            const double *o = acc64 + 3 * (int64_t)pix;
            rd = o[0];
            gd = o[1];
            bd = o[2];
        } else {
            const float *o = acc + 3 * (int64_t)pix;
            rf = o[0];
            gf = o[1];
            bf = o[2];
        }
    }
    for (int32_t s0 = 0; s0 < kk; s0 += kResolveChunk) {
        __syncthreads();
        stage_chunk(t, L, Lsh, q0, npix, kk, s0);
        __syncthreads();
        const int32_t m = min(kResolveChunk, kk - s0);
        for (int j = 0; j < m; ++j) {
            const float4 v = t.v[threadIdx.x][j];
            if (acc64) {
                rd += (double)v.y;
                gd += (double)v.z;
                bd += (double)v.w;
            } else {
                rf += v.y;
                gf += v.z;
                bf += v.w;
            }
        }
    }
    if (!live) return;
    if (acc64) {
        double *o = acc64 + 3 * (int64_t)pix;
        o[0] = rd;
        o[1] = gd;
        o[2] = bd;
    } else {
        float *o = acc + 3 * (int64_t)pix;
        o[0] = rf;
        o[1] = gf;
        o[2] = bf;
    }
}

// Training rows {inputs (9), wi (3), luminance(L_i) |cos theta_i| / pdf_i, BSDF density}
// from the path records after a pass: the network learns the cosine-weighted incident
// radiance, the product a diffuse vertex integrates (the shading normal is recovered from
// the normal inputs). A record is a row when that weight is positive and finite. Rows
// are written in record order, the same rows in the same order on every run (no
// atomics): k_collect_count counts the rows of each block of 256 records, k_collect_scan
// turns the counts into row offsets and the total (one block, fixed order), and
// k_collect_write writes each block's rows at its offset and clears the records'
// densities for the next pass. (One atomic append per wave on a single counter took
// 4.8 ms per pass of 33.5M records, serialised in L2.)
__device__ __forceinline__ bool record_row(const float *rec, int64_t e, int64_t ne, float *w) {
    *w = 0.0f;
    if (e >= ne) return false;
    const float *r = rec + e * kRecFloats;
    if (!(r[12] > 0.0f)) return false;
    const float nx = 2.0f * r[3] - 1.0f, ny = 2.0f * r[4] - 1.0f, nz = 2.0f * r[5] - 1.0f;
    const float cosw = fabsf(nx * r[9] + ny * r[10] + nz * r[11]);
    *w = (0.2126f * (r[16] + r[kRecNext]) + 0.7152f * (r[17] + r[kRecNext + 1]) + 0.0722f * (r[18] + r[kRecNext + 2])) *
         cosw / r[12];
    return *w > 0.0f && isfinite(*w);
}

__global__ __launch_bounds__(256) void k_collect_count(const float *rec, int64_t ne, int32_t *off) {
    __shared__ int32_t wc[4];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float w;
    const unsigned long long m = __ballot(record_row(rec, e, ne, &w));
    if (__lane_id() == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) off[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// One block: off[0, nb) block counts -> exclusive row offsets; off[nb] = the row total.
__global__ __launch_bounds__(1024) void k_collect_scan(int32_t *off, int64_t nb) { scan_counts(off, nb, off + nb); }

__global__ __launch_bounds__(256) void k_collect_write(float *rec, int64_t ne, const int32_t *off, float *buf,
                                                       int32_t cap) {
    __shared__ int32_t wc[4];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    float w;
    const bool valid = record_row(rec, e, ne, &w);
    const unsigned long long m = __ballot(valid);
    if (lane == 0) wc[wave] = __popcll(m);
    __syncthreads();
    int32_t slot = off[blockIdx.x] + __popcll(m & ((1ull << lane) - 1ull));
    for (int k = 0; k < wave; ++k) slot += wc[k];
    if (valid && slot < cap) {
        // 16-byte pieces: records are 96 bytes, rows 64
        const float4 *r = reinterpret_cast<const float4 *>(rec + e * kRecFloats);
        float4 *o = reinterpret_cast<float4 *>(buf + (int64_t)slot * 16);
        const float4 r0 = r[0], r1 = r[1], r2 = r[2], r4 = r[4];
        o[0] = r0;
        o[1] = r1;
        o[2] = r2;
        o[3] = make_float4(w, r4.w, 0.0f, 0.0f);
    }
    // clear the densities this pass recorded (most records of a pass stay empty)
    if (e < ne && rec[e * kRecFloats + 12] != 0.0f) rec[e * kRecFloats + 12] = 0.0f;
}

__global__ __launch_bounds__(256) void k_stage_train(const float *buf, int32_t n, float *in, float *dirs, float *wts,
                                                     float *pb) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *b = buf + (int64_t)i * 16;
    for (int k = 0; k < 9; ++k) in[(int64_t)i * 9 + k] = b[k];
    dirs[3 * (int64_t)i] = b[9];
    dirs[3 * (int64_t)i + 1] = b[10];
    dirs[3 * (int64_t)i + 2] = b[11];
    wts[i] = b[12];
    pb[i] = b[13];
}

__global__ void k_zero(float *acc, double *acc64, unsigned long long *stats, int64_t n3) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) {
        if (acc) acc[i] = 0.0f;
        if (acc64) acc64[i] = 0.0;
    }
    if (i < 2 && stats) stats[i] = 0ull;
}

// End of a path-mode render: the pixel sums become means in place, with the
// arithmetic the host would use (a float multiply by 1 / spp, or an FP64 division
// rounded to float), so the image can be copied straight into the caller's buffers.
__global__ void k_finish(float *acc, double *acc64, int64_t n3, int32_t spp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    if (acc64) {
        const double m = acc64[i] / (double)spp;
        acc64[i] = m;
        acc[i] = (float)m;
    } else {
        acc[i] = acc[i] * (1.0f / (float)spp);
    }
}

__global__ void k_debug_bsdf(tt_material m, f3 wo, bool front, int64_t n, uint32_t seed, float *out, float *out_eval) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float u[4];
    draws((uint32_t)i, (uint32_t)(i >> 32), 0x0D0Bu, 2u, seed, u);
    f3 wi = F3(0, 0, 0), w = F3(0, 0, 0);
    float pdf = 0.0f;
    bool delta = false;
    bool ok = bsdf_sample(m, wo, front, u, &wi, &w, &pdf, &delta);
    if (!ok) {
        wi = F3(0, 0, 0);
        w = F3(0, 0, 0);
        pdf = 0.0f;
    }
    float *o = out + 8 * i;
    o[0] = wi.x;
    o[1] = wi.y;
    o[2] = wi.z;
    o[3] = w.x;
    o[4] = w.y;
    o[5] = w.z;
    o[6] = pdf;
    o[7] = delta ? 1.0f : 0.0f;
    if (out_eval) {
        f3 fc;
        float pe;
        bsdf_eval(m, wo, wi, &fc, &pe);
        float *e = out_eval + 4 * i;
        e[0] = fc.x;
        e[1] = fc.y;
        e[2] = fc.z;
        e[3] = pe;
    }
}

__global__ void k_debug_env_invert(SceneView sv, int64_t n, const float *u, int32_t *out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int j, i;
    float4 t;
    env_texel(sv, u[2 * k], u[2 * k + 1], &j, &i, &t);
    out[2 * k] = j;
    out[2 * k + 1] = i;
}

__global__ void k_debug_env(SceneView sv, int64_t n, uint32_t seed, float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float u[4];
    draws((uint32_t)i, (uint32_t)(i >> 32), 0x0E0Bu, 1u, seed, u);
    f3 d;
    float pdf;
    f3 le = env_sample(sv, u, &d, &pdf);
    float pe;
    env_eval(sv, d, &pe);
    float *o = out + 8 * i;
    o[0] = d.x;
    o[1] = d.y;
    o[2] = d.z;
    o[3] = pdf;
    o[4] = le.x;
    o[5] = le.y;
    o[6] = le.z;
    o[7] = pe;
}

__global__ void k_debug_guide(const float *raw, int64_t n, uint32_t seed, float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Mixture lobes;
    decode_mixture(raw, lobes);
    float u[4], d[3];
    draws((uint32_t)i, (uint32_t)(i >> 32), 0x0F0Bu, 3u, seed, u);
    mixture_sample(lobes, u[1], u[2], u[3], d);
    float *o = out + 4 * i;
    o[0] = d[0];
    o[1] = d[1];
    o[2] = d[2];
    o[3] = mixture_density(lobes, d[0], d[1], d[2]);
}

__global__ void k_debug_math(int op, const float *x, const float *y, float *out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float c;
    if (op == 0) out[i] = tm_sincos(x[i], &c);
    else if (op == 1) {
        tm_sincos(x[i], &c);
        out[i] = c;
    } else if (op == 2) out[i] = tm_pow(x[i], y[i]);
    else if (op == 3) out[i] = tm_atan2(x[i], y[i]);
    else out[i] = tm_acos(x[i]);
}

// Host-side camera basis: the same float operations as the device code and the
// host restatement (RenderKernel.cu:131-146, evaluated once per render).
struct HF3 {
    float x, y, z;
};
HF3 hnorm(HF3 a) {
    float n = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / n, a.y / n, a.z / n};
}
HF3 hcross(HF3 a, HF3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
HF3 hadd(HF3 a, HF3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HF3 hscale(HF3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
f3 tof3(HF3 a) { return {a.x, a.y, a.z}; }

// Host-side camera basis: the same float operations as the host restatement
// (RenderKernel.cu:131-146, evaluated once per render instead of per sample).
CamArgs camera_args(const tt_camera &c, float pi, float two_pi) {
    HF3 pos = {c.eye[0], c.eye[1], c.eye[2]};
    HF3 view = hnorm({c.view[0], c.view[1], c.view[2]});
    HF3 up = hnorm({c.up[0], c.up[1], c.up[2]});
    HF3 hax = hnorm(hcross(view, up));
    HF3 vax = hnorm(hcross(hax, view));
    HF3 middle = hadd(pos, view);
    HF3 horizontal = hscale(hax, tanf(c.fov[0] * 0.5f * (pi / 180)));
    HF3 vertical = hscale(vax, tanf(-c.fov[1] * 0.5f * (pi / 180)));
    CamArgs a;
    a.pos = tof3(pos);
    a.hax = tof3(hax);
    a.vax = tof3(vax);
    a.middle = tof3(middle);
    a.horizontal = tof3(horizontal);
    a.vertical = tof3(vertical);
    a.focal = c.focal;
    a.aperture = c.aperture;
    a.two_pi = two_pi;
    return a;
}

// Guiding state of one render call: the network, its sampling share and the
// online-training schedule (paths recorded per pass, rows per Adam step).
struct GuideRun {
    tt_guide *g = nullptr;
    GuideShade gd{};          // raw / rec filled per launch
    bool train = false;
    int32_t batch_rows = 1 << 18;
    int32_t train_spp = 0;    // train only while fewer samples than this are done (0: always)
    int32_t max_bounce = 0;   // the network guides bounces below this (0: every bounce)
    int32_t growth = 1;       // each training pass this many times wider than the one before (up to the widest)
    int64_t steps = 0, rows = 0;
    double train_ms = 0.0;
    float first_loss = 0.0f;  // loss of the first Adam step
};

tt_status ensure_state(tt_scene *s, int32_t cap, int32_t npix, bool fp64) {
    RenderState *rs = s->rs;
    if (rs && rs->cap >= cap && rs->pix_cap >= npix && (!fp64 || rs->acc64_cap >= npix)) return TT_OK;
    free_render_state(s);
    rs = new (std::nothrow) RenderState();
    if (!rs) return fail(TT_ERR_NOMEM, "render state: host allocation failed");
    s->rs = rs;
    rs->cap = cap;
    rs->pix_cap = npix;
    tt_status st = TT_OK;
    auto al = [&](void **p, size_t bytes, const char *what) {
        if (st == TT_OK) st = dev_alloc(s->ctx, p, bytes, what);
    };
    for (int k = 0; k < 2; ++k) {
        al((void **)&rs->rays[k], (size_t)cap * sizeof(tt_ray), "render rays");
        al((void **)&rs->path[k], (size_t)cap * 4, "render path ids");
        al((void **)&rs->root[k], (size_t)cap * 4, "render ray roots");
    }
    al((void **)&rs->hits, (size_t)cap * sizeof(tt_hit), "render hits");
    al((void **)&rs->thr, (size_t)cap * 16, "render throughput");
    al((void **)&rs->L, (size_t)cap * 16, "render radiance");
    al((void **)&rs->Lsh, (size_t)cap * 16, "render next-event radiance");
    for (int k = 0; k < 2; ++k) {
        al((void **)&rs->srays[k], (size_t)cap * sizeof(tt_ray), "render shadow rays");
        al((void **)&rs->sroot[k], (size_t)cap * 4, "render shadow roots");
        al((void **)&rs->scontrib[k], (size_t)cap * 16, "render shadow contributions");
        al((void **)&rs->socc[k], (size_t)cap * 4, "render shadow flags");
    }
    al((void **)&rs->counts, 32, "render queue counts");
    al((void **)&rs->stats, 16, "render stats");
    al((void **)&rs->sample_dev, 16, "render sample index");
    al((void **)&rs->table, sizeof(SceneTable), "render scene table");
    al((void **)&rs->acc, (size_t)npix * 12, "render accumulation");
    if (fp64) {
        al((void **)&rs->acc64, (size_t)npix * 24, "render FP64 accumulation");
        rs->acc64_cap = npix;
    }
    if (st != TT_OK) free_render_state(s);
    return st;
}

// One bounce of the wavefront: extend -> shade (-> shadow -> apply).
// as (path mode, may be null): the stream of the next-event work. A bounce's shadow
// trace, occlusion and apply run there once its shade kernel is done, beside the next
// bounces' extension traces, guide inference and shade kernels on hs. The next-event
// buffers alternate by bounce (set b & 1), and the apply adds into the paths' own
// next-event radiance (Lsh, summed with L at the end of the pass) and credits training
// records in fields of their own, so a shade kernel waits only for the next-event work
// two bounces back, the last user of its set. The queue length of bounce b is counts[2 + (b & 3)]: the extension trace zeroes the one
// its shade kernel appends to, three bounces past the last next-event work that read
// it, which a shade kernel before it has waited for. Small late-bounce queues leave
// most of the GPU idle behind their slowest rays; the launches fill it together.
hipError_t launch_bounce(tt_scene *s, RenderState *rs, const tt_render_params *p, const SceneView &sv, int32_t cap,
                         int32_t npix, int b, bool fp64, hipStream_t hs, const GuideRun *gr, hipStream_t as) {
    const int cur = b & 1, nxt = cur ^ 1;
    const int sc = 2 + (b & 3), set = b & 1;
    const int tri_mode = p->mode == TT_RENDER_REFERENCE ? kTriReference : kTriTwoSided;
    // reference mode: the exact loop (hits identical to the host restatement);
    // path mode: the speculative while-while loop unless told otherwise
    const int variant = p->mode == TT_RENDER_REFERENCE ? 0 : (p->variant >= 0 ? p->variant : 1);
    const int lds = p->lds_stack > 0 ? p->lds_stack : 16;
    const int block = p->block > 0 ? p->block : 64;
    // the small queue kernels stride over the device-side count with 16 blocks of 256
    // threads per CU; the shade kernels take one thread per slot of the capacity
    static_assert(kAppendBlock == 256, "shade launches use the 256-thread grid");
    const unsigned gblocks =
        (unsigned)std::min<int64_t>(((int64_t)cap + 255) / 256, 16LL * std::max(1, s->ctx->num_cus));
    const unsigned sblocks = (unsigned)(((int64_t)cap + kAppendBlock - 1) / kAppendBlock);
    TraceArgs ta{};
    ta.nodes = (const float4 *)s->nodes;
    ta.tris = (const float4 *)s->tris;
    ta.rays = rs->rays[cur];
    ta.hits = rs->hits;
    ta.n = cap;
    ta.num_nodes = (int32_t)s->num_nodes;
    ta.n_dev = rs->counts + cur;
    ta.reset0 = rs->counts + nxt;
    ta.reset1 = rs->counts + sc;
    ta.ray_root = rs->root[cur];
    // camera rays (bounce 0) are coherent and their queue is full: one wave per
    // 64-ray group; later extension and shadow queues: persistent waves that refill
    // once 16 lanes are idle
    const int refill_later = p->refill > 0 ? p->refill : 16;
    const int refill_ext = b == 0 ? p->refill_camera : refill_later;
    hipError_t e = launch_trace(ta, variant, false, lds, block, 0, refill_ext, s->ctx->num_cus, hs, tri_mode);
    if (e != hipSuccess) return e;
    if (sv.num_geoms > 0) {
        GeomArgs ga{rs->rays[cur], rs->hits, nullptr, rs->counts + cur, sv.geoms, sv.num_geoms};
        hipLaunchKernelGGL(k_geom_hits, dim3(gblocks), dim3(256), 0, hs, ga);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    ShadeArgs sa{};
    // the network guides bounces below max_bounce; a training pass records every
    // bounce (the guided shade kernel with no mixture samples the BSDF alone)
    const bool guide_here = gr && (gr->max_bounce <= 0 || b < gr->max_bounce);
    const bool records = gr && gr->gd.rec != nullptr;
    if (guide_here) {
        // guide inference over the hit queue: inputs, flags and counts, the guided slots
        // in slot order, then the MLP over that list
        tt_guide *g = gr->g;
        if ((int64_t)cap > g->icap) return hipErrorInvalidValue;   // the inference rows cover the queue
        const unsigned rblocks = (unsigned)std::min<int64_t>(((int64_t)cap + 255) / 256, g->iblocks);
        int32_t *rows_dev = g->iblock_n + g->iblocks;
        GuideInArgs ga{gr->gd, sv, rs->rays[cur], rs->hits, rs->counts, cur, g->iin, g->imask, g->iblock_n, nullptr};
        hipLaunchKernelGGL(k_guide_inputs<false>, dim3(rblocks), dim3(256), 0, hs, ga);
        hipLaunchKernelGGL(k_guide_slots, dim3(rblocks), dim3(256), 0, hs, (const uint64_t *)g->imask,
                           (const int32_t *)g->iblock_n, rs->counts + cur, g->islot, rows_dev);
        e = hipGetLastError();
        if (e == hipSuccess) e = guide_infer(g, g->iin, g->iraw, cap, rows_dev, hs, g->islot);
        if (e != hipSuccess) return e;
        sa.gd = gr->gd;
        sa.gd.raw = gr->g->iraw;
    } else if (records) {
        sa.gd = gr->gd;
        sa.gd.raw = nullptr;
    }
    sa.sv = sv;
    sa.tab = (const SceneTable *)rs->table;
    sa.sample_dev = rs->sample_dev;
    sa.rays_in = rs->rays[cur];
    sa.rays_out = rs->rays[nxt];
    sa.path_in = rs->path[cur];
    sa.path_out = rs->path[nxt];
    sa.root_in = rs->root[cur];
    sa.root_out = rs->root[nxt];
    sa.sroot = rs->sroot[set];
    sa.hits = rs->hits;
    sa.thr = rs->thr;
    sa.L = rs->L;
    sa.Lsh = rs->Lsh;
    sa.counts = rs->counts;
    sa.cur = cur;
    sa.nxt = nxt;
    sa.scount = sc;
    sa.srays = rs->srays[set];
    sa.scontrib = rs->scontrib[set];
    sa.acc = rs->acc;
    sa.acc64 = fp64 ? rs->acc64 : nullptr;
    sa.stats = rs->stats;
    sa.bounce = b;
    sa.max_depth = p->max_depth;
    sa.rr_depth = p->rr_depth;
    sa.npix = npix;
    sa.k = cap / npix;
    sa.seed = p->seed;
    sa.inv_spp = 1.0f / (float)p->spp;
    sa.pi = p->pi;
    sa.two_pi = p->two_pi;
    if (p->mode == TT_RENDER_REFERENCE) {
        hipLaunchKernelGGL(k_shade_reference, dim3(sblocks), dim3(kAppendBlock), 0, hs, sa);
        return hipGetLastError();
    }
    // the next-event work that last read this bounce's set (two bounces back)
    if (as && b >= 2) e = hipStreamWaitEvent(hs, s->ctx->shadow_done[set], 0);
    if (e != hipSuccess) return e;
    if (guide_here || records) hipLaunchKernelGGL(k_shade_path<true>, dim3(sblocks), dim3(kAppendBlock), 0, hs, sa);
    else hipLaunchKernelGGL(k_shade_path<false>, dim3(sblocks), dim3(kAppendBlock), 0, hs, sa);
    e = hipGetLastError();
    if (e == hipSuccess && as) e = hipEventRecord(s->ctx->shade_done, hs);
    if (e == hipSuccess && as) e = hipStreamWaitEvent(as, s->ctx->shade_done, 0);
    if (e != hipSuccess) return e;
    const hipStream_t ss = as ? as : hs;
    TraceArgs sh{};
    sh.nodes = (const float4 *)s->nodes;
    sh.tris = (const float4 *)s->tris;
    sh.rays = rs->srays[set];
    sh.flags = rs->socc[set];
    sh.n = cap;
    sh.num_nodes = (int32_t)s->num_nodes;
    sh.n_dev = rs->counts + sc;
    sh.ray_root = rs->sroot[set];
    e = launch_trace(sh, variant, true, lds, block, 0, refill_later, s->ctx->num_cus, ss, kTriTwoSided);
    if (e != hipSuccess) return e;
    if (sv.num_geoms > 0) {
        GeomArgs go{rs->srays[set], nullptr, rs->socc[set], rs->counts + sc, sv.geoms, sv.num_geoms};
        hipLaunchKernelGGL(k_geom_occluded, dim3(gblocks), dim3(256), 0, ss, go);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    ApplyArgs aa{rs->counts + sc, rs->socc[set], rs->scontrib[set], rs->Lsh, rs->stats, b, p->max_depth,
                 gr ? gr->gd : GuideShade{}};
    hipLaunchKernelGGL(k_apply_shadow, dim3(gblocks), dim3(256), 0, ss, aa);
    e = hipGetLastError();
    if (e == hipSuccess && as) e = hipEventRecord(s->ctx->shadow_done[set], as);
    return e;
}

// One sample pass (kk samples per pixel in flight) for all scenes of the table.
// overlap: path mode runs each bounce's next-event work on the context's second stream
// (off under graph capture, which records hs alone); the pass ends with hs waiting
// for the last of it.
hipError_t launch_pass(tt_scene *s, RenderState *rs, const tt_render_params *p, const SceneView &sv, int32_t npix,
                       int32_t kk, bool fp64, hipStream_t hs, bool overlap, const GuideRun *gr = nullptr) {
    const hipStream_t as = overlap && p->mode == TT_RENDER_PATH ? s->ctx->aux : nullptr;
    const int32_t cap = npix * kk;
    GenArgs g{};
    g.tab = (const SceneTable *)rs->table;
    g.npix = npix;
    g.mode = p->mode;
    g.k = kk;
    g.sample_dev = rs->sample_dev;
    g.seed = p->seed;
    g.tmin = p->mode == TT_RENDER_REFERENCE ? p->tmin : 0.0f;
    g.rays = rs->rays[0];
    g.path = rs->path[0];
    g.root = rs->root[0];
    g.thr = rs->thr;
    g.L = rs->L;
    g.counts = rs->counts;
    hipLaunchKernelGGL(k_generate, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, hs, g);
    hipError_t e = hipGetLastError();
    for (int32_t b = 0; e == hipSuccess && b < p->max_depth; ++b)
        e = launch_bounce(s, rs, p, sv, cap, npix, b, fp64, hs, gr, as);
    if (as) {   // whatever was queued on as (in order: the last bounce's event covers it), also after an error
        const hipError_t j = hipStreamWaitEvent(hs, s->ctx->shadow_done[(p->max_depth - 1) & 1], 0);
        if (e == hipSuccess) e = j;
    }
    if (e == hipSuccess && p->mode == TT_RENDER_PATH) {
        hipLaunchKernelGGL(k_resolve, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, hs, g.tab,
                           (const float4 *)rs->L, (const float4 *)rs->Lsh, npix, kk, fp64 ? nullptr : rs->acc,
                           fp64 ? rs->acc64 : nullptr);
        e = hipGetLastError();
    } else if (e == hipSuccess) {
        hipLaunchKernelGGL(k_resolve_reference, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, hs, g.tab,
                           (const float4 *)rs->L, npix, kk, 1.0f / (float)p->spp, rs->acc);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_advance_sample, dim3(1), dim3(1), 0, hs, rs->sample_dev, kk);
        e = hipGetLastError();
    }
    return e;
}

SceneView scene_view(const tt_scene *s, const tt_render_params *p) {
    SceneView sv{};
    sv.tri_normals = (const float4 *)s->tri_normals;
    sv.tri_geo = (const float4 *)s->tri_geo;
    sv.tri_mat = s->tri_mat;
    sv.shade_tri = (const uint4 *)s->shade_tri;
    sv.mats = s->mats;
    sv.num_mats = s->num_mats;
    sv.env = (const float4 *)s->env_rgb;
    sv.env_marg = s->env_marg_cdf;
    sv.env_cond = s->env_cond_cdf;
    sv.env_gmarg = s->env_marg_cells;
    sv.env_cells = s->env_cells;
    sv.env_w = s->env_w;
    sv.env_h = s->env_h;
    sv.env_scale = s->env_scale;
    sv.env_inv_total = s->env_total > 0.0 ? (float)(1.0 / s->env_total) : 0.0f;
    if (p->mode == TT_RENDER_PATH && s->env_total <= 0.0) sv.env = nullptr;  // black map: nothing to sample
    sv.emit_tri = s->emit_tri;
    sv.emit_cdf = s->emit_cdf;
    sv.emit_verts = (const float4 *)s->emit_verts;
    sv.num_emit = s->num_emit;
    sv.emit_inv_area = s->emit_area > 0.0f ? 1.0f / s->emit_area : 0.0f;
    sv.geoms = s->geoms;
    sv.num_geoms = s->num_geoms;
    return sv;
}

tt_status validate_params(const tt_render_params *p) {
    if (p->spp <= 0 || p->max_depth <= 0 || p->max_depth > 1024 || p->sample_offset < 0)
        return fail(TT_ERR_INVALID, "render: spp / max_depth out of range");
    if (p->mode != TT_RENDER_REFERENCE && p->mode != TT_RENDER_PATH) return fail(TT_ERR_INVALID, "render: bad mode");
    if (p->guided) return fail(TT_ERR_UNSUPPORTED, "render: guided rendering needs a guide (tt_guide_render)");
    if (p->mode == TT_RENDER_REFERENCE && p->accumulate_fp64)
        return fail(TT_ERR_INVALID, "render: FP64 accumulation is a path-mode option");
    if (p->lds_stack != 0 && p->lds_stack != 8 && p->lds_stack != 16 && p->lds_stack != 24 && p->lds_stack != 32)
        return fail(TT_ERR_INVALID, "render: lds_stack must be 0, 8, 16, 24 or 32");
    if (p->block != 0 && (p->block < 64 || p->block > 256 || p->block % 64))
        return fail(TT_ERR_INVALID, "render: block must be 0, 64, 128, 192 or 256");
    if (p->variant < -1 || p->variant > 1) return fail(TT_ERR_INVALID, "render: variant must be -1, 0 or 1");
    if (p->paths_in_flight < 0) return fail(TT_ERR_INVALID, "render: paths_in_flight must be >= 0");
    if (p->tile < 0 || p->tile > 32) return fail(TT_ERR_INVALID, "render: tile must be 0..32");
    if (p->refill < 0 || p->refill > 64 || p->refill_camera < 0 || p->refill_camera > 64)
        return fail(TT_ERR_INVALID, "render: refill thresholds must be 0..64");
    return TT_OK;
}

// Sizes the guide's record / training buffers for ntrain paths of max_depth vertices.
tt_status guide_train_buffers(tt_guide *g, int32_t ntrain, int32_t max_depth, hipStream_t hs) {
    const int64_t ne = (int64_t)ntrain * max_depth;
    // row offsets and counts are 32-bit
    if (ne >= (int64_t)INT32_MAX) return fail(TT_ERR_INVALID, "guide: train_paths x max_depth must stay below 2^31");
    if (g->rec_cap < ne * kRecFloats || g->train_cap < ne) {
        if (g->rec) hipFree(g->rec);
        if (g->train_buf) hipFree(g->train_buf);
        g->rec = g->train_buf = nullptr;
        g->rec_cap = g->train_cap = 0;
        tt_status st = dev_alloc(g->ctx, (void **)&g->rec, (size_t)ne * kRecFloats * 4, "guide path records");
        if (st == TT_OK) st = dev_alloc(g->ctx, (void **)&g->train_buf, (size_t)ne * 16 * 4, "guide training rows");
        if (st != TT_OK) return st;
        g->rec_cap = ne * kRecFloats;
        g->train_cap = ne;
    }
    const int64_t nb = (ne + 255) / 256;
    if (g->train_off_cap < nb + 1) {
        if (g->train_off) hipFree(g->train_off);
        g->train_off = nullptr;
        g->train_off_cap = 0;
        tt_status st = dev_alloc(g->ctx, (void **)&g->train_off, (size_t)(nb + 1) * 4, "guide training row offsets");
        if (st != TT_OK) return st;
        g->train_off_cap = nb + 1;
    }
    hipError_t e = hipMemsetAsync(g->rec, 0, (size_t)ne * kRecFloats * 4, hs);
    return e == hipSuccess ? TT_OK : fail(TT_ERR_DEVICE, std::string("guide buffers: ") + hipGetErrorString(e));
}

// After a pass: the recorded vertices become training rows, and the network
// takes Adam steps over them in batches of batch_rows (a remainder under a
// quarter batch is dropped, except when it is the only batch).
hipError_t guide_train_pass(GuideRun *gr, int32_t max_depth, hipStream_t hs) {
    tt_guide *g = gr->g;
    const int64_t ne = (int64_t)gr->gd.ntrain * max_depth;
    const int64_t nb = (ne + 255) / 256;
    hipLaunchKernelGGL(k_collect_count, dim3((unsigned)nb), dim3(256), 0, hs, (const float *)g->rec, ne, g->train_off);
    hipLaunchKernelGGL(k_collect_scan, dim3(1), dim3(1024), 0, hs, g->train_off, nb);
    hipLaunchKernelGGL(k_collect_write, dim3((unsigned)nb), dim3(256), 0, hs, g->rec, ne,
                       (const int32_t *)g->train_off, g->train_buf, (int32_t)g->train_cap);
    int32_t cnt = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&cnt, g->train_off + nb, 4, hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    cnt = (int32_t)std::min<int64_t>(cnt, g->train_cap);
    for (int32_t off = 0; e == hipSuccess && off < cnt; off += gr->batch_rows) {
        const int32_t n = std::min(gr->batch_rows, cnt - off);
        if (off > 0 && n < gr->batch_rows / 4) break;
        hipLaunchKernelGGL(k_stage_train, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, hs,
                           (const float *)(g->train_buf + (int64_t)off * 16), n, g->batch.in, g->batch.dirs,
                           g->batch.wts, g->batch.pb);
        e = hipGetLastError();
        g->train_alpha = gr->gd.learn_alpha != 0;
        if (e == hipSuccess) e = guide_train_rows(g, n, hs);
        g->train_alpha = false;
        gr->steps++;
        gr->rows += n;
        if (e == hipSuccess && gr->steps == 1) {   // the first step's loss (one wait per render)
            e = hipMemcpyAsync(&gr->first_loss, g->scalars, 4, hipMemcpyDeviceToHost, hs);
            if (e == hipSuccess) e = hipStreamSynchronize(hs);
        }
    }
    return e;
}

// Renders nsc scenes sharing one node / triangle array (roots[k] = root node of
// scene k) in one wavefront. out: host, concatenated images in scene order.
// Where a render's image goes: `count` floats (and FP64 values, if rgb64 is set) of
// the concatenated image starting at `offset`.
struct OutSeg {
    float *rgb;
    double *rgb64;
    int64_t offset, count;
};

tt_status render_impl(tt_scene *s, int32_t nsc, const tt_camera *cams, const int32_t *roots,
                      const tt_render_params *p, const OutSeg *segs, int32_t nseg, tt_render_stats *stats,
                      GuideRun *gr = nullptr) {
    tt_status st = validate_params(p);
    if (st != TT_OK) return st;
    if (gr && p->mode != TT_RENDER_PATH) return fail(TT_ERR_INVALID, "render: guiding needs path mode");
    if (nsc <= 0 || nsc > kMaxScenes) return fail(TT_ERR_INVALID, "render: 1..64 scenes");
    if (!s->tri_normals || !s->tri_mat || !s->mats || !s->tri_geo)
        return fail(TT_ERR_INVALID, "render: scene has no shading data");
    SceneTable tab{};
    tab.n = nsc;
    tab.tile = p->tile > 0 ? p->tile : kSlotTile;

    int64_t npix64 = 0;
    for (int32_t k = 0; k < nsc; ++k) {
        const tt_camera &c = cams[k];
        if (c.width < 2 || c.height < 2 || (int64_t)c.width * c.height > (1 << 26))
            return fail(TT_ERR_INVALID, "render: resolution out of range");
        tab.pix_off[k] = (int32_t)npix64;
        tab.w[k] = c.width;
        tab.h[k] = c.height;
        tab.root[k] = roots[k];
        tab.cam[k] = camera_args(c, p->pi, p->two_pi);
        npix64 += (int64_t)c.width * c.height;
    }
    if (npix64 > (1 << 26)) return fail(TT_ERR_INVALID, "render: total resolution out of range");
    const int32_t npix = (int32_t)npix64;
    tab.pix_off[nsc] = npix;
    const bool fp64 = p->accumulate_fp64 != 0;
    tt_ctx *ctx = s->ctx;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(TT_ERR_DEVICE, "hipSetDevice failed");
    // k samples per pixel are in flight in one pass (each path's radiance is kept in
    // its own slot and the pass ends by adding a pixel's samples in sample order, in
    // both modes). Each bounce's launches last as long as their slowest waves, so the
    // fewer the passes the less that tail costs: by default up to 2^27 paths (256 B of
    // state each, 34.4 GB) are in flight, guided renders with a frozen network included
    // (about 232 B more per path for the network's inputs, mixture and slot list). Renders that train the network
    // default to about 8M paths: it learns between passes.
    int32_t k = 1;
    {
        const int32_t target = p->paths_in_flight > 0 ? p->paths_in_flight : (gr && gr->train) ? (8 << 20) : (1 << 27);
        k = std::max(1, target / npix);
        k = std::min(k, p->spp);
        while (k > 1 && (int64_t)k * npix > (1LL << 27)) --k;
    }
    // a render that stops training part-way finishes with passes as wide as an
    // untrained one (paths_in_flight then sizes the training passes only)
    int32_t k_after = k;
    if (gr && gr->train && ((gr->train_spp > 0 && gr->train_spp < p->spp) || gr->growth > 1)) {
        k_after = std::max(1, std::min(p->spp, (1 << 27) / npix));
        while (k_after > 1 && (int64_t)k_after * npix > (1LL << 27)) --k_after;
        k_after = std::max(k_after, k);
    }
    st = ensure_state(s, npix * k_after, npix, fp64);
    if (st != TT_OK) return st;
    RenderState *rs = s->rs;
    hipStream_t hs = ctx->stream;
    if (gr) {
        // inference rows: one per queued path; training rows: batch_rows
        st = guide_alloc_infer(gr->g, (int64_t)npix * k_after);
        if (st == TT_OK && gr->train) st = guide_alloc_batch(gr->g, gr->batch_rows);
        if (st == TT_OK && gr->train) {
            const int32_t want = std::max(1, gr->gd.ntrain);
            gr->gd.stride = std::max(1, (npix * k) / want);
            gr->gd.ntrain = (npix * k) / gr->gd.stride + 1;
            st = guide_train_buffers(gr->g, gr->gd.ntrain, p->max_depth, hs);
        }
        if (st != TT_OK) return st;
        for (int q = 0; q < 3; ++q) {
            const float ext = s->bmax[q] - s->bmin[q];
            gr->gd.bmin[q] = s->bmin[q];
            gr->gd.binv[q] = ext > 0.0f ? 1.0f / ext : 1.0f;
        }
    }
    const SceneView sv = scene_view(s, p);
    const int64_t n3 = (int64_t)npix * 3;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipError_t e = hipMemcpyAsync(rs->table, &tab, sizeof(tab), hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, hs);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_zero, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, hs, rs->acc,
                           fp64 ? rs->acc64 : nullptr, rs->stats, n3);
        hipLaunchKernelGGL(k_set_sample, dim3(1), dim3(1), 0, hs, rs->sample_dev, p->sample_offset);
        e = hipGetLastError();
    }
    int32_t iters = 0;
    const int32_t full = p->spp / k, rest = p->spp % k;
    if (gr) {
        // guided: host-driven passes, training between them (not graph-captured)
        int32_t done = 0, pass = 0, k_train = k;
        while (e == hipSuccess && done < p->spp) {
            const bool train = gr->train && (gr->train_spp <= 0 || done < gr->train_spp);
            const int32_t kk = std::min(train ? k_train : k_after, p->spp - done);
            if (train && gr->growth > 1) k_train = (int32_t)std::min<int64_t>((int64_t)k_train * gr->growth, k_after);
            gr->gd.rec = train ? gr->g->rec : nullptr;
            // the recorded paths (every stride-th id, at most ntrain) spread over this
            // pass's whole width
            if (train) gr->gd.stride = std::max<int32_t>(1, (int32_t)(((int64_t)npix * kk) / std::max(1, gr->gd.ntrain)));
            gr->gd.toff = gr->gd.stride > 1 ? (int32_t)(((int64_t)pass * 7919) % gr->gd.stride) : 0;
            e = launch_pass(s, rs, p, sv, npix, kk, fp64, hs, true, gr);
            if (e == hipSuccess && train) {
                hipEvent_t t0 = nullptr, t1 = nullptr;
                e = hipEventCreate(&t0);
                if (e == hipSuccess) e = hipEventCreate(&t1);
                if (e == hipSuccess) e = hipEventRecord(t0, hs);
                if (e == hipSuccess) e = guide_train_pass(gr, p->max_depth, hs);
                if (e == hipSuccess) e = hipEventRecord(t1, hs);
                if (e == hipSuccess) e = hipEventSynchronize(t1);
                float tms = 0.0f;
                if (e == hipSuccess && hipEventElapsedTime(&tms, t0, t1) == hipSuccess) gr->train_ms += tms;
                if (t0) hipEventDestroy(t0);
                if (t1) hipEventDestroy(t1);
            }
            done += kk;
            ++pass;
            iters += p->max_depth;
        }
    } else if (e == hipSuccess && p->use_graph && full > 1) {
        // capture one pass (generate, max_depth x (extend, shade, shadow, apply),
        // advance the sample index) and replay it
        e = hipStreamBeginCapture(hs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t ce = launch_pass(s, rs, p, sv, npix, k, fp64, hs, false);
            e = hipStreamEndCapture(hs, &graph);
            if (e == hipSuccess) e = ce;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        for (int32_t i = 0; e == hipSuccess && i < full; ++i) e = hipGraphLaunch(gexec, hs);
        iters += full * p->max_depth;
    } else {
        for (int32_t i = 0; e == hipSuccess && i < full; ++i) e = launch_pass(s, rs, p, sv, npix, k, fp64, hs, true);
        iters += full * p->max_depth;
    }
    if (!gr && e == hipSuccess && rest > 0) {
        e = launch_pass(s, rs, p, sv, npix, rest, fp64, hs, true);
        iters += p->max_depth;
    }
    if (e == hipSuccess) e = hipEventRecord(e1, hs);
    // reference mode's sums are already the image (acc + L / spp per sample)
    if (e == hipSuccess && p->mode != TT_RENDER_REFERENCE) {
        hipLaunchKernelGGL(k_finish, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, hs, rs->acc,
                           fp64 ? rs->acc64 : nullptr, n3, p->spp);
        e = hipGetLastError();
    }
    unsigned long long hstats[2] = {0, 0};
    for (int32_t k = 0; e == hipSuccess && k < nseg; ++k) {
        e = hipMemcpyAsync(segs[k].rgb, rs->acc + segs[k].offset, (size_t)segs[k].count * 4, hipMemcpyDeviceToHost, hs);
        if (e == hipSuccess && fp64 && segs[k].rgb64)
            e = hipMemcpyAsync(segs[k].rgb64, rs->acc64 + segs[k].offset, (size_t)segs[k].count * 8,
                               hipMemcpyDeviceToHost, hs);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hstats, rs->stats, 16, hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (gexec) hipGraphExecDestroy(gexec);
    if (graph) hipGraphDestroy(graph);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("render: ") + hipGetErrorString(e));
    if (stats) {
        stats->rays_extension = (int64_t)hstats[0];
        stats->rays_shadow = (int64_t)hstats[1];
        stats->samples = (int64_t)npix * p->spp;
        stats->seconds = ms * 1e-3;
        stats->iterations = iters;
        stats->paths_in_flight = npix * k_after;
    }
    return TT_OK;
}

}  // namespace

void free_render_state(tt_scene *s) {
    RenderState *rs = s->rs;
    if (!rs) return;
    void *ptrs[] = {rs->rays[0], rs->rays[1], rs->path[0], rs->path[1], rs->root[0], rs->root[1], rs->hits, rs->thr,
                    rs->L, rs->Lsh, rs->srays[0], rs->srays[1], rs->sroot[0], rs->sroot[1], rs->scontrib[0],
                    rs->scontrib[1], rs->socc[0], rs->socc[1], rs->counts, rs->stats,
                    rs->sample_dev, rs->table, rs->acc, rs->acc64};
    for (void *q : ptrs)
        if (q) hipFree(q);
    delete rs;
    s->rs = nullptr;
}

}  // namespace tt

using namespace tt;

#define TT_HIPR(call, what)                                                                       \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(TT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct tt_batch {
    tt_scene *scene = nullptr;   // all scenes concatenated
    tt_mesh *mesh = nullptr;
    tt_bvh *bvh = nullptr;
    std::vector<int32_t> roots;
};

extern "C" {

tt_status tt_camera_from_orbit(const float center[3], float yaw, float pitch, float radius, float fovx_deg,
                               int32_t width, int32_t height, float focal, float aperture, tt_camera *out) {
    if (!center || !out || width <= 0 || height <= 0) return fail(TT_ERR_INVALID, "tt_camera_from_orbit: bad arguments");
    const float PI = 3.14156265f;  // Camera.cpp:5
    // Camera::setFovx (Camera.cpp:115-119) with radiansToDegrees / degreesToRadians (:107-113)
    float fovy_rad = std::atan(std::tan(fovx_deg / 180.0f * PI * 0.5f) * (float)height / (float)width) * 2.0f;
    float fovy = fovy_rad * 180.0f / PI;
    // Camera::getCameraMeta (Camera.cpp:121-140)
    float xd = std::sin(yaw) * std::cos(pitch);
    float yd = std::sin(pitch);
    float zd = std::cos(yaw) * std::cos(pitch);
    float m1 = (float)(-1.0);
    tt_camera c{};
    c.view[0] = xd * m1;
    c.view[1] = yd * m1;
    c.view[2] = zd * m1;
    c.eye[0] = center[0] + xd * radius;
    c.eye[1] = center[1] + yd * radius;
    c.eye[2] = center[2] + zd * radius;
    c.up[0] = 0.0f;
    c.up[1] = 1.0f;
    c.up[2] = 0.0f;
    c.fov[0] = fovx_deg;
    c.fov[1] = fovy;
    c.aperture = aperture;
    c.focal = focal;
    c.width = width;
    c.height = height;
    *out = c;
    return TT_OK;
}

tt_status tt_render(tt_scene *s, const tt_camera *cam, const tt_render_params *p, float *rgb_out, double *rgb64_out,
                    tt_render_stats *stats) {
    if (!s || !cam || !p || !rgb_out) return fail(TT_ERR_INVALID, "tt_render: null argument");
    const int32_t root = 0;
    const int64_t n3 = (int64_t)std::max(0, cam->width) * std::max(0, cam->height) * 3;
    const OutSeg seg{rgb_out, rgb64_out, 0, n3};
    return render_impl(s, 1, cam, &root, p, &seg, 1, stats);
}

tt_status tt_guide_render(tt_scene *s, tt_guide *g, const tt_camera *cam, const tt_render_params *p,
                          const tt_guide_params *gp, float *rgb_out, double *rgb64_out, tt_render_stats *stats,
                          tt_guide_stats *gstats) {
    if (!s || !g || !cam || !p || !gp || !rgb_out) return fail(TT_ERR_INVALID, "tt_guide_render: null argument");
    if (g->ctx != s->ctx) return fail(TT_ERR_INVALID, "tt_guide_render: guide and scene on different contexts");
    if (!(gp->alpha >= 0.0f && gp->alpha <= 1.0f)) return fail(TT_ERR_INVALID, "tt_guide_render: alpha outside [0, 1]");
    if (gp->train_paths < 0 || gp->train_batch < 0 || gp->train_spp < 0 || gp->train_batch > (1 << 22) ||
        gp->max_bounce < 0 || gp->pass_growth < 0 || gp->pass_growth > 16)
        return fail(TT_ERR_INVALID, "tt_guide_render: bad training sizes");
    tt_render_params q = *p;
    q.guided = 0;
    GuideRun gr;
    gr.g = g;
    gr.gd.alpha = gp->alpha;
    gr.gd.min_rough = gp->min_roughness;
    gr.train = gp->train != 0;
    gr.gd.ntrain = gp->train_paths > 0 ? gp->train_paths : (1 << 16);
    gr.batch_rows = gp->train_batch > 0 ? gp->train_batch : (1 << 18);
    gr.train_spp = gp->train_spp;
    gr.max_bounce = gp->max_bounce;
    gr.growth = std::max(1, gp->pass_growth);
    gr.gd.learn_alpha = gp->learn_alpha ? 1 : 0;
    const int32_t root = 0;
    const int64_t n3 = (int64_t)std::max(0, cam->width) * std::max(0, cam->height) * 3;
    const OutSeg seg{rgb_out, rgb64_out, 0, n3};
    tt_status st = render_impl(s, 1, cam, &root, &q, &seg, 1, stats, &gr);
    if (st == TT_OK && gstats) {
        gstats->train_steps = gr.steps;
        gstats->train_rows = gr.rows;
        gstats->train_seconds = gr.train_ms * 1e-3;
        float loss = 0.0f;
        if (gr.steps > 0 && hipMemcpy(&loss, g->scalars, 4, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(TT_ERR_DEVICE, "tt_guide_render: reading the loss failed");
        gstats->last_loss = loss;
        gstats->first_loss = gr.first_loss;
    }
    return st;
}

tt_status tt_batch_create(tt_ctx *ctx, int32_t n, const tt_mesh *const *meshes, const tt_bvh *const *bvhs,
                          const tt_material *const *mats, const int32_t *nmats, const float *env_rgb, int32_t env_w,
                          int32_t env_h, float env_scale, tt_batch **out) {
    if (!ctx || !out || n <= 0 || n > kMaxScenes || !meshes || !bvhs || !mats || !nmats)
        return fail(TT_ERR_INVALID, "tt_batch_create: bad arguments (1..64 scenes)");
    try {
        auto b = std::make_unique<tt_batch>();
        b->mesh = new tt_mesh();
        b->bvh = new tt_bvh();
        Mesh &M = b->mesh->m;
        Flat &F = b->bvh->b.flat;
        std::vector<tt_material> allm;
        for (int32_t k = 0; k < n; ++k) {
            if (!meshes[k] || !bvhs[k] || nmats[k] <= 0 || !mats[k]) return fail(TT_ERR_INVALID, "tt_batch_create: null scene");
            const Mesh &m = meshes[k]->m;
            const Flat &f = bvhs[k]->b.flat;
            const int32_t vo = (int32_t)M.verts.size(), no = (int32_t)M.normals.size(), to = M.ntris();
            const int32_t mo = (int32_t)allm.size();
            const int32_t node_o = (int32_t)(F.dev_nodes.size() / 16), rec_o = (int32_t)(F.dev_tris.size() / 12);
            if (f.dev_nodes.empty()) return fail(TT_ERR_INVALID, "tt_batch_create: empty tree");
            const tt_status cs = check_tri_ids(f, m.ntris());
            if (cs != TT_OK) return cs;
            M.verts.insert(M.verts.end(), m.verts.begin(), m.verts.end());
            M.normals.insert(M.normals.end(), m.normals.begin(), m.normals.end());
            for (int32_t x : m.tri_v) M.tri_v.push_back(x + vo);
            for (int32_t x : m.tri_n) M.tri_n.push_back(x + no);
            for (int32_t x : m.tri_mat) {
                if (x < 0 || x >= nmats[k]) return fail(TT_ERR_INVALID, "tt_batch_create: material id out of range");
                M.tri_mat.push_back(x + mo);
            }
            allm.insert(allm.end(), mats[k], mats[k] + nmats[k]);
            b->roots.push_back(node_o);
            for (size_t q = 0; q < f.dev_nodes.size(); q += 16) {
                int32_t w[16];
                std::memcpy(w, &f.dev_nodes[q], 64);
                for (int c = 0; c < 2; ++c) w[12 + c] = w[12 + c] >= 0 ? w[12 + c] + node_o : ~(~w[12 + c] + rec_o);
                F.dev_nodes.insert(F.dev_nodes.end(), w, w + 16);
            }
            for (size_t q = 0; q < f.dev_tris.size(); q += 12) {
                float r[12];
                std::memcpy(r, &f.dev_tris[q], 48);
                int32_t id, mid;
                std::memcpy(&id, &r[3], 4);
                std::memcpy(&mid, &r[7], 4);
                if (id >= 0) id += to;
                mid += mo;
                std::memcpy(&r[3], &id, 4);
                std::memcpy(&r[7], &mid, 4);
                F.dev_tris.insert(F.dev_tris.end(), r, r + 12);
            }
        }
        tt_status st = tt_scene_create(ctx, b->mesh, b->bvh, allm.data(), (int32_t)allm.size(), env_rgb, env_w, env_h,
                                       env_scale, &b->scene);
        if (st != TT_OK) {
            tt_batch_destroy(b.release());
            return st;
        }
        *out = b.release();
        return TT_OK;
    } catch (const std::bad_alloc &) {
        return fail(TT_ERR_NOMEM, "tt_batch_create: host allocation failed");
    }
}

void tt_batch_destroy(tt_batch *b) {
    if (!b) return;
    if (b->scene) tt_scene_destroy(b->scene);
    delete b->mesh;
    delete b->bvh;
    delete b;
}

tt_status tt_batch_render(tt_batch *b, const tt_camera *cams, const tt_render_params *p, float *const *rgb_out,
                          tt_render_stats *stats) {
    if (!b || !cams || !p || !rgb_out) return fail(TT_ERR_INVALID, "tt_batch_render: null argument");
    const int32_t n = (int32_t)b->roots.size();
    // each scene's part of the concatenated image goes straight to its buffer
    std::vector<OutSeg> segs((size_t)n);
    int64_t off = 0;
    for (int32_t k = 0; k < n; ++k) {
        if (!rgb_out[k]) return fail(TT_ERR_INVALID, "tt_batch_render: null image");
        const int64_t cnt = (int64_t)std::max(0, cams[k].width) * std::max(0, cams[k].height) * 3;
        segs[(size_t)k] = OutSeg{rgb_out[k], nullptr, off, cnt};
        off += cnt;
    }
    return render_impl(b->scene, n, cams, b->roots.data(), p, segs.data(), n, stats);
}

tt_status tt_debug_bsdf_sample(tt_ctx *ctx, const tt_material *m, const float wo[3], int32_t front, int64_t n,
                               uint32_t seed, float *out, float *out_eval) {
    if (!ctx || !m || !wo || !out || n <= 0 || n > (1LL << 26)) return fail(TT_ERR_INVALID, "tt_debug_bsdf_sample: bad arguments");
    TT_HIPR(hipSetDevice(ctx->device), "hipSetDevice");
    float *d_out = nullptr, *d_eval = nullptr;
    tt_status st = dev_alloc(ctx, (void **)&d_out, (size_t)n * 32, "debug bsdf");
    if (st == TT_OK && out_eval) st = dev_alloc(ctx, (void **)&d_eval, (size_t)n * 16, "debug bsdf eval");
    if (st != TT_OK) {
        if (d_out) hipFree(d_out);
        return st;
    }
    hipLaunchKernelGGL(k_debug_bsdf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, *m,
                       f3{wo[0], wo[1], wo[2]}, front != 0, n, seed, d_out, d_eval);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n * 32, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && out_eval) e = hipMemcpyAsync(out_eval, d_eval, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d_out);
    if (d_eval) hipFree(d_eval);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("tt_debug_bsdf_sample: ") + hipGetErrorString(e));
    return TT_OK;
}

tt_status tt_debug_env_sample(tt_scene *s, int64_t n, uint32_t seed, float *out) {
    if (!s || !out || n <= 0 || n > (1LL << 26)) return fail(TT_ERR_INVALID, "tt_debug_env_sample: bad arguments");
    if (!s->env_rgb || s->env_total <= 0.0) return fail(TT_ERR_INVALID, "tt_debug_env_sample: scene has no env map");
    tt_ctx *ctx = s->ctx;
    TT_HIPR(hipSetDevice(ctx->device), "hipSetDevice");
    SceneView sv{};
    sv.env = (const float4 *)s->env_rgb;
    sv.env_marg = s->env_marg_cdf;
    sv.env_cond = s->env_cond_cdf;
    sv.env_gmarg = s->env_marg_cells;
    sv.env_cells = s->env_cells;
    sv.env_w = s->env_w;
    sv.env_h = s->env_h;
    sv.env_scale = s->env_scale;
    sv.env_inv_total = (float)(1.0 / s->env_total);
    float *d_out = nullptr;
    tt_status st = dev_alloc(ctx, (void **)&d_out, (size_t)n * 32, "debug env");
    if (st != TT_OK) return st;
    hipLaunchKernelGGL(k_debug_env, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, sv, n, seed, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n * 32, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d_out);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("tt_debug_env_sample: ") + hipGetErrorString(e));
    return TT_OK;
}

tt_status tt_debug_env_invert(tt_scene *s, const float *u, int64_t n, int32_t *out) {
    if (!s || !u || !out || n <= 0 || n > (1LL << 26)) return fail(TT_ERR_INVALID, "tt_debug_env_invert: bad arguments");
    if (!s->env_rgb || s->env_total <= 0.0) return fail(TT_ERR_INVALID, "tt_debug_env_invert: scene has no env map");
    for (int64_t k = 0; k < 2 * n; ++k)
        if (!(u[k] >= 0.0f && u[k] < 1.0f)) return fail(TT_ERR_INVALID, "tt_debug_env_invert: u must be in [0, 1)");
    tt_ctx *ctx = s->ctx;
    TT_HIPR(hipSetDevice(ctx->device), "hipSetDevice");
    SceneView sv{};
    sv.env_marg = s->env_marg_cdf;
    sv.env_cond = s->env_cond_cdf;
    sv.env_gmarg = s->env_marg_cells;
    sv.env_cells = s->env_cells;
    sv.env_w = s->env_w;
    sv.env_h = s->env_h;
    void *d = nullptr;
    tt_status st = dev_alloc(ctx, &d, (size_t)n * 16, "debug env invert");
    if (st != TT_OK) return st;
    float *d_u = (float *)d;
    int32_t *d_out = (int32_t *)((char *)d + (size_t)n * 8);
    hipError_t e = hipMemcpyAsync(d_u, u, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_debug_env_invert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, sv, n,
                           (const float *)d_u, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("tt_debug_env_invert: ") + hipGetErrorString(e));
    return TT_OK;
}

tt_status tt_debug_guide_sample(tt_ctx *ctx, const float *raw, int64_t n, uint32_t seed, float *out) {
    if (!ctx || !raw || !out || n <= 0 || n > (1LL << 26)) return fail(TT_ERR_INVALID, "tt_debug_guide_sample: bad arguments");
    TT_HIPR(hipSetDevice(ctx->device), "hipSetDevice");
    float *d_buf = nullptr;
    tt_status st = dev_alloc(ctx, (void **)&d_buf, (size_t)n * 16 + 256, "debug guide");
    if (st != TT_OK) return st;
    float *d_raw = d_buf + 4 * n;
    hipError_t e = hipMemcpyAsync(d_raw, raw, kGuideNOut * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_debug_guide, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                           (const float *)d_raw, n, seed, d_buf);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_buf, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d_buf);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("tt_debug_guide_sample: ") + hipGetErrorString(e));
    return TT_OK;
}

tt_status tt_debug_math(tt_ctx *ctx, int32_t op, const float *x, const float *y, float *out, int64_t n) {
    if (!ctx || op < 0 || op > 4 || n < 0 || (n > 0 && (!x || !out || ((op == 2 || op == 3) && !y))))
        return fail(TT_ERR_INVALID, "tt_debug_math: bad arguments");
    if (n == 0) return TT_OK;
    TT_HIPR(hipSetDevice(ctx->device), "hipSetDevice");
    hipLaunchKernelGGL(k_debug_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, op, x, y, out, n);
    TT_HIPR(hipGetLastError(), "debug math launch");
    TT_HIPR(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TT_OK;
}

tt_status tt_debug_guide_infer_slots(tt_guide *g, const float *inputs, const int32_t *guided, int64_t n,
                                     float *raw_out, int64_t *rows) {
    if (!g || !inputs || !guided || !raw_out || !rows || n <= 0 || n > (1LL << 24))
        return fail(TT_ERR_INVALID, "tt_debug_guide_infer_slots: bad arguments");
    TT_HIPR(hipSetDevice(g->ctx->device), "hipSetDevice");
    tt_status st = guide_alloc_infer(g, n);
    if (st != TT_OK) return st;
    int32_t *d_given = nullptr;
    st = dev_alloc(g->ctx, (void **)&d_given, (size_t)n * 4 + 256, "debug guide flags");
    if (st != TT_OK) return st;
    int32_t *d_n = d_given + ((n + 63) & ~63LL);
    const int32_t n32 = (int32_t)n;
    hipStream_t hs = g->ctx->stream;
    hipError_t e = hipMemcpyAsync(g->iin, inputs, (size_t)n * kGuideIn * 4, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemcpyAsync(d_given, guided, (size_t)n * 4, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemcpyAsync(d_n, &n32, 4, hipMemcpyHostToDevice, hs);
    // every output word starts as 0xFFFFFFFF: the rows the inference skips keep it
    if (e == hipSuccess) e = hipMemsetAsync(g->iraw, 0xFF, (size_t)n * kGuideNOutPad16 * 4, hs);
    int32_t *rows_dev = g->iblock_n + g->iblocks;
    if (e == hipSuccess) {
        const unsigned rblocks = (unsigned)std::min<int64_t>((n + 255) / 256, g->iblocks);
        GuideInArgs ga{};
        ga.counts = d_n;
        ga.cur = 0;
        ga.in = g->iin;
        ga.mask = g->imask;
        ga.block_n = g->iblock_n;
        ga.given = d_given;
        hipLaunchKernelGGL(k_guide_inputs<true>, dim3(rblocks), dim3(256), 0, hs, ga);
        hipLaunchKernelGGL(k_guide_slots, dim3(rblocks), dim3(256), 0, hs, (const uint64_t *)g->imask,
                           (const int32_t *)g->iblock_n, (const int32_t *)d_n, g->islot, rows_dev);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = guide_infer(g, g->iin, g->iraw, n, rows_dev, hs, g->islot);
    std::vector<float> pad;
    int32_t r32 = 0;
    if (e == hipSuccess) {
        pad.resize((size_t)n * kGuideNOutPad16);
        e = hipMemcpyAsync(pad.data(), g->iraw, pad.size() * 4, hipMemcpyDeviceToHost, hs);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&r32, rows_dev, 4, hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    hipFree(d_given);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("tt_debug_guide_infer_slots: ") + hipGetErrorString(e));
    for (int64_t r = 0; r < n; ++r)
        std::memcpy(raw_out + r * kGuideNOut, pad.data() + r * kGuideNOutPad16, kGuideNOut * 4);
    *rows = r32;
    return TT_OK;
}

}  // extern "C"
