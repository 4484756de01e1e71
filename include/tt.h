/*
 * tt.h — C ABI of the MI355X (gfx950) rendering hot path.
 *
 * Plain C types only: pointers, sizes, POD structs with asserted sizes. No torch
 * or HIP types cross this boundary (device buffers are passed as void* / typed
 * pointers to device memory, streams as void*).
 *
 * Every entry point returns a tt_status (0 = TT_OK). A failing call never
 * throws; the message of the last failure on the calling thread is returned by
 * tt_last_error(). Host objects (tt_mesh, tt_bvh) need no GPU. Device objects
 * (tt_ctx, tt_scene) own all device memory they allocate; the caller owns every
 * buffer it passes in. A tt_ctx is not thread-safe: use one per host thread.
 *
 * Reference seams replaced (paths relative to the reference repository):
 *   - BVH::BVH + BVH::createHolder          src/bvh/BVH.cpp:3-63,358      -> tt_bvh_build
 *   - BVHCompact::createCompact             src/bvh/BVHCompact.cpp:88-250 -> tt_bvh_export_compact
 *   - BVHCompact(FILE*) / BVHCompact::save  src/bvh/BVHCompact.cpp:23-79,319-375 -> tt_bvh_save / tt_bvh_load
 *   - Renderer::Renderer                    src/core/Renderer.cpp:8-57    -> tt_ctx_create + tt_scene_create
 *   - intersect(ray, bvhCompact, true)      include/geometry/IntersectKernel.cuh:78-296 -> tt_trace_closest
 *   - intersect(ray, bvhCompact, false)     include/geometry/IntersectKernel.cuh:265-269 -> tt_trace_any
 *   - Renderer::render / pathTracingKernel  src/core/RenderKernel.cu:97-233 -> tt_render
 *   - hdrSample                             include/hdr/HDRKernel.cuh:6-42 -> tt_env_* (lookup + importance sampling)
 */
#ifndef TT_H
#define TT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef enum {
    TT_OK = 0,
    TT_ERR_INVALID = 1,     /* bad argument (null pointer, size out of range) */
    TT_ERR_NOMEM = 2,       /* host or device allocation failed */
    TT_ERR_DEVICE = 3,      /* HIP runtime error */
    TT_ERR_IO = 4,          /* file could not be read / written / is corrupt */
    TT_ERR_UNSUPPORTED = 5  /* feature not available (e.g. no GPU) */
} tt_status;

/* Message of the last failing call on this thread ("" if none). */
const char *tt_last_error(void);
/* ABI version; bumped on any layout change. */
int tt_abi_version(void);

/* ------------------------------------------------------------------ records */
/* Ray, 32 B: {origin, tmin, direction, tmax}. Same fields as Ray
 * (include/geometry/Ray.h:10-13); hits need tmin < t < tmax (strict). */
typedef struct tt_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} tt_ray;

/* Hit, 16 B: original triangle id (-1 = miss), distance t and the barycentrics
 * (u, v) of the accepted triangle. The reference Hit (include/geometry/Hit.h:9-15)
 * stores index + distance; u and v are added here. */
typedef struct tt_hit {
    int32_t tri_id;
    float t, u, v;
} tt_hit;

/* Builder statistics (reference prints them at src/bvh/BVH.cpp:8-9,45-54 and
 * src/bvh/BVHCompact.cpp:40-42). */
typedef struct tt_bvh_stats {
    int32_t num_tris;          /* input triangles */
    int32_t num_refs;          /* triangle references in leaves (>= num_tris with spatial splits) */
    int32_t num_duplicates;    /* numDuplicates of the reference builder */
    int32_t num_inner;         /* inner nodes */
    int32_t num_leaves;        /* leaf nodes */
    int32_t max_depth;         /* deepest leaf level (root = 0) */
    float   sah;               /* top-down SAH cost (BVHNode::computeSubtreeProbabilities) */
    float   root_lo[3], root_hi[3];
    double  build_seconds;
} tt_bvh_stats;

/* Sizes of the reference's compact arrays (BVHCompact.h:42-48), in elements. */
typedef struct tt_compact_sizes {
    uint32_t nodes;        /* float4 entries (4 per inner node) */
    uint32_t vertices;     /* float4 entries (3 per triangle + 1 terminator per leaf) */
    uint32_t normals;      /* float4 entries (same as vertices) */
    uint32_t tri_indices;  /* int32 entries (same as vertices) */
    uint32_t mat_indices;  /* int32 entries (same as vertices) */
    uint32_t leaf_count;
    uint32_t tri_count;
} tt_compact_sizes;

/* ------------------------------------------------------------------ host mesh */
typedef struct tt_mesh tt_mesh;

/* Flat triangle mesh, as Scene (src/core/Scene.cpp:7-23): vertices, normals,
 * per-triangle vertex indices, normal indices and material index. normals /
 * tri_normals / tri_mats may be NULL (face normals / material 0 are used). */
tt_status tt_mesh_create(const float *verts, int32_t nverts,
                         const int32_t *tris, int32_t ntris,
                         const float *normals, int32_t nnormals,
                         const int32_t *tri_normals, const int32_t *tri_mats,
                         tt_mesh **out);
void tt_mesh_destroy(tt_mesh *m);
tt_status tt_mesh_info(const tt_mesh *m, int32_t *nverts, int32_t *ntris, int32_t *nnormals);
/* Copies out any of the arrays (pass NULL to skip one). */
tt_status tt_mesh_get(const tt_mesh *m, float *verts, int32_t *tris, float *normals,
                      int32_t *tri_normals, int32_t *tri_mats);
/* Appends src (vertex/normal indices rebased, material ids offset by mat_offset). */
tt_status tt_mesh_append(tt_mesh *dst, const tt_mesh *src, int32_t mat_offset);

/* Procedural generators (deterministic for a given seed) of the measured workloads;
 * no reference counterpart (the reference loads its scenes from files).
 * Soup: ntris equilateral triangles, centroid ~ U[0,1]^3, side ~ LogNormal(mu, sigma),
 *       uniformly random orientation; face normals. */
tt_status tt_mesh_gen_soup(int32_t ntris, uint64_t seed, float mu, float sigma, tt_mesh **out);
/* Icosphere subdivided `subdiv` times (20*4^subdiv tris), outward CCW winding,
 * radius*(1 + amp*fbm(p)) radial displacement (amp = 0 disables; `octaves` of
 * value noise), centred at c; smooth vertex normals. */
tt_status tt_mesh_gen_icosphere(int32_t subdiv, float radius, const float c[3],
                                float amp, int32_t octaves, uint64_t seed,
                                int32_t mat, tt_mesh **out);
/* Parallelogram p0,p1,p2 as two triangles (p0,p1,p2) and (p1,p1+p2-p0,p2)
 * (reference Plane, src/geometry/Plane.cpp; winding kept). */
tt_status tt_mesh_gen_quad(const float p0[3], const float p1[3], const float p2[3],
                           int32_t mat, tt_mesh **out);
/* Per-face material assignment: fractions f[0..k) of faces get mats[0..k). */
tt_status tt_mesh_assign_materials(tt_mesh *m, int32_t k, const float *fractions,
                                   const int32_t *mats, uint64_t seed);

/* Ray sets of the traversal workload (no reference counterpart: the reference
 * generates its camera rays inside pathTracingKernel, src/core/RenderKernel.cu:128-197).
 * Primary: width*height*spp pinhole rays (eye, look-at target, fovx degrees),
 *   jittered in-pixel, ordered in 8x8 pixel tiles (spp-major inside a tile).
 * Incoherent: n rays from uniform points on uniformly chosen triangles, cosine
 *   distributed about the front normal normalize(e1 x e2). */
tt_status tt_rays_primary(int32_t width, int32_t height, int32_t spp, const float eye[3],
                          const float target[3], float fovx_deg, float tmin, float tmax,
                          uint64_t seed, tt_ray *out);
tt_status tt_rays_incoherent(const tt_mesh *m, int64_t n, float tmin, float tmax,
                             uint64_t seed, tt_ray *out);

/* ------------------------------------------------------------------ host SBVH */
typedef struct tt_bvh tt_bvh;

/* SBVH build (Stich et al. 2009), decision-for-decision as BVH::BVH
 * (src/bvh/BVH.cpp:3-341): SAH node/tri cost 1/1, 32 spatial bins, max depth 64,
 * spatial splits above depth 48 with overlap >= split_alpha * root area.
 * nthreads <= 0 uses all cores; the result does not depend on nthreads. */
tt_status tt_bvh_build(const tt_mesh *m, float split_alpha, int32_t nthreads, tt_bvh **out);
void tt_bvh_destroy(tt_bvh *b);
tt_status tt_bvh_get_stats(const tt_bvh *b, tt_bvh_stats *out);

/* Reference compact layout (BVHCompact::createCompact, src/bvh/BVHCompact.cpp:88-250):
 * query sizes with all pointers NULL, then copy. nodes: 16 int32 per inner node
 * (child = float4 offset, leaf = ~vertex offset); vertices/normals: float4 per
 * vertex, leaf terminator {0x80000000,0,0,0}; tri/mat indices padded 3 per tri. */
tt_status tt_bvh_export_compact(const tt_bvh *b, tt_compact_sizes *sizes,
                                int32_t *nodes, float *vertices, float *normals,
                                int32_t *tri_indices, int32_t *mat_indices);

/* Device layout: nodes (64 B each: c0.xy bounds, c1.xy bounds, z bounds, {c0,c1,n0,n1})
 * with inner child = node index, leaf child = ~first triangle, n = leaf triangle
 * count; triangles 48 B each: {v0, id}, {v1-v0, mat}, {v2-v0, 0}. */
tt_status tt_bvh_export_device_layout(const tt_bvh *b, int64_t *num_nodes, int64_t *num_tris,
                                      int32_t *nodes, float *tris);

/* Cache file: the reference's bvh.cache format (7 x u32 header + the compact
 * arrays, src/bvh/BVHCompact.cpp:319-375) followed by a trailer holding the
 * device layout. tt_bvh_load restores a tt_bvh that can be passed to
 * tt_scene_create; it fails with TT_ERR_IO on a short or corrupt file. */
tt_status tt_bvh_save(const tt_bvh *b, const char *path);
tt_status tt_bvh_load(const char *path, tt_bvh **out);

/* ------------------------------------------------------------------ device */
typedef struct tt_ctx tt_ctx;
typedef struct tt_scene tt_scene;

/* Device, stream and per-frame device state of the renderer: replaces the buffers
 * Renderer::Renderer allocates for a frame (accumulation buffer, camera and render
 * metadata, src/core/Renderer.cpp:25-41); destroy scenes before their context. */
tt_status tt_ctx_create(int32_t device, tt_ctx **out);
void tt_ctx_destroy(tt_ctx *ctx);
/* Synchronises the context's stream. */
tt_status tt_ctx_sync(tt_ctx *ctx);

/* Device memory and timing on the context's device / stream (plumbing for callers
 * that hold no other HIP runtime; no reference counterpart). Copies are synchronous. Timer: `n` HIP event
 * slots recorded on the context stream; elapsed = ms between two recorded slots. */
tt_status tt_dev_alloc(tt_ctx *ctx, size_t bytes, void **out);
tt_status tt_dev_free(tt_ctx *ctx, void *p);
tt_status tt_memcpy_h2d(tt_ctx *ctx, void *dst, const void *src, size_t bytes);
tt_status tt_memcpy_d2h(tt_ctx *ctx, void *dst, const void *src, size_t bytes);
tt_status tt_memset_d(tt_ctx *ctx, void *dst, int value, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte loads/stores); its
 * HBM counters calibrate traffic measurements of the other kernels. bytes must
 * be a multiple of 16 and both pointers 16-byte aligned. Asynchronous. */
tt_status tt_dev_copy(tt_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Node-fetch roofline probe: `fetches` independent 64-byte reads (4 x 16 B) at
 * uniformly random 64-byte-aligned offsets of buf[0, bytes), 16 per lane; returns
 * the kernel time in ms. bytes >= 64 and buf 64-byte aligned. Synchronous. */
tt_status tt_bench_gather64(tt_ctx *ctx, const void *buf, size_t bytes, int64_t fetches, uint32_t seed, float *ms);
typedef struct tt_timer tt_timer;
tt_status tt_timer_create(tt_ctx *ctx, int32_t n, tt_timer **out);
tt_status tt_timer_record(tt_timer *t, int32_t slot);
tt_status tt_timer_elapsed(tt_timer *t, int32_t a, int32_t b, float *ms);
void tt_timer_destroy(tt_timer *t);
/* The context's stream (hipStream_t) as an opaque pointer. */
void *tt_ctx_stream(tt_ctx *ctx);

/* Materials. REF_* reproduce the reference samplers (reference render mode);
 * the others are the physically based set (path-tracer mode). */
typedef enum {
    TT_MAT_DIFFUSE = 0,     /* Lambert; ref: diffSample (include/material/DiffImpl.cuh:12-36) */
    TT_MAT_CONDUCTOR = 1,   /* GGX conductor (VNDF sampling); ref mode: metalSample Phong lobe */
    TT_MAT_DIELECTRIC = 2,  /* smooth dielectric, Fresnel; ref mode: refrSample (IOR 1.4) */
    TT_MAT_EMITTER = 3,     /* emission only (black BSDF) */
    TT_MAT_SPECULAR = 4,    /* perfect mirror; ref: specSample (include/material/SpecImpl.cuh:12-27) */
    TT_MAT_COAT = 5         /* Fresnel-chosen mirror or Lambert layer (n 1.000293 / 1.291);
                               ref: coatSample (include/material/CoatImpl.cuh:75-125) */
} tt_mat_type;

typedef struct tt_material {
    int32_t type;           /* tt_mat_type */
    float albedo[3];        /* diffuse colour / conductor F0 / dielectric tint / mirror colour */
    float emission[3];
    float roughness;        /* GGX alpha = roughness^2 (conductor) */
    float ior;              /* dielectric */
    float specular[3];      /* coat: colour of the mirror reflection (albedo = its diffuse colour) */
} tt_material;              /* 48 B */

/* Analytic primitives tested after the triangle traversal, as the reference's
 * per-ray loop over its geometry list (src/core/RenderKernel.cu:38-61):
 *   sphere (include/geometry/SphereImpl.cuh:10-32): centre a, radius b[0]; nearest
 *     root above 0.01, accepted below the ray's tmax;
 *   plane (include/geometry/PlaneImpl.cuh:10-52): the parallelogram p0 = a,
 *     p1 = b, p2 = c (u, v in [0, 1] along p1 - p0, p2 - p0), both sides, |det| >=
 *     1e-5, t > 1e-5 and below tmax; normal (p1 - p0) x (p2 - p0).
 * The nearest primitive replaces the triangle hit when strictly closer. Hits report
 * tri_id = -2 - (geometry index). */
typedef enum { TT_GEOM_SPHERE = 0, TT_GEOM_PLANE = 1 } tt_geom_type;

typedef struct tt_geometry {
    int32_t type;           /* tt_geom_type */
    int32_t material;       /* index into the scene's materials */
    float a[3], b[3], c[3];
    float pad;
} tt_geometry;              /* 48 B */

/* Uploads the SBVH device layout, per-triangle shading normals and materials
 * (replaces the uploads of Renderer::Renderer, src/core/Renderer.cpp:12-23).
 * The env map (RGB float, width x height, equirectangular, may be NULL = black)
 * gets its marginal/conditional CDF built on the GPU (FP64 prefix sums). */
tt_status tt_scene_create(tt_ctx *ctx, const tt_mesh *mesh, const tt_bvh *bvh,
                          const tt_material *mats, int32_t nmats,
                          const float *env_rgb, int32_t env_w, int32_t env_h,
                          float env_scale, tt_scene **out);
void tt_scene_destroy(tt_scene *s);
tt_status tt_scene_info(const tt_scene *s, int64_t *num_nodes, int64_t *num_tris, int64_t *device_bytes);
/* Replaces the scene's analytic primitives (n <= 4096; n = 0 removes them). tt_render
 * tests them in both modes (extension and shadow rays); tt_trace_closest /
 * tt_trace_any test triangles only, as intersect() does. */
tt_status tt_scene_set_geometries(tt_scene *s, const tt_geometry *g, int32_t n);
/* Copies the env-map sampling distribution built on the GPU: marginal CDF over
 * rows (env_h + 1 floats), conditional CDFs (env_h x (env_w + 1) floats) and the
 * FP64 total weight (sum of luminance * sin(theta) over texels). */
tt_status tt_env_export_cdf(const tt_scene *s, float *marg, float *cond, double *total);

/* Traversal launch tuning. variant: 0 = per-lane exact (single-lane order; hit ids
 * bit-identical to the host reference traversal), 1 = speculative while-while
 * (Aila-Laine postponed leaf + wave ballot). lds_stack: LDS stack entries per lane
 * (8, 16, 24 or 32); block: threads per block (64..256, multiple of 64).
 * refill: 0 (default) = one wave per 64-ray group; 1..64 = persistent waves: the
 * launch holds as many waves as the GPU keeps resident, wave w of W traces the
 * 64-ray groups w, w + W, ..., and a lane whose ray is done takes the wave's next ray
 * once at least `refill` of its 64 lanes are idle (keeps incoherent waves full; on
 * the 1M-soup traversal workload, half coherent rays, one wave per group is faster).
 * xcd_order: 0 = waves in launch order; 1 = blocks are renumbered so that each of the
 * 8 XCDs (which take blocks round-robin) holds one contiguous range of waves.
 * reorder: 0 (default) = rays are traced in the caller's order; 1 = the call first
 * stable-sorts the rays by a 24-bit Morton code of their origins (8 bits per axis over
 * the scene's mesh bounds), traces them in that order and writes each result at the
 * ray's own index. Results are identical; waves of rays that start near each other
 * share node fetches (on the 1M-soup traversal workload the 4.1M incoherent rays trace
 * 18% faster; camera rays, which share one origin, keep their order). Needs n < 2^31
 * and a context-owned workspace of 16 bytes per ray plus the sort's temporary storage,
 * allocated on first use and grown when a larger n arrives; a call with reorder = 1
 * waits (on its stream) for the previous such call of the context to finish with that
 * workspace, whatever stream it ran on. The sort issues no memset; still, a call is
 * not meant for HIP graph capture, since the workspace may grow with hipMalloc and
 * calls order themselves through a context event (tt_trace_closest_noalloc is the
 * entry for callers that cannot allocate). The Morton cells span the scene's mesh
 * bounds, or the root node's box when the scene was created without a mesh. */
typedef struct tt_trace_cfg {
    int32_t variant;
    int32_t lds_stack;
    int32_t block;
    int32_t xcd_order;
    int32_t refill;
    int32_t reorder;
} tt_trace_cfg;

/* Closest hit for n rays (replaces intersect(ray, bvh, true),
 * include/geometry/IntersectKernel.cuh:78-296). rays / hits are DEVICE pointers; stream is a hipStream_t
 * (NULL = the context stream). Asynchronous: synchronise before reading hits.
 * cfg may be NULL: exact loop, 16 LDS entries, 128-thread blocks, one wave per 64-ray
 * group (trenchanttracer_amd.autotune.tune_trace picks a configuration per workload). */
tt_status tt_trace_closest(tt_scene *s, const tt_ray *rays, int64_t n, tt_hit *hits,
                           const tt_trace_cfg *cfg, void *stream);
/* Any hit (shadow rays; intersect(ray, bvh, false), IntersectKernel.cuh:265-269):
 * hit_flags[i] = 1 if some triangle is hit in (tmin, tmax). */
tt_status tt_trace_any(tt_scene *s, const tt_ray *rays, int64_t n, int32_t *hit_flags,
                       const tt_trace_cfg *cfg, void *stream);
/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same results and arguments as tt_trace_closest, but with cfg->reorder = 1 the
 * Morton keys, permutation and sort storage live in a fixed 72-MiB device array of the
 * library (one copy per GPU, held until the process ends) instead of the context's
 * workspace, and the call allocates nothing. At most 4,194,304 rays per call (larger n
 * returns TT_ERR_INVALID). Every call on one GPU shares that array: calls that overlap
 * on one GPU (any context, any stream) race, so callers order them. */
tt_status tt_trace_closest_noalloc(tt_scene *s, const tt_ray *rays, int64_t n, tt_hit *hits,
                                   const tt_trace_cfg *cfg, void *stream);
/* Node-fetch roofline probe for a ray set (no reference counterpart): the time of
 * the traversal's own loop without its arithmetic. A recording launch of the
 * closest-hit loop (cfg->variant, LDS stack 16, cfg's block / refill / reorder) stores
 * the outcome of every node's slab test (4 bits: which children are entered, in which
 * order); the replay launch then runs the same loop over the same rays in the same
 * launch form (same waves, stack, leaf postponement, ballots and refills) with each
 * outcome read from that record, fetching every node (64 B) and every triangle
 * record (48 B) and folding their words into one 16-B result per ray, but with no
 * slab or triangle arithmetic. With reorder = 1 every replay is preceded by the
 * call's own key pass and sort, so the time bounds a whole tt_trace_closest call.
 * *ms = average over reps (HIP events, after one warm-up). node_fetches /
 * tri_fetches (may be NULL): the loop's node visits and triangle tests (bytes: 64
 * per node + 48 per triangle record + 48 per ray). counts_out (host, may be NULL):
 * n x 2 {node visits, triangle tests} per ray. Allocates and frees its record (half a
 * byte per node visit); synchronous. */
tt_status tt_bench_fetch_replay(tt_scene *s, const tt_ray *rays, int64_t n, const tt_trace_cfg *cfg, int32_t reps,
                                float *ms, int64_t *node_fetches, int64_t *tri_fetches, int32_t *counts_out);
/* Host-buffer convenience wrappers (copy in, trace, copy out, synchronise). */
tt_status tt_trace_closest_host(tt_scene *s, const tt_ray *rays, int64_t n, tt_hit *hits,
                                const tt_trace_cfg *cfg);
tt_status tt_trace_any_host(tt_scene *s, const tt_ray *rays, int64_t n, int32_t *hit_flags,
                            const tt_trace_cfg *cfg);

/* ------------------------------------------------------------------ render */
/* Pinhole / thin-lens camera with the fields of CameraMeta (include/core/Camera.h:8-16). */
typedef struct tt_camera {
    float eye[3], view[3], up[3];
    float fov[2];            /* degrees: fov[0] horizontal, fov[1] vertical */
    float aperture, focal;   /* aperture radius (lens sampled if > 1e-5), focal distance */
    int32_t width, height;
} tt_camera;

/* Orbit camera as Camera::Camera + setFovx + getCameraMeta (src/core/Camera.cpp:12-44,
 * 115-140): eye = center + (sin yaw cos pitch, sin pitch, cos yaw cos pitch) * radius,
 * view = -(that direction), up = +y, fov[1] from fov[0] and the aspect ratio with the
 * reference's PI = 3.14156265f in the degree conversions. */
tt_status tt_camera_from_orbit(const float center[3], float yaw, float pitch, float radius, float fovx_deg,
                               int32_t width, int32_t height, float focal, float aperture, tt_camera *out);

typedef enum {
    TT_RENDER_REFERENCE = 0,  /* the reference megakernel's semantics: emission on hit, BSDF sampling only,
                                 reference samplers (Lambert / Phong-30 metal / IOR dielectric), front-face
                                 triangle test; bit-identical to the host restatement for equal inputs */
    TT_RENDER_PATH = 1        /* path tracer: Lambert / GGX conductor (VNDF) / smooth dielectric, next-event
                                 estimation of emitters and the env map with MIS (power heuristic), Russian
                                 roulette, two-sided triangle test */
} tt_render_mode;

typedef struct tt_render_params {
    int32_t spp;              /* samples per pixel */
    int32_t max_depth;        /* path segments (the reference hard-codes 10 bounces, RenderKernel.cu:35) */
    int32_t rr_depth;         /* Russian roulette from this segment on (path mode) */
    uint32_t seed;            /* Philox key */
    int32_t mode;             /* tt_render_mode */
    int32_t guided;           /* path mode: 1 = learned guiding proposal (one-sample MIS with the BSDF) */
    int32_t sample_offset;    /* index of the first sample (progressive rendering) */
    int32_t accumulate_fp64;  /* accumulate pixel sums in FP64 (oracle renders) */
    float pi, two_pi;         /* camera / env constants (reference: 3.14156265f and 6.2831853f) */
    float tmin;               /* reference mode ray tmin (0.01, RenderKernel.cu:110) */
    int32_t use_graph;        /* capture one sample pass in a hipGraph and replay it (the captured
                                 pass runs on one stream; otherwise path mode runs each bounce's
                                 next-event work on the context's second stream) */
    int32_t paths_in_flight;  /* target paths per pass, rounded to whole samples (0 = up to
                                 2^27 paths, 256 B of state each, 488 B with a guide;
                                 guided renders that train the network: about 8M) */
    int32_t lds_stack;        /* traversal LDS stack entries per lane (0 = 16) */
    int32_t block;            /* traversal block size (0 = 64) */
    int32_t variant;          /* path-mode traversal loop: -1 speculative (default), 0 exact, 1 speculative */
    int32_t refill_camera;    /* traversal refill (see tt_trace_cfg) for the camera rays (0 = one wave per group) */
    int32_t refill;           /* ... for every later extension and shadow launch (0 = persistent, 16) */
    int32_t tile;             /* path slots walk tile x tile pixel blocks (0 = 8; 1 = pixel rows; 2..32);
                                 the image does not depend on it */
} tt_render_params;

typedef struct tt_render_stats {
    int64_t rays_extension;   /* closest-hit rays traced */
    int64_t rays_shadow;      /* any-hit rays traced */
    int64_t samples;          /* width * height * spp */
    double seconds;           /* GPU time of the render (HIP events) */
    int32_t iterations;       /* wavefront iterations launched */
    int32_t paths_in_flight;  /* paths per pass */
} tt_render_stats;

/* Renders one image (replaces Renderer::render and renderKernel,
 * src/core/RenderKernel.cu:97-233, src/core/Renderer.cpp:82-134). rgb_out (host, width*height*3 floats, row 0 = top) receives the
 * per-pixel mean over the spp samples; rgb64_out (host, may be NULL) the FP64 mean
 * when accumulate_fp64 is set. Synchronous. */
tt_status tt_render(tt_scene *s, const tt_camera *cam, const tt_render_params *p, float *rgb_out,
                    double *rgb64_out, tt_render_stats *stats);

/* Progressive accumulation and 8-bit display of one rendered frame (replaces the tail of
 * pathTracingKernel, src/core/RenderKernel.cu:202-217, and the buffer reset of
 * Renderer::display, src/core/Renderer.cpp:86-93). frame: device, npix x 3 floats (one
 * frame's per-pixel mean, as tt_render returns it); sum: device, npix x 3 floats, the
 * running sum, which frame_number 1 restarts from the reference's cudaMemset(buffer, 1)
 * pattern (every float 0x01010101); rgba: device, npix packed colours, r | g << 8 |
 * b << 16 | 1 << 24 with channel = (uint8)(pow(clamp(sum / frame_number, 0, 1), 1 / 2.2f)
 * * 255) (NaN clamps to 1). frame_number >= 1. Enqueued on stream (NULL: the context
 * stream); allocates nothing. */
tt_status tt_display_frame(tt_ctx *ctx, const float *frame, float *sum, int64_t npix, int32_t frame_number,
                           uint32_t *rgba, void *stream);

/* Batched rendering (no reference counterpart; the reference renders one scene):
 * up to 64 scenes in one wavefront (one node / triangle array,
 * per-ray root nodes, per-scene camera and resolution), sharing one env map.
 * tt_batch_render fills rgb_out[k] (width_k * height_k * 3 floats) for camera k;
 * the RNG stream of each scene equals that of rendering it alone. */
typedef struct tt_batch tt_batch;
tt_status tt_batch_create(tt_ctx *ctx, int32_t n, const tt_mesh *const *meshes, const tt_bvh *const *bvhs,
                          const tt_material *const *mats, const int32_t *nmats, const float *env_rgb,
                          int32_t env_w, int32_t env_h, float env_scale, tt_batch **out);
void tt_batch_destroy(tt_batch *b);
tt_status tt_batch_render(tt_batch *b, const tt_camera *cams, const tt_render_params *p, float *const *rgb_out,
                          tt_render_stats *stats);

/* Learned path-guiding proposal (no reference counterpart): MLP (frequency encoding of position / normal /
 * outgoing direction -> 3 x 64 -> 8-lobe vMF mixture), BF16 MFMA forward and
 * backward kernels with FP32 accumulation, FP32 master weights, Adam.
 * inputs: n x 9 floats in [0, 1]; raw outputs: n x 40 floats (lobe k: mean
 * raw[3k..3k+2], log kappa raw[24+k], logit raw[32+k]). Training rows: direction
 * (n x 3, unit) and weight (n, radiance / sampling pdf); the loss is
 * sum_r w_r (-log q(d_r)) / sum_r w_r. Parameters: tt_guide_param_count() floats,
 * per layer W[out][in] then b[out] (layers 224->64, 64->64, 64->64, 64->48). */
typedef struct tt_guide tt_guide;
tt_status tt_guide_create(tt_ctx *ctx, uint32_t seed, float lr, tt_guide **out);
void tt_guide_destroy(tt_guide *g);
int64_t tt_guide_param_count(void);
tt_status tt_guide_get_params(tt_guide *g, float *out);
tt_status tt_guide_set_params(tt_guide *g, const float *in);
tt_status tt_guide_forward(tt_guide *g, const float *inputs, int64_t n, float *raw_out);
tt_status tt_guide_grad(tt_guide *g, const float *inputs, const float *dirs, const float *wts, int64_t n,
                        float *loss, float *grad_out);
tt_status tt_guide_train_step(tt_guide *g, const float *inputs, const float *dirs, const float *wts, int64_t n,
                              float *loss);

/* ------------------------------------------------------------------ loaders */
/* Radiance RGBE (.hdr), as HDRLoader::load (src/loader/HDRLoader.cpp:5-153): header
 * "#?RADIANCE", lines up to an empty line, resolution "-Y h +X w"; flat, old-style
 * and adaptive run-length scanlines; pixel (r, g, b, e) -> (m / 256) * 2^(e - 128);
 * rows in file order (row 0 = top, the row the env lookup reads for +y).
 * tt_hdr_read always returns the size; pass rgb = NULL to query it, else rgb needs
 * capacity >= width * height * 3 floats. tt_hdr_write writes adaptive-RLE RGBE. */
tt_status tt_hdr_read(const char *path, int32_t *width, int32_t *height, float *rgb, int64_t capacity);
tt_status tt_hdr_write(const char *path, int32_t width, int32_t height, const float *rgb);

/* Wavefront OBJ, as OBJLoader::load (src/loader/OBJLoader.cpp:3-105): v / vn (vt,
 * vp ignored), faces v, v/t, v//n, v/t/n (1-based, fan-triangulated), groups (g),
 * usemtl per group, mtllib = a JSON material library beside the file
 * (MaterialLoader, src/loader/MaterialLoader.cpp:9-36: Diff / Metal / Spec / Refr /
 * Coat with optional emission; materials in name order). Vertices are centred on
 * their bounding box and scaled to a largest coordinate of 1 (Object::postProcess,
 * src/core/Object.cpp:40-87). Without vn, each triangle gets its face normal.
 * Without a library, material 0 is a grey diffuse. *num_mats receives the count;
 * mats (may be NULL) needs mats_capacity >= it. */
tt_status tt_obj_load(const char *path, tt_mesh **mesh, tt_material *mats, int32_t mats_capacity, int32_t *num_mats);

/* JSON scene configuration, as Config (src/util/Config.cpp:3-22), SceneLoader
 * (src/loader/SceneLoader.cpp:8-44) and Camera (src/core/Camera.cpp:12-44): Mesh
 * objects (OBJ files, then "transform" {scale, translate, orientation in degrees}
 * applied as S * T * (Ry * Rx * Rz), Transform.cpp:13-34) become one mesh; Plane and
 * Sphere objects become tt_geometry records with materials of the master library.
 * Materials: the master library ("material-file") in name order, then each mesh's
 * library in object order. Relative paths resolve against base_dir (may be NULL). */
typedef struct tt_config tt_config;
typedef struct tt_config_info {
    int32_t width, height, samples, bvh_reload, material_reload;
    int32_t num_materials, num_geometries, num_tris;
    float position[3];      /* orbit centre */
    float yaw, pitch, radius, focal, aperture, fovx;
    char work_dir[256], hdr_file[256], material_file[256];
} tt_config_info;
tt_status tt_config_load(const char *path, const char *base_dir, tt_config **out);
void tt_config_destroy(tt_config *c);
/* Any output may be NULL; mats / geoms need num_materials / num_geometries entries;
 * *mesh receives a new mesh (NULL when the scene has no triangles). */
tt_status tt_config_get(const tt_config *c, tt_config_info *info, tt_material *mats, tt_geometry *geoms,
                        tt_mesh **mesh);

/* Guided path tracing (path mode only; no counterpart in the reference, whose
 * renderer samples BSDFs only, RenderKernel.cu:40-60). On diffuse vertices and
 * conductors with roughness >= min_roughness, the direction is drawn from the
 * network's mixture with probability alpha and from the BSDF otherwise, weighted
 * by the combined density alpha q + (1 - alpha) p_bsdf (one-sample MIS, so the
 * image is unbiased for any network state); next-event MIS uses the same density.
 * With train set, train_paths paths of every pass record their vertices (inputs,
 * sampled direction, density, the radiance later found along the path) and the
 * network takes Adam steps on batches of train_batch rows between passes, until
 * train_spp samples per pixel are done (0: the whole render); the remaining samples
 * then run in passes as wide as a render without training (paths_in_flight, if
 * set, sizes the training passes only). The training rows are collected in record
 * order and every reduction of a training step runs in a fixed order: the image and
 * the trained network are the same on every run for the same inputs and seed. */
typedef struct tt_guide_params {
    float alpha;          /* probability of sampling the network on guided vertices */
    float min_roughness;  /* conductors at least this rough are guided */
    int32_t train;        /* 1: train online between passes */
    int32_t train_paths;  /* paths recorded per pass (0: 65536) */
    int32_t train_batch;  /* rows per Adam step (0: 262144, at most 4194304) */
    int32_t train_spp;    /* samples per pixel after which training stops (0: never) */
    int32_t max_bounce;   /* the network guides bounces 0 .. max_bounce - 1 (0: every bounce);
                             later vertices sample the BSDF alone */
    int32_t pass_growth;  /* 0 or 1: training passes of constant width; g >= 2: each training pass g times
                             wider than the one before, up to the width of an untrained render */
    int32_t learn_alpha;  /* 1: the selection probability is the network's own per-vertex output
                             (0.02 .. 0.95), trained on the cross-entropy of the combined density
                             alpha q + (1 - alpha) p_bsdf; alpha is then ignored */
} tt_guide_params;

typedef struct tt_guide_stats {
    int64_t train_steps;  /* Adam steps taken */
    int64_t train_rows;   /* rows trained on */
    double train_seconds; /* GPU time of the training between passes */
    float last_loss;      /* loss of the last step */
    float first_loss;     /* loss of the first step */
} tt_guide_stats;

tt_status tt_guide_render(tt_scene *s, tt_guide *g, const tt_camera *cam, const tt_render_params *p,
                          const tt_guide_params *gp, float *rgb_out, double *rgb64_out, tt_render_stats *stats,
                          tt_guide_stats *gstats);

/* Test hooks (no reference counterpart).
 * Sampling hook for the guiding mixture: decodes one row of 40 raw network outputs
 * (host), draws n directions with the device's mixture sampler and returns n x 4
 * floats {direction.xyz, mixture density at the direction}. Synchronous. */
tt_status tt_debug_guide_sample(tt_ctx *ctx, const float *raw, int64_t n, uint32_t seed, float *out);

/* Check hook for the fused inference kernel: the same n x 40 outputs as
 * tt_guide_forward, computed by the training path's separate encoding and layer
 * kernels. The two agree bit for bit. Synchronous. */
tt_status tt_debug_guide_forward_unfused(tt_guide *g, const float *inputs, int64_t n, float *raw_out);
/* Check hook for the renderer's guided-row inference: n queue slots with the given
 * inputs (n x 9) and guided flags (n, nonzero = guided) go through the render's own
 * flag and count, slot-list and inference kernels. raw_out (n x 40) gets each guided
 * slot's outputs, which equal tt_guide_forward's for its inputs bit for bit, and
 * 0xFFFFFFFF words in every other row (the rows the renderer does not infer stay
 * unwritten); *rows = the number of guided slots. Synchronous. */
tt_status tt_debug_guide_infer_slots(tt_guide *g, const float *inputs, const int32_t *guided, int64_t n,
                                     float *raw_out, int64_t *rows);
/* Check hook for the fused training kernels: the loss and gradient tt_guide_grad
 * returns, computed with the separate encoding, layer and backward-product kernels.
 * The two agree bit for bit. Synchronous. */
tt_status tt_debug_guide_grad_unfused(tt_guide *g, const float *inputs, const float *dirs, const float *wts, int64_t n,
                                      float *loss, float *grad_out);

/* Reproducibility hook for the render math: applies op (0 sin, 1 cos, 2 pow(x, y),
 * 3 atan2(x, y) with x the ordinate, 4 acos) of the device sampler and env-lookup math
 * to n device values; y is read for ops 2 and 3. Synchronous. */
tt_status tt_debug_math(tt_ctx *ctx, int32_t op, const float *x, const float *y, float *out, int64_t n);

/* Sampling hooks for the path-mode BSDFs and the env-map distribution (the same
 * device functions the shade stage calls), for distribution tests.
 * tt_debug_bsdf_sample: n samples for a fixed local wo (z = shading normal), front =
 *   ray on the side the geometric normal points to; out: n x 8 floats
 *   {wi.xyz, weight.xyz (f cos / pdf), pdf, delta}. out_eval (may be NULL): n x 4
 *   {f cos .xyz, pdf} of bsdf_eval(wo, wi) for the sampled wi.
 * tt_debug_env_sample: n directions from the scene's env distribution; out: n x 8
 *   {dir.xyz, pdf, Le.xyz, pdf of env_eval(dir)}.
 * Host buffers; synchronous. */
tt_status tt_debug_bsdf_sample(tt_ctx *ctx, const tt_material *m, const float wo[3], int32_t front, int64_t n,
                               uint32_t seed, float *out, float *out_eval);
tt_status tt_debug_env_sample(tt_scene *s, int64_t n, uint32_t seed, float *out);
/* The env texel (row, column) the sampler picks for given uniforms: u = n x 2 floats
 * in [0, 1) (marginal, conditional), out = n x 2 int32 {row, column}. The inversion
 * of the marginal and conditional CDFs (tt_env_export_cdf) through their guide
 * tables: the largest index k with cdf[k] < u. Host buffers; synchronous. */
tt_status tt_debug_env_invert(tt_scene *s, const float *u, int64_t n, int32_t *out);
/* The order a reordering trace call (tt_trace_cfg.reorder = 1) traces the rays in:
 * order[r] = index of the r-th ray, the stable sort of the ray indices by the 24-bit
 * Morton code of the origins over the scene's bounds (8 bits per axis, x in bits
 * 3k + 2, y in 3k + 1, z in 3k; cells clamped to 0..255). rays: device, n records;
 * order: host, n int32. cap > n sorts as a queue of capacity cap whose length n is
 * read on the device (the path tracer's form); cap <= n, with the length from the
 * host. Synchronous. */
tt_status tt_debug_ray_order(tt_scene *s, const tt_ray *rays, int64_t n, int64_t cap, int32_t *order);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* TT_H */
