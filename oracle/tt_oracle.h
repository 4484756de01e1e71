/*
 * tt_oracle.h — CPU restatement of the reference rendering hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * never by the product library (trenchanttracer_amd), which has no CPU
 * fallback. Built with gcc -O2 -ffp-contract=off (no -march), so each float
 * operation is one IEEE binary32 operation in source order.
 *
 * Pinning (what checks each function, see tests/):
 *   o_philox4x32_10      Random123 known-answer vectors (tests/test_oracle.py)
 *   o_intersect          == o_brute_closest on seeded scenes (argmin over all
 *                        triangles of the same triangle test; independent of
 *                        the tree) + closed-form triangle cases
 *   o_tri_test           closed forms: det = 2*area*cos, back faces rejected,
 *                        edge inclusive, t = tmin exclusive, |det| <= 1e-5 rejected
 *   o_brute_closest      brute force itself; checked on hand-made scenes with
 *                        known answers
 *   o_trace_batch        == o_intersect per ray (threads only split the work)
 *   o_sinf/o_cosf/...    float64 numpy within stated ulp bounds
 *   o_render_ref         restated megakernel; checked by white furnace and
 *                        closed-form cases (tests/test_render_oracle.py)
 *   o_ref_bounce         each reference sampler's law: direction cosines and
 *                        azimuth against the uniforms, Phong lobe exponent,
 *                        Snell / total internal reflection, the dielectric's
 *                        0.2 reflection share and Re/P, Tr/(1-P) weights with
 *                        the reference's R0, origin offsets along nl, chi-square
 *                        of drawn directions (tests/test_oracle_samplers.py);
 *                        every check is also run on builds with a planted
 *                        mistake (-DO_PLANT=k) and must fail there
 *   o_env_lookup         texel selection against closed-form longitude /
 *                        latitude of directions, the x2 scale, the linear-index
 *                        wrap and the end clamp (tests/test_oracle_env.py)
 *   o_atan2f/o_acosf     float64 numpy within 2 ulp, C99 special values
 *   o_display_frame      closed forms: 8-bit value = floor(255 * c^(1/2.2)) in
 *                        float64 away from integer boundaries, 0 and 255 at the
 *                        clamp ends, NaN -> 255, alpha 1, the byte-1 starting
 *                        sum, running sum / frame count (tests/test_oracle_display.py);
 *                        planted rounding, starting-sum and exponent mistakes fail
 */
#ifndef TT_ORACLE_H
#define TT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float ox, oy, oz, tmin, dx, dy, dz, tmax; } o_ray;
typedef struct { int32_t tri; float t, u, v; } o_hit;
typedef struct { int64_t inner, tris, leaves; } o_counts;

/* Reference compact BVH (BVHCompact.h:30-48): nodes as int32 (4 per float4),
 * vertices / normals as float4, index arrays int32. */
typedef struct {
    const int32_t *nodes;
    const float *vertices;
    const float *normals;
    const int32_t *tri_idx;
    const int32_t *mat_idx;
    int64_t n_nodes_f4;   /* float4 entries in nodes */
    int64_t n_vertices;   /* float4 entries in vertices */
} o_bvh;

void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Möller–Trumbore as IntersectKernel.cuh:230-254. Returns 1 if accepted with
 * tmin < t < tmax. */
int o_tri_test(const float v0[3], const float v1[3], const float v2[3], const o_ray *r, float tmax,
               float *t, float *u, float *v);

/* Single-lane restatement of intersect() (IntersectKernel.cuh:78-296). any != 0
 * stops at the first accepted triangle. normal_out (may be NULL) receives the
 * interpolated unnormalised normal. counts (may be NULL) accumulate node visits,
 * triangle tests and processed leaves. Returns 0, or -1 on stack overflow. */
int o_intersect(const o_bvh *b, const o_ray *r, int any, o_hit *out, float normal_out[3], o_counts *counts);

/* Argmin over all triangles (v0.x -0.0 rewritten to +0.0 as the compact layout
 * does), first index wins ties. */
void o_brute_closest(const float *verts, const int32_t *tris, int32_t ntris, const o_ray *r, o_hit *out);

void o_brute_batch(const float *verts, const int32_t *tris, int32_t ntris, const o_ray *rays, int64_t n,
                   o_hit *hits);
void o_trace_counts(const o_bvh *b, const o_ray *rays, int64_t n, int any, o_hit *hits, int64_t *counts3);

/* Batch traversal on nthreads host threads (static contiguous chunks).
 * counts_out (may be NULL) receives the totals. Returns wall seconds. */
double o_trace_batch(const o_bvh *b, const o_ray *rays, int64_t n, int any, o_hit *hits,
                     int nthreads, o_counts *counts_out);

/* Reference render restatement (tt_oracle_render.c). */
typedef struct { float eye[3], view[3], up[3]; float fov[2]; float aperture, focal; int32_t width, height; } o_camera;
typedef struct { int32_t type; float albedo[3]; float emission[3]; float roughness, ior; float specular[3]; } o_material;
/* analytic primitive: type 0 sphere (centre a, radius b[0]), 1 parallelogram (p0 a, p1 b, p2 c) */
typedef struct { int32_t type, material; float a[3], b[3], c[3]; float pad; } o_geometry;
typedef struct { int32_t spp, max_depth; uint32_t seed; int32_t sample_offset; float pi, two_pi, tmin, tmax; } o_ref_params;
/* equirectangular env map: width x height RGB floats, row 0 = +y (latitude 0) */
typedef struct { const float *rgb; int32_t width, height; float scale; } o_env;

/* Renders rows [y0, y1) of the image (row 0 = top) into out_rgb (W*H*3 floats,
 * the average over spp samples). tri_normals: 12 floats (3 float4) per original
 * triangle; tri_mat: material per original triangle; env (may be NULL = black)
 * is looked up on a miss as hdrSample does. */
int o_render_ref(const o_bvh *b, const float *tri_normals, const int32_t *tri_mat, const o_material *mats,
                 const o_geometry *geoms, int32_t num_geoms, const o_env *env, const o_camera *cam,
                 const o_ref_params *p, int32_t y0, int32_t y1, float *out_rgb, int nthreads);
/* One reference-mode surface interaction (RenderKernel.cu:72-92 and the material's
 * sampler): incoming direction d, hit point, unnormalised normal nrm, uniforms u.
 * out: next origin (3), next direction (3), mask multiplier (3, from 1). Returns 0
 * when the material ends the path. */
int o_ref_bounce(const o_material *m, const float u[4], const float d[3], const float point[3], const float nrm[3],
                 float out[9]);
/* hdrSample (HDRKernel.cuh:6-42) for direction d (out: RGB). */
void o_env_lookup(const o_env *e, const float d[3], float pi, float two_pi, float out[3]);
/* distance at which ray r hits primitive g under the reference's rules
 * (SphereImpl.cuh:10-32, PlaneImpl.cuh:10-52), +inf when it does not */
float o_geom_t(const o_geometry *g, const o_ray *r);

/* Float math used by the restated samplers (Cody–Waite reduction + Taylor
 * polynomials; |error| <= 2 ulp on the documented ranges). */
float o_sinf(float x);
float o_cosf(float x);
float o_logf(float x);
float o_expf(float x);
float o_powf(float x, float y);
float o_atan2f(float y, float x);
float o_acosf(float x);
/* op over n values: 0 sin x, 1 cos x, 2 pow(x, y), 3 atan2(x, y) (x the ordinate), 4 acos x */
void o_math_batch(int op, const float *x, const float *y, float *out, int64_t n);
/* Progressive display (RenderKernel.cu:202-217, Renderer.cpp:86-93): sum (npix x 3) +=
 * frame (npix x 3), starting from the reference's byte-1 pattern when frame_number == 1;
 * rgba[i] = the gamma-2.2 8-bit colour of sum / frame_number, alpha 1. */
void o_display_frame(const float *frame, float *sum, int64_t npix, int32_t frame_number, uint32_t *rgba);

#ifdef __cplusplus
}
#endif

#endif
