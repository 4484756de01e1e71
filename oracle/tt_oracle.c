/*
 * tt_oracle.c — CPU restatement of the reference traversal (test infrastructure:
 * the checker for the HIP kernels, never part of the product). See tt_oracle.h
 * for what pins each function.
 */
#include "tt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <string.h>
#include <time.h>

#define O_SENTINEL 0x76543210 /* IntersectKernel.cuh:9 */
#define O_STACK 256           /* reference: 64, unchecked (IntersectKernel.cuh:8); deeper here, checked */

/* ---------------------------------------------------------------- Philox */
/* Philox-4x32-10 (Salmon et al., SC'11), written from the paper's round
 * function: two 32x32->64 multiplies, xor with key, Weyl key schedule. */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ---------------------------------------------------------------- helpers */
static inline int32_t f2b(float f) { int32_t i; memcpy(&i, &f, 4); return i; }
static inline float b2f(int32_t i) { float f; memcpy(&f, &i, 4); return f; }
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

/* fminf / fmaxf of the slab code (IntersectKernel.cuh:55,59). Numerically the
 * smaller / larger operand; for a pair of zeros of opposite sign -0 counts as
 * smaller (IEEE 754-2019 minimum/maximum). The reference's choice for that pair
 * is NVIDIA hardware behaviour and is not pinned. */
static inline float o_fmin(float a, float b) {
    if (a != a || b != b) return b2f(0x7FC00000);
    if (a < b) return a;
    if (b < a) return b;
    return signbit(a) ? a : b;
}
static inline float o_fmax(float a, float b) {
    if (a != a || b != b) return b2f(0x7FC00000);
    if (a > b) return a;
    if (b > a) return b;
    return signbit(a) ? b : a;
}

/* PTX vmin/vmax.s32.s32.s32.{min,max} d, a, b, c: secondary op applied to the
 * primary result and c (IntersectKernel.cuh:22-36), on the float bit patterns. */
static inline float fmin_fmin(float a, float b, float c) { return b2f(imin(imin(f2b(a), f2b(b)), f2b(c))); }
static inline float fmin_fmax(float a, float b, float c) { return b2f(imax(imin(f2b(a), f2b(b)), f2b(c))); }
static inline float fmax_fmin(float a, float b, float c) { return b2f(imin(imax(f2b(a), f2b(b)), f2b(c))); }
static inline float fmax_fmax(float a, float b, float c) { return b2f(imax(imax(f2b(a), f2b(b)), f2b(c))); }

/* spanBeginKepler / spanEndKepler (IntersectKernel.cuh:54-60). */
static inline float span_begin(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
    return fmax_fmax(o_fmin(a0, a1), o_fmin(b0, b1), fmin_fmax(c0, c1, d));
}
static inline float span_end(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
    return fmin_fmin(o_fmax(a0, a1), o_fmax(b0, b1), fmax_fmin(c0, c1, d));
}

/* dot / cross with the operand order of LinearMath.h:260-266. */
static inline void vsub(const float a[3], const float b[3], float o[3]) {
    o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}
static inline void vcross(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline float vdot(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Triangle test, IntersectKernel.cuh:230-254 (including the det < -EPSILON
 * branch whose first condition is true for every finite u, :248-250). Returns 1
 * if the triangle is accepted with tmin < t < tmax. */
int o_tri_test(const float v0[3], const float v1[3], const float v2[3], const o_ray *r, float tmax,
               float *t_out, float *u_out, float *v_out) {
    const float EPSILON = 0.00001f;
    float o[3] = {r->ox, r->oy, r->oz}, d[3] = {r->dx, r->dy, r->dz};
    float e1[3], e2[3], tv[3], pv[3], qv[3];
    vsub(v1, v0, e1);
    vsub(v2, v0, e2);
    vsub(o, v0, tv);
    vcross(d, e2, pv);
    float det = vdot(e1, pv);
    float inv = 1.0f / det;
    float u = vdot(tv, pv) * inv;
    vcross(tv, e1, qv);
    float v = vdot(d, qv) * inv;
    if (det > EPSILON) {
        if (u < 0.0f || u > 1.0f) return 0;
        if (v < 0.0f || (u + v) > 1.0f) return 0;
    } else if (det < -EPSILON) {
        if (u > 0.0f || u < 1.0f) return 0;
        if (v > 0.0f || (u + v) < 1.0f) return 0;
    } else {
        return 0;
    }
    float t = vdot(e2, qv) * inv;
    if (r->tmin < t && t < tmax) {
        *t_out = t; *u_out = u; *v_out = v;
        return 1;
    }
    return 0;
}

/* ---------------------------------------------------------------- traversal */
int o_intersect(const o_bvh *b, const o_ray *r, int any, o_hit *out, float normal_out[3], o_counts *counts) {
    int32_t stack[O_STACK];
    int sp = 0;
    int32_t leaf_addr = 0, node_addr = 0;
    float hit_t = r->tmax, hit_u = 0.0f, hit_v = 0.0f;
    int32_t hit_index = -1;
    float nrm[3] = {0.0f, 0.0f, 0.0f};
    int64_t c_inner = 0, c_tris = 0, c_leaves = 0;
    int overflow = 0;

    /* ray precompute, IntersectKernel.cuh:92-101 (exp2f(-80) is exactly 2^-80) */
    const float ooeps = 0x1p-80f;
    float idirx = 1.0f / (fabsf(r->dx) > ooeps ? r->dx : copysignf(ooeps, r->dx));
    float idiry = 1.0f / (fabsf(r->dy) > ooeps ? r->dy : copysignf(ooeps, r->dy));
    float idirz = 1.0f / (fabsf(r->dz) > ooeps ? r->dz : copysignf(ooeps, r->dz));
    float oodx = r->ox * idirx, oody = r->oy * idiry, oodz = r->oz * idirz;

    stack[0] = O_SENTINEL;

    if (b->n_nodes_f4 == 0) node_addr = O_SENTINEL; /* empty scene */

    while (node_addr != O_SENTINEL) {
        while (node_addr >= 0 && node_addr != O_SENTINEL) {
            const int32_t *nd = b->nodes + (int64_t)node_addr * 4;
            float n0x0 = b2f(nd[0]), n0x1 = b2f(nd[1]), n0y0 = b2f(nd[2]), n0y1 = b2f(nd[3]);
            float n1x0 = b2f(nd[4]), n1x1 = b2f(nd[5]), n1y0 = b2f(nd[6]), n1y1 = b2f(nd[7]);
            float nz0 = b2f(nd[8]), nz1 = b2f(nd[9]), nz2 = b2f(nd[10]), nz3 = b2f(nd[11]);
            int32_t ch0 = nd[12], ch1 = nd[13];
            c_inner++;

            float c0lox = n0x0 * idirx - oodx;
            float c0hix = n0x1 * idirx - oodx;
            float c0loy = n0y0 * idiry - oody;
            float c0hiy = n0y1 * idiry - oody;
            float c0loz = nz0 * idirz - oodz;
            float c0hiz = nz1 * idirz - oodz;
            float c1loz = nz2 * idirz - oodz;
            float c1hiz = nz3 * idirz - oodz;
            float c0min = span_begin(c0lox, c0hix, c0loy, c0hiy, c0loz, c0hiz, r->tmin);
            float c0max = span_end(c0lox, c0hix, c0loy, c0hiy, c0loz, c0hiz, hit_t);
            float c1lox = n1x0 * idirx - oodx;
            float c1hix = n1x1 * idirx - oodx;
            float c1loy = n1y0 * idiry - oody;
            float c1hiy = n1y1 * idiry - oody;
            float c1min = span_begin(c1lox, c1hix, c1loy, c1hiy, c1loz, c1hiz, r->tmin);
            float c1max = span_end(c1lox, c1hix, c1loy, c1hiy, c1loz, c1hiz, hit_t);

            const float ray_tmax = 1e20f;
            int tr0 = (c0min <= c0max) && (c0min >= r->tmin) && (c0min <= ray_tmax);
            int tr1 = (c1min <= c1max) && (c1min >= r->tmin) && (c1min <= ray_tmax);

            if (!tr0 && !tr1) {
                node_addr = sp >= 0 ? stack[sp] : O_SENTINEL;
                sp--;
            } else {
                int32_t far = ch1;
                node_addr = tr0 ? ch0 : ch1;
                if (tr0 && tr1) {
                    if (c1min < c0min) {
                        int32_t tmp = node_addr; node_addr = far; far = tmp;
                    }
                    sp++;
                    if (sp >= O_STACK) { overflow = 1; node_addr = O_SENTINEL; break; }
                    stack[sp] = far;
                }
            }
            /* first leaf: postpone (IntersectKernel.cuh:172-178) */
            if (node_addr < 0 && leaf_addr >= 0) {
                leaf_addr = node_addr;
                node_addr = sp >= 0 ? stack[sp] : O_SENTINEL;
                sp--;
            }
            /* ballot over a single lane: break as soon as this lane holds a leaf */
            if (!(leaf_addr >= 0)) break;
        }

        while (leaf_addr < 0) {
            c_leaves++;
            for (int32_t tri_addr = ~leaf_addr;; tri_addr += 3) {
                const float *p0 = b->vertices + (int64_t)tri_addr * 4;
                if (f2b(p0[0]) == (int32_t)0x80000000) break;
                const float *p1 = p0 + 4, *p2 = p0 + 8;
                c_tris++;
                float t, u, v;
                float v0[3] = {p0[0], p0[1], p0[2]}, v1[3] = {p1[0], p1[1], p1[2]}, v2[3] = {p2[0], p2[1], p2[2]};
                if (o_tri_test(v0, v1, v2, r, hit_t, &t, &u, &v)) {
                    hit_t = t;
                    hit_u = u;
                    hit_v = v;
                    hit_index = tri_addr;
                    if (b->normals) {
                        /* n0 * (1 - u - v) + n1 * u + n2 * v (IntersectKernel.cuh:264) */
                        const float *q0 = b->normals + (int64_t)tri_addr * 4, *q1 = q0 + 4, *q2 = q0 + 8;
                        float w = 1 - u - v;
                        for (int k = 0; k < 3; ++k) nrm[k] = q0[k] * w + q1[k] * u + q2[k] * v;
                    }
                    if (any) {
                        node_addr = O_SENTINEL;
                        break;
                    }
                }
            }
            leaf_addr = node_addr;
            if (node_addr < 0) {
                node_addr = sp >= 0 ? stack[sp] : O_SENTINEL;
                sp--;
            }
        }
    }

    out->tri = -1;
    out->t = hit_t;
    out->u = hit_u;
    out->v = hit_v;
    if (hit_index != -1) out->tri = b->tri_idx[hit_index]; /* remap, IntersectKernel.cuh:285-290 */
    else { out->u = 0.0f; out->v = 0.0f; }
    if (normal_out) { normal_out[0] = nrm[0]; normal_out[1] = nrm[1]; normal_out[2] = nrm[2]; }
    if (counts) { counts->inner += c_inner; counts->tris += c_tris; counts->leaves += c_leaves; }
    return overflow ? -1 : 0;
}

void o_brute_closest(const float *verts, const int32_t *tris, int32_t ntris, const o_ray *r, o_hit *out) {
    float best = r->tmax, bu = 0.0f, bv = 0.0f;
    int32_t bi = -1;
    for (int32_t i = 0; i < ntris; ++i) {
        const float *a = verts + 3 * (int64_t)tris[3 * i], *b = verts + 3 * (int64_t)tris[3 * i + 1],
                    *c = verts + 3 * (int64_t)tris[3 * i + 2];
        float v0[3] = {a[0], a[1], a[2]}, v1[3] = {b[0], b[1], b[2]}, v2[3] = {c[0], c[1], c[2]};
        if (v0[0] == 0.0f) v0[0] = 0.0f; /* BVHCompact.cpp:143 */
        float t, u, v;
        if (o_tri_test(v0, v1, v2, r, best, &t, &u, &v)) {
            best = t; bu = u; bv = v; bi = i;
        }
    }
    out->tri = bi;
    out->t = best;
    out->u = bi >= 0 ? bu : 0.0f;
    out->v = bi >= 0 ? bv : 0.0f;
}

/* ---------------------------------------------------------------- batches */
typedef struct {
    const o_bvh *b;
    const o_ray *rays;
    o_hit *hits;
    int64_t lo, hi;
    int any;
    o_counts counts;
} o_job;

static void *o_worker(void *p) {
    o_job *j = (o_job *)p;
    memset(&j->counts, 0, sizeof(j->counts));
    for (int64_t i = j->lo; i < j->hi; ++i) o_intersect(j->b, &j->rays[i], j->any, &j->hits[i], NULL, &j->counts);
    return NULL;
}

double o_trace_batch(const o_bvh *b, const o_ray *rays, int64_t n, int any, o_hit *hits, int nthreads,
                     o_counts *counts_out) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    o_job jobs[256];
    pthread_t th[256];
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int k = 0; k < nthreads; ++k) {
        jobs[k].b = b; jobs[k].rays = rays; jobs[k].hits = hits; jobs[k].any = any;
        jobs[k].lo = n * k / nthreads;
        jobs[k].hi = n * (k + 1) / nthreads;
        pthread_create(&th[k], NULL, o_worker, &jobs[k]);
    }
    o_counts tot = {0, 0, 0};
    for (int k = 0; k < nthreads; ++k) {
        pthread_join(th[k], NULL);
        tot.inner += jobs[k].counts.inner;
        tot.tris += jobs[k].counts.tris;
        tot.leaves += jobs[k].counts.leaves;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (counts_out) *counts_out = tot;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

void o_brute_batch(const float *verts, const int32_t *tris, int32_t ntris, const o_ray *rays, int64_t n,
                   o_hit *hits) {
    for (int64_t i = 0; i < n; ++i) o_brute_closest(verts, tris, ntris, &rays[i], &hits[i]);
}

/* Per-ray counts variant (for the bytes-per-ray figure and tests). */
void o_trace_counts(const o_bvh *b, const o_ray *rays, int64_t n, int any, o_hit *hits, int64_t *counts3) {
    for (int64_t i = 0; i < n; ++i) {
        o_counts c = {0, 0, 0};
        o_intersect(b, &rays[i], any, &hits[i], NULL, &c);
        counts3[3 * i] = c.inner;
        counts3[3 * i + 1] = c.tris;
        counts3[3 * i + 2] = c.leaves;
    }
}
