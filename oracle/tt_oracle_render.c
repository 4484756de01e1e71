/*
 * tt_oracle_render.c — CPU restatement of the reference path tracer
 * (src/core/RenderKernel.cu:27-218 with DiffImpl.cuh, MetalImpl.cuh, RefrImpl.cuh,
 * SpecImpl.cuh, CoatImpl.cuh, HDRKernel.cuh) for image parity of the reference
 * render mode. TEST INFRASTRUCTURE: the checker, never part of the product.
 *
 * Differences from the reference that the GPU reference mode shares:
 *   - random numbers: Philox-4x32-10 addressed by (pixel, sample, bounce), uniform
 *     in (0, 1] like curand_uniform; cuRAND's XORWOW stream is not reproduced;
 *   - sinf / cosf / powf / atan2f / acosf: the functions below (the reference calls
 *     CUDA's);
 *   - the env lookup clamps its texel index to the last texel: the reference reads
 *     past the end of its texture there (u2 + v2 * width >= width * height, reached
 *     when acosf(d.y) / PI exceeds 1 with PI = 3.14156265f < pi, or when the wrapped
 *     longitude rounds to TWO_PI in the last row);
 *   - a render returns the mean of its samples; the reference's running sum over
 *     frames, which starts at bytes of 1 (2.3694278e-38 per channel,
 *     Renderer.cpp:27,88), and its 8-bit output are o_display_frame below.
 *
 * Planted mistakes: building with -DO_PLANT=<k> (k > 0) compiles one deliberate
 * error into this file (see the PLANT() sites below). Only the tests build such a
 * library, to show that their checks of the sampler and env laws fail on it.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "tt_oracle.h"

#ifndef O_PLANT
#define O_PLANT 0
#endif
#define PLANT(k) (O_PLANT == (k))

/* ---------------------------------------------------------------- float math */
/* Cody–Waite reduction by pi/2 with a three-part constant (12 + 12 + 24 bits),
 * then Taylor polynomials of sin (to r^9) and cos (to r^10) on |r| <= pi/4. */
static float o_sincos(float x, float *c_out) {
    float k = rintf(x * 0x1.45f306p-1f);
    float r = x - k * 0x1.92p+0f;
    r = r - k * 0x1.fb4p-12f;
    r = r - k * 0x1.4442d2p-24f;
    float r2 = r * r;
    float ps = -0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f));
    float s = r + r * r2 * ps;
    float pc = 0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f));
    float c = 1.0f - 0.5f * r2 + r2 * r2 * pc;
    int q = ((int)k) & 3;
    float sn, cs;
    if (q == 0) { sn = s; cs = c; }
    else if (q == 1) { sn = c; cs = -s; }
    else if (q == 2) { sn = -s; cs = -c; }
    else { sn = -c; cs = s; }
    *c_out = cs;
    return sn;
}

float o_sinf(float x) { float c; return o_sincos(x, &c); }
float o_cosf(float x) { float c; o_sincos(x, &c); return c; }

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* log2 by x = m 2^e, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)) series to f^11 */
static float o_log2f(float x) {
    int e = 0;
    if (x < 0x1p-126f) { x = x * 0x1p24f; e = -24; }
    uint32_t b = fbits(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = bitsf((b & 0x7fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e = e + 1; }
    float f = (m - 1.0f) / (m + 1.0f);
    float f2 = f * f;
    float p = f2 * (0x1.555556p-2f + f2 * (0x1.99999ap-3f + f2 * (0x1.24924ap-3f + f2 * (0x1.c71c72p-4f + f2 * 0x1.745d18p-4f))));
    float ln_m = (f + f * p) * 2.0f;
    return (float)e + ln_m * 0x1.715476p+0f;
}

/* 2^z = 2^n e^(g ln 2), |g| <= 1/2, Taylor to t^8 */
static float o_exp2f(float z) {
    if (z < -150.0f) return 0.0f;
    if (z >= 128.0f) return INFINITY;
    float n = rintf(z);
    float g = z - n;
    float t = g * 0x1.62e430p-1f;
    float p = 1.0f + t * (1.0f + t * (0.5f + t * (0x1.555556p-3f + t * (0x1.555556p-5f + t * (0x1.111112p-7f +
              t * (0x1.6c16c2p-10f + t * (0x1.a01a02p-13f + t * 0x1.a01a02p-16f)))))));
    int ni = (int)n;
    if (ni < -126) { p = p * 0x1p-64f; ni += 64; }
    return p * bitsf((uint32_t)(ni + 127) << 23);
}

float o_logf(float x) { return o_log2f(x) * 0x1.62e430p-1f; }
float o_expf(float x) { return o_exp2f(x * 0x1.715476p+0f); }
/* x in [0, 1], y > 0 */
float o_powf(float x, float y) {
    if (x == 0.0f) return 0.0f;
    return o_exp2f(y * o_log2f(x));
}

/* atan2 by octant reduction: a = min(|x|, |y|) / max(|x|, |y|) in [0, 1]; above
 * tan(pi/8), atan a = pi/4 + atan((a - 1) / (a + 1)); atan t by its Taylor series
 * t - t^3/3 + ... - t^19/19 on |t| <= tan(pi/8); then pi/2 - r when |y| > |x|,
 * pi - r when x is negative, and the sign of y. pi/4, pi/2 and pi are
 * each a float plus a float correction, the correction added first. Zeros follow C99 atan2
 * (atan2(+-0, +0) = +-0, atan2(+-0, -0) = +-pi); NaN propagates. */
float o_atan2f(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y);
    int swap = ay > ax;
    float num = swap ? ax : ay, den = swap ? ay : ax;
    float r;
    if (den == 0.0f) {
        r = 0.0f;
    } else {
        float a = num / den;
        float hi = 0.0f, lo = 0.0f, t = a;
        if (a > 0x1.a8279ap-2f) {
            t = (a - 1.0f) / (a + 1.0f);
            hi = 0x1.921fb6p-1f;
            lo = -0x1.777a5cp-26f;
        }
        float t2 = t * t;
        float p = t2 * (-0x1.555556p-2f + t2 * (0x1.99999ap-3f + t2 * (-0x1.24924ap-3f + t2 * (0x1.c71c72p-4f +
                  t2 * (-0x1.745d18p-4f + t2 * (0x1.3b13b2p-4f + t2 * (-0x1.111112p-4f + t2 * (0x1.e1e1e2p-5f +
                  t2 * -0x1.af286cp-5f))))))));
        r = hi + (lo + (t + t * p));
    }
    if (swap) r = 0x1.921fb6p+0f + (-0x1.777a5cp-25f - r);
    if (x < 0.0f || (den == 0.0f && signbit(x))) r = 0x1.921fb6p+1f + (-0x1.777a5cp-24f - r);
    return signbit(y) ? -r : r;
}

/* acos by the half-angle identity acos x = 2 atan2(sqrt(1 - x), sqrt(1 + x)) */
float o_acosf(float x) { return 2.0f * o_atan2f(sqrtf(1.0f - x), sqrtf(1.0f + x)); }

/* op over n values, as the device math hook: 0 sin x, 1 cos x, 2 pow(x, y),
 * 3 atan2(x, y) (x the ordinate), 4 acos x */
void o_math_batch(int op, const float *x, const float *y, float *out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        float v = x[i];
        out[i] = op == 0 ? o_sinf(v) : op == 1 ? o_cosf(v) : op == 2 ? o_powf(v, y[i]) : op == 3 ? o_atan2f(v, y[i])
                                                                                                 : o_acosf(v);
    }
}

/* ---------------------------------------------------------------- Vec3f ops
 * operand order of include/math/LinearMath.h */
typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub3(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline float vdot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross3(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
/* Vec3f::normalize (LinearMath.h:101-107): three divisions by the norm */
static inline v3 vnorm(v3 a) {
    float n = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return V(a.x / n, a.y / n, a.z / n);
}

/* ---------------------------------------------------------------- RNG */
/* uniform in (0, 1] from the top 24 bits */
static inline float u01c(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

static void draws(uint32_t pix, uint32_t sample, uint32_t slot, uint32_t seed, float u[4]) {
    uint32_t ctr[4] = {pix, sample, slot, 0u}, key[2] = {seed, 0x5EEDu}, o[4];
    o_philox4x32_10(ctr, key, o);
    for (int i = 0; i < 4; ++i) u[i] = u01c(o[i]);
}

/* ---------------------------------------------------------------- samplers */
typedef struct {
    v3 o, d;
    float tmin, tmax;
} ray3;

/* diffSample, DiffImpl.cuh:12-36 */
static ray3 diff_sample(const o_material *m, const float u[4], ray3 r, v3 point, v3 n, v3 nl, v3 *mask) {
    float phi = (float)(2.0 * 3.14159265358979323846 * (double)u[0]);
    float r2 = u[1];
    float r2s = sqrtf(r2);
    v3 w = vnorm(nl);
    v3 a = fabsf(w.x) >= 0.1f ? V(0, 1, 0) : V(1, 0, 0); /* (double)|w.x| > .1 */
    v3 uu = vnorm(vcross3(a, w));
    v3 vv = vcross3(w, uu);
    float c, s = o_sincos(phi, &c);
    if (PLANT(1)) { float t = c; c = s; s = t; }
    float wz = sqrtf(1 - r2);
    if (PLANT(2)) { wz = r2s; r2s = sqrtf(1 - r2); }
    v3 d = vadd(vadd(vscale(vscale(uu, c), r2s), vscale(vscale(vv, s), r2s)), vscale(w, wz));
    r.d = vnorm(d);
    r.o = vadd(point, vscale(PLANT(15) ? n : nl, 0.001f));
    *mask = vmul(*mask, V(m->albedo[0], m->albedo[1], m->albedo[2]));
    return r;
}

/* metalSample, MetalImpl.cuh:13-41 (Phong lobe, exponent 30) */
static ray3 metal_sample(const o_material *m, const float u[4], ray3 r, v3 point, v3 n, v3 nl, v3 *mask) {
    float phi = (float)(2.0 * 3.14159265358979323846 * (double)u[0]);
    float r2 = u[1];
    float cos_t = o_powf(1 - r2, 1.0f / (30.0f + (PLANT(3) ? 0 : 1)));
    float sin_t = sqrtf(1 - cos_t * cos_t);
    v3 w = vnorm(vsub3(r.d, vscale(vscale(n, 2.0f), vdot3(n, r.d))));
    v3 a = fabsf(w.x) >= 0.1f ? V(0, 1, 0) : V(1, 0, 0);
    v3 uu = vnorm(vcross3(a, w));
    v3 vv = vcross3(w, uu);
    float c, s = o_sincos(phi, &c);
    v3 d = vadd(vadd(vscale(vscale(uu, c), sin_t), vscale(vscale(vv, s), sin_t)), vscale(w, cos_t));
    r.d = vnorm(d);
    r.o = vadd(point, vscale(PLANT(4) ? n : nl, 0.0001f));
    *mask = vmul(*mask, V(m->albedo[0], m->albedo[1], m->albedo[2]));
    return r;
}

/* refrSample, RefrImpl.cuh:12-61 (IOR from the material; the reference fixes 1.4) */
static ray3 refr_sample(const o_material *m, const float u[4], ray3 r, v3 point, v3 n, v3 nl, v3 *mask) {
    int into = vdot3(n, nl) > 0;
    float nc = 1.0f, nt = m->ior;
    float nnt = into ? nc / nt : nt / nc;
    float ddn = vdot3(r.d, nl);
    float cos2t = 1.0f - nnt * nnt * (1.f - ddn * ddn);
    if (cos2t < 0.0f) {
        r.d = vnorm(vsub3(r.d, vscale(vscale(n, 2.0f), vdot3(n, r.d))));
        r.o = vadd(point, vscale(nl, 0.001f));
    } else {
        v3 tdir = vscale(r.d, nnt);
        tdir = vsub3(tdir, vscale(PLANT(5) ? nl : n, (float)(into ? 1 : -1) * (ddn * nnt + sqrtf(cos2t))));
        tdir = vnorm(tdir);
        float R0 = (nt - nc) * (nt - nc) / (nt + nc) * (nt + nc);
        if (PLANT(6)) R0 = ((nt - nc) / (nt + nc)) * ((nt - nc) / (nt + nc));
        float c = 1.f - (into ? -ddn : vdot3(tdir, n));
        float Re = R0 + (1.f - R0) * c * c * c * c * c;
        float Tr = 1 - Re;
        float P = .25f + .5f * Re;
        float RP = PLANT(8) ? Re : Re / P;
        float TP = Tr / (1.f - P);
        if (u[0] < (PLANT(7) ? 0.25f : 0.2f)) { /* (double)u < 0.2 */
            *mask = vscale(*mask, RP);
            r.d = vnorm(vsub3(r.d, vscale(vscale(n, 2.0f), vdot3(n, r.d))));
            r.o = vadd(point, vscale(nl, 0.001f));
        } else {
            *mask = vscale(*mask, TP);
            r.d = vnorm(tdir);
            r.o = vadd(point, vscale(nl, 0.001f));
        }
    }
    return r;
}

/* specSample, SpecImpl.cuh:12-27: mirror about n, origin off nl */
static ray3 spec_sample(const o_material *m, ray3 r, v3 point, v3 n, v3 nl, v3 *mask) {
    r.d = vnorm(vsub3(r.d, vscale(vscale(n, vdot3(n, r.d)), 2.0f)));
    r.o = vadd(point, vscale(nl, 0.001f));
    if (!PLANT(14)) *mask = vmul(*mask, V(m->albedo[0], m->albedo[1], m->albedo[2]));
    return r;
}

/* coatSample, CoatImpl.cuh:12-125: reflect(), refract() and fresnel() with air
 * 1.000293 over coat 1.291 (powf(x, 2) as the rounded square), then the Fresnel
 * coin picks the mirror (specular colour) or a cosine-weighted Lambert bounce
 * (diffuse colour). Uniforms: coin u[0], then u[1], u[2] for the Lambert bounce. */
static ray3 coat_sample(const o_material *m, const float u[4], ray3 r, v3 point, v3 n, v3 nl, v3 *mask) {
    const float n_air = PLANT(13) ? 1.291f : 1.000293f, n_coat = PLANT(13) ? 1.000293f : 1.291f;
    v3 reflect_dir = vsub3(r.d, vscale(vscale(n, 2.0f), vdot3(n, r.d)));
    float cos1 = vdot3(n, vscale(r.d, -1.0f));
    float ratio = n_air / n_coat;
    float radicand = 1 - (ratio * ratio) * (1 - cos1 * cos1);
    v3 transmit = V(0, 0, 0);
    if (!(radicand < 0)) {
        float cos2 = sqrtf(radicand);
        if (cos1 > 0) transmit = vadd(vscale(r.d, ratio), vscale(n, cos1 * ratio - cos2));
        else transmit = vadd(vscale(r.d, ratio), vscale(n, cos1 * ratio + cos2));
    }
    float threshold;
    float tlen = sqrtf(vdot3(transmit, transmit));
    if ((double)tlen <= 0.12345 || vdot3(n, transmit) > 0) {
        threshold = 1.0f; /* total internal reflection */
    } else {
        float ci = vdot3(n, vscale(r.d, -1.0f));
        float ct = vdot3(vscale(n, -1.0f), transmit);
        float rs = (n_air * ci - n_coat * ct) / (n_air * ci + n_coat * ct);
        float rp = (n_air * ct - n_coat * ci) / (n_air * ct + n_coat * ci);
        threshold = (rs * rs + rp * rp) / 2.0f;
    }
    if (u[0] < threshold) {
        *mask = vmul(*mask, V(m->specular[0], m->specular[1], m->specular[2]));
        r.d = vnorm(reflect_dir);
    } else {
        float r1 = (float)(2.0 * 3.14159265358979323846 * (double)u[1]);
        float r2 = u[2];
        float r2s = sqrtf(r2);
        v3 w = vnorm(nl);
        v3 a = fabsf(w.x) >= 0.1f ? V(0, 1, 0) : V(1, 0, 0);
        v3 uu = vnorm(vcross3(a, w));
        v3 vv = vcross3(w, uu);
        float c, s = o_sincos(r1, &c);
        v3 d = vadd(vadd(vscale(vscale(uu, c), r2s), vscale(vscale(vv, s), r2s)), vscale(w, sqrtf(1 - r2)));
        r.d = vnorm(d);
        *mask = vmul(*mask, V(m->albedo[0], m->albedo[1], m->albedo[2]));
    }
    r.o = vadd(point, vscale(nl, 0.001f));
    return r;
}

/* The surface interaction of renderKernel (RenderKernel.cu:72-92) after the
 * emission is added: n = normalize(hit.normal), nl faces the incoming ray
 * (IntersectKernel.cuh:68-72), then the material's sampler. Returns 0 when the
 * material ends the path (an emitter: emission only). */
static int ref_bounce(const o_material *m, const float u[4], ray3 *r, v3 point, v3 nrm, v3 *mask) {
    v3 n = vnorm(nrm);
    v3 nl = vdot3(n, r->d) < 0 ? n : vscale(n, -1);
    if (m->type == 0) *r = diff_sample(m, u, *r, point, n, nl, mask);
    else if (m->type == 1) *r = metal_sample(m, u, *r, point, n, nl, mask);
    else if (m->type == 2) *r = refr_sample(m, u, *r, point, n, nl, mask);
    else if (m->type == 4) *r = spec_sample(m, *r, point, n, nl, mask);
    else if (m->type == 5) *r = coat_sample(m, u, *r, point, n, nl, mask);
    else return 0;
    return 1;
}

int o_ref_bounce(const o_material *m, const float u[4], const float d[3], const float point[3], const float nrm[3],
                 float out[9]) {
    ray3 r;
    r.o = V(0, 0, 0);
    r.d = V(d[0], d[1], d[2]);
    r.tmin = 0.0f;
    r.tmax = 1e20f;
    v3 mask = V(1, 1, 1);
    int cont = ref_bounce(m, u, &r, V(point[0], point[1], point[2]), V(nrm[0], nrm[1], nrm[2]), &mask);
    out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z;
    out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
    out[6] = mask.x; out[7] = mask.y; out[8] = mask.z;
    return cont;
}

/* ---------------------------------------------------------------- environment */
/* hdrSample, HDRKernel.cuh:6-42: longitude atan2f(d.x, d.z) wrapped into [0, TWO_PI),
 * latitude acosf(d.y); u = lon / TWO_PI, v = lat / PI (the reference's PI, slightly
 * below pi); nearest texel (int)(u * width) + (int)(v * height) * width, read as a
 * linear index (u * width reaching width reads the next row's first texel), times
 * 2 (and the scene's env scale). An index past the last texel is clamped to it. */
static v3 env_ref(const o_env *e, v3 d, float pi, float two_pi) {
    if (!e || !e->rgb || e->width <= 0 || e->height <= 0) return V(0, 0, 0);
    float lx = PLANT(9) ? o_atan2f(d.z, d.x) : o_atan2f(d.x, d.z);
    lx = lx < 0.f ? lx + two_pi : lx;
    float ly = o_acosf(d.y);
    float u = lx / two_pi, v = ly / (PLANT(11) ? two_pi : pi);
    float fu = u * (float)e->width, fv = v * (float)e->height;
    int64_t u2 = fu >= 0.0f ? (int64_t)fu : 0, v2 = fv >= 0.0f ? (int64_t)fv : 0;
    int64_t idx = PLANT(12) ? v2 + u2 * e->height : u2 + v2 * e->width;
    const int64_t last = (int64_t)e->width * e->height - 1;
    if (idx > last) idx = last;
    const float *t = e->rgb + 3 * idx;
    float s = (PLANT(10) ? 1.0f : 2.0f) * e->scale;
    return V(t[0] * s, t[1] * s, t[2] * s);
}

void o_env_lookup(const o_env *e, const float d[3], float pi, float two_pi, float out[3]) {
    v3 c = env_ref(e, V(d[0], d[1], d[2]), pi, two_pi);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

/* ---------------------------------------------------------------- analytic primitives */
/* planeIntersect, PlaneImpl.cuh:10-52 (not culling; double literals) */
static float plane_t(const o_geometry *g, v3 o, v3 d) {
    v3 p0 = V(g->a[0], g->a[1], g->a[2]);
    v3 e1 = vsub3(V(g->b[0], g->b[1], g->b[2]), p0);
    v3 e2 = vsub3(V(g->c[0], g->c[1], g->c[2]), p0);
    v3 P = vcross3(d, e2);
    float det = vdot3(e1, P);
    if ((double)det > -0.00001 && (double)det < 0.00001) return INFINITY;
    float inv_det = 1.f / det;
    v3 T = vsub3(o, p0);
    v3 Q = vcross3(T, e1);
    float u = vdot3(T, P) * inv_det;
    float v = vdot3(d, Q) * inv_det;
    if (u < 0.f || u > 1.f || v < 0.f || v > 1.f) return INFINITY;
    float t = vdot3(e2, Q) * inv_det;
    return (double)t > 0.00001 ? t : INFINITY;
}

/* sphereIntersect, SphereImpl.cuh:10-32 */
static float sphere_t(const o_geometry *g, v3 o, v3 d, float tmax) {
    v3 op = vsub3(V(g->a[0], g->a[1], g->a[2]), o);
    float eps = 0.01f;
    float b = vdot3(op, d);
    float disc = b * b - vdot3(op, op) + g->b[0] * g->b[0];
    if (disc >= 0) {
        disc = sqrtf(disc);
        float t, dist = (t = b - disc) > eps ? t : ((t = b + disc) > eps ? t : 0.0f);
        if (0.01f < dist && dist < tmax) return dist;
    }
    return INFINITY;
}

float o_geom_t(const o_geometry *g, const o_ray *r) {
    v3 o = V(r->ox, r->oy, r->oz), d = V(r->dx, r->dy, r->dz);
    return g->type == 0 ? sphere_t(g, o, d, r->tmax) : plane_t(g, o, d);
}

/* ---------------------------------------------------------------- render */
typedef struct {
    const o_bvh *b;
    const float *tri_normals;
    const int32_t *tri_mat;
    const o_material *mats;
    const o_geometry *geoms;
    int32_t num_geoms;
    const o_camera *cam;
    const o_ref_params *p;
    const o_env *env;
    float *out;
    int32_t y0, y1;
} o_rjob;

static void render_rows(const o_rjob *j) {
    const o_camera *cam = j->cam;
    const o_ref_params *p = j->p;
    const int W = cam->width, H = cam->height;
    /* camera basis and extents, RenderKernel.cu:131-146 */
    v3 pos = V(cam->eye[0], cam->eye[1], cam->eye[2]);
    v3 view = vnorm(V(cam->view[0], cam->view[1], cam->view[2]));
    v3 up = vnorm(V(cam->up[0], cam->up[1], cam->up[2]));
    v3 hax = vnorm(vcross3(view, up));
    v3 vax = vnorm(vcross3(hax, view));
    v3 middle = vadd(pos, view);
    v3 horizontal = vscale(hax, tanf(cam->fov[0] * 0.5f * (p->pi / 180)));
    v3 vertical = vscale(vax, tanf(-cam->fov[1] * 0.5f * (p->pi / 180)));
    const float inv_spp = 1.0f / (float)p->spp;

    for (int32_t Y = j->y0; Y < j->y1; ++Y)
        for (int32_t X = 0; X < W; ++X) {
            uint32_t pix = (uint32_t)(Y * W + X);
            v3 final = V(0, 0, 0);
            for (int s = 0; s < p->spp; ++s) {
                uint32_t sg = (uint32_t)(p->sample_offset + s);
                float u[4];
                draws(pix, sg, 0u, p->seed, u);
                float jx = u[0] - 0.5f, jy = u[1] - 0.5f;
                float sx = (jx + (float)X) / (float)(W - 1);
                float sy = (jy + (float)Y) / (float)(H - 1);
                v3 on_plane = vadd(vadd(middle, vscale(horizontal, (2 * sx) - 1)), vscale(vertical, (2 * sy) - 1));
                v3 on_image = vadd(pos, vscale(vsub3(on_plane, pos), cam->focal));
                v3 ap = pos;
                if (cam->aperture > 1e-5f) { /* (double)aperture > 0.00001 */
                    float angle = p->two_pi * u[2];
                    float dist = cam->aperture * sqrtf(u[3]);
                    float ca, sa = o_sincos(angle, &ca);
                    ap = vadd(vadd(pos, vscale(hax, ca * dist)), vscale(vax, sa * dist));
                }
                ray3 r;
                r.d = vnorm(vnorm(vsub3(on_image, ap)));
                r.o = ap;
                r.tmin = p->tmin;
                r.tmax = p->tmax;

                /* renderKernel, RenderKernel.cu:27-95 */
                v3 mask = V(1, 1, 1), acc = V(0, 0, 0);
                for (int bounce = 0; bounce < p->max_depth; ++bounce) {
                    o_ray q = {r.o.x, r.o.y, r.o.z, r.tmin, r.d.x, r.d.y, r.d.z, r.tmax};
                    o_hit h;
                    o_intersect(j->b, &q, 0, &h, NULL, NULL);
                    /* geometry list after the triangles, RenderKernel.cu:38-61 */
                    int gsel = -1;
                    for (int32_t k = 0; k < j->num_geoms; ++k) {
                        float t = o_geom_t(&j->geoms[k], &q);
                        if (t < q.tmax && t < h.t) {
                            h.t = t;
                            gsel = k;
                        }
                    }
                    if ((gsel < 0 && h.tri < 0) || h.t > 1e19f) { /* miss: hdrSample, RenderKernel.cu:65-69 */
                        acc = vadd(acc, vmul(mask, env_ref(j->env, r.d, p->pi, p->two_pi)));
                        break;
                    }
                    v3 point = vadd(r.o, vscale(r.d, h.t));
                    v3 nrm;
                    const o_material *m;
                    if (gsel >= 0) {
                        const o_geometry *g = &j->geoms[gsel];
                        if (g->type == 0) {
                            nrm = V(point.x - g->a[0], point.y - g->a[1], point.z - g->a[2]);
                        } else {
                            v3 p0 = V(g->a[0], g->a[1], g->a[2]);
                            nrm = vcross3(vsub3(V(g->b[0], g->b[1], g->b[2]), p0), vsub3(V(g->c[0], g->c[1], g->c[2]), p0));
                        }
                        m = &j->mats[g->material];
                    } else {
                        const float *tn = j->tri_normals + 12 * (int64_t)h.tri;
                        float w = 1 - h.u - h.v;
                        nrm = V(tn[0] * w + tn[4] * h.u + tn[8] * h.v, tn[1] * w + tn[5] * h.u + tn[9] * h.v,
                                tn[2] * w + tn[6] * h.u + tn[10] * h.v);
                        m = &j->mats[j->tri_mat[h.tri]];
                    }
                    acc = vadd(acc, vmul(mask, V(m->emission[0], m->emission[1], m->emission[2])));
                    draws(pix, sg, (uint32_t)(bounce + 1), p->seed, u);
                    if (!ref_bounce(m, u, &r, point, nrm, &mask)) break; /* emitter: emission only */
                }
                final = vadd(final, vscale(acc, inv_spp));
            }
            float *o = j->out + 3 * (int64_t)pix;
            o[0] = final.x;
            o[1] = final.y;
            o[2] = final.z;
        }
}

static void *render_worker(void *arg) {
    render_rows((const o_rjob *)arg);
    return NULL;
}

int o_render_ref(const o_bvh *b, const float *tri_normals, const int32_t *tri_mat, const o_material *mats,
                 const o_geometry *geoms, int32_t num_geoms, const o_env *env, const o_camera *cam,
                 const o_ref_params *p, int32_t y0, int32_t y1, float *out_rgb, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (y0 < 0 || y1 > cam->height || y0 > y1) return -1;
    o_rjob jobs[256];
    pthread_t th[256];
    int32_t rows = y1 - y0;
    for (int k = 0; k < nthreads; ++k) {
        o_rjob jb = {b, tri_normals, tri_mat, mats, geoms, num_geoms, cam, p, env, out_rgb, y0 + rows * k / nthreads,
                     y0 + rows * (k + 1) / nthreads};
        jobs[k] = jb;
        pthread_create(&th[k], NULL, render_worker, &jobs[k]);
    }
    for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
    return 0;
}

/* ---------------------------------------------------------------- display */
/* The tail of pathTracingKernel (RenderKernel.cu:202-217) and the buffer reset of
 * Renderer::display (Renderer.cpp:86-93): frame_number 1 starts the running sum
 * from the cudaMemset(buffer, 1, ...) pattern (every float 0x01010101); then
 * sum += frame, mean = sum / frame_number, c = fmaxf(0, fminf(mean, 1)) (clamp in
 * CutilMath.h:44-46, so NaN gives 1), channel = (unsigned char)(powf(c, 1 / 2.2f) *
 * 255), alpha 1, packed r | g << 8 | b << 16 | a << 24 as the Color union's uchar4. */
static uint32_t display_channel(float sum, float frames) {
    float mean = sum / frames;
    float c = fmaxf(0.0f, fminf(mean, 1.0f));
    float p = o_powf(c, PLANT(18) ? 2.2f : 1.0f / 2.2f) * 255.0f;
    return (uint32_t)(uint8_t)(PLANT(16) ? p + 0.5f : p);
}

void o_display_frame(const float *frame, float *sum, int64_t npix, int32_t frame_number, uint32_t *rgba) {
    const uint32_t bits = PLANT(17) ? 0u : 0x01010101u;
    float fresh;
    memcpy(&fresh, &bits, 4);
    const float frames = (float)frame_number;
    for (int64_t i = 0; i < npix; ++i) {
        uint32_t c = 1u << 24;
        for (int k = 0; k < 3; ++k) {
            float v = (frame_number == 1 ? fresh : sum[3 * i + k]) + frame[3 * i + k];
            sum[3 * i + k] = v;
            c |= display_channel(v, frames) << (8 * k);
        }
        rgba[i] = c;
    }
}
