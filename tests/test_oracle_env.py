"""The oracle's hdrSample restatement (include/hdr/HDRKernel.cuh:6-42) and its
atan2 / acos against closed forms.

Texel selection: for a direction of longitude phi = atan2(d.x, d.z) in [0, 2 pi)
and colatitude theta = acos(d.y), the lookup reads column floor(W phi / 2pi) and row
floor(H theta / PI) (PI = 3.14156265, the reference's constant) as the linear index
col + row * W, and returns that texel times 2 times the env scale. A map whose texels
hold their own (column, row) checks this at texel centres over the whole sphere, at
the poles and axes, at the longitude wrap (u * W reaching W reads the next row's
first texel) and at the south pole (v * H reaching H, past the end, is clamped to the
last texel). Planted mistakes (swapped atan2 operands, the missing x2, TWO_PI for PI,
a transposed index) must each fail one of these checks.
"""
import math

import numpy as np
import pytest

F32 = np.float32
PI_REF = 3.14156265
TWO_PI_REF = 6.2831853071795864769


def index_env(w=64, h=32):
    img = np.zeros((h, w, 3), np.float32)
    img[:, :, 0] = np.arange(w)[None, :]
    img[:, :, 1] = np.arange(h)[:, None]
    img[:, :, 2] = 7.0
    return img


def sph(theta, phi):
    return np.array([math.sin(theta) * math.sin(phi), math.cos(theta), math.sin(theta) * math.cos(phi)])


def law_texel_centres(O, L):
    w, h, scale = 64, 32, 1.5
    env = O.Env(index_env(w, h), scale)
    dirs, want = [], []
    for r in range(h):
        for c in range(w):
            dirs.append(sph(math.pi * (r + 0.5) / h, 2 * math.pi * (c + 0.5) / w))
            want.append((c, r, 7.0))
    got = env.lookup(np.array(dirs), lib_=L)
    assert np.array_equal(got, np.float32(np.array(want) * 2 * scale)), "texel centres"


def law_axes_and_edges(O, L):
    w, h = 64, 32
    env = O.Env(index_env(w, h), 1.0)
    cases = [
        ((0.0, 1.0, 0.0), (0, 0)),          # north pole: colatitude 0, atan2(0, 0) = 0
        ((0.0, 0.0, 1.0), (0, 16)),         # +z: longitude 0; colatitude pi/2 / PI -> row 16
        ((1.0, 0.0, 0.0), (16, 16)),        # +x: longitude pi/2 -> column W/4
        ((0.0, 0.0, -1.0), (32, 16)),       # -z: longitude pi -> column W/2
        ((-1.0, 0.0, 0.0), (48, 16)),       # -x: longitude 3pi/2 (wrapped) -> column 3W/4
        ((-1e-9, 0.0, 1.0), (0, 17)),       # longitude -1e-9 + TWO_PI rounds to TWO_PI: u*W = W, next row
        ((0.0, -1.0, 0.0), (w - 1, h - 1)), # south pole: v*H = H (PI < pi), past the end: last texel
    ]
    for d, (c, r) in cases:
        got = env.lookup(np.array([d]), lib_=L)[0]
        assert np.array_equal(got, np.float32([2 * c, 2 * r, 14.0])), (d, got, (c, r))


def law_scale_and_colour(O, L):
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 3, (8, 16, 3)).astype(np.float32)
    env = O.Env(img, 0.25)
    got = env.lookup(np.array([sph(math.pi * 2.5 / 8, 2 * math.pi * 4.5 / 16)]), lib_=L)[0]
    assert np.array_equal(got, img[2, 4] * F32(0.5)), "x2 times scale"


LAWS = {"centres": law_texel_centres, "axes": law_axes_and_edges, "scale": law_scale_and_colour}


@pytest.mark.parametrize("name", sorted(LAWS))
def test_env_law(oracle, name):
    LAWS[name](oracle, oracle.lib)


PLANTS = {9: ("centres", "axes"), 10: ("centres", "scale"), 11: ("centres", "axes"), 12: ("centres", "axes")}


@pytest.mark.parametrize("k", sorted(PLANTS))
def test_planted_env_mistake_fails_a_law(oracle, k):
    L = oracle.load_planted(k)
    failed = []
    for name in PLANTS[k]:
        try:
            LAWS[name](oracle, L)
        except AssertionError:
            failed.append(name)
    assert failed, f"planted mistake {k} passed {PLANTS[k]}"


def _ulps(got, ref):
    ulp = np.spacing(np.maximum(np.abs(ref), 1e-38).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / ulp


def test_atan2_within_2p5_ulp(oracle):
    rng = np.random.default_rng(0)
    y = np.concatenate([rng.standard_normal(200000), rng.uniform(-1, 1, 100000) * 10.0 ** rng.uniform(-30, 30, 100000),
                        np.sin(np.linspace(0, 2 * np.pi, 100001))]).astype(np.float32)
    x = np.concatenate([rng.standard_normal(200000), rng.uniform(-1, 1, 100000) * 10.0 ** rng.uniform(-30, 30, 100000),
                        np.cos(np.linspace(0, 2 * np.pi, 100001))]).astype(np.float32)
    got = oracle.math(3, y, x)
    e = _ulps(got, np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert e.max() <= 2.5, e.max()


def test_atan2_special_values(oracle):
    a = oracle.lib.o_atan2f
    assert a(0.0, 0.0) == 0.0 and math.copysign(1, a(-0.0, 0.0)) < 0
    assert a(0.0, -0.0) == F32(math.pi) and a(-0.0, -0.0) == -F32(math.pi)
    assert a(1.0, 0.0) == F32(math.pi / 2) and a(-1.0, -0.0) == -F32(math.pi / 2)
    assert a(1.0, 1.0) == F32(math.pi / 4) and a(1.0, -1.0) == F32(3 * math.pi / 4)
    assert math.isnan(a(float("nan"), 1.0)) and math.isnan(a(1.0, float("nan")))


def test_acos_within_3p5_ulp(oracle):
    rng = np.random.default_rng(1)
    x = np.concatenate([np.linspace(-1, 1, 200001), 1 - rng.uniform(0, 1e-3, 100000),
                        -1 + rng.uniform(0, 1e-3, 100000), [1.0, -1.0, 0.0, -0.0]]).astype(np.float32)
    got = oracle.math(4, x)
    e = _ulps(got, np.arccos(x.astype(np.float64)))
    assert e.max() <= 3.5, e.max()
    assert oracle.lib.o_acosf(1.0) == 0.0 and oracle.lib.o_acosf(-1.0) == F32(math.pi)
    assert math.isnan(oracle.lib.o_acosf(1.5))


def test_render_in_uniform_env_closed_form(tt, oracle):
    """A convex diffuse sphere of albedo a under a constant env L (reference mode): a
    camera ray that misses sees 2L; one that hits bounces once off the convex surface
    and escapes, so every sample carries a * 2L. Pixel = spp * (x / spp) summed."""
    sphere = tt.Mesh.icosphere(3, 0.6, (0, 0, 0))
    b = tt.BVH.build(sphere)
    a, L, spp = 0.625, 0.75, 4
    mats = [tt.material(tt.MAT_DIFFUSE, (a, a, a))]
    cam = tt.camera((0, 0, 3), (0, 0, -1), fov=(40, 40), width=40, height=40)
    env = oracle.Env(np.full((16, 32, 3), L, np.float32), 1.0)
    img = oracle.render_ref(oracle.CompactBVH(b.export_compact()), sphere.arrays(), mats, cam, spp=spp, max_depth=6,
                            seed=2, env=env)
    hit = np.float32(a * 2 * L)
    vals = np.unique(img)
    # each pixel is either 2L (all samples miss), a*2L (all hit) or a mix on the silhouette
    assert np.isclose(img[0, 0, 0], 2 * L) and np.isclose(img[20, 20, 0], hit, rtol=1e-6)
    inside = np.abs(img - hit) <= 1e-6 * hit
    outside = np.abs(img - 2 * L) <= 1e-6 * 2 * L
    assert inside.mean() > 0.15 and outside.mean() > 0.5
    mixed = ~(inside | outside)
    # silhouette pixels: k of spp samples hit: value = (k a + (spp - k)) 2L / spp
    mixes = np.array([(k * a + (spp - k)) * 2 * L / spp for k in range(spp + 1)])
    assert np.all(np.min(np.abs(img[mixed][:, None] - mixes[None, :]), axis=1) <= 1e-6), vals[:20]
