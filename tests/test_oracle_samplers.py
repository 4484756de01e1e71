"""The oracle's reference-mode samplers against their closed-form laws.

Each sampler of the reference (include/material/DiffImpl.cuh:12-36,
MetalImpl.cuh:13-41, RefrImpl.cuh:12-61, SpecImpl.cuh:12-27, CoatImpl.cuh:75-125)
is checked through the oracle hook o_ref_bounce, which runs the same code as the
oracle's render loop, against what the mathematics fixes:

  diffuse   cos(theta) about nl = sqrt(1 - u1); azimuth 2 pi u0 measured from the
            frame axis U = normalize(cross(Y or X, w)) towards V = w x U; chi-square
            of the drawn directions against cos(theta) / pi
  Phong     the lobe axis is the mirror direction; cos(alpha) = (1 - u1)^(1/31),
            i.e. density (31 / 2 pi) cos^30(alpha); chi-square of cos^31(alpha)
  smallpt   reflection exactly when u0 < 0.2; weights Re / P and Tr / (1 - P) with
            P = 1/4 + Re/2 and Schlick's Re = R0 + (1 - R0)(1 - cos)^5 on the air
            side, R0 = (nt - nc)^2 (the reference's operator precedence); Snell's law
            in and out; total internal reflection above asin(1 / ior)
  mirror    reflection about n, colour product
  coat      the Fresnel reflectance of air (1.000293) over the coat (1.291) is the
            mirror probability; otherwise the diffuse law above
  all       the next origin is point + nl * offset (the incident side)

Every check is also run on oracle builds with one planted mistake each (a swapped
sin/cos, the inverted elevation map, a 1/30 Phong exponent, n in place of nl, the
textbook R0, a 0.25 reflection share, an unweighted reflection, swapped coat
indices, a missing colour product): each of those builds must fail at least one.
"""
import math
from types import SimpleNamespace

import numpy as np
import pytest

F32 = np.float32


def mat(type, albedo=(0.7, 0.5, 0.3), ior=1.4, specular=(0.9, 0.8, 0.6)):
    return SimpleNamespace(type=type, albedo=albedo, emission=(0, 0, 0), roughness=0.0, ior=ior, specular=specular)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def frame(w):
    """The reference's frame about w (DiffImpl.cuh:21-24): U from Y (|w.x| > 0.1) or X."""
    a = np.array([0.0, 1.0, 0.0]) if abs(F32(w[0])) >= F32(0.1) else np.array([1.0, 0.0, 0.0])
    U = unit(np.cross(a, w))
    return U, np.cross(w, U)


def azimuth_err(d, w, phi_want):
    U, V = frame(w)
    phi = math.atan2(float(np.dot(d, V)), float(np.dot(d, U)))
    return abs((phi - phi_want + math.pi) % (2 * math.pi) - math.pi)


NORMALS = [(0, 1, 0), (0, 0, -2.5), (0.3, -0.2, 0.9), (0.1, 0.7, -0.7), (0.0999, 0.5, 0.86), (-0.8, 0.1, 0.3)]
POINTS = [(0.25, -0.5, 1.0), (3.0, 2.0, -7.0)]


def incoming(n, side, tilt):
    """A unit direction making angle `tilt` with -n (side = +1: arrives at the front, n . d < 0)."""
    n = unit(n)
    U, _ = frame(n)
    return unit(-side * math.cos(tilt) * n + math.sin(tilt) * U)


def uniforms_grid():
    g = [2.0 ** -24, 0.013, 0.25, 0.49, 0.5, 0.77, 0.999, 1.0]
    return [(a, b) for a in g for b in g]


def check_origin(o, point, nl, offset):
    want = np.asarray(point, np.float64) + offset * nl
    assert np.all(np.abs(o - want) <= 2e-6 * (1 + np.abs(want))), (o, want)


# ---------------------------------------------------------------------- laws
def law_diffuse(O, L):
    m = mat(0)
    for nrm in NORMALS:
        n = unit(nrm)
        for side in (1, -1):
            d = incoming(nrm, side, 0.4)
            nl = n if np.dot(n, d) < 0 else -n
            for point in POINTS:
                for u0, u1 in uniforms_grid():
                    cont, o, dd, mk = O.ref_bounce(m, (u0, u1, 0.5, 0.5), d, point, nrm, lib_=L)
                    assert cont
                    dd = dd.astype(np.float64)
                    assert abs(np.linalg.norm(dd) - 1) < 2e-6
                    assert abs(np.dot(dd, nl) - math.sqrt(1 - u1)) < 3e-6, ("elevation", u1)
                    if u1 > 1e-4:
                        assert azimuth_err(dd, nl, 2 * math.pi * u0) < 2e-4 / math.sqrt(u1), ("azimuth", u0, u1)
                    check_origin(o, point, nl, 1e-3)
                    assert np.array_equal(mk, np.float32(m.albedo))


def law_diffuse_chi2(O, L, n=20000, seed=3):
    """Equal-probability bins of cos^2(theta) x azimuth under the density cos(theta) / pi."""
    m = mat(0)
    rng = np.random.default_rng(seed)
    nrm = (0.3, -0.2, 0.9)
    n_ = unit(nrm)
    d = incoming(nrm, 1, 0.3)
    U, V = frame(n_)
    k = 8
    cnt = np.zeros((k, k))
    for u0, u1 in rng.uniform(0, 1, (n, 2)).astype(np.float32):
        _, _, dd, _ = O.ref_bounce(m, (u0, u1, 0.5, 0.5), d, (0, 0, 0), nrm, lib_=L)
        c = float(np.dot(dd, n_))
        phi = math.atan2(float(np.dot(dd, V)), float(np.dot(dd, U))) % (2 * math.pi)
        cnt[min(k - 1, int(c * c * k)), min(k - 1, int(phi / (2 * math.pi) * k))] += 1
    e = n / k / k
    chi2 = ((cnt - e) ** 2 / e).sum()
    assert chi2 < 120, chi2   # 63 dof: p ~ 1e-5


def law_phong(O, L):
    m = mat(1)
    for nrm in NORMALS:
        n = unit(nrm)
        for side in (1, -1):
            d = incoming(nrm, side, 0.6)
            nl = n if np.dot(n, d) < 0 else -n
            w = unit(d - 2 * np.dot(n, d) * n)
            for point in POINTS:
                for u0, u1 in uniforms_grid():
                    cont, o, dd, mk = O.ref_bounce(m, (u0, u1, 0.5, 0.5), d, point, nrm, lib_=L)
                    dd = dd.astype(np.float64)
                    ca = (1 - float(F32(u1))) ** (1 / 31)
                    assert abs(np.dot(dd, w) - ca) < 3e-6, ("lobe", u1)
                    sa = math.sqrt(max(0.0, 1 - ca * ca))
                    if sa > 1e-2:
                        assert azimuth_err(dd, w, 2 * math.pi * u0) < 2e-4 / sa, ("azimuth", u0, u1)
                    check_origin(o, point, nl, 1e-4)
                    assert np.array_equal(mk, np.float32(m.albedo))


def law_phong_chi2(O, L, n=20000, seed=4):
    """cos^31(alpha) about the mirror direction is uniform: density (31 / 2 pi) cos^30."""
    m = mat(1)
    rng = np.random.default_rng(seed)
    nrm = (0.1, 0.7, -0.7)
    n_ = unit(nrm)
    d = incoming(nrm, 1, 0.5)
    w = unit(d - 2 * np.dot(n_, d) * n_)
    k = 20
    cnt = np.zeros(k)
    for u0, u1 in rng.uniform(0, 1, (n, 2)).astype(np.float32):
        _, _, dd, _ = O.ref_bounce(m, (u0, u1, 0.5, 0.5), d, (0, 0, 0), nrm, lib_=L)
        c = min(1.0, float(np.dot(dd, w)))
        cnt[min(k - 1, int(c ** 31 * k))] += 1
    e = n / k
    assert ((cnt - e) ** 2 / e).sum() < 55     # 19 dof: p ~ 3e-5


def schlick(cos_air, nt, nc=1.0):
    r0 = (nt - nc) ** 2          # (nt - nc)^2 / (nt + nc) * (nt + nc), RefrImpl.cuh:36
    return r0 + (1 - r0) * (1 - cos_air) ** 5


def law_refr(O, L):
    for ior in (1.4, 1.5):
        m = mat(2, ior=ior)
        crit = math.asin(1 / ior)
        for nrm in NORMALS[:4]:
            n = unit(nrm)
            for side, angles in ((1, (0.0, 0.2, 0.6, 1.1, 1.45)), (-1, (0.0, 0.3, 0.7, crit - 0.01))):
                for th in angles:
                    d = incoming(nrm, side, th)
                    nl = n if np.dot(n, d) < 0 else -n
                    into = side == 1
                    s_t = math.sin(th) / ior if into else math.sin(th) * ior
                    th_t = math.asin(s_t)
                    re = schlick(math.cos(th) if into else math.cos(th_t), ior)
                    p = 0.25 + 0.5 * re
                    mirror = unit(d - 2 * np.dot(n, d) * n)
                    for u0 in (2.0 ** -24, 0.1, float(np.nextafter(F32(0.2), F32(0)))):
                        _, o, dd, mk = O.ref_bounce(m, (u0, 0.5, 0.5, 0.5), d, (0.25, -0.5, 1.0), nrm, lib_=L)
                        assert np.allclose(dd, mirror, atol=3e-6), ("reflect", ior, side, th, u0)
                        assert np.allclose(mk, re / p, rtol=2e-5), ("Re/P", ior, side, th, mk, re / p)
                        check_origin(o, (0.25, -0.5, 1.0), nl, 1e-3)
                    for u0 in (0.2, 0.24, 0.6, 1.0):
                        _, o, dd, mk = O.ref_bounce(m, (u0, 0.5, 0.5, 0.5), d, (0.25, -0.5, 1.0), nrm, lib_=L)
                        dd = dd.astype(np.float64)
                        # Snell: the transmitted ray crosses the surface, in the plane of d and n
                        assert np.dot(dd, nl) < 0, ("side", ior, side, th)
                        assert abs(math.sqrt(max(0.0, 1 - np.dot(dd, nl) ** 2)) - s_t) < 3e-6, ("snell", ior, side, th)
                        assert abs(np.dot(dd, np.cross(n, d))) < 3e-6
                        assert np.allclose(mk, (1 - re) / (1 - p), rtol=2e-5), ("Tr/(1-P)", ior, side, th)
                        check_origin(o, (0.25, -0.5, 1.0), nl, 1e-3)
            # leaving the medium above the critical angle: always reflected, unweighted
            for th in (crit + 0.01, 1.0, 1.4):
                d = incoming(nrm, -1, th)
                mirror = unit(d - 2 * np.dot(n, d) * n)
                for u0 in (0.1, 0.5, 0.99):
                    _, o, dd, mk = O.ref_bounce(m, (u0, 0.5, 0.5, 0.5), d, (0.25, -0.5, 1.0), nrm, lib_=L)
                    assert np.allclose(dd, mirror, atol=3e-6), ("TIR", ior, th)
                    assert np.array_equal(mk, np.ones(3, np.float32))
                    check_origin(o, (0.25, -0.5, 1.0), n, 1e-3 * -1)


def law_refr_normal_incidence(O, L):
    """Entering along -n at IOR 1.4: Re = R0 = 0.16, P = 0.33: weights 0.16 / 0.33 and 0.84 / 0.67."""
    m = mat(2, ior=1.4)
    _, _, _, mk = O.ref_bounce(m, (0.1, 0.5, 0.5, 0.5), (0, 0, -1), (0, 0, 0), (0, 0, 1), lib_=L)
    assert np.allclose(mk, 0.16 / 0.33, rtol=1e-6)
    _, _, dd, mk = O.ref_bounce(m, (0.5, 0.5, 0.5, 0.5), (0, 0, -1), (0, 0, 0), (0, 0, 1), lib_=L)
    assert np.allclose(mk, 0.84 / 0.67, rtol=1e-6)
    assert np.allclose(dd, (0, 0, -1), atol=1e-7)


def law_mirror(O, L):
    m = mat(4)
    for nrm in NORMALS:
        n = unit(nrm)
        for side in (1, -1):
            d = incoming(nrm, side, 0.7)
            nl = n if np.dot(n, d) < 0 else -n
            _, o, dd, mk = O.ref_bounce(m, (0.3, 0.3, 0.3, 0.3), d, POINTS[0], nrm, lib_=L)
            assert np.allclose(dd, unit(d - 2 * np.dot(n, d) * n), atol=3e-6)
            assert np.array_equal(mk, np.float32(m.albedo))
            check_origin(o, POINTS[0], nl, 1e-3)


def fresnel(ci, n1, n2):
    st = n1 / n2 * math.sqrt(max(0.0, 1 - ci * ci))
    if st >= 1:
        return 1.0
    ct = math.sqrt(1 - st * st)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n1 * ct - n2 * ci) / (n1 * ct + n2 * ci)
    return (rs * rs + rp * rp) / 2


def law_coat(O, L):
    m = mat(5)
    n1, n2 = 1.000293, 1.291
    assert abs(fresnel(1.0, n1, n2) - ((n2 - n1) / (n2 + n1)) ** 2) < 1e-12
    nrm = (0.3, -0.2, 0.9)
    n = unit(nrm)
    for th in (0.0, 0.5, 1.047, 1.3):
        d = incoming(nrm, 1, th)
        thr = fresnel(math.cos(th), n1, n2)
        mirror = unit(d - 2 * np.dot(n, d) * n)
        _, o, dd, mk = O.ref_bounce(m, (thr * (1 - 1e-4), 0.3, 0.6, 0.5), d, POINTS[0], nrm, lib_=L)
        assert np.allclose(dd, mirror, atol=3e-6) and np.array_equal(mk, np.float32(m.specular)), ("coat mirror", th)
        check_origin(o, POINTS[0], n, 1e-3)
        _, o, dd, mk = O.ref_bounce(m, (thr * (1 + 1e-4), 0.3, 0.6, 0.5), d, POINTS[0], nrm, lib_=L)
        assert np.array_equal(mk, np.float32(m.albedo)), ("coat diffuse", th)
        assert abs(np.dot(dd, n) - math.sqrt(1 - 0.6)) < 3e-6
        assert azimuth_err(dd.astype(np.float64), n, 2 * math.pi * 0.3) < 1e-4


LAWS = {"diffuse": law_diffuse, "diffuse_chi2": law_diffuse_chi2, "phong": law_phong,
        "phong_chi2": law_phong_chi2, "refr": law_refr, "refr_normal": law_refr_normal_incidence,
        "mirror": law_mirror, "coat": law_coat}


@pytest.mark.parametrize("name", sorted(LAWS))
def test_sampler_law(oracle, name):
    LAWS[name](oracle, oracle.lib)


# planted mistake -> the laws that must catch it (tt_oracle_render.c PLANT() sites)
PLANTS = {1: ("diffuse",), 2: ("diffuse",), 3: ("phong", "phong_chi2"), 4: ("phong",), 5: ("refr",),
          6: ("refr", "refr_normal"), 7: ("refr",), 8: ("refr", "refr_normal"), 13: ("coat",), 14: ("mirror",),
          15: ("diffuse",)}


@pytest.mark.parametrize("k", sorted(PLANTS))
def test_planted_mistake_fails_a_law(oracle, k):
    L = oracle.load_planted(k)
    failed = []
    for name in PLANTS[k]:
        try:
            LAWS[name](oracle, L)
        except AssertionError:
            failed.append(name)
    assert failed, f"planted mistake {k} passed every law {PLANTS[k]}"
