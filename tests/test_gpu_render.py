"""GPU render parity: reference mode vs the oracle restatement; path mode vs
closed forms (furnace) and vs the reference mode's converged mean."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_device_math_bit_identical_to_oracle(tt, oracle, ctx):
    _say("device math")
    rng = np.random.default_rng(1)
    x = np.concatenate([np.float32(np.linspace(0, 6.2831853, 20001)), rng.uniform(0, 6.3, 20000).astype(np.float32)])
    for op, fn in ((0, oracle.lib.o_sinf), (1, oracle.lib.o_cosf)):
        got = tt.debug_math(ctx, op, x)
        want = np.array([fn(float(v)) for v in x], np.float32)
        assert np.array_equal(got.view(np.int32), want.view(np.int32)), op
    k = rng.integers(0, 1 << 24, 40000)
    px = (1.0 - (k + 1) * 2.0 ** -24).astype(np.float32)
    py = np.full_like(px, np.float32(1.0) / np.float32(31.0))
    got = tt.debug_math(ctx, 2, px, py)
    want = np.array([oracle.lib.o_powf(float(a), float(b)) for a, b in zip(px, py)], np.float32)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))
    # atan2 / acos of the env lookup: unit directions, axes, signed zeros, tiny and huge ratios
    d = rng.standard_normal((60000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ay = np.concatenate([d[:, 0], [0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1e-30, 3e30, np.nan]]).astype(np.float32)
    ax = np.concatenate([d[:, 2], [0.0, 0.0, -0.0, -0.0, -0.0, 0.0, -1.0, 1e-20, 1.0]]).astype(np.float32)
    got = tt.debug_math(ctx, 3, ay, ax)
    assert np.array_equal(got.view(np.int32), oracle.math(3, ay, ax).view(np.int32))
    ac = np.concatenate([d[:, 1], np.linspace(-1, 1, 20001), [1.0, -1.0, 1 - 2 ** -24, -1 + 2 ** -24]]).astype(np.float32)
    got = tt.debug_math(ctx, 4, ac)
    assert np.array_equal(got.view(np.int32), oracle.math(4, ac).view(np.int32))


def _ref_case(tt, oracle, ctx, mesh, mats, cam, spp, depth, seed, env=None, env_scale=1.0):
    b = tt.BVH.build(mesh)
    s = tt.DeviceScene(ctx, mesh, b, mats, env=env, env_scale=env_scale)
    img, st = s.render(cam, tt.render_params(spp=spp, max_depth=depth, seed=seed, mode=tt.RENDER_REFERENCE))
    ref = oracle.render_ref(oracle.CompactBVH(b.export_compact()), mesh.arrays(), mats, cam, spp=spp,
                            max_depth=depth, seed=seed,
                            env=oracle.Env(env, env_scale) if env is not None else None)
    s.close()
    return img, ref, st


def test_reference_render_cornell_with_sky_matches_oracle(tt, oracle, ctx):
    """The reference's miss path, hdrSample (HDRKernel.cuh:6-42): Cornell box + icosphere
    (open at the front) under the procedural sky with its sun, 128x128, 16 spp, depth 8:
    bit-identical to the oracle's restatement of the lookup."""
    from trenchanttracer_amd import workloads
    _say("reference render cornell + sky")
    mesh, mats, cam = workloads.cornell(subdiv=3, width=128, height=128)
    env = workloads.sky_env(512, 256, sun_radiance=50.0)
    img, ref, st = _ref_case(tt, oracle, ctx, mesh, mats, cam, 16, 8, 11, env=env, env_scale=0.75)
    same = np.mean(img.view(np.int32) == ref.view(np.int32))
    _say(f"  {same * 100:.4f}% bit-identical, mean {img.mean():.4f}")
    assert same == 1.0


def test_reference_render_open_scene_in_sky_matches_oracle(tt, oracle, ctx):
    """Diffuse, Phong-metal and glass icospheres over a diffuse floor quad in the open sky
    (most paths end on the env map, many after a glass or metal bounce; directions cover
    the whole sphere, the poles and the longitude seam): bit-identical to the oracle.
    The map is 37 x 23 so texel boundaries fall at odd positions."""
    from trenchanttracer_amd import workloads
    _say("reference render open scene + sky")
    mesh = tt.Mesh.icosphere(3, 0.45, (-0.9, 0.0, 0.0), mat=0)
    mesh.append(tt.Mesh.icosphere(3, 0.45, (0.0, 0.0, 0.3), mat=1))
    mesh.append(tt.Mesh.icosphere(3, 0.45, (0.9, 0.0, 0.0), mat=2))
    mesh.append(tt.Mesh.quad((-4, -0.46, 4), (4, -0.46, 4), (-4, -0.46, -4), mat=3))
    mats = [tt.material(tt.MAT_DIFFUSE, (0.8, 0.3, 0.2)), tt.material(tt.MAT_CONDUCTOR, (0.9, 0.9, 0.7)),
            tt.material(tt.MAT_DIELECTRIC, (1, 1, 1), ior=1.4), tt.material(tt.MAT_DIFFUSE, (0.5, 0.5, 0.5))]
    cam = tt.camera((0.0, 0.6, 3.2), (0.0, -0.2, -1.0), fov=(70, 55), width=160, height=120)
    env = workloads.sky_env(37, 23, sun_radiance=20.0, sun_diam_deg=12.0)
    img, ref, st = _ref_case(tt, oracle, ctx, mesh, mats, cam, 8, 10, 4, env=env)
    same = np.mean(img.view(np.int32) == ref.view(np.int32))
    _say(f"  {same * 100:.4f}% bit-identical, mean {img.mean():.4f}")
    assert same == 1.0
    assert img.mean() > 0.1


def test_reference_render_cornell_matches_oracle(tt, oracle, ctx):
    """Cornell box + icosphere-4, 256x256, 16 spp, depth 8, seed 1 (the CPU-reference case).
    Tolerance |dL| <= 1e-3 * max(1, L) per pixel; expected bit-identical."""
    from trenchanttracer_amd import workloads
    _say("reference render cornell")
    mesh, mats, cam = workloads.cornell()
    img, ref, st = _ref_case(tt, oracle, ctx, mesh, mats, cam, 16, 8, 1)
    tol = 1e-3 * np.maximum(1.0, np.abs(ref))
    assert np.all(np.abs(img - ref) <= tol)
    same = np.mean(img.view(np.int32) == ref.view(np.int32))
    _say(f"cornell reference render: {same * 100:.3f}% of channels bit-identical, stats {st}")
    assert same == 1.0
    assert img.mean() > 0.01


@pytest.mark.parametrize("w,h", [(37, 23), (2, 2), (9, 66), (130, 7)])
def test_reference_render_ragged_sizes_match_oracle(tt, oracle, ctx, w, h):
    """Paths walk the image in 8x8 tiles (narrower or shorter tiles at the right and
    bottom edges) while each pixel keeps its own random numbers and accumulation slot:
    at sizes that are not multiples of 8 the image equals the oracle's bit for bit."""
    from trenchanttracer_amd import workloads
    _say(f"reference render ragged {w}x{h}")
    mesh, mats, cam = workloads.cornell(subdiv=2, width=w, height=h)
    img, ref, st = _ref_case(tt, oracle, ctx, mesh, mats, cam, 4, 6, 3)
    assert np.array_equal(img.view(np.int32), ref.view(np.int32))


@pytest.mark.parametrize("paths", [0, 256 * 256, 3 * 256 * 256])
def test_reference_render_any_pass_structure_matches_oracle(tt, oracle, ctx, paths):
    """Reference mode keeps each path's radiance in its slot and ends a pass by adding
    a pixel's samples in sample order (acc + L * 1/spp, as the reference accumulates):
    16 spp in one pass, in 16 passes of one sample, or in passes of 3 samples (the
    last pass partial) give the oracle's image bit for bit."""
    from trenchanttracer_amd import workloads
    _say(f"reference render pass structure, paths in flight {paths}")
    mesh, mats, cam = workloads.cornell(subdiv=3)
    b = tt.BVH.build(mesh)
    s = tt.DeviceScene(ctx, mesh, b, mats)
    img, st = s.render(cam, tt.render_params(spp=16, max_depth=8, seed=5, mode=tt.RENDER_REFERENCE,
                                             paths_in_flight=paths))
    ref = oracle.render_ref(oracle.CompactBVH(b.export_compact()), mesh.arrays(), mats, cam, spp=16, max_depth=8,
                            seed=5)
    s.close()
    _say(f"  {st}")
    assert st["paths_in_flight"] == (256 * 256 * 16 if paths == 0 else paths)
    assert np.array_equal(img.view(np.int32), ref.view(np.int32))


def test_reference_render_large_frame_matches_oracle(tt, oracle, ctx):
    """1024x640 (655,360 paths in one queue): more rays than the traversal kernel's
    grid holds at once (32 one-wave blocks per CU), so its blocks stride over the
    queue more than once. Bit-identical to the oracle at 1 spp."""
    from trenchanttracer_amd import workloads
    _say("reference render cornell 1024x640")
    mesh, mats, cam = workloads.cornell(width=1024, height=640)
    img, ref, st = _ref_case(tt, oracle, ctx, mesh, mats, cam, 1, 8, 2)
    same = np.mean(img.view(np.int32) == ref.view(np.int32))
    _say(f"cornell 1024x640 reference render: {same * 100:.4f}% of channels bit-identical, stats {st}")
    assert same == 1.0
    assert img.mean() > 0.01


def test_reference_render_metal_and_glass_match_oracle(tt, oracle, ctx):
    _say("reference render metal/glass")
    from trenchanttracer_amd import workloads
    mesh, mats, cam = workloads.cornell(subdiv=3, width=96, height=80)
    mesh.append(tt.Mesh.icosphere(3, 0.35, (-0.7, -0.65, 0.8), mat=5))
    mesh.append(tt.Mesh.icosphere(3, 0.35, (0.7, -0.65, 1.0), mat=6))
    mats = mats + [tt.material(tt.MAT_CONDUCTOR, (0.9, 0.8, 0.6)), tt.material(tt.MAT_DIELECTRIC, (1, 1, 1), ior=1.4)]
    img, ref, _ = _ref_case(tt, oracle, ctx, mesh, mats, cam, 8, 10, 7)
    assert np.array_equal(img.view(np.int32), ref.view(np.int32))


def _furnace_scene(tt, ctx, mat, env_value=1.0):
    sphere = tt.Mesh.icosphere(5, 1.0, (0, 0, 0), mat=0)
    env = np.full((16, 32, 3), env_value, np.float32)
    b = tt.BVH.build(sphere)
    return tt.DeviceScene(ctx, sphere, b, [mat], env=env)


@pytest.mark.parametrize("kind", ["diffuse", "glass", "smooth_metal"])
def test_path_mode_furnace(tt, ctx, kind):
    """Uniform environment of radiance 1 around a unit sphere: a convex diffuse sphere
    reflects albedo * 1; a clear glass sphere (absorption-free, Fresnel-split) and a
    near-smooth conductor with F0 = 1 return 1."""
    _say(f"furnace {kind}")
    if kind == "diffuse":
        mat, want = tt.material(tt.MAT_DIFFUSE, (0.5, 0.5, 0.5)), 0.5
    elif kind == "glass":
        mat, want = tt.material(tt.MAT_DIELECTRIC, (1, 1, 1), ior=1.5), 1.0
    else:
        mat, want = tt.material(tt.MAT_CONDUCTOR, (1, 1, 1), roughness=0.05), 1.0
    s = _furnace_scene(tt, ctx, mat)
    cam = tt.camera((0, 0, 3.0), (0, 0, -1), fov=(60, 60), width=64, height=64)
    img, st = s.render(cam, tt.render_params(spp=256, max_depth=64, rr_depth=8, seed=5, mode=tt.RENDER_PATH))
    yy, xx = np.mgrid[0:64, 0:64]
    r = np.hypot(xx - 31.5, yy - 31.5)
    # the sphere subtends asin(1/3) = 19.47 deg: radius tan(19.47)/tan(30) * 32 = 19.6 px
    inside = r < 15
    outside = r > 24
    assert abs(img[outside].mean() - 1.0) < 1e-5
    m = img[inside].mean()
    _say(f"furnace {kind}: sphere mean {m:.4f} (want {want}), stats {st}")
    assert abs(m - want) < 0.01 * max(want, 0.5)
    s.close()


def test_path_mode_agrees_with_reference_mode(tt, ctx):
    """Diffuse-only Cornell box: BSDF-only sampling (reference mode) and NEE + MIS +
    Russian roulette (path mode) estimate the same radiance."""
    from trenchanttracer_amd import workloads
    _say("path vs reference mean")
    mesh, mats, cam = workloads.cornell(subdiv=3, width=48, height=48)
    b = tt.BVH.build(mesh)
    s = tt.DeviceScene(ctx, mesh, b, mats)
    ref, _ = s.render(cam, tt.render_params(spp=4096, max_depth=8, seed=1, mode=tt.RENDER_REFERENCE))
    pt, st = s.render(cam, tt.render_params(spp=1024, max_depth=8, rr_depth=3, seed=2, mode=tt.RENDER_PATH,
                                            pi=tt.REF_PI))
    # 8x8 blocks: means agree within a few percent
    rb = ref.reshape(6, 8, 6, 8, 3).mean(axis=(1, 3))
    pb = pt.reshape(6, 8, 6, 8, 3).mean(axis=(1, 3))
    lit = rb.sum(-1) > 0.05
    rel = np.abs(rb - pb)[lit] / np.maximum(rb[lit], 1e-3)
    _say(f"path vs reference: global {pt.mean():.4f} vs {ref.mean():.4f}, block rel err median {np.median(rel):.4f} "
         f"max {rel.max():.4f}, stats {st}")
    assert abs(pt.mean() - ref.mean()) < 0.02 * ref.mean()
    assert np.median(rel) < 0.03
    s.close()


def test_env_cdf_matches_fp64_numpy(tt, ctx):
    _say("env cdf")
    rng = np.random.default_rng(4)
    h, w = 64, 128
    env = rng.gamma(0.5, 1.0, (h, w, 3)).astype(np.float32)
    env[10:12, 50:52] = 1e4
    sphere = tt.Mesh.icosphere(1, 0.1, (0, 0, 0))
    s = tt.DeviceScene(ctx, sphere, tt.BVH.build(sphere), [tt.material(0)], env=env)
    marg, cond, total = s.env_cdf(w, h)
    e64 = env.astype(np.float64)
    lum = 0.2126 * e64[..., 0] + 0.7152 * e64[..., 1] + 0.0722 * e64[..., 2]
    wgt = lum * np.sin(np.pi * (np.arange(h) + 0.5) / h)[:, None]
    rows = wgt.sum(1)
    assert abs(total - rows.sum()) <= 1e-12 * rows.sum()
    want_m = np.concatenate([[0], np.cumsum(rows) / rows.sum()]).astype(np.float32)
    assert np.max(np.abs(marg - want_m)) <= 1.2e-7
    want_c = np.concatenate([np.zeros((h, 1)), np.cumsum(wgt, 1) / rows[:, None]], 1).astype(np.float32)
    assert np.max(np.abs(cond - want_c)) <= 1.2e-7
    s.close()


def _ff_corner(a, b, d):
    """Form factor from a differential area to a parallel rectangle [0, a] x [0, b]
    whose corner lies on the area's normal at distance d (closed form)."""
    A, B = a / d, b / d
    sa, sb = np.sqrt(1 + A * A), np.sqrt(1 + B * B)
    return (A / sa * np.arctan(B / sa) + B / sb * np.arctan(A / sb)) / (2 * np.pi)


def _ff_rect(y0, y1, z0, z1, d):
    """Parallel rectangle [y0, y1] x [z0, z1] in coordinates centred on the foot of
    the normal, by superposition of corner rectangles (signed extents)."""
    def F(y, z):
        return np.sign(y) * np.sign(z) * _ff_corner(abs(y), abs(z), d)
    return F(y1, z1) - F(y0, z1) - F(y1, z0) + F(y0, z0)


@pytest.mark.parametrize("point", [(2.0, 1.0, -0.5), (2.0, 1.6, 0.3), (2.0, 0.4, 1.2)])
def test_direct_light_matches_form_factor(tt, ctx, point):
    """Direct lighting (path depth 2: next-event estimation + BSDF hits under MIS)
    of a diffuse wall point by a parallel rectangular light 0.98 away equals
    albedo * Le * F, F the closed-form differential-area-to-rectangle form factor.
    The camera looks at the point through a 0.02 degree field of view, so every
    pixel sees it. Coordinates reach 2, so the shadow-ray origin offset is twice
    the minimum: a shadow ray measured from the unoffset point would stop past the
    light and report it occluded."""
    from trenchanttracer_amd.workloads import _rect
    _say(f"direct light at {point}")
    wall = _rect(0, 2.0, 0, 2, -2, 2, -1, 0)               # x = 2, facing -x
    light = _rect(0, 1.02, 1.3, 1.9, -1.5, 1.5, +1, 1)     # x = 1.02, facing +x
    mesh = wall.append(light)
    rho, le = 0.7, 60.0
    mats = [tt.material(tt.MAT_DIFFUSE, (rho,) * 3), tt.material(tt.MAT_DIFFUSE, (0, 0, 0), emission=(le,) * 3)]
    sc = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats)
    P = np.array(point)
    cam = tt.camera(P - [0.5, 0, 0], (1, 0, 0), fov=(0.02, 0.02), width=16, height=16)
    img, st = sc.render(cam, tt.render_params(spp=2048, max_depth=2, rr_depth=8, seed=3, mode=tt.RENDER_PATH))
    want = rho * le * _ff_rect(1.3 - P[1], 1.9 - P[1], -1.5 - P[2], 1.5 - P[2], 2.0 - 1.02)
    got = img.mean()
    _say(f"  radiance {got:.5f}, closed form {want:.5f}")
    assert abs(got - want) < 0.01 * want


def test_reference_render_analytic_cornell_matches_oracle(tt, oracle, ctx):
    """The reference's Cornell configuration as it renders it (walls and light as
    analytic planes, spheres with mirror / coat / metal / glass materials, a mesh):
    the GPU reference mode is bit-identical to the CPU restatement."""
    from trenchanttracer_amd import workloads
    _say("reference render, analytic cornell")
    mesh, mats, geoms, cam = workloads.cornell_analytic(112, 96)
    b = tt.BVH.build(mesh)
    s = tt.DeviceScene(ctx, mesh, b, mats)
    s.set_geometries(geoms)
    img, st = s.render(cam, tt.render_params(spp=8, max_depth=10, seed=5, mode=tt.RENDER_REFERENCE))
    ref = oracle.render_ref(oracle.CompactBVH(b.export_compact()), mesh.arrays(), mats, cam, spp=8, max_depth=10,
                            seed=5, geoms=geoms)
    same = np.mean(img.view(np.int32) == ref.view(np.int32))
    _say(f"  {same * 100:.3f}% of channels bit-identical; mean {img.mean():.4f}")
    assert np.all(np.abs(img - ref) <= 1e-3 * np.maximum(1.0, np.abs(ref)))
    assert same == 1.0
    assert img.mean() > 0.05
    s.close()


def _far_mesh(tt):
    return tt.Mesh.quad([100, 100, 100], [100.001, 100, 100], [100, 100.001, 100], mat=0)


@pytest.mark.parametrize("kind", ["mirror", "coat", "diffuse"])
def test_path_mode_furnace_analytic_sphere(tt, ctx, kind):
    """Analytic unit sphere in a uniform environment of radiance 1 (path mode): a
    mirror of colour 1 and a coat whose two colours are 1 return 1 (what the coat
    does not reflect enters the Lambert layer), a diffuse sphere of albedo 0.5
    returns 0.5."""
    _say(f"analytic furnace {kind}")
    mat = {"mirror": tt.material(tt.MAT_SPECULAR, (1, 1, 1)),
           "coat": tt.material(tt.MAT_COAT, (1, 1, 1), specular=(1, 1, 1)),
           "diffuse": tt.material(tt.MAT_DIFFUSE, (0.5, 0.5, 0.5))}[kind]
    want = 0.5 if kind == "diffuse" else 1.0
    mesh = _far_mesh(tt)
    s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), [tt.material(tt.MAT_DIFFUSE, (0, 0, 0)), mat],
                       env=np.ones((16, 32, 3), np.float32))
    s.set_geometries([tt.sphere((0, 0, 0), 1.0, 1)])
    cam = tt.camera((0, 0, 3), (0, 0, -1), fov=(40, 40), width=64, height=64)
    img, st = s.render(cam, tt.render_params(spp=256, max_depth=16, rr_depth=8, seed=2, mode=tt.RENDER_PATH))
    lum = img.mean(2)
    yy, xx = np.mgrid[0:64, 0:64]
    inside = (yy - 31.5) ** 2 + (xx - 31.5) ** 2 < 14 ** 2
    _say(f"  inside mean {lum[inside].mean():.5f} (want {want}), outside {lum[~inside].mean():.5f}")
    assert abs(lum[inside].mean() - want) < 0.01 * want
    s.close()


def test_path_mode_emissive_box_of_planes(tt, ctx):
    """A closed box of six analytic planes facing inwards (path-mode emitters emit on
    their front side), albedo 0.5, emission 0.5; analytic emitters are reached by
    BSDF sampling only: radiance 1 everywhere."""
    _say("analytic emissive box")
    mesh = _far_mesh(tt)
    mats = [tt.material(tt.MAT_DIFFUSE, (0, 0, 0)),
            tt.material(tt.MAT_DIFFUSE, (0.5, 0.5, 0.5), emission=(0.5, 0.5, 0.5))]
    s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats)
    geoms = []
    for ax in range(3):
        for c in (-1.0, 1.0):
            u, v = (ax + 1) % 3, (ax + 2) % 3
            p0 = [0.0, 0.0, 0.0]; p0[ax], p0[u], p0[v] = c, -1.0, -1.0
            p1 = list(p0); p1[u] = 1.0
            p2 = list(p0); p2[v] = 1.0
            if c > 0:      # normal (p1 - p0) x (p2 - p0) = +axis: swap to face inwards
                p1, p2 = p2, p1
            geoms.append(tt.plane(p0, p1, p2, 1))
    s.set_geometries(geoms)
    cam = tt.camera((0.1, 0.2, 0.3), (0.3, -0.2, -1.0), fov=(90, 90), width=64, height=64)
    img, st = s.render(cam, tt.render_params(spp=256, max_depth=40, rr_depth=4, seed=2, mode=tt.RENDER_PATH))
    _say(f"  mean {img.mean():.5f}")
    assert abs(img.mean() - 1.0) < 0.01
    s.close()


def test_path_mode_is_deterministic(tt, ctx):
    """Path-mode radiance is summed per path and resolved per pixel in sample order
    (no floating-point atomics): two renders with one seed are bit-identical, in
    float and in FP64 accumulation, and so are one-pass and multi-pass renders'
    sample sets (the same paths, only summed in different groupings)."""
    from trenchanttracer_amd import workloads
    _say("path-mode determinism")
    mesh, mats, cam = workloads.cornell(subdiv=3, width=64, height=48)
    s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats)
    p = tt.render_params(spp=16, max_depth=8, rr_depth=3, seed=4, mode=tt.RENDER_PATH)
    a, _ = s.render(cam, p)
    b, _ = s.render(cam, p)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    p64 = tt.render_params(spp=16, max_depth=8, rr_depth=3, seed=4, mode=tt.RENDER_PATH, fp64=True)
    _, a64, _ = s.render(cam, p64, want_fp64=True)
    _, b64, _ = s.render(cam, p64, want_fp64=True)
    assert np.array_equal(a64, b64)
    p4 = tt.render_params(spp=16, max_depth=8, rr_depth=3, seed=4, mode=tt.RENDER_PATH, fp64=True,
                          paths_in_flight=64 * 48 * 4)
    _, c64, _ = s.render(cam, p4, want_fp64=True)
    assert np.allclose(c64, a64, rtol=1e-12, atol=1e-12)
    s.close()


def test_path_mode_large_queues_match_small_passes(tt, ctx):
    """256x256 at 32 spp in one pass is 2,097,152 paths: the queue kernels' blocks
    stride over each bounce's queue several times (traversal grid 32 one-wave blocks
    per CU, shade grid 16 blocks of 256 per CU). The same samples rendered as 32
    passes of 65,536 paths (one stride each) give the same image bit for bit: each
    pixel's sum takes its samples in sample order across passes."""
    from trenchanttracer_amd import workloads
    _say("path-mode large queues")
    mesh, mats, cam = workloads.cornell(subdiv=3, width=256, height=256)
    s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats)
    big = tt.render_params(spp=32, max_depth=8, rr_depth=3, seed=9, mode=tt.RENDER_PATH, fp64=True)
    small = tt.render_params(spp=32, max_depth=8, rr_depth=3, seed=9, mode=tt.RENDER_PATH, fp64=True,
                             paths_in_flight=256 * 256)
    _, a64, sa = s.render(cam, big, want_fp64=True)
    _, b64, sb = s.render(cam, small, want_fp64=True)
    _say(f"one pass: {sa}; 32 passes: {sb}")
    assert sa["paths_in_flight"] == 256 * 256 * 32 and sb["paths_in_flight"] == 256 * 256
    assert sa["rays_extension"] == sb["rays_extension"] and sa["rays_shadow"] == sb["rays_shadow"]
    assert np.array_equal(a64, b64)
    assert a64.mean() > 0.01
    fa, _ = s.render(cam, tt.render_params(spp=32, max_depth=8, rr_depth=3, seed=9, mode=tt.RENDER_PATH))
    fb, _ = s.render(cam, tt.render_params(spp=32, max_depth=8, rr_depth=3, seed=9, mode=tt.RENDER_PATH,
                                           paths_in_flight=3 * 256 * 256))
    assert np.array_equal(fa.view(np.int32), fb.view(np.int32))
    s.close()


def test_pathtrace_case_full_size_invariants(tt, ctx):
    """The path-trace case at its full size (1,310,720-tri displaced icosphere, 2048x1024
    sky + sun, 1920x1080, 64 spp, depth 16: 132,710,400 paths in one pass). Each path's
    arithmetic does not depend on how the traversal is launched, so the image and the
    ray counts are bit-identical for the exact and the speculative loop, for other
    block sizes, refill thresholds and LDS stack depths, and for a pass replayed from
    a hipGraph; FP64 accumulation of the same samples agrees with the float sums to
    float rounding. The pixel tile the path slots walk (16 x 16, pixel rows) changes
    only the order of the paths, not the image."""
    from trenchanttracer_amd import workloads
    _say("path-trace case full size")
    mesh, mats, env, cam = workloads.pathtrace()
    s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats, env=env)
    base = dict(spp=64, max_depth=16, rr_depth=4, seed=1, mode=tt.RENDER_PATH)
    img, st = s.render(cam, tt.render_params(**base))
    _say(f"full size: {st}")
    assert st["paths_in_flight"] == 1920 * 1080 * 64
    assert st["rays_extension"] > 1920 * 1080 * 64 and st["rays_shadow"] > 0
    assert np.isfinite(img).all() and 0.1 < img.mean() < 10.0
    for extra in (dict(variant=0), dict(block=256, refill=8), dict(lds_stack=8, refill=32, refill_camera=16),
                  dict(use_graph=True), dict(tile=16), dict(tile=1)):
        other, so = s.render(cam, tt.render_params(**base, **extra))
        _say(f"{extra}: {so['seconds']:.4f} s")
        assert so["rays_extension"] == st["rays_extension"] and so["rays_shadow"] == st["rays_shadow"], extra
        assert np.array_equal(other.view(np.int32), img.view(np.int32)), extra
    _, img64, _ = s.render(cam, tt.render_params(**base, fp64=True), want_fp64=True)
    assert np.allclose(img64, img, rtol=1e-5, atol=1e-6)
    s.close()
