"""Synthetic workloads (the reference ships no scene assets).

traversal(): the traversal microbenchmark — a triangle soup (centroids uniform in
the unit cube, equilateral triangles with log-normal side, mu = ln 0.01,
sigma = 0.5, uniformly random orientation), SBVH with alpha = 1e-5, and a ray set
of coherent primary rays (pinhole at (0.5, 0.5, -1) looking +z, fovx 45 deg,
1920x1080 x 2 jittered samples, tmin 0) followed by the same number of
incoherent cosine-hemisphere rays from random surface points (tmin 1e-4).
"""
import math
import time

import numpy as np

from . import BVH, Mesh, rays_incoherent, rays_primary

TRAVERSAL_FULL = dict(ntris=1_000_000, width=1920, height=1080, spp=2)


def traversal(ntris=1_000_000, width=1920, height=1080, spp=2, geo_seed=1, ray_seed=2, split_alpha=1e-5,
              nthreads=0, log=None):
    t0 = time.time()
    mesh = Mesh.soup(ntris, seed=geo_seed, mu=math.log(0.01), sigma=0.5)
    t1 = time.time()
    bvh = BVH.build(mesh, split_alpha=split_alpha, nthreads=nthreads)
    t2 = time.time()
    prim = rays_primary(width, height, spp, (0.5, 0.5, -1.0), (0.5, 0.5, 0.0), fovx=45.0, tmin=0.0, tmax=1e20,
                        seed=ray_seed)
    inc = rays_incoherent(mesh, len(prim), tmin=1e-4, tmax=1e20, seed=ray_seed)
    rays = np.concatenate([prim, inc])
    t3 = time.time()
    if log:
        log(f"workload: {ntris} tris generated in {t1 - t0:.1f}s, SBVH in {t2 - t1:.1f}s "
            f"({bvh.stats()['num_refs']} refs), {len(rays)} rays in {t3 - t2:.1f}s")
    return mesh, bvh, rays, len(prim)


def bytes_per_ray(inner, tris, nrays):
    """Algorithmic bytes of the device layout: ray (32) + hit (16) + 64 per node
    visit + 48 per triangle test, with visits counted by the single-lane traversal."""
    return 32.0 + 16.0 + (64.0 * inner + 48.0 * tris) / max(1, nrays)


# Cornell box of the reference config (config/cornell_dragon.json:19-62): back,
# left, right, ceiling and floor quads plus the ceiling light quad, each as two
# triangles (p0, p1, p2), (p1, p1 + p2 - p0, p2); every wall faces the interior.
CORNELL_QUADS = [
    ("back", [-1.5, -1, -1], [1.5, -1, -1], [-1.5, 1, -1]),
    ("left", [-1.5, -1, -1], [-1.5, 1, -1], [-1.5, -1, 4]),
    ("right", [1.5, -1, -1], [1.5, -1, 4], [1.5, 1, -1]),
    ("ceiling", [-1.5, 1, -1], [1.5, 1, -1], [-1.5, 1, 4]),
    ("floor", [-30, -1, -30], [-30, -1, 30], [30, -1, -30]),
    ("light", [-0.3, 0.9999, 1.2], [0.3, 0.9999, 1.2], [-0.3, 0.9999, 1.8]),
]


def cornell(subdiv=4, sphere_radius=0.3, sphere_center=(0.0, -0.7, 1.5), width=256, height=256):
    """Cornell box + icosphere (the CPU-reference render case).

    Materials: 0 grey diffuse 0.75, 1 red (0.75, 0.25, 0.25), 2 green (0.25, 0.75, 0.25),
    3 light (diffuse albedo 0, emission 15), 4 sphere diffuse 0.8. Camera: the orbit
    camera of the config (centre (0, 0.2, 3), yaw 0, pitch 0.1, radius 4, focal 1) with
    fovx 45 and no aperture. Returns (mesh, materials, camera).
    """
    from . import MAT_DIFFUSE, camera_from_orbit, material
    mat_of = {"back": 0, "left": 1, "right": 2, "ceiling": 0, "floor": 0, "light": 3}
    mesh = None
    for name, p0, p1, p2 in CORNELL_QUADS:
        q = Mesh.quad(p0, p1, p2, mat=mat_of[name])
        mesh = q if mesh is None else mesh.append(q)
    mesh.append(Mesh.icosphere(subdiv, sphere_radius, sphere_center, mat=4))
    mats = [material(MAT_DIFFUSE, (0.75, 0.75, 0.75)),
            material(MAT_DIFFUSE, (0.75, 0.25, 0.25)),
            material(MAT_DIFFUSE, (0.25, 0.75, 0.25)),
            material(MAT_DIFFUSE, (0.0, 0.0, 0.0), emission=(15.0, 15.0, 15.0)),
            material(MAT_DIFFUSE, (0.8, 0.8, 0.8))]
    cam = camera_from_orbit((0.0, 0.2, 3.0), 0.0, 0.1, radius=4.0, fovx=45.0, width=width, height=height,
                            focal=1.0, aperture=0.0)
    return mesh, mats, cam


def sky_env(width=2048, height=1024, sun_elev_deg=35.0, sun_azim_deg=40.0, sun_diam_deg=0.5, sun_radiance=1e4):
    """Procedural HDR environment (equirectangular, row 0 = +y): sky gradient from a
    pale horizon (radiance ~1) to a blue zenith, dim ground below the horizon, and a
    sun disk of the given angular diameter and radiance. Direction of texel (i, j):
    theta = pi (j + 0.5) / H from +y, phi = 2 pi (i + 0.5) / W measured from +z
    towards +x (the map's u = atan2(d.x, d.z) / 2 pi)."""
    j = (np.arange(height) + 0.5) / height
    i = (np.arange(width) + 0.5) / width
    theta = np.pi * j[:, None]
    phi = 2 * np.pi * i[None, :]
    d = np.stack([np.sin(theta) * np.sin(phi), np.cos(theta) * np.ones_like(phi), np.sin(theta) * np.cos(phi)], -1)
    up = np.clip(d[..., 1], -1, 1)
    horizon = np.array([1.0, 0.95, 0.85])
    zenith = np.array([0.25, 0.45, 0.95])
    t = np.clip(up, 0, 1)[..., None] ** 0.5
    sky = horizon * (1 - t) + zenith * t
    ground = np.array([0.18, 0.16, 0.14]) * np.ones_like(sky)
    env = np.where((up >= 0)[..., None], sky, ground)
    el, az = np.radians(sun_elev_deg), np.radians(sun_azim_deg)
    sdir = np.array([np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)])
    cosang = d @ sdir
    env[cosang >= np.cos(np.radians(sun_diam_deg) / 2)] = sun_radiance
    return env.astype(np.float32)


def pathtrace(subdiv=8, width=1920, height=1080, env_w=2048, env_h=1024, seed=1):
    """Full path-trace case: icosphere subdivided `subdiv` times (1,310,720 tris at 8)
    with seeded fBm radial displacement (6 octaves, amplitude 0.05); per-face
    materials 40% diffuse / 40% GGX conductor (each face its own roughness drawn from
    U[0.05, 0.5], gold or silver F0 alternating) / 20% dielectric IOR 1.5; the
    procedural sky; pinhole camera. Returns (mesh, materials, env, camera); the
    materials are a ctypes array of tt_material (one per conductor face plus the
    diffuse and dielectric ones)."""
    from . import camera
    from ._native import tt_material
    mesh = Mesh.icosphere(subdiv, 1.0, (0.0, 0.0, 0.0), amp=0.05, octaves=6, seed=seed, mat=0)
    mesh.assign_materials([0.4, 0.4, 0.2], [0, 1, 2], seed=seed)
    a = mesh.arrays()
    tm = a["tri_mats"]
    cond = np.flatnonzero(tm == 1)
    k = len(cond)
    rng = np.random.default_rng(seed + 7)
    rec = np.zeros(k + 2, dtype=[("type", "<i4"), ("albedo", "<f4", 3), ("emission", "<f4", 3),
                                 ("roughness", "<f4"), ("ior", "<f4"), ("specular", "<f4", 3)])
    assert rec.dtype.itemsize == 48
    rec["ior"] = 1.5
    rec["specular"] = 1.0
    rec[0]["type"], rec[0]["albedo"] = 0, (0.8, 0.7, 0.6)
    rec[1:k + 1]["type"] = 1
    gold = (np.arange(k) % 2 == 1)[:, None]
    rec["albedo"][1:k + 1] = np.where(gold, np.float32([0.95, 0.78, 0.4]), np.float32([0.92, 0.92, 0.9]))
    rec["roughness"][1:k + 1] = rng.uniform(0.05, 0.5, k).astype(np.float32)
    rec[k + 1]["type"], rec[k + 1]["albedo"] = 2, (1.0, 1.0, 1.0)
    new_tm = np.where(tm == 0, 0, k + 1).astype(np.int32)
    new_tm[cond] = 1 + np.arange(k, dtype=np.int32)
    mesh = Mesh.from_arrays(a["verts"], a["tris"], a["normals"], a["tri_normals"], new_tm)
    mats = (tt_material * (k + 2)).from_buffer_copy(rec.tobytes())
    env = sky_env(env_w, env_h)
    cam = camera((0.0, 0.35, 3.3), (0.0, -0.1, -1.0), fov=(40.0, 40.0 * height / width), width=width,
                 height=height, focal=1.0)
    return mesh, mats, env, cam


def batched(n=32, seed=1, tri_lo=1_000, tri_hi=2_000_000, resolutions=(256, 512, 1024), log=None):
    """The batched case: n seeded scenes, alternately displaced icospheres and
    triangle soups, triangle counts log-uniform in [tri_lo, tri_hi] (icospheres take
    the subdivision level nearest in log), square resolutions drawn from
    `resolutions`, path-mode materials, one shared sky. Returns
    (meshes, materials, cameras, env, tri_counts)."""
    from . import MAT_CONDUCTOR, MAT_DIELECTRIC, MAT_DIFFUSE, camera, material
    rng = np.random.default_rng(seed)
    meshes, mats, cams, counts = [], [], [], []
    for k in range(n):
        target = float(np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi))))
        res = int(rng.choice(resolutions))
        if k % 2 == 0:
            sub = int(np.clip(np.round(np.log(target / 20.0) / np.log(4.0)), 2, 8))
            m = Mesh.icosphere(sub, 1.0, (0, 0, 0), amp=0.05, octaves=6, seed=seed * 100 + k)
            mm = [material(MAT_DIFFUSE, (0.8, 0.7, 0.6)),
                  material(MAT_CONDUCTOR, (0.95, 0.78, 0.4), roughness=float(rng.uniform(0.05, 0.5))),
                  material(MAT_DIELECTRIC, (1, 1, 1), ior=1.5)]
            m.assign_materials([0.4, 0.4, 0.2], [0, 1, 2], seed=seed * 100 + k)
            cam = camera((0.0, 0.35, 3.3), (0.0, -0.1, -1.0), fov=(40.0, 40.0), width=res, height=res)
        else:
            m = Mesh.soup(int(target), seed=seed * 100 + k, mu=float(np.log(0.01 * (1e6 / target) ** (1 / 3))),
                          sigma=0.5)
            mm = [material(MAT_DIFFUSE, (0.75, 0.75, 0.75)), material(MAT_CONDUCTOR, (0.9, 0.9, 0.9), roughness=0.3)]
            m.assign_materials([0.7, 0.3], [0, 1], seed=seed * 100 + k)
            cam = camera((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), fov=(50.0, 50.0), width=res, height=res)
        meshes.append(m)
        mats.append(mm)
        cams.append(cam)
        counts.append(m.ntris)
        if log:
            log(f"scene {k}: {m.ntris} tris, {res}x{res}")
    env = sky_env(1024, 512)
    return meshes, mats, cams, env, counts


def _rect(axis, c, a0, a1, b0, b1, sign, mat):
    """Axis-aligned rectangle in the plane coord[axis] = c spanning [a0, a1] x [b0, b1]
    along the next two axes in cyclic order, geometric normal sign * e_axis."""
    u, v = (axis + 1) % 3, (axis + 2) % 3

    def pt(a, b):
        p = [0.0, 0.0, 0.0]
        p[axis], p[u], p[v] = c, a, b
        return p

    p0, p1, p2 = pt(a0, b0), pt(a1, b0), pt(a0, b1)
    if sign < 0:
        p1, p2 = p2, p1
    return Mesh.quad(p0, p1, p2, mat=mat)


def door_room(width=256, height=256, emission=60.0):
    """Indirect-lighting case for guiding: a closed room x in [-2, 2], y in [0, 2],
    z in [-2, 2] (walls facing inwards) split by a partition at x = 1 with a door
    (|z| < 0.2, y < 1.2). The only light is a panel just behind the partition
    (x = 1.02, y in [1.3, 1.9], z in [-1.5, 1.5]) facing +x, away from the door:
    no light passes the door directly, so the camera's part of the room sees only
    what the lit far wall of the light's part reflects through the door.
    Materials: 0 walls grey diffuse 0.7, 1 light (albedo 0, the given emission),
    2 partition diffuse 0.6. Returns (mesh, materials, camera)."""
    from . import MAT_DIFFUSE, camera, material
    parts = [
        _rect(1, 0.0, -2, 2, -2, 2, +1, 0),    # floor (y = 0, up): axes z, x
        _rect(1, 2.0, -2, 2, -2, 2, -1, 0),    # ceiling
        _rect(0, -2.0, 0, 2, -2, 2, +1, 0),    # x = -2 wall: axes y, z
        _rect(0, 2.0, 0, 2, -2, 2, -1, 0),     # x = +2 wall
        _rect(2, -2.0, -2, 2, 0, 2, +1, 0),    # z = -2 wall: axes x, y
        _rect(2, 2.0, -2, 2, 0, 2, -1, 0),     # z = +2 wall
        _rect(0, 1.0, 0, 2, -2, -0.2, -1, 2),  # partition, left of the door
        _rect(0, 1.0, 0, 2, 0.2, 2, -1, 2),    # partition, right of the door
        _rect(0, 1.0, 1.2, 2, -0.2, 0.2, -1, 2),  # lintel above the door
        _rect(0, 1.02, 1.3, 1.9, -1.5, 1.5, +1, 1),  # light facing +x: axes y, z
    ]
    mesh = parts[0]
    for q in parts[1:]:
        mesh.append(q)
    mats = [material(MAT_DIFFUSE, (0.7, 0.7, 0.7)),
            material(MAT_DIFFUSE, (0.0, 0.0, 0.0), emission=(emission, emission, emission)),
            material(MAT_DIFFUSE, (0.6, 0.6, 0.6))]
    cam = camera((-1.7, 1.0, 1.7), (2.7, -0.6, -2.0), fov=(70.0, 70.0 * height / width), width=width, height=height)
    return mesh, mats, cam


def emissive_box(albedo=0.5, emission=0.5, width=64, height=64):
    """Closed box [-1, 1]^3, walls facing inwards, every wall diffuse with the given
    albedo and emission: the radiance everywhere inside is emission / (1 - albedo)
    (up to the path-depth truncation). Returns (mesh, materials, camera)."""
    from . import MAT_DIFFUSE, camera, material
    parts = [_rect(a, s * -1.0, -1, 1, -1, 1, s, 0) for a in range(3) for s in (+1, -1)]
    mesh = parts[0]
    for q in parts[1:]:
        mesh.append(q)
    mats = [material(MAT_DIFFUSE, (albedo,) * 3, emission=(emission,) * 3)]
    cam = camera((0.1, 0.2, 0.3), (0.3, -0.2, -1.0), fov=(90.0, 90.0), width=width, height=height)
    return mesh, mats, cam


def cornell_analytic(width=128, height=128, subdiv=3):
    """The reference's own Cornell configuration (config/cornell_dragon.json:19-62) as
    it renders it: the five walls and the light as analytic planes (both sides), a
    mesh in the box (an icosphere of `subdiv`, standing in for data/dragon.obj,
    which does not ship) and four analytic spheres with the remaining materials:
    mirror, coat, Phong metal and glass. Camera as in the config (orbit centre
    (0, 0.2, 3), pitch 0.1, radius 4, focal 1, no aperture). Materials: 0 grey
    diffuse 0.75, 1 light (albedo 0, emission 15), 2 sphere diffuse 0.8, 3 mirror
    0.95, 4 coat (diffuse (0.6, 0.2, 0.2), mirror 0.9), 5 metal (0.9, 0.8, 0.6),
    6 glass IOR 1.4. Returns (mesh, materials, geometries, camera)."""
    from . import (MAT_COAT, MAT_CONDUCTOR, MAT_DIELECTRIC, MAT_DIFFUSE, MAT_SPECULAR, camera_from_orbit, material,
                   plane, sphere)
    geoms = [plane(p0, p1, p2, 1 if name == "light" else 0) for name, p0, p1, p2 in CORNELL_QUADS]
    geoms += [sphere((-0.8, -0.7, 0.6), 0.3, 3), sphere((0.8, -0.7, 0.7), 0.3, 4),
              sphere((-0.6, -0.75, 2.0), 0.25, 5), sphere((0.6, -0.75, 2.1), 0.25, 6)]
    mesh = Mesh.icosphere(subdiv, 0.3, (0.0, -0.7, 1.3), mat=2)
    mats = [material(MAT_DIFFUSE, (0.75, 0.75, 0.75)),
            material(MAT_DIFFUSE, (0.0, 0.0, 0.0), emission=(15.0, 15.0, 15.0)),
            material(MAT_DIFFUSE, (0.8, 0.8, 0.8)),
            material(MAT_SPECULAR, (0.95, 0.95, 0.95)),
            material(MAT_COAT, (0.6, 0.2, 0.2), specular=(0.9, 0.9, 0.9)),
            material(MAT_CONDUCTOR, (0.9, 0.8, 0.6)),
            material(MAT_DIELECTRIC, (1.0, 1.0, 1.0), ior=1.4)]
    cam = camera_from_orbit((0.0, 0.2, 3.0), 0.0, 0.1, radius=4.0, fovx=45.0, width=width, height=height,
                            focal=1.0, aperture=0.0)
    return mesh, mats, geoms, cam
