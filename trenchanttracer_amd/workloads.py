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
