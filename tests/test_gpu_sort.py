"""The ray reorder's radix sort (tt_debug_ray_order, the permutation a reordering trace
call traces in) against numpy: the stable argsort of the 24-bit Morton keys of the ray
origins over the scene's bounds, computed here in float32 with the same operations.
Sizes around the 4096-key tile (1, 63, 4095, 4096, 4097), ragged large ones, all keys
equal (camera rays: the identity), origins outside the box, infinite and NaN origins,
and keys whose digits collide in every pass; each with the length given on the host and
as a queue of larger capacity whose length is read on the device (the path tracer's
form), where the tiles past the length are skipped."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, file=sys.stderr, flush=True)


def _spread3(v):
    v = v.astype(np.uint32) & 0xFF
    v = (v | (v << 8)) & 0x0000F00F
    v = (v | (v << 4)) & 0x000C30C3
    v = (v | (v << 2)) & 0x00249249
    return v


def morton_keys(rays, lo, hi):
    f = np.float32
    ext = (hi - lo).astype(f)
    scale = np.where(ext > 0, f(256.0) / np.where(ext > 0, ext, f(1)), f(0)).astype(f)
    o = rays[:, :3].astype(f)
    with np.errstate(invalid="ignore", over="ignore"):
        c = np.fmin(np.fmax((o - lo) * scale, f(0)), f(255)).astype(np.uint32)
    return (_spread3(c[:, 0]) << 2) | (_spread3(c[:, 1]) << 1) | _spread3(c[:, 2])


@pytest.fixture(scope="module")
def soup(tt, ctx):
    mesh = tt.Mesh.soup(5000, seed=3)
    s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh))
    a = mesh.arrays()
    v = a["verts"][a["tris"]].reshape(-1, 3).astype(np.float32)
    yield tt, s, mesh, v.min(0), v.max(0)
    s.close()


def _check(tt, ctx, s, rays, lo, hi, what):
    d = tt.DeviceBuffer.from_array(ctx, np.ascontiguousarray(rays, np.float32))
    want = np.argsort(morton_keys(rays, lo, hi), kind="stable").astype(np.int32)
    # the length from the host, then as a queue whose length is read on the device
    # (capacities past the last tile and inside it)
    for cap in (0, len(rays) + 1, len(rays) + 9000):
        _say(f"ray order: {what}, {len(rays)} rays, capacity {cap}")
        got = s.ray_order(d, len(rays), cap)
        assert np.array_equal(got, want), f"{what}, cap {cap}: {np.count_nonzero(got != want)} positions differ"
    d.free()


@pytest.mark.parametrize("n", [1, 63, 4095, 4096, 4097, 100_003, 1_000_003])
def test_ray_order_is_stable_morton_sort(soup, ctx, n):
    tt, s, mesh, lo, hi = soup
    k = n // 2
    rays = np.concatenate([tt.rays_primary(max(1, k), 1, 1, (0.5, 0.5, -1.0), (0.5, 0.5, 0.0))[:k],
                           tt.rays_incoherent(mesh, n - k, seed=n)])
    _check(tt, ctx, s, rays, lo, hi, "camera + incoherent")


def test_ray_order_equal_keys_is_identity(soup, ctx):
    tt, s, mesh, lo, hi = soup
    rays = tt.rays_primary(640, 480, 1, (0.5, 0.5, -1.0), (0.5, 0.5, 0.0))
    _check(tt, ctx, s, rays, lo, hi, "one origin")
    d = tt.DeviceBuffer.from_array(ctx, rays)
    assert np.array_equal(s.ray_order(d, len(rays)), np.arange(len(rays), dtype=np.int32))
    assert np.array_equal(s.ray_order(d, len(rays), 3 * len(rays)), np.arange(len(rays), dtype=np.int32))
    d.free()


def test_ray_order_odd_origins(soup, ctx):
    tt, s, mesh, lo, hi = soup
    rng = np.random.default_rng(5)
    n = 300_000
    rays = tt.rays_incoherent(mesh, n, seed=11)
    o = rays[:, :3]
    o[: n // 4] = rng.uniform(-5, 6, (n // 4, 3))            # far outside the box on every side
    o[n // 4: n // 4 + 1000] = np.inf
    o[n // 4 + 1000: n // 4 + 2000] = -np.inf
    o[n // 4 + 2000: n // 4 + 3000, 1] = np.nan
    # a few cells only: long runs of equal digits in every pass
    o[n // 2:] = lo + (hi - lo) * rng.integers(0, 4, (n - n // 2, 3)).astype(np.float32) / np.float32(4)
    _check(tt, ctx, s, rays, lo, hi, "outside / inf / NaN / clustered origins")
