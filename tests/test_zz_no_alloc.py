"""tt_trace_closest_noalloc (reorder workspace in the library's fixed device array)
against tt_trace_closest with its context-owned workspace: bit-identical hit records
on two calls in a row, the second smaller, then at the size limit; a call above the
limit is refused before any device work."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 22


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _trace(tt, ctx, scene, rays, noalloc, cfg):
    n = len(rays)
    rb = tt.DeviceBuffer.from_array(ctx, np.ascontiguousarray(rays, np.float32))
    hb = tt.DeviceBuffer(ctx, n * 16)
    if noalloc:
        scene.trace_closest_noalloc(rb, n, hb, cfg)
    else:
        scene.trace_closest(rb, n, hb, cfg)
    ctx.sync()
    out = hb.download((n, 4), np.float32)
    rb.free()
    hb.free()
    return out


def test_noalloc_trace_equals_workspace_trace(tt, ctx):
    mesh = tt.Mesh.soup(60000, seed=7)
    bvh = tt.BVH.build(mesh)
    scene = tt.DeviceScene(ctx, mesh, bvh)
    rays = np.concatenate([tt.rays_primary(320, 240, 2, (0.5, 0.5, -1.0), (0.5, 0.5, 0.0)),
                           tt.rays_incoherent(mesh, 150000, seed=3)])
    rng = np.random.default_rng(1)
    rays = rays[rng.permutation(len(rays))]
    for cfg in (tt.trace_cfg(1, 16, 256, refill=32, reorder=1), tt.trace_cfg(0, 16, 128, reorder=1)):
        for n in (len(rays), 41835):
            _say(f"noalloc trace: {n} rays, variant {cfg.variant}, block {cfg.block}")
            want = _trace(tt, ctx, scene, rays[:n], False, cfg)
            got = _trace(tt, ctx, scene, rays[:n], True, cfg)
            assert np.array_equal(got.view(np.int32), want.view(np.int32))
            assert (got[:, 0].view(np.int32) >= 0).mean() > 0.3
    # the size limit itself
    big = tt.rays_incoherent(mesh, MAX_RAYS, seed=5)
    cfg = tt.trace_cfg(1, 16, 256, refill=32, reorder=1)
    _say(f"noalloc trace: {MAX_RAYS} rays (the limit)")
    want = _trace(tt, ctx, scene, big, False, cfg)
    got = _trace(tt, ctx, scene, big, True, cfg)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))
    # above the limit: refused on the host, nothing launched
    rb = tt.DeviceBuffer(ctx, 32)
    hb = tt.DeviceBuffer(ctx, 16)
    with pytest.raises(tt.TTError, match="at most 4194304 rays"):
        scene.trace_closest_noalloc(rb, MAX_RAYS + 1, hb, cfg)
    rb.free()
    hb.free()
    scene.close()
