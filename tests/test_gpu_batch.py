"""Batched rendering and hipGraph replay: same results as single-scene eager renders."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_graph_replays_equal_eager(tt, ctx):
    """One sample pass captured in a hipGraph and replayed spp times, three renders
    with eager renders in between: every replayed render equals the eager one
    (reference mode: bit-identical)."""
    from trenchanttracer_amd import workloads
    _say("graph vs eager")
    mesh, mats, cam = workloads.cornell(subdiv=3, width=64, height=48)
    s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats)
    eager, _ = s.render(cam, tt.render_params(spp=8, max_depth=8, seed=3))
    for rep in range(3):
        g, st = s.render(cam, tt.render_params(spp=8, max_depth=8, seed=3, use_graph=True))
        assert np.array_equal(g.view(np.int32), eager.view(np.int32)), rep
        e2, _ = s.render(cam, tt.render_params(spp=8, max_depth=8, seed=3))
        assert np.array_equal(e2.view(np.int32), eager.view(np.int32))
    # path mode, FP64 sums: equal up to the order of the atomic additions
    pe, pe64, _ = s.render(cam, tt.render_params(spp=16, max_depth=8, seed=4, mode=tt.RENDER_PATH, fp64=True,
                                                  paths_in_flight=64 * 48 * 4), want_fp64=True)
    for rep in range(2):
        pg, pg64, st = s.render(cam, tt.render_params(spp=16, max_depth=8, seed=4, mode=tt.RENDER_PATH, fp64=True,
                                                      paths_in_flight=64 * 48 * 4, use_graph=True), want_fp64=True)
        assert np.allclose(pg64, pe64, rtol=1e-12, atol=1e-12)
    s.close()


def test_batch_equals_individual_renders(tt, ctx):
    _say("batch vs individual")
    from trenchanttracer_amd import workloads
    scenes = []
    m1, mats1, cam1 = workloads.cornell(subdiv=3, width=64, height=48)
    scenes.append((m1, mats1, cam1))
    m2 = tt.Mesh.soup(3000, seed=5, mu=float(np.log(0.05)))
    mats2 = [tt.material(tt.MAT_DIFFUSE, (0.7, 0.7, 0.7), emission=(0.5, 0.4, 0.3))]
    cam2 = tt.camera((0.5, 0.5, -1.2), (0, 0, 1), fov=(50, 40), width=41, height=29)   # partial edge tiles
    scenes.append((m2, mats2, cam2))
    m3 = tt.Mesh.icosphere(4, 0.6, (0, 0, 0), amp=0.05, octaves=4, seed=2)
    mats3 = [tt.material(tt.MAT_DIFFUSE, (0.2, 0.6, 0.9), emission=(1, 1, 1))]
    cam3 = tt.camera((0, 0, 2.5), (0, 0, -1), fov=(40, 40), width=32, height=32)
    scenes.append((m3, mats3, cam3))
    bvhs = [tt.BVH.build(m) for m, _, _ in scenes]
    batch = tt.Batch(ctx, [m for m, _, _ in scenes], bvhs, [mm for _, mm, _ in scenes])
    cams = [c for _, _, c in scenes]
    prm = tt.render_params(spp=8, max_depth=6, seed=9)
    imgs, st = batch.render(cams, prm)
    _say(f"batch stats {st}")
    for (m, mm, c), b, img in zip(scenes, bvhs, imgs):
        s = tt.DeviceScene(ctx, m, b, mm)
        single, _ = s.render(c, prm)
        assert np.array_equal(single.view(np.int32), img.view(np.int32))
        s.close()
    # graph-captured batch equals the eager batch
    imgs_g, _ = batch.render(cams, tt.render_params(spp=8, max_depth=6, seed=9, use_graph=True))
    for a, b in zip(imgs, imgs_g):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))
    batch.close()
