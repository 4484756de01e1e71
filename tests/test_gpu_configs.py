"""The batched and learned-guiding configurations at their stated sizes.

Batched: 32 seeded scenes (displaced icospheres and triangle soups of 1k-2M
triangles, 256^2 / 512^2 / 1024^2), 16 spp, path tracing, rendered in one batched
wavefront: each scene's image equals its single-scene render bit for bit (the RNG
streams and the per-path arithmetic follow the pixel, not the batch), graph replay
equals eager launches, and the autotuner's choice gives the same images.

Guided: the path-trace case (1,310,720-tri displaced icosphere, 1920x1080, depth 16)
rendered with the network trained online from scratch (Adam lr 1e-3, 2^18-row
batches, one-sample MIS with the BSDF, learned selection probability, training passes
growing 4x from 1 spp): the image is finite, the training loss falls,
and the image agrees in the mean with an FP64-accumulated plain render once the
worst 0.1% of pixels of either image (the sun caustics' fireflies) are set aside.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_batched_32_scenes_equal_single_renders(tt, ctx):
    from trenchanttracer_amd import autotune, workloads
    _say("batched case: 32 scenes, 16 spp")
    meshes, mats, cams, env, counts = workloads.batched(32)
    bvhs = [tt.BVH.build(m, nthreads=16) for m in meshes]
    _say(f"  {sum(counts)} triangles, {sum(c.width * c.height for c in cams)} pixels")
    base = dict(spp=16, max_depth=8, rr_depth=4, seed=1, mode=tt.RENDER_PATH)
    batch = tt.Batch(ctx, meshes, bvhs, mats, env=env)
    imgs, st = batch.render(cams, tt.render_params(**base))
    _say(f"  batched: {st}")
    assert st["samples"] == 16 * sum(c.width * c.height for c in cams)
    for k, (m, b, mm, c) in enumerate(zip(meshes, bvhs, mats, cams)):
        s = tt.DeviceScene(ctx, m, b, mm, env=env)
        single, _ = s.render(c, tt.render_params(**base))
        s.close()
        assert np.isfinite(single).all()
        assert np.array_equal(single.view(np.int32), imgs[k].view(np.int32)), k
    g, _ = batch.render(cams, tt.render_params(**base, use_graph=True))
    for k in range(32):
        assert np.array_equal(g[k].view(np.int32), imgs[k].view(np.int32)), ("graph", k)
    tuned, info = autotune.tune_render(lambda p: batch.render(cams, p)[1], base, rounds=1)
    _say(f"  tuned {tuned} {info}")
    t, _ = batch.render(cams, tt.render_params(**base, **tuned))
    for k in range(32):
        assert np.array_equal(t[k].view(np.int32), imgs[k].view(np.int32)), ("tuned", k)
    batch.close()


def test_guided_pathtrace_case_trains_and_agrees_with_plain(tt, ctx):
    from trenchanttracer_amd import workloads
    _say("guided path-trace case, 1080p, lr 1e-3, 2^18-row batches")
    mesh, mats, env, cam = workloads.pathtrace()
    s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh, nthreads=16), mats, env=env)
    _, ref, _ = s.render(cam, tt.render_params(spp=256, max_depth=16, rr_depth=4, seed=1234, mode=tt.RENDER_PATH,
                                               fp64=True), want_fp64=True)
    g = tt.Guide(ctx, seed=1, lr=1e-3)
    # the bench's configuration: passes growing 8x, the camera bounce's vertices guided
    img, st, gs = g.render(s, cam, tt.render_params(spp=64, max_depth=16, rr_depth=4, seed=77, mode=tt.RENDER_PATH,
                                                    paths_in_flight=1920 * 1080),
                           train=True, train_paths=2 << 20, train_batch=1 << 18, pass_growth=8, learn_alpha=True,
                           max_bounce=1)
    _say(f"  render {st}; training {gs}")
    assert gs["train_steps"] >= 4
    assert gs["last_loss"] < gs["first_loss"], gs
    # then the trained network frozen, guiding every bounce
    img_f, st_f, _ = g.render(s, cam, tt.render_params(spp=64, max_depth=16, rr_depth=4, seed=78, mode=tt.RENDER_PATH),
                              train=False, learn_alpha=True)
    _say(f"  frozen, every bounce guided: {st_f}")
    for name, im in (("online, first bounce guided", img), ("frozen, every bounce guided", img_f)):
        assert np.isfinite(im).all()
        lum_g, lum_r = im.mean(-1), ref.mean(-1)
        cut = max(np.quantile(lum_g, 0.999), np.quantile(lum_r, 0.999))
        keep = (lum_g <= cut) & (lum_r <= cut)
        mg, mr = float(im[keep].mean()), float(ref[keep].mean())
        _say(f"  {name}: mean over {keep.mean() * 100:.2f}% of pixels: guided {mg:.5f}, plain FP64 {mr:.5f}")
        assert abs(mg - mr) <= 0.02 * mr
    g.close()
    s.close()
