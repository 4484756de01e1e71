"""Progressive display on the GPU (tt_display_frame) against the oracle's restatement of
the reference's accumulation and 8-bit output (RenderKernel.cu:202-217,
Renderer.cpp:86-93): the running sum and the packed colours bit-identical over
successive reference-mode frames of the config scene, over synthetic frames with
negative, large and non-finite values, and across a reset."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, file=sys.stderr, flush=True)


def test_display_frames_match_oracle(tt, oracle, ctx, scene_files):
    sc = tt.SceneConfig(str(scene_files), base_dir=str(scene_files.parent))
    s = tt.DeviceScene(ctx, sc.mesh, tt.BVH.build(sc.mesh), sc.materials)
    s.set_geometries(sc.geometries)
    cam = sc.camera()
    disp = tt.Display(ctx, cam.width, cam.height)
    ref_sum = np.zeros((cam.height, cam.width, 3), np.float32)
    for k in range(1, 5):
        _say(f"display frame {k} ({cam.width}x{cam.height})")
        img, _ = s.render(cam, tt.render_params(spp=2, max_depth=10, seed=1000 + k, mode=tt.RENDER_REFERENCE))
        rgba = disp.frame(img)
        want = oracle.display_frame(img, ref_sum, k)
        assert np.array_equal(disp.running_sum().view(np.uint32), ref_sum.view(np.uint32)), f"sum, frame {k}"
        assert np.array_equal(rgba.reshape(-1, 4).view(np.uint32).ravel(), want.ravel()), f"colours, frame {k}"
    assert rgba[..., :3].mean() > 5 and np.all(rgba[..., 3] == 1)
    # reset: the next frame restarts from the byte-1 pattern
    disp.reset()
    img, _ = s.render(cam, tt.render_params(spp=1, max_depth=10, seed=77, mode=tt.RENDER_REFERENCE))
    rgba = disp.frame(img)
    want = oracle.display_frame(img, ref_sum, 1)
    assert np.array_equal(disp.running_sum().view(np.uint32), ref_sum.view(np.uint32))
    assert np.array_equal(rgba.reshape(-1, 4).view(np.uint32).ravel(), want.ravel())
    disp.close()
    s.close()


def test_display_synthetic_values_match_oracle(tt, oracle, ctx):
    rng = np.random.default_rng(3)
    h, w = 97, 131                       # npix not a multiple of the block size
    special = np.array([0.0, -0.0, 1.0, -1.0, 2.0, 1e30, -1e30, np.inf, -np.inf, np.nan, 1e-40, 0.5], np.float32)
    disp = tt.Display(ctx, w, h)
    ref_sum = np.zeros((h, w, 3), np.float32)
    for k in range(1, 4):
        f = (rng.standard_normal((h, w, 3)) * 0.7 + 0.4).astype(np.float32)
        f.reshape(-1)[: special.size * k] = np.tile(special, k)
        _say(f"synthetic display frame {k}")
        rgba = disp.frame(f)
        want = oracle.display_frame(f, ref_sum, k)
        got_sum = disp.running_sum()
        assert np.array_equal(np.isnan(got_sum), np.isnan(ref_sum))
        fin = ~np.isnan(ref_sum)
        assert np.array_equal(got_sum[fin].view(np.uint32), ref_sum[fin].view(np.uint32)), f"sum, frame {k}"
        assert np.array_equal(rgba.reshape(-1, 4).view(np.uint32).ravel(), want.ravel()), f"colours, frame {k}"
    disp.close()


def test_display_rejects_bad_arguments(tt, ctx):
    d = tt.DeviceBuffer(ctx, 64)
    with pytest.raises(tt.TTError):
        tt.check(tt.lib.tt_display_frame(ctx._h, tt.C.c_void_p(d.ptr), tt.C.c_void_p(d.ptr), 4, 0,
                                         tt.C.c_void_p(d.ptr), None))
    d.free()
