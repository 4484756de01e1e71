"""A scene in the reference's file formats (JSON configuration, JSON material
libraries, OBJ, RGBE environment) loaded and rendered on the GPU: reference mode
bit-identical to the CPU restatement; path mode with the environment read from
an .hdr file."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, file=sys.stderr, flush=True)


def test_config_reference_render_matches_oracle(tt, oracle, ctx, scene_files):
    sc = tt.SceneConfig(str(scene_files), base_dir=str(scene_files.parent))
    _say(f"config: {sc.info.num_tris} tris, {len(sc.geometries)} primitives, {len(sc.materials)} materials")
    b = tt.BVH.build(sc.mesh)
    s = tt.DeviceScene(ctx, sc.mesh, b, sc.materials)
    s.set_geometries(sc.geometries)
    cam = sc.camera()
    img, st = s.render(cam, tt.render_params(spp=sc.samples, max_depth=10, seed=9, mode=tt.RENDER_REFERENCE))
    ref = oracle.render_ref(oracle.CompactBVH(b.export_compact()), sc.mesh.arrays(), sc.materials, cam,
                            spp=sc.samples, max_depth=10, seed=9, geoms=sc.geometries)
    same = np.mean(img.view(np.int32) == ref.view(np.int32))
    _say(f"  {same * 100:.3f}% bit-identical, mean {img.mean():.4f}")
    assert same == 1.0 and img.mean() > 0.02
    s.close()


def test_config_path_render_with_hdr_environment(tt, ctx, scene_files):
    from trenchanttracer_amd import workloads
    sky = workloads.sky_env(256, 128, sun_radiance=1e3, sun_diam_deg=2.0)
    tt.hdr_write(str(scene_files.parent / "sky.hdr"), sky)
    sc = tt.SceneConfig(str(scene_files), base_dir=str(scene_files.parent))
    env = tt.hdr_read(str(scene_files.parent / sc.hdr_file))
    assert env.shape == sky.shape
    s = tt.DeviceScene(ctx, sc.mesh, tt.BVH.build(sc.mesh), sc.materials, env=env)
    s.set_geometries(sc.geometries)
    img, st = s.render(sc.camera(), tt.render_params(spp=16, max_depth=8, seed=3, mode=tt.RENDER_PATH))
    _say(f"  path render mean {img.mean():.4f}, rays {st['rays_extension']}")
    assert np.all(np.isfinite(img)) and img.mean() > 0.02
    s.close()
