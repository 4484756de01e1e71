import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def tt():
    import trenchanttracer_amd as m
    return m


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    return o


@pytest.fixture(scope="session")
def ctx(tt):
    c = tt.Context(0)
    yield c
    c.close()


@pytest.fixture
def scene_files(tmp_path):
    """A scene in the reference's file formats under tmp_path: a JSON master material
    library, an OBJ with groups and its own JSON library, and a configuration with
    analytic planes and a sphere plus the transformed OBJ. Returns the config path."""
    import json
    (tmp_path / "lib.json").write_text(json.dumps({
        "body": {"type": "Diff", "diffuseColor": [0.8, 0.7, 0.6]},
        "trim": {"type": "Coat", "diffuseColor": [0.3, 0.3, 0.6], "specularColor": [0.9, 0.9, 0.9]}}))
    (tmp_path / "thing.obj").write_text(
        "mtllib lib.json\n"
        "v -1 -1 -1\nv 1 -1 -1\nv 1 1 -1\nv -1 1 -1\nv -1 -1 1\nv 1 -1 1\nv 1 1 1\nv -1 1 1\n"
        "g sides\nusemtl body\n"
        "f 1 2 6 5\nf 2 3 7 6\nf 3 4 8 7\nf 4 1 5 8\n"
        "g caps\nusemtl trim\n"
        "f 1 4 3 2\nf 5 6 7 8\n")
    (tmp_path / "master.json").write_text(json.dumps({
        "diff1": {"type": "Diff", "diffuseColor": [0.75, 0.75, 0.75]},
        "light": {"type": "Diff", "diffuseColor": [0, 0, 0], "emission": [15, 15, 15]},
        "chrome": {"type": "Spec", "color": [0.9, 0.9, 0.9]},
        "glass": {"type": "Refr"},
        "brass": {"type": "Metal", "color": [0.9, 0.8, 0.5]}}))
    objects = [
        {"type": "Plane", "p0": [-1.5, -1, -1], "p1": [1.5, -1, -1], "p2": [-1.5, 1, -1], "material": "diff1"},
        {"type": "Plane", "p0": [-1.5, -1, -1], "p1": [-1.5, 1, -1], "p2": [-1.5, -1, 4], "material": "diff1"},
        {"type": "Plane", "p0": [1.5, -1, -1], "p1": [1.5, -1, 4], "p2": [1.5, 1, -1], "material": "diff1"},
        {"type": "Plane", "p0": [-1.5, 1, -1], "p1": [1.5, 1, -1], "p2": [-1.5, 1, 4], "material": "diff1"},
        {"type": "Plane", "p0": [-30, -1, -30], "p1": [-30, -1, 30], "p2": [30, -1, -30], "material": "diff1"},
        {"type": "Plane", "p0": [-0.3, 0.9999, 1.2], "p1": [0.3, 0.9999, 1.2], "p2": [-0.3, 0.9999, 1.8],
         "material": "light"},
        {"type": "Sphere", "position": [-0.8, -0.7, 0.9], "radius": 0.3, "material": "chrome"},
        {"type": "Sphere", "position": [0.8, -0.7, 1.6], "radius": 0.25, "material": "glass"},
        {"type": "Mesh", "file": "thing.obj",
         "transform": {"scale": [0.3, 0.3, 0.3], "orientation": [0, 30, 0], "translate": [0, -2.3, 4.5]}}]
    cfg = {"work-dir": "work", "samples": 4, "material-file": "master.json", "hdr-file-name": "sky.hdr",
           "bvh-reload": True, "material-reload": True,
           "camera": {"width": 96, "height": 72, "position": [0, 0.2, 3], "yaw": 0, "pitch": 0.1, "radius": 4,
                      "focal": 1, "aperture": 0.0},
           "objects": objects}
    (tmp_path / "scene.json").write_text(json.dumps(cfg))
    return tmp_path / "scene.json"
