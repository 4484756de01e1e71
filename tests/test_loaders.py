"""Loaders of the reference's scene pipeline: RGBE (.hdr) decoding against the closed
form (m / 256) * 2^(e - 128) on hand-built byte streams of every scanline kind,
OBJ parsing (groups, face formats, JSON material libraries, centring and scaling),
and scene configurations (material order, analytic geometry, transforms)."""
import json
import os

import numpy as np
import pytest


def _rgbe(r, g, b, e):
    return np.array([r, g, b], np.float64) / 256.0 * 2.0 ** (e - 128)


def _hdr_bytes(w, h, body, header=b"FORMAT=32-bit_rle_rgbe\n"):
    return b"#?RADIANCE\n" + header + b"\n" + f"-Y {h} +X {w}\n".encode() + body


def test_hdr_flat_and_old_rle(tt, tmp_path):
    # width 4 < 8: flat pixels; (1, 1, 1, n) repeats the previous pixel n times
    px = [(128, 64, 32, 129), (255, 0, 1, 100), (1, 1, 1, 2)]       # row 0: a, b, b, b
    body = bytes(v for p in px for v in p)
    body += bytes([10, 20, 30, 128, 1, 1, 1, 3])                      # row 1: c, c, c, c
    f = tmp_path / "a.hdr"
    f.write_bytes(_hdr_bytes(4, 2, body))
    img = tt.hdr_read(str(f))
    assert img.shape == (2, 4, 3)
    a, b = _rgbe(*px[0]).astype(np.float32), _rgbe(*px[1]).astype(np.float32)
    assert np.array_equal(img[0, 0], a)
    assert np.all(img[0, 1:] == b)
    assert np.array_equal(a, np.float32([1.0, 0.5, 0.25]))            # (m / 256) * 2^(e - 128)
    assert np.all(img[1] == _rgbe(10, 20, 30, 128).astype(np.float32))


def test_hdr_adaptive_rle(tt, tmp_path):
    w = 10
    rng = np.random.default_rng(0)
    comps = rng.integers(0, 256, (4, w)).astype(np.uint8)
    comps[1, 2:9] = 77          # a run inside the green component
    comps[3] = 131              # one exponent for the row
    body = bytes([2, 2, 0, w])
    for c in range(4):
        row = comps[c]
        if c == 1:
            body += bytes([2]) + bytes(row[:2]) + bytes([128 + 7, 77]) + bytes([1]) + bytes(row[9:])
        elif c == 3:
            body += bytes([128 + w, 131])
        else:
            body += bytes([w]) + bytes(row)
    f = tmp_path / "b.hdr"
    f.write_bytes(_hdr_bytes(w, 1, body))
    img = tt.hdr_read(str(f))
    want = np.stack([_rgbe(comps[0, x], comps[1, x], comps[2, x], comps[3, x]) for x in range(w)])
    assert np.array_equal(img[0], want.astype(np.float32))


def test_hdr_scanline_without_rle_marker(tt, tmp_path):
    # width >= 8 but the first bytes are not (2, 2, ...): the scanline is flat
    w = 8
    px = [(2, 3, 4, 129)] + [(k, k + 1, k + 2, 128) for k in range(10, 17)]
    f = tmp_path / "c.hdr"
    f.write_bytes(_hdr_bytes(w, 1, bytes(v for p in px for v in p)))
    img = tt.hdr_read(str(f))
    want = np.stack([_rgbe(*p) for p in px]).astype(np.float32)
    assert np.array_equal(img[0], want)


def test_hdr_write_read_round_trip(tt, tmp_path):
    rng = np.random.default_rng(1)
    img = (rng.random((13, 37, 3)) * np.exp(rng.normal(0, 3, (13, 37, 1)))).astype(np.float32)
    img[3, 5:20] = 0.25        # runs
    img[7] = 0.0
    f = tmp_path / "d.hdr"
    tt.hdr_write(str(f), img)
    back = tt.hdr_read(str(f))
    assert back.shape == img.shape
    mx = img.max(axis=2, keepdims=True)
    assert np.all(np.abs(back - img) <= mx * 2.0 ** -7 + 1e-30)   # RGBE: 8-bit mantissa of the largest channel


def test_hdr_errors(tt, tmp_path):
    bad = tmp_path / "bad.hdr"
    bad.write_bytes(b"#?NOTRADIANCE\n\n-Y 1 +X 1\n\x80\x80\x80\x80")
    with pytest.raises(RuntimeError):
        tt.hdr_read(str(bad))
    short = tmp_path / "short.hdr"
    short.write_bytes(_hdr_bytes(4, 2, bytes(4 * 5)))
    with pytest.raises(RuntimeError):
        tt.hdr_read(str(short))
    with pytest.raises(RuntimeError):
        tt.hdr_read(str(tmp_path / "missing.hdr"))


MTL = {
    "zeta": {"type": "Diff", "diffuseColor": [0.1, 0.2, 0.3]},
    "alpha": {"type": "Coat", "diffuseColor": [0.5, 0.4, 0.3], "specularColor": [0.9, 0.8, 0.7], "emission": [1, 2, 3]},
    "mirror": {"type": "Spec", "color": [0.7, 0.7, 0.7]},
    "glass": {"type": "Refr"},
    "gold": {"type": "Metal", "color": [1.0, 0.8, 0.3]},
}

OBJ = """# test object
mtllib lib.json
v 0 0 0
v 2 0 0
v 2 4 0
v 0 4 0
v 1 2 6
vn 0 0 1
vn 0 1 0
vt 0.5 0.5
g first
usemtl zeta
f 1//1 2//1 3//1 4//1
g second
usemtl mirror
f 1/1/2 2/1/2 5/1/2
usemtl gold
f 3 4 5
"""


def _write_obj(tmp_path):
    (tmp_path / "lib.json").write_text(json.dumps(MTL))
    (tmp_path / "obj.obj").write_text(OBJ)
    return str(tmp_path / "obj.obj")


def test_obj_groups_faces_and_materials(tt, tmp_path):
    mesh, mats = tt.obj_load(_write_obj(tmp_path))
    a = mesh.arrays()
    # materials in name order: alpha, glass, gold, mirror, zeta
    assert [m.type for m in mats] == [tt.MAT_COAT, tt.MAT_DIELECTRIC, tt.MAT_CONDUCTOR, tt.MAT_SPECULAR, tt.MAT_DIFFUSE]
    assert list(mats[0].specular) == pytest.approx([0.9, 0.8, 0.7]) and list(mats[0].emission) == [1, 2, 3]
    # the quad becomes a fan (1,2,3), (1,3,4); the group's last usemtl applies to all its faces
    assert a["tris"].tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4], [2, 3, 4]]
    assert a["tri_mats"].tolist() == [4, 4, 2, 2]
    # centred on the bounding box (1, 2, 3) and scaled by 1 / 3
    v = a["verts"]
    assert np.allclose(v[0], [-1 / 3, -2 / 3, -1]) and np.allclose(v[4], [0, 0, 1])
    assert a["tri_normals"][0].tolist() == [0, 0, 0] and a["tri_normals"][2].tolist() == [1, 1, 1]


def test_obj_without_normals_gets_face_normals(tt, tmp_path):
    (tmp_path / "n.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    mesh, mats = tt.obj_load(str(tmp_path / "n.obj"))
    a = mesh.arrays()
    assert len(mats) == 1 and mats[0].type == tt.MAT_DIFFUSE
    assert np.allclose(a["normals"][a["tri_normals"][0, 0]], [0, 0, 1])


def _rot(ax, ang):
    c, s = np.cos(ang), np.sin(ang)
    if ax == "x":
        return np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]])
    if ax == "y":
        return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]])
    return np.array([[c, -s, 0, 0], [s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])


def test_scene_config(tt, tmp_path):
    _write_obj(tmp_path)
    (tmp_path / "master.json").write_text(json.dumps({
        "light": {"type": "Diff", "diffuseColor": [0, 0, 0], "emission": [15, 15, 15]},
        "diff1": {"type": "Diff", "diffuseColor": [0.75, 0.75, 0.75]},
        "chrome": {"type": "Spec", "color": [0.9, 0.9, 0.9]}}))
    xf = {"scale": [2, 1, 0.5], "orientation": [10, 90, -20], "translate": [0, -0.5, 1]}
    cfg = {"work-dir": "data/x", "samples": 3, "material-file": "master.json", "hdr-file-name": "sky.hdr",
           "bvh-reload": True, "material-reload": False,
           "camera": {"width": 64, "height": 48, "position": [0, 0.2, 3], "pitch": 0.1, "focal": 1, "aperture": 0.0001},
           "objects": [{"type": "Plane", "p0": [-1.5, -1, -1], "p1": [1.5, -1, -1], "p2": [-1.5, 1, -1], "material": "diff1"},
                       {"type": "Sphere", "position": [0, 1, 2], "radius": 0.5, "material": "chrome"},
                       {"type": "Mesh", "file": "obj.obj", "transform": xf},
                       {"type": "Plane", "p0": [-0.3, 0.9999, 1.2], "p1": [0.3, 0.9999, 1.2], "p2": [-0.3, 0.9999, 1.8],
                        "material": "light"}]}
    (tmp_path / "scene.json").write_text(json.dumps(cfg))
    sc = tt.SceneConfig(str(tmp_path / "scene.json"), base_dir=str(tmp_path))
    i = sc.info
    assert (i.width, i.height, i.samples, i.bvh_reload, i.material_reload) == (64, 48, 3, 1, 0)
    assert list(i.position) == pytest.approx([0, 0.2, 3]) and i.yaw == 0 and i.pitch == pytest.approx(0.1)
    assert i.radius == 4 and i.fovx == 45 and i.focal == 1 and i.aperture == pytest.approx(1e-4)
    assert sc.hdr_file == "sky.hdr" and sc.work_dir == "data/x"
    # master library in name order (chrome, diff1, light), then the OBJ library (5)
    assert len(sc.materials) == 8
    assert [m.type for m in sc.materials[:3]] == [tt.MAT_SPECULAR, tt.MAT_DIFFUSE, tt.MAT_DIFFUSE]
    assert list(sc.materials[2].emission) == [15, 15, 15]
    g = sc.geometries
    assert [(x.type, x.material) for x in g] == [(1, 1), (0, 0), (1, 2)]
    assert list(g[1].a) == [0, 1, 2] and g[1].b[0] == 0.5
    # mesh: postProcess, then M = S T (Ry Rx Rz) applied to (v, 1); materials offset by 3
    mesh, _ = tt.obj_load(str(tmp_path / "obj.obj"))
    base = mesh.arrays()["verts"].astype(np.float64)
    a = np.radians(xf["orientation"])
    S = np.diag(xf["scale"] + [1])
    T = np.eye(4)
    T[:3, 3] = xf["translate"]
    M = S @ T @ (_rot("y", a[1]) @ _rot("x", a[0]) @ _rot("z", a[2]))
    want = (M @ np.c_[base, np.ones(len(base))].T).T[:, :3]
    got = sc.mesh.arrays()
    assert np.allclose(got["verts"], want, atol=2e-5)
    assert got["tri_mats"].tolist() == [3 + 4, 3 + 4, 3 + 2, 3 + 2]
    n = got["normals"]
    assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    cam = sc.camera()
    assert (cam.width, cam.height) == (64, 48)


def test_scene_config_errors(tt, tmp_path):
    (tmp_path / "m.json").write_text(json.dumps({"a": {"type": "Diff", "diffuseColor": [1, 1, 1]}}))
    base = {"work-dir": "w", "samples": 1, "material-file": "m.json", "hdr-file-name": "h", "bvh-reload": True,
            "material-reload": True, "camera": {"width": 8, "height": 8}}
    for objs in ([{"type": "Plane", "p0": [0, 0, 0], "p1": [1, 0, 0], "p2": [0, 1, 0], "material": "nope"}],
                 [{"type": "Cube"}], [{"type": "Mesh", "file": "missing.obj"}]):
        (tmp_path / "s.json").write_text(json.dumps(dict(base, objects=objs)))
        with pytest.raises(RuntimeError):
            tt.SceneConfig(str(tmp_path / "s.json"), base_dir=str(tmp_path))
    (tmp_path / "s.json").write_text("{ not json")
    with pytest.raises(RuntimeError):
        tt.SceneConfig(str(tmp_path / "s.json"), base_dir=str(tmp_path))


def _trailer_layout(data):
    """Offsets in a saved cache: start of the device node words and tri floats."""
    import ctypes
    import struct
    from trenchanttracer_amd import _native as N
    at = data.index(b"TTDL")
    dn, dt = struct.unpack_from("<qq", data, at + 8)
    nodes = at + 24 + ctypes.sizeof(N.tt_bvh_stats)
    return dn, dt, nodes, nodes + dn * 64


def test_cache_with_broken_tree_is_refused(tt, tmp_path):
    """The traversal kernel follows the cache's device layout on the GPU, so a loaded
    layout must be a tree from node 0 (no cycle, every child in range), with leaves
    inside the triangle array and the last record ending a leaf; each broken form is
    refused with TT_ERR_IO while the intact file loads."""
    import struct
    m = tt.Mesh.soup(2000, seed=5)
    b = tt.BVH.build(m)
    good = tmp_path / "good.cache"
    b.save(good)
    data = good.read_bytes()
    dn, dt, nodes, tris = _trailer_layout(data)
    assert dn > 2 and dt > 2
    tt.BVH.load(good)

    def child_word(i, c):
        return nodes + (i * 16 + 12 + c) * 4

    # find an inner child of the root and a leaf slot anywhere
    c0 = struct.unpack_from("<i", data, child_word(0, 0))[0]
    assert c0 > 0
    leaf_at = next(child_word(i, c) for i in range(dn) for c in (0, 1)
                   if struct.unpack_from("<i", data, child_word(i, c))[0] < 0)
    cases = {
        "cycle to the root": (child_word(c0, 0), struct.pack("<i", 0)),
        "child past the end": (child_word(0, 0), struct.pack("<i", dn)),
        "node reached twice": (child_word(0, 1), struct.pack("<i", c0)),
        "leaf past the triangles": (leaf_at, struct.pack("<i", ~(dt + 5))),
        "last record not ending a leaf": (tris + ((dt - 1) * 12 + 11) * 4, struct.pack("<i", 0)),
    }
    for name, (off, word) in cases.items():
        bad = bytearray(data)
        bad[off:off + 4] = word
        p = tmp_path / "bad.cache"
        p.write_bytes(bytes(bad))
        with pytest.raises(tt.TTError):
            tt.BVH.load(p)
