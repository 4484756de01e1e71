"""The oracle pinned against known answers, closed forms and brute force."""
import numpy as np
import pytest


def test_philox_known_answers(oracle):
    # Random123 kat_vectors for philox4x32-10.
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(oracle.philox(ctr, key)) == want


def _ray(o, d, tmin=0.0, tmax=1e20):
    return np.array([*o, tmin, *d, tmax], np.float32)


def test_tri_test_closed_forms(oracle):
    v0, v1, v2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]      # e1 x e2 = +z
    # ray travelling -z hits the front face: det = 2*area*cos = 1 * 1
    ok, t, u, v = oracle.tri_test(v0, v1, v2, _ray([0.25, 0.25, 2], [0, 0, -1]))
    assert ok and t == 2.0 and u == 0.25 and v == 0.25
    # back face (travelling +z) is never accepted (IntersectKernel.cuh:248-250)
    ok, *_ = oracle.tri_test(v0, v1, v2, _ray([0.25, 0.25, -2], [0, 0, 1]))
    assert not ok
    # edge u + v = 1 inclusive
    ok, t, u, v = oracle.tri_test(v0, v1, v2, _ray([0.5, 0.5, 1], [0, 0, -1]))
    assert ok and u + v == 1.0
    # just outside
    ok, *_ = oracle.tri_test(v0, v1, v2, _ray([0.5, 0.50001, 1], [0, 0, -1]))
    assert not ok
    # t == tmin is exclusive, t == tmax exclusive
    assert not oracle.tri_test(v0, v1, v2, _ray([0.25, 0.25, 2], [0, 0, -1], tmin=2.0))[0]
    assert not oracle.tri_test(v0, v1, v2, _ray([0.25, 0.25, 2], [0, 0, -1], tmax=2.0))[0]
    assert oracle.tri_test(v0, v1, v2, _ray([0.25, 0.25, 2], [0, 0, -1], tmin=1.9999))[0]
    # |det| = 2 * area * |cos| <= 1e-5 is rejected (absolute epsilon)
    s = 2e-3  # area = s^2/2 = 2e-6 -> det = 4e-6 at normal incidence
    ok, *_ = oracle.tri_test([0, 0, 0], [s, 0, 0], [0, s, 0], _ray([s / 4, s / 4, 1], [0, 0, -1]))
    assert not ok
    s = 5e-3  # det = 2.5e-5 > 1e-5
    ok, *_ = oracle.tri_test([0, 0, 0], [s, 0, 0], [0, s, 0], _ray([s / 4, s / 4, 1], [0, 0, -1]))
    assert ok
    # grazing: det = s^2 * cos; s = 1, cos = 5e-6 -> rejected
    d = np.array([0, np.sqrt(1 - 25e-12), -5e-6])
    ok, *_ = oracle.tri_test([0, 0, 0], [1, 0, 0], [0, 1, 0], _ray([0.2, -0.5, 1e-6 * 0 + 3.5e-6], d))
    assert not ok


def test_tri_test_random_vs_float64(oracle):
    """Acceptance agrees with an exact float64 barycentric test away from edges."""
    rng = np.random.default_rng(3)
    agree = 0
    for _ in range(2000):
        v = rng.uniform(-1, 1, (3, 3))
        o = rng.uniform(-2, 2, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        e1, e2 = v[1] - v[0], v[2] - v[0]
        p = np.cross(d, e2)
        det = e1 @ p
        tv = o - v[0]
        u = (tv @ p) / det
        q = np.cross(tv, e1)
        vv = (d @ q) / det
        t = (e2 @ q) / det
        margin = min(abs(u), abs(vv), abs(1 - u - vv), abs(u - 1), abs(vv - 1))
        if margin < 1e-4 or abs(det) < 1e-4 or abs(t) < 1e-4:
            continue
        want = det > 1e-5 and u >= 0 and vv >= 0 and u + vv <= 1 and t > 0
        ok, tt_, uu, vw = oracle.tri_test(v[0], v[1], v[2], _ray(o, d))
        assert ok == want
        if ok:
            assert abs(tt_ - t) <= 1e-5 * max(1, abs(t))
        agree += 1
    assert agree > 1000


@pytest.mark.parametrize("ntris,seed", [(1, 5), (2, 6), (300, 7), (5000, 8)])
def test_bvh_traversal_equals_brute_force(tt, oracle, ntris, seed):
    m = tt.Mesh.soup(ntris, seed=seed, mu=float(np.log(0.2 if ntris < 100 else 0.03)))
    b = tt.BVH.build(m)
    cb = oracle.CompactBVH(b.export_compact())
    rays = np.concatenate([tt.rays_primary(48, 40, 2, (0.5, 0.5, -1.0), (0.5, 0.5, 0.0), seed=seed),
                           tt.rays_incoherent(m, 3000, seed=seed)])
    h, st = cb.trace(rays)
    a = m.arrays()
    hb = oracle.brute(a["verts"], a["tris"], rays)
    ids, idb = oracle.hit_ids(h), oracle.hit_ids(hb)
    assert (ids >= 0).sum() > 0
    assert np.array_equal(ids, idb)
    assert np.array_equal(h[:, 1:].view(np.int32), hb[:, 1:].view(np.int32))


def test_cornell_hand_checked(tt, oracle):
    """Reference Cornell quads face the interior: rays from inside hit them."""
    back = tt.Mesh.quad([-1.5, -1, -1], [1.5, -1, -1], [-1.5, 1, -1])
    left = tt.Mesh.quad([-1.5, -1, -1], [-1.5, 1, -1], [-1.5, -1, 4])
    b = tt.BVH.build(back.append(left))
    cb = oracle.CompactBVH(b.export_compact())
    rays = np.stack([_ray([0, 0, 1], [0, 0, -1]), _ray([0, 0, 1], [-1, 0, 0]), _ray([0, 0, 1], [0, 0, 1])])
    h, _ = cb.trace(rays)
    ids = oracle.hit_ids(h)
    assert ids[0] in (0, 1) and abs(h[0, 1] - 2.0) < 1e-6
    assert ids[1] in (2, 3) and abs(h[1, 1] - 1.5) < 1e-6
    assert ids[2] == -1


def test_any_hit_consistent_with_closest(tt, oracle):
    m = tt.Mesh.soup(3000, seed=11, mu=float(np.log(0.05)))
    cb = oracle.CompactBVH(tt.BVH.build(m).export_compact())
    rays = tt.rays_incoherent(m, 5000, seed=12)
    hc, _ = cb.trace(rays)
    ha, _ = cb.trace(rays, any_hit=True)
    assert np.array_equal(oracle.hit_ids(hc) >= 0, oracle.hit_ids(ha) >= 0)


def test_batch_threads_equal_serial(tt, oracle):
    m = tt.Mesh.soup(4000, seed=2)
    cb = oracle.CompactBVH(tt.BVH.build(m).export_compact())
    rays = tt.rays_incoherent(m, 4000, seed=3)
    h8, c8 = cb.trace(rays, nthreads=8)
    h1, c1 = cb.trace(rays, nthreads=1)
    hs, cnt = cb.trace_counts(rays)
    assert np.array_equal(h8.view(np.int32), h1.view(np.int32))
    assert np.array_equal(h8.view(np.int32), hs.view(np.int32))
    assert c8["inner"] == c1["inner"] == cnt[:, 0].sum()
