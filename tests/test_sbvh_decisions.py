"""The product SBVH builder (csrc/host/sbvh.cpp) against the oracle's restatement of
the reference builder (oracle/sbvh_ref.py, from src/bvh/BVH.cpp:3-356), node by node.

Both trees are compared in the reference's compact form (BVHCompact.cpp:88-250): the
child boxes of every inner node (bit for bit), the child order, and each leaf's
triangle ids in leaf order; the duplicate counts must match too. The scenes are small
and chosen so that every decision rule of the reference is taken somewhere: leaves by
SAH, object splits (with ties between equal centroids and equal SAH values), spatial
splits with unsplit-left, unsplit-right and duplicated references, a flat axis
(0 * inf = NaN bin indices), and SAH sweeps across empty bins. The restatement counts
the decisions it takes; the test requires each to occur. (The fallback from a spatial
partition with an empty side to the object split, BVH.cpp:101-104, was not reached by
any of 900 seeded mixed-size scenes searched with the restatement; it stays covered
by the restatement's code only.) Each planted mistake in the restatement (a <= tie rule,
leaves in stack order, growth by an empty box skipped, centroid ties kept in stack
order) must change the tree of at least one scene.

Three rules of the reference cannot change its output, so no scene can pin them:
  - building the right child first (BVH.cpp:108-109) only orders the builder's
    internal triangle-index array; the compact arrays follow the flattening's own
    order, so a left-first build gives the same trees (the last test shows it);
  - the bin of a NaN index on a flat axis (BVH.cpp:183-186): every reference of the
    node falls in one bin of that axis, so with the other 31 bins empty every plane
    on it has infinite SAH (growth by an empty bin spans +-FLT_MAX), whichever bin
    the NaN goes to;
  - spatial over object on an exact SAH tie (BVH.cpp:97): equal SAH arises when the
    plane separates the same references as the object split does, and both then give
    the same children with the same boxes.
"""
import numpy as np
import pytest

from oracle import sbvh_ref


def _mesh(tt, verts, tris):
    return tt.Mesh.from_arrays(np.asarray(verts, np.float32), np.asarray(tris, np.int32))


def _random_tris(rng, n, scale=1.0, size=1.0):
    c = rng.uniform(0, scale, (n, 1, 3))
    v = (c + rng.uniform(-size, size, (n, 3, 3))).reshape(-1, 3)
    return v, np.arange(3 * n).reshape(n, 3)


def scenes(tt):
    rng = np.random.default_rng(20240)
    out = []
    for n in range(2, 9):                       # tiny soups of small triangles: object splits, leaves
        for s in range(4):
            m = tt.Mesh.soup(n, seed=100 * n + s)
            a = m.arrays()
            out.append((f"soup{n}_{s}", a["verts"], a["tris"], 1e-5))
    for n in (3, 5, 8, 12, 20):                 # large overlapping triangles: spatial splits
        for s in range(3):
            v, t = _random_tris(rng, n, 1.0, 0.8)
            out.append((f"big{n}_{s}", v, t, 1e-5))
    for s in range(3):                          # long thin slivers across small triangles
        v1, t1 = _random_tris(rng, 40, 1.0, 0.02)
        k = 6
        a = rng.uniform(0, 1, (k, 3))
        b = a.copy()
        b[:, s % 3] = 1.0 - a[:, s % 3] * 0.1
        c = a + rng.uniform(-0.01, 0.01, (k, 3))
        v2 = np.stack([a, b, c], 1).reshape(-1, 3)
        v = np.concatenate([v1, v2])
        t = np.concatenate([t1, len(v1) + np.arange(3 * k).reshape(k, 3)])
        out.append((f"slivers_{s}", v, t, 1e-5))
    for ax in range(3):                         # flat along one axis: NaN spatial bins
        v, t = _random_tris(rng, 24, 1.0, 0.3)
        v[:, ax] = 0.5
        out.append((f"flat{ax}", v, t, 1e-5))
    # equal centroids and equal SAH values: identical triangles, a regular grid
    base = np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0.05]], np.float32)
    v = np.concatenate([base + np.float32([0.25 * (i % 2), 0.0, 0.0]) for i in range(8)])
    out.append(("identical_pairs", v, np.arange(24).reshape(8, 3), 1e-5))
    g = []
    for i in range(4):
        for j in range(4):
            g.append(base + np.float32([0.3 * i, 0.3 * j, 0.0]))
    out.append(("grid16", np.concatenate(g), np.arange(48).reshape(16, 3), 1e-5))
    # sizes log-uniform over two decades: spatial splits with unsplit-left, unsplit-right
    # and duplicated references (seeds found by running the restatement)
    for n, seed in ((24, 13), (24, 27), (40, 11), (40, 23), (40, 4)):
        r = np.random.default_rng(seed)
        c = r.uniform(0, 1, (n, 1, 3))
        sz = np.exp(r.uniform(np.log(0.01), np.log(1.0), (n, 1, 1)))
        v = (c + sz * r.uniform(-1, 1, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
        out.append((f"mixed{n}_{seed}", v, np.arange(3 * n).reshape(n, 3), 1e-5))
    ico = tt.Mesh.icosphere(1, 0.7, (0.1, -0.2, 0.3))
    a = ico.arrays()
    out.append(("icosphere1", a["verts"], a["tris"], 1e-5))
    v, t = _random_tris(rng, 30, 1.0, 0.6)
    out.append(("big30_alpha0", v, t, 0.0))
    out.append(("big30_alpha1", v, t, 1.0))
    return out


@pytest.fixture(scope="module")
def built(tt):
    res = []
    for name, v, t, alpha in scenes(tt):
        prod = tt.BVH.build(_mesh(tt, v, t), split_alpha=alpha, nthreads=1)
        ref = sbvh_ref.build(v, t, alpha)
        res.append((name, v, t, alpha, prod, ref))
    return res


def test_product_tree_equals_restatement(built):
    bad = []
    for name, v, t, alpha, prod, ref in built:
        got = sbvh_ref.canonical_from_compact(prod.export_compact())
        if got != sbvh_ref.canonical(ref.root):
            bad.append(name)
            continue
        assert prod.stats()["num_duplicates"] == ref.duplicates, name
    assert not bad, bad


def test_every_decision_rule_is_exercised(built):
    total = {}
    for *_, ref in built:
        for k, x in ref.decisions.items():
            total[k] = total.get(k, 0) + x
    print(total)
    for k in ("leaf", "object", "spatial", "unsplit_left", "unsplit_right", "duplicate", "flat_bins"):
        assert total[k] > 0, (k, total)


@pytest.mark.parametrize("plant", [p for p in sbvh_ref.PLANTS if p not in ("none", "left_first")])
def test_planted_rule_changes_a_tree(built, plant):
    for name, v, t, alpha, prod, ref in built:
        want = sbvh_ref.canonical_from_compact(prod.export_compact())
        if sbvh_ref.canonical(sbvh_ref.build(v, t, alpha, plant=plant).root) != want:
            return
    pytest.fail(f"planted rule '{plant}' gives the product's tree on every scene")


def test_build_order_is_not_observable_in_the_compact_arrays(built):
    for name, v, t, alpha, prod, ref in built:
        b = sbvh_ref.build(v, t, alpha, plant="left_first")
        assert sbvh_ref.canonical(b.root) == sbvh_ref.canonical(ref.root), name
