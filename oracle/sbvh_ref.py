"""Restatement of the reference SBVH builder (src/bvh/BVH.cpp:3-356 with
SAHHelper.cpp:19-25, AABB.cpp:11-43 and the reference-stack discipline of
BVH.h:43-91), in numpy float32 scalar arithmetic, for small meshes.

TEST INFRASTRUCTURE ONLY: the checker of the product builder
(trenchanttracer_amd/csrc/host/sbvh.cpp); imported by tests/ alone. It keeps the
reference's single reference stack (children are built right first; a leaf pops its
references from the stack top) and its in-place partition of a spatial split, so the
order of references at every node is the reference's, not derived from the product.

Reference behaviours reproduced on purpose (several can be switched off by a planted
mistake, see PLANTS, so that tests can show their checks notice it):
  - AABB::grow(AABB) grows by the other box's two corner points (AABB.cpp:16-19):
    growing by an empty box (min = +FLT_MAX, max = -FLT_MAX) yields [-FLT_MAX,
    FLT_MAX], so a spatial plane with an empty bin on either side has infinite SAH;
  - bin indices: (int) of (min - origin) * (1 / binSize), with x86's conversion of
    NaN or out-of-range values to INT_MIN (a flat axis gives 0 * inf = NaN), then
    clamped to [0, 31] and [firstBin, 31] (BVH.cpp:183-186);
  - min1f3 (LinearMath.h:66-72) and the test order leaf, spatial, object: a leaf
    wins a tie with a split, a spatial split wins a tie with the object split
    (BVH.cpp:89-104); an empty side after the spatial partition falls back to the
    object split;
  - object sweeps per axis x, y, z over the sort by (min + max, triangle id) with a
    strict < (BVH.cpp:128-151, 343-351); spatial sweeps likewise (BVH.cpp:204-228);
  - duplicate / unsplit-left / unsplit-right by min1f3 in that order of tests
    (BVH.cpp:263-298).

SAH costs are the defaults (node and triangle cost 1, batch 1, leaf sizes 1 and
0x7FFFFFF; SAHHelper.h:11-19).
"""
import numpy as np

F = np.float32
FMAX = F(np.finfo(np.float32).max)
INT_MIN = -(1 << 31)
BINS = 32
MAX_DEPTH = 64
MAX_SPATIAL_DEPTH = 48

# planted mistakes (tests only): each changes one decision rule
PLANTS = ("none", "tie_leq", "left_first", "leaf_order", "grow_skips_empty", "no_triidx_tiebreak")


class Box:
    __slots__ = ("lo", "hi")

    def __init__(self, lo=None, hi=None):
        self.lo = list(lo) if lo is not None else [FMAX, FMAX, FMAX]
        self.hi = list(hi) if hi is not None else [-FMAX, -FMAX, -FMAX]

    def copy(self):
        return Box(self.lo, self.hi)

    def grow_point(self, p):
        # min3f / max3f (LinearMath.h:252-258)
        self.lo = [a if a < b else b for a, b in zip(self.lo, p)]
        self.hi = [a if a > b else b for a, b in zip(self.hi, p)]

    def valid(self):
        return all(a <= b for a, b in zip(self.lo, self.hi))

    def area(self):
        if not self.valid():
            return F(0)
        dx, dy, dz = (b - a for a, b in zip(self.lo, self.hi))
        return (dx * dy + dy * dz + dz * dx) * F(2)

    def intersect(self, o):
        self.lo = [a if a > b else b for a, b in zip(self.lo, o.lo)]
        self.hi = [a if a < b else b for a, b in zip(self.hi, o.hi)]

    def key(self):
        return tuple(int(np.float32(v).view(np.int32)) for v in self.lo + self.hi)


def min1f(a, b):
    return b if a > b else a


def cvt_int(x):
    """(int) of a float as x86 cvttss2si: NaN or out of range -> INT_MIN."""
    if not np.isfinite(x) or x >= F(2147483648.0) or x < F(-2147483648.0):
        return INT_MIN
    return int(x)


def clamp_i(v, lo, hi):
    return lo if v < lo else (hi if v > hi else v)


class Ref:
    __slots__ = ("tri", "b")

    def __init__(self, tri, b):
        self.tri = tri
        self.b = b


class Builder:
    def __init__(self, verts, tris, split_alpha=1e-5, plant="none"):
        assert plant in PLANTS
        self.plant = plant
        self.V = [[F(c) for c in v] for v in np.asarray(verts, np.float32)]
        self.T = [tuple(int(i) for i in t) for t in np.asarray(tris, np.int32)]
        self.stack = []
        root = Box()
        for i, t in enumerate(self.T):
            b = Box()
            for j in t:
                b.grow_point(self.V[j])
            self.stack.append(Ref(i, b))
            self.grow(root, b)
        self.min_overlap = root.area() * F(split_alpha)
        self.right_bounds = [Box() for _ in range(max(len(self.T), BINS) - 1)]
        self.duplicates = 0
        self.tri_indices = []
        self.decisions = {"leaf": 0, "object": 0, "spatial": 0, "spatial_fallback": 0, "unsplit_left": 0,
                          "unsplit_right": 0, "duplicate": 0, "flat_bins": 0}
        self.root = self.build_node(len(self.T), root, 0)

    # AABB::grow(const AABB &): the two corner points of the other box
    def grow(self, box, other):
        if self.plant == "grow_skips_empty" and not other.valid():
            return
        box.grow_point(other.lo)
        box.grow_point(other.hi)

    def sort_range(self, dim):
        n0 = self.n0
        seg = self.stack[n0:]
        if self.plant == "no_triidx_tiebreak":
            seg.sort(key=lambda r: float(r.b.lo[dim] + r.b.hi[dim]))   # stable: keeps the current order on ties
        else:
            seg.sort(key=lambda r: (float(r.b.lo[dim] + r.b.hi[dim]), r.tri))
        self.stack[n0:] = seg

    def better(self, sah, best):
        return sah <= best if self.plant == "tie_leq" else sah < best

    # ------------------------------------------------------------------ buildNode, BVH.cpp:70-111
    def build_node(self, n, bounds, level):
        if n <= 1 or level >= MAX_DEPTH:
            return self.create_leaf(n, bounds)
        self.n0 = len(self.stack) - n
        area = bounds.area()
        leaf_sah = area * F(n)
        node_sah = area * F(2)
        obj = self.find_object_split(n, node_sah)
        sp_sah, sp_dim, sp_pos = FMAX, 0, F(0)
        if level < MAX_SPATIAL_DEPTH:
            ov = obj["left"].copy()
            ov.intersect(obj["right"])
            if ov.area() >= self.min_overlap:
                sp_sah, sp_dim, sp_pos = self.find_spatial_split(n, bounds, node_sah)
        min_sah = min1f(min1f(leaf_sah, obj["sah"]), sp_sah)
        if min_sah == leaf_sah and n <= 0x7FFFFFF:
            self.decisions["leaf"] += 1
            return self.create_leaf(n, bounds)
        left_n, left_b, right_n, right_b = 0, Box(), 0, Box()
        spatial = min_sah == sp_sah
        if spatial:
            left_n, left_b, right_n, right_b = self.perform_spatial_split(n, sp_dim, sp_pos)
        if not left_n or not right_n:
            if spatial:
                self.decisions["spatial_fallback"] += 1
            self.decisions["object"] += 1
            self.n0 = len(self.stack) - n
            self.sort_range(obj["dim"])
            left_n, left_b = obj["num_left"], obj["left"]
            right_n, right_b = n - obj["num_left"], obj["right"]
        else:
            self.decisions["spatial"] += 1
        self.duplicates += left_n + right_n - n
        if self.plant == "left_first":
            # build the left child first: the left child's references are not on the stack top
            top = self.stack[len(self.stack) - right_n:]
            del self.stack[len(self.stack) - right_n:]
            lnode = self.build_node(left_n, left_b, level + 1)
            self.stack.extend(top)
            rnode = self.build_node(right_n, right_b, level + 1)
        else:
            rnode = self.build_node(right_n, right_b, level + 1)
            lnode = self.build_node(left_n, left_b, level + 1)
        return ("I", bounds, lnode, rnode)

    # ------------------------------------------------------------------ createLeaf, BVH.cpp:113-120
    def create_leaf(self, n, bounds):
        tris = [self.stack.pop().tri for _ in range(n)]
        if self.plant == "leaf_order":
            tris = tris[::-1]
        self.tri_indices.extend(tris)
        return ("L", bounds, tuple(tris))

    # ------------------------------------------------------------------ findObjectSplit, BVH.cpp:122-154
    def find_object_split(self, n, node_sah):
        best = {"sah": FMAX, "dim": 0, "num_left": 0, "left": Box(), "right": Box()}
        for dim in range(3):
            self.sort_range(dim)
            refs = self.stack[self.n0:]
            rb = Box()
            for i in range(n - 1, 0, -1):
                self.grow(rb, refs[i].b)
                self.right_bounds[i - 1] = rb.copy()
            lb = Box()
            for i in range(1, n):
                self.grow(lb, refs[i - 1].b)
                sah = node_sah + lb.area() * F(i) + self.right_bounds[i - 1].area() * F(n - i)
                if self.better(sah, best["sah"]):
                    best = {"sah": sah, "dim": dim, "num_left": i, "left": lb.copy(),
                            "right": self.right_bounds[i - 1].copy()}
        return best

    # ------------------------------------------------------------------ findSpatialSplit, BVH.cpp:166-231
    def find_spatial_split(self, n, bounds, node_sah):
        origin = list(bounds.lo)
        size = [(h - o) * F(1.0 / BINS) for h, o in zip(bounds.hi, origin)]
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            inv = [F(1) / s for s in size]
            bins = [[[Box(), 0, 0] for _ in range(BINS)] for _ in range(3)]
            for ref in self.stack[self.n0:]:
                first = [cvt_int((ref.b.lo[d] - origin[d]) * inv[d]) for d in range(3)]
                last = [cvt_int((ref.b.hi[d] - origin[d]) * inv[d]) for d in range(3)]
                if INT_MIN in first + last:
                    self.decisions["flat_bins"] += 1
                first = [clamp_i(v, 0, BINS - 1) for v in first]
                last = [clamp_i(v, f, BINS - 1) for v, f in zip(last, first)]
                for dim in range(3):
                    cur = ref
                    for i in range(first[dim], last[dim]):
                        lref, rref = self.split_reference(cur, dim, origin[dim] + size[dim] * F(i + 1))
                        self.grow(bins[dim][i][0], lref.b)
                        cur = rref
                    self.grow(bins[dim][last[dim]][0], cur.b)
                    bins[dim][first[dim]][1] += 1
                    bins[dim][last[dim]][2] += 1
            best = (FMAX, 0, F(0))
            for dim in range(3):
                rb = Box()
                for i in range(BINS - 1, 0, -1):
                    self.grow(rb, bins[dim][i][0])
                    self.right_bounds[i - 1] = rb.copy()
                lb = Box()
                ln, rn = 0, n
                for i in range(1, BINS):
                    self.grow(lb, bins[dim][i - 1][0])
                    ln += bins[dim][i - 1][1]
                    rn -= bins[dim][i - 1][2]
                    sah = node_sah + lb.area() * F(ln) + self.right_bounds[i - 1].area() * F(rn)
                    if self.better(sah, best[0]):
                        best = (sah, dim, origin[dim] + size[dim] * F(i))
        return best

    # ------------------------------------------------------------------ performSpatialSplit, BVH.cpp:233-302
    def perform_spatial_split(self, n, dim, pos):
        refs = self.stack
        left_start = len(refs) - n
        left_end = left_start
        right_start = len(refs)
        lb, rb = Box(), Box()
        i = left_end
        while i < right_start:
            if refs[i].b.hi[dim] <= pos:
                self.grow(lb, refs[i].b)
                refs[i], refs[left_end] = refs[left_end], refs[i]
                left_end += 1
            elif refs[i].b.lo[dim] >= pos:
                self.grow(rb, refs[i].b)
                right_start -= 1
                refs[i], refs[right_start] = refs[right_start], refs[i]
                i -= 1
            i += 1
        while left_end < right_start:
            lref, rref = self.split_reference(refs[left_end], dim, pos)
            lub, rub, ldb, rdb = lb.copy(), rb.copy(), lb.copy(), rb.copy()
            self.grow(lub, refs[left_end].b)
            self.grow(rub, refs[left_end].b)
            self.grow(ldb, lref.b)
            self.grow(rdb, rref.b)
            lac, rac = F(left_end - left_start), F(len(refs) - right_start)
            lbc, rbc = F(left_end - left_start + 1), F(len(refs) - right_start + 1)
            unsplit_left = lub.area() * lbc + rb.area() * rac
            unsplit_right = lb.area() * lac + rub.area() * rbc
            dup = ldb.area() * lbc + rdb.area() * rbc
            m = min1f(min1f(unsplit_left, unsplit_right), dup)
            if m == unsplit_left:
                self.decisions["unsplit_left"] += 1
                lb = lub
                left_end += 1
            elif m == unsplit_right:
                self.decisions["unsplit_right"] += 1
                rb = rub
                right_start -= 1
                refs[left_end], refs[right_start] = refs[right_start], refs[left_end]
            else:
                self.decisions["duplicate"] += 1
                lb, rb = ldb, rdb
                refs[left_end] = lref
                left_end += 1
                refs.append(rref)
        return left_end - left_start, lb, len(refs) - right_start, rb

    # ------------------------------------------------------------------ splitReference, BVH.cpp:304-341
    def split_reference(self, ref, dim, pos):
        lb, rb = Box(), Box()
        i0, i1, i2 = self.T[ref.tri]
        v1 = self.V[i2]
        for ind in (i0, i1, i2):
            v0 = v1
            v1 = self.V[ind]
            v0p, v1p = v0[dim], v1[dim]
            if v0p <= pos:
                lb.grow_point(v0)
            if v0p >= pos:
                rb.grow_point(v0)
            if (v0p < pos and v1p > pos) or (v0p > pos and v1p < pos):
                t = (pos - v0p) / (v1p - v0p)
                t = F(0) if t < F(0) else (F(1) if t > F(1) else t)
                p = [a * (F(1) - t) + b * t for a, b in zip(v0, v1)]
                lb.grow_point(p)
                rb.grow_point(p)
        lb.hi[dim] = pos
        rb.lo[dim] = pos
        lb.intersect(ref.b)
        rb.intersect(ref.b)
        return Ref(ref.tri, lb), Ref(ref.tri, rb)


def canonical(node, root=True):
    """Nested form of a tree for comparison: inner -> ("I", [(child box bits, child) x 2]);
    leaf -> ("L", triangle ids in leaf order). The root's own box is not part of it (the
    compact layout stores child boxes only). A root that is a leaf (the reference's
    flattening asserts on it, BVHCompact.cpp:104) is stored as an inner node holding
    that leaf and an empty leaf with an empty box, as the product stores it."""
    if node[0] == "L":
        if root:
            return ("I", ((node[1].key(), ("L", node[2])), (Box().key(), ("L", ()))))
        return ("L", node[2])
    return ("I", ((node[2][1].key(), canonical(node[2], False)), (node[3][1].key(), canonical(node[3], False))))


def canonical_from_compact(c):
    """The same nested form decoded from the reference compact arrays
    (BVHCompact.cpp:88-250): child index >= 0 -> inner node at that float4 offset, < 0
    -> leaf whose triangles start at vertex float4 ~index, 3 per triangle, up to the
    {0x80000000, 0, 0, 0} terminator; triangle ids from tri_idx at each start."""
    nodes = np.asarray(c["nodes"], np.int32).reshape(-1)
    verts = np.asarray(c["vertices"], np.float32).reshape(-1, 4)
    tri_idx = np.asarray(c["tri_idx"], np.int32)
    vbits = verts.view(np.int32)

    def leaf(v):
        out = []
        while vbits[v, 0] != np.int32(-0x80000000):
            out.append(int(tri_idx[v]))
            v += 3
        return ("L", tuple(out))

    def inner(off):
        w = nodes[4 * off: 4 * off + 16]
        boxes = ((w[0], w[2], w[8], w[1], w[3], w[9]), (w[4], w[6], w[10], w[5], w[7], w[11]))
        kids = []
        for i in range(2):
            p = int(w[12 + i])
            kids.append((tuple(int(x) for x in boxes[i]), inner(p) if p >= 0 else leaf(~p)))
        return ("I", tuple(kids))

    return inner(0)


def build(verts, tris, split_alpha=1e-5, plant="none"):
    return Builder(verts, tris, split_alpha, plant)
