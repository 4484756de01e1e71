"""ctypes wrapper of the oracle (oracle/_build/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. Never imported by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


class o_bvh(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("vertices", C.c_void_p), ("normals", C.c_void_p),
                ("tri_idx", C.c_void_p), ("mat_idx", C.c_void_p), ("n_nodes_f4", C.c_int64),
                ("n_vertices", C.c_int64)]


class o_counts(C.Structure):
    _fields_ = [("inner", C.c_int64), ("tris", C.c_int64), ("leaves", C.c_int64)]


class o_camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("view", C.c_float * 3), ("up", C.c_float * 3), ("fov", C.c_float * 2),
                ("aperture", C.c_float), ("focal", C.c_float), ("width", C.c_int32), ("height", C.c_int32)]


class o_material(C.Structure):
    _fields_ = [("type", C.c_int32), ("albedo", C.c_float * 3), ("emission", C.c_float * 3),
                ("roughness", C.c_float), ("ior", C.c_float), ("specular", C.c_float * 3)]


class o_geometry(C.Structure):
    _fields_ = [("type", C.c_int32), ("material", C.c_int32), ("a", C.c_float * 3), ("b", C.c_float * 3),
                ("c", C.c_float * 3), ("pad", C.c_float)]


class o_ref_params(C.Structure):
    _fields_ = [("spp", C.c_int32), ("max_depth", C.c_int32), ("seed", C.c_uint32), ("sample_offset", C.c_int32),
                ("pi", C.c_float), ("two_pi", C.c_float), ("tmin", C.c_float), ("tmax", C.c_float)]


class o_env(C.Structure):
    _fields_ = [("rgb", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("scale", C.c_float)]


def _load(path=LIB_PATH):
    if not os.path.exists(path):
        from . import build as _b
        _b.build()
    lib = C.CDLL(path)
    P = C.c_void_p
    lib.o_philox4x32_10.argtypes = [P, P, P]
    lib.o_philox4x32_10.restype = None
    lib.o_intersect.argtypes = [C.POINTER(o_bvh), P, C.c_int, P, P, C.POINTER(o_counts)]
    lib.o_intersect.restype = C.c_int
    lib.o_brute_batch.argtypes = [P, P, C.c_int32, P, C.c_int64, P]
    lib.o_brute_batch.restype = None
    lib.o_trace_counts.argtypes = [C.POINTER(o_bvh), P, C.c_int64, C.c_int, P, P]
    lib.o_trace_counts.restype = None
    lib.o_trace_batch.argtypes = [C.POINTER(o_bvh), P, C.c_int64, C.c_int, P, C.c_int, C.POINTER(o_counts)]
    lib.o_trace_batch.restype = C.c_double
    lib.o_tri_test.argtypes = [P, P, P, P, C.c_float, P, P, P]
    lib.o_tri_test.restype = C.c_int
    lib.o_render_ref.argtypes = [C.POINTER(o_bvh), P, P, P, P, C.c_int32, C.POINTER(o_env), C.POINTER(o_camera),
                                 C.POINTER(o_ref_params), C.c_int32, C.c_int32, P, C.c_int]
    lib.o_ref_bounce.argtypes = [C.POINTER(o_material), P, P, P, P, P]
    lib.o_ref_bounce.restype = C.c_int
    lib.o_env_lookup.argtypes = [C.POINTER(o_env), P, C.c_float, C.c_float, P]
    lib.o_env_lookup.restype = None
    lib.o_atan2f.argtypes = [C.c_float, C.c_float]
    lib.o_atan2f.restype = C.c_float
    lib.o_acosf.argtypes = [C.c_float]
    lib.o_acosf.restype = C.c_float
    lib.o_math_batch.argtypes = [C.c_int, P, P, P, C.c_int64]
    lib.o_math_batch.restype = None
    lib.o_display_frame.argtypes = [P, P, C.c_int64, C.c_int32, P]
    lib.o_display_frame.restype = None
    lib.o_render_ref.restype = C.c_int
    lib.o_geom_t.argtypes = [C.POINTER(o_geometry), P]
    lib.o_geom_t.restype = C.c_float
    for name in ("o_sinf", "o_cosf", "o_logf", "o_expf"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = [C.c_float]
            getattr(lib, name).restype = C.c_float
    if hasattr(lib, "o_powf"):
        lib.o_powf.argtypes = [C.c_float, C.c_float]
        lib.o_powf.restype = C.c_float
    return lib


lib = _load()


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.empty(4, np.uint32)
    lib.o_philox4x32_10(_p(c), _p(k), _p(o))
    return o


class CompactBVH:
    """Holds the reference compact arrays (numpy) and the C view of them."""

    def __init__(self, compact):
        self.a = {k: np.ascontiguousarray(v) for k, v in compact.items() if isinstance(v, np.ndarray)}
        self.c = o_bvh(_p(self.a["nodes"]), _p(self.a["vertices"]), _p(self.a["normals"]), _p(self.a["tri_idx"]),
                       _p(self.a["mat_idx"]), len(self.a["nodes"]), len(self.a["vertices"]))

    def trace(self, rays, any_hit=False, nthreads=None):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.empty((len(r), 4), np.float32)
        cnt = o_counts()
        nt = nthreads or min(8, os.cpu_count() or 1)
        secs = lib.o_trace_batch(C.byref(self.c), _p(r), len(r), 1 if any_hit else 0, _p(hits), nt, C.byref(cnt))
        return hits, {"inner": cnt.inner, "tris": cnt.tris, "leaves": cnt.leaves, "seconds": secs, "threads": nt}

    def trace_counts(self, rays, any_hit=False):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.empty((len(r), 4), np.float32)
        cnt = np.empty((len(r), 3), np.int64)
        lib.o_trace_counts(C.byref(self.c), _p(r), len(r), 1 if any_hit else 0, _p(hits), _p(cnt))
        return hits, cnt


def brute(verts, tris, rays):
    v = np.ascontiguousarray(verts, dtype=np.float32)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    hits = np.empty((len(r), 4), np.float32)
    lib.o_brute_batch(_p(v), _p(t), len(t), _p(r), len(r), _p(hits))
    return hits


def tri_test(v0, v1, v2, ray, tmax=None):
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (v0, v1, v2)]
    r = np.ascontiguousarray(ray, dtype=np.float32)
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    ok = lib.o_tri_test(_p(a[0]), _p(a[1]), _p(a[2]), _p(r), float(r[7] if tmax is None else tmax),
                        C.byref(t), C.byref(u), C.byref(v))
    return bool(ok), t.value, u.value, v.value


def hit_ids(h):
    return np.ascontiguousarray(h[:, 0]).view(np.int32)


def tri_normal_table(mesh_arrays):
    """12 floats (3 float4) per triangle: the shading normals of its vertices."""
    n = mesh_arrays["normals"][mesh_arrays["tri_normals"]]  # (T, 3, 3)
    out = np.zeros((len(n), 3, 4), np.float32)
    out[:, :, :3] = n
    return out.reshape(-1, 12)


def geometries(geoms):
    """ctypes array of o_geometry from objects with type, material, a, b, c fields."""
    arr = (o_geometry * max(1, len(geoms)))()
    for k, g in enumerate(geoms):
        arr[k].type, arr[k].material = int(g.type), int(g.material)
        arr[k].a[:] = list(g.a)
        arr[k].b[:] = list(g.b)
        arr[k].c[:] = list(g.c)
    return arr


def geom_t(geom, ray):
    """Distance at which `ray` (8 floats) hits one analytic primitive (+inf: none)."""
    g = geometries([geom])
    r = np.ascontiguousarray(ray, np.float32).reshape(8)
    return float(lib.o_geom_t(C.byref(g[0]), _p(r)))


def materials(mats):
    """ctypes array of o_material from objects with type, albedo, emission, roughness, ior, specular."""
    arr = (o_material * max(1, len(mats)))()
    for k, m in enumerate(mats):
        arr[k].type = m.type
        arr[k].albedo[:] = list(m.albedo)
        arr[k].emission[:] = list(m.emission)
        arr[k].roughness, arr[k].ior = m.roughness, m.ior
        arr[k].specular[:] = list(m.specular)
    return arr


class Env:
    """An env map (H, W, 3) float32 held for the oracle's hdrSample restatement."""

    def __init__(self, rgb, scale=1.0):
        self.rgb = np.ascontiguousarray(rgb, np.float32)
        assert self.rgb.ndim == 3 and self.rgb.shape[2] == 3
        self.c = o_env(self.rgb.ctypes.data, self.rgb.shape[1], self.rgb.shape[0], float(scale))

    def lookup(self, dirs, pi=3.14156265, two_pi=6.2831853071795864769, lib_=None):
        L = lib_ or lib
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.empty_like(d)
        for k in range(len(d)):
            L.o_env_lookup(C.byref(self.c), C.c_void_p(d[k:].ctypes.data), pi, two_pi,
                           C.c_void_p(out[k:].ctypes.data))
        return out


def ref_bounce(mat, u, d, point, nrm, lib_=None):
    """One reference-mode surface interaction: (continues, next origin, next direction, mask)."""
    L = lib_ or lib
    m = materials([mat])
    uu = np.ascontiguousarray(u, np.float32).reshape(4)
    dd, pp, nn = (np.ascontiguousarray(x, np.float32).reshape(3) for x in (d, point, nrm))
    out = np.empty(9, np.float32)
    c = L.o_ref_bounce(C.byref(m[0]), _p(uu), _p(dd), _p(pp), _p(nn), _p(out))
    return bool(c), out[0:3].copy(), out[3:6].copy(), out[6:9].copy()


def math(op, x, y=None, lib_=None):
    """Oracle float math over arrays: op 0 sin, 1 cos, 2 pow(x, y), 3 atan2(x, y), 4 acos."""
    L = lib_ or lib
    xa = np.ascontiguousarray(x, np.float32).ravel()
    ya = np.ascontiguousarray(y if y is not None else np.zeros_like(xa), np.float32).ravel()
    out = np.empty_like(xa)
    L.o_math_batch(int(op), _p(xa), _p(ya), _p(out), len(xa))
    return out


def display_frame(frame, sum_, frame_number, lib_=None):
    """The reference's progressive display step: updates sum_ ((..., 3) float32, in place) by
    frame and returns the packed 8-bit colours (uint32, r in the low byte)."""
    L = lib_ or lib
    f = np.ascontiguousarray(frame, np.float32)
    assert sum_.dtype == np.float32 and sum_.flags.c_contiguous and sum_.shape == f.shape and f.shape[-1] == 3
    out = np.empty(f.shape[:-1], np.uint32)
    L.o_display_frame(_p(f), _p(sum_), f.size // 3, int(frame_number), _p(out))
    return out


def load_planted(k):
    """The oracle built with planted mistake k (-DO_PLANT=k), for checks that must fail on it."""
    from . import build as _b
    return _load(_b.build(plant=k))


def render_ref(compact_bvh, mesh_arrays, materials, cam, spp, max_depth, seed, pi=3.14156265,
               two_pi=6.2831853071795864769, tmin=0.01, rows=None, nthreads=None, sample_offset=0, geoms=(),
               env=None):
    """Reference-semantics render on the CPU: (H, W, 3) float32 image (rows outside `rows` stay 0);
    `geoms` are the analytic primitives tested after the triangles; `env` (an Env, or None = black)
    is looked up on a miss."""
    oc = o_camera()
    for f in ("eye", "view", "up", "fov"):
        getattr(oc, f)[:] = list(getattr(cam, f))
    oc.aperture, oc.focal, oc.width, oc.height = cam.aperture, cam.focal, cam.width, cam.height
    mats = (o_material * len(materials))()
    for k, m in enumerate(materials):
        mats[k].type = m.type
        mats[k].albedo[:] = list(m.albedo)
        mats[k].emission[:] = list(m.emission)
        mats[k].roughness, mats[k].ior = m.roughness, m.ior
        mats[k].specular[:] = list(m.specular)
    prm = o_ref_params(int(spp), int(max_depth), int(seed), int(sample_offset), float(pi), float(two_pi),
                       float(tmin), 1e20)
    tn = tri_normal_table(mesh_arrays)
    tm = np.ascontiguousarray(mesh_arrays["tri_mats"], np.int32)
    out = np.zeros((cam.height, cam.width, 3), np.float32)
    y0, y1 = rows if rows is not None else (0, cam.height)
    nt = nthreads or min(8, os.cpu_count() or 1)
    ga = geometries(list(geoms))
    envp = C.byref(env.c) if env is not None else None
    rc = lib.o_render_ref(C.byref(compact_bvh.c), _p(tn), _p(tm), C.cast(mats, C.c_void_p), C.cast(ga, C.c_void_p),
                          len(geoms), envp, C.byref(oc), C.byref(prm), int(y0), int(y1), _p(out), nt)
    if rc != 0:
        raise RuntimeError("o_render_ref failed")
    return out
