"""MI355X-native rendering hot path of TrenchantTracer.

Python mirror of the reference's host interface over the C ABI (include/tt.h):

    Mesh        ~ Scene (src/core/Scene.cpp): vertices, normals, triangles, materials
    BVH         ~ BVH + BVHCompact (src/bvh/BVH.cpp, src/bvh/BVHCompact.cpp)
    Context     ~ device selection + stream (src/main.cpp:97-98)
    DeviceScene ~ Renderer's device copies (src/core/Renderer.cpp:8-57)
    DeviceScene.intersect / trace_closest / trace_any ~ intersect() (IntersectKernel.cuh:78)

Device entry points run the HIP kernels of libtt.so; they raise if the library
or a GPU is missing (there is no CPU fallback).
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import _native as N
from ._native import TTError, check, lib, tt_trace_cfg

__all__ = ["Mesh", "BVH", "Context", "DeviceScene", "DeviceBuffer", "Timer", "Batch", "Guide", "TTError", "rays_primary", "rays_incoherent",
           "trace_cfg", "MAT_DIFFUSE", "MAT_CONDUCTOR", "MAT_DIELECTRIC", "MAT_EMITTER", "material"]

MAT_DIFFUSE, MAT_CONDUCTOR, MAT_DIELECTRIC, MAT_EMITTER, MAT_SPECULAR, MAT_COAT = 0, 1, 2, 3, 4, 5


def _fptr(a):
    return a.ctypes.data_as(N.PF)


def _vp(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _f3(x):
    return (C.c_float * 3)(*[float(v) for v in x])


class Mesh:
    """Host triangle mesh (flat, as the reference Scene)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @classmethod
    def from_arrays(cls, verts, tris, normals=None, tri_normals=None, tri_mats=None):
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        n = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        tn = None if tri_normals is None else np.ascontiguousarray(tri_normals, dtype=np.int32).reshape(-1, 3)
        tm = None if tri_mats is None else np.ascontiguousarray(tri_mats, dtype=np.int32).reshape(-1)
        h = C.c_void_p()
        check(lib.tt_mesh_create(_vp(v), len(v), _vp(t), len(t), _vp(n), 0 if n is None else len(n), _vp(tn),
                                 _vp(tm), C.byref(h)))
        return cls(h)

    @classmethod
    def soup(cls, ntris, seed=1, mu=float(np.log(0.01)), sigma=0.5):
        h = C.c_void_p()
        check(lib.tt_mesh_gen_soup(int(ntris), int(seed), float(mu), float(sigma), C.byref(h)))
        return cls(h)

    @classmethod
    def icosphere(cls, subdiv, radius=1.0, center=(0, 0, 0), amp=0.0, octaves=6, seed=1, mat=0):
        h = C.c_void_p()
        check(lib.tt_mesh_gen_icosphere(int(subdiv), float(radius), _f3(center), float(amp), int(octaves),
                                        int(seed), int(mat), C.byref(h)))
        return cls(h)

    @classmethod
    def quad(cls, p0, p1, p2, mat=0):
        h = C.c_void_p()
        check(lib.tt_mesh_gen_quad(_f3(p0), _f3(p1), _f3(p2), int(mat), C.byref(h)))
        return cls(h)

    def append(self, other, mat_offset=0):
        check(lib.tt_mesh_append(self._h, other._h, int(mat_offset)))
        return self

    def assign_materials(self, fractions, mats, seed=1):
        f = np.ascontiguousarray(fractions, dtype=np.float32)
        m = np.ascontiguousarray(mats, dtype=np.int32)
        check(lib.tt_mesh_assign_materials(self._h, len(f), _fptr(f), m.ctypes.data_as(N.PI32), int(seed)))

    def info(self):
        nv, nt, nn = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib.tt_mesh_info(self._h, C.byref(nv), C.byref(nt), C.byref(nn)))
        return nv.value, nt.value, nn.value

    @property
    def ntris(self):
        return self.info()[1]

    def arrays(self):
        nv, nt, nn = self.info()
        v = np.empty((nv, 3), np.float32)
        t = np.empty((nt, 3), np.int32)
        n = np.empty((nn, 3), np.float32)
        tn = np.empty((nt, 3), np.int32)
        tm = np.empty(nt, np.int32)
        check(lib.tt_mesh_get(self._h, _vp(v), _vp(t), _vp(n), _vp(tn), _vp(tm)))
        return {"verts": v, "tris": t, "normals": n, "tri_normals": tn, "tri_mats": tm}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.tt_mesh_destroy(h)
            self._h = None


class BVH:
    """Host SBVH: BVH::BVH (src/bvh/BVH.cpp:3-63) + BVHCompact (BVHCompact.cpp:88-250)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def build(cls, mesh, split_alpha=1.0e-5, nthreads=0):
        h = C.c_void_p()
        check(lib.tt_bvh_build(mesh._h, float(split_alpha), int(nthreads), C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, path):
        h = C.c_void_p()
        check(lib.tt_bvh_load(str(path).encode(), C.byref(h)))
        return cls(h)

    def save(self, path):
        check(lib.tt_bvh_save(self._h, str(path).encode()))

    def stats(self):
        s = N.tt_bvh_stats()
        check(lib.tt_bvh_get_stats(self._h, C.byref(s)))
        return {"num_tris": s.num_tris, "num_refs": s.num_refs, "num_duplicates": s.num_duplicates,
                "num_inner": s.num_inner, "num_leaves": s.num_leaves, "max_depth": s.max_depth,
                "sah": s.sah, "root_lo": list(s.root_lo), "root_hi": list(s.root_hi),
                "build_seconds": s.build_seconds}

    def export_compact(self):
        """Reference compact arrays (nodes int32 [n,4], vertices/normals float32 [n,4], tri/mat int32)."""
        sz = N.tt_compact_sizes()
        check(lib.tt_bvh_export_compact(self._h, C.byref(sz), None, None, None, None, None))
        nodes = np.empty((sz.nodes, 4), np.int32)
        verts = np.empty((sz.vertices, 4), np.float32)
        norms = np.empty((sz.normals, 4), np.float32)
        tri = np.empty(sz.tri_indices, np.int32)
        mat = np.empty(sz.mat_indices, np.int32)
        check(lib.tt_bvh_export_compact(self._h, C.byref(sz), _vp(nodes), _vp(verts), _vp(norms), _vp(tri),
                                        _vp(mat)))
        return {"nodes": nodes, "vertices": verts, "normals": norms, "tri_idx": tri, "mat_idx": mat,
                "leaf_count": sz.leaf_count, "tri_count": sz.tri_count}

    def device_layout(self):
        nn, nt = C.c_int64(), C.c_int64()
        check(lib.tt_bvh_export_device_layout(self._h, C.byref(nn), C.byref(nt), None, None))
        nodes = np.empty((nn.value, 16), np.int32)
        tris = np.empty((nt.value, 12), np.float32)
        check(lib.tt_bvh_export_device_layout(self._h, None, None, _vp(nodes), _vp(tris)))
        return nodes, tris

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.tt_bvh_destroy(h)
            self._h = None


def rays_primary(width, height, spp, eye, target, fovx=45.0, tmin=0.0, tmax=1e20, seed=2):
    out = np.empty((width * height * spp, 8), np.float32)
    check(lib.tt_rays_primary(int(width), int(height), int(spp), _f3(eye), _f3(target), float(fovx), float(tmin),
                              float(tmax), int(seed), _vp(out)))
    return out


def rays_incoherent(mesh, n, tmin=1e-4, tmax=1e20, seed=2):
    out = np.empty((int(n), 8), np.float32)
    check(lib.tt_rays_incoherent(mesh._h, int(n), float(tmin), float(tmax), int(seed), _vp(out)))
    return out


def trace_cfg(variant=0, lds_stack=16, block=128, xcd_order=0, refill=0, reorder=0):
    return tt_trace_cfg(int(variant), int(lds_stack), int(block), int(xcd_order), int(refill), int(reorder))


def material(type, albedo=(0.8, 0.8, 0.8), emission=(0, 0, 0), roughness=0.0, ior=1.5, specular=(1.0, 1.0, 1.0)):
    """tt_material; `specular` is the mirror colour of a coat (albedo its diffuse colour)."""
    m = N.tt_material()
    m.type = int(type)
    m.albedo[:] = [float(x) for x in albedo]
    m.emission[:] = [float(x) for x in emission]
    m.roughness = float(roughness)
    m.ior = float(ior)
    m.specular[:] = [float(x) for x in specular]
    return m


GEOM_SPHERE, GEOM_PLANE = 0, 1


def sphere(center, radius, mat):
    """Analytic sphere (tt_geometry), as the reference's Sphere object."""
    g = N.tt_geometry()
    g.type, g.material = GEOM_SPHERE, int(mat)
    g.a[:] = [float(x) for x in center]
    g.b[:] = [float(radius), 0.0, 0.0]
    return g


def plane(p0, p1, p2, mat):
    """Analytic parallelogram p0 + u (p1 - p0) + v (p2 - p0), u, v in [0, 1] (the
    reference's Plane object, tested on both sides)."""
    g = N.tt_geometry()
    g.type, g.material = GEOM_PLANE, int(mat)
    g.a[:] = [float(x) for x in p0]
    g.b[:] = [float(x) for x in p1]
    g.c[:] = [float(x) for x in p2]
    return g


REF_PI = 3.14156265        # the reference's PI (src/core/Renderer.cpp:5, Camera.cpp:5)
REF_TWO_PI = 6.2831853071795864769
RENDER_REFERENCE, RENDER_PATH = 0, 1


def camera(eye, view, up=(0, 1, 0), fov=(45.0, 45.0), width=256, height=256, aperture=0.0, focal=1.0):
    c = N.tt_camera()
    c.eye[:] = [float(x) for x in eye]
    c.view[:] = [float(x) for x in view]
    c.up[:] = [float(x) for x in up]
    c.fov[:] = [float(x) for x in fov]
    c.aperture = float(aperture)
    c.focal = float(focal)
    c.width, c.height = int(width), int(height)
    return c


def camera_from_orbit(center, yaw, pitch, radius=4.0, fovx=45.0, width=256, height=256, focal=1.0, aperture=0.0):
    """Camera::Camera + setFovx + getCameraMeta (src/core/Camera.cpp:12-44, 115-140)."""
    c = N.tt_camera()
    check(lib.tt_camera_from_orbit(_f3(center), float(yaw), float(pitch), float(radius), float(fovx), int(width),
                                   int(height), float(focal), float(aperture), C.byref(c)))
    return c


def render_params(spp=16, max_depth=8, rr_depth=4, seed=1, mode=RENDER_REFERENCE, guided=0, sample_offset=0,
                  fp64=False, pi=None, two_pi=None, tmin=0.01, use_graph=False, paths_in_flight=0, lds_stack=0,
                  block=0, variant=-1, refill_camera=0, refill=0, tile=0):
    p = N.tt_render_params()
    p.refill_camera, p.refill, p.tile = int(refill_camera), int(refill), int(tile)
    p.use_graph, p.paths_in_flight, p.lds_stack, p.block, p.variant = (int(bool(use_graph)), int(paths_in_flight),
                                                                       int(lds_stack), int(block), int(variant))
    p.spp, p.max_depth, p.rr_depth, p.seed = int(spp), int(max_depth), int(rr_depth), int(seed)
    p.mode, p.guided, p.sample_offset, p.accumulate_fp64 = int(mode), int(guided), int(sample_offset), int(bool(fp64))
    if pi is None:
        pi = REF_PI if mode == RENDER_REFERENCE else float(np.pi)
    p.pi = float(pi)
    p.two_pi = float(REF_TWO_PI if two_pi is None else two_pi)
    p.tmin = float(tmin)
    return p


def _stats(st):
    return {"rays_extension": st.rays_extension, "rays_shadow": st.rays_shadow, "samples": st.samples,
            "seconds": st.seconds, "iterations": st.iterations, "paths_in_flight": st.paths_in_flight}


class Batch:
    """Up to 64 scenes rendered in one wavefront (tt_batch): one node array with
    per-scene roots, per-scene cameras / resolutions, one shared env map."""

    def __init__(self, ctx, meshes, bvhs, materials, env=None, env_scale=1.0):
        n = len(meshes)
        self.ctx = ctx
        self._mats = [m if isinstance(m, C.Array) else (N.tt_material * len(m))(*m) for m in materials]
        mesh_arr = (C.c_void_p * n)(*[m._h.value for m in meshes])
        bvh_arr = (C.c_void_p * n)(*[b._h.value for b in bvhs])
        mat_arr = (C.c_void_p * n)(*[C.cast(m, C.c_void_p).value for m in self._mats])
        nm = (C.c_int32 * n)(*[len(m) for m in materials])
        env_arr = None
        ew = eh = 0
        if env is not None:
            env_arr = np.ascontiguousarray(env, dtype=np.float32)
            eh, ew = env_arr.shape[:2]
        h = C.c_void_p()
        check(lib.tt_batch_create(ctx._h, n, mesh_arr, bvh_arr, mat_arr, nm,
                                  _vp(env_arr) if env_arr is not None else None, ew, eh, float(env_scale), C.byref(h)))
        self._h = h
        self.n = n
        ctx._adopt(self)

    def render(self, cams, params):
        imgs = [np.empty((c.height, c.width, 3), np.float32) for c in cams]
        cam_arr = (N.tt_camera * len(cams))(*cams)
        ptrs = (C.c_void_p * len(cams))(*[im.ctypes.data for im in imgs])
        st = N.tt_render_stats()
        check(lib.tt_batch_render(self._h, C.cast(cam_arr, C.c_void_p), C.byref(params), ptrs, C.byref(st)))
        return imgs, _stats(st)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.tt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Guide:
    """Path-guiding network (tt_guide): BF16 MFMA MLP -> 8-lobe vMF mixture."""

    def __init__(self, ctx, seed=1, lr=1e-3):
        h = C.c_void_p()
        check(lib.tt_guide_create(ctx._h, int(seed), float(lr), C.byref(h)))
        self._h = h
        self.ctx = ctx
        self.nparams = int(lib.tt_guide_param_count())
        ctx._adopt(self)

    def params(self):
        out = np.empty(self.nparams, np.float32)
        check(lib.tt_guide_get_params(self._h, _vp(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        check(lib.tt_guide_set_params(self._h, _vp(p)))

    def forward(self, inputs):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        out = np.empty((len(x), 40), np.float32)
        check(lib.tt_guide_forward(self._h, _vp(x), len(x), _vp(out)))
        return out

    def forward_unfused(self, inputs):
        """Same outputs as forward() from the training path's separate encoding and
        layer kernels (tt_debug_guide_forward_unfused)."""
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        out = np.empty((len(x), 40), np.float32)
        check(lib.tt_debug_guide_forward_unfused(self._h, _vp(x), len(x), _vp(out)))
        return out

    def infer_slots(self, inputs, guided):
        """The renderer's guided-row inference over n queue slots with these inputs and
        guided flags (tt_debug_guide_infer_slots): (outputs n x 40 with the guided rows
        filled and 0xFFFFFFFF words elsewhere, number of guided rows)."""
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        f = np.ascontiguousarray(guided, np.int32).reshape(-1)
        if len(f) != len(x):
            raise ValueError("one guided flag per input row")
        out = np.empty((len(x), 40), np.float32)
        rows = C.c_int64()
        check(lib.tt_debug_guide_infer_slots(self._h, _vp(x), _vp(f), len(x), _vp(out), C.byref(rows)))
        return out, rows.value

    def grad(self, inputs, dirs, wts):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(wts, np.float32).reshape(-1)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_float()
        check(lib.tt_guide_grad(self._h, _vp(x), _vp(d), _vp(w), len(x), C.byref(loss), _vp(g)))
        return loss.value, g

    def grad_unfused(self, inputs, dirs, wts):
        """grad() through the separate (unfused) training kernels, for checking the fused ones."""
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(wts, np.float32).reshape(-1)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_float()
        check(lib.tt_debug_guide_grad_unfused(self._h, _vp(x), _vp(d), _vp(w), len(x), C.byref(loss), _vp(g)))
        return loss.value, g

    def train_step(self, inputs, dirs, wts):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(wts, np.float32).reshape(-1)
        loss = C.c_float()
        check(lib.tt_guide_train_step(self._h, _vp(x), _vp(d), _vp(w), len(x), C.byref(loss)))
        return loss.value

    def render(self, scene, cam, params, alpha=0.5, train=True, train_paths=0, train_batch=0, train_spp=0,
               min_roughness=0.3, want_fp64=False, max_bounce=0, pass_growth=0, learn_alpha=False):
        """Guided path-traced image of `scene` (tt_guide_render): returns (image, stats,
        guide stats) or (image, image64, stats, guide stats) with want_fp64."""
        gp = N.tt_guide_params(float(alpha), float(min_roughness), int(bool(train)), int(train_paths),
                               int(train_batch), int(train_spp), int(max_bounce), int(pass_growth),
                               int(bool(learn_alpha)))
        out = np.empty((cam.height, cam.width, 3), np.float32)
        out64 = np.empty((cam.height, cam.width, 3), np.float64) if want_fp64 else None
        st = N.tt_render_stats()
        gs = N.tt_guide_stats()
        check(lib.tt_guide_render(scene._h, self._h, C.byref(cam), C.byref(params), C.byref(gp), _vp(out), _vp(out64),
                                  C.byref(st), C.byref(gs)))
        gst = dict(train_steps=gs.train_steps, train_rows=gs.train_rows, train_seconds=gs.train_seconds,
                   last_loss=gs.last_loss, first_loss=gs.first_loss)
        return (out, out64, _stats(st), gst) if want_fp64 else (out, _stats(st), gst)

    def sample(self, raw, n, seed=1):
        """n directions drawn from the mixture of one raw output row (40 floats) by the
        device sampler, with the mixture density at each: (n, 4) float32."""
        r = np.ascontiguousarray(raw, np.float32).reshape(40)
        out = np.empty((int(n), 4), np.float32)
        check(lib.tt_debug_guide_sample(self.ctx._h, _vp(r), int(n), int(seed), _vp(out)))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.tt_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Context:
    """One GPU and its stream (tt_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib.tt_ctx_create(int(device), C.byref(h)))
        self._h = h
        self._children = weakref.WeakSet()  # scenes, buffers, timers: released before the context

    def _adopt(self, obj):
        self._children.add(obj)

    def sync(self):
        check(lib.tt_ctx_sync(self._h))

    @property
    def stream(self):
        return lib.tt_ctx_stream(self._h)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            for c in list(self._children):
                c.close()
            lib.tt_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class DeviceBuffer:
    """Device allocation owned by a Context (tt_dev_alloc)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib.tt_dev_alloc(ctx._h, self.nbytes, C.byref(p)))
        self.ptr = p.value or 0
        ctx._adopt(self)

    @classmethod
    def from_array(cls, ctx, a):
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        check(lib.tt_memcpy_h2d(self.ctx._h, C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data), a.nbytes))

    def download(self, shape, dtype):
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested more bytes than the device buffer holds")
        check(lib.tt_memcpy_d2h(self.ctx._h, C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), out.nbytes))
        return out

    def copy_from(self, other, nbytes=None):
        n = self.nbytes if nbytes is None else int(nbytes)
        check(lib.tt_dev_copy(self.ctx._h, C.c_void_p(self.ptr), C.c_void_p(other.ptr), n))

    def zero(self):
        check(lib.tt_memset_d(self.ctx._h, C.c_void_p(self.ptr), 0, self.nbytes))

    def free(self):
        if getattr(self, "ptr", 0) and self.ctx._h is not None:
            lib.tt_dev_free(self.ctx._h, C.c_void_p(self.ptr))
        self.ptr = 0

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Display:
    """Progressive accumulation and 8-bit display of rendered frames (tt_display_frame):
    the reference's accumulatedBuffer / frameNumber average and gamma-2.2 colour output
    (src/core/RenderKernel.cu:202-217, src/core/Renderer.cpp:86-93). frame(rgb) adds one
    frame (host (h, w, 3) float32 as DeviceScene.render returns it, or a DeviceBuffer of
    npix x 3 floats) and returns the (h, w, 4) uint8 image; reset() restarts the average
    (the reference's bufferReset)."""

    def __init__(self, ctx, width, height):
        self.ctx, self.width, self.height = ctx, int(width), int(height)
        npix = self.width * self.height
        self._frame = DeviceBuffer(ctx, npix * 12)
        self._sum = DeviceBuffer(ctx, npix * 12)
        self._rgba = DeviceBuffer(ctx, npix * 4)
        self.frame_number = 0

    def reset(self):
        self.frame_number = 0

    def frame(self, rgb, stream=None):
        npix = self.width * self.height
        if isinstance(rgb, DeviceBuffer):
            src = rgb
        else:
            a = np.ascontiguousarray(rgb, np.float32)
            if a.size != npix * 3:
                raise ValueError(f"frame has {a.size} floats, expected {npix * 3}")
            self._frame.upload(a)
            src = self._frame
        self.frame_number += 1
        check(lib.tt_display_frame(self.ctx._h, C.c_void_p(src.ptr), C.c_void_p(self._sum.ptr), npix,
                                   self.frame_number, C.c_void_p(self._rgba.ptr), stream))
        return self._rgba.download((self.height, self.width, 4), np.uint8)

    def running_sum(self):
        return self._sum.download((self.height, self.width, 3), np.float32)

    def close(self):
        for b in (self._frame, self._sum, self._rgba):
            b.free()


def gather64_gbps(ctx, buf, nbytes, fetches=1 << 26, seed=1):
    """Random 64-byte gather bandwidth (GB/s) over the first nbytes of a DeviceBuffer."""
    ms = C.c_float()
    check(lib.tt_bench_gather64(ctx._h, C.c_void_p(buf.ptr), int(nbytes), int(fetches), int(seed), C.byref(ms)))
    return fetches * 64 / (ms.value * 1e-3) / 1e9


def debug_math(ctx, op, x, y=None):
    """Device sampler math (op 0 sin, 1 cos, 2 pow(x, y), 3 atan2(x, y), 4 acos) on host arrays,
    through device buffers."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(x if y is None else y, dtype=np.float32)
    bx, by = DeviceBuffer.from_array(ctx, x), DeviceBuffer.from_array(ctx, y)
    bo = DeviceBuffer(ctx, x.nbytes)
    check(lib.tt_debug_math(ctx._h, int(op), C.c_void_p(bx.ptr), C.c_void_p(by.ptr), C.c_void_p(bo.ptr), len(x)))
    out = bo.download(x.shape, np.float32)
    for b in (bx, by, bo):
        b.free()
    return out


def debug_bsdf_sample(ctx, mat, wo, n, front=True, seed=1):
    """n samples of the path-mode BSDF for local wo: (samples n x 8, eval n x 4)."""
    out = np.empty((int(n), 8), np.float32)
    ev = np.empty((int(n), 4), np.float32)
    check(lib.tt_debug_bsdf_sample(ctx._h, C.byref(mat), _f3(wo), int(bool(front)), int(n), int(seed), _vp(out),
                                   _vp(ev)))
    return out, ev


class Timer:
    """HIP events recorded on the context stream (tt_timer)."""

    def __init__(self, ctx, n):
        h = C.c_void_p()
        check(lib.tt_timer_create(ctx._h, int(n), C.byref(h)))
        self._h = h
        ctx._adopt(self)

    def record(self, slot):
        check(lib.tt_timer_record(self._h, int(slot)))

    def elapsed(self, a, b):
        ms = C.c_float()
        check(lib.tt_timer_elapsed(self._h, int(a), int(b), C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.tt_timer_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _dev_ptr(x):
    """Device address of a DeviceBuffer or an int."""
    if isinstance(x, DeviceBuffer):
        return C.c_void_p(x.ptr)
    return C.c_void_p(int(x))


class DeviceScene:
    """Device copy of a built scene (tt_scene): SBVH layout, normals, materials, env map."""

    def __init__(self, ctx, mesh, bvh, materials=None, env=None, env_scale=1.0):
        self.ctx = ctx
        mats = None
        nm = 0
        if isinstance(materials, C.Array) and len(materials):
            nm, mats = len(materials), materials
        elif materials:
            nm = len(materials)
            mats = (N.tt_material * nm)(*materials)
        env_arr = None
        ew = eh = 0
        if env is not None:
            env_arr = np.ascontiguousarray(env, dtype=np.float32)
            eh, ew = env_arr.shape[:2]
        h = C.c_void_p()
        check(lib.tt_scene_create(ctx._h, mesh._h if mesh is not None else None, bvh._h,
                                  C.cast(mats, C.c_void_p) if mats is not None else None, nm,
                                  _vp(env_arr) if env_arr is not None else None, ew, eh, float(env_scale),
                                  C.byref(h)))
        self._h = h
        ctx._adopt(self)

    def info(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib.tt_scene_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"num_nodes": a.value, "num_tris": b.value, "device_bytes": c.value}

    def env_cdf(self, w, h):
        marg = np.empty(h + 1, np.float32)
        cond = np.empty((h, w + 1), np.float32)
        tot = C.c_double()
        check(lib.tt_env_export_cdf(self._h, _vp(marg), _vp(cond), C.byref(tot)))
        return marg, cond, tot.value

    def trace_closest(self, rays, n, hits, cfg=None, stream=None):
        """rays / hits: DeviceBuffers (n x 32 B / n x 16 B) or device addresses; stream: hipStream_t or None."""
        check(lib.tt_trace_closest(self._h, _dev_ptr(rays), int(n), _dev_ptr(hits),
                                   C.byref(cfg) if cfg is not None else None,
                                   C.c_void_p(stream) if stream else None))

    def trace_closest_noalloc(self, rays, n, hits, cfg=None, stream=None):
        """As trace_closest; with cfg.reorder = 1 the reorder workspace is the library's fixed
        device array (no allocation; at most 4,194,304 rays per call)."""
        check(lib.tt_trace_closest_noalloc(self._h, _dev_ptr(rays), int(n), _dev_ptr(hits),
                                           C.byref(cfg) if cfg is not None else None,
                                           C.c_void_p(stream) if stream else None))

    def fetch_replay(self, rays, n, cfg=None, reps=3, want_counts=False):
        """Node-fetch roofline probe (tt_bench_fetch_replay) over device rays: dict with the
        replay's ms, node / triangle-record fetch totals, its bytes and GB/s; with
        want_counts also the per-ray (n, 2) fetch counts."""
        ms = C.c_float()
        nf, tf = C.c_int64(), C.c_int64()
        counts = np.empty((int(n), 2), np.int32) if want_counts else None
        check(lib.tt_bench_fetch_replay(self._h, _dev_ptr(rays), int(n), C.byref(cfg) if cfg is not None else None,
                                        int(reps), C.byref(ms), C.byref(nf), C.byref(tf), _vp(counts)))
        nbytes = 64 * nf.value + 48 * tf.value + 48 * int(n)
        out = {"ms": ms.value, "node_fetches": nf.value, "tri_fetches": tf.value, "bytes": nbytes,
               "gbps": nbytes / (ms.value * 1e-3) / 1e9 if ms.value > 0 else 0.0}
        if want_counts:
            out["counts"] = counts
        return out

    def trace_any(self, rays, n, flags, cfg=None, stream=None):
        check(lib.tt_trace_any(self._h, _dev_ptr(rays), int(n), _dev_ptr(flags),
                               C.byref(cfg) if cfg is not None else None,
                               C.c_void_p(stream) if stream else None))

    def intersect(self, rays, cfg=None):
        """Host arrays in and out: rays (n, 8) float32 -> (tri_id int32, t, u, v float32)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.empty((len(r), 4), np.float32)
        check(lib.tt_trace_closest_host(self._h, _vp(r), len(r), _vp(out), C.byref(cfg) if cfg is not None else None))
        return out[:, 0].view(np.int32).copy(), out[:, 1].copy(), out[:, 2].copy(), out[:, 3].copy()

    def render(self, cam, params, want_fp64=False):
        """Host image (height, width, 3) float32 (+ float64 image when FP64 accumulation is on) and stats."""
        out = np.empty((cam.height, cam.width, 3), np.float32)
        out64 = np.empty((cam.height, cam.width, 3), np.float64) if want_fp64 else None
        st = N.tt_render_stats()
        check(lib.tt_render(self._h, C.byref(cam), C.byref(params), _vp(out), _vp(out64), C.byref(st)))
        return (out, out64, _stats(st)) if want_fp64 else (out, _stats(st))

    def set_geometries(self, geoms):
        """Analytic primitives (list of `sphere` / `plane`) tested after the triangles."""
        arr = (N.tt_geometry * max(1, len(geoms)))(*geoms)
        check(lib.tt_scene_set_geometries(self._h, C.cast(arr, C.c_void_p), len(geoms)))

    def env_invert(self, u):
        """Env texels (row, column) the sampler picks for uniforms u: (n, 2) float32 in [0, 1)."""
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 2)
        out = np.empty((len(u), 2), np.int32)
        check(lib.tt_debug_env_invert(self._h, _vp(u), len(u), _vp(out)))
        return out

    def ray_order(self, d_rays, n, cap=0):
        """The order a reordering trace call traces the rays of a DeviceBuffer in
        (tt_debug_ray_order): int32 array, order[r] = index of the r-th ray. cap > n:
        sorted as a queue of that capacity whose length is read on the device."""
        out = np.empty(int(n), np.int32)
        check(lib.tt_debug_ray_order(self._h, C.c_void_p(d_rays.ptr), int(n), int(cap), _vp(out)))
        return out

    def env_samples(self, n, seed=1):
        out = np.empty((int(n), 8), np.float32)
        check(lib.tt_debug_env_sample(self._h, int(n), int(seed), _vp(out)))
        return out

    def occluded(self, rays, cfg=None):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.empty(len(r), np.int32)
        check(lib.tt_trace_any_host(self._h, _vp(r), len(r), _vp(out), C.byref(cfg) if cfg is not None else None))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.tt_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


# ---------------------------------------------------------------- loaders
def hdr_read(path):
    """Radiance RGBE file -> (height, width, 3) float32, row 0 = top (tt_hdr_read)."""
    w, h = C.c_int32(), C.c_int32()
    check(lib.tt_hdr_read(os.fsencode(path), C.byref(w), C.byref(h), None, 0))
    out = np.empty((h.value, w.value, 3), np.float32)
    check(lib.tt_hdr_read(os.fsencode(path), C.byref(w), C.byref(h), _vp(out), out.size))
    return out


def hdr_write(path, rgb):
    """(height, width, 3) float image -> adaptive-RLE Radiance RGBE file."""
    a = np.ascontiguousarray(rgb, np.float32)
    check(lib.tt_hdr_write(os.fsencode(path), a.shape[1], a.shape[0], _vp(a)))


def obj_load(path):
    """OBJ file (+ its JSON material library) -> (Mesh, [tt_material])."""
    h = C.c_void_p()
    n = C.c_int32()
    check(lib.tt_obj_load(os.fsencode(path), C.byref(h), None, 0, C.byref(n)))
    lib.tt_mesh_destroy(h)
    mats = (N.tt_material * n.value)()
    check(lib.tt_obj_load(os.fsencode(path), C.byref(h), C.cast(mats, C.c_void_p), n.value, C.byref(n)))
    return Mesh(h), list(mats)


class SceneConfig:
    """A reference scene configuration (tt_config): camera, materials, analytic
    geometries and the mesh of its OBJ objects."""

    def __init__(self, path, base_dir=None):
        h = C.c_void_p()
        check(lib.tt_config_load(os.fsencode(path), os.fsencode(base_dir) if base_dir else None, C.byref(h)))
        try:
            info = N.tt_config_info()
            check(lib.tt_config_get(h, C.byref(info), None, None, None))
            mats = (N.tt_material * max(1, info.num_materials))()
            geoms = (N.tt_geometry * max(1, info.num_geometries))()
            mh = C.c_void_p()
            check(lib.tt_config_get(h, None, C.cast(mats, C.c_void_p), C.cast(geoms, C.c_void_p), C.byref(mh)))
        finally:
            lib.tt_config_destroy(h)
        self.info = info
        self.materials = list(mats)[:info.num_materials]
        self.geometries = list(geoms)[:info.num_geometries]
        self.mesh = Mesh(mh) if mh.value else None
        self.width, self.height, self.samples = info.width, info.height, info.samples
        self.hdr_file = info.hdr_file.decode()
        self.work_dir = info.work_dir.decode()

    def camera(self, width=None, height=None):
        """Orbit camera of the configuration (Camera.cpp defaults, fovx 45)."""
        i = self.info
        return camera_from_orbit(tuple(i.position), i.yaw, i.pitch, radius=i.radius, fovx=i.fovx,
                                 width=width or i.width, height=height or i.height, focal=i.focal,
                                 aperture=i.aperture)
