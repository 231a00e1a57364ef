"""ctypes binding of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker (or the timed CPU baseline).  See
oracle.h for what each function restates and how it is pinned.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

ORC_DIFFUSE, ORC_REFLECT, ORC_REFRACT = 0, 1, 2


class Thing(ctypes.Structure):
    _fields_ = [("type", c_int32), ("albedo", c_float * 3), ("fuzz", c_float), ("index", c_float),
                ("center", c_float * 3), ("radius", c_float), ("kind", c_int32)]


class CameraC(ctypes.Structure):
    _fields_ = [("eye", c_float * 3), ("pat", c_float * 3), ("vup", c_float * 3), ("fov", c_float),
                ("aspratio", c_float), ("aperture", c_float), ("fostance", c_float), ("u", c_float * 3),
                ("v", c_float * 3), ("w", c_float * 3), ("hvec", c_float * 3), ("wvec", c_float * 3),
                ("dvec", c_float * 3)]


class RtowOpts(ctypes.Structure):
    _fields_ = [("fp32", c_int), ("nthreads", c_int), ("argorder", c_int), ("seed_base", c_uint32)]


class SceneC(ctypes.Structure):
    _fields_ = [("tris", c_void_p), ("tri_inst", c_void_p), ("ntris", c_uint32), ("things", c_void_p),
                ("nthings", c_uint32), ("bvh", c_void_p), ("ambient", c_float * 3), ("iterative", c_int32)]


_lib = None


def build():
    """Compiles liboracle.so with the committed Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = c_void_p
        sig = {
            "orc_frand48_draws": (None, [P, c_uint32, c_uint32, P]),
            "orc_camera_set": (None, [POINTER(CameraC), P, P, P, c_float, c_float, c_float, c_float]),
            "orc_camera_ray": (None, [POINTER(CameraC), c_float, c_float, P, P, P]),
            "orc_primary_rays": (None, [POINTER(CameraC), c_uint32, c_uint32, P]),
            "orc_sphere_mesh": (c_int, [c_float, c_uint32, P, POINTER(c_uint32), P, POINTER(c_uint32)]),
            "orc_reduce": (c_int, [P, c_uint32, P, POINTER(c_uint32), P]),
            "orc_scene_rtow": (c_uint32, [c_uint32, c_int32, P, c_uint32]),
            "orc_transform_point": (None, [P, P, P]),
            "orc_closest_f64_brute": (None, [P, c_uint32, P, c_uint32, P, P, P, P, P, c_int]),
            "orc_bvh_build": (c_void_p, [P, c_uint32]),
            "orc_bvh_free": (None, [P]),
            "orc_bvh_node_count": (c_uint32, [P]),
            "orc_closest_f64_bvh": (None, [P, P, c_uint32, P, P, P, P, P, c_int]),
            "orc_closest_f32_bvh": (None, [P, P, c_uint32, P, P, P, P, c_int]),
            "orc_decidable": (c_int, [c_double, c_double, c_double, c_double, c_double, c_double]),
            "orc_render": (c_uint64, [POINTER(SceneC), POINTER(CameraC), c_uint32, c_uint32, c_uint32, c_uint32,
                                      c_uint32, c_uint32, c_int, P, P, P, P, P, c_uint32]),
            "orc_render_frozen": (None, [POINTER(SceneC), POINTER(CameraC), c_uint32, c_uint32, c_uint32, c_uint32,
                                         c_uint32, c_uint32, c_int, P, P]),
            "orc_rtow_render": (c_uint64, [c_int, c_int, c_int, c_int, c_int, P]),
            "orc_rtow_rays": (c_uint64, [c_int, c_int, c_int, c_int, c_int, c_uint64]),
            "orc_rtow_scene": (c_uint32, [c_int, P, c_uint32]),
            "orc_rtow_render_ex": (c_uint64, [POINTER(RtowOpts), c_int, c_int, c_int, c_int, P, P, P, P]),
            "orc_pp_srgb": (None, [P, P, c_int, c_int]),
            "orc_pp_srgb_f64": (None, [P, P, c_int, c_int]),
            "orc_pp_none": (None, [P, P, c_int, c_int]),
            "orc_reflect": (None, [P, P, P]),
            "orc_refract": (None, [P, P, c_float, P]),
            "orc_schlick": (c_float, [c_float, c_float]),
            "orc_sky": (None, [P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(c_void_p)


def frand48_draws(seeds, k):
    seeds = np.ascontiguousarray(seeds, np.uint32)
    out = np.empty((len(seeds), k), np.float32)
    lib().orc_frand48_draws(_p(seeds), len(seeds), k, _p(out))
    return out


def camera(eye, pat, vup, fov, aspratio, aperture, fostance):
    c = CameraC()
    f = lambda v: np.ascontiguousarray(v, np.float32)
    e, p, u = f(eye), f(pat), f(vup)
    lib().orc_camera_set(ctypes.byref(c), _p(e), _p(p), _p(u), fov, aspratio, aperture, fostance)
    return c


def primary_rays(cam, w, h):
    rays = np.empty((h * w, 8), np.float32)
    lib().orc_primary_rays(ctypes.byref(cam), w, h, _p(rays))
    return rays


def sphere_mesh(radius, ndiv):
    nv_cap = 2 * 4 ** ndiv + 2
    ni_cap = 4 * 4 ** ndiv
    v = np.empty((nv_cap, 3), np.float32)
    i = np.empty((ni_cap, 3), np.uint32)
    nv, ni = c_uint32(0), c_uint32(0)
    rc = lib().orc_sphere_mesh(radius, ndiv, _p(v), ctypes.byref(nv), _p(i), ctypes.byref(ni))
    assert rc == 0
    return v[:nv.value].copy(), i[:ni.value].copy()


def reduce(vtmp):
    vtmp = np.ascontiguousarray(vtmp, np.float32).reshape(-1, 3)
    v = np.empty_like(vtmp)
    idx = np.empty(len(vtmp), np.uint32)
    nv = c_uint32(0)
    lib().orc_reduce(_p(vtmp), len(vtmp), _p(v), ctypes.byref(nv), _p(idx))
    return v[:nv.value].copy(), idx.reshape(-1, 3)


def scene_rtow(seed=0, half=11):
    cap = 4 * half * half + 8
    arr = (Thing * cap)()
    n = lib().orc_scene_rtow(seed, half, arr, cap)
    return [arr[i] for i in range(n)]


def rtow_scene_glibc(argorder=1):
    arr = (Thing * 1024)()
    n = lib().orc_rtow_scene(argorder, arr, 1024)
    return [arr[i] for i in range(n)]


def transform_point(xf, v):
    xf = np.ascontiguousarray(xf, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    out = np.empty(3, np.float32)
    lib().orc_transform_point(_p(xf), _p(v), _p(out))
    return out


def flatten(meshes, things, levels):
    """World-space soup float32[N,9] and instance ids for spheres at levels
    (list per thing) from meshes {level: (v, i)}; transform per thing is
    scale radius + translate center, applied with orc_transform_point."""
    tris, inst = [], []
    for k, (t, lv) in enumerate(zip(things, levels)):
        v, i = meshes[lv]
        r = t["radius"] if isinstance(t, dict) else t.radius
        c = t["center"] if isinstance(t, dict) else list(t.center)
        xf = np.array([r, 0, 0, c[0], 0, r, 0, c[1], 0, 0, r, c[2]], np.float32)
        w = transform_points(xf, v)
        tris.append(w[i].reshape(-1, 9))
        inst.append(np.full(len(i), k, np.uint32))
    return np.ascontiguousarray(np.concatenate(tris)), np.concatenate(inst)


def transform_points(xf, v):
    """orc_transform_point semantics, vectorised in numpy float32 (same
    operation order, one rounding per operation)."""
    xf = np.asarray(xf, np.float32)
    v = np.asarray(v, np.float32)
    out = np.empty_like(v)
    for r in range(3):
        acc = v[:, 0] * xf[4 * r]
        acc = acc + v[:, 1] * xf[4 * r + 1]
        acc = acc + v[:, 2] * xf[4 * r + 2]
        out[:, r] = acc + xf[4 * r + 3]
    return out


def closest_f64_brute(tris, rays, nthreads=8):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = len(rays)
    t, u, v, t2 = (np.empty(n, np.float64) for _ in range(4))
    prim = np.empty(n, np.uint32)
    lib().orc_closest_f64_brute(_p(tris), len(tris), _p(rays), n, _p(t), _p(prim), _p(u), _p(v), _p(t2),
                                nthreads)
    return dict(t=t, prim=prim, u=u, v=v, t2=t2)


class BVH:
    def __init__(self, tris):
        self.tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self.ptr = lib().orc_bvh_build(_p(self.tris), len(self.tris))

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_bvh_free(self.ptr)
            self.ptr = None

    def closest_f64(self, rays, nthreads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        t, u, v, t2 = (np.empty(n, np.float64) for _ in range(4))
        prim = np.empty(n, np.uint32)
        lib().orc_closest_f64_bvh(self.ptr, _p(rays), n, _p(t), _p(prim), _p(u), _p(v), _p(t2), nthreads)
        return dict(t=t, prim=prim, u=u, v=v, t2=t2)

    def closest_f32(self, rays, nthreads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        t, u, v = (np.empty(n, np.float32) for _ in range(3))
        prim = np.empty(n, np.uint32)
        lib().orc_closest_f32_bvh(self.ptr, _p(rays), n, _p(t), _p(prim), _p(u), _p(v), nthreads)
        return dict(t=t, prim=prim, u=u, v=v)


def decidable(res, rays, tris, k=64.0):
    """Rays whose FP64 closest hit is far enough from every ambiguity for an
    exact primitive-id comparison: the hit point is at least delta from each
    edge of its triangle, the next distinct hit and tmin are at least delta
    away along the ray, with delta = k * 2^-24 * max(|ori|, |hit|, 1) in world
    units (k float32 ulps of the coordinate magnitude).  Misses are decidable
    when no triangle lies within delta of the ray."""
    rays = np.asarray(rays, np.float32).reshape(-1, 8).astype(np.float64)
    tris = np.asarray(tris, np.float32).reshape(-1, 9).astype(np.float64)
    o, d = rays[:, 0:3], rays[:, 4:7]
    t, prim = res["t"], res["prim"]
    hit = np.isfinite(t)
    dn = np.linalg.norm(d, axis=1)
    ok = np.ones(len(t), bool)
    tt = np.where(hit, t, 0.0)
    p = o + tt[:, None] * d
    mag = np.maximum(np.maximum(np.abs(o).max(axis=1), np.abs(p).max(axis=1)), 1.0)
    delta = k * 2.0 ** -24 * mag
    hi = np.where(hit)[0]
    if len(hi):
        T = tris[prim[hi]].reshape(-1, 3, 3)
        A, B, C = T[:, 0], T[:, 1], T[:, 2]
        P = p[hi]

        def edge_dist(P, X, Y):
            e = Y - X
            n2 = (e * e).sum(axis=1)
            s = np.clip(((P - X) * e).sum(axis=1) / np.where(n2 > 0, n2, 1.0), 0.0, 1.0)
            q = X + s[:, None] * e
            return np.linalg.norm(P - q, axis=1)

        de = np.minimum(np.minimum(edge_dist(P, A, B), edge_dist(P, B, C)), edge_dist(P, C, A))
        gap = (res["t2"][hi] - t[hi]) * dn[hi]
        tgap = (t[hi] - rays[hi, 3]) * dn[hi]
        ok[hi] = (de >= delta[hi]) & (gap >= delta[hi]) & (tgap >= delta[hi])
    return ok


class Scene:
    """Flattened scene for orc_render: soup + per-tri instance + things + BVH."""

    def __init__(self, tris, tri_inst, things, ambient=(.5, .7, 1.), iterative=False):
        self.tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self.tri_inst = np.ascontiguousarray(tri_inst, np.uint32)
        self.things = (Thing * len(things))()
        for k, t in enumerate(things):
            T = self.things[k]
            if isinstance(t, dict):
                T.type = int(t["type"])
                T.albedo[:] = [float(x) for x in t["albedo"]]
                T.fuzz = float(t["fuzz"]); T.index = float(t["index"])
                T.center[:] = [float(x) for x in t["center"]]
                T.radius = float(t["radius"]); T.kind = int(t["kind"])
            else:
                ctypes.memmove(ctypes.byref(T), ctypes.byref(t), ctypes.sizeof(Thing))
        self.bvh = BVH(self.tris)
        self.c = SceneC(_p(self.tris).value, _p(self.tri_inst).value, len(self.tris),
                        ctypes.cast(self.things, c_void_p).value, len(things), self.bvh.ptr,
                        (c_float * 3)(*ambient), 1 if iterative else 0)

    def render(self, cam, w, h, spp, depth, y0=0, y1=None, nthreads=8, want_sumsq=False, seed_base=0):
        y1 = h if y1 is None else y1
        raw = np.zeros((h, w, 3), np.float32)
        rpp = np.zeros((h, w), np.uint32)
        nrm = np.zeros((h, w, 3), np.float32)
        alb = np.zeros((h, w, 3), np.float32)
        sq = np.zeros((h, w, 3), np.float32) if want_sumsq else None
        rays = lib().orc_render(ctypes.byref(self.c), ctypes.byref(cam), w, h, spp, depth, y0, y1, nthreads,
                                _p(raw), _p(rpp), _p(nrm), _p(alb), _p(sq) if sq is not None else None,
                                seed_base)
        return dict(raw=raw, rpp=rpp, normals=nrm, albedos=alb, sumsq=sq, rays=rays)


    def materials(self):
        """float32[n,4]: albedo rgb and fuzz of every thing"""
        m = np.zeros((len(self.things), 4), np.float32)
        for k in range(len(self.things)):
            m[k, :3] = list(self.things[k].albedo)
            m[k, 3] = self.things[k].fuzz
        return m

    def render_frozen(self, cam, w, h, spp, depth, params, y0=0, y1=None, nthreads=8):
        """Paths traced with the scene's own materials, radiance evaluated in
        FP64 with params (float32[n,4]) along the same vertices -> float64[h,w,3]."""
        y1 = h if y1 is None else y1
        params = np.ascontiguousarray(params, np.float32).reshape(-1, 4)
        assert len(params) == len(self.things)
        raw = np.zeros((h, w, 3), np.float64)
        lib().orc_render_frozen(ctypes.byref(self.c), ctypes.byref(cam), w, h, spp, depth, y0, y1, nthreads,
                                _p(params), _p(raw))
        return raw


def rtow_render(w=1280, h=720, spp=10, depth=50, argorder=1):
    out = np.empty((h, w, 3), np.uint8)
    rays = lib().orc_rtow_render(w, h, spp, depth, argorder, _p(out))
    return out, rays


def rtow_render_ex(w, h, spp, depth, fp32=False, nthreads=0, seed_base=0, argorder=1, rgb=False):
    """The rtow.cxx tracer with options (orc_rtow_render_ex): nthreads 0 =
    the reference's glibc stream on one thread, else per-pixel drand48
    streams.  Returns dict(mean=float64[h,w,3] (y = 0 the bottom row),
    sumsq, rpp=uint32[h,w], rays, rgb=uint8[h,w,3] in PPM order or None)."""
    o = RtowOpts(1 if fp32 else 0, int(nthreads), int(argorder), int(seed_base))
    mean = np.zeros((h, w, 3), np.float64)
    sq = np.zeros((h, w, 3), np.float64)
    rpp = np.zeros((h, w), np.uint32)
    out = np.zeros((h, w, 3), np.uint8) if rgb else None
    rays = lib().orc_rtow_render_ex(ctypes.byref(o), w, h, spp, depth, _p(out) if rgb else None, _p(mean), _p(sq),
                                    _p(rpp))
    return dict(mean=mean, sumsq=sq, rpp=rpp, rays=int(rays), rgb=out)


def rtow_rays(npix, w=1280, h=720, spp=10, depth=50, argorder=1):
    """Rays the reference tracer casts for its first npix output pixels."""
    return int(lib().orc_rtow_rays(w, h, spp, depth, argorder, int(npix)))


def pp_srgb(raw):
    raw = np.ascontiguousarray(raw, np.float32)
    h, w = raw.shape[:2]
    out = np.empty((h, w, 4), np.uint8)
    lib().orc_pp_srgb(_p(raw), _p(out), w, h)
    return out


def pp_srgb_f64(raw):
    raw = np.ascontiguousarray(raw, np.float32)
    h, w = raw.shape[:2]
    out = np.empty((h, w, 3), np.float64)
    lib().orc_pp_srgb_f64(_p(raw), _p(out), w, h)
    return out


def pp_none(raw):
    raw = np.ascontiguousarray(raw, np.float32)
    h, w = raw.shape[:2]
    out = np.empty((h, w, 4), np.uint8)
    lib().orc_pp_none(_p(raw), _p(out), w, h)
    return out


def reflect(d, n):
    d = np.ascontiguousarray(d, np.float32); n = np.ascontiguousarray(n, np.float32)
    out = np.empty(3, np.float32)
    lib().orc_reflect(_p(d), _p(n), _p(out))
    return out


def refract(d, n, ratio):
    d = np.ascontiguousarray(d, np.float32); n = np.ascontiguousarray(n, np.float32)
    out = np.empty(3, np.float32)
    lib().orc_refract(_p(d), _p(n), ratio, _p(out))
    return out


def schlick(cos_t, ratio):
    return lib().orc_schlick(cos_t, ratio)


def sky(d, ambient=(.5, .7, 1.)):
    d = np.ascontiguousarray(d, np.float32); a = np.ascontiguousarray(ambient, np.float32)
    out = np.empty(3, np.float32)
    lib().orc_sky(_p(d), _p(a), _p(out))
    return out
