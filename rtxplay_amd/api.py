"""Host-side mirror of the reference's hot-path API over the C-ABI.

Names and argument meaning follow the reference classes:
  Camera     optx/camera.h:28-94        (set + per-field setters that re-run set)
  Thing      optx/thing.h:47-52         (vces/ices filled by Scene.add)
  Sphere     optx/sphere.h:20-38        (radius, ndiv -> indexed mesh)
  Scene      optx/scene.h:19-63         (load/add/set/get/build/update/[])
  LpGeneral  optx/rtwo.h:26-52          (launch parameters)
  Launcher   optx/launcher.h:17-31      (resize, ignite(stream, once))
  pp_sRGB / pp_none                     optx/postproc.cu:49-61
Failures raise RtxError (a RuntimeError), as CUDA_CHECK/OPTX_CHECK throw in the
reference (optx/util_cpu.h:18-53).
"""
import ctypes

import numpy as np

from ._capi import (CameraC, Float3, HitC, LpGeneralC, OpticsC, RayC, RtxError, SceneStatsC, ThingC, TuningC, byref,
                    c_float, c_int, c_uint32, c_uint64, c_void_p, call, lib)

TYPE_DIFFUSE, TYPE_REFLECT, TYPE_REFRACT, TYPE_NUM = 0, 1, 2, 3


def _f3(v):
    return Float3(float(v[0]), float(v[1]), float(v[2]))


# ---------------------------------------------------------------------------
# device memory
# ---------------------------------------------------------------------------
class DeviceBuffer:
    """A device allocation (rtx_malloc) with numpy upload/download."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = c_void_p()
        call("rtx_malloc", byref(p), self.nbytes)
        self._ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @property
    def ptr(self):
        return self._ptr.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than the buffer")
        call("rtx_memcpy_htod", self._ptr, a.ctypes.data_as(c_void_p), a.nbytes)

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("download larger than the buffer")
        call("rtx_memcpy_dtoh", out.ctypes.data_as(c_void_p), self._ptr, out.nbytes)
        return out

    def zero(self):
        call("rtx_memset", self._ptr, 0, self.nbytes)

    def free(self):
        if self._ptr.value:
            call("rtx_free", self._ptr)
            self._ptr = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    call("rtx_synchronize")


def _stream_ptr(stream):
    if stream is None:
        return None
    return getattr(stream, "ptr", stream)


# ---------------------------------------------------------------------------
# Camera: optx/camera.h:28-94
# ---------------------------------------------------------------------------
class Camera:
    def __init__(self):
        self.c = CameraC()

    def set(self, eye, pat, vup, fov, aspratio, aperture, fostance):
        call("rtx_camera_set", byref(self.c), _f3(eye), _f3(pat), _f3(vup), float(fov), float(aspratio),
             float(aperture), float(fostance))
        return self

    def _reset(self, **kw):
        c = self.c
        args = dict(eye=tuple(c.eye), pat=tuple(c.pat), vup=tuple(c.vup), fov=c.fov, aspratio=c.aspratio,
                    aperture=c.aperture, fostance=c.fostance)
        args.update(kw)
        self.set(**args)

    # accessors re-run set(), optx/camera.h:50-62
    def eye(self, v=None):
        return tuple(self.c.eye) if v is None else self._reset(eye=v)

    def pat(self, v=None):
        return tuple(self.c.pat) if v is None else self._reset(pat=v)

    def vup(self, v=None):
        return tuple(self.c.vup) if v is None else self._reset(vup=v)

    def fov(self, v=None):
        return self.c.fov if v is None else self._reset(fov=v)

    def aspratio(self, v):
        self._reset(aspratio=v)

    def aperture(self, v=None):
        return self.c.aperture if v is None else self._reset(aperture=v)

    def fostance(self, v=None):
        return self.c.fostance if v is None else self._reset(fostance=v)


# ---------------------------------------------------------------------------
# Thing / Optics: optx/thing.h:17-52
# ---------------------------------------------------------------------------
class Thing:
    def __init__(self, optics_type=TYPE_DIFFUSE, albedo=(0.5, 0.5, 0.5), fuzz=0.0, index=1.5):
        self.c = ThingC()
        self.c.optics.type = optics_type
        if optics_type == TYPE_DIFFUSE:
            self.c.optics.diffuse.albedo = _f3(albedo)
        elif optics_type == TYPE_REFLECT:
            self.c.optics.reflect.albedo = _f3(albedo)
            self.c.optics.reflect.fuzz = float(fuzz)
        elif optics_type == TYPE_REFRACT:
            self.c.optics.refract.index = float(index)
        else:
            raise ValueError("unknown optics type")

    @classmethod
    def diffuse(cls, albedo):
        return cls(TYPE_DIFFUSE, albedo=albedo)

    @classmethod
    def reflect(cls, albedo, fuzz):
        return cls(TYPE_REFLECT, albedo=albedo, fuzz=fuzz)

    @classmethod
    def refract(cls, index):
        return cls(TYPE_REFRACT, index=index)

    @property
    def optics(self):
        return self.c.optics

    @property
    def vces(self):
        return self.c.vces

    @property
    def ices(self):
        return self.c.ices


# ---------------------------------------------------------------------------
# Sphere: optx/sphere.h:20-38
# ---------------------------------------------------------------------------
class Sphere:
    """Tessellated sphere: a tetrahedron subdivided ndiv times (host)."""

    def __init__(self, radius=1.0, ndiv=6):
        nv, ni = c_uint32(0), c_uint32(0)
        call("rtx_sphere_mesh", float(radius), int(ndiv), None, 0, byref(nv), None, 0, byref(ni))
        v = np.empty((nv.value, 3), np.float32)
        i = np.empty((ni.value, 3), np.uint32)
        call("rtx_sphere_mesh", float(radius), int(ndiv), v.ctypes.data_as(c_void_p), nv.value, byref(nv),
             i.ctypes.data_as(c_void_p), ni.value, byref(ni))
        self.vces = v[:nv.value]
        self.ices = i[:ni.value]

    def mesh(self):
        """(vertices float32[nv,3], indices uint32[ni,3]) -- the Mesh tuple of optx/sphere.h:18"""
        return self.vces, self.ices


class Object:
    """A list of submeshes (optx/object.h:20-33); Scene.add concatenates them.
    Object(meshes) takes (vertices, indices) pairs; Object(path) reads a
    Wavefront OBJ file (Object(wavefront), optx/object.cxx:23-93)."""

    def __init__(self, meshes):
        if isinstance(meshes, (str, bytes)):
            meshes = read_obj(meshes)
        self.meshes = list(meshes)

    def __len__(self):
        return len(self.meshes)

    def __getitem__(self, m):
        return self.meshes[m]


# ---------------------------------------------------------------------------
# Scene: optx/scene.h:19-63
# ---------------------------------------------------------------------------
class Scene:
    """Owns meshes, instances and the device acceleration structure."""

    def __init__(self, device=0):
        p = c_void_p()
        call("rtx_ctx_create", int(device), byref(p))
        self.ctx = p
        self._things = []

    def close(self):
        if self.ctx.value:
            call("rtx_ctx_destroy", self.ctx)
            self.ctx = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self):
        """Fills the scene; returns the number of things (optx/scene.h:26)."""
        raise NotImplementedError

    def add(self, obj, object_id=None):
        """add(Object|mesh) -> mesh id (optx/scene.cxx:38-167);
        add(Thing, mesh id) -> thing id (optx/scene.cxx:169-196)."""
        if isinstance(obj, Thing):
            tid = c_uint32(0)
            call("rtx_thing_add", self.ctx, byref(obj.c), int(object_id), byref(tid))
            return tid.value
        if isinstance(obj, Object):
            vs, is_, off = [], [], 0
            for v, i in obj.meshes:
                vs.append(np.asarray(v, np.float32).reshape(-1, 3))
                is_.append(np.asarray(i, np.uint32).reshape(-1, 3) + np.uint32(off))
                off += len(vs[-1])
            v = np.concatenate(vs) if vs else np.zeros((0, 3), np.float32)
            i = np.concatenate(is_) if is_ else np.zeros((0, 3), np.uint32)
        else:
            v, i = obj
            v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
            i = np.ascontiguousarray(i, np.uint32).reshape(-1, 3)
        mid = c_uint32(0)
        call("rtx_mesh_add", self.ctx, v.ctypes.data_as(c_void_p), len(v), i.ctypes.data_as(c_void_p), len(i),
             byref(mid))
        return mid.value

    def set(self, thing, transform):
        """3x4 row-major object-to-world transform (optx/scene.cxx:198-215)."""
        t = (c_float * 12)(*[float(x) for x in np.asarray(transform, np.float32).reshape(12)])
        try:
            call("rtx_thing_set", self.ctx, int(thing), t)
        except RtxError:
            return False
        return True

    def get(self, thing):
        t = (c_float * 12)()
        try:
            call("rtx_thing_get", self.ctx, int(thing), t)
        except RtxError:
            return None
        return np.array(list(t), np.float32)

    def __getitem__(self, i):
        t = ThingC()
        call("rtx_thing_at", self.ctx, int(i), byref(t))
        return t

    def __len__(self):
        n = c_uint32(0)
        call("rtx_thing_count", self.ctx, byref(n))
        return n.value

    FLAT, TWO_LEVEL = 0, 1

    def build_mode(self, mode=None):
        """Reads (and with an argument sets) the layout of the next build:
        Scene.FLAT (one BVH over world-space triangles, the default) or
        Scene.TWO_LEVEL (per-mesh BVHs under an instance BVH, as the
        reference's IAS over GASes, optx/scene.cxx:38-273)."""
        if mode is not None:
            call("rtx_scene_set_build_mode", self.ctx, int(mode))
        m = c_int(0)
        call("rtx_scene_get_build_mode", self.ctx, byref(m))
        return m.value

    def build(self, stream=None):
        """Device build; returns the traversable handle (optx/scene.cxx:227-273)."""
        h = c_uint64(0)
        call("rtx_scene_build", self.ctx, _stream_ptr(stream), byref(h))
        self.handle = h.value
        return h.value

    def update(self, handle, stream=None):
        """Refreshes the structure after set(); the handle stays valid
        (optx/scene.cxx:275-294).  Two-level scenes refit their top level."""
        call("rtx_scene_update", self.ctx, c_uint64(handle), _stream_ptr(stream))
        return handle

    def stats(self, handle=None):
        st = SceneStatsC()
        call("rtx_scene_stats_get", self.ctx, c_uint64(handle or self.handle), byref(st))
        return {k: getattr(st, k) for k, _ in SceneStatsC._fields_}

    def triangles(self, handle=None):
        """World-space triangles float32[N,3,3] and instance ids uint32[N] (original order)."""
        n = self.stats(handle)["num_prims"]
        tris = np.empty((n, 3, 3), np.float32)
        inst = np.empty(n, np.uint32)
        call("rtx_scene_triangles", self.ctx, c_uint64(handle or self.handle), tris.ctypes.data_as(c_void_p),
             inst.ctypes.data_as(c_void_p))
        return tris, inst

    def tuning(self, **kw):
        """Reads (and with keyword arguments sets) the launch tuning, e.g. shade_batch=40."""
        t = TuningC()
        call("rtx_tuning_get", self.ctx, byref(t))
        if kw:
            for k, v in kw.items():
                setattr(t, k, int(v))
            call("rtx_tuning_set", self.ctx, byref(t))
        return {k: getattr(t, k) for k in ("shade_batch", "inner_frac", "tile_w", "stack_lds", "blocks_per_cu",
                                           "max_blocks_per_cu")}

    def set_materials(self, params, handle=None, stream=None):
        """Material parameters float32[n,4] (albedo rgb, fuzz per thing) into the
        built scene; params: a numpy array, or an int device pointer."""
        h = c_uint64(handle or self.handle)
        if isinstance(params, (int, np.integer)):
            call("rtx_scene_set_materials", self.ctx, h, c_void_p(int(params)), _stream_ptr(stream))
            return
        p = np.ascontiguousarray(params, np.float32).reshape(-1, 4)
        if len(p) != len(self):
            raise RtxError(1, "set_materials: %d rows for %d things" % (len(p), len(self)))
        d = DeviceBuffer.from_array(p)
        call("rtx_scene_set_materials", self.ctx, h, d.ptr, _stream_ptr(stream))
        synchronize()
        d.free()

    def get_materials(self, handle=None):
        """float32[n,4] material parameters of the built scene (dielectric: 0,0,0,ior)."""
        n = len(self)
        d = DeviceBuffer(16 * max(n, 1))
        call("rtx_scene_get_materials", self.ctx, c_uint64(handle or self.handle), d.ptr, None)
        synchronize()
        out = d.download(np.float32, (n, 4))
        d.free()
        return out

    def set_miss_color(self, ambient):
        call("rtx_set_miss_color", self.ctx, _f3(ambient))

    RECURSIVE, ITERATIVE = 0, 1

    def semantics(self, mode=None):
        """Reads (and with an argument sets) which program set launches follow:
        Scene.RECURSIVE (optx/camera_r.cu + optx/optics_r.cu, the default
        here) or Scene.ITERATIVE (optx/camera_i.cu + optx/optics_i.cu, the
        reference's default build, optx/rtwo.cxx:44-50)."""
        if mode is not None:
            call("rtx_set_semantics", self.ctx, int(mode))
        m = c_int(0)
        call("rtx_get_semantics", self.ctx, byref(m))
        return m.value

    def trace(self, rays, handle=None, stream=None):
        """Closest hits for rays float32[n,8] (ori, tmin, dir, tmax) -> structured hits."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        dr = DeviceBuffer.from_array(rays)
        dh = DeviceBuffer(32 * max(n, 1))
        call("rtx_trace", self.ctx, c_uint64(handle or self.handle), dr.ptr, n, dh.ptr, _stream_ptr(stream))
        synchronize()
        dt = np.dtype([("t", np.float32), ("u", np.float32), ("v", np.float32), ("instance_id", np.uint32),
                       ("primitive_index", np.uint32), ("prim", np.uint32), ("pad", np.uint32, 2)])
        out = dh.download(np.uint8, 32 * n).view(dt)
        dr.free()
        dh.free()
        return out


# ---------------------------------------------------------------------------
# LpGeneral + Launcher: optx/rtwo.h:26-52, optx/launcher.cxx
# ---------------------------------------------------------------------------
class LpGeneral(LpGeneralC):
    pass


class Launcher:
    """Owns rawRGB/rpp/normals/albedos/pick_id and writes their pointers into
    the caller's LpGeneral (optx/launcher.cxx:24-47)."""

    def __init__(self, scene, lp):
        self.scene = scene
        self.lp = lp
        self._bufs = {}
        self._npix = 0
        self.resize(lp.image_w, lp.image_h)

    def resize(self, w, h):
        self.free()
        n = int(w) * int(h)
        self._npix = n
        self._bufs = {"rawRGB": DeviceBuffer(12 * n), "rpp": DeviceBuffer(4 * n), "normals": DeviceBuffer(12 * n),
                      "albedos": DeviceBuffer(12 * n), "pick_id": DeviceBuffer(4)}
        for k, b in self._bufs.items():
            setattr(self.lp, k, b.ptr)

    def ignite(self, stream=None, once=False, sync=True):
        """optixLaunch replacement; synchronous like the reference unless sync=False."""
        if not once and self.lp.image_w * self.lp.image_h > self._npix:
            raise RtxError(1, "image larger than the Launcher's buffers (resize first)")
        call("rtx_launch", self.scene.ctx, byref(self.lp), 1 if once else 0, _stream_ptr(stream))
        if sync:
            synchronize()

    def ignite_counted(self, stream=None, cycles=False):
        """Render once more with traversal counters: (node visits, triangle tests) per pixel
        [, (traversal cycles, loop cycles, inner-node wave iterations, triangle wave iterations)]."""
        n = self.lp.image_w * self.lp.image_h
        dn, dt, dc = DeviceBuffer(8 * n), DeviceBuffer(8 * n), DeviceBuffer(32)
        dc.zero()
        call("rtx_launch_counted", self.scene.ctx, byref(self.lp), dn.ptr, dt.ptr, dc.ptr, _stream_ptr(stream))
        synchronize()
        shape = (self.lp.image_h, self.lp.image_w)
        out = dn.download(np.uint64, shape), dt.download(np.uint64, shape)
        cyc = tuple(int(x) for x in dc.download(np.uint64, (4,)))
        dn.free()
        dt.free()
        dc.free()
        return out + (cyc,) if cycles else out

    def adjoint(self, dldi, stream=None):
        """dL/d material params float32[n,4] given dL/d rawRGB float32[h,w,3];
        the Launcher's rawRGB must hold the forward image of the same lp."""
        g = np.ascontiguousarray(dldi, np.float32).reshape(self.lp.image_h, self.lp.image_w, 3)
        n = len(self.scene)
        dg = DeviceBuffer.from_array(g)
        dp = DeviceBuffer(16 * max(n, 1))
        call("rtx_launch_adjoint", self.scene.ctx, byref(self.lp), dg.ptr, dp.ptr, _stream_ptr(stream))
        synchronize()
        out = dp.download(np.float32, (n, 4))
        dg.free()
        dp.free()
        return out

    def free(self):
        for b in self._bufs.values():
            b.free()
        self._bufs = {}

    def raw_rgb(self):
        return self._bufs["rawRGB"].download(np.float32, (self.lp.image_h, self.lp.image_w, 3))

    def rays_per_pixel(self):
        return self._bufs["rpp"].download(np.uint32, (self.lp.image_h, self.lp.image_w))

    def normals(self):
        return self._bufs["normals"].download(np.float32, (self.lp.image_h, self.lp.image_w, 3))

    def albedos(self):
        return self._bufs["albedos"].download(np.float32, (self.lp.image_h, self.lp.image_w, 3))

    def pick_id(self):
        return int(self._bufs["pick_id"].download(np.uint32, (1,))[0])


class Batch:
    """Many frames (each its own LpGeneral: scene handle, camera, size, spp,
    depth, output buffers) rendered by one persistent launch; with
    use_graph the first launch on a stream is captured into a hipGraph and
    replayed after that.  Every frame equals Launcher.ignite of its LpGeneral.
    The frames' buffers stay owned by the caller (e.g. one Launcher each)."""

    def __init__(self, scene, lps, use_graph=True):
        self.scene = scene
        arr = (LpGeneralC * len(lps))()
        for i, lp in enumerate(lps):
            ctypes.memmove(ctypes.byref(arr, i * ctypes.sizeof(LpGeneralC)), ctypes.byref(lp),
                           ctypes.sizeof(LpGeneralC))
        p = c_void_p()
        call("rtx_batch_create", scene.ctx, arr, len(lps), 1 if use_graph else 0, byref(p))
        self.ptr = p

    def launch(self, stream=None, sync=True):
        call("rtx_batch_launch", self.ptr, _stream_ptr(stream))
        if sync:
            synchronize()

    def free(self):
        if self.ptr is not None and self.ptr.value:
            call("rtx_batch_destroy", self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def read_obj(path):
    """[(vertices float32[nv,3], indices uint32[ni,3]), ...], one per OBJ shape."""
    o = c_void_p()
    call("rtx_object_load", path.encode() if isinstance(path, str) else path, byref(o))
    try:
        n = c_uint32(0)
        call("rtx_object_size", o, byref(n))
        out = []
        for m in range(n.value):
            pv, pi, nv, ni = c_void_p(), c_void_p(), c_uint32(0), c_uint32(0)
            call("rtx_object_mesh", o, m, byref(pv), byref(nv), byref(pi), byref(ni))
            v = np.ctypeslib.as_array(ctypes.cast(pv, ctypes.POINTER(c_float)), (nv.value * 3,)).reshape(-1, 3).copy()
            i = np.ctypeslib.as_array(ctypes.cast(pi, ctypes.POINTER(c_uint32)), (ni.value * 3,)).reshape(-1, 3).copy()
            out.append((v, i))
        return out
    finally:
        call("rtx_object_free", o)


def write_obj(path, meshes, name="object"):
    """Writes (vertices, indices) meshes as one OBJ shape (optx/sphere.cxx:108-140)."""
    vs = [np.ascontiguousarray(v, np.float32).reshape(-1, 3) for v, _ in meshes]
    is_ = [np.ascontiguousarray(i, np.uint32).reshape(-1, 3) for _, i in meshes]
    n = len(vs)
    pv = (c_void_p * max(n, 1))(*[x.ctypes.data for x in vs])
    pi = (c_void_p * max(n, 1))(*[x.ctypes.data for x in is_])
    nv = (c_uint32 * max(n, 1))(*[len(x) for x in vs])
    ni = (c_uint32 * max(n, 1))(*[len(x) for x in is_])
    call("rtx_object_write", path.encode() if isinstance(path, str) else path, name.encode(), pv, nv, pi, ni, n)


def pp_sRGB(src, dst, w, h, stream=None):
    """optx/postproc.cu:56-61 (device pointers)"""
    call("rtx_pp_srgb", c_void_p(src), c_void_p(dst), int(w), int(h), _stream_ptr(stream))


def pp_none(src, dst, w, h, stream=None):
    """optx/postproc.cu:49-54 (device pointers)"""
    call("rtx_pp_none", c_void_p(src), c_void_p(dst), int(w), int(h), _stream_ptr(stream))


def rng_draws(seeds, k):
    """k device Frand48 draws for each seed -> float32[n,k]"""
    seeds = np.ascontiguousarray(seeds, np.uint32)
    n = len(seeds)
    ds = DeviceBuffer.from_array(seeds)
    do = DeviceBuffer(4 * n * k)
    call("rtx_rng_draws", ds.ptr, n, int(k), do.ptr, None)
    synchronize()
    out = do.download(np.float32, (n, k))
    ds.free()
    do.free()
    return out


def device_count():
    n = ctypes.c_int(0)
    call("rtx_device_count", byref(n))
    return n.value


__all__ = ["Camera", "Thing", "Sphere", "Object", "Scene", "LpGeneral", "Launcher", "Batch", "DeviceBuffer", "pp_sRGB",
           "read_obj", "write_obj",
           "pp_none", "rng_draws", "device_count", "synchronize", "TYPE_DIFFUSE", "TYPE_REFLECT", "TYPE_REFRACT",
           "RtxError", "lib", "HitC", "RayC", "OpticsC"]
