"""ctypes declarations of the C-ABI in include/rtx.h (librtx.so).

Loading fails loudly (RuntimeError) when the built library is missing: there
is no fallback path.
"""
import ctypes
import os
import sys
from ctypes import (POINTER, Structure, Union, byref, c_bool, c_char_p, c_float, c_int, c_int32,
                    c_size_t, c_uint8, c_uint32, c_uint64, c_void_p)

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "librtx.so")


class Float3(Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def __repr__(self):
        return "Float3(%g, %g, %g)" % (self.x, self.y, self.z)


class Uint3(Structure):
    _fields_ = [("x", c_uint32), ("y", c_uint32), ("z", c_uint32)]


class Uchar4(Structure):
    _fields_ = [("x", c_uint8), ("y", c_uint8), ("z", c_uint8), ("w", c_uint8)]


# optx/thing.h:17-45
class Diffuse(Structure):
    _fields_ = [("albedo", Float3)]


class Reflect(Structure):
    _fields_ = [("albedo", Float3), ("fuzz", c_float)]


class Refract(Structure):
    _fields_ = [("index", c_float)]


class _OpticsUnion(Union):
    _fields_ = [("diffuse", Diffuse), ("reflect", Reflect), ("refract", Refract)]


class OpticsC(Structure):
    _anonymous_ = ("_u",)
    _fields_ = [("type", c_int32), ("_u", _OpticsUnion)]


# optx/thing.h:47-52
class ThingC(Structure):
    _fields_ = [("vces", c_void_p), ("ices", c_void_p), ("optics", OpticsC)]


# optx/camera.h:81-93
class CameraC(Structure):
    _fields_ = [("eye", Float3), ("pat", Float3), ("vup", Float3),
                ("fov", c_float), ("aspratio", c_float), ("aperture", c_float), ("fostance", c_float),
                ("u", Float3), ("v", Float3), ("w", Float3),
                ("hvec", Float3), ("wvec", Float3), ("dvec", Float3)]


# optx/rtwo.h:26-52
class LpGeneralC(Structure):
    _fields_ = [("image", c_void_p), ("image_w", c_uint32), ("image_h", c_uint32),
                ("rawRGB", c_void_p), ("normals", c_void_p), ("albedos", c_void_p),
                ("camera", CameraC), ("spp", c_uint32), ("depth", c_uint32),
                ("is_handle", c_uint64), ("rpp", c_void_p),
                ("picker", c_bool), ("pick_x", c_uint32), ("pick_y", c_uint32), ("pick_id", c_void_p)]


class RayC(Structure):
    _fields_ = [("ori", Float3), ("tmin", c_float), ("dir", Float3), ("tmax", c_float)]


class HitC(Structure):
    _fields_ = [("t", c_float), ("u", c_float), ("v", c_float), ("instance_id", c_uint32),
                ("primitive_index", c_uint32), ("prim", c_uint32), ("pad", c_uint32 * 2)]


class SceneStatsC(Structure):
    _fields_ = [("num_prims", c_uint32), ("num_nodes", c_uint32), ("num_leaves", c_uint32),
                ("max_depth", c_uint32), ("max_leaf", c_uint32), ("sah_cost", c_float),
                ("build_ms", c_float), ("device_bytes", c_uint64), ("wide_nodes", c_uint32),
                ("wide_depth", c_uint32), ("two_level", c_uint32), ("unique_prims", c_uint32),
                ("tlas_nodes", c_uint32), ("num_blas", c_uint32)]


class TuningC(Structure):
    _fields_ = [("shade_batch", c_uint32), ("inner_frac", c_uint32), ("tile_w", c_uint32), ("stack_lds", c_uint32),
                ("blocks_per_cu", c_uint32), ("max_blocks_per_cu", c_uint32), ("reserved", c_uint32 * 2)]


# every exported symbol: name -> (restype, argtypes)
SIGNATURES = {
    "rtx_last_error": (c_char_p, []),
    "rtx_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
    "rtx_ctx_destroy": (c_int, [c_void_p]),
    "rtx_camera_set": (c_int, [POINTER(CameraC), Float3, Float3, Float3, c_float, c_float, c_float, c_float]),
    "rtx_sphere_mesh": (c_int, [c_float, c_uint32, c_void_p, c_uint32, POINTER(c_uint32), c_void_p, c_uint32,
                                POINTER(c_uint32)]),
    "rtx_mesh_add": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_uint32, POINTER(c_uint32)]),
    "rtx_thing_add": (c_int, [c_void_p, POINTER(ThingC), c_uint32, POINTER(c_uint32)]),
    "rtx_thing_set": (c_int, [c_void_p, c_uint32, POINTER(c_float)]),
    "rtx_thing_get": (c_int, [c_void_p, c_uint32, POINTER(c_float)]),
    "rtx_thing_at": (c_int, [c_void_p, c_uint32, POINTER(ThingC)]),
    "rtx_thing_count": (c_int, [c_void_p, POINTER(c_uint32)]),
    "rtx_scene_set_build_mode": (c_int, [c_void_p, c_int]),
    "rtx_scene_get_build_mode": (c_int, [c_void_p, POINTER(c_int)]),
    "rtx_scene_build": (c_int, [c_void_p, c_void_p, POINTER(c_uint64)]),
    "rtx_scene_update": (c_int, [c_void_p, c_uint64, c_void_p]),
    "rtx_scene_stats_get": (c_int, [c_void_p, c_uint64, POINTER(SceneStatsC)]),
    "rtx_scene_triangles": (c_int, [c_void_p, c_uint64, c_void_p, c_void_p]),
    "rtx_set_miss_color": (c_int, [c_void_p, Float3]),
    "rtx_set_semantics": (c_int, [c_void_p, c_int]),
    "rtx_get_semantics": (c_int, [c_void_p, POINTER(c_int)]),
    "rtx_tuning_get": (c_int, [c_void_p, POINTER(TuningC)]),
    "rtx_tuning_set": (c_int, [c_void_p, POINTER(TuningC)]),
    "rtx_launch": (c_int, [c_void_p, POINTER(LpGeneralC), c_int, c_void_p]),
    "rtx_launch_noalloc": (c_int, [c_void_p, POINTER(LpGeneralC), c_int, c_void_p]),
    "rtx_scene_set_materials": (c_int, [c_void_p, c_uint64, c_void_p, c_void_p]),
    "rtx_scene_get_materials": (c_int, [c_void_p, c_uint64, c_void_p, c_void_p]),
    "rtx_launch_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "rtx_object_load": (c_int, [c_char_p, POINTER(c_void_p)]),
    "rtx_object_size": (c_int, [c_void_p, POINTER(c_uint32)]),
    "rtx_object_mesh": (c_int, [c_void_p, c_uint32, POINTER(c_void_p), POINTER(c_uint32), POINTER(c_void_p),
                                POINTER(c_uint32)]),
    "rtx_object_free": (c_int, [c_void_p]),
    "rtx_object_write": (c_int, [c_char_p, c_char_p, c_void_p, c_void_p, c_void_p, c_void_p, c_uint32]),
    "rtx_batch_create": (c_int, [c_void_p, c_void_p, c_uint32, c_int, POINTER(c_void_p)]),
    "rtx_batch_launch": (c_int, [c_void_p, c_void_p]),
    "rtx_batch_destroy": (c_int, [c_void_p]),
    "rtx_launch_counted": (c_int, [c_void_p, POINTER(LpGeneralC), c_void_p, c_void_p, c_void_p, c_void_p]),
    "rtx_trace": (c_int, [c_void_p, c_uint64, c_void_p, c_uint32, c_void_p, c_void_p]),
    "rtx_rng_draws": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p]),
    "rtx_pp_srgb": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "rtx_pp_none": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "rtx_malloc": (c_int, [POINTER(c_void_p), c_size_t]),
    "rtx_free": (c_int, [c_void_p]),
    "rtx_memcpy_htod": (c_int, [c_void_p, c_void_p, c_size_t]),
    "rtx_memcpy_dtoh": (c_int, [c_void_p, c_void_p, c_size_t]),
    "rtx_memset": (c_int, [c_void_p, c_int, c_size_t]),
    "rtx_synchronize": (c_int, []),
    "rtx_device_count": (c_int, [POINTER(c_int)]),
}

STATUS = {0: "RTX_SUCCESS", 1: "RTX_ERROR_INVALID_VALUE", 2: "RTX_ERROR_INVALID_HANDLE",
          3: "RTX_ERROR_NOT_BUILT", 4: "RTX_ERROR_OUT_OF_MEMORY", 5: "RTX_ERROR_HIP", 6: "RTX_ERROR_LIMIT"}


class RtxError(RuntimeError):
    """A failed C-ABI call (status code + message); the reference throws
    std::runtime_error from CUDA_CHECK/OPTX_CHECK (optx/util_cpu.h:18-53)."""

    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


_lib = None
_after_torch = False


def lib():
    """The loaded librtx.so with prototypes set (RuntimeError if not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librtx.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        global _after_torch
        # with torch in the process, its bundled HIP runtime must be the one
        # librtx binds to (one runtime, shared device pointers and streams):
        # that holds when torch was imported first
        _after_torch = "torch" in sys.modules
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise RtxError(status, lib().rtx_last_error().decode())
    return status


def call(name, *args):
    """Calls an rtx_* entry point and raises RtxError on a non-zero status."""
    return check(getattr(lib(), name)(*args))


__all__ = ["Float3", "Uint3", "Uchar4", "OpticsC", "ThingC", "CameraC", "LpGeneralC", "RayC", "HitC",
           "SceneStatsC", "TuningC", "SIGNATURES", "RtxError", "lib", "check", "call", "byref", "c_void_p", "c_uint32",
           "c_uint64", "c_float", "c_int"]
