"""rtxplay_amd -- MI355X (gfx950) ray-tracing hot path of RTXplay.

The product is librtx.so (hand-written HIP kernels behind the C-ABI in
include/rtx.h).  This package is its host-side mirror of the reference API
(Scene, Thing, Camera, Sphere, LpGeneral, Launcher, pp_sRGB) over ctypes.
"""
from ._capi import RtxError, lib, LIB_PATH  # noqa: F401
from .api import (Batch, Camera, DeviceBuffer, Launcher, LpGeneral, Object, Scene, Sphere, Thing,  # noqa: F401
                  TYPE_DIFFUSE, TYPE_REFLECT, TYPE_REFRACT, device_count, pp_none, pp_sRGB, read_obj, rng_draws, write_obj,
                  synchronize)
