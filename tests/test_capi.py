"""librtx.so without a GPU: it loads, exports every entry point of
include/rtx.h, and its host-side functions (camera, tessellation, argument
checks) agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "rtx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:rtx_status|const char\*)\s+(rtx_\w+)\s*\(", txt, re.M)))


def test_exports_every_declared_symbol(rtx):
    from rtxplay_amd._capi import SIGNATURES
    names = _declared()
    assert len(names) >= 25
    L = ctypes.CDLL(rtx.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(SIGNATURES), set(names) ^ set(SIGNATURES)


def test_structure_sizes_match_c_layout(rtx):
    from rtxplay_amd import _capi as C
    # optx/thing.h: int + union{float3,float} -> 20 B; Thing = 2 pointers + Optics (padded) -> 40 B
    assert ctypes.sizeof(C.OpticsC) == 20 and ctypes.sizeof(C.ThingC) == 40
    assert ctypes.sizeof(C.CameraC) == 4 * (9 + 4 + 18)
    assert ctypes.sizeof(C.RayC) == 32 and ctypes.sizeof(C.HitC) == 32
    # LpGeneral: field offsets as a C compiler lays out optx/rtwo.h:26-52
    off = {n: getattr(C.LpGeneralC, n).offset for n, _ in C.LpGeneralC._fields_}
    assert off["image_w"] == 8 and off["rawRGB"] == 16 and off["camera"] == 40
    assert off["is_handle"] % 8 == 0 and off["pick_x"] == off["picker"] + 4


def test_camera_set_matches_oracle(rtx, orc):
    args = ((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, 1280 / 720, 0.1, 10.0)
    a = rtx.Camera().set(*args).c
    b = orc.camera(*args)
    for k in ("u", "v", "w", "hvec", "wvec", "dvec"):
        assert np.array_equal(np.array(list(getattr(a, k)), np.float32), np.array(list(getattr(b, k)), np.float32)), k


def test_camera_setters_rerun_set(rtx):
    c = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, 16 / 9, 0.1, 10.0)
    before = tuple(c.c.hvec)
    c.fov(40.0)
    assert c.fov() == 40.0 and tuple(c.c.hvec) != before


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 6])
def test_sphere_mesh_matches_oracle(rtx, orc, n):
    v, i = rtx.Sphere(1.0, n).mesh()
    ov, oi = orc.sphere_mesh(1.0, n)
    assert np.array_equal(v.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(i, oi)


def test_sphere_mesh_limits(rtx):
    from rtxplay_amd._capi import RtxError, c_uint32, byref, call
    nv, ni = c_uint32(0), c_uint32(0)
    with pytest.raises(RtxError) as e:
        call("rtx_sphere_mesh", 1.0, 12, None, 0, byref(nv), None, 0, byref(ni))
    assert e.value.status == 6
    v = np.empty((3, 3), np.float32)
    with pytest.raises(RtxError) as e:   # capacity too small
        call("rtx_sphere_mesh", 1.0, 1, v.ctypes.data, 3, byref(nv), v.ctypes.data, 3, byref(ni))
    assert e.value.status == 1


def test_null_arguments_are_errors(rtx):
    L = rtx.lib()
    assert L.rtx_ctx_create(0, None) == 1
    assert b"NULL" in L.rtx_last_error()
    assert L.rtx_launch(None, None, 0, None) == 1
    assert L.rtx_trace(None, 0, None, 0, None, None) == 1
    assert L.rtx_pp_srgb(None, None, -1, 1, None) == 1
    assert L.rtx_camera_set(None, *([rtx._capi.Float3()] * 3), 1.0, 1.0, 0.0, 1.0) == 1
    assert L.rtx_scene_set_build_mode(None, 1) == 1
    assert L.rtx_scene_get_build_mode(None, None) == 1
    assert L.rtx_batch_create(None, None, 0, 0, None) == 1
    assert L.rtx_batch_launch(None, None) == 1
    assert L.rtx_object_load(None, None) == 1


def test_build_modes_are_the_documented_values(rtx):
    """RTX_BUILD_FLAT = 0 and RTX_BUILD_TWO_LEVEL = 1 in include/rtx.h, as
    the Python mirror's Scene.FLAT / Scene.TWO_LEVEL."""
    txt = open(os.path.join(ROOT, "include", "rtx.h")).read()
    m = re.search(r"enum\s*\{\s*RTX_BUILD_FLAT\s*=\s*(\d+)\s*,\s*RTX_BUILD_TWO_LEVEL\s*=\s*(\d+)\s*\}", txt)
    assert m and (int(m.group(1)), int(m.group(2))) == (rtx.Scene.FLAT, rtx.Scene.TWO_LEVEL) == (0, 1)
    from rtxplay_amd import _capi as C
    names = [n for n, _ in C.SceneStatsC._fields_]
    assert names[-4:] == ["two_level", "unique_prims", "tlas_nodes", "num_blas"]
    assert ctypes.sizeof(C.SceneStatsC) == 7 * 4 + 4 + 8 + 6 * 4


def test_unknown_handle_rejected_without_gpu(rtx):
    """rtx_launch validates the traversable handle before touching the device."""
    from rtxplay_amd._capi import LpGeneralC
    L = rtx.lib()
    lp = LpGeneralC()
    lp.is_handle = 0xDEADBEEF
    # a non-NULL but bogus context pointer is never dereferenced before the handle check
    assert L.rtx_launch(ctypes.c_void_p(1), ctypes.byref(lp), 0, None) == 2
    assert b"is_handle" in L.rtx_last_error()


def test_ctx_create_without_device_fails_cleanly(rtx):
    if rtx.device_count() > 0:
        pytest.skip("a GPU is present")
    p = ctypes.c_void_p()
    st = rtx.lib().rtx_ctx_create(0, ctypes.byref(p))
    assert st != 0 and p.value is None
