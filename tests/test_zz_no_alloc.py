"""rtx_launch_noalloc through the C-ABI: the launch whose scratch lives in the
library's static device array renders the same pixels as rtx_launch, on two
calls in a row (a frame in probe order, which uses every part of the
scratch, then a smaller frame in slot order), and rejects a frame beyond the
static tile capacity before launching anything."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def scene(rtx):
    from rtxplay_amd.scene import RTOWScene
    s = RTOWScene(seed=0, half=11, level=3)
    s.load()
    say("build level 3")
    s.build()
    yield s
    s.close()


def _lp(rtx, s, w, h, spp, depth=16):
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, spp, depth, s.handle
    lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
    return lp


def _outputs(rtx, lp):
    """Device output buffers (pre-filled with a sentinel) wired into lp."""
    w, h = lp.image_w, lp.image_h
    bufs = {k: rtx.DeviceBuffer.from_array(np.full((h, w, 3), np.nan, np.float32))
            for k in ("rawRGB", "normals", "albedos")}
    bufs["rpp"] = rtx.DeviceBuffer.from_array(np.full((h, w), 0xFFFFFFFF, np.uint32))
    for k, b in bufs.items():
        setattr(lp, k, b.ptr)
    return bufs


def _read(bufs, w, h):
    out = tuple(bufs[k].download(np.float32, (h, w, 3)) for k in ("rawRGB", "normals", "albedos"))
    return out + (bufs["rpp"].download(np.uint32, (h, w)),)


def _free(bufs):
    for b in bufs.values():
        b.free()


def test_noalloc_matches_launch_on_two_calls(rtx, scene):
    from rtxplay_amd._capi import call
    frames = [(1280, 720, 16), (200, 113, 5)]       # probe order (2+ pixels per lane), then smaller, slot order
    refs = []
    for w, h, spp in frames:
        lp = _lp(rtx, scene, w, h, spp)
        bufs = _outputs(rtx, lp)
        say("rtx_launch %dx%d spp %d" % (w, h, spp))
        call("rtx_launch", scene.ctx, ctypes.byref(lp), 0, None)
        rtx.synchronize()
        refs.append(_read(bufs, w, h))
        _free(bufs)
    # the two no-allocation calls in a row, the second with the smaller frame
    gots = []
    for w, h, spp in frames:
        lp = _lp(rtx, scene, w, h, spp)
        bufs = _outputs(rtx, lp)
        say("rtx_launch_noalloc %dx%d spp %d" % (w, h, spp))
        call("rtx_launch_noalloc", scene.ctx, ctypes.byref(lp), 0, None)
        rtx.synchronize()
        gots.append(_read(bufs, w, h))
        _free(bufs)
    for (w, h, spp), ref, got in zip(frames, refs, gots):
        say("frame %dx%d: rays %d vs %d" % (w, h, int(got[3].sum()), int(ref[3].sum())))
        for a, b in zip(got, ref):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert (got[3] >= spp).all()


def test_noalloc_picker_matches_launch(rtx, scene):
    from rtxplay_amd._capi import call
    w, h = 320, 180
    ids = {}
    for entry in ("rtx_launch", "rtx_launch_noalloc"):
        got = []
        pick = rtx.DeviceBuffer(4)
        for x, y in ((160, 90), (10, 170), (300, 40)):
            lp = _lp(rtx, scene, w, h, 1)
            lp.picker, lp.pick_x, lp.pick_y, lp.pick_id = True, x, y, pick.ptr
            call(entry, scene.ctx, ctypes.byref(lp), 1, None)
            rtx.synchronize()
            got.append(int(pick.download(np.uint32, 1)[0]))
        pick.free()
        ids[entry] = got
    say("picked", ids)
    assert ids["rtx_launch"] == ids["rtx_launch_noalloc"]


def test_noalloc_rejects_frames_beyond_the_static_tiles(rtx, scene):
    """4200 x 4200 at 16 spp renders in probe order with 525 x 525 tiles of
    8 x 8, more than the 2^18 the static scratch holds: the call fails with
    RTX_ERROR_LIMIT before launching (the output buffer is real all the same)."""
    from rtxplay_amd._capi import RtxError, call
    w, h = 4200, 4200
    lp = _lp(rtx, scene, w, h, 16)
    raw = rtx.DeviceBuffer(12 * w * h)
    lp.rawRGB = raw.ptr
    with pytest.raises(RtxError) as e:
        call("rtx_launch_noalloc", scene.ctx, ctypes.byref(lp), 0, None)
    rtx.synchronize()
    say("rejected:", e.value)
    assert e.value.status == 6
    raw.free()
