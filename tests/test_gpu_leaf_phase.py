"""The render kernels' leaf phase deals the (lane, triangle) pairs of all
parked leaves out over the wave and keeps, per lane, the minimum of
(t, global primitive id) (leaves_compact in csrc/traverse.h).  Coincident
copies of one sphere make every hit on it an exact tie in t, between copies
whose triangles share leaves of up to 8 (the builder cannot split equal
centroids), so the leaf phases carry many pairs per lane: the closest hit must
go to the lowest primitive id, the first copy, as the per-lane loop and the
oracle decide it (optx/camera_r.cu:82-99 asks OptiX for the closest hit; the
tie rule is this build's, shared by the oracle)."""
import numpy as np
import pytest

import cases as C

pytestmark = pytest.mark.gpu

COLOURS = [(.9, .1, .1), (.1, .1, .9), (.1, .9, .1), (.9, .9, .1)]


def say(*a):
    print("[gpu-leaf]", *a, flush=True)


def _parts(rtx, copies, level):
    v, i = rtx.Sphere(1.0, level).mesh()
    xf = C.xform(t=(0, 0, -3))
    return [C.part((v, i), xf, C.DIFFUSE, COLOURS[k]) for k in range(copies)]


def _first_copy(albedos):
    """First-hit guides (iterative semantics): a pixel's albedo is the mean
    over its samples of the first hit's albedo (0 on a miss), so if every hit
    on the sphere went to the first (red) copy, each pixel is a multiple of
    (.9, .1, .1): green equals blue and red is nine times green."""
    a = albedos.reshape(-1, 3).astype(np.float64)
    seen = a[:, 0] > 0
    assert seen.mean() > 0.2
    assert np.abs(a[:, 1] - a[:, 2]).max() < 1e-6
    assert np.allclose(a[seen, 0], 9 * a[seen, 1], rtol=1e-5)


def _gpu(rtx, parts, two_level=False):
    s = rtx.Scene()
    if two_level:
        s.build_mode(s.TWO_LEVEL)
    s.set_miss_color(C.AMBIENT)
    for p in parts:
        mid = s.add(p["mesh"])
        tid = s.add(rtx.Thing(p["type"], albedo=p["albedo"], fuzz=p["fuzz"], index=p["index"]), mid)
        s.set(tid, p["xf"])
    s.semantics(s.ITERATIVE)
    s.build()
    return s


@pytest.mark.parametrize("copies,level", [(2, 3), (4, 2)])
def test_coincident_copies_hit_the_first(rtx, orc, copies, level):
    parts = _parts(rtx, copies, level)
    cam = C.camera(160, 120, (0, 0, 0), (0, 0, -1), 40)
    spp, depth = 4, 8
    say("%d coincident level-%d spheres, %dx%d spp %d" % (copies, level, cam["w"], cam["h"], spp))
    s = _gpu(rtx, parts)
    st = s.stats()
    say("leaves", st["num_leaves"], "largest", st["max_leaf"], "triangles", st["num_prims"])
    assert st["max_leaf"] >= copies
    g = C.gpu_render(rtx, s, cam, spp, depth)
    s.close()
    o = C.oracle_scene(orc, parts, iterative=True).render(C.oracle_camera(orc, cam), cam["w"], cam["h"], spp, depth,
                                                            nthreads=8)
    for k in ("raw", "normals", "albedos"):
        assert np.array_equal(g[k].view(np.uint32), o[k].view(np.uint32)), k
    assert np.array_equal(g["rpp"], o["rpp"])
    _first_copy(g["albedos"])


def test_coincident_instances_hit_the_first(rtx):
    """The two-level layout: exact ties go to the lower instance."""
    parts = _parts(rtx, 4, 2)
    cam = C.camera(160, 120, (0, 0, 0), (0, 0, -1), 40)
    s = _gpu(rtx, parts, two_level=True)
    g = C.gpu_render(rtx, s, cam, 4, 8)
    s.close()
    _first_copy(g["albedos"])
