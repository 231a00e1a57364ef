"""The oracle's path loop (orc_render, the composition of the per-material
steps that every pixel-parity test relies on) against images the
mathematics fixes, for both of the reference's program sets: the recursive
programs (optx/camera_r.cu + optx/optics_r.cu) and the iterative ones
(optx/camera_i.cu + optx/optics_i.cu).  The scenes and closed forms are in
cases.py; the GPU runs the same checks in test_gpu_closed_forms.py.

What each case pins: the miss program and the ray count (empty scene); black
at the depth limit versus the kept throughput (depth 1, mirror + sphere at
depth 2); the Lambertian direction distribution, albedo and sky together
(ground at depth 2); the mirror direction and the metal's albedo (fuzz-0
mirror); the dielectric's side test, the 1/ior ratio on entry and Schlick's
reflectance (glass slab at 60 degrees: a swapped ratio reflects every ray,
a wrong power or cosine moves the reflected share); the guide values of both
program sets, including a black albedo behind a mirror."""
import numpy as np
import pytest

import cases as C


def _render(orc, parts, cam, spp, depth, iterative=False, ambient=C.AMBIENT):
    sc = C.oracle_scene(orc, parts, iterative, ambient)
    ocam = C.oracle_camera(orc, cam)
    r = sc.render(ocam, cam["w"], cam["h"], spp, depth, nthreads=8, want_sumsq=True)
    r["rays0"] = orc.primary_rays(ocam, cam["w"], cam["h"]).reshape(cam["h"], cam["w"], 8)
    return r


@pytest.mark.parametrize("iterative", [False, True])
def test_empty_scene_is_the_sky(orc, iterative):
    parts, cam = C.empty_scene()
    r = _render(orc, parts, cam, 1, 16, iterative)
    want = C.sky64(r["rays0"][..., 4:7])
    assert (r["rpp"] == 1).all()
    assert np.abs(r["raw"] - want).max() < 2e-6
    r4 = _render(orc, parts, cam, 4, 16, iterative)
    assert (r4["rpp"] == 4).all()
    assert (r4["normals"] == 0).all() and (r4["albedos"] == 0).all()


def test_depth_one_hit_is_black_recursive_and_albedo_iterative(orc):
    parts, cam = C.ground_scene()
    r = _render(orc, parts, cam, 2, 1, iterative=False)
    assert (r["rpp"] == 2).all()
    assert (r["raw"] == 0).all()                          # optx/optics_r.cu:31-42
    assert (r["albedos"] == 0).all()                      # the limit's hit sets no guide
    r = _render(orc, parts, cam, 2, 1, iterative=True)
    assert (r["rpp"] == 2).all()
    assert (r["raw"] == np.array(C.GROUND_ALBEDO, np.float32)).all()   # optx/camera_i.cu:75-95
    assert (r["albedos"] == np.array(C.GROUND_ALBEDO, np.float32)).all()
    assert np.abs(r["normals"] - np.array([0, 1, 0])).max() < 1e-6


@pytest.mark.parametrize("depth", [2, 3])
def test_lambertian_ground_gives_the_cosine_weighted_sky_mean(orc, depth):
    parts, cam = C.ground_scene()
    spp = 64
    r = _render(orc, parts, cam, spp, depth)
    assert (r["rpp"] == 2 * spp).all()                   # every bounce leaves to the sky
    got = r["raw"].reshape(-1, 3).astype(np.float64).mean(axis=0)
    want = C.lambert_sky_mean(C.GROUND_ALBEDO)
    sig = C.mean_sigma(r["sumsq"], r["raw"], spp)
    print("ground mean", got, "closed form", want, "sigma", sig)
    assert (np.abs(got - want) <= 3 * sig + 1e-6).all()   # blue: amb.z = 1, every sample is exactly a.z


def test_fuzz0_mirror_reflects_the_sky(orc):
    parts, cam = C.mirror_scene()
    r = _render(orc, parts, cam, 1, 2)
    assert (r["rpp"] == 2).all()
    d = r["rays0"][..., 4:7]
    n = C.MIRROR_NORMAL                                  # facing the rays (d . n < 0)
    want = np.asarray(C.MIRROR_ALBEDO) * C.sky64(C.reflect64(d, n))
    assert np.abs(r["raw"] - want).max() < 1e-5
    assert np.abs(r["normals"] - n).max() < 1e-6
    assert (r["albedos"] == np.array(C.MIRROR_ALBEDO, np.float32)).all()


def test_fuzzy_mirror_absorbs_the_fuzzed_directions_below_it(orc):
    """With a white sky every sample is the mirror's albedo or black
    (absorbed, optx/optics_r.cu:242): the mean is albedo x (1 - the share of
    the fuzz ball below the surface)."""
    parts, cam = C.fuzzy_mirror_scene()
    spp = 256
    r = _render(orc, parts, cam, spp, 2, ambient=(1., 1., 1.))
    d = r["rays0"][..., 4:7].astype(np.float64)
    n = C.rot_y(C.FUZZ_DEG) @ np.array([0, 0, 1.0])
    cos_r = -((d / np.linalg.norm(d, axis=-1, keepdims=True)) @ n)   # dot(mirror direction, N) = -dot(unit d, N)
    Pa = C.absorbed_share(cos_r, C.FUZZ)
    alb = np.asarray(C.MIRROR_ALBEDO)
    exp = alb * (1 - Pa.mean())
    got = r["raw"].astype(np.float64).reshape(-1, 3).mean(axis=0)
    sig = alb * np.sqrt((Pa * (1 - Pa)).sum() / spp) / Pa.size
    print("fuzzy mirror mean", got, "closed form", exp, "sigma", sig, "absorbed share", Pa.mean())
    assert 0.2 < Pa.mean() < 0.3
    assert (np.abs(got - exp) <= 3 * sig).all()


def _slab_expectation(r, cam):
    d = r["rays0"][..., 4:7].astype(np.float64)
    u = d / np.linalg.norm(d, axis=-1, keepdims=True)
    n = C.rot_y(C.SLAB_DEG) @ np.array([0, 0, 1.0])     # front face, facing the camera
    cos_t = -(u @ n)
    P = C.schlick64(cos_t, 1 / 1.5)                       # entering: ratio 1/ior
    return P, C.sky64(C.reflect64(d, n))


@pytest.mark.parametrize("iterative", [False, True])
def test_glass_slab_enters_with_one_over_ior(orc, iterative):
    """Depth 2: a sample reflected at the front face reaches the sky; a
    refracted one hits the back face with the last ray (black in the
    recursive programs, white throughput in the iterative ones)."""
    parts, cam = C.slab_scene()
    spp = 256
    r = _render(orc, parts, cam, spp, 2, iterative)
    assert (r["rpp"] == 2 * spp).all()
    P, sky = _slab_expectation(r, cam)
    want = P[..., None] * sky + ((1 - P)[..., None] if iterative else 0)
    got = r["raw"].astype(np.float64).reshape(-1, 3).mean(axis=0)
    exp = want.reshape(-1, 3).mean(axis=0)
    sig = np.sqrt((P * (1 - P))[..., None] * (sky - (1 if iterative else 0)) ** 2 / spp).reshape(-1, 3)
    sig = np.sqrt((sig ** 2).sum(axis=0)) / sig.shape[0]
    print("slab mean", got, "closed form", exp, "sigma", sig, "reflected share", P.mean())
    assert 0.06 < P.mean() < 0.08
    assert (np.abs(got - exp) <= 3 * sig).all()


@pytest.fixture(scope="module")
def sphere3(orc):
    return orc.sphere_mesh(1.0, 3)


def test_guides_and_depth_limit_behind_a_mirror(orc, sphere3):
    """Mirror, then a diffuse sphere: the central pixels' samples all take
    that path, so the guide values and the depth-limit colours are exact."""
    Ad = (.6, .5, .4)
    Am = np.array(C.MIRROR_ALBEDO, np.float32)
    parts, cam = C.mirror_sphere_scene(sphere3, Ad)
    m = C.central_pixels(cam, 3.5)
    assert m.sum() > 400
    # recursive, depth 3: the deepest hit's (the sphere's) albedo
    r = _render(orc, parts, cam, 2, 3)
    assert (r["rpp"][m] >= 4).all()
    assert (r["albedos"][m] == np.array(Ad, np.float32)).all()
    # iterative, depth 3: the first hit's (the mirror's) albedo and normal
    r = _render(orc, parts, cam, 2, 3, iterative=True)
    assert (r["albedos"][m] == Am).all()
    assert np.abs(r["normals"][m] - C.MIRROR_NORMAL).max() < 1e-6
    # depth 2: the sphere is hit by the last ray: black, or the throughput
    r = _render(orc, parts, cam, 2, 2)
    assert (r["rpp"][m] == 4).all() and (r["raw"][m] == 0).all()
    r = _render(orc, parts, cam, 2, 2, iterative=True)
    assert (r["rpp"][m] == 4).all()
    assert (r["raw"][m] == Am * np.array(Ad, np.float32)).all()
    # a black sphere: the recursive programs keep the deepest NON-ZERO albedo
    parts, cam = C.mirror_sphere_scene(sphere3, (0, 0, 0))
    r = _render(orc, parts, cam, 2, 3)
    assert (r["albedos"][m] == Am).all()
    assert (r["raw"][m] == 0).all()
