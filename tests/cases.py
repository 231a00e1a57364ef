"""Scenes whose images the mathematics fixes, shared by the oracle tests (CPU)
and the GPU tests, and the closed forms they are checked against.

Each scene is a list of parts (an object-space mesh, a 3x4 object-to-world
transform and a material) plus a camera.  The oracle takes the parts as a
world-space soup (the transform applied as the device flatten does it); the
GPU takes them through the C-ABI (rtx_mesh_add, rtx_thing_add,
rtx_thing_set).  A dielectric's centre is its transform's translation, as
the side test of optx/optics_r.cu:376-384 takes it.

The closed forms are FP64 evaluations of the reference's definitions: the
sky of the miss program (optx/camera_r.cu:135-140), the mirror direction
(optx/optics_r.cu:235-236), Schlick's reflectance (optx/optics_r.cu:387-388)
and the cosine-weighted mean of the sky over a Lambertian bounce
(optx/optics_r.cu:108: the direction N + a point on the unit sphere has the
density cos(theta)/pi, so E[unit(d).y] = 2/3 for an upward N).
"""
import math

import numpy as np

F = np.float32
DIFFUSE, REFLECT, REFRACT = 0, 1, 2
AMBIENT = (.5, .7, 1.)


# ---- meshes and transforms --------------------------------------------------
def quad(half=1.0):
    """Two triangles spanning [-half, half]^2 in the z = 0 plane."""
    v = np.array([[-half, -half, 0], [half, -half, 0], [half, half, 0], [-half, half, 0]], F)
    i = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return v, i


def slab(half=20.0, thick=1.0):
    """A box of [-half, half]^2 x [-thick, thick]: front and back faces only
    (z = +thick and z = -thick), 4 triangles."""
    v = np.array([[-half, -half, thick], [half, -half, thick], [half, half, thick], [-half, half, thick],
                  [-half, -half, -thick], [half, -half, -thick], [half, half, -thick], [-half, half, -thick]], F)
    i = np.array([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6]], np.uint32)
    return v, i


def rot_y(deg):
    a = math.radians(deg)
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


def rot_x(deg):
    a = math.radians(deg)
    return np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])


def xform(R=None, t=(0, 0, 0), s=1.0):
    """3x4 row-major [s R | t] as float32[12]."""
    R = np.eye(3) if R is None else np.asarray(R, np.float64)
    M = np.zeros((3, 4))
    M[:, :3] = s * R
    M[:, 3] = t
    return M.astype(F).reshape(12)


def part(mesh, xf, typ, albedo=(0, 0, 0), fuzz=0.0, index=1.5):
    return dict(mesh=mesh, xf=np.asarray(xf, F), type=typ, albedo=tuple(F(a) for a in albedo), fuzz=F(fuzz),
                index=F(index))


def camera(w, h, eye, pat, fov, vup=(0, 1, 0), aperture=0.0, fostance=10.0):
    return dict(w=w, h=h, eye=eye, pat=pat, vup=vup, fov=float(fov), aperture=float(aperture),
                fostance=float(fostance))


# ---- the scenes ---------------------------------------------------------------
def empty_scene():
    """One small triangle far behind the camera: every ray misses."""
    parts = [part(quad(.01), xform(t=(0, 0, 1000)), DIFFUSE, (.5, .5, .5))]
    return parts, camera(64, 48, (0, 0, 0), (0, 0, -1), 40)


GROUND_ALBEDO = (.5, .4, .3)


def ground_scene():
    """A flat diffuse ground y = 0 under a camera that looks down at it:
    every primary ray hits it, every bounce leaves it upwards to the sky."""
    parts = [part(quad(50.0), xform(R=rot_x(-90)), DIFFUSE, GROUND_ALBEDO)]
    return parts, camera(64, 48, (0, 1, 0), (0, 0, -1), 40)


MIRROR_ALBEDO = (.9, .8, .7)
MIRROR_NORMAL = np.array([1, 0, 1]) / math.sqrt(2)


def mirror_scene():
    """A fuzz-0 metal quad at 45 degrees in front of the camera; the
    reflected rays leave to the sky."""
    parts = [part(quad(8.0), xform(R=rot_y(45), t=(0, 0, -10)), REFLECT, MIRROR_ALBEDO, fuzz=0.0)]
    return parts, camera(64, 48, (0, 0, 0), (0, 0, -10), 20)


FUZZ, FUZZ_DEG = 0.5, 80.0


def fuzzy_mirror_scene():
    """A metal quad of fuzz 0.5 turned by 80 degrees about y in front of a
    narrow camera: the rays meet it at about 80 degrees of incidence, and a
    share of the fuzzed directions point into it (absorbed: black)."""
    parts = [part(quad(8.0), xform(R=rot_y(FUZZ_DEG), t=(0, 0, -10)), REFLECT, MIRROR_ALBEDO, fuzz=FUZZ)]
    return parts, camera(32, 32, (0, 0, 0), (0, 0, -10), 2)


def absorbed_share(cos_r, fuzz):
    """P(dot(r + fuzz u, N) <= 0) for u uniform in the unit ball and a unit
    mirror direction r with dot(r, N) = cos_r: the ball's cap below
    -cos_r / fuzz, of height h = 1 - cos_r / fuzz, holds h^2 (3 - h) / 4 of
    its volume."""
    h = np.clip(1 - np.asarray(cos_r, np.float64) / fuzz, 0, 2)
    return h * h * (3 - h) / 4


SLAB_DEG = 60.0


def slab_scene():
    """A glass slab (ior 1.5, 2 units thick) turned by 60 degrees about y in
    front of a narrow camera: every primary ray enters its front face at
    about 60 degrees of incidence, near the slab's centre."""
    parts = [part(slab(20.0, 1.0), xform(R=rot_y(SLAB_DEG), t=(0, 0, -10)), REFRACT, index=1.5)]
    return parts, camera(32, 32, (0, 0, 0), (0, 0, -10), 2)


SPHERE_CENTRE = (6.0, 0.0, -10.0)


def mirror_sphere_scene(sphere_mesh, diffuse_albedo):
    """The fuzz-0 mirror of mirror_scene reflecting the camera's central rays
    into a diffuse sphere (radius 1.5) that the camera does not see directly:
    the mirror image of the eye is (-10, 0, -10), from where the sphere
    spans 5.4 degrees, so every ray within 3.5 degrees of the view axis
    meets the mirror, then the sphere."""
    parts = [part(quad(8.0), xform(R=rot_y(45), t=(0, 0, -10)), REFLECT, MIRROR_ALBEDO, fuzz=0.0),
             part(sphere_mesh, xform(t=SPHERE_CENTRE, s=1.5), DIFFUSE, diffuse_albedo)]
    return parts, camera(48, 48, (0, 0, 0), (0, 0, -10), 10)


def central_pixels(cam, max_deg):
    """Mask of pixels whose centre ray is within max_deg of the view axis."""
    w, h = cam["w"], cam["h"]
    half = math.tan(math.radians(cam["fov"]) / 2)
    ys, xs = np.mgrid[0:h, 0:w]
    sx = (2 * (xs + .5) / w - 1) * half * w / h
    sy = (2 * (ys + .5) / h - 1) * half
    return np.degrees(np.arctan(np.hypot(sx, sy))) < max_deg


# ---- closed forms (FP64) -------------------------------------------------------
def sky64(d, ambient=AMBIENT):
    d = np.asarray(d, np.float64)
    u = d / np.linalg.norm(d, axis=-1, keepdims=True)
    t = .5 * (u[..., 1:2] + 1.0)
    return (1.0 - t) + t * np.asarray(ambient, np.float64)


def reflect64(d, n):
    d = np.asarray(d, np.float64)
    u = d / np.linalg.norm(d, axis=-1, keepdims=True)
    n = np.asarray(n, np.float64)
    return u - 2.0 * (u @ n)[..., None] * n


def schlick64(cos_t, ratio):
    r0 = ((1 - ratio) / (1 + ratio)) ** 2
    return r0 + (1 - r0) * (1 - cos_t) ** 5


def lambert_sky_mean(albedo, ambient=AMBIENT):
    """Mean colour of a Lambertian bounce off an upward-facing surface whose
    bounce ray always reaches the sky: albedo * (1/6 + 5/6 ambient)."""
    return np.asarray(albedo, np.float64) * (1 / 6 + 5 / 6 * np.asarray(ambient, np.float64))


def mean_sigma(sumsq, raw, spp):
    """Standard error of the image mean per channel from per-pixel sums of
    squared sample colours."""
    var = np.maximum(np.asarray(sumsq, np.float64) / spp - np.asarray(raw, np.float64) ** 2, 0) / spp
    n = var.shape[0] * var.shape[1]
    return np.sqrt(var.reshape(-1, 3).sum(axis=0)) / n


# ---- building the scenes -------------------------------------------------------
def world_soup(orc, parts):
    """World-space triangles float32[N,9], the thing id of each and the
    oracle's thing records."""
    tris, inst, things = [], [], []
    for k, p in enumerate(parts):
        v, i = p["mesh"]
        wv = orc.transform_points(p["xf"], v)
        tris.append(wv[i].reshape(-1, 9))
        inst.append(np.full(len(i), k, np.uint32))
        xf = p["xf"]
        things.append(dict(type=p["type"], albedo=p["albedo"], fuzz=p["fuzz"], index=p["index"],
                           center=(xf[3], xf[7], xf[11]), radius=F(1), kind=1))
    return np.concatenate(tris), np.concatenate(inst), things


def oracle_scene(orc, parts, iterative=False, ambient=AMBIENT):
    tris, inst, things = world_soup(orc, parts)
    return orc.Scene(tris, inst, things, ambient=ambient, iterative=iterative)


def oracle_camera(orc, cam):
    return orc.camera(cam["eye"], cam["pat"], cam["vup"], cam["fov"], cam["w"] / cam["h"], cam["aperture"],
                      cam["fostance"])


def gpu_scene(rtx, parts, iterative=False, ambient=AMBIENT):
    s = rtx.Scene()
    s.set_miss_color(ambient)
    for p in parts:
        mid = s.add(p["mesh"])
        tid = s.add(rtx.Thing(p["type"], albedo=p["albedo"], fuzz=p["fuzz"], index=p["index"]), mid)
        s.set(tid, p["xf"])
    s.semantics(s.ITERATIVE if iterative else s.RECURSIVE)
    s.build()
    return s


def gpu_render(rtx, s, cam, spp, depth, rcam=None):
    """rawRGB, rpp, normals, albedos of one rtx_launch (rcam: an rtx Camera
    to use instead of setting one from cam)."""
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = cam["w"], cam["h"], spp, depth, s.handle
    if rcam is None:
        rcam = rtx.Camera().set(cam["eye"], cam["pat"], cam["vup"], cam["fov"], cam["w"] / cam["h"],
                                cam["aperture"], cam["fostance"])
    lp.camera = rcam.c
    L = rtx.Launcher(s, lp)
    L.ignite()
    out = dict(raw=L.raw_rgb().copy(), rpp=L.rays_per_pixel().copy(), normals=L.normals().copy(),
               albedos=L.albedos().copy())
    L.free()
    return out


# ---- rtow.cxx's own scene and camera on meshes ----------------------------------
def rtow_camera(orc, w, h):
    """The camera of rtow.cxx:85-94 (camera.h:10-23, set in FP64) as an
    oracle camera whose pixel mapping s' = 2(x+u)/w - 1 (optx/camera_r.cu:
    69-70) gives the rays of rtow's s = 2(x+u)/(w-1) - 1 (rtow.cxx:112-113):
    s = k s' + (k - 1) with k = w/(w-1), so wvec' = k wvec and dvec' =
    dvec - (k - 1) wvec, likewise for t and hvec."""
    eye, pat, vup = np.array([13., 2., 3.]), np.zeros(3), np.array([0., 1., 0.])
    cw = (eye - pat) / np.linalg.norm(eye - pat)
    cu = np.cross(vup, cw)
    cu /= np.linalg.norm(cu)
    cv = np.cross(cw, cu)
    hh = 2 * math.tan(.5 * 20 * math.pi / 180)
    ww = hh * 16 / 9
    hvec, wvec, dvec = 10 * hh / 2 * cv, 10 * ww / 2 * cu, 10 * cw
    kw, kh = w / (w - 1), h / (h - 1)
    c = orc.CameraC()
    c.eye[:], c.pat[:], c.vup[:] = eye, pat, vup
    c.fov, c.aspratio, c.aperture, c.fostance = 20., 16 / 9, .1, 10.
    c.u[:], c.v[:], c.w[:] = cu, cv, cw
    c.hvec[:], c.wvec[:] = kh * hvec, kw * wvec
    c.dvec[:] = dvec - (kw - 1) * wvec - (kh - 1) * hvec
    return c


def rtow_mesh_scene(orc, ground, small, big, iterative=False):
    """rtow.cxx's own scene (glibc rand() after srand(1), rtow.cxx:51-80)
    with the ground, the small and the big spheres tessellated at the given
    levels: an oracle scene and its triangle count."""
    things = orc.rtow_scene_glibc(1)
    lv = [(ground, small, big)[t.kind] for t in things]
    meshes = {k: orc.sphere_mesh(1.0, k) for k in set(lv)}
    tris, inst = orc.flatten(meshes, things, lv)
    return orc.Scene(tris, inst, things, iterative=iterative), len(tris)


def tile_means(a, t):
    h, w = a.shape[:2]
    return np.asarray(a, np.float64).reshape(h // t, t, w // t, t, -1).mean(axis=(1, 3))


def pixel_var(sumsq, mean, spp):
    """Variance of a pixel's mean colour from the sum of squared samples."""
    return np.maximum(np.asarray(sumsq, np.float64) / spp - np.asarray(mean, np.float64) ** 2, 0) / spp


def gpu_rtow_mesh_scene(rtx, orc, ground, small, big, iterative=False):
    """rtx_* counterpart of rtow_mesh_scene: the same things, meshes and
    transforms through the C-ABI."""
    from rtxplay_amd.scene import transform_for
    things = orc.rtow_scene_glibc(1)
    s = rtx.Scene()
    mids = {}
    for t in things:
        lv = (ground, small, big)[t.kind]
        if lv not in mids:
            mids[lv] = s.add(rtx.Sphere(1.0, lv).mesh())
        tid = s.add(rtx.Thing(t.type, albedo=tuple(t.albedo), fuzz=t.fuzz, index=t.index), mids[lv])
        s.set(tid, transform_for(dict(radius=F(t.radius), center=tuple(F(c) for c in t.center))))
    s.semantics(s.ITERATIVE if iterative else s.RECURSIVE)
    s.build()
    return s


def gpu_camera_from(rtx, ocam):
    """An rtx camera with the oracle camera's fields (bit for bit)."""
    c = rtx.Camera()
    for k in ("eye", "pat", "vup", "u", "v", "w", "hvec", "wvec", "dvec"):
        src = getattr(ocam, k)
        dst = getattr(c.c, k)
        dst.x, dst.y, dst.z = float(src[0]), float(src[1]), float(src[2])
    for k in ("fov", "aspratio", "aperture", "fostance"):
        setattr(c.c, k, getattr(ocam, k))
    return c
