"""The RTOW final scene as RTWO::load builds it (optx/rtwo.cxx:137-244), drawn
from one seeded Frand48 stream instead of glibc rand().

Draw order (the reference leaves the albedo products' order to the compiler):
per grid cell select, center x, center z; then for a diffuse sphere two RGB
triples whose product is the albedo, for a metal sphere an RGB triple in
[.5,1) and the fuzz in [0,.5), for glass nothing.  All arithmetic is FP32
with one rounding per operation, as the C++ recipe does it.
"""
import numpy as np

from .api import Scene, Sphere, Thing, TYPE_DIFFUSE, TYPE_REFLECT, TYPE_REFRACT

F = np.float32
_M48 = (1 << 48) - 1
_TWO48 = F(2.0 ** 48)


class Frand48:
    """optx/frand48.h:17-35 on the host (float results bit-identical to the device)."""

    def __init__(self, seed=4711):
        self.state = int(seed) & 0xFFFFFFFFFFFFFFFF
        for _ in range(16):
            self()

    def __call__(self):
        self.state = (0x5DEECE66D * self.state + 0xB) & 0xFFFFFFFFFFFFFFFF
        return F(self.state & _M48) / _TWO48

    def range(self, lo, hi):
        """util::rnd(min,max), optx/util_gpu.h:28"""
        return F(lo) + self() * (F(hi) - F(lo))


def sphere_things(seed=0, half=11):
    """The sphere list: dicts with type, albedo, fuzz, index, center, radius, kind
    (kind 0 ground, 1 small, 2 big), in instance order."""
    rng = Frand48(seed)
    out = [dict(type=TYPE_DIFFUSE, albedo=(F(.5), F(.5), F(.5)), fuzz=F(0), index=F(0),
                center=(F(0), F(-1000), F(0)), radius=F(1000), kind=0)]
    for a in range(-half, half):
        for b in range(-half, half):
            select = rng()
            cx = F(a) + F(.9) * rng()
            cz = F(b) + F(.9) * rng()
            c = (cx, F(.2), cz)
            dx, dy, dz = cx - F(4), F(.2) - F(.2), cz - F(0)
            if np.sqrt(dx * dx + dy * dy + dz * dz) > F(.9):
                if select < F(.8):
                    r1, g1, b1 = rng(), rng(), rng()
                    r2, g2, b2 = rng(), rng(), rng()
                    out.append(dict(type=TYPE_DIFFUSE, albedo=(r1 * r2, g1 * g2, b1 * b2), fuzz=F(0), index=F(0),
                                    center=c, radius=F(.2), kind=1))
                elif select < F(.95):
                    alb = (rng.range(.5, 1.), rng.range(.5, 1.), rng.range(.5, 1.))
                    fz = rng.range(0., .5)
                    out.append(dict(type=TYPE_REFLECT, albedo=alb, fuzz=fz, index=F(0), center=c, radius=F(.2),
                                    kind=1))
                else:
                    out.append(dict(type=TYPE_REFRACT, albedo=(F(0), F(0), F(0)), fuzz=F(0), index=F(1.5),
                                    center=c, radius=F(.2), kind=1))
    out.append(dict(type=TYPE_REFRACT, albedo=(F(0),) * 3, fuzz=F(0), index=F(1.5), center=(F(0), F(1), F(0)),
                    radius=F(1), kind=2))
    out.append(dict(type=TYPE_DIFFUSE, albedo=(F(.4), F(.2), F(.1)), fuzz=F(0), index=F(0),
                    center=(F(-4), F(1), F(0)), radius=F(1), kind=2))
    out.append(dict(type=TYPE_REFLECT, albedo=(F(.7), F(.6), F(.5)), fuzz=F(0), index=F(0),
                    center=(F(4), F(1), F(0)), radius=F(1), kind=2))
    return out


class GlibcRand:
    """util::rnd() of the reference's host code (optx/util_cpu.h:90-91): the
    C library's rand() over (RAND_MAX + 1), in FP32, from srand(1) -- the
    stream RTWO::load draws from, since it never seeds.  (rand() keeps one
    state per process; the recipe below resets it.)"""

    def __init__(self):
        import ctypes
        self.libc = ctypes.CDLL("libc.so.6")
        self.libc.srand(1)
        self.scale = F(2147483647) + F(1)

    def __call__(self):
        return F(self.libc.rand()) / self.scale

    def range(self, lo, hi):
        return F(lo) + self() * (F(hi) - F(lo))

    def v3(self, lo=None, hi=None):
        """V::rnd() / V::rnd(min, max) (optx/v.h:80-81): make_float3 of three
        draws, whose arguments g++ evaluates right to left (z drawn first),
        as the byte-exact restatement of rtow.cxx confirms for the same
        pattern there"""
        d = (lambda: self()) if lo is None else (lambda: self.range(lo, hi))
        z = d()
        y = d()
        x = d()
        return x, y, z


def sphere_things_glibc(half=11):
    """RTWO::load (optx/rtwo.cxx:154-243) on the reference's own stream: the
    same sphere list the reference's batch binary renders (at half 11: 392
    diffuse, 63 metal and 27 glass small spheres plus the ground and three
    big ones).  The brace-initialised center draws x before z; a diffuse
    albedo is the component product of two V::rnd() triples."""
    rng = GlibcRand()
    out = [dict(type=TYPE_DIFFUSE, albedo=(F(.5), F(.5), F(.5)), fuzz=F(0), index=F(0),
                center=(F(0), F(-1000), F(0)), radius=F(1000), kind=0)]
    for a in range(-half, half):
        for b in range(-half, half):
            select = rng()
            cx = F(a) + F(.9) * rng()
            cz = F(b) + F(.9) * rng()
            c = (cx, F(.2), cz)
            dx, dy, dz = cx - F(4), F(.2) - F(.2), cz - F(0)
            if np.sqrt(dx * dx + dy * dy + dz * dz) > F(.9):
                if select < F(.8):
                    p = rng.v3()
                    q = rng.v3()
                    alb = (p[0] * q[0], p[1] * q[1], p[2] * q[2])
                    out.append(dict(type=TYPE_DIFFUSE, albedo=alb, fuzz=F(0), index=F(0), center=c, radius=F(.2),
                                    kind=1))
                elif select < F(.95):
                    alb = rng.v3(.5, 1.)
                    fz = rng.range(0., .5)
                    out.append(dict(type=TYPE_REFLECT, albedo=alb, fuzz=fz, index=F(0), center=c, radius=F(.2),
                                    kind=1))
                else:
                    out.append(dict(type=TYPE_REFRACT, albedo=(F(0), F(0), F(0)), fuzz=F(0), index=F(1.5),
                                    center=c, radius=F(.2), kind=1))
    out.extend(sphere_things(0, 0)[1:])     # the three big spheres
    return out


# the reference's own per-sphere tessellation levels (optx/rtwo.cxx:138-241)
def mixed_level(thing, index, count):
    if thing["kind"] == 0:
        return 9
    if thing["kind"] == 2:
        return {TYPE_REFRACT: 8, TYPE_DIFFUSE: 6, TYPE_REFLECT: 3}[thing["type"]]
    return 3 if thing["type"] == TYPE_REFRACT else 6


def transform_for(thing):
    r = thing["radius"]
    c = thing["center"]
    return np.array([r, 0, 0, c[0], 0, r, 0, c[1], 0, 0, r, c[2]], np.float32)


class RTOWScene(Scene):
    """RTWO::load (optx/rtwo.cxx:133-245) over unit-sphere meshes.

    level: an int (every sphere at that subdivision level) or "mixed" (the
    reference's per-sphere levels); seed: the Frand48 seed of the sphere
    list, or "glibc" for the reference's own stream (sphere_things_glibc:
    with level "mixed", the scene the reference's batch binary renders);
    spec: a sphere_things-style list to load instead; two_level: build the
    reference's layout (one structure per mesh under an instance structure)
    instead of the flat one."""

    def __init__(self, seed=0, half=11, level=4, device=0, spec=None, two_level=False):
        super().__init__(device)
        if two_level:
            self.build_mode(Scene.TWO_LEVEL)
        self.seed, self.half, self.level = seed, half, level
        self.things_spec = []
        self._spec = spec

    def load(self):
        if self._spec is not None:
            spec = self._spec
        elif self.seed == "glibc":
            spec = sphere_things_glibc(self.half)
        else:
            spec = sphere_things(self.seed, self.half)
        levels = [self.level if not isinstance(self.level, str) else mixed_level(t, i, len(spec))
                  for i, t in enumerate(spec)]
        meshes = {}
        for lv in sorted(set(levels)):
            meshes[lv] = self.add(Sphere(1.0, lv).mesh())
        for t, lv in zip(spec, levels):
            th = Thing(t["type"], albedo=t["albedo"], fuzz=t["fuzz"], index=t["index"])
            tid = self.add(th, meshes[lv])
            self.set(tid, transform_for(t))
        self.things_spec = spec
        self.levels = levels
        return len(spec)


# camera of optx/rtwo.cxx:100-108
RTOW_CAMERA = dict(eye=(13., 2., 3.), pat=(0., 0., 0.), vup=(0., 1., 0.), fov=20., aperture=.1, fostance=10.)


# ---------------------------------------------------------------------------
# the 64-frame batch workload: independent scenes and cameras of uneven size
# ---------------------------------------------------------------------------
BATCH_PARAM_SEED = 0x0BA7C400
_BATCH_DRAWS = 6


def batch_frame(i):
    """Parameters of frame i of the batch workload.

    The scene is sphere_things(seed=i, half) with every sphere at one
    subdivision level.  The other parameters come from one Frand48 stream
    (seed BATCH_PARAM_SEED) of which frame i takes draws 6i .. 6i+5 (one
    stream, because drand48 streams of consecutive seeds move in lock step),
    in this order: half-extent U{3..11}, level U{2..6}, resolution scale s
    log-uniform in [1, 6] (w x h = 2*round(160 s) x 2*round(90 s), i.e.
    320x180 .. 1920x1080), spp log-uniform in [4, 64], lookfrom angle
    U[0, 2 pi) on a radius-13 ring, lookfrom height U[1, 4].  Camera: lookat
    (0,0,0), vup y, vfov 20, aperture 0.1, focus distance 10; depth 16."""
    r = Frand48(BATCH_PARAM_SEED)
    for _ in range(_BATCH_DRAWS * int(i)):
        r()
    half = 3 + min(int(r() * 9), 8)
    level = 2 + min(int(r() * 5), 4)
    s = 6.0 ** float(r())
    w, h = min(2 * int(round(160 * s)), 1920), min(2 * int(round(90 * s)), 1080)
    spp = int(round(4.0 * 16.0 ** float(r())))
    ang = 2.0 * np.pi * float(r())
    height = 1.0 + 3.0 * float(r())
    eye = (13.0 * np.cos(ang), height, 13.0 * np.sin(ang))
    return dict(index=int(i), seed=int(i), half=half, level=level, w=w, h=h, spp=spp, depth=16,
                camera=dict(eye=eye, pat=(0., 0., 0.), vup=(0., 1., 0.), fov=20., aperture=.1, fostance=10.))


class BatchFrames:
    """Builds the scenes of a list of batch frames (one Scene, i.e. one
    context and acceleration structure, per frame), one Launcher per frame
    for its output buffers, and the LpGeneral records."""

    def __init__(self, frames, device=0, spp_scale=1.0, two_level=False):
        from .api import Camera, Launcher, LpGeneral
        self.frames = frames
        self.scenes, self.launchers, self.lps = [], [], []
        for f in frames:
            sc = RTOWScene(seed=f["seed"], half=f["half"], level=f["level"], device=device, two_level=two_level)
            sc.load()
            handle = sc.build()
            lp = LpGeneral()
            lp.image_w, lp.image_h, lp.depth, lp.is_handle = f["w"], f["h"], f["depth"], handle
            lp.spp = max(1, int(round(f["spp"] * spp_scale)))
            c = f["camera"]
            lp.camera = Camera().set(c["eye"], c["pat"], c["vup"], c["fov"], f["w"] / f["h"], c["aperture"],
                                     c["fostance"]).c
            self.scenes.append(sc)
            self.launchers.append(Launcher(sc, lp))
            self.lps.append(lp)

    def triangles(self):
        return sum(sc.stats()["num_prims"] for sc in self.scenes)

    def samples(self):
        return sum(lp.image_w * lp.image_h * lp.spp for lp in self.lps)

    def rays(self):
        return sum(int(l.rays_per_pixel().sum(dtype=np.uint64)) for l in self.launchers)

    def close(self):
        for l in self.launchers:
            l.free()
        for sc in self.scenes:
            sc.close()
