"""The host recipe of RTWO::load (optx/rtwo.cxx:137-244) on the reference's
own stream: glibc rand() from srand(1) through util::rnd() (optx/util_cpu.h:
90-91), V::rnd() triples drawn z first (g++ evaluates make_float3's
arguments right to left).  The sphere counts are the ones the reference's
recipe gives with that stream (392 diffuse, 63 metal and 27 glass small
spheres, 486 things with the ground and the three big spheres)."""
import collections

import numpy as np

from rtxplay_amd.api import TYPE_DIFFUSE, TYPE_REFLECT, TYPE_REFRACT
from rtxplay_amd.scene import mixed_level, sphere_things, sphere_things_glibc

F = np.float32


def test_glibc_recipe_counts():
    s = sphere_things_glibc(11)
    assert len(s) == 486
    small = collections.Counter(t["type"] for t in s if t["kind"] == 1)
    assert (small[TYPE_DIFFUSE], small[TYPE_REFLECT], small[TYPE_REFRACT]) == (392, 63, 27)
    # the ground first and the three big spheres last, in the reference's order
    assert s[0]["radius"] == F(1000) and s[0]["center"] == (F(0), F(-1000), F(0))
    assert [t["type"] for t in s[-3:]] == [TYPE_REFRACT, TYPE_DIFFUSE, TYPE_REFLECT]
    assert [t["center"] for t in s[-3:]] == [(F(0), F(1), F(0)), (F(-4), F(1), F(0)), (F(4), F(1), F(0))]
    # the reference's tessellation levels (gas_9 ground, gas_6 / gas_3 small, gas_8 / 6 / 3 big)
    lv = [mixed_level(t, i, len(s)) for i, t in enumerate(s)]
    assert lv[0] == 9 and lv[-3:] == [8, 6, 3]
    assert all(lv[i] == (3 if t["type"] == TYPE_REFRACT else 6) for i, t in enumerate(s) if t["kind"] == 1)


def test_glibc_recipe_ranges_and_repeat():
    s = sphere_things_glibc(11)
    assert s == sphere_things_glibc(11)           # srand(1) again: the same list
    cells = set()
    for t in s[1:-3]:
        x, y, z = t["center"]
        a, b = int(np.floor(x)), int(np.floor(z))
        assert y == F(.2) and t["radius"] == F(.2) and (a, b) not in cells
        cells.add((a, b))
        assert 0 <= x - a <= .9 + 1e-5 and 0 <= z - b <= .9 + 1e-5
        alb = np.array(t["albedo"], np.float64)
        if t["type"] == TYPE_DIFFUSE:
            assert (alb >= 0).all() and (alb <= 1).all()
        elif t["type"] == TYPE_REFLECT:
            assert (alb >= .5).all() and (alb <= 1).all() and 0 <= t["fuzz"] <= .5
        else:
            assert t["index"] == F(1.5)


def test_glibc_recipe_differs_from_the_seeded_one():
    """The seeded recipe (sphere_things) is a different stream of the same
    recipe: same grid, different draws."""
    g, f = sphere_things_glibc(11), sphere_things(0, 11)
    assert [t["center"] for t in g[1:6]] != [t["center"] for t in f[1:6]]
