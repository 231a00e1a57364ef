"""orc_render -- the oracle's path loop with the GPU programs' semantics, which
the HIP frame equals bit for bit -- against the restatement of the
reference's own CPU tracer rtow.cxx, whose image is byte-exact with the
reference binary (test_oracle.py::test_rtow_restatement_bit_exact), on
rtow.cxx's own scene (glibc rand() after srand(1), rtow.cxx:51-80).

The two renderers share no code: rtow traces analytic spheres in FP64 with
the recursion of rtow.cxx:34-49 and the materials of optics.h; orc_render
traces the spheres tessellated (optx/sphere.cxx) in FP32 with the loop of
optx/camera_r.cu + optx/optics_r.cu.  Their images agree in expectation up
to three known differences, each accounted for:

  * depth: rtow's trace(ray, D) allows D scatters, i.e. D + 1 rays, which is
    lp.depth = D + 1 (rtow.cxx:34-43 against optx/optics_r.cu:31-42);
  * pixel mapping: rtow divides by w-1 and h-1 (rtow.cxx:112-113), the GPU by
    w and h (optx/camera_r.cu:69-70); the oracle's camera is rtow's camera
    with wvec, hvec and dvec rescaled so that both cast the same rays
    (cases.rtow_camera);
  * geometry: flat facets inside the analytic spheres.  The bias shrinks with
    the tessellation level; the tests measure it as the change from a coarser
    tessellation and allow it on top of the Monte Carlo error.

Checks at two depths (rtow depth 15 = lp.depth 16, the benchmark frames' depth; and
rtow depth 2 = lp.depth 3, where the depth limit decides most paths): the
per-channel image means agree within 3 sigma plus the tessellation term,
the 25 x 25-pixel tile means agree (chi^2 per tile near 1, no tile beyond
5 sigma), and the rays per sample agree."""
import numpy as np
import pytest

import cases as C

W, H, TILE = 400, 225, 25
# ground, small spheres, big spheres: the finer level set and a coarser one
FINE, COARSE = (9, 5, 8), (8, 4, 7)
RUNS = {"depth16": (10, 15), "depth3": (16, 2)}       # (spp, rtow depth)


@pytest.fixture(scope="module")
def renders(orc):
    cam = C.rtow_camera(orc, W, H)
    out = {}
    for key, (spp, depth) in RUNS.items():
        out[key] = {"spp": spp, "rtow": orc.rtow_render_ex(W, H, spp, depth)}
    for name, lv in (("fine", FINE), ("coarse", COARSE)):
        sc, n = C.rtow_mesh_scene(orc, *lv)
        print("tessellation", lv, n, "triangles")
        for key, (spp, depth) in RUNS.items():
            out[key][name] = sc.render(cam, W, H, spp, depth + 1, nthreads=8, want_sumsq=True)
        del sc
    return out


@pytest.mark.parametrize("key", sorted(RUNS))
def test_mesh_render_matches_rtow_in_the_mean(renders, key):
    r = renders[key]
    spp, R, G, Gc = r["spp"], r["rtow"], r["fine"], r["coarse"]
    vg, vr = C.pixel_var(G["sumsq"], G["raw"], spp), C.pixel_var(R["sumsq"], R["mean"], spp)
    n = W * H
    diff = G["raw"].reshape(-1, 3).astype(np.float64).mean(axis=0) - R["mean"].reshape(-1, 3).mean(axis=0)
    sigma = np.sqrt(vg.reshape(-1, 3).sum(axis=0) + vr.reshape(-1, 3).sum(axis=0)) / n
    tess = np.abs(G["raw"].reshape(-1, 3).astype(np.float64).mean(axis=0) -
                  Gc["raw"].reshape(-1, 3).astype(np.float64).mean(axis=0))
    print(key, "mean diff", diff, "3 sigma", 3 * sigma, "tessellation term", tess)
    assert (np.abs(diff) <= 3 * sigma + tess).all()
    # rays per sample: rtow counts every trace() call, the oracle every closest-hit query
    rays_g, rays_r = int(G["rpp"].sum(dtype=np.uint64)), int(R["rpp"].sum(dtype=np.uint64))
    sr = np.sqrt(np.var(G["rpp"].astype(np.float64)) + np.var(R["rpp"].astype(np.float64))) * np.sqrt(n)
    tr = abs(int(G["rpp"].sum(dtype=np.uint64)) - int(Gc["rpp"].sum(dtype=np.uint64)))
    print(key, "rays", rays_g, rays_r, "3 sigma", 3 * sr, "tessellation term", tr)
    assert rays_r == R["rays"]
    assert abs(rays_g - rays_r) <= 3 * sr + tr


@pytest.mark.parametrize("key", sorted(RUNS))
def test_mesh_render_matches_rtow_tile_by_tile(renders, key):
    r = renders[key]
    spp, R, G = r["spp"], r["rtow"], r["fine"]
    vg, vr = C.pixel_var(G["sumsq"], G["raw"], spp), C.pixel_var(R["sumsq"], R["mean"], spp)
    d = C.tile_means(G["raw"], TILE) - C.tile_means(R["mean"], TILE)
    sd = np.sqrt((C.tile_means(vg, TILE) + C.tile_means(vr, TILE)) / (TILE * TILE))
    live = sd > 0                           # sky tiles of the blue channel are exact on both sides
    z = d[live] / sd[live]
    chi2 = float((z ** 2).mean())
    print(key, "tiles", z.size, "chi2/dof %.2f" % chi2, "max |z| %.2f" % np.abs(z).max())
    assert chi2 < 1.6 and np.abs(z).max() < 5.0
    assert (np.abs(d[~live]) < 1e-6).all()
