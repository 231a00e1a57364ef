"""Batched launches (many scenes and cameras in one persistent launch,
hipGraph-captured): every frame of a batch must be bit-identical to a plain
launch of the same launch parameters, under any launch tuning, on graph
replays, and for the 64-frame workload."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-batch]", *a, flush=True)


# small uneven frames: (seed, half, level, w, h, spp, camera angle, height)
SMALL = [(0, 3, 2, 96, 54, 6, 0.3, 2.0), (1, 4, 3, 40, 23, 3, 1.9, 1.2), (2, 5, 2, 130, 70, 2, 3.0, 3.5),
         (3, 3, 3, 17, 9, 9, 4.4, 1.0), (4, 6, 2, 64, 64, 4, 5.5, 2.7), (5, 3, 2, 1, 1, 5, 0.0, 4.0)]


def _frames():
    out = []
    for seed, half, level, w, h, spp, ang, hgt in SMALL:
        out.append(dict(index=seed, seed=seed, half=half, level=level, w=w, h=h, spp=spp, depth=16,
                        camera=dict(eye=(13 * np.cos(ang), hgt, 13 * np.sin(ang)), pat=(0., 0., 0.),
                                    vup=(0., 1., 0.), fov=20., aperture=.1, fostance=10.)))
    return out


def _outputs(bf):
    return [(l.raw_rgb(), l.rays_per_pixel(), l.normals(), l.albedos()) for l in bf.launchers]


def _clear(bf):
    for l in bf.launchers:
        for k in ("rawRGB", "rpp", "normals", "albedos"):
            l._bufs[k].zero()


def _same(a, b):
    return all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))


@pytest.fixture(scope="module")
def small_batch(rtx):
    from rtxplay_amd.scene import BatchFrames
    say("building", len(SMALL), "scenes")
    bf = BatchFrames(_frames())
    for l in bf.launchers:
        l.ignite()
    ref = _outputs(bf)
    yield bf, ref
    bf.close()


def test_batch_equals_plain_launches(rtx, small_batch):
    bf, ref = small_batch
    for use_graph in (False, True):
        _clear(bf)
        b = rtx.Batch(bf.scenes[0], bf.lps, use_graph=use_graph)
        from rtxplay_amd.hip import Stream
        st = Stream()
        say("batch launch, graph", use_graph)
        b.launch(st)
        st.synchronize()
        got = _outputs(bf)
        for i, (g, r) in enumerate(zip(got, ref)):
            assert _same(g, r), "frame %d differs (graph=%s)" % (i, use_graph)
        b.free()


def test_batch_graph_replays(rtx, small_batch):
    """The captured graph is replayed several times with eager work between
    the replays; each replay reproduces the plain launches."""
    from rtxplay_amd.hip import Stream
    bf, ref = small_batch
    st = Stream()
    b = rtx.Batch(bf.scenes[0], bf.lps, use_graph=True)
    for rep in range(3):
        _clear(bf)
        bf.launchers[1].ignite(stream=st)       # eager launch on the same stream between replays
        say("replay", rep)
        b.launch(st)
        st.synchronize()
        assert all(_same(g, r) for g, r in zip(_outputs(bf), ref)), "replay %d" % rep
    b.free()


def test_batch_iterative_semantics_equals_plain_launches(rtx):
    """Depth 3 under the iterative programs' semantics (set on every frame's
    context; the batch takes its first context's): each frame of the graph
    launch equals its plain launch bit for bit."""
    from rtxplay_amd.hip import Stream
    from rtxplay_amd.scene import BatchFrames
    bf = BatchFrames(_frames())
    try:
        for sc, lp in zip(bf.scenes, bf.lps):
            sc.semantics(sc.ITERATIVE)
            lp.depth = 3
        for l in bf.launchers:
            l.ignite()
        ref = _outputs(bf)
        _clear(bf)
        b = rtx.Batch(bf.scenes[0], bf.lps, use_graph=True)
        st = Stream()
        say("iterative batch, depth 3")
        b.launch(st)
        st.synchronize()
        for i, (g, r) in enumerate(zip(_outputs(bf), ref)):
            assert _same(g, r), "frame %d differs" % i
        b.free()
    finally:
        bf.close()


@pytest.mark.parametrize("tuning", [dict(tile_w=16), dict(tile_w=2, stack_lds=16), dict(shade_batch=8, inner_frac=48),
                                    dict(blocks_per_cu=1, tile_w=64), dict(shade_batch=64, inner_frac=1)])
def test_tuning_does_not_change_pixels(rtx, small_batch, tuning):
    bf, ref = small_batch
    sc = bf.scenes[0]
    old = sc.tuning()
    try:
        sc.tuning(**tuning)
        say("tuning", sc.tuning())
        _clear(bf)
        b = rtx.Batch(sc, bf.lps, use_graph=False)
        b.launch()
        b.free()
        assert all(_same(g, r) for g, r in zip(_outputs(bf), ref)), "batch under %s" % tuning
        # the plain launch under the same tuning
        _clear(bf)
        bf.launchers[0].ignite()
        assert _same(_outputs(bf)[0], ref[0]), "plain launch under %s" % tuning
    finally:
        sc.tuning(**{k: old[k] for k in ("tile_w", "stack_lds", "blocks_per_cu", "shade_batch", "inner_frac")})


def test_tuning_rejects_bad_values(rtx, small_batch):
    bf, _ = small_batch
    sc = bf.scenes[0]
    mx = sc.tuning()["max_blocks_per_cu"]
    for bad in (dict(tile_w=3), dict(tile_w=128), dict(stack_lds=24), dict(inner_frac=65),
                dict(blocks_per_cu=mx + 1)):
        with pytest.raises(rtx.RtxError):
            sc.tuning(**bad)


def test_batch_rejects_bad_frames(rtx, small_batch):
    import ctypes
    bf, _ = small_batch
    bad = type(bf.lps[0])()
    ctypes.memmove(ctypes.byref(bad), ctypes.byref(bf.lps[0]), ctypes.sizeof(bad))
    bad.is_handle = 12345
    with pytest.raises(rtx.RtxError):
        rtx.Batch(bf.scenes[0], [bf.lps[1], bad])
    bad.is_handle = bf.lps[0].is_handle
    bad.spp = 0
    with pytest.raises(rtx.RtxError):
        rtx.Batch(bf.scenes[0], [bad])
    with pytest.raises(rtx.RtxError):
        rtx.Batch(bf.scenes[0], [])


def test_batch_64_frames_equal_plain(rtx):
    """The 64-frame workload (spp cut to 1/8, at least 1): batch == 64 plain launches."""
    from rtxplay_amd.scene import BatchFrames, batch_frame
    say("building the 64 batch scenes")
    bf = BatchFrames([batch_frame(i) for i in range(64)], spp_scale=0.125)
    try:
        say("triangles", bf.triangles(), "samples", bf.samples())
        for l in bf.launchers:
            l.ignite(sync=False)
        rtx.synchronize()
        ref = [(l.raw_rgb(), l.rays_per_pixel()) for l in bf.launchers]
        _clear(bf)
        b = rtx.Batch(bf.scenes[0], bf.lps, use_graph=True)
        from rtxplay_amd.hip import Stream
        st = Stream()
        b.launch(st)
        st.synchronize()
        say("batch done")
        for i, l in enumerate(bf.launchers):
            assert np.array_equal(l.raw_rgb().view(np.uint32), ref[i][0].view(np.uint32)), "frame %d" % i
            assert np.array_equal(l.rays_per_pixel(), ref[i][1]), "frame %d rpp" % i
        b.free()
    finally:
        bf.close()
