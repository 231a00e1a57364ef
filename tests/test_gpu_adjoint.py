"""Differentiable render: material parameters on the device, the adjoint
path-replay kernel against central finite differences (h = 1e-3, common
random numbers), and the torch.autograd harness.

Albedo never changes a path, so its gradient is checked against finite
differences of the real GPU render.  Fuzz moves directions, and a path whose
direction crosses a triangle edge or a silhouette jumps; the adjoint gives
the derivative with the discrete path structure fixed, so its check is the
oracle's frozen-structure render (paths traced once, radiance re-evaluated
in FP64 along the same vertices), whose finite differences are that
derivative."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H_FD = 1e-3
REL = 1e-3


def say(*a):
    print("[gpu-adj]", *a, flush=True)


def _lp(rtx, handle, w, h, spp, depth):
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, spp, depth, handle
    lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
    return lp


def _target_params(m0, types, seed=7):
    """Seeded 'ground truth' materials: albedo scaled by U(.6, 1.3), clipped
    to [0.02, 1]; metal fuzz U(0, .5)."""
    from rtxplay_amd import TYPE_DIFFUSE, TYPE_REFLECT
    g = np.random.default_rng(seed)
    mt = m0.copy()
    for k, t in enumerate(types):
        if t in (TYPE_DIFFUSE, TYPE_REFLECT):
            mt[k, :3] = np.clip(m0[k, :3] * g.uniform(.6, 1.3, 3), .02, 1.).astype(np.float32)
        if t == TYPE_REFLECT:
            mt[k, 3] = np.float32(g.uniform(0, .5))
    return mt


class Frame:
    def __init__(self, rtx, level, w, h, spp, depth):
        from rtxplay_amd.scene import RTOWScene
        self.sc = RTOWScene(seed=0, half=11, level=level)
        self.sc.load()
        self.handle = self.sc.build()
        self.lp = _lp(rtx, self.handle, w, h, spp, depth)
        self.la = rtx.Launcher(self.sc, self.lp)
        self.types = [t["type"] for t in self.sc.things_spec]
        self.m0 = self.sc.get_materials()
        self.w, self.h, self.spp, self.depth = w, h, spp, depth

    def render(self, m):
        self.sc.set_materials(m)
        self.la.ignite()
        return self.la.raw_rgb()

    def close(self):
        self.sc.set_materials(self.m0)
        self.la.free()
        self.sc.close()


@pytest.fixture(scope="module")
def dframe(rtx):
    say("level-4 scene, 640x360, 64 spp, depth 8")
    f = Frame(rtx, 4, 640, 360, 64, 8)
    f.mt = _target_params(f.m0, f.types)
    f.T = f.render(f.mt).astype(np.float64)
    f.I = f.render(f.m0)
    yield f
    f.close()


def _loss(img, T):
    return float(((img.astype(np.float64) - T) ** 2).sum())


def _adjoint(f, I, T):
    f.sc.set_materials(f.m0)
    f.la.ignite()
    assert np.array_equal(f.la.raw_rgb().view(np.uint32), I.view(np.uint32))
    return f.la.adjoint((2.0 * (I.astype(np.float64) - T)).astype(np.float32)).astype(np.float64)


def test_materials_roundtrip(dframe, rtx):
    from rtxplay_amd import TYPE_DIFFUSE, TYPE_REFLECT
    spec = dframe.sc.things_spec
    for k, t in enumerate(spec):
        if t["type"] in (TYPE_DIFFUSE, TYPE_REFLECT):
            assert np.array_equal(dframe.m0[k, :3], np.array(t["albedo"], np.float32)), k
        if t["type"] == TYPE_REFLECT:
            assert dframe.m0[k, 3] == np.float32(t["fuzz"])
    dframe.sc.set_materials(dframe.mt)
    got = dframe.sc.get_materials()
    dframe.sc.set_materials(dframe.m0)
    for k, t in enumerate(dframe.types):
        if t == TYPE_DIFFUSE:
            assert np.array_equal(got[k, :3], dframe.mt[k, :3])
        elif t == TYPE_REFLECT:
            assert np.array_equal(got[k], dframe.mt[k])
        else:
            assert np.array_equal(got[k], dframe.m0[k])          # dielectric: untouched


def test_set_materials_equals_a_scene_built_with_them(dframe, rtx):
    from rtxplay_amd.scene import RTOWScene
    from rtxplay_amd import TYPE_DIFFUSE, TYPE_REFLECT
    spec = [dict(t) for t in dframe.sc.things_spec]
    for k, t in enumerate(spec):
        if t["type"] in (TYPE_DIFFUSE, TYPE_REFLECT):
            t["albedo"] = tuple(np.float32(x) for x in dframe.mt[k, :3])
        if t["type"] == TYPE_REFLECT:
            t["fuzz"] = np.float32(dframe.mt[k, 3])
    s2 = RTOWScene(level=4, spec=spec)
    s2.load()
    lp2 = _lp(rtx, s2.build(), 160, 90, 8, 8)
    l2 = rtx.Launcher(s2, lp2)
    l2.ignite()
    fresh = l2.raw_rgb()
    l2.free()
    s2.close()
    f = Frame(rtx, 4, 160, 90, 8, 8)
    try:
        got = f.render(dframe.mt)
    finally:
        f.close()
    assert np.array_equal(got.view(np.uint32), fresh.view(np.uint32))


def test_albedo_gradient_matches_finite_differences(dframe, rtx):
    """The full 640x360 level-4 frame: the 16 albedo channels with the largest gradients, the
    ground among them, against central differences of the GPU render."""
    g = _adjoint(dframe, dframe.I, dframe.T)
    from rtxplay_amd import TYPE_REFRACT
    assert np.all(g[[k for k, t in enumerate(dframe.types) if t == TYPE_REFRACT]] == 0)
    cand = [(abs(g[k, c]), k, c) for k in range(len(g)) for c in range(3)]
    cand.sort(reverse=True)
    picks = [(k, c) for _, k, c in cand[:16]]
    assert any(k == 0 for k, _ in picks)
    worst = 0.0
    for k, c in picks:
        mp, mm = dframe.m0.copy(), dframe.m0.copy()
        mp[k, c] = np.float32(dframe.m0[k, c] + H_FD)
        mm[k, c] = np.float32(dframe.m0[k, c] - H_FD)
        fd = (_loss(dframe.render(mp), dframe.T) - _loss(dframe.render(mm), dframe.T)) / (float(mp[k, c]) - float(mm[k, c]))
        rel = abs(g[k, c] - fd) / abs(fd)
        worst = max(worst, rel)
        say("albedo", k, c, "adjoint %.6e fd %.6e rel %.2e" % (g[k, c], fd, rel))
        assert rel <= REL, (k, c, g[k, c], fd)
    dframe.sc.set_materials(dframe.m0)
    say("worst albedo rel err %.2e" % worst)


def test_fuzz_gradient_matches_frozen_differences_on_the_full_frame(dframe, rtx, orc):
    """The full 640x360 level-4 frame: the 3 metals with the largest fuzz
    gradients against central differences (h = 1e-3) of the oracle's
    frozen-structure render, the interior derivative the adjoint computes
    (the jumps of paths that cross a facet edge, a silhouette or the
    absorption boundary are not in it; the benchmark reports their size
    against differences of the real render).  The loss is taken over the
    pixels where the GPU image and the oracle's agree."""
    from rtxplay_amd import TYPE_REFLECT
    meshes = {4: orc.sphere_mesh(1.0, 4)}
    tris, inst = orc.flatten(meshes, dframe.sc.things_spec, [4] * len(dframe.types))
    sc = orc.Scene(tris, inst, dframe.sc.things_spec)
    cam = orc.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, dframe.w / dframe.h, 0.1, 10.0)
    say("oracle frozen render of the 640x360 frame")
    agree = np.all(np.abs(sc.render_frozen(cam, dframe.w, dframe.h, dframe.spp, dframe.depth, dframe.m0) - dframe.I) <= 1e-5, axis=2)
    say("pixels where GPU and oracle agree: %.4f" % agree.mean())
    assert agree.mean() > 0.99
    M = agree[..., None].astype(np.float64)
    g = _adjoint(dframe, dframe.I, dframe.T + (1.0 - M) * (dframe.I - dframe.T))
    # central differences need fuzz >= h (the big metal sphere's fuzz is 0)
    metals = sorted([k for k, t in enumerate(dframe.types) if t == TYPE_REFLECT and dframe.m0[k, 3] >= 2 * H_FD],
                    key=lambda k: -abs(g[k, 3]))

    def loss(m):
        return float((M * (sc.render_frozen(cam, dframe.w, dframe.h, dframe.spp, dframe.depth, m) - dframe.T) ** 2).sum())

    for k in metals[:3]:
        mp, mm = dframe.m0.copy(), dframe.m0.copy()
        mp[k, 3] = np.float32(dframe.m0[k, 3] + H_FD)
        mm[k, 3] = np.float32(dframe.m0[k, 3] - H_FD)
        fd = (loss(mp) - loss(mm)) / (float(mp[k, 3]) - float(mm[k, 3]))
        rel = abs(g[k, 3] - fd) / max(abs(fd), 1e-12)
        say("fuzz", k, "adjoint %.6e frozen fd %.6e rel %.2e" % (g[k, 3], fd, rel))
        assert rel <= REL, (k, g[k, 3], fd)


def test_albedo_gradient_iterative_semantics(rtx):
    """Under the iterative programs' semantics a sample that reaches the
    depth limit keeps its throughput (a product of albedos): at depth 3 most
    samples do, and the adjoint's albedo gradients match central differences
    of the GPU render."""
    f = Frame(rtx, 4, 160, 90, 16, 3)
    f.sc.semantics(f.sc.ITERATIVE)
    try:
        mt = _target_params(f.m0, f.types)
        T = f.render(mt).astype(np.float64)
        I = f.render(f.m0)
        g = _adjoint(f, I, T)
        cand = sorted(((abs(g[k, c]), k, c) for k in range(len(g)) for c in range(3)), reverse=True)[:8]
        for _, k, c in cand:
            mp, mm = f.m0.copy(), f.m0.copy()
            mp[k, c] = np.float32(f.m0[k, c] + H_FD)
            mm[k, c] = np.float32(f.m0[k, c] - H_FD)
            fd = (_loss(f.render(mp), T) - _loss(f.render(mm), T)) / (float(mp[k, c]) - float(mm[k, c]))
            rel = abs(g[k, c] - fd) / abs(fd)
            say("iterative albedo", k, c, "adjoint %.6e fd %.6e rel %.2e" % (g[k, c], fd, rel))
            assert rel <= REL, (k, c, g[k, c], fd)
    finally:
        f.sc.semantics(f.sc.RECURSIVE)
        f.close()


@pytest.fixture(scope="module")
def small(rtx, orc):
    f = Frame(rtx, 4, 160, 90, 16, 8)
    f.mt = _target_params(f.m0, f.types)
    f.T = f.render(f.mt).astype(np.float64)
    f.I = f.render(f.m0)
    say("oracle scene for the frozen-structure checker")
    meshes = {4: orc.sphere_mesh(1.0, 4)}
    tris, inst = orc.flatten(meshes, f.sc.things_spec, [4] * len(f.types))
    f.orc = orc.Scene(tris, inst, f.sc.things_spec)
    f.cam = orc.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, 160 / 90, 0.1, 10.0)
    yield f
    f.close()


def test_oracle_frozen_matches_gpu_image(small):
    img = small.orc.render_frozen(small.cam, 160, 90, 16, 8, small.m0)
    d = np.abs(img - small.I)
    say("frozen vs gpu: max %.2e, frac > 1e-4 %.4f" % (d.max(), (d > 1e-4).mean()))
    assert (d > 1e-4).mean() < 2e-3


def test_fuzz_gradient_matches_frozen_finite_differences(small, rtx):
    """The loss is taken over the pixels where the GPU image and the oracle's
    agree (FP32 traversal near triangle edges sends a few paths elsewhere in
    one of them); both the adjoint and the differences use that loss."""
    from rtxplay_amd import TYPE_REFLECT
    agree = np.all(np.abs(small.orc.render_frozen(small.cam, 160, 90, 16, 8, small.m0) - small.I) <= 1e-5, axis=2)
    say("pixels where GPU and oracle agree: %.4f" % agree.mean())
    assert agree.mean() > 0.99
    M = agree[..., None].astype(np.float64)
    g = _adjoint(small, small.I, small.T + (1.0 - M) * (small.I - small.T))
    for k, t in enumerate(small.types):
        if t != TYPE_REFLECT:
            assert g[k, 3] == 0.0
    metals = sorted([k for k, t in enumerate(small.types) if t == TYPE_REFLECT], key=lambda k: -abs(g[k, 3]))
    picks = metals[:8]

    def loss(m):
        return float((M * (small.orc.render_frozen(small.cam, 160, 90, 16, 8, m) - small.T) ** 2).sum())

    for k in picks:
        mp, mm = small.m0.copy(), small.m0.copy()
        mp[k, 3] = np.float32(small.m0[k, 3] + H_FD)
        mm[k, 3] = np.float32(small.m0[k, 3] - H_FD)
        fd = (loss(mp) - loss(mm)) / (float(mp[k, 3]) - float(mm[k, 3]))
        rel = abs(g[k, 3] - fd) / max(abs(fd), 1e-12)
        say("fuzz", k, "adjoint %.6e fd %.6e rel %.2e" % (g[k, 3], fd, rel))
        assert rel <= REL, (k, g[k, 3], fd)
    # albedo through the same checker
    cand = sorted(((abs(g[k, c]), k, c) for k in range(len(g)) for c in range(3)), reverse=True)[:6]
    for _, k, c in cand:
        mp, mm = small.m0.copy(), small.m0.copy()
        mp[k, c] = np.float32(small.m0[k, c] + H_FD)
        mm[k, c] = np.float32(small.m0[k, c] - H_FD)
        fd = (loss(mp) - loss(mm)) / (float(mp[k, c]) - float(mm[k, c]))
        rel = abs(g[k, c] - fd) / abs(fd)
        say("albedo (frozen)", k, c, "adjoint %.6e fd %.6e rel %.2e" % (g[k, c], fd, rel))
        assert rel <= REL


def test_adjoint_rejects_bad_arguments(small, rtx):
    lp = rtx.LpGeneral()
    ctypes.memmove(ctypes.byref(lp), ctypes.byref(small.lp), ctypes.sizeof(lp))
    lp.is_handle = 999
    d = rtx.DeviceBuffer(16 * len(small.types))
    gi = rtx.DeviceBuffer(12 * small.w * small.h)
    with pytest.raises(rtx.RtxError):
        rtx.api.call("rtx_launch_adjoint", small.sc.ctx, ctypes.byref(lp), gi.ptr, d.ptr, None)
    lp.is_handle = small.lp.is_handle
    lp.rawRGB = None
    with pytest.raises(rtx.RtxError):
        rtx.api.call("rtx_launch_adjoint", small.sc.ctx, ctypes.byref(lp), gi.ptr, d.ptr, None)
    d.free()
    gi.free()


def test_torch_harness(small, rtx):
    import torch
    from rtxplay_amd.diff import DiffRender
    r = DiffRender(small.sc, small.lp)
    p = torch.tensor(small.m0, device="cuda", requires_grad=True)
    T = torch.tensor(small.T, dtype=torch.float32, device="cuda")
    img = r(p)
    torch.cuda.synchronize()
    assert np.array_equal(img.detach().cpu().numpy().view(np.uint32), small.I.view(np.uint32))
    loss = ((img - T) ** 2).sum()
    loss.backward()
    g_t = p.grad.detach().cpu().numpy().astype(np.float64)
    # dL/dI in float32 from torch vs the float64-built one: compare loosely
    g = _adjoint(small, small.I, small.T)
    scale = np.abs(g).max()
    assert np.abs(g_t - g).max() <= 1e-4 * scale
    # a few Adam steps from all-0.5 albedo / 0.25 fuzz lower the loss
    from rtxplay_amd import TYPE_DIFFUSE, TYPE_REFLECT
    init = small.m0.copy()
    for k, t in enumerate(small.types):
        if t in (TYPE_DIFFUSE, TYPE_REFLECT):
            init[k, :3] = .5
        if t == TYPE_REFLECT:
            init[k, 3] = .25
    mask = torch.tensor([[1, 1, 1, 1 if t == TYPE_REFLECT else 0] if t in (TYPE_DIFFUSE, TYPE_REFLECT)
                         else [0, 0, 0, 0] for t in small.types], dtype=torch.float32, device="cuda")
    q = torch.tensor(init, device="cuda", requires_grad=True)
    opt = torch.optim.Adam([q], lr=0.02)
    losses = []
    for it in range(25):
        opt.zero_grad()
        loss = ((r(q) - T) ** 2).sum()
        loss.backward()
        q.grad.mul_(mask)
        opt.step()
        with torch.no_grad():
            q.clamp_(0.0, 1.0)
        losses.append(float(loss.detach()))
    say("adam losses", ["%.2f" % x for x in losses[::6]], "->", "%.2f" % losses[-1])
    assert losses[-1] < 0.5 * losses[0]
    small.sc.set_materials(small.m0)


def test_torch_harness_two_forwards_before_backward(small, rtx):
    """loss = f(p1) + f(p2): the second forward leaves p2 in the scene, and
    the first forward's backward must still replay its paths with p1; the
    gradients equal those of the two losses taken one at a time."""
    import torch
    from rtxplay_amd.diff import DiffRender
    r = DiffRender(small.sc, small.lp)
    T = torch.tensor(small.T, dtype=torch.float32, device="cuda")
    p1 = torch.tensor(small.m0, device="cuda")
    p2 = torch.tensor(small.mt, device="cuda")
    grads = []
    for ps in ((p1,), (p2,)):
        q = ps[0].clone().requires_grad_(True)
        ((r(q) - T) ** 2).sum().backward()
        torch.cuda.synchronize()
        grads.append(q.grad.detach().cpu().numpy().astype(np.float64))
    a = p1.clone().requires_grad_(True)
    b = p2.clone().requires_grad_(True)
    (((r(a) - T) ** 2).sum() + ((r(b) - T) ** 2).sum()).backward()
    torch.cuda.synchronize()
    ga, gb = a.grad.detach().cpu().numpy(), b.grad.detach().cpu().numpy()
    scale = max(np.abs(grads[0]).max(), np.abs(grads[1]).max())
    say("two forwards: max |dg| %.2e of %.2e" % (max(np.abs(ga - grads[0]).max(), np.abs(gb - grads[1]).max()), scale))
    assert np.abs(ga - grads[0]).max() <= 1e-4 * scale
    assert np.abs(gb - grads[1]).max() <= 1e-4 * scale
    assert np.abs(grads[0] - grads[1]).max() > 1e-2 * scale      # the two parameter sets differ
    small.sc.set_materials(small.m0)
