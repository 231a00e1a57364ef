"""Inverse rendering with the adjoint kernel: recover per-sphere Lambertian
albedo and metal fuzz of the RTOW final scene (level-4 spheres) from one
target image.

The target is rendered with seeded ground-truth materials (albedo of every
diffuse and metal sphere scaled by U(.6, 1.3) and clipped to [.02, 1], metal
fuzz U(0, .5)); the optimisation starts from albedo 0.5 and fuzz 0.25 and
minimises the L2 image loss with Adam, gradients from the adjoint
path-replay kernel through the torch.autograd harness.  Both renders use the
same per-pixel random streams.

    python -m rtxplay_amd.apps.recover_materials [--iters 200] [--w 640 --h 360 --spp 64 --depth 8]
                                                 [--fuzz-spsa 0.02]

Prints one line per report interval and a JSON summary: loss, and the mean
absolute parameter error over the parameters the image constrains (those
with a nonzero gradient at the start) before and after.
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from rtxplay_amd.diff import DiffRender  # noqa: E402  (torch first, then librtx)
import rtxplay_amd as rtx
from rtxplay_amd import TYPE_DIFFUSE, TYPE_REFLECT
from rtxplay_amd.scene import RTOWScene


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--w", type=int, default=640)
    ap.add_argument("--h", type=int, default=360)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--level", type=int, default=4)
    ap.add_argument("--lr", type=float, default=0.02)
    ap.add_argument("--lr-end", type=float, default=0.001, help="learning rate at the last iteration (geometric decay)")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--report", type=int, default=20)
    ap.add_argument("--fuzz-spsa", type=float, default=0.0,
                    help="c > 0: fuzz gradients from simultaneous perturbation of all fuzz values by +-c (two "
                         "more renders per step, differences of the real render, so with the boundary terms "
                         "the adjoint's interior derivative leaves out) instead of the adjoint")
    a = ap.parse_args(argv)

    sc = RTOWScene(seed=0, half=11, level=a.level)
    sc.load()
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = a.w, a.h, a.spp, a.depth, sc.build()
    lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, a.w / a.h, 0.1, 10.0).c
    r = DiffRender(sc, lp)
    types = [t["type"] for t in sc.things_spec]
    m0 = sc.get_materials()
    g = np.random.default_rng(a.seed)
    truth = m0.copy()
    for k, t in enumerate(types):
        if t in (TYPE_DIFFUSE, TYPE_REFLECT):
            truth[k, :3] = np.clip(m0[k, :3] * g.uniform(.6, 1.3, 3), .02, 1.)
        if t == TYPE_REFLECT:
            truth[k, 3] = g.uniform(0, .5)
    mask = np.zeros_like(m0)
    for k, t in enumerate(types):
        if t in (TYPE_DIFFUSE, TYPE_REFLECT):
            mask[k, :3] = 1
        if t == TYPE_REFLECT:
            mask[k, 3] = 1
    init = np.where(mask > 0, np.where(np.arange(4)[None, :] < 3, .5, .25), m0).astype(np.float32)

    dev = torch.device("cuda")
    with torch.no_grad():
        target = r(torch.tensor(truth, device=dev))
    mask_t = torch.tensor(mask, device=dev)
    q = torch.tensor(init, device=dev, requires_grad=True)
    opt = torch.optim.Adam([q], lr=a.lr)
    sched = torch.optim.lr_scheduler.LambdaLR(opt, lambda it: (a.lr_end / a.lr) ** (it / max(a.iters - 1, 1)))

    # the parameters the image constrains: nonzero gradient at the start
    loss0 = ((r(q) - target) ** 2).sum()
    loss0.backward()
    g0 = q.grad.abs().cpu().numpy()
    seen = (g0 > 0) & (mask > 0)
    strong = seen & (g0 >= np.median(g0[seen]))          # the better-constrained half
    q.grad = None
    err0 = float(np.abs(init - truth)[seen].mean())
    err0s = float(np.abs(init - truth)[strong].mean())
    print("%d material parameters, %d constrained by the image; start loss %.4f, mean |error| %.4f "
          "(better-constrained half %.4f)" % (int(mask.sum()), int(seen.sum()), float(loss0.detach()), err0, err0s),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    loss = loss0
    for it in range(1, a.iters + 1):
        opt.zero_grad()
        loss = ((r(q) - target) ** 2).sum()
        loss.backward()
        if a.fuzz_spsa > 0:
            with torch.no_grad():
                fz = mask_t[:, 3] > 0
                delta = torch.where(torch.rand(q.shape[0], device=dev) < .5, -1.0, 1.0) * fz
                hi, lo = q.detach().clone(), q.detach().clone()
                hi[:, 3] += a.fuzz_spsa * delta
                lo[:, 3] -= a.fuzz_spsa * delta
                d = float(((r.forward(hi) - target) ** 2).sum() - ((r.forward(lo) - target) ** 2).sum())
                q.grad[:, 3] = torch.where(fz, d / (2 * a.fuzz_spsa * torch.where(fz, delta, 1.0)), 0.0)
        q.grad.mul_(mask_t)
        opt.step()
        sched.step()
        with torch.no_grad():
            q.clamp_(0.0, 1.0)
        if it % a.report == 0 or it == a.iters:
            e = np.abs(q.detach().cpu().numpy() - truth)
            print("iter %4d  loss %.5f  mean |error| %.5f (better-constrained half %.5f)  (%.1f s)"
                  % (it, float(loss.detach()), e[seen].mean(), e[strong].mean(), time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    qn = q.detach().cpu().numpy()
    alb, fzm = seen.copy(), seen.copy()
    alb[:, 3] = False
    fzm[:, :3] = False
    out = {"iters": a.iters, "image": [a.w, a.h], "spp": a.spp, "depth": a.depth, "level": a.level,
           "fuzz_spsa": a.fuzz_spsa,
           "albedo_abs_err_start": float(np.abs(init - truth)[alb].mean()),
           "albedo_abs_err_end": float(np.abs(qn - truth)[alb].mean()),
           "fuzz_abs_err_start": float(np.abs(init - truth)[fzm].mean()) if fzm.any() else None,
           "fuzz_abs_err_end": float(np.abs(qn - truth)[fzm].mean()) if fzm.any() else None,
           "parameters": int(mask.sum()), "constrained": int(seen.sum()),
           "loss_start": float(loss0.detach()), "loss_end": float(loss.detach()),
           "mean_abs_err_start": err0, "mean_abs_err_end": float(np.abs(qn - truth)[seen].mean()),
           "mean_abs_err_start_strong": err0s, "mean_abs_err_end_strong": float(np.abs(qn - truth)[strong].mean()),
           "seconds": round(time.perf_counter() - t0, 2)}
    print(json.dumps(out), flush=True)
    sc.close()
    return out


if __name__ == "__main__":
    main()
