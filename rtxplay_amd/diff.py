"""torch.autograd harness over the differentiable render (forward: rtx_launch
with the material parameters written into the scene; backward:
rtx_launch_adjoint, the path-replay adjoint kernel).

    r = DiffRender(scene, lp)                 # lp: size, spp, depth, camera, handle
    img = r(params)                           # params: cuda float32 [things, 4]
    loss = ((img - target) ** 2).sum(); loss.backward()   # params.grad

torch is plumbing here (device memory, streams, the optimiser); the render
and its adjoint run in librtx.  torch must be imported before librtx is
loaded, so that both use torch's HIP runtime; importing this module first
takes care of that.
"""
import ctypes

import torch

from . import _capi
from ._capi import LpGeneralC, call


def _check_runtime():
    if _capi._lib is not None and not _capi._after_torch:
        raise RuntimeError("librtx was loaded before torch: import rtxplay_amd.diff (or torch) before "
                           "using the library, so that both share one HIP runtime")


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, r):
        p = r.prepare(params)
        img = r.render(p)
        ctx.r = r
        ctx.save_for_backward(img, p)
        return img

    @staticmethod
    def backward(ctx, grad_img):
        img, p = ctx.saved_tensors
        # the adjoint replays the forward's paths with the scene's materials:
        # write this forward's parameters back first (another forward may
        # have set the scene's materials since)
        return ctx.r.adjoint(img, grad_img, p, prepared=True), None


class DiffRender:
    """Renders scene with lp (a LpGeneral: image_w/h, spp, depth, camera,
    is_handle) for material parameters given as a torch tensor on the GPU."""

    def __init__(self, scene, lp, device="cuda"):
        _check_runtime()
        _capi.lib()
        self.scene = scene
        self.n = len(scene)
        self.lp = LpGeneralC()
        ctypes.memmove(ctypes.byref(self.lp), ctypes.byref(lp), ctypes.sizeof(LpGeneralC))
        for k in ("rpp", "normals", "albedos", "pick_id"):
            setattr(self.lp, k, None)
        self.lp.picker = 0
        self.device = torch.device(device)
        self.shape = (self.lp.image_h, self.lp.image_w, 3)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def prepare(self, params):
        """params as a contiguous float32 [things, 4] tensor on the device (a
        copy: later in-place changes of params do not reach it)."""
        p = params.detach().to(self.device, torch.float32).clone().contiguous()
        if p.shape != (self.n, 4):
            raise ValueError("params must be [%d, 4]" % self.n)
        return p

    def _set_materials(self, p, s):
        call("rtx_scene_set_materials", self.scene.ctx, ctypes.c_uint64(self.lp.is_handle),
             ctypes.c_void_p(p.data_ptr()), s)

    def forward(self, params):
        """The image for params (no autograd record; __call__ records one)."""
        return self.render(self.prepare(params))

    def render(self, p):
        """The image for prepared parameters p (see prepare)."""
        s = self._stream()
        self._set_materials(p, s)
        img = torch.empty(self.shape, dtype=torch.float32, device=self.device)
        self.lp.rawRGB = img.data_ptr()
        call("rtx_launch", self.scene.ctx, ctypes.byref(self.lp), 0, s)
        return img

    def adjoint(self, img, grad_img, params=None, prepared=False):
        """dL/d params for the forward image img rendered with params (the
        scene's current materials when None)."""
        s = self._stream()
        if params is not None:
            self._set_materials(params if prepared else self.prepare(params), s)
        g = grad_img.detach().to(self.device, torch.float32).contiguous()
        d = torch.empty((self.n, 4), dtype=torch.float32, device=self.device)
        self.lp.rawRGB = img.data_ptr()
        call("rtx_launch_adjoint", self.scene.ctx, ctypes.byref(self.lp), ctypes.c_void_p(g.data_ptr()),
             ctypes.c_void_p(d.data_ptr()), s)
        return d

    def __call__(self, params):
        return _Render.apply(params, self)


def material_params(scene_things, dtype=torch.float32):
    """float[n,4] (albedo rgb, fuzz) of a sphere_things list."""
    rows = [[float(t["albedo"][0]), float(t["albedo"][1]), float(t["albedo"][2]), float(t["fuzz"])]
            for t in scene_things]
    return torch.tensor(rows, dtype=dtype)
