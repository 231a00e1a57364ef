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
        img = r.forward(params)
        ctx.r = r
        ctx.save_for_backward(img)
        return img

    @staticmethod
    def backward(ctx, grad_img):
        img, = ctx.saved_tensors
        return ctx.r.adjoint(img, grad_img), None


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

    def forward(self, params):
        p = params.detach().to(self.device, torch.float32).contiguous()
        if p.shape != (self.n, 4):
            raise ValueError("params must be [%d, 4]" % self.n)
        s = self._stream()
        call("rtx_scene_set_materials", self.scene.ctx, ctypes.c_uint64(self.lp.is_handle),
             ctypes.c_void_p(p.data_ptr()), s)
        img = torch.empty(self.shape, dtype=torch.float32, device=self.device)
        self.lp.rawRGB = img.data_ptr()
        call("rtx_launch", self.scene.ctx, ctypes.byref(self.lp), 0, s)
        return img

    def adjoint(self, img, grad_img):
        g = grad_img.detach().to(self.device, torch.float32).contiguous()
        d = torch.empty((self.n, 4), dtype=torch.float32, device=self.device)
        self.lp.rawRGB = img.data_ptr()
        call("rtx_launch_adjoint", self.scene.ctx, ctypes.byref(self.lp), ctypes.c_void_p(g.data_ptr()),
             ctypes.c_void_p(d.data_ptr()), self._stream())
        return d

    def __call__(self, params):
        return _Render.apply(params, self)


def material_params(scene_things, dtype=torch.float32):
    """float[n,4] (albedo rgb, fuzz) of a sphere_things list."""
    rows = [[float(t["albedo"][0]), float(t["albedo"][1]), float(t["albedo"][2]), float(t["fuzz"])]
            for t in scene_things]
    return torch.tensor(rows, dtype=dtype)
