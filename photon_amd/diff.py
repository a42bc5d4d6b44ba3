"""Differentiable rendering harness: the gfx950 renderer as a torch function.

`DiffRenderer` holds a scene whose first `n_train` materials have trainable
base colour, roughness and metallic, and whose texture 0 (an RGB albedo
texture) has trainable texels.  `DiffRenderer.render(params, texture)` is a
torch.autograd.Function: the forward pass is phx_render (the primal samples),
the backward pass is phx_adjoint, the path-replay adjoint, on a disjoint set
of samples (decorrelated gradient).  All tensors live on the GPU; parameters
reach the renderer by device-to-device copies.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import lib
from .backend import HipBackend, PathTracer

PARAMS_PER_MATERIAL = 5  # base r, g, b, roughness, metallic


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class DiffRenderer:
    def __init__(self, scene, camera, width, height, spp, max_depth, adjoint_spp=None, adjoint_offset=None,
                 n_train=None, device=0):
        self.pt = PathTracer(HipBackend(device=device))
        self.pt.set_scene(scene)
        self.pt.camera = camera
        self.pt.params = abi.default_render_params(width, height, spp, max_depth)
        self.adj_params = abi.default_render_params(width, height, adjoint_spp or spp, max_depth)
        # decorrelated: the adjoint's sample indices follow the primal's
        self.adj_params.sample_offset = spp if adjoint_offset is None else adjoint_offset
        self.n_train = len(scene.materials) if n_train is None else n_train
        self.device = torch.device("cuda", device)
        tex = getattr(scene, "texture_pixels", None)
        self.texture_shape = None
        if tex is not None and len(scene.texture_infos):
            off, tw, th = scene.texture_infos[0]
            self.texture_shape = (int(th), int(tw), 3)
            self.texture_offset = int(off)
        self.atlas_pixels = 0 if tex is None else len(np.asarray(tex).reshape(-1)) // 3
        self.primal_ms = self.adjoint_ms = 0.0

    def initial_params(self, scene):
        """(n_train, 5) tensor of the scene's base colour, roughness, metallic and
        the (h, w, 3) texture 0 tensor (None without a texture)."""
        p = np.array([[m.base_color[0], m.base_color[1], m.base_color[2], m.roughness, m.metallic]
                      for m in scene.materials[:self.n_train]], np.float32)
        tex = None
        if self.texture_shape is not None:
            px = np.asarray(scene.texture_pixels, np.float32).reshape(-1, 3)
            n = self.texture_shape[0] * self.texture_shape[1]
            tex = torch.tensor(px[self.texture_offset:self.texture_offset + n].reshape(self.texture_shape),
                               device=self.device)
        return torch.tensor(p, device=self.device), tex

    def set_params(self, params, texture=None):
        params = params.detach().to(self.device, torch.float32).contiguous()
        if params.shape != (self.n_train, PARAMS_PER_MATERIAL):
            raise ValueError(f"params must be ({self.n_train}, 5)")
        torch.cuda.synchronize()
        self.pt.ctx.check(lib().phx_set_material_params(self.pt.ctx.h, _ptr(params), self.n_train))
        if texture is not None:
            texture = texture.detach().to(self.device, torch.float32).contiguous()
            if tuple(texture.shape) != self.texture_shape:
                raise ValueError(f"texture must be {self.texture_shape}")
            n = self.texture_shape[0] * self.texture_shape[1]
            self.pt.ctx.check(lib().phx_set_texture_pixels(self.pt.ctx.h, _ptr(texture), self.texture_offset, n))

    def primal(self):
        """Renders the primal samples; returns the (h, w, 3) colour tensor."""
        self.pt.render(read=False)
        self.primal_ms = self.pt.ctx.stats().last_render_ms
        h, w = self.pt.params.height, self.pt.params.width
        out = torch.empty((h, w, 3), dtype=torch.float32, device=self.device)
        torch.cuda.synchronize()
        # phx_read_result copies device to device when given device memory
        fp = C.cast(C.c_void_p(out.data_ptr()), C.POINTER(C.c_float))
        self.pt.ctx.check(lib().phx_read_result(self.pt.ctx.h, fp, None, None, None, None))
        return out

    def adjoint(self, grad_image):
        """phx_adjoint with dLoss/dImage = grad_image: (params grad, texture grad)."""
        g = grad_image.detach().to(self.device, torch.float32).contiguous()
        nmat = self.pt._desc.material_count
        gm = torch.zeros(nmat * PARAMS_PER_MATERIAL, dtype=torch.float32, device=self.device)
        gt = torch.zeros(max(self.atlas_pixels, 1) * 3, dtype=torch.float32, device=self.device)
        torch.cuda.synchronize()
        self.pt.ctx.check(lib().phx_adjoint(self.pt.ctx.h, C.byref(self.pt.camera), C.byref(self.adj_params),
                                            _ptr(g), _ptr(gm), _ptr(gt) if self.atlas_pixels else None))
        self.adjoint_ms = self.pt.ctx.stats().last_render_ms
        gp = gm.view(nmat, PARAMS_PER_MATERIAL)[:self.n_train].clone()
        gtex = None
        if self.texture_shape is not None:
            n = self.texture_shape[0] * self.texture_shape[1]
            gtex = gt.view(-1, 3)[self.texture_offset:self.texture_offset + n].reshape(self.texture_shape).clone()
        return gp, gtex

    def render(self, params, texture=None):
        return _RenderFunction.apply(params, texture, self)


class _RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, texture, renderer):
        renderer.set_params(params, texture)
        ctx.renderer = renderer
        ctx.has_texture = texture is not None
        return renderer.primal()

    @staticmethod
    def backward(ctx, grad_image):
        gp, gt = ctx.renderer.adjoint(grad_image)
        return gp, (gt if ctx.has_texture else None), None
