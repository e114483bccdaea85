"""torch.autograd.Function over the differentiable direct-illumination render.

The reference's Python package (virapy, source/virapy/init_bindings.cpp:75-80)
binds no renderer and vira has no differentiable path; this module is the new
differentiable entry point.  Forward: DEM heights and an albedo texture ->
radiance image (H, W, 8) of the fixed-visibility direct-illumination model
(``vira_gpu_di_render``).  Backward: the hand-written adjoint
(``vira_gpu_adjoint``) gives dL/dheights and dL/dtexels; visibility (hit
triangles, barycentrics, shadow-ray results) is held fixed at the forward
pass's values.  The adjoint reads the forward pass's hit records when they fit
in ``DIOptions.record_budget`` and re-traces them otherwise.
"""
from __future__ import annotations

import torch

from . import scene as S


class DifferentiableDem:
    """A DEM mesh on the GPU whose heights and albedo texture are torch leaves.

    ``ctx`` is a ``scene.Context`` whose TLAS holds ``mesh`` (from
    ``Context.dem_mesh`` + ``blas_build`` + ``tlas_build``); lights must be
    point lights.  Calling the object renders and records the graph.
    """

    def __init__(self, ctx: S.Context, mesh: int, camera: S.Camera, options: S.DIOptions | None = None,
                 albedo=1.0, smooth_shading: bool = False):
        self.ctx = ctx
        self.mesh = mesh
        self.camera = camera
        self.options = options or S.DIOptions()
        self.albedo = albedo
        self.smooth_shading = smooth_shading
        self.version = 0          # bumped whenever the GPU mesh takes new heights / texture
        self._state = None        # (heights, texture) currently on the GPU

    def _apply(self, heights: torch.Tensor, texture: torch.Tensor | None):
        h = heights.detach().contiguous()
        t = None if texture is None else texture.detach().contiguous()
        self.ctx.dem_update(self.mesh, h)
        self.ctx.material_set(self.mesh, albedo=self.albedo, texture=t, smooth_shading=self.smooth_shading)
        self._state = (h, t)
        self.version += 1

    def __call__(self, heights: torch.Tensor, texture: torch.Tensor | None = None) -> torch.Tensor:
        return DemDIRender.apply(heights, texture, self)


class DemDIRender(torch.autograd.Function):
    @staticmethod
    def forward(fctx, heights, texture, dem: DifferentiableDem):
        dem._apply(heights, texture)
        img = dem.ctx.di_render(dem.camera, dem.options)
        fctx.dem = dem
        fctx.version = dem.version
        fctx.has_texture = texture is not None
        fctx.save_for_backward(heights.detach(), None if texture is None else texture.detach())
        return img

    @staticmethod
    def backward(fctx, grad_image):
        dem: DifferentiableDem = fctx.dem
        heights, texture = fctx.saved_tensors
        if dem.version != fctx.version:
            # another forward pass ran since: restore this pass's geometry (the
            # adjoint then re-traces the hit records)
            dem._apply(heights, texture)
        gh = torch.empty_like(heights) if fctx.needs_input_grad[0] else None
        gt = torch.empty_like(texture) if (fctx.has_texture and fctx.needs_input_grad[1]) else None
        if gh is not None or gt is not None:
            dem.ctx.adjoint(dem.camera, dem.options, grad_image.contiguous(), gh, gt)
        return gh, gt, None
