"""torch.autograd wrapper of the differentiable points -> image path.

Forward (all on the GPU through libgvdbx.so):
    SetPoints / SetPointWeights -> RebuildTopology -> FinishTopology ->
    UpdateAtlas -> SplatCubic (rho = sum w (1-|v-p|^2/r^2)^3) ->
    UpdateApron(boundval) -> RaycastEA (I, T per pixel).
Backward:
    RaycastEABackward (dL/d voxel, replayed from per-segment transmittance
    checkpoints) -> ApronAdjoint -> SplatCubicBackward (dL/dp, dL/dw).

The topology depends on the positions only through which bricks are active,
a piecewise-constant function, so it carries no gradient.  The volume must
not be rebuilt or re-splatted between a forward and its backward.
"""
from dataclasses import dataclass, field

import torch

from . import VolumeGVDB, camera_ortho, ea_params


@dataclass
class SplatRenderConfig:
    radius: float = 2.0                 # cubic splat radius (voxels)
    topo_radius: float = 4.0            # radius handed to RebuildTopology
    trans: tuple = (0.0, 0.0, 0.0)
    chan: int = 0
    boundval: float = 0.0               # UpdateApron value outside the tree
    cam: object = None                  # gvdbx.Camera
    rp: object = None                   # gvdbx.RenderParams (ea_params())
    extra: dict = field(default_factory=dict)


class _SplatRenderEA(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, weights, vol, cfg):
        n = positions.shape[0]
        vol.SetStream(torch.cuda.current_stream())
        vol.SetPoints(positions)
        vol.SetPointWeights(weights)
        vol.RebuildTopology(n, cfg.topo_radius, cfg.trans)
        vol.FinishTopology(False, True)
        vol.UpdateAtlas()
        vol.SplatCubic(n, cfg.radius, cfg.trans, cfg.chan, False)
        vol.UpdateApron(cfg.chan, cfg.boundval)
        record = any(ctx.needs_input_grad[:2])
        I, T = vol.RaycastEA(cfg.chan, cfg.cam, cfg.rp, record=record)
        ctx.vol, ctx.cfg, ctx.n = vol, cfg, n
        ctx.want_w = ctx.needs_input_grad[1]
        ctx.save_for_backward(I, T)
        return I, T

    @staticmethod
    def backward(ctx, gI, gT):
        I, T = ctx.saved_tensors
        vol, cfg = ctx.vol, ctx.cfg
        vol.SetStream(torch.cuda.current_stream())
        gI = torch.zeros_like(I) if gI is None else gI.contiguous().float()
        gT = None if gT is None else gT.contiguous().float()
        grad = vol.RaycastEABackward(cfg.chan, cfg.cam, cfg.rp, I, T, gI, gT)
        vol.ApronAdjoint(grad)
        dpos, dw = vol.SplatCubicBackward(ctx.n, cfg.radius, cfg.trans, grad, want_dw=ctx.want_w)
        return dpos, dw, None, None


def splat_render_ea(positions, weights, vol, cfg):
    """(I, T) of the emission-absorption image of the splatted points;
    differentiable in positions (N,3) and weights (N,)."""
    if not positions.is_cuda:
        raise RuntimeError("splat_render_ea needs CUDA tensors (there is no CPU path)")
    return _SplatRenderEA.apply(positions.contiguous(), weights.contiguous(), vol, cfg)


def ortho_view(lo_xy, hi_xy, z0, width, height, step=0.5, kappa=0.05):
    """Orthographic +z view of the rectangle [lo_xy, hi_xy) starting at depth z0."""
    cam = camera_ortho(width, height, (lo_xy[0], lo_xy[1], z0), (hi_xy[0] - lo_xy[0], 0, 0),
                       (0, hi_xy[1] - lo_xy[1], 0), (0, 0, 1))
    return cam, ea_params(step=step, kappa=kappa)


__all__ = ["SplatRenderConfig", "splat_render_ea", "ortho_view", "VolumeGVDB"]
