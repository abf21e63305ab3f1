"""torch.autograd.Function over the differentiable path tracer (a thin test binding).

``render_loss(kernel, maps, target, passes, seed, albedo, emission, turbidity, elevation, azimuth,
checker=None)`` returns the L2 loss sum((film - target)^2) of a `passes`-sample render with the given
parameters, as a float64 tensor whose backward is the path-replay adjoint of the C ABI
(df_loss_gradient).  `maps` is (surface_albedo, surface_emission) or, with checkerboard groups,
(surface_albedo, surface_emission, surface_checker) and `checker` the groups' colours (groups x 2 x 3).
The angles may be floats (held fixed) or tensors.  Every call renders from the same per-pixel seeds
(common random numbers), so the loss is a deterministic function of the parameters.
"""

import numpy as np
import torch


def _np(t):
    return t.detach().cpu().numpy().astype(np.float32)


class RenderLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, emission, checker, turbidity, elevation, azimuth, kernel, maps, target, passes, seed):
        sc = maps[2] if len(maps) > 2 else None
        kernel.setDifferentiable(maps[0], maps[1], _np(albedo), _np(emission), float(turbidity), float(elevation),
                                 float(azimuth), surface_checker=sc, checker=_np(checker) if sc is not None else None)
        kernel.seed(seed)
        loss, grad, _, _ = kernel.lossGradient(target, passes, gradient=True)
        g = torch.from_numpy(grad)
        na, ne = albedo.numel(), emission.numel()
        nc = checker.numel() if sc is not None else 0
        parts = (g[:na].reshape(albedo.shape).to(albedo.dtype),
                 g[na:na + ne].reshape(emission.shape).to(emission.dtype),
                 (g[na + ne:na + ne + nc].reshape(checker.shape) if nc else torch.zeros_like(checker)).to(checker.dtype),
                 g[-3].to(turbidity.dtype), g[-2].to(elevation.dtype), g[-1].to(azimuth.dtype))
        ctx.save_for_backward(*parts)
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        s = gout.to(torch.float64)
        return tuple(s.to(p.dtype) * p for p in ctx.saved_tensors) + (None,) * 5


def render_loss(kernel, maps, target, passes, seed, albedo, emission, turbidity, elevation, azimuth=30.0,
                checker=None):
    def as_tensor(v):
        return v if torch.is_tensor(v) else torch.tensor(float(v), dtype=torch.float64)

    if checker is None:
        checker = torch.zeros(0, dtype=torch.float64)
    return RenderLoss.apply(albedo, emission, checker, as_tensor(turbidity), as_tensor(elevation),
                            as_tensor(azimuth), kernel, maps, target, passes, seed)
