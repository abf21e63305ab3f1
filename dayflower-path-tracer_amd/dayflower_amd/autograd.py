"""torch.autograd.Function over the differentiable path tracer (a thin test binding).

``render_loss(kernel, maps, target, passes, seed, albedo, emission, turbidity, elevation, azimuth)``
returns the L2 loss sum((film - target)^2) of a `passes`-sample render with the given parameters, as a
float64 tensor whose backward is the path-replay adjoint of the C ABI (df_loss_gradient).  Every call
renders from the same per-pixel seeds (common random numbers), so the loss is a deterministic function of
the parameters.
"""

import numpy as np
import torch


class RenderLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, emission, turbidity, elevation, kernel, maps, target, passes, seed, azimuth):
        kernel.setDifferentiable(maps[0], maps[1], albedo.detach().cpu().numpy().astype(np.float32),
                                 emission.detach().cpu().numpy().astype(np.float32), float(turbidity),
                                 float(elevation), float(azimuth))
        kernel.seed(seed)
        loss, grad, _, _ = kernel.lossGradient(target, passes, gradient=True)
        na = albedo.numel()
        ne = emission.numel()
        g = torch.from_numpy(grad)
        ctx.save_for_backward(g[:na].reshape(albedo.shape).to(albedo.dtype),
                              g[na:na + ne].reshape(emission.shape).to(emission.dtype),
                              g[na + ne].to(turbidity.dtype), g[na + ne + 1].to(elevation.dtype))
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        ga, ge, gt, gel = ctx.saved_tensors
        s = gout.to(ga.dtype)
        return s * ga, s * ge, s * gt, s * gel, None, None, None, None, None, None


def render_loss(kernel, maps, target, passes, seed, albedo, emission, turbidity, elevation, azimuth=30.0):
    return RenderLoss.apply(albedo, emission, turbidity, elevation, kernel, maps, target, passes, seed, azimuth)
