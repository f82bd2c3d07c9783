"""The adjoint as a torch.autograd.Function (verification harness only):
forward returns the L2 image loss of a scene whose outline points and
substance colours are the inputs, backward returns the HIP adjoint's
gradients."""
from __future__ import annotations

import copy

import numpy as np
import torch

import gudni_amd as G


class RasterL2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pairs, rgba, rasterizer, scene, target):
        sc = copy.copy(scene)
        sc.pairs = pairs.detach().cpu().numpy().astype(np.float32).reshape(scene.pairs.shape)
        sc.substances = rgba.detach().cpu().numpy().astype(np.float32)
        adj = G.Adjoint(rasterizer, [(sc, target)])
        try:
            adj.run()
            loss, dp, dc, _ = adj.result(0)
        finally:
            adj.close()
        ctx.save_for_backward(torch.from_numpy(dp).to(pairs.device).reshape(pairs.shape),
                              torch.from_numpy(dc).to(rgba.device))
        return torch.tensor(loss, dtype=torch.float64, device=pairs.device)

    @staticmethod
    def backward(ctx, grad):
        dp, dc = ctx.saved_tensors
        return (grad.to(dp.dtype) * dp, grad.to(dc.dtype) * dc, None, None, None)


def raster_l2(pairs, rgba, rasterizer, scene, target):
    return RasterL2.apply(pairs, rgba, rasterizer, scene, target)
