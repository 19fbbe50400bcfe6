"""The calibration's forward / backward pair (gs_calib_forward /
gs_calib_backward) as a torch.autograd.Function, so the 14 parameters can be
fitted with any torch optimizer.  torch is plumbing here: the loss and its
gradient come from the library's kernels.

    cal = Calibration(ensemble, CalibConfig(...))
    theta = torch.tensor(calib_params(), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=0.05)
    for _ in range(20):
        opt.zero_grad()
        loss = calibration_loss(cal, theta)
        loss.backward()
        opt.step()

The gradient is that of the last forward pass (the library keeps one set of
checkpoints), which is the one autograd differentiates.
"""
from __future__ import annotations

import numpy as np
import torch


class CalibrationLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, calib) -> torch.Tensor:
        th = theta.detach().to("cpu", torch.float64).numpy()
        loss = calib.forward(np.ascontiguousarray(th))
        ctx.calib = calib
        return torch.tensor(loss, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g = torch.from_numpy(ctx.calib.backward()).to(grad_out.device, grad_out.dtype)
        return grad_out * g, None


def calibration_loss(calib, theta: torch.Tensor) -> torch.Tensor:
    """Loss of the calibration model at theta (14 parameters, gs_calib_params
    order), differentiable with respect to theta."""
    return CalibrationLoss.apply(theta, calib)
