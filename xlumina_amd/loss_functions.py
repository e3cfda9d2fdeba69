"""Loss functions of the optical-table optimisers on the device (drop-in for
xlumina/loss_functions.py).

Each loss runs as one fused native reduction (xl_field_loss, xl_small_area):
the forward pass also writes the cotangent of its input when that input
requires a gradient, and the backward pass scales it by the incoming
gradient.  Targets are constants.  Inputs may be torch tensors or NumPy
arrays; results are float32 tensors on the input's GPU.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as N
from . import ops
from .wave_optics import default_device


def _c64(a, device=None) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=device if device is not None else a.device, dtype=torch.complex64)
    return torch.as_tensor(np.array(a), device=device if device is not None else default_device()).to(
        torch.complex64)


def _f32(a, device=None) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=device if device is not None else a.device, dtype=torch.float32)
    return torch.as_tensor(np.asarray(a, dtype=np.float32), device=device if device is not None else default_device())


class _FieldLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, target, mode):
        loss, cot = ops.field_loss(field.detach(), target, mode, want_grad=ctx.needs_input_grad[0])
        ctx.save_for_backward(cot)
        return loss

    @staticmethod
    def backward(ctx, gl):
        (cot,) = ctx.saved_tensors
        if cot is None:
            return None, None, None
        return cot * gl.to(cot.real.dtype).view(-1, *([1] * (cot.dim() - 1))), None, None


class _SmallArea(torch.autograd.Function):
    @staticmethod
    def forward(ctx, intensity):
        loss, g = ops.small_area(intensity.detach(), want_grad=ctx.needs_input_grad[0])
        ctx.save_for_backward(g)
        return loss

    @staticmethod
    def backward(ctx, gl):
        (g,) = ctx.saved_tensors
        return g * gl.view(-1, *([1] * (g.dim() - 1)))


def _field_loss(inp, target, mode):
    f = _c64(inp)
    t = _c64(target, f.device)
    return _FieldLoss.apply(f[None], t, mode)[0]


# -- hybrid (topology + parameters) optimisation -----------------------------------

def small_area_hybrid(detected_intensity):
    """loss_functions.py:24-40: (fraction of the peak 0.7) area of the bright
    spot over its share of the total intensity; gradient with the mask held
    constant."""
    d = _f32(detected_intensity)
    return _SmallArea.apply(d[None])[0]


def vectorized_loss_hybrid(detected_intensities):
    """loss_functions.py:42-49: small_area_hybrid of each (M, N, N) detector -> (M,)."""
    return _SmallArea.apply(_f32(detected_intensities))


# -- field MSEs ----------------------------------------------------------------------

def mean_batch_MSE_Intensity(optimized, target):
    """loss_functions.py:51-63: (mean, per-pair MSE_Intensity) over (B, N, N) pairs."""
    f = _c64(optimized)
    mse = _FieldLoss.apply(f, _c64(target, f.device), N.XL_LOSS_INTENSITY)
    return mse.mean(), mse


def _stack3(light):
    return light.E if hasattr(light, "E") else _c64(light)


def vMSE_Amplitude(input_light, target_light):
    """loss_functions.py:65-80: [MSEx, MSEy, MSEz] in amplitude."""
    return _FieldLoss.apply(_stack3(input_light), _stack3(target_light), N.XL_LOSS_AMPLITUDE)


def vMSE_Phase(input_light, target_light):
    """loss_functions.py:82-97: [MSEx, MSEy, MSEz] in phase."""
    return _FieldLoss.apply(_stack3(input_light), _stack3(target_light), N.XL_LOSS_PHASE)


def vMSE_Intensity(input_light, target_light):
    """loss_functions.py:99-105 (vmap of MSE_Intensity): [MSEx, MSEy, MSEz]."""
    return _FieldLoss.apply(_stack3(input_light), _stack3(target_light), N.XL_LOSS_INTENSITY)


def MSE_Amplitude(input_light, target_light):
    """loss_functions.py:107-120: sum (|E| - |T|)^2 / num_pix."""
    return _field_loss(input_light, target_light, N.XL_LOSS_AMPLITUDE)


def MSE_Phase(input_light, target_light):
    """loss_functions.py:122-135: sum (angle E - angle T)^2 / num_pix."""
    return _field_loss(input_light, target_light, N.XL_LOSS_PHASE)


def MSE_Intensity(input_light, target_light):
    """loss_functions.py:137-149: sum (|E|^2 - |T|^2)^2 / num_pix."""
    return _field_loss(input_light, target_light, N.XL_LOSS_INTENSITY)
