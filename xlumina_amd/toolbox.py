"""Grid and bookkeeping helpers the propagation path needs (toolbox.py subset)."""
from __future__ import annotations

import numpy as np
import torch


def space(x_total, num_pix):
    """toolbox.py:49-61: x, y = linspace(-x_total, x_total, num_pix) in float64."""
    x = np.linspace(-x_total, x_total, num_pix, dtype=np.float64)
    y = np.linspace(-x_total, x_total, num_pix, dtype=np.float64)
    return x, y


def build_LCD_cell(eta, theta, shape):
    """toolbox.py:121-135: constant (or per-pixel) eta / theta arrays of (shape, shape)."""
    if isinstance(eta, torch.Tensor) or isinstance(theta, torch.Tensor):
        dev = eta.device if isinstance(eta, torch.Tensor) else theta.device
        ones = torch.ones((shape, shape), dtype=torch.float32, device=dev)
        return eta * ones, theta * ones
    return eta * np.ones((shape, shape)), theta * np.ones((shape, shape))


def wrap_phase(phase):
    """toolbox.py:63-72: arctan2(sin, cos)."""
    if isinstance(phase, torch.Tensor):
        return torch.atan2(torch.sin(phase), torch.cos(phase))
    return np.arctan2(np.sin(phase), np.cos(phase))


def is_conserving_energy(light_source, propagated_light):
    """toolbox.py:74-96: total propagated intensity / source intensity."""
    def total(l):
        if l.info in ("Wave optics light", "Wave optics light source"):
            return torch.sum(torch.abs(l.field) ** 2)
        return sum(torch.sum(torch.abs(c) ** 2) for c in (l.Ex, l.Ey, l.Ez))
    return total(propagated_light) / total(light_source)


def softmin(args, beta=90):
    """toolbox.py:98-103: -logsumexp(-beta args) / beta (differentiable min)."""
    a = args if isinstance(args, torch.Tensor) else torch.as_tensor(np.asarray(args, dtype=np.float64))
    return -torch.logsumexp(-beta * a.reshape(-1), dim=0) / beta


def rotate_mask(X, Y, angle, origin=None):
    """toolbox.py:269-289: the frame (X, Y) rotated by `angle` (radians) about
    `origin`; with no origin, about ((X[-1] + X[0]) / 2, (Y[-1] + Y[0]) / 2),
    which for meshgrids is the mean of the first and last rows, as written."""
    X = X if isinstance(X, torch.Tensor) else torch.as_tensor(np.asarray(X, dtype=np.float64))
    Y = Y if isinstance(Y, torch.Tensor) else torch.as_tensor(np.asarray(Y, dtype=np.float64), device=X.device)
    if origin is None:
        x0, y0 = (X[-1] + X[0]) / 2, (Y[-1] + Y[0]) / 2
    else:
        x0, y0 = origin
    a = torch.as_tensor(angle, dtype=torch.float64, device=X.device)
    c, s = torch.cos(a), torch.sin(a)
    return (X - x0) * c + (Y - y0) * s, -(X - x0) * s + (Y - y0) * c
