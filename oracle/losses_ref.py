"""Float64 NumPy restatement of XLuminA's loss functions (TEST ORACLE).

Test infrastructure only: imported by ``tests/``; never by the product
package.  Each function follows one reference function of
``xlumina/loss_functions.py`` / ``xlumina/toolbox.py`` literally.
"""
from __future__ import annotations

import numpy as np
from scipy.special import logsumexp


def small_area_hybrid(detected_intensity):
    """loss_functions.py:24-40."""
    epsilon = 0.7
    eps = 1e-08
    I = detected_intensity / (np.sum(detected_intensity) + eps)
    mask = np.where(I > epsilon * np.max(I), 1, 0)
    return np.sum(mask) / (np.sum(mask * I) + eps)


def small_area_hybrid_grad(detected_intensity):
    """Gradient of small_area_hybrid with the mask held constant (jnp.where
    passes no gradient to its condition): differentiated by hand."""
    eps = 1e-08
    D = np.asarray(detected_intensity, dtype=np.float64)
    S = D.sum() + eps
    I = D / S
    m = (I > 0.7 * I.max()).astype(np.float64)
    A = np.sum(m * I)
    return -m.sum() / (A + eps) ** 2 * (m - A) / S


def vectorized_loss_hybrid(detected_intensities):
    """loss_functions.py:42-49."""
    return np.array([small_area_hybrid(d) for d in detected_intensities])


def MSE_Amplitude(input_light, target_light):
    """loss_functions.py:107-120."""
    num_pix = input_light.shape[0] * input_light.shape[1]
    return np.sum((np.abs(input_light) - np.abs(target_light)) ** 2) / num_pix


def MSE_Phase(input_light, target_light):
    """loss_functions.py:122-135."""
    num_pix = input_light.shape[0] * input_light.shape[1]
    return np.sum((np.angle(input_light) - np.angle(target_light)) ** 2) / num_pix


def MSE_Intensity(input_light, target_light):
    """loss_functions.py:137-149."""
    num_pix = input_light.shape[0] * input_light.shape[1]
    return np.sum((np.abs(input_light) ** 2 - np.abs(target_light) ** 2) ** 2) / num_pix


def vMSE(fn, E_in, E_target):
    """loss_functions.py:65-105: fn of each component (Ex, Ey, Ez)."""
    return np.array([fn(E_in[i], E_target[i]) for i in range(3)])


def mean_batch_MSE_Intensity(optimized, target):
    """loss_functions.py:51-63."""
    mse = np.array([MSE_Intensity(o, t) for o, t in zip(optimized, target)])
    return np.mean(mse), mse


def softmin(args, beta=90):
    """toolbox.py:98-103."""
    return -logsumexp(-beta * np.asarray(args, dtype=np.float64)) / beta
