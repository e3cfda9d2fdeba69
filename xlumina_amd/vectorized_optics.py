"""Vectorial fields (Ex, Ey, Ez) and their propagation (drop-in for
xlumina/vectorized_optics.py).

The three components live in one (3, N, N) complex64 tensor; ``Ex``, ``Ey``,
``Ez`` are views.  Jones masks applied with optical_elements (sSLM, LCD,
linear_polarizer) stay pending on the returned light (up to two in a row) and
are fused, together with the reference's Ez construction, into the first pass
of the next VRS / VCZT / objective-lens propagation.
"""
from __future__ import annotations

import math
import time
from typing import Optional, Tuple

import numpy as np
import torch

from . import _native as N
from . import autograd as AG
from . import ops
from .ops import MaskSpec
from .wave_optics import _GridMixin, _field_tensor, _grid, _zf, default_device, gaussian_envelope, minimum_z
from . import wave_optics as _wo


class VectorizedLight(_GridMixin):
    """Vectorial EM field (vectorized_optics.py:36-49)."""

    def __init__(self, x=None, y=None, wavelength=None, device: Optional[torch.device] = None):
        self.x = _grid(x)
        self.y = _grid(y)
        self.wavelength = wavelength
        self.k = 2 * math.pi / wavelength
        self.n = 1
        self.device = torch.device(device) if device is not None else default_device()
        self._X = self._Y = None
        self._E: Optional[torch.Tensor] = None          # zeros, allocated on first read
        self._src: Optional[torch.Tensor] = None
        self._masks: Tuple[MaskSpec, ...] = ()
        self.info = "Vectorized light"

    # -- components ---------------------------------------------------------------
    @property
    def E(self) -> torch.Tensor:
        """(3, N, N) complex64 stack of (Ex, Ey, Ez)."""
        if self._masks:
            self._E = AG.pre(self._src, self.x, self.y, masks=self._masks, nin=self._src.shape[0], ncomp=3)
            self._src, self._masks = None, ()
        elif self._E is None:
            self._E = torch.zeros((3, len(self.x), len(self.y)), dtype=torch.complex64, device=self.device)
        return self._E

    @E.setter
    def E(self, v) -> None:
        self._E = _field_tensor(v, None, self.device).reshape(3, len(self.y), len(self.x))
        self._src, self._masks = None, ()

    def _set_comp(self, i, v):
        v = _field_tensor(v, None, self.device)
        cur = self.E
        self._E = torch.stack([v if j == i else cur[j] for j in range(3)])

    Ex = property(lambda self: self.E[0], lambda self, v: self._set_comp(0, v))
    Ey = property(lambda self: self.E[1], lambda self, v: self._set_comp(1, v))
    Ez = property(lambda self: self.E[2], lambda self, v: self._set_comp(2, v))

    def _pending(self):
        if self._masks:
            return self._src, self._masks
        return self.E, ()

    def _with_mask(self, spec: MaskSpec) -> "VectorizedLight":
        src, masks = self._pending()
        if len(masks) >= 2:
            src, masks = self.E, ()
        out = VectorizedLight(self.x, self.y, self.wavelength, self.device)
        out._src, out._masks = src, tuple(masks) + (spec,)
        return out

    # -- propagation ------------------------------------------------------------------
    def VRS_propagation(self, z):
        """vectorized_optics.py:244-284: Ez_in = (Ex X + Ey Y)/r, RS of (Ex, Ey, Ez_in)
        with one shared kernel -> (light, quality factor)."""
        tic = time.perf_counter()
        src, masks = self._pending()
        qf = ops.quality_factor_buffer(z)   # a device distance: the propagation launch reports it
        out = AG.rs(src, self.x, self.y, z, self.wavelength, masks=masks, ez_mode=N.XL_EZ_VRS, nin=2,
                    quality_factor=qf)
        light = VectorizedLight(self.x, self.y, self.wavelength, self.device)
        light._E = out
        if qf is None:
            qf = ops.quality_factor(self.x, self.y, self.wavelength, z)
        if _wo.VERBOSE:
            print(f"Time taken to perform one VRS propagation (in seconds): {(time.perf_counter() - tic):.4f}")
        return light, qf

    def get_VRS_minimum_z(self, n=1, quality_factor=1):
        """vectorized_optics.py:286-319; prints and returns the distance."""
        z_min = minimum_z(self.x, self.y, self.wavelength, n, quality_factor)
        print("Minimum distance to propagate (in um):", z_min)
        return z_min

    def VCZT(self, z, xout=None, yout=None):
        """vectorized_optics.py:321-361: Ez_in = (Ex X/r + Ey Y/r) z/r, CZT of the 3 components."""
        tic = time.perf_counter()
        xout = self.x if xout is None else _grid(xout)
        yout = self.y if yout is None else _grid(yout)
        src, masks = self._pending()
        zt = z if isinstance(z, torch.Tensor) and z.requires_grad else None    # dL/dz wanted
        geom = ops.czt_geometry(self.x, self.y, xout, yout, _zf(z), self.wavelength)
        out = AG.czt(src, geom, masks=masks, ez_mode=N.XL_EZ_VCZT, nin=2, z=zt)
        light = VectorizedLight(xout, yout, self.wavelength, self.device)
        light._E = out
        if _wo.VERBOSE:
            print(f"Time taken to perform one VCZT propagation (in seconds):  {(time.perf_counter() - tic):.4f}")
        return light


def VRS_propagation_jit(input_field, z, nx, ny, dx, dy, Xext, Yext, k):
    """vectorized_optics.py:364-373: RS of each of the 3 components (no Ez rebuild)."""
    xr = _wo._grid_from_lags(np.asarray(Xext)[0, :], nx)
    yr = _wo._grid_from_lags(np.asarray(Yext)[:, 0], ny)
    f = input_field if isinstance(input_field, torch.Tensor) else torch.as_tensor(input_field)
    f = f.to(default_device()).to(torch.complex64)
    return AG.rs(f, xr, yr, float(z), 2 * math.pi / k)


def VCZT_jit(field, z, wavelength, k, nx, ny, dx, dy, Xout, Yout, X, Y, Dm, fy_1, fy_2, fx_1, fx_2):
    """vectorized_optics.py:375-384: CZT_jit of each of the 3 components."""
    xo, yo = _grid(np.asarray(Xout)[0, :]), _grid(np.asarray(Yout)[:, 0])
    xi = _grid(X[0, :].cpu() if isinstance(X, torch.Tensor) else np.asarray(X)[0, :])
    yi = _grid(Y[:, 0].cpu() if isinstance(Y, torch.Tensor) else np.asarray(Y)[:, 0])
    geom = ops.CztGeometry(xi, yi, xo, yo, float(z), float(wavelength), float(Dm), N.XL_CZT_RS,
                           complex(z * dx * dy * wavelength))
    f = field if isinstance(field, torch.Tensor) else torch.as_tensor(field)
    return AG.czt(f.to(default_device()).to(torch.complex64), geom)


def vectorized_CZT_for_high_NA(field, nx, ny, Dm, fy_1, fy_2, fx_1, fx_2):
    """vectorized_optics.py:386-394."""
    return _wo.CZT_for_high_NA_jit(field, nx, ny, Dm, fy_1, fy_2, fx_1, fx_2)


class PolarizedLightSource(VectorizedLight):
    """Polarized light sources (vectorized_optics.py:396-490)."""

    def __init__(self, x, y, wavelength, device: Optional[torch.device] = None):
        super().__init__(x, y, wavelength, device)
        self.info = "Vectorized light source"

    def gaussian_beam(self, w0, jones_vector, center=(0, 0), z_w0=(0, 0), alpha=0):
        """vectorized_optics.py:402-465."""
        amp, phase = gaussian_envelope(self.X, self.Y, self.k, self.n, w0, center, z_w0, alpha)
        nj = np.array(jones_vector, dtype=np.float64)
        nj = nj / np.linalg.norm(nj)
        Ex = (float(nj[0]) * amp) * phase
        Ey = (float(nj[1]) * amp) * phase
        self.E = torch.stack([Ex, Ey, torch.zeros_like(Ex)])

    def plane_wave(self, jones_vector, theta=0, phi=0, z0=0):
        """vectorized_optics.py:467-490."""
        nj = np.array(jones_vector, dtype=np.float64)
        nj = nj / np.linalg.norm(nj)
        pw = torch.exp(1j * self.k * (self.X * math.sin(theta) * math.cos(phi)
                                      + self.Y * math.sin(theta) * math.sin(phi) + z0 * math.cos(theta)))
        self.E = torch.stack([float(nj[0]) * pw, float(nj[1]) * pw, torch.zeros_like(pw)])

    def radial_beam(self, w0):
        """Radially polarised Gaussian Ex = G X/rho, Ey = G Y/rho, G = exp(-rho^2/w0^2)
        (not in the reference, which builds radial polarisation with a quadrant sSLM)."""
        rho = torch.sqrt(self.X ** 2 + self.Y ** 2)
        G = torch.exp(-rho ** 2 / w0 ** 2)
        self.E = torch.stack([G * self.X / rho, G * self.Y / rho, torch.zeros_like(rho)])
