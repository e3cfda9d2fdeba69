"""Scalar fields and their propagation (drop-in for xlumina/wave_optics.py).

``ScalarLight.RS_propagation`` and ``ScalarLight.CZT`` run the HIP kernels of
libxlumina_hip.so; SLM phase masks applied to a light (optical_elements.SLM)
stay pending on the returned light and are fused into the next propagation's
first pass.  Fields are complex64 tensors on a ROCm device; grids ``x, y`` stay
float64 NumPy arrays as in the reference.
"""
from __future__ import annotations

import math
import os
import time
from typing import Optional, Tuple

import numpy as np
import torch

from . import _native as N
from . import autograd as AG
from . import ops
from .ops import MaskSpec

VERBOSE = os.environ.get("XL_VERBOSE", "0") == "1"


def set_verbose(flag: bool) -> None:
    """Print the per-call timing line the reference prints (off by default)."""
    global VERBOSE
    VERBOSE = bool(flag)


def default_device() -> torch.device:
    if torch.cuda.is_available():
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device("cpu")


def _grid(a) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def _field_tensor(v, shape, device) -> torch.Tensor:
    if not isinstance(v, torch.Tensor):
        v = torch.as_tensor(np.asarray(v))
    if v.dtype != torch.complex64:
        v = v.to(torch.complex64)
    return v.to(device)


class _GridMixin:
    """Meshgrid X, Y (rows = y) as float64 device tensors, built on first use."""

    @property
    def X(self) -> torch.Tensor:
        if getattr(self, "_X", None) is None:
            self._build_mesh()
        return self._X

    @property
    def Y(self) -> torch.Tensor:
        if getattr(self, "_Y", None) is None:
            self._build_mesh()
        return self._Y

    def _build_mesh(self):
        xt = torch.as_tensor(self.x, dtype=torch.float64, device=self.device)
        yt = torch.as_tensor(self.y, dtype=torch.float64, device=self.device)
        self._Y, self._X = torch.meshgrid(yt, xt, indexing="ij")


class ScalarLight(_GridMixin):
    """Scalar field U(x, y) (wave_optics.py:43-53)."""

    def __init__(self, x, y, wavelength, device: Optional[torch.device] = None):
        self.x = _grid(x)
        self.y = _grid(y)
        self.wavelength = wavelength
        self.k = 2 * math.pi / wavelength
        self.n = 1
        self.device = torch.device(device) if device is not None else default_device()
        self._X = self._Y = None
        self._field: Optional[torch.Tensor] = None     # zeros, allocated on first read
        self._src: Optional[torch.Tensor] = None
        self._masks: Tuple[MaskSpec, ...] = ()
        self.info = "Wave optics light"

    # -- field with pending (fused) SLM masks ---------------------------------
    @property
    def field(self) -> torch.Tensor:
        if self._masks:
            self._field = AG.pre(self._src, self.x, self.y, masks=self._masks)[0]
            self._src, self._masks = None, ()
        elif self._field is None:
            self._field = torch.zeros((len(self.x), len(self.y)), dtype=torch.complex64, device=self.device)
        return self._field

    @field.setter
    def field(self, v) -> None:
        self._field = _field_tensor(v, (len(self.y), len(self.x)), self.device)
        self._src, self._masks = None, ()

    def _pending(self):
        """(source (1, n, n), mask chain) for a fused propagation."""
        if self._masks:
            return self._src, self._masks
        f = self.field
        return f.reshape((1,) + tuple(f.shape)), ()

    def _with_mask(self, spec: MaskSpec) -> "ScalarLight":
        src, masks = self._pending()
        if len(masks) >= 2:
            src, masks = self.field.reshape((1,) + tuple(self.field.shape)), ()
        out = ScalarLight(self.x, self.y, self.wavelength, self.device)
        out._src, out._masks = src, tuple(masks) + (spec,)
        return out

    # -- aperture masks (wave_optics.py:96-171) ------------------------------------
    def _times(self, mask) -> "ScalarLight":
        out = ScalarLight(self.x, self.y, self.wavelength, self.device)
        out._field = self.field * mask.to(self.device)
        return out

    def apply_circular_mask(self, r):
        """wave_optics.py:96-109: field times the ellipse X^2/rx^2 + Y^2/ry^2 < 1."""
        from .optical_elements import circular_mask
        return self._times(circular_mask(self.X, self.Y, r))

    def apply_triangular_mask(self, r, angle, m, height):
        """wave_optics.py:111-133: field times the triangle of optical_elements'
        triangular_mask, with `angle` in radians here (degrees there)."""
        from .optical_elements import _triangle
        return self._times(_triangle(self.X, self.Y, r, angle, m, height))

    def apply_rectangular_mask(self, center, width, height, angle):
        """wave_optics.py:135-153: field times the rectangle (angle in degrees)."""
        from .optical_elements import rectangular_mask
        return self._times(rectangular_mask(self.X, self.Y, center, width, height, angle))

    def apply_annular_aperture(self, di, do):
        """wave_optics.py:155-171: field times the ring between diameters di and do."""
        from .optical_elements import annular_aperture
        return self._times(annular_aperture(di, do, self.X, self.Y))

    # -- propagation ------------------------------------------------------------
    def RS_propagation(self, z):
        """wave_optics.py:173-196: Rayleigh-Sommerfeld to distance z -> (light, quality factor)."""
        tic = time.perf_counter()
        src, masks = self._pending()
        qf = ops.quality_factor_buffer(z)   # a device distance: the propagation launch reports it
        out = AG.rs(src, self.x, self.y, z, self.wavelength, masks=masks, quality_factor=qf)
        light = ScalarLight(self.x, self.y, self.wavelength, self.device)
        light._field = out[0]
        if qf is None:
            qf = ops.quality_factor(self.x, self.y, self.wavelength, z)
        if VERBOSE:
            print(f"Time taken to perform one RS propagation (in seconds): {(time.perf_counter() - tic):.4f}")
        return light, qf

    def get_RS_minimum_z(self, n=1, quality_factor=1):
        """wave_optics.py:198-231; prints the distance like the reference and also returns it."""
        z_min = minimum_z(self.x, self.y, self.wavelength, n, quality_factor)
        print("Minimum distance to propagate (in microns):", z_min)
        return z_min

    def CZT(self, z, xout=None, yout=None):
        """wave_optics.py:233-263: chirp z-transform propagation onto (xout, yout)."""
        tic = time.perf_counter()
        xout = self.x if xout is None else _grid(xout)
        yout = self.y if yout is None else _grid(yout)
        src, masks = self._pending()
        zt = z if isinstance(z, torch.Tensor) and z.requires_grad else None    # dL/dz wanted
        geom = ops.czt_geometry(self.x, self.y, xout, yout, _zf(z), self.wavelength)
        out = AG.czt(src, geom, masks=masks, z=zt)
        light = ScalarLight(xout, yout, self.wavelength, self.device)
        light._field = out[0]
        if VERBOSE:
            print(f"Time taken to perform one CZT propagation (in seconds): {(time.perf_counter() - tic):.4f}")
        return light


def _zf(z) -> float:
    return float(z.detach().reshape(-1)[0]) if isinstance(z, torch.Tensor) else float(z)


def minimum_z(x, y, wavelength, n=1, quality_factor=1) -> float:
    """Minimum distance for a given quality factor (wave_optics.py:211-229)."""
    x = _grid(x)
    y = _grid(y)
    range_x = x[-1] - x[0]
    range_y = y[-1] - y[0]
    dx = range_x / np.size(x)
    dy = range_y / np.size(y)
    dr_real = math.sqrt(dx ** 2 + dy ** 2)
    rmax = math.sqrt(range_x ** 2 + range_y ** 2)
    factor = (((quality_factor * dr_real + rmax) ** 2 - (wavelength / n) ** 2 - rmax ** 2)
              / (2 * wavelength / n)) ** 2 - rmax ** 2
    return math.sqrt(factor) if factor > 0 else 0


class LightSource(ScalarLight):
    """Scalar light sources (wave_optics.py:462-542)."""

    def __init__(self, x, y, wavelength, device: Optional[torch.device] = None):
        super().__init__(x, y, wavelength, device)
        self.info = "Wave optics light source"

    def gaussian_beam(self, w0, E0, center=(0, 0), z_w0=(0, 0), alpha=0):
        """wave_optics.py:468-523 (float64 on the device, stored as complex64)."""
        amp, phase = gaussian_envelope(self.X, self.Y, self.k, self.n, w0, center, z_w0, alpha)
        self.field = (E0 * amp) * phase

    def plane_wave(self, A=1, theta=0, phi=0, z0=0):
        """wave_optics.py:525-542."""
        self.field = A * torch.exp(1j * self.k * (self.X * math.sin(theta) * math.cos(phi)
                                                  + self.Y * math.sin(theta) * math.sin(phi)
                                                  + z0 * math.cos(theta)))


def gaussian_envelope(X, Y, k, n, w0, center, z_w0, alpha):
    """Amplitude and phase of the reference's Gaussian beam (wave_optics.py:481-520)."""
    w0_x, w0_y = w0
    x0, y0 = center
    z_w0x, z_w0y = z_w0
    Rayleigh_x = k * w0_x ** 2 * n / 2
    Rayleigh_y = k * w0_y ** 2 * n / 2
    Gouy_phase_x = math.atan2(z_w0x, Rayleigh_x)
    Gouy_phase_y = math.atan2(z_w0y, Rayleigh_y)
    w_x = w0_x * math.sqrt(1 + (z_w0x / Rayleigh_x) ** 2)
    w_y = w0_y * math.sqrt(1 + (z_w0y / Rayleigh_y) ** 2)
    R_x = 1e12 if z_w0x == 0 else z_w0x * (1 + (Rayleigh_x / z_w0x) ** 2)
    # the reference tests z_w0x for R_y too (wave_optics.py:507)
    R_y = 1e12 if z_w0x == 0 else z_w0y * (1 + (Rayleigh_y / z_w0y) ** 2)
    x_rot = X * math.cos(alpha) + Y * math.sin(alpha)
    y_rot = -X * math.sin(alpha) + Y * math.cos(alpha)
    phase = torch.exp(-1j * ((k * z_w0x + k * X ** 2 / (2 * R_x) - Gouy_phase_x)
                             + (k * z_w0y + k * Y ** 2 / (2 * R_y) - Gouy_phase_y)))
    amp = (w0_x / w_x) * (w0_y / w_y) * torch.exp(-(x_rot - x0) ** 2 / (w_x ** 2) - (y_rot - y0) ** 2 / (w_y ** 2))
    return amp, phase


# ---------------------------------------------------------------------------
# module-level functions of wave_optics.py
# ---------------------------------------------------------------------------

def build_grid(x, y):
    """wave_optics.py:265-279: (nx, ny, dx, dy, Xext, Yext) on the monotone lag grid."""
    x = _grid(x)
    y = _grid(y)
    nx, ny = len(x), len(y)
    dx, dy = x[1] - x[0], y[1] - y[0]
    x_padded = np.pad((x[0] - x[::-1]), (0, np.size(x) - 1), "reflect")
    y_padded = np.pad((y[0] - y[::-1]), (0, np.size(y) - 1), "reflect")
    III = np.concatenate((np.ones(len(x_padded) // 2 + 1), -np.ones(len(x_padded) // 2)))
    Xext, Yext = np.meshgrid(x_padded * III, y_padded * III)
    return nx, ny, dx, dy, Xext, Yext


def _grid_from_lags(ext_row, n):
    """Samples relative to the first one, recovered from a lag row of build_grid."""
    ext_row = _grid(ext_row)
    return ext_row[n - 1:]   # lags 0 .. (n-1) d == x - x[0]


def RS_propagation_jit(input_field, z, nx, ny, dx, dy, Xext, Yext, k):
    """wave_optics.py:281-289 on the lag grid of build_grid; returns the (ny, nx) field."""
    xr = _grid_from_lags(np.asarray(Xext)[0, :], nx)
    yr = _grid_from_lags(np.asarray(Yext)[:, 0], ny)
    field = input_field if isinstance(input_field, torch.Tensor) else torch.as_tensor(input_field)
    field = field.to(default_device()).to(torch.complex64)
    return AG.rs(field.reshape(1, ny, nx), xr, yr, float(z), 2 * math.pi / k)[0]


def transfer_function_RS(z, X, Y, k):
    """wave_optics.py:291-297 on a meshgrid (X varies along rows, Y along columns)."""
    x = _grid(X[0, :] if not isinstance(X, torch.Tensor) else X[0, :].cpu())
    y = _grid(Y[:, 0] if not isinstance(Y, torch.Tensor) else Y[:, 0].cpu())
    dev = X.device if isinstance(X, torch.Tensor) else default_device()
    return ops.transfer_function_rs(x, y, float(z), 2 * math.pi / k, dev)


def build_CZT_grid(z, wavelength, xin, yin, xout, yout):
    """wave_optics.py:299-331."""
    xin, yin, xout, yout = _grid(xin), _grid(yin), _grid(xout), _grid(yout)
    nx, ny = len(xout), len(yout)
    Xout, Yout = np.meshgrid(xout, yout)
    dx = xin[1] - xin[0]
    dy = yin[1] - yin[0]
    Dm = wavelength * z / dx
    return nx, ny, dx, dy, Xout, Yout, Dm, yout[0] + Dm / 2, yout[-1] + Dm / 2, xout[0] + Dm / 2, xout[-1] + Dm / 2


def compute_np2(x):
    """wave_optics.py:373-383."""
    return int(2 ** int(np.ceil(np.log2(x))))


def _czt_core_geometry(x, y, xout, yout, Dm, wavelength, mode, cte, z=0.0):
    return ops.CztGeometry(_grid(x), _grid(y), _grid(xout), _grid(yout), float(z), float(wavelength), float(Dm),
                           mode, complex(cte))


def CZT_jit(field, z, wavelength, k, nx, ny, dx, dy, Xout, Yout, X, Y, Dm, fy_1, fy_2, fx_1, fx_2):
    """wave_optics.py:333-357 (Xout/Yout/X/Y meshgrids give the sample vectors;
    the window limits follow from them and Dm as in build_CZT_grid)."""
    xo, yo = _grid(np.asarray(Xout)[0, :]), _grid(np.asarray(Yout)[:, 0])
    xi = _grid(X[0, :].cpu() if isinstance(X, torch.Tensor) else np.asarray(X)[0, :])
    yi = _grid(Y[:, 0].cpu() if isinstance(Y, torch.Tensor) else np.asarray(Y)[:, 0])
    geom = ops.CztGeometry(xi, yi, xo, yo, float(z), float(wavelength), float(Dm), N.XL_CZT_RS,
                           complex(z * dx * dy * wavelength))
    f = field if isinstance(field, torch.Tensor) else torch.as_tensor(field)
    f = f.to(default_device()).to(torch.complex64)
    return AG.czt(f.reshape((1,) + tuple(f.shape[-2:])), geom)[0]


def CZT_for_high_NA_jit(field, nx, ny, Dm, fy_1, fy_2, fx_1, fx_2):
    """wave_optics.py:359-371: the two Bluestein passes alone.  The output
    samples are recovered from the window limits: out[0] = f_1 - Dm/2,
    out[-1] = f_2 - Dm/2, evenly spaced (as build_high_NA_VCZT_grid defines)."""
    xo = np.linspace(fx_1 - Dm / 2, fx_2 - Dm / 2, nx)
    yo = np.linspace(fy_1 - Dm / 2, fy_2 - Dm / 2, ny)
    f = field if isinstance(field, torch.Tensor) else torch.as_tensor(field)
    f = f.to(default_device()).to(torch.complex64)
    f = f.reshape((-1,) + tuple(f.shape[-2:]))
    n_in_y, n_in_x = f.shape[-2:]
    geom = ops.CztGeometry(np.arange(n_in_x, dtype=np.float64), np.arange(n_in_y, dtype=np.float64), xo, yo, 0.0,
                           1.0, float(Dm), N.XL_CZT_PLAIN, 1.0)
    out = AG.czt(f, geom)
    return out[0] if out.shape[0] == 1 else out
