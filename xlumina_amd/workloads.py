"""Seeded synthetic inputs of the propagation workloads (shared by the
benchmark, the tests and the serving generator).

  scalar_rs_1024   1024^2, x,y in [-1.2, 1.2] mm, lambda 650 nm, Gaussian w0 0.4 mm
                   times a seeded SLM phase U[-pi, pi], RS to z = 50 mm
  vrs_vczt_2048    2048^2 radially polarised Gaussian (lambda 635 nm, window
                   2.4 mm) through an LCD with seeded per-pixel eta, theta in
                   U[0, 2 pi], VRS to z = 30 mm, then VCZT to a 512^2 ROI of
                   0.2 mm at z = 30 mm
  high_na_2048     2048^2 pupil of radius 1.8 mm, f = 2 mm (NA 0.9), lambda
                   650 nm, Gaussian w0 1.2 mm Jones (1, 1), seeded sSLM phases
                   U[-pi, pi] and a constant LCD, focused by VCZT_objective_lens
                   to a 256^2 window of 4 um; loss = intensity second moment
  SharpFocusTable  the hybrid sharp-focus optical table (1024^2 -> six 400^2
                   high-NA detectors) with seeded raw parameters, forward +
                   backward of its softmin small-area loss
Lengths in microns.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N
from . import autograd as AG
from . import ops
from .ops import MaskSpec
from .toolbox import space


def _c64(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a).astype(np.complex64)).to(dev)


def _f32(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def gaussian(x, y, wavelength, w0):
    """Gaussian beam at its waist with the reference's R = 1e12 phase (wave_optics.py:468-523)."""
    X, Y = np.meshgrid(x, y)
    k = 2 * np.pi / wavelength
    phase = np.exp(-1j * (k * X ** 2 / (2 * 1e12) + k * Y ** 2 / (2 * 1e12)))
    return np.exp(-X ** 2 / w0 ** 2 - Y ** 2 / w0 ** 2) * phase


@dataclass
class ScalarRS:
    x: np.ndarray
    y: np.ndarray
    wavelength: float
    z: float
    field: torch.Tensor      # (1, n, n)
    phase: torch.Tensor      # (n, n) float32

    def run(self, out=None, workspace=None):
        return ops.rs_propagate(self.field, self.x, self.y, self.z, self.wavelength,
                                masks=[MaskSpec(N.XL_MASK_SLM, self.phase)], intensity=True, out=out,
                                workspace=workspace)


def scalar_rs_1024(device, n: int = 1024) -> ScalarRS:
    x, y = space(1200.0, n)
    u = gaussian(x, y, 0.65, 400.0)
    phi = np.random.default_rng(1).uniform(-np.pi, np.pi, (n, n)).astype(np.float32)
    return ScalarRS(x, y, 0.65, 5e4, _c64(u[None], device), _f32(phi, device))


@dataclass
class VrsVczt:
    x: np.ndarray
    y: np.ndarray
    wavelength: float
    z_vrs: float
    z_czt: float
    xout: np.ndarray
    yout: np.ndarray
    field: torch.Tensor      # (2, n, n) Ex, Ey
    eta: torch.Tensor
    theta: torch.Tensor

    def geometry(self):
        return ops.czt_geometry(self.x, self.y, self.xout, self.yout, self.z_czt, self.wavelength)

    def run(self, geom=None, czt_fn=None):
        m = MaskSpec(N.XL_MASK_LCD, self.eta, self.theta)
        vrs = ops.rs_propagate(self.field, self.x, self.y, self.z_vrs, self.wavelength, masks=[m],
                               ez_mode=N.XL_EZ_VRS, nin=2)
        g = geom or self.geometry()
        f = czt_fn or ops.czt
        return vrs, f(vrs, g, ez_mode=N.XL_EZ_VCZT, nin=2)


def radial_field(x, y, w0):
    X, Y = np.meshgrid(x, y)
    rho = np.sqrt(X ** 2 + Y ** 2)
    G = np.exp(-rho ** 2 / w0 ** 2)
    return np.stack([G * X / rho, G * Y / rho])


def vrs_vczt_2048(device, n: int = 2048, m_out: int = 512) -> VrsVczt:
    x, y = space(1200.0, n)
    E = radial_field(x, y, 1200.0)
    eta, th = np.random.default_rng(2).uniform(0, 2 * np.pi, (2, n, n)).astype(np.float32)
    xo = np.linspace(-100.0, 100.0, m_out)
    return VrsVczt(x, y, 0.635, 3e4, 3e4, xo, xo.copy(), _c64(E, device), _f32(eta, device), _f32(th, device))


@dataclass
class HighNA:
    x: np.ndarray
    y: np.ndarray
    wavelength: float
    radius: float
    f: float
    xout: np.ndarray
    yout: np.ndarray
    field: torch.Tensor       # (2, n, n)
    alpha: torch.Tensor       # sSLM phase on Ex
    phi: torch.Tensor         # sSLM phase on Ey
    eta: float
    theta: float

    def geometry(self):
        return ops.hna_geometry(self.x, self.y, self.xout, self.yout, self.wavelength, self.radius, self.f)

    def masks(self, alpha=None, phi=None):
        return [MaskSpec(N.XL_MASK_SSLM, self.alpha if alpha is None else alpha, self.phi if phi is None else phi),
                MaskSpec(N.XL_MASK_LCD_SCALAR, None, None, self.eta, self.theta)]

    def loss(self, alpha, phi, geom=None):
        """Differentiable focal-spot second moment sum I (u^2+v^2) / sum I."""
        E = AG.czt(self.field, geom or self.geometry(), masks=self.masks(alpha, phi), nin=2)
        return focal_second_moment(E, self.xout, self.yout)


_r2_cache: dict = {}


class _FocalMoment(torch.autograd.Function):
    """sum I r2 / sum I through the fused native reduction and its gradient
    kernel (the incoming gradient stays on the device)."""

    @staticmethod
    def forward(ctx, E, r2):
        loss, stats = ops.focal_moment(E.detach(), r2)
        ctx.save_for_backward(E, r2, stats)
        return loss

    @staticmethod
    def backward(ctx, gl):
        E, r2, stats = ctx.saved_tensors
        return ops.focal_moment_grad(E.detach(), r2, stats, gl), None


def focal_second_moment(E: torch.Tensor, xo, yo) -> torch.Tensor:
    """sum I (u^2 + v^2) / sum I over the focal window, I = sum_c |E_c|^2.  The
    u^2 + v^2 map is built once per window and device (a host-to-device copy
    per step would stall the stream)."""
    xo, yo = np.asarray(xo, dtype=np.float64), np.asarray(yo, dtype=np.float64)
    key = (xo.tobytes(), yo.tobytes(), str(E.device))
    r2 = _r2_cache.get(key)
    if r2 is None:
        yy, xx = np.meshgrid(yo, xo, indexing="ij")
        r2 = torch.from_numpy((xx ** 2 + yy ** 2).astype(np.float32)).to(E.device)
        _r2_cache[key] = r2
    return _FocalMoment.apply(E.contiguous(), r2)


def high_na_2048(device, n: int = 2048, m_out: int = 256) -> HighNA:
    radius = 1800.0
    f = radius / 0.9
    x, y = space(radius, n)
    X, Y = np.meshgrid(x, y)
    g = np.exp(-(X ** 2 + Y ** 2) / 1200.0 ** 2) / math.sqrt(2)
    E = np.stack([g, g]).astype(np.complex128)
    rng = np.random.default_rng(3)
    a, p = rng.uniform(-np.pi, np.pi, (2, n, n)).astype(np.float32)
    eta, th = rng.uniform(0, 2 * np.pi, 2)
    xo = np.linspace(-2.0, 2.0, m_out)
    return HighNA(x, y, 0.65, radius, f, xo, xo.copy(), _c64(E, device), _f32(a, device), _f32(p, device),
                  float(eta), float(th))


class CapturedStep:
    """One forward + backward of `loss_fn()` (a scalar) with respect to
    `params`, captured into a HIP graph: `replay()` re-runs every kernel of the
    step with no Python or launch overhead and returns (loss, grads) -- static
    tensors overwritten by each replay.  The parameters are updated in place
    between replays by the caller (their storage is what the graph reads).

    Capture follows torch's whole-step recipe: warm-up iterations on a side
    stream, the parameters' .grad cleared so that the captured backward
    allocates the gradient buffers from the graph's pool (and a replay
    overwrites them rather than accumulating)."""

    def __init__(self, loss_fn, params, warmup: int = 3):
        self.params = list(params)
        dev = self.params[0].device
        self.stream = torch.cuda.Stream(device=dev)
        self.stream.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(self.stream):
            for _ in range(warmup):
                for p in self.params:
                    p.grad = None
                loss_fn().backward()
        torch.cuda.current_stream(dev).wait_stream(self.stream)
        torch.cuda.synchronize(dev)
        for p in self.params:
            p.grad = None
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self.loss = loss_fn()
            self.loss.backward()
        self.grads = [p.grad for p in self.params]
        self.device = dev

    def replay(self):
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.stream):
            self.graph.replay()
        torch.cuda.current_stream(self.device).wait_stream(self.stream)
        return self.loss, self.grads


class SharpFocusTable:
    """The optical table of experiments/hybrid_sharp_optical_table.py at n^2
    (635 nm, 2500 um half-width, three building blocks, nine splitters, six
    high-NA detectors of m^2 at NA 0.9) with seeded random raw parameters;
    the loss is softmin of the small-area losses of the six detectors
    (hybrid_sharp_optical_table.py:50-58).  step() runs one forward +
    backward."""

    def __init__(self, dev=None, n: int = 1024, m: int = 400, seed: int = 9):
        from . import loss_functions as LF
        from . import optical_elements as OE
        from .toolbox import softmin, space
        from .vectorized_optics import PolarizedLightSource
        dev = torch.device(dev) if dev is not None else torch.device("cuda", torch.cuda.current_device())
        xs, ys = space(2500.0, n)
        self.src = PolarizedLightSource(xs, ys, 0.635, dev)
        self.src.gaussian_beam(w0=(1200.0, 1200.0), jones_vector=(1, 1))
        xo, yo = space(10.0, m)
        self.fixed = [1800.0, 1800.0 / 0.9, xo, yo]
        rng = np.random.default_rng(seed)
        raw = []
        for _ in range(3):
            raw += [rng.uniform(0, 1, (n, n)), rng.uniform(0, 1, (n, n))] + [rng.uniform(0, 1, 1) for _ in range(4)]
        raw += [rng.uniform(0, 1, 1) for _ in range(11)]
        self.params = [torch.tensor(v, device=dev, requires_grad=True) for v in raw]
        self._OE, self._LF, self._softmin = OE, LF, softmin

    def loss(self):
        s = self.src
        inten, _ = self._OE.hybrid_setup_sharp_focus(s, s, s, s, s, s, self.params, self.fixed)
        return self._softmin(self._LF.vectorized_loss_hybrid(inten))

    def step(self):
        for p in self.params:
            p.grad = None
        L = self.loss()
        L.backward()
        return L
