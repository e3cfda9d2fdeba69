"""Jones masks and the high-NA objective lens on the propagation path
(drop-in for the corresponding functions of xlumina/optical_elements.py).

SLM / sSLM / LCD / linear_polarizer return a light whose mask is pending: the
next propagation fuses it into its first pass (two masks in a row fuse; a
third materialises the first two).  Reading the field of such a light
materialises it with the standalone pre-op kernel.
"""
from __future__ import annotations

import math
import time
from typing import Optional

import numpy as np
import torch

from . import _native as N
from . import autograd as AG
from . import ops
from .ops import MaskSpec
from .vectorized_optics import VectorizedLight
from .wave_optics import ScalarLight, _grid
from . import wave_optics as _wo


class RawAngle:
    """A mask angle given by its raw value p, the angle being 2 pi p - pi (the
    map of the reference's optimisers, optical_elements.py:1556-1559).  With
    p a float64 tensor the map runs inside the mask kernels and gradients
    reach p directly; sSLM, sSLM_with_amplitude and LCD accept it for their
    angles."""
    __slots__ = ("p",)

    def __init__(self, p):
        self.p = p


def _param(p, light):
    if p is None:
        return None
    if isinstance(p, torch.Tensor):
        return p.to(light.device)
    return torch.as_tensor(np.asarray(p, dtype=np.float32), device=light.device)


def _angle(p, light):
    """(tensor, is_raw) of a mask angle: a RawAngle keeps its float64 raw values."""
    if isinstance(p, RawAngle):
        t = p.p if isinstance(p.p, torch.Tensor) else torch.as_tensor(np.asarray(p.p, dtype=np.float64))
        return t.to(light.device), True
    return _param(p, light), False


# -- (1) scalar -------------------------------------------------------------------------

def phase_scalar_SLM(phase):
    """optical_elements.py:76-85: exp(i phase)."""
    if isinstance(phase, torch.Tensor):
        return torch.polar(torch.ones_like(phase), phase)
    return np.exp(1j * np.asarray(phase))


def SLM(input_field: ScalarLight, phase_array, shape: int):
    """optical_elements.py:87-103: per-pixel phase mask -> (light, exp(i phase)).

    The light's mask stays pending and is fused into its next propagation."""
    phase = _param(phase_array, input_field)
    if tuple(phase.shape) != (shape, shape):
        raise ValueError(f"SLM phase array must be ({shape}, {shape})")
    light = input_field._with_mask(MaskSpec(N.XL_MASK_SLM, phase))
    return light, phase_scalar_SLM(phase)


# -- (2) Jones matrices (2 x 2, scalar parameters) -------------------------------------------

def jones_LP(alpha):
    """optical_elements.py:109-120."""
    a = torch.as_tensor(alpha, dtype=torch.float64)
    c, s = torch.cos(a), torch.sin(a)
    return torch.stack([torch.stack([c ** 2, c * s]), torch.stack([c * s, s ** 2])])


def jones_general_retarder(eta, theta, delta):
    """optical_elements.py:122-139."""
    eta = torch.as_tensor(eta, dtype=torch.float64)
    theta = torch.as_tensor(theta, dtype=torch.float64)
    delta = torch.as_tensor(delta, dtype=torch.float64)
    em, ep = torch.exp(-(eta / 2) * 1j), torch.exp((eta / 2) * 1j)
    c, s = torch.cos(theta), torch.sin(theta)
    return torch.stack([torch.stack([em * c ** 2 + ep * s ** 2, (em - ep) * torch.exp(-delta * 1j) * s * c]),
                        torch.stack([(em - ep) * torch.exp(delta * 1j) * s * c, em * s ** 2 + ep * c ** 2])])


def jones_sSLM(alpha, phi):
    """optical_elements.py:141-152."""
    a = torch.as_tensor(alpha, dtype=torch.float64)
    p = torch.as_tensor(phi, dtype=torch.float64)
    z = torch.zeros((), dtype=torch.complex128)
    return torch.stack([torch.stack([torch.exp(1j * a), z]), torch.stack([z, torch.exp(1j * p)])])


def jones_LCD(eta, theta):
    """optical_elements.py:169-180."""
    return jones_general_retarder(eta, theta, 0.0)


def jones_sSLM_with_amplitude(alpha, phi, a1, a2):
    """optical_elements.py:155-168: diag(a1 e^{i alpha}, a2 e^{i phi})."""
    j = jones_sSLM(alpha, phi)
    amp = torch.stack([torch.as_tensor(a1, dtype=torch.complex128), torch.as_tensor(a2, dtype=torch.complex128)])
    return j * amp[:, None]


# -- (3) polarisation devices ----------------------------------------------------------------

def sSLM(input_field: VectorizedLight, alpha_array=None, phi_array=None) -> VectorizedLight:
    """optical_elements.py:186-222: diag(e^{i alpha}, e^{i phi}) per pixel, Ez kept."""
    a, ra = _angle(alpha_array, input_field)
    b, rb = _angle(phi_array, input_field)
    return input_field._with_mask(MaskSpec(N.XL_MASK_SSLM, a, b, unit=int(ra) | 2 * int(rb)))


def sSLM_with_amplitude(input_field: VectorizedLight, alpha_array=None, phi_array=None, A1_array=None,
                        A2_array=None) -> VectorizedLight:
    """optical_elements.py:224-264: diag(A1 e^{i alpha}, A2 e^{i phi}) per pixel, Ez kept
    (two fused stages: the phases, then the real amplitudes)."""
    light = sSLM(input_field, alpha_array, phi_array)
    return light._with_mask(MaskSpec(N.XL_MASK_AMP, _param(A1_array, input_field), _param(A2_array, input_field)))


def LCD(input_field: VectorizedLight, eta, theta) -> VectorizedLight:
    """optical_elements.py:266-305: retarder with eta/theta constant (floats) or per pixel; Ez kept."""
    plain = lambda v: not isinstance(v, (torch.Tensor, RawAngle)) and np.ndim(v) == 0
    if plain(eta) and plain(theta):
        spec = MaskSpec(N.XL_MASK_LCD_SCALAR, None, None, float(eta), float(theta))
    else:
        # one-value float64 tensors stay one value (read by the kernels from device memory)
        e, re_ = _angle(eta, input_field)
        t, rt = _angle(theta, input_field)
        spec = MaskSpec(N.XL_MASK_LCD, e, t, unit=int(re_) | 2 * int(rt))
    return input_field._with_mask(spec)


def linear_polarizer(input_field: VectorizedLight, alpha) -> VectorizedLight:
    """optical_elements.py:307-332: [[c^2, cs], [cs, s^2]] per pixel; Ez becomes 0."""
    return input_field._with_mask(MaskSpec(N.XL_MASK_LP, _param(alpha, input_field)))


# -- high-NA objective lens ---------------------------------------------------------------------

def high_NA_objective_lens(input_field: VectorizedLight, radius: float, f: float):
    """optical_elements.py:515-553: (3, N, N) field after the lens and sin(theta_max)."""
    sin_theta_max = radius / math.sqrt(radius ** 2 + f ** 2)
    src, masks = input_field._pending()
    out = AG.pre(src, input_field.x, input_field.y, masks=masks, hna=True, nin=2, ncomp=3, hna_radius=radius,
                 hna_f=f)
    return out, sin_theta_max


def build_high_NA_VCZT_grid(f, r, wavelength, xin, xout, yout):
    """optical_elements.py:640-672."""
    xin, xout, yout = _grid(xin), _grid(xout), _grid(yout)
    nx, ny = len(xout), len(yout)
    Dm = f * wavelength * (len(xin) - 1) / (2 * r)
    return nx, ny, Dm, yout[0] + Dm / 2, yout[-1] + Dm / 2, xout[0] + Dm / 2, xout[-1] + Dm / 2


def VCZT_objective_lens(input_field: VectorizedLight, r: float, f: float, xout, yout) -> VectorizedLight:
    """optical_elements.py:600-638: lens map + two Bluestein passes + -i sin^2(theta_max)/(f lambda),
    all on the device with pending masks fused into the first pass."""
    tic = time.perf_counter()
    xout, yout = _grid(xout), _grid(yout)
    src, masks = input_field._pending()
    geom = ops.hna_geometry(input_field.x, input_field.y, xout, yout, input_field.wavelength, r, f)
    out = AG.czt(src, geom, masks=masks, nin=2)
    light = VectorizedLight(xout, yout, input_field.wavelength, input_field.device)
    light._E = out
    if _wo.VERBOSE:
        print("Time taken to perform one VCZT propagation through objective lens (in seconds):  "
              f"{(time.perf_counter() - tic):.4f}")
    return light


# -- (4) lenses and building blocks ------------------------------------------------------------

def _grid(A, like=None):
    if isinstance(A, torch.Tensor):
        return A.to(torch.float64)
    return torch.as_tensor(np.asarray(A, dtype=np.float64), device=None if like is None else like.device)


def _mesh(light):
    """(X (1, Nx), Y (Ny, 1)) float64 on the light's device, broadcasting to the grid."""
    dev = light.device
    X = torch.as_tensor(np.asarray(light.x, dtype=np.float64), device=dev)[None, :]
    Y = torch.as_tensor(np.asarray(light.y, dtype=np.float64), device=dev)[:, None]
    return X, Y


def _phase_mask(phase_turns, amplitude=None):
    """exp(2 pi i t) as complex64 from float64 turns t, reduced to [-1/2, 1/2]
    first: lens phases reach 10^4 rad, beyond float32's reach."""
    t = phase_turns - torch.round(phase_turns)
    amp = torch.ones_like(t) if amplitude is None else amplitude.to(t.dtype).expand_as(t)
    return torch.polar(amp, 2 * math.pi * t).to(torch.complex64)


def _masked(input_field, mask):
    """The light times a complex mask; a vectorized light's Ez is dropped, as
    every mask element of optical_elements.py:678-779 builds a new light
    from Ex and Ey only."""
    dev = input_field.device
    if isinstance(input_field, VectorizedLight):
        out = VectorizedLight(input_field.x, input_field.y, input_field.wavelength, dev)
        E = input_field.E
        out._E = torch.stack([E[0] * mask, E[1] * mask, torch.zeros_like(E[2])])
    elif isinstance(input_field, ScalarLight):
        out = ScalarLight(input_field.x, input_field.y, input_field.wavelength, dev)
        out._field = input_field.field * mask
    else:
        raise ValueError("Invalid input. Please use ScalarLight or VectorizedLight object.")
    return out


def lens(input_field, radius, focal):
    """optical_elements.py:678-703: circular pupil of radius (rx, ry) times
    exp(-i k (X^2 / (2 fx) + Y^2 / (2 fy))) -> (light, lens mask).

    The phase reaches ~k R^2 / (2 f) (10^4 rad for a 1 mm pupil at f = 25 mm),
    so it is formed in float64 on the device and reduced before the complex64
    cast."""
    fx, fy = focal
    X, Y = _mesh(input_field)
    pupil = circular_mask(X, Y, radius).to(torch.float64)
    turns = -(input_field.k / (2 * math.pi)) * (X ** 2 / (2 * fx) + Y ** 2 / (2 * fy))
    mask = _phase_mask(turns, pupil)
    return _masked(input_field, mask), mask


def cylindrical_lens(input_field, focal_length, refractive_index=1.5, angle=0):
    """optical_elements.py:705-744: plano-convex cylindrical lens with its phase
    along the rotated x axis: R = f (n - 1), thickness R - sqrt(R^2 - Xrot^2),
    phase k (n - 1) (Xrot^2 / (2 f) + thickness) -> (light, exp(-i phase)).
    Outside |Xrot| < R the square root is NaN, as in the reference."""
    X, Y = _mesh(input_field)
    a = torch.as_tensor(angle, dtype=torch.float64, device=X.device)
    Xrot = X * torch.cos(a) + Y * torch.sin(a)
    R = focal_length * (refractive_index - 1)
    thickness = R - torch.sqrt(R ** 2 - Xrot ** 2)
    phase = input_field.k * (refractive_index - 1) * (Xrot ** 2 / (2 * focal_length) + thickness)
    mask = _phase_mask(-phase / (2 * math.pi))
    return _masked(input_field, mask), mask


def axicon_lens(input_field, alpha, n=1.5):
    """optical_elements.py:746-779: phase k r (n - 1) sin(alpha) tan(alpha),
    r = sqrt(X^2 + Y^2) -> (light, exp(-i phase)) (a Bessel-beam generator)."""
    X, Y = _mesh(input_field)
    r = torch.sqrt(X ** 2 + Y ** 2)
    phase = input_field.k * r * (n - 1) * math.sin(alpha) * math.tan(alpha)
    mask = _phase_mask(-phase / (2 * math.pi))
    return _masked(input_field, mask), mask


# -- amplitude masks (optical_elements.py:781-892) ------------------------------------------------

def circular_mask(X, Y, r):
    """optical_elements.py:781-795: 1 inside X^2 / rx^2 + Y^2 / ry^2 < 1, else 0."""
    rx, ry = r
    X = _grid(X)
    Y = _grid(Y, X)
    return torch.where(X ** 2 / rx ** 2 + Y ** 2 / ry ** 2 < 1, 1, 0)


def _triangle(X, Y, r, angle, m, height):
    from .toolbox import rotate_mask
    x0, y0 = r
    X = _grid(X)
    Xrot, Yrot = rotate_mask(X, _grid(Y, X), angle, r)
    edge = -m * torch.abs(Xrot - x0) + y0
    return torch.where((Yrot < edge) & (Yrot > y0 - height), 1, 0)


def triangular_mask(X, Y, r, angle, m, height):
    """optical_elements.py:797-817: triangle with top corner r = (x0, y0), edge
    slope m and the given height, rotated by `angle` degrees about r; as in the
    reference, the corner is subtracted again after the rotation about it."""
    return _triangle(X, Y, r, angle * (math.pi / 180), m, height)


def rectangular_mask(X, Y, center, width, height, angle):
    """optical_elements.py:819-836: rectangle of width x height about `center`,
    rotated by `angle` degrees."""
    from .toolbox import rotate_mask
    X = _grid(X)
    Xrot, Yrot = rotate_mask(X, _grid(Y, X), angle * (math.pi / 180), center)
    return torch.where((Xrot < width / 2) & (Xrot > -width / 2) & (Yrot < height / 2) & (Yrot > -height / 2), 1, 0)


def annular_aperture(di, do, X, Y):
    """optical_elements.py:838-854: ring between the inner and outer diameters
    di and do (radii di / 2, do / 2)."""
    X = _grid(X)
    Y = _grid(Y, X)
    r2 = X ** 2 + Y ** 2
    stop = torch.where(r2 / (di / 2) ** 2 < 1, 0, 1)
    ring = torch.where(r2 / (do / 2) ** 2 < 1, 1, 0)
    return stop * ring


def forked_grating(X, Y, center, angle, period, l, kind=''):
    """optical_elements.py:856-892: binary forked grating, 0 where
    cos(l atan2(Xrot, Yrot) - 2 pi Xrot / period) < 0 and 1 elsewhere, rotated
    by `angle` degrees about `center` ('Amplitude': the 0/1 mask; 'Phase':
    exp(i pi mask); any other kind returns None, as in the reference).  The
    reference thresholds angle(exp(i c)), which equals c for |c| <= 1."""
    from .toolbox import rotate_mask
    X = _grid(X)
    Xrot, Yrot = rotate_mask(X, _grid(Y, X), angle * (math.pi / 180), center)
    c = torch.cos(l * torch.atan2(Xrot, Yrot) - 2 * math.pi / period * Xrot)
    binary = torch.where(c < 0, 0, 1)
    if kind == 'Amplitude':
        return binary
    if kind == 'Phase':
        return torch.polar(torch.ones_like(c), math.pi * binary.to(torch.float64))
    return None


def building_block(input_light: VectorizedLight, alpha, phi, z, eta, theta) -> VectorizedLight:
    """optical_elements.py:919-939: sSLM(alpha, phi) -> VRS(z) -> LCD(eta, theta).
    The sSLM is fused into the propagation and the LCD stays pending for the
    next one; z may be a tensor with requires_grad."""
    l_propagated, _ = sSLM(input_light, alpha, phi).VRS_propagation(z)
    return LCD(l_propagated, eta, theta)


def bb_amplitude_and_phase_mod(input_light: VectorizedLight, alpha, phi, amp1, amp2, z, eta, theta) -> VectorizedLight:
    """optical_elements.py:894-917: sSLM_with_amplitude -> VRS(z) -> LCD(eta, theta)."""
    l_propagated, _ = sSLM_with_amplitude(input_light, alpha, phi, amp1, amp2).VRS_propagation(z)
    return LCD(l_propagated, eta, theta)


# -- beam splitters (optical_elements.py:334-513) --------------------------------------------

class _BeamSplitter(torch.autograd.Function):
    """c = ca a + cb b, d = da a + db b on (Ex, Ey) with coef = (ca, cb, da, db)
    (a complex128 tensor of 4 that may carry a gradient); Ez of c and d is 0.
    Backward: the same kernel with the conjugate-transposed coefficients, and
    the coefficient cotangents conj(sum conj(g_out) in) from xl_bs_dots.
    Coefficients on the GPU are read by the kernels from device memory, so a
    table of splitters runs without host synchronisation."""

    @staticmethod
    def forward(ctx, a, b, coef):
        k = coef.detach()
        c, d = ops.beam_splitter(a.detach(), None if b is None else b.detach(), k)
        ctx.k = k
        ctx.has_b = b is not None
        ctx.save_for_backward(a, b if b is not None else a)
        return c, d

    @staticmethod
    def backward(ctx, gc, gd):
        a, b = ctx.saved_tensors
        b = b if ctx.has_b else None
        k = ctx.k
        gc = gc.contiguous() if gc is not None else None
        gd = gd.contiguous() if gd is not None else None
        z = torch.zeros_like(a) if gc is None or gd is None else None
        gcc = gc if gc is not None else z
        gdd = gd if gd is not None else z
        ga = gb = gk = None
        need_b = ctx.has_b and ctx.needs_input_grad[1]
        if ctx.needs_input_grad[0] or need_b:
            # (conj ca, conj da, conj cb, conj db); basic indexing keeps the step free of
            # host-to-device index uploads, so it can be captured in a HIP graph
            kh = torch.stack([k[0], k[2], k[1], k[3]]).conj()
            ga, gb = ops.beam_splitter(gcc, gdd, kh, want_c=True, want_d=need_b, ncomp_out=a.shape[0])
        if ctx.needs_input_grad[2]:
            s = ops.bs_dots(gc, gd, a.detach(), None if b is None else b.detach())
            gk = s.to(device=k.device, dtype=k.dtype).conj()
        return (ga if ctx.needs_input_grad[0] else None), (gb if need_b else None), gk


class _SplitterTheta(torch.autograd.Function):
    """The lossy symmetric / single-input splitters with their coefficients
    formed by the kernels from the float64 device angle theta (theta_mode 1:
    symmetric, 2 / 3: one input at port 0 / 1).  Backward: the same kernel
    with the conjugate-transposed coefficients for the field cotangents and
    xl_bs_dots writing dL/dtheta, so a splitter adds no scalar torch kernels
    to a step."""

    @staticmethod
    def forward(ctx, a, b, theta, mode):
        th = theta.detach()
        c, d = ops.beam_splitter(a.detach(), None if b is None else b.detach(), None, theta=th, theta_mode=mode)
        ctx.mode, ctx.has_b = mode, b is not None
        ctx.save_for_backward(a, b if b is not None else a, th)
        return c, d

    @staticmethod
    def backward(ctx, gc, gd):
        a, b, th = ctx.saved_tensors
        b = b if ctx.has_b else None
        gc = gc.contiguous() if gc is not None else None
        gd = gd.contiguous() if gd is not None else None
        ga = gb = gt = None
        need_b = ctx.has_b and ctx.needs_input_grad[1]
        if ctx.needs_input_grad[2] and (ctx.needs_input_grad[0] or need_b):
            # dL/dtheta and the input cotangents in one pass over gc, gd, a, b
            gt, ga, gb = ops.bs_dots(gc, gd, a.detach(), None if b is None else b.detach(), theta=th,
                                     theta_mode=ctx.mode, want_a=ctx.needs_input_grad[0], want_b=need_b,
                                     ncomp_out=a.shape[0])
            return ga, gb, gt.reshape(th.shape), None
        z = torch.zeros_like(a) if gc is None or gd is None else None
        if ctx.needs_input_grad[0] or need_b:
            ga, gb = ops.beam_splitter(gc if gc is not None else z, gd if gd is not None else z, None,
                                       want_c=True, want_d=need_b, ncomp_out=a.shape[0], theta=th,
                                       theta_mode=ctx.mode, adjoint=True)
        if ctx.needs_input_grad[2]:
            gt = ops.bs_dots(gc, gd, a.detach(), None if b is None else b.detach(), theta=th,
                             theta_mode=ctx.mode).reshape(th.shape)
        return (ga if ctx.needs_input_grad[0] else None), (gb if need_b else None), gt, None


def _splitter(a: VectorizedLight, b: Optional[VectorizedLight], theta, mode):
    if isinstance(theta, RawAngle):   # the kernels map the raw parameter (XL_BS_THETA_RAW)
        theta, mode = theta.p, mode | 4
    th = _scalar(theta, a.device)
    if th.device != a.device:
        th = th.to(a.device)
    out_c = VectorizedLight(a.x, a.y, a.wavelength, a.device)
    out_d = VectorizedLight(a.x, a.y, a.wavelength, a.device)
    out_c._E, out_d._E = _SplitterTheta.apply(a.E, None if b is None else b.E, th.reshape(1), mode)
    return out_c, out_d


def _scalar(v, device=None):
    """A one-element parameter (float, 0-d or (1,) tensor) as a 0-d float64
    tensor, on the tensor's own device (a float goes to `device`)."""
    if isinstance(v, torch.Tensor):
        if v.numel() != 1:
            raise ValueError("beam-splitter parameters are scalars")
        return v.to(dtype=torch.float64).reshape(())
    return torch.tensor(float(np.asarray(v).reshape(-1)[0]), dtype=torch.float64, device=device)


def _bs_apply(a: VectorizedLight, b: Optional[VectorizedLight], coef):
    out_c = VectorizedLight(a.x, a.y, a.wavelength, a.device)
    out_d = VectorizedLight(a.x, a.y, a.wavelength, a.device)
    out_c._E, out_d._E = _BeamSplitter.apply(a.E, None if b is None else b.E, coef)
    return out_c, out_d


def BS_symmetric(a: VectorizedLight, b: VectorizedLight, theta):
    """optical_elements.py:334-392: lossy symmetric splitter [[R, iT], [iT, R]]
    (R = |sin theta|, T = |cos theta| less 1% of T) -> lights at ports c, d."""
    return _splitter(a, b, theta, 1)


def BS_symmetric_SI(a: VectorizedLight, theta, port=0):
    """optical_elements.py:394-460: single-input splitter; port 0: c = R a,
    d = iT a; port 1: c = iT a, d = R a."""
    return _splitter(a, None, theta, 2 if int(port) == 0 else 3)


def BS(a1: VectorizedLight, a2: VectorizedLight, R, T, phase):
    """optical_elements.py:462-513: [[sqrt T, e^{i phase} sqrt R],
    [-e^{-i phase} sqrt R, sqrt T]] -> lights at ports b1, b2."""
    dev = next((v.device for v in (R, T, phase) if isinstance(v, torch.Tensor)), None)
    sR, sT, ph = torch.sqrt(_scalar(R, dev)), torch.sqrt(_scalar(T, dev)), _scalar(phase, dev)
    zero = torch.zeros((), dtype=torch.float64, device=sT.device)
    st = torch.complex(sT, zero)
    e = torch.polar(torch.ones((), dtype=torch.float64, device=ph.device), ph)
    return _bs_apply(a1, a2, torch.stack([st, sR * e, -sR * e.conj(), st]))


# -- fluorescence and a pre-built table (optical_elements.py:941-962, 1503-1649) ----------------

def fluorescence(i_ex, i_dep, beta=1):
    """optical_elements.py:941-955: effective STED intensity
    i_ex (1 - beta (1 - exp(-i_dep / (i_ex + 1e-8))))."""
    return i_ex * (1 - beta * (1 - torch.exp(-(i_dep / (i_ex + 1e-8)))))


def vectorized_fluorophores(i_ex, i_dep):
    """optical_elements.py:957-962: fluorescence over a stack of detectors."""
    return fluorescence(i_ex, i_dep)


class _ComponentIntensity(torch.autograd.Function):
    """torch.stack([|E_j[c]|^2 for each field E_j]) of (3, My, Mx) complex64
    fields, written straight into the stack; backward: each field's
    cotangent 2 g_j E_j[c] in plane c and zeros in the other two (one fill
    and one product per field instead of the select / abs / pow / stack
    backward chain)."""

    @staticmethod
    def forward(ctx, comp, *fields):
        out = torch.empty((len(fields),) + tuple(fields[0].shape[-2:]), dtype=torch.float32,
                          device=fields[0].device)
        for j, e in enumerate(fields):
            torch.abs(e[comp], out=out[j])
            out[j].square_()
        ctx.comp = comp
        ctx.save_for_backward(*fields)
        return out

    @staticmethod
    def backward(ctx, g):
        c = ctx.comp
        g2 = g * 2.0
        grads = []
        for j, e in enumerate(ctx.saved_tensors):
            if not ctx.needs_input_grad[1 + j]:
                grads.append(None)
                continue
            ge = torch.empty(e.shape, dtype=e.dtype, device=e.device)
            if c == e.shape[0] - 1:
                ge[:c].zero_()
            else:
                for i in range(e.shape[0]):
                    if i != c:
                        ge[i].zero_()
            torch.mul(e[c], g2[j], out=ge[c])
            grads.append(ge)
        return (None, *grads)


def _ez_intensities(dets):
    """(6, My, Mx) stack of |Ez|^2 of the detector lights."""
    return _ComponentIntensity.apply(2, *[d.E for d in dets])


def _unit_angle(p):
    return p * (2 * math.pi) - math.pi          # raw (0, 1) -> (-pi, pi)


def _distance(p, offset):
    return (torch.abs(p) * 100 + offset) * 1e4  # cm -> um


def _distances(ps, offset):
    """[_distance(p, offset) for p in ps], mapped together: one chain of
    element-wise launches for all of a table's distances (each result a
    one-element view of it) where the raw values are one-element tensors on
    one device."""
    if not ps or not all(isinstance(p, torch.Tensor) and p.numel() == 1 for p in ps) \
            or len({p.device for p in ps}) != 1:
        return [_distance(p, offset) for p in ps]
    z = (torch.cat([p.reshape(1) for p in ps]).abs() * 100 + offset) * 1e4
    return [v.reshape(p.shape) for v, p in zip(z.split(1), ps)]


def _bs_pair(a, b, theta):
    """BS_symmetric(a, b); where one input port carries no light (a source the
    colour of a two-colour table does not have) the single-input splitter fed
    at the other port, as optical_elements.py:1741-1828 writes it."""
    if a is None:
        return BS_symmetric_SI(b, theta, port=1)
    if b is None:
        return BS_symmetric_SI(a, theta, port=0)
    return BS_symmetric(a, b, theta)


def _as_tensors(parameters):
    return [p if isinstance(p, torch.Tensor) else torch.as_tensor(np.asarray(p, dtype=np.float64))
            for p in parameters]


def _normalised_intensity(dets):
    """(6, My, Mx) of (|Ex|^2 + |Ey|^2) / sum per detector."""
    out = []
    for d in dets:
        i = d.Ex.abs() ** 2 + d.Ey.abs() ** 2
        out.append(i / i.sum())
    return torch.stack(out)


def _table_3x3(ls, block, zs, bs, z4, z5, lens):
    """Topology shared by hybrid_setup_sharp_focus, robust_discovery and
    hybrid_setup_fluorophores (optical_elements.py:1594-1634, 1062-1102,
    1741-1828): sources ls[0..5] (None where a colour has no source), block(light,
    i) the i-th building block (which includes its first propagation), zs[i] =
    (z_i1, z_i2), nine splitter angles, the mid-space distances z4, z5 and the
    objective lens map.  Returns the six detector lights."""
    ls1, ls2, ls3, ls4, ls5, ls6 = ls
    (z11, z12), (z21, z22), (z31, z32) = zs
    s1, s2, s3 = z11 + z12, z21 + z22, z31 + z32   # each used twice

    def vrs(light, z):
        return light.VRS_propagation(z)[0]

    c1, d1 = _bs_pair(ls1, ls4, bs[0])
    b2 = vrs(block(c1, 0), z12)
    c2, d2 = _bs_pair(ls2, b2, bs[1])
    b3 = vrs(c2, s2)
    c3, d3 = _bs_pair(ls3, b3, bs[2])
    det_1 = lens(vrs(c3, s3))
    a5, a6 = vrs(d2, z4), vrs(d3, z4)
    c4, d4 = _bs_pair(d1, ls5, bs[3])
    b5 = vrs(c4, s1)
    c5, d5 = BS_symmetric(a5, b5, bs[4])
    b6 = vrs(block(c5, 1), z22)
    c6, d6 = BS_symmetric(a6, b6, bs[5])
    det_2 = lens(vrs(c6, s3))
    a8, a9 = vrs(d5, z5), vrs(d6, z5)
    c7, d7 = _bs_pair(d4, ls6, bs[6])
    b8 = vrs(c7, s1)
    c8, d8 = BS_symmetric(a8, b8, bs[7])
    b9 = vrs(c8, s2)
    c9, d9 = BS_symmetric(a9, b9, bs[8])
    det_3 = lens(vrs(block(c9, 2), z32))
    return [det_1, det_2, det_3, lens(d7), lens(d8), lens(d9)]


def hybrid_setup_sharp_focus(ls1, ls2, ls3, ls4, ls5, ls6, parameters, fixed_params, distance_offset=8.9):
    """optical_elements.py:1503-1649: the 3 x 3 table of three building blocks
    and nine symmetric beam splitters, six high-NA detectors.

    parameters (29 tensors): per block i = 1..3 [phase_i_1, phase_i_2 (N x N),
    eta_i, theta_i, z_i_1, z_i_2], then bs1..bs9, z4, z5 -- raw values mapped
    as the reference does (angles p 2pi - pi, distances (|p| 100 + offset) cm).
    fixed_params = [r, f, xout, yout] of the objective lens.
    Returns (intensities (6, My, Mx) = |Ez|^2 per detector, detector lights).
    Every stage runs on the device; gradients reach all 29 parameters."""
    r, f, xout, yout = fixed_params[:4]
    blocks, bs, z4, z5 = _table_parameters(parameters, distance_offset)

    def block(light, i):
        ph1, ph2, eta, th, za, _ = blocks[i]
        return building_block(light, ph1, ph2, za, eta, th)

    dets = _table_3x3((ls1, ls2, ls3, ls4, ls5, ls6), block, [(b[4], b[5]) for b in blocks], bs, z4, z5,
                      lambda light: VCZT_objective_lens(light, r, f, xout, yout))
    return _ez_intensities(dets), dets


def robust_discovery(ls1, ls2, ls3, ls4, ls5, ls6, parameters, fixed_params, noise_distances, noise_slms, noise_wps,
                     noise_amps, distance_offset=15):
    """optical_elements.py:964-1117: the sharp-focus table with amplitude-and-
    phase building blocks (bb_amplitude_and_phase_mod) and additive noise on
    every physical parameter: noise_slms (6 phase masks), noise_amps (6
    amplitude masks), noise_wps (6 retarder angles), noise_distances (8:
    z1_1, z1_2, z2_1, z2_2, z3_1, z3_2, z4, z5).  parameters (35): per block
    [phase_1, phase_2, A_1, A_2, eta, theta, z_1, z_2], then bs1..bs9, z4, z5.
    Returns the intensities (6, My, Mx) = |Ez|^2 per detector."""
    r, f, xout, yout = fixed_params[:4]
    P = _as_tensors(parameters)
    dev = P[0].device

    def on(v):   # noise as a tensor next to the parameters
        return v.to(dev) if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v, dtype=np.float64), device=dev)

    noise_slms, noise_amps = [on(v) for v in noise_slms], [on(v) for v in noise_amps]
    noise_wps, noise_distances = [on(v) for v in noise_wps], [on(v) for v in noise_distances]
    blocks, zs = [], []
    for i in range(3):
        q = P[8 * i: 8 * i + 8]
        blocks.append((_unit_angle(q[0]) + noise_slms[2 * i], _unit_angle(q[1]) + noise_slms[2 * i + 1],
                       q[2] + noise_amps[2 * i], q[3] + noise_amps[2 * i + 1],
                       _unit_angle(q[4]) + noise_wps[2 * i], _unit_angle(q[5]) + noise_wps[2 * i + 1]))
        zs.append((_distance(q[6], distance_offset) + noise_distances[2 * i],
                   _distance(q[7], distance_offset) + noise_distances[2 * i + 1]))
    bs = [RawAngle(p) for p in P[24:33]]
    z4 = _distance(P[33], distance_offset) + noise_distances[6]
    z5 = _distance(P[34], distance_offset) + noise_distances[7]

    def block(light, i):
        ph1, ph2, a1, a2, eta, th = blocks[i]
        return bb_amplitude_and_phase_mod(light, ph1, ph2, a1, a2, zs[i][0], eta, th)

    dets = _table_3x3((ls1, ls2, ls3, ls4, ls5, ls6), block, zs, bs, z4, z5,
                      lambda light: VCZT_objective_lens(light, r, f, xout, yout))
    return _ez_intensities(dets)


def _table_parameters(parameters, distance_offset):
    """The 29 raw parameters of the 3 x 3 tables (optical_elements.py:1561-1596):
    three blocks [phase_1, phase_2, eta, theta, z_1, z_2], nine splitter
    angles, z4, z5, mapped to radians and microns."""
    P = [p if isinstance(p, torch.Tensor) else torch.as_tensor(np.asarray(p, dtype=np.float64)) for p in parameters]
    z = _distances([P[6 * i + k] for i in range(3) for k in (4, 5)] + [P[27], P[28]], distance_offset)
    blocks = []
    for i in range(3):
        ph1, ph2, eta, th = P[6 * i: 6 * i + 4]
        # the angles stay raw: their map to (-pi, pi) runs in the mask and splitter kernels
        blocks.append((RawAngle(ph1), RawAngle(ph2), RawAngle(eta), RawAngle(th), z[2 * i], z[2 * i + 1]))
    bs = [RawAngle(p) for p in P[18:27]]
    return blocks, bs, z[6], z[7]


def hybrid_setup_fluorophores(ls1, ls2, ls3, ls4, ls5, ls6, parameters, fixed_params, distance_offset=8.9):
    """optical_elements.py:1651-1853: the 3 x 3 table run for a depletion
    colour (sources ls1, ls3, ls5) and an excitation colour (ls2, ls4, ls6)
    through the same elements, single-input splitters where only one colour
    enters; six high-NA detectors; returns the STED effective intensity
    fluorescence(I_ex / sum I_ex, I_dep / sum I_dep) per detector (6, My, Mx),
    with I = |Ex|^2 + |Ey|^2.  Parameters as hybrid_setup_sharp_focus."""
    r, f, xout, yout = fixed_params[:4]
    blocks, bs, z4, z5 = _table_parameters(parameters, distance_offset)
    zs = [(b[4], b[5]) for b in blocks]

    def block(light, i):
        ph1, ph2, eta, th, za, _ = blocks[i]
        return building_block(light, ph1, ph2, za, eta, th)

    def olens(light):
        return VCZT_objective_lens(light, r, f, xout, yout)

    dep = _table_3x3((ls1, None, ls3, None, ls5, None), block, zs, bs, z4, z5, olens)
    exc = _table_3x3((None, ls2, None, ls4, None, ls6), block, zs, bs, z4, z5, olens)
    return vectorized_fluorophores(_normalised_intensity(exc), _normalised_intensity(dep))


def _fixed_parameters(parameters, fixed_params, distance_offset):
    """The 25 raw parameters of the fixed-mask tables (optical_elements.py:1174-1215,
    1405-1446): eight distances z1_1..z4_2, nine splitter angles, four (eta,
    theta) pairs; the fixed masks fixed_params[4:10] are phase1_1, phase1_2,
    phase2 (both masks), phase3 (both masks), phase4_1, phase4_2.  Returns the
    per-block (phase_1, phase_2, eta, theta, z_1, z_2) and the splitter angles."""
    ph11, ph12, ph2, ph3, ph41, ph42 = fixed_params[4:10]
    P = _as_tensors(parameters)
    z = _distances(P[0:8], distance_offset)
    bs = [RawAngle(p) for p in P[8:17]]
    wp = [RawAngle(p) for p in P[17:25]]
    masks = ((ph11, ph12), (ph2, ph2), (ph3, ph3), (ph41, ph42))
    blocks = [(masks[i][0], masks[i][1], wp[2 * i], wp[2 * i + 1], z[2 * i], z[2 * i + 1]) for i in range(4)]
    return blocks, bs


def _table_fixed(ls, blocks, bs, lens):
    """Topology of the fixed-mask tables (optical_elements.py:1448-1486,
    1217-1321): four building blocks placed on the first row, the second row,
    above the third splitter column and below the first; sources ls[0..5]
    (None where a colour has no source).  Returns the six detector lights."""
    ls1, ls2, ls3, ls4, ls5, ls6 = ls
    (z11, z12), (z21, z22), (z31, z32), (z41, z42) = [(b[4], b[5]) for b in blocks]

    def vrs(light, z):
        return light.VRS_propagation(z)[0]

    def block(light, i):
        ph1, ph2, eta, th, za, zb = blocks[i]
        return vrs(building_block(light, ph1, ph2, za, eta, th), zb)

    c1, d1 = _bs_pair(ls1, ls4, bs[0])
    b2 = block(c1, 0)
    c2, d2 = _bs_pair(ls2, b2, bs[1])
    b3 = vrs(c2, z21 + z22)
    c3, d3 = _bs_pair(ls3, b3, bs[2])
    det_1 = lens(c3)
    a5 = vrs(d2, z31 + z32)
    a6 = block(d3, 2)
    c4, d4 = _bs_pair(d1, ls5, bs[3])
    b5 = vrs(c4, z11 + z12)
    c5, d5 = BS_symmetric(a5, b5, bs[4])
    b6 = block(c5, 1)
    c6, d6 = BS_symmetric(a6, b6, bs[5])
    det_2 = lens(c6)
    a7 = block(d4, 3)
    a8, a9 = vrs(d5, z41 + z42), vrs(d6, z41 + z42)
    c7, d7 = _bs_pair(a7, ls6, bs[6])
    b8 = vrs(c7, z11 + z12)
    c8, d8 = BS_symmetric(a8, b8, bs[7])
    b9 = vrs(c8, z21 + z22)
    c9, d9 = BS_symmetric(a9, b9, bs[8])
    det_3 = lens(c9)
    return [det_1, det_2, det_3, lens(d7), lens(d8), lens(d9)]


def hybrid_setup_fixed_slms(ls1, ls2, ls3, ls4, ls5, ls6, parameters, fixed_params, distance_offset=8.9):
    """optical_elements.py:1347-1501: the 3 x 3 table with four building
    blocks whose phase masks are fixed (fixed_params[4:10]: phase1_1,
    phase1_2, phase2 (both masks), phase3 (both masks), phase4_1, phase4_2,
    in radians) and 25 trained raw parameters: eight distances z1_1..z4_2,
    nine splitter angles, four (eta, theta) pairs.  Returns (intensities
    (6, My, Mx) = |Ez|^2 per detector, detector lights)."""
    r, f, xout, yout = fixed_params[:4]
    blocks, bs = _fixed_parameters(parameters, fixed_params, distance_offset)
    dets = _table_fixed((ls1, ls2, ls3, ls4, ls5, ls6), blocks, bs,
                        lambda light: VCZT_objective_lens(light, r, f, xout, yout))
    return _ez_intensities(dets), dets


def hybrid_setup_fixed_slms_fluorophores(ls1, ls2, ls3, ls4, ls5, ls6, parameters, fixed_params,
                                         distance_offset=8.9):
    """optical_elements.py:1119-1345: the fixed-mask table run for a depletion
    colour (ls1, ls3, ls5) and an excitation colour (ls2, ls4, ls6), single-input
    splitters where only one colour enters; returns the STED effective
    intensity fluorescence(I_ex / sum I_ex, I_dep / sum I_dep) per detector
    (6, My, Mx), I = |Ex|^2 + |Ey|^2.  Parameters as hybrid_setup_fixed_slms."""
    r, f, xout, yout = fixed_params[:4]
    blocks, bs = _fixed_parameters(parameters, fixed_params, distance_offset)

    def olens(light):
        return VCZT_objective_lens(light, r, f, xout, yout)

    dep = _table_fixed((ls1, None, ls3, None, ls5, None), blocks, bs, olens)
    exc = _table_fixed((None, ls2, None, ls4, None, ls6), blocks, bs, olens)
    return vectorized_fluorophores(_normalised_intensity(exc), _normalised_intensity(dep))


def six_times_six_ansatz(ls1, ls2, ls3, ls4, ls5, ls6, ls7, ls8, ls9, ls10, ls11, ls12, parameters, fixed_params,
                         distance_offset=7.6):
    """optical_elements.py:1855-2125: a 6 x 6 grid of 36 symmetric splitters fed
    by ls1..ls6 from the top and ls7..ls12 from the left, two building blocks
    with fixed masks (fixed_params[4:8] = phase1_1, phase1_2, phase2_1,
    phase2_2) and 12 high-NA detectors.

    parameters (52 raw): z1..z12, bs1..bs36 (row-major), eta1, theta1, eta2,
    theta2.  Light leaving a splitter to the right travels z1, z2 + z3, z4,
    z5 + z6, z7 after columns 1..5; light leaving downwards travels z8..z12
    after rows 1..5.  Block 1 sits after row 2, column 2 (z2 | z3), block 2
    after row 4, column 4 (z5 | z6).  Detectors 1..5 take the last column of
    rows 1..5, detectors 7..11 the bottom row's splitters 1..5 (whose light
    also travels on), detectors 6 and 12 the two outputs of the last splitter.
    Returns the intensities (12, My, Mx) = |Ez|^2 per detector."""
    r, f, xout, yout = fixed_params[:4]
    ph11, ph12, ph21, ph22 = fixed_params[4:8]
    P = _as_tensors(parameters)
    z = _distances(P[0:12], distance_offset)
    bs = [RawAngle(p) for p in P[12:48]]
    e1, t1, e2, t2 = [RawAngle(p) for p in P[48:52]]
    across = [z[0], z[1] + z[2], z[3], z[4] + z[5], z[6]]
    down = z[7:12]

    def vrs(light, d):
        return light.VRS_propagation(d)[0]

    def olens(light):
        return VCZT_objective_lens(light, r, f, xout, yout)

    def right(light, i, j):
        if (i, j) == (1, 1):
            return vrs(building_block(light, ph11, ph12, z[1], e1, t1), z[2])
        if (i, j) == (3, 3):
            return vrs(building_block(light, ph21, ph22, z[4], e2, t2), z[5])
        return vrs(light, across[j])

    dets = [None] * 12
    above = [ls1, ls2, ls3, ls4, ls5, ls6]
    for i, left in enumerate((ls7, ls8, ls9, ls10, ls11, ls12)):
        below, b = [], left
        for j in range(6):
            c, d = BS_symmetric(above[j], b, bs[6 * i + j])
            if i < 5:
                below.append(vrs(d, down[i]))
            elif j < 5:
                dets[6 + j] = olens(c)
            else:
                dets[11] = olens(d)
            if j < 5:
                b = right(c, i, j)
            else:
                dets[i] = olens(c)
        above = below
    return _ez_intensities(dets)


# -- noise for robust (noise-aware) tables (optical_elements.py:2127-2300) ---------------------

_NOISE_LEVELS = {            # (distance min, max) in microns, (phase min, max) in radians
    'low': ((10.0, 50.0), (0.01, 0.05)),
    'mild': ((50.0, 500.0), (0.05, 0.5)),
    'high': ((500.0, 1000.0), (0.5, 1.0)),
    'all': ((10.0, 1000.0), (0.01, 1.0)),
}


def _noise_bounds(NS):
    level = NS['noise_level']
    if level == 'tunable':
        return tuple(NS['misalignment']), tuple(NS['phase_noise'])
    if level not in _NOISE_LEVELS:
        raise ValueError(f"noise_level must be one of {sorted(_NOISE_LEVELS) + ['tunable']}, got {level!r}")
    return _NOISE_LEVELS[level]


def shake_setup(key, resolution: int, NS: dict, device=None):
    """optical_elements.py:2127-2217: noise for every physical variable of an
    optical table, drawn with torch's generators (the reference draws from
    JAX's threefry streams, so the values differ; the distributions are the
    same).

    key: an integer seed.  NS: 'number of distances', 'number of sSLM',
    'number of wps', 'noise_level' ('low', 'mild', 'high', 'all' or 'tunable'
    with 'misalignment' and 'phase_noise' (min, max)), 'discretize'.
    Returns (noise_distances (n_d,), noise_slms (2 n_slm, res, res), noise_wps
    (2 n_wp,), noise_amps (2 n_slm, res, res), key0, key): distances uniform in
    [min_d, max_d) with no sign; the others a random sign times a uniform
    magnitude in [min_phase, max_phase).  As in the reference, the SLM phase
    and amplitude noise share their magnitudes pixel by pixel (both come from
    one key) with independent signs, and with 'discretize' the sign is cast to
    uint8, so a negative sign becomes 255 for the SLM and amplitude noise.
    key0 is the seed for the next call.  All arrays are float64 on `device`."""
    (d_lo, d_hi), (p_lo, p_hi) = _noise_bounds(NS)
    dev = torch.device(device) if device is not None else torch.device('cpu')
    seeds = torch.randint(0, 2 ** 62, (5,), generator=torch.Generator().manual_seed(int(key)))
    key0, k1, k2, k3, k4 = (int(s) for s in seeds)

    def gen(seed):
        return torch.Generator(device=dev).manual_seed(seed)

    def uniform(shape, lo, hi, g):
        return torch.rand(shape, dtype=torch.float64, device=dev, generator=g) * (hi - lo) + lo

    def sign(shape, g, wrap):
        s = torch.randint(0, 2, shape, device=dev, generator=g).to(torch.float64) * 2 - 1
        return torch.where(s < 0, torch.full_like(s, 255.0), s) if wrap else s

    n_slm = 2 * int(NS['number of sSLM'])
    n_wp = 2 * int(NS['number of wps'])
    wrap = bool(NS['discretize'])
    masks = (n_slm, resolution, resolution)
    g2 = gen(k2)
    slm_sign = sign(masks, g2, wrap)
    magnitude = uniform(masks, p_lo, p_hi, g2)
    distances = uniform((int(NS['number of distances']),), d_lo, d_hi, gen(k1))
    amps = sign(masks, gen(k4), wrap) * magnitude
    slms = slm_sign * magnitude
    g3 = gen(k3)
    wps = sign((n_wp,), g3, False) * uniform((n_wp,), p_lo, p_hi, g3)
    return distances, slms, wps, amps, key0, int(key)


def batch_shake_setup(keys, resolution: int, NS: dict, device=None):
    """The per-table noise of a noise-aware batch (noisy_optimization.ipynb,
    cells 20-25: shake_setup vmapped over one key per table): stacks of
    (T, n_d), (T, 2 n_slm, res, res), (T, 2 n_wp), (T, 2 n_slm, res, res), the
    next keys and the keys used."""
    out = [shake_setup(k, resolution, NS, device) for k in keys]
    d, s, w, a = (torch.stack([o[i] for o in out]) for i in range(4))
    return d, s, w, a, [o[4] for o in out], [o[5] for o in out]


def batch_robust_discovery(ls1, ls2, ls3, ls4, ls5, ls6, parameters, fixed_params, noise_distances, noise_slms,
                           noise_wps, noise_amps, distance_offset=15):
    """robust_discovery over a batch of tables that share sources and
    parameters and differ in their noise (noisy_optimization.ipynb, cell 7:
    vmap over axis 0 of the four noise stacks).  Every table runs its own
    propagation distances, so the tables run one after another on the device.
    Returns (T, 6, My, Mx)."""
    return torch.stack([robust_discovery(ls1, ls2, ls3, ls4, ls5, ls6, parameters, fixed_params, noise_distances[t],
                                         noise_slms[t], noise_wps[t], noise_amps[t], distance_offset)
                        for t in range(len(noise_distances))])
